"""numpy restatement of the reference's own test oracles
(tests/helper.py:4-73, tests/test_field.py:12-129), written from their text.

Arrays are in the C++ index order (last index = x, fastest), which is what
the reference tests obtain with `.transpose(2, 1, 0)` in 3D
(tests/test_field.py:29).  Used as an independent check of the oracle and
of the GPU path; it shares no code with either.
"""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_solitons(path=os.path.join(GOLDEN, "solitons.txt")):
    rows = []
    with open(path) as f:
        for line in f:
            v = line.split()
            if len(v) == 5:
                rows.append([float(x) for x in v])
    return np.array(rows)


def solitons(dim, N, L=1.0, rows=None, bkg=1.0):
    """tests/helper.py:21-73 with box length L (positions i L/N, centres c L)."""
    if rows is None:
        rows = read_solitons()
    x = np.arange(N) * (L / N)
    if dim == 2:
        yy, xx = np.meshgrid(x, x, indexing="ij")
        zz = np.zeros_like(xx)
    else:
        zz, yy, xx = np.meshgrid(x, x, x, indexing="ij")
    rho = np.full(xx.shape, bkg)
    for amp, sg, xc, yc, zc in rows:
        nrm = 1.0 / (sg ** 3 * 2 * np.pi)
        r2 = (xx - xc * L) ** 2 + (yy - yc * L) ** 2
        if dim == 3:
            r2 = r2 + (zz - zc * L) ** 2
        rho = rho + amp * np.exp(-r2 / (2 * sg ** 2)) * nrm
    rho = rho / rho.mean()
    return np.sqrt(rho).astype(complex)


def k2(dim, N, L=1.0):
    """tests/helper.py:4-19: (2 pi/L)^2 |n|^2 with n the FFT frequencies."""
    n = np.fft.fftfreq(N, 1.0 / N)
    k = 2 * np.pi / L * n
    grids = np.meshgrid(*([k] * dim), indexing="ij")
    return sum(g ** 2 for g in grids)


def potential(psi, L=1.0, norm=1.0):
    """tests/test_field.py:32-33: V = Re ifftn(-fftn(4 pi norm (|psi|^2-1))/(k^2 + (k^2==0)))"""
    K2 = k2(psi.ndim, psi.shape[0], L)
    Vh = -np.fft.fftn(4.0 * np.pi * norm * (np.abs(psi) ** 2 - 1.0)) / (K2 + (K2 == 0))
    return np.real(np.fft.ifftn(Vh))


def kick(psi, dt, L=1.0, norm=1.0):
    """tests/test_field.py:32-34"""
    return psi * np.exp(-1.0j * dt * potential(psi, L, norm))


def drift(psi, dt, L=1.0):
    """tests/test_field.py:60-62"""
    K2 = k2(psi.ndim, psi.shape[0], L)
    return np.fft.ifftn(np.exp(dt * (-1.0j * K2 * 0.5)) * np.fft.fftn(psi))


def kick_drift_step(psi, dt, L=1.0, norm=1.0):
    """tests/test_field.py:114-126"""
    V = potential(psi, L, norm)
    psi = np.exp(-1.0j * dt / 2.0 * V) * psi
    psi = drift(psi, dt, L)
    V = potential(psi, L, norm)
    return np.exp(-1.0j * dt / 2.0 * V) * psi


PARAMS = [(n, N, dt, idx) for n in (2, 3) for N in (4, 8) for dt in (0.05, 0.1) for idx in (1, 9)]
PARAMS_POT = [(n, N, idx) for n in (2, 3) for N in (4, 8) for idx in (1, 9)]
