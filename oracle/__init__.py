"""FP64 CPU oracle for the KDK Schroedinger-Poisson path (TEST INFRASTRUCTURE).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package, and only to compare against.  Product code
(``qaxions_amd``) never imports it.

Thin ctypes/numpy wrappers over ``liboqx.so`` built from ``qx_oracle.c``
(``make -C oracle``).  Arrays follow the reference's layouts: psi is complex128
of shape ``(N,)*dim``, V float64 of the same shape, Vhat complex128 of shape
``(N,)*(dim-1) + (N//2+1,)``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboqx.so")
_lib = None

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


class Physics(C.Structure):
    _fields_ = [("pref", C.c_double), ("rho_bar", C.c_double),
                ("k0_mode", C.c_int), ("rho_bar_mode", C.c_int)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        i, d, u64 = C.c_int, C.c_double, C.c_uint64
        sig = {
            "qxo_fft_c2c": (None, [i, i, _dp, i]),
            "qxo_fft_r2c": (None, [i, i, _dp, _dp]),
            "qxo_fft_c2r": (None, [i, i, _dp, _dp]),
            "qxo_fft_line": (None, [i, _dp, i]),
            "qxo_kick": (None, [i, i, _dp, _dp, d]),
            "qxo_drift": (None, [i, i, d, _dp, d]),
            "qxo_poisson": (None, [i, i, d, _dp, i]),
            "qxo_update_potential": (d, [i, i, d, _dp, _dp, _dp, d, d, i, i]),
            "qxo_compute_energy": (d, [i, i, _dp, _dp]),
            "qxo_density_stats": (None, [i, i, _dp, _dp, _dp]),
            "qxo_mass_energy": (None, [i, i, d, _dp, _dp, _dp, _dp, _dp]),
            "qxo_spectrum": (None, [i, i, d, _dp, i, i, _dp, _i64p, _dp]),
            "qxo_leapfrog": (None, [i, i, d, C.POINTER(Physics), _dp, _dp, _dp, d, i, i, i]),
            "qxo_propagate": (None, [i, i, d, C.POINTER(Physics), _dp, _dp, _dp, d]),
            "qxo_mre_schedule": (None, [d, d, i, i, i, _dp, _dp, _dp, _dp, _dp]),
            "qxo_leapfrog_schedule": (None, [i, i, d, C.POINTER(Physics), _dp, _dp, _dp, i,
                                             _dp, _dp, _dp]),
            "qxo_meas_list": (i, [C.c_int64, C.c_int64, _i64p]),
            "qxo_ic_solitons": (None, [i, i, d, d, _dp, i, _dp]),
            "qxo_rng_u64": (u64, [u64, u64, u64]),
            "qxo_rng_uniform": (d, [u64, u64, u64]),
            "qxo_rng_normal": (d, [u64, u64, u64]),
            "qxo_ic_grf": (None, [i, i, d, u64, d, _dp]),
            "qxo_ic_blob_table": (None, [i, u64, d, _dp]),
            "qxo_ic_blobs": (None, [i, i, d, d, _dp, i, _dp]),
            "qxo_set_threads": (None, [i]),
            "qxo_get_threads": (i, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _shape(dim, N):
    return (N,) * dim


def _half_shape(dim, N):
    return (N,) * (dim - 1) + (N // 2 + 1,)


def _dimN(a, half=False):
    dim = a.ndim
    N = a.shape[0]
    return dim, N


def set_threads(n: int) -> None:
    lib().qxo_set_threads(int(n))


def get_threads() -> int:
    return lib().qxo_get_threads()


# ---- transforms -------------------------------------------------------------
def fft_c2c(x: np.ndarray, sign: int) -> np.ndarray:
    y = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().qxo_fft_c2c(y.ndim, y.shape[0], _p(y.view(np.float64)), int(sign))
    return y


def fft_r2c(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    dim, N = x.ndim, x.shape[0]
    out = np.zeros(_half_shape(dim, N), np.complex128)
    lib().qxo_fft_r2c(dim, N, _p(x), _p(out.view(np.float64)))
    return out


def fft_c2r(xh: np.ndarray) -> np.ndarray:
    xh = np.ascontiguousarray(xh, dtype=np.complex128)
    dim, N = xh.ndim, xh.shape[0]
    out = np.zeros(_shape(dim, N), np.float64)
    lib().qxo_fft_c2r(dim, N, _p(xh.view(np.float64)), _p(out))
    return out


def fft_line(x: np.ndarray, sign: int) -> np.ndarray:
    y = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().qxo_fft_line(y.shape[0], _p(y.view(np.float64)), int(sign))
    return y


# ---- operators --------------------------------------------------------------
def kick(psi: np.ndarray, V: np.ndarray, dt: float) -> np.ndarray:
    y = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    V = np.ascontiguousarray(V, dtype=np.float64)
    lib().qxo_kick(y.ndim, y.shape[0], _p(y.view(np.float64)), _p(V), float(dt))
    return y


def drift(psi: np.ndarray, L: float, dt: float) -> np.ndarray:
    y = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    lib().qxo_drift(y.ndim, y.shape[0], float(L), _p(y.view(np.float64)), float(dt))
    return y


def poisson(Vhat: np.ndarray, L: float, k0_mode: int = 1) -> np.ndarray:
    y = np.ascontiguousarray(Vhat, dtype=np.complex128).copy()
    lib().qxo_poisson(y.ndim, y.shape[0], float(L), _p(y.view(np.float64)), int(k0_mode))
    return y


def update_potential(psi, L, pref, rho_bar=1.0, k0_mode=1, rho_bar_mode=0):
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    dim, N = psi.ndim, psi.shape[0]
    V = np.zeros(_shape(dim, N))
    Vh = np.zeros(_half_shape(dim, N), np.complex128)
    vmax = lib().qxo_update_potential(dim, N, float(L), _p(psi.view(np.float64)), _p(V),
                                      _p(Vh.view(np.float64)), float(pref), float(rho_bar),
                                      int(k0_mode), int(rho_bar_mode))
    return V, Vh, vmax


def compute_energy(psi):
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    V = np.zeros(psi.shape)
    r = lib().qxo_compute_energy(psi.ndim, psi.shape[0], _p(psi.view(np.float64)), _p(V))
    return V, r


def density_stats(psi):
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    a, b = C.c_double(), C.c_double()
    lib().qxo_density_stats(psi.ndim, psi.shape[0], _p(psi.view(np.float64)), C.byref(a),
                            C.byref(b))
    return a.value, b.value


def mass_energy(psi, V, L):
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    Vp = None if V is None else _p(np.ascontiguousarray(V, dtype=np.float64))
    m, k, p = C.c_double(), C.c_double(), C.c_double()
    lib().qxo_mass_energy(psi.ndim, psi.shape[0], float(L), _p(psi.view(np.float64)), Vp,
                          C.byref(m), C.byref(k), C.byref(p))
    return m.value, k.value, p.value


def spectrum(field, L, nbins=None, spb=1):
    field = np.ascontiguousarray(field, dtype=np.float64)
    dim, N = field.ndim, field.shape[0]
    if nbins is None:
        nbins = N // 2 + 1
    Pk = np.zeros(nbins)
    cnt = np.zeros(nbins, np.int64)
    k = np.zeros(nbins)
    lib().qxo_spectrum(dim, N, float(L), _p(field), int(nbins), int(spb), _p(Pk),
                       cnt.ctypes.data_as(_i64p), _p(k))
    return Pk, cnt, k


def leapfrog(psi, V, L, ds, nsteps, pref, rho_bar=1.0, k0_mode=1, rho_bar_mode=0,
             first_half=1, last_half=1):
    psi = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    V = np.ascontiguousarray(V, dtype=np.float64).copy()
    dim, N = psi.ndim, psi.shape[0]
    Vh = np.zeros(_half_shape(dim, N), np.complex128)
    ph = Physics(pref, rho_bar, k0_mode, rho_bar_mode)
    lib().qxo_leapfrog(dim, N, float(L), C.byref(ph), _p(psi.view(np.float64)), _p(V),
                       _p(Vh.view(np.float64)), float(ds), int(nsteps), int(first_half),
                       int(last_half))
    return psi, V


def leapfrog_inplace(psi, V, Vh, L, ds, nsteps, pref, rho_bar=1.0, k0_mode=1, rho_bar_mode=0,
                     first_half=1, last_half=1):
    """leapfrog on caller arrays, updated in place (no copies: timing of the
    C loop alone); psi complex128, V float64 and Vh complex128 (half
    spectrum), all C-contiguous"""
    for a, t in ((psi, np.complex128), (V, np.float64), (Vh, np.complex128)):
        if a.dtype != t or not a.flags.c_contiguous:
            raise ValueError("leapfrog_inplace needs C-contiguous complex128/float64/complex128 arrays")
    ph = Physics(pref, rho_bar, k0_mode, rho_bar_mode)
    lib().qxo_leapfrog(psi.ndim, psi.shape[0], float(L), C.byref(ph), _p(psi.view(np.float64)), _p(V),
                       _p(Vh.view(np.float64)), float(ds), int(nsteps), int(first_half), int(last_half))


def propagate(psi, V, L, ds, pref, rho_bar=1.0, k0_mode=1, rho_bar_mode=0):
    psi = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    V = np.ascontiguousarray(V, dtype=np.float64).copy()
    dim, N = psi.ndim, psi.shape[0]
    Vh = np.zeros(_half_shape(dim, N), np.complex128)
    ph = Physics(pref, rho_bar, k0_mode, rho_bar_mode)
    lib().qxo_propagate(dim, N, float(L), C.byref(ph), _p(psi.view(np.float64)), _p(V),
                        _p(Vh.view(np.float64)), float(ds))
    return psi, V


def mre_schedule(ai, dt, nsteps, first_half=1, last_half=1):
    """per-step constants of an MRE run: (drift_dt[n], kick_dt[n+1], pref[n], a, s)"""
    n = int(nsteps)
    dd, kk, pp = np.zeros(n), np.zeros(n + 1), np.zeros(n)
    a, s = C.c_double(), C.c_double()
    lib().qxo_mre_schedule(float(ai), float(dt), n, int(first_half), int(last_half), _p(dd),
                           _p(kk), _p(pp), C.byref(a), C.byref(s))
    return dd, kk, pp, a.value, s.value


def leapfrog_schedule(psi, V, L, drift_dt, kick_dt, pref, rho_bar=1.0, k0_mode=1,
                      rho_bar_mode=0):
    psi = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    V = np.ascontiguousarray(V, dtype=np.float64).copy()
    dim, N = psi.ndim, psi.shape[0]
    Vh = np.zeros(_half_shape(dim, N), np.complex128)
    dd = np.ascontiguousarray(drift_dt, np.float64)
    kk = np.ascontiguousarray(kick_dt, np.float64)
    pp = np.ascontiguousarray(pref, np.float64)
    assert kk.size == dd.size + 1 and pp.size == dd.size
    ph = Physics(0.0, rho_bar, k0_mode, rho_bar_mode)
    lib().qxo_leapfrog_schedule(dim, N, float(L), C.byref(ph), _p(psi.view(np.float64)), _p(V),
                                _p(Vh.view(np.float64)), int(dd.size), _p(dd), _p(kk), _p(pp))
    return psi, V


def meas_list(nsteps, nmeas):
    out = np.zeros(max(int(nmeas), 1), np.int64)
    n = lib().qxo_meas_list(int(nsteps), int(nmeas), out.ctypes.data_as(_i64p))
    return out[:n].tolist()


# ---- initial conditions -------------------------------------------------------
def ic_solitons(dim, N, L, solitons, bkg=1.0):
    sol = np.ascontiguousarray(np.asarray(solitons, np.float64).reshape(-1, 5))
    psi = np.zeros(_shape(dim, N), np.complex128)
    lib().qxo_ic_solitons(dim, N, float(L), float(bkg), _p(sol), sol.shape[0],
                          _p(psi.view(np.float64)))
    return psi


def ic_grf(dim, N, L, seed, amp=0.1):
    psi = np.zeros(_shape(dim, N), np.complex128)
    lib().qxo_ic_grf(dim, N, float(L), int(seed), float(amp), _p(psi.view(np.float64)))
    return psi


def blob_table(nblobs, seed, L):
    t = np.zeros((nblobs, 6))
    lib().qxo_ic_blob_table(int(nblobs), int(seed), float(L), _p(t))
    return t


def ic_blobs(dim, N, L, table, bkg=0.1):
    t = np.ascontiguousarray(table, np.float64)
    psi = np.zeros(_shape(dim, N), np.complex128)
    lib().qxo_ic_blobs(dim, N, float(L), float(bkg), _p(t), t.shape[0],
                       _p(psi.view(np.float64)))
    return psi


def rng_u64(seed, stream, idx):
    return lib().qxo_rng_u64(seed, stream, idx)


def rng_normal(seed, stream, idx):
    return lib().qxo_rng_normal(seed, stream, idx)


def rng_uniform(seed, stream, idx):
    return lib().qxo_rng_uniform(seed, stream, idx)
