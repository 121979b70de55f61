"""Pins the FP64 CPU oracle (oracle/qx_oracle.c) before anything is compared
with it: the reference's own known-answer tests (restated in tests/refmath.py
and frozen in tests/golden/golden_field.npz), brute-force DFTs, closed forms
and invariants.  Each check is chosen so that a plausible slip in the oracle
(a dropped 1/S, a wrong sign, a transposed axis, the wrong k = 0 rule) fails.
"""
import math

import numpy as np
import pytest

import oracle as o
import refmath as rm

rng = np.random.default_rng(1234)


def brute_dft(x, sign):
    """O(S^2) multi-dimensional DFT, straight from the definition."""
    y = np.asarray(x, complex)
    for ax in range(y.ndim):
        n = y.shape[ax]
        j = np.arange(n)
        W = np.exp(sign * 2j * np.pi * np.outer(j, j) / n)
        y = np.moveaxis(np.tensordot(W, np.moveaxis(y, ax, 0), axes=(1, 0)), 0, ax)
    return y


# ---------------------------------------------------------------- transforms
@pytest.mark.parametrize("dim,N", [(2, 2), (2, 4), (2, 8), (2, 6), (3, 4), (3, 8), (3, 5), (2, 16)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_c2c_matches_brute_force(dim, N, sign):
    x = rng.standard_normal((N,) * dim) + 1j * rng.standard_normal((N,) * dim)
    got = o.fft_c2c(x, sign)
    want = brute_dft(x, sign)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("dim,N", [(2, 64), (2, 256), (3, 32)])
def test_c2c_matches_numpy(dim, N):
    x = rng.standard_normal((N,) * dim) + 1j * rng.standard_normal((N,) * dim)
    assert np.abs(o.fft_c2c(x, -1) - np.fft.fftn(x)).max() < 1e-13 * np.abs(x).sum() ** 0.5 * N
    assert np.allclose(o.fft_c2c(x, +1), np.fft.ifftn(x) * x.size, rtol=0, atol=1e-11)


@pytest.mark.parametrize("dim,N", [(2, 8), (3, 8), (2, 32)])
def test_impulse_plane_wave_parseval_roundtrip(dim, N):
    S = N ** dim
    d = np.zeros((N,) * dim, complex)
    d.flat[0] = 1.0
    assert np.allclose(o.fft_c2c(d, -1), 1.0, atol=1e-15)  # impulse -> flat
    n0 = (1, 2, 3)[:dim]
    grid = np.indices((N,) * dim)
    pw = np.exp(2j * np.pi * sum(n0[a] * grid[a] for a in range(dim)) / N)
    F = o.fft_c2c(pw, -1)
    expect = np.zeros_like(F)
    expect[n0] = S
    assert np.abs(F - expect).max() < 1e-11
    x = rng.standard_normal((N,) * dim) + 1j * rng.standard_normal((N,) * dim)
    X = o.fft_c2c(x, -1)
    assert math.isclose((np.abs(X) ** 2).sum() / S, (np.abs(x) ** 2).sum(), rel_tol=1e-13)
    assert np.abs(o.fft_c2c(X, +1) / S - x).max() < 1e-13


@pytest.mark.parametrize("dim,N", [(2, 2), (2, 8), (3, 4), (3, 8), (2, 6), (2, 64)])
def test_r2c_c2r(dim, N):
    x = rng.standard_normal((N,) * dim)
    H = o.fft_r2c(x)
    assert H.shape == (N,) * (dim - 1) + (N // 2 + 1,)
    assert np.abs(H - np.fft.rfftn(x)).max() < 1e-12 * N ** dim
    back = o.fft_c2r(H)
    assert np.abs(back - x * x.size).max() < 1e-11 * N ** dim


# ---------------------------------------------------------------- operators
@pytest.mark.parametrize("n,N,dt,idx", rm.PARAMS)
def test_reference_kick_drift_kds_cases(n, N, dt, idx, golden_dir):
    """tests/test_field.py:12-65,94-129 against the oracle."""
    g = np.load(f"{golden_dir}/golden_field.npz")
    psi = o.ic_solitons(n, N, 1.0, rm.read_solitons(), bkg=1.0)
    pref = 4 * math.pi  # norm = 1 (src/field.cxx:79)
    V, _, _ = o.update_potential(psi, 1.0, pref)
    assert o.kick(psi, V, dt).flat[idx].real == pytest.approx(g[f"kick_{n}_{N}_{dt}_{idx}"], abs=1e-12)
    assert o.drift(psi, 1.0, dt).flat[idx].real == pytest.approx(g[f"drift_{n}_{N}_{dt}_{idx}"], abs=1e-12)
    p2, _ = o.propagate(psi, V, 1.0, dt, pref)
    assert p2.flat[idx].real == pytest.approx(g[f"kds_{n}_{N}_{dt}_{idx}"], abs=1e-12)


@pytest.mark.parametrize("n,N,idx", rm.PARAMS_POT)
def test_reference_potential_cases(n, N, idx, golden_dir):
    g = np.load(f"{golden_dir}/golden_field.npz")
    psi = o.ic_solitons(n, N, 1.0, rm.read_solitons(), bkg=1.0)
    V, _, vmax = o.update_potential(psi, 1.0, 4 * math.pi)
    assert V.flat[idx] == pytest.approx(g[f"potential_{n}_{N}_{idx}"], abs=1e-12)
    assert vmax == pytest.approx(np.abs(V).max(), rel=1e-15)


@pytest.mark.parametrize("n,N", [(2, 8), (3, 8), (2, 16)])
def test_whole_arrays_against_golden(n, N, golden_dir):
    g = np.load(f"{golden_dir}/golden_field.npz")
    psi = o.ic_solitons(n, N, 1.0, rm.read_solitons(), bkg=1.0)
    assert np.abs(psi - g[f"arr_psi0_{n}_{N}"]).max() < 1e-14
    V, _, _ = o.update_potential(psi, 1.0, 4 * math.pi)
    assert np.abs(V - g[f"arr_potential_{n}_{N}"]).max() < 1e-12
    assert np.abs(o.kick(psi, V, 0.1) - g[f"arr_kick_{n}_{N}"]).max() < 1e-12
    assert np.abs(o.drift(psi, 1.0, 0.1) - g[f"arr_drift_{n}_{N}"]).max() < 1e-12
    p2, _ = o.propagate(psi, V, 1.0, 0.1, 4 * math.pi)
    assert np.abs(p2 - g[f"arr_kds_{n}_{N}"]).max() < 1e-12


@pytest.mark.parametrize("dim,N,L", [(2, 16, 10.0), (3, 8, 3.0)])
def test_drift_plane_wave_exact(dim, N, L):
    """A plane wave e^{i k0.x} acquires exactly e^{-i dt k0^2/2}."""
    n0 = np.array((2, -3, 1)[:dim])
    grid = np.indices((N,) * dim)
    phase = 2 * np.pi * sum(n0[a] * grid[a] for a in range(dim)) / N
    psi = np.exp(1j * phase)
    dt = 0.37
    k2 = (2 * np.pi / L) ** 2 * (n0 ** 2).sum()
    got = o.drift(psi, L, dt)
    assert np.abs(got - psi * np.exp(-0.5j * dt * k2)).max() < 1e-13


@pytest.mark.parametrize("dim,N,L", [(2, 16, 10.0), (3, 8, 2.0)])
def test_potential_of_cosine_density(dim, N, L):
    """rho = 1 + eps cos(k0 x) -> V = -pref eps cos(k0 x)/k0^2 (rho_bar = 1)."""
    eps, pref = 0.3, 1.7
    grid = np.indices((N,) * dim)
    n0 = (1, 2, 0)[:dim]
    ph = 2 * np.pi * sum(n0[a] * grid[a] for a in range(dim)) / N
    psi = np.sqrt(1 + eps * np.cos(ph)).astype(complex)
    k02 = (2 * np.pi / L) ** 2 * sum(v * v for v in n0)
    V, _, _ = o.update_potential(psi, L, pref, rho_bar=1.0, k0_mode=1)
    assert np.abs(V - (-pref * eps * np.cos(ph) / k02)).max() < 1e-12


def test_k0_conventions():
    """Reference: the k = 0 mode is multiplied by -1/S (src/field.cxx:253,282);
    with <rho> != rho_bar that leaves V offset by -pref(<rho>-rho_bar) (1/(k^2=1))."""
    N, L, pref = 8, 1.0, 2.0
    psi = np.full((N, N), np.sqrt(1.5), complex)  # uniform rho = 1.5
    V_ref, _, _ = o.update_potential(psi, L, pref, rho_bar=1.0, k0_mode=1)
    assert np.allclose(V_ref, -pref * 0.5, atol=1e-14)
    V_zero, _, _ = o.update_potential(psi, L, pref, rho_bar=1.0, k0_mode=0)
    assert np.abs(V_zero).max() < 1e-14
    V_mean, _, _ = o.update_potential(psi, L, pref, rho_bar=7.0, k0_mode=1, rho_bar_mode=1)
    assert np.abs(V_mean).max() < 1e-14


def test_poisson_factor_layout():
    """Poisson factor on the half spectrum: half axis unfolded, others folded."""
    N, L = 8, 2.0
    Vh = np.ones((N, N, N // 2 + 1), complex)
    out = o.poisson(Vh, L, k0_mode=1)
    dk = 2 * np.pi / L
    for (i0, i1, i2) in [(0, 0, 1), (5, 3, 4), (4, 4, 0), (7, 1, 2), (0, 0, 0)]:
        n0 = i0 if i0 <= N // 2 else i0 - N
        n1 = i1 if i1 <= N // 2 else i1 - N
        m = n0 * n0 + n1 * n1 + i2 * i2
        want = -1 / N ** 3 if m == 0 else -1.0 / (dk * dk * m) / N ** 3
        assert out[i0, i1, i2].real == pytest.approx(want, rel=1e-15)


def test_leapfrog_equals_repeated_propagate():
    """Merged-kick loop (src/main.cxx:65-85) == nsteps x propagate up to roundoff."""
    N, L, dt, pref = 16, 10.0, 1e-2, 1.0
    psi = o.ic_grf(2, N, L, 3)
    V, _, _ = o.update_potential(psi, L, pref)
    a, Va = o.leapfrog(psi, V, L, dt, 5, pref)
    b, Vb = psi, V
    for _ in range(5):
        b, Vb = o.propagate(b, Vb, L, dt, pref)
    assert np.abs(a - b).max() < 1e-13
    assert np.abs(Va - Vb).max() < 1e-13


def test_leapfrog_segments_chain():
    """Segments chained with first_half=0 / last_half=0 equal one segment."""
    N, L, dt, pref = 16, 10.0, 1e-2, 1.0
    psi = o.ic_grf(2, N, L, 5)
    V, _, _ = o.update_potential(psi, L, pref)
    a, _ = o.leapfrog(psi, V, L, dt, 6, pref)
    b, Vb = o.leapfrog(psi, V, L, dt, 2, pref, first_half=1, last_half=0)
    b, Vb = o.leapfrog(b, Vb, L, dt, 4, pref, first_half=0, last_half=1)
    assert np.abs(a - b).max() < 1e-14


MRE_RHO_BAR = 1.5 * (2 * math.pi) ** 2 * 1.27 ** 2   # src/field.cxx:86


def test_mre_schedule_recursion_and_continuum_limit():
    """Field::setCosmo MRE + updateTime (src/field.cxx:82-89, 349-361),
    unrolled.  Checked in forms other than the oracle's: every step advances
    conformal time a ds by exactly 2.5e-4, so a grows by C sqrt(1 + a) 2.5e-4
    (C = 1.27^2); the continuum solution of da/dtau = C sqrt(1 + a) is
    a(tau) = (sqrt(1 + a_i) + C tau / 2)^2 - 1, and forward Euler on a convex
    solution lags it by about tau h C^2 / 4."""
    ai, dt, n = 0.1, 1e-4, 1000
    dd, kk, pp, a, s = o.mre_schedule(ai, dt, n)
    C, h = 1.27 ** 2, 2.5e-4
    assert dd[0] == dt / ai and kk[0] == 0.5 * dt / ai          # ds = dt / a_i, half opening kick
    a_prev = np.concatenate([[ai], pp[:-1]])                      # a before each updateTime
    assert np.allclose(pp - a_prev, C * h * np.sqrt(1 + a_prev), rtol=1e-10, atol=0)
    assert np.allclose(dd[1:] * a_prev[:-1], h, rtol=1e-14, atol=0)   # ds = 2.5e-4 / a
    assert np.array_equal(kk[1:-1], dd[1:])                       # merged kicks use the new ds
    assert kk[-1] == pytest.approx(0.5 * h / pp[-2], rel=1e-15)   # closing half kick
    assert a == pp[-1]
    assert s == pytest.approx(np.sum(h / a_prev), rel=1e-13)     # s sums the new ds
    tau = n * h
    exact = (math.sqrt(1 + ai) + C * tau / 2) ** 2 - 1
    lag = exact - a
    bound = tau * h * C * C / 4
    assert 0.5 * bound < lag < 1.5 * bound
    # the segment flags
    _, kk2, _, _, _ = o.mre_schedule(ai, dt, 3, first_half=0, last_half=0)
    assert kk2[0] == dt / ai and kk2[-1] == 0.0


def test_leapfrog_schedule_constant_equals_leapfrog():
    """With constant per-step values the scheduled loop is the pinned
    merged-kick loop, bit for bit."""
    N, L, dt, pref = 16, 10.0, 1e-2, 1.3
    psi = o.ic_grf(2, N, L, 8)
    V, _, _ = o.update_potential(psi, L, pref)
    a, Va = o.leapfrog(psi, V, L, dt, 5, pref)
    b, Vb = o.leapfrog_schedule(psi, V, L, [dt] * 5, [dt / 2] + [dt] * 4 + [dt / 2], [pref] * 5)
    assert np.array_equal(a, b) and np.array_equal(Va, Vb)


def test_leapfrog_schedule_is_the_operator_sequence():
    """The scheduled loop composes the pinned operators in the reference's
    order (src/main.cxx:65-85) with each step's own constants."""
    N, L = 8, 1.0
    psi = o.ic_grf(3, N, L, 4) * 0.5
    dd, kk, pp, _, _ = o.mre_schedule(0.1, 1e-4, 4)
    V, _, _ = o.update_potential(psi, L, 0.1, MRE_RHO_BAR)
    got, Vg = o.leapfrog_schedule(psi, V, L, dd, kk, pp, rho_bar=MRE_RHO_BAR)
    p = o.kick(psi, V, kk[0])
    for j in range(4):
        p = o.drift(p, L, dd[j])
        V, _, _ = o.update_potential(p, L, pp[j], MRE_RHO_BAR)
        p = o.kick(p, V, kk[j + 1])
    assert np.abs(got - p).max() < 1e-14
    assert np.abs(Vg - V).max() < 1e-12 * np.abs(V).max()


def test_mass_conserved_energy_bounded():
    N, L, dt, pref = 32, 10.0, 1e-3, 1.0
    psi = o.ic_grf(2, N, L, 11)
    V, _, _ = o.update_potential(psi, L, pref)
    m0, k0, p0 = o.mass_energy(psi, V, L)
    p, Vp = o.leapfrog(psi, V, L, dt, 50, pref)
    m1, k1, p1 = o.mass_energy(p, Vp, L)
    assert abs(m1 - m0) / m0 < 1e-13
    assert abs((k1 + p1) - (k0 + p0)) / abs(k0 + p0) < 1e-4


def _two_level_density(dim, N, lo, hi, seed):
    """psi with |psi|^2 = lo on the first half of axis 0 and hi on the second,
    with random phases (both real and imaginary parts nonzero) and the two
    halves shuffled, so a formula that drops Im psi, squares the wrong part or
    depends on site order fails."""
    r = np.random.default_rng(seed)
    rho = np.full((N,) * dim, float(lo))
    rho[N // 2:] = hi
    rho = r.permutation(rho.reshape(-1)).reshape(rho.shape)
    theta = r.uniform(0.1, 2 * np.pi - 0.1, rho.shape)
    return np.sqrt(rho) * np.exp(1j * theta), rho


@pytest.mark.parametrize("dim,N", [(2, 8), (2, 16), (3, 8)])
def test_compute_energy_closed_form(dim, N):
    """computeEnergy (src/field.cxx:200-227): V <- |psi|^2 and max/mean.  With
    |psi|^2 in {1, 3} on two halves, max/mean = 3/2 and V is the two levels."""
    psi, rho = _two_level_density(dim, N, 1.0, 3.0, 5)
    V, r = o.compute_energy(psi)
    assert np.abs(V - rho).max() < 1e-14
    assert r == pytest.approx(1.5, rel=1e-14)
    # a single spike of height A on a unit background: max/mean = A S/(S-1+A)
    S = N ** dim
    spike = np.ones((N,) * dim, complex) * np.exp(0.7j)
    spike.reshape(-1)[S // 3] = math.sqrt(9.0) * np.exp(-1.1j)
    V, r = o.compute_energy(spike)
    assert r == pytest.approx(9.0 * S / (S - 1 + 9.0), rel=1e-14)
    assert V.reshape(-1)[S // 3] == pytest.approx(9.0, rel=1e-15)


@pytest.mark.parametrize("dim,N", [(2, 8), (3, 8)])
def test_density_stats_closed_form(dim, N):
    """The same statistics without touching V: max/mean = hi/((lo+hi)/2) and
    sum |psi|^2 = S (lo + hi)/2 (mass = (L/N)^dim times that)."""
    S = N ** dim
    for lo, hi in [(1.0, 3.0), (0.1, 10.0), (2.0, 2.0)]:
        psi, _ = _two_level_density(dim, N, lo, hi, 7)
        r, s = o.density_stats(psi)
        assert r == pytest.approx(hi / ((lo + hi) / 2), rel=1e-14)
        assert s == pytest.approx(S * (lo + hi) / 2, rel=1e-14)
    # agrees with computeEnergy's max/mean on any state, and mass_energy's mass
    psi = o.ic_grf(dim, N, 6.0, 3)
    r, s = o.density_stats(psi)
    _, r2 = o.compute_energy(psi)
    m, _, _ = o.mass_energy(psi, None, 6.0)
    assert r == pytest.approx(r2, rel=1e-15)
    assert m == pytest.approx(s * (6.0 / N) ** dim, rel=1e-14)


def test_kinetic_energy_closed_form():
    N, L = 16, 4.0
    grid = np.indices((N, N))
    psi = 2.0 * np.exp(2j * np.pi * (3 * grid[0] + grid[1]) / N)
    m, k, _ = o.mass_energy(psi, np.zeros((N, N)), L)
    assert m == pytest.approx(4.0 * L * L, rel=1e-14)
    assert k == pytest.approx(0.5 * m * (2 * np.pi / L) ** 2 * 10, rel=1e-13)


# ---------------------------------------------------------------- spectrum
def test_spectrum_single_cosine():
    N, L = 16, 10.0
    grid = np.indices((N, N))
    f = 1.0 + 0.5 * np.cos(2 * np.pi * 3 * grid[1] / N)
    P, cnt, k = o.spectrum(f, L)
    S = N * N
    # k = 0 shell: a single mode with |F|^2 = S^2 -> P0 = 1 (mean^2)
    assert P[0] == pytest.approx(1.0, rel=1e-14) and cnt[0] == 1
    # |n| = 3: two stored modes (ix = 3) each weight 2 -> shell sum 2*(S/4)^2/S^2
    shell_count = sum(1 for a in range(-N // 2 + 1, N // 2 + 1) for b in range(-N // 2 + 1, N // 2 + 1)
                      if round(math.sqrt(a * a + b * b)) == 3)
    assert cnt[3] == shell_count
    assert P[3] * cnt[3] == pytest.approx(2 * (0.25 ** 2), rel=1e-12)
    assert k[3] == pytest.approx(3 * 2 * np.pi / L)


@pytest.mark.parametrize("dim,N", [(2, 16), (3, 8), (2, 10)])
def test_spectrum_counts_and_weights_vs_full_spectrum(dim, N):
    """Hermitian half-spectrum weights reproduce the full-spectrum shell sums."""
    f = rng.standard_normal((N,) * dim)
    P, cnt, _ = o.spectrum(f, 1.0)
    F = np.fft.fftn(f)
    n = np.fft.fftfreq(N, 1.0 / N)
    grids = np.meshgrid(*([n] * dim), indexing="ij")
    shell = np.rint(np.sqrt(sum(g ** 2 for g in grids))).astype(int)
    # the half spectrum double counts ix in (0, N/2) and counts ix in {0, N/2} once:
    # over the full spectrum that equals all modes except the mirrored Nyquist half
    last = np.abs(grids[-1])
    S = N ** dim
    for b in range(N // 2 + 1):
        sel = (shell == b)
        # modes whose last index is -N/2 (only for even N) are not in the half spectrum
        hx = np.rint(np.moveaxis(np.indices((N,) * dim), 0, -1)[..., -1])
        full_sel = sel & ~((hx > N // 2))
        w = np.where((hx == 0) | (hx == N // 2), 1, 2)
        want_cnt = (w * full_sel).sum()
        assert cnt[b] == want_cnt
        if want_cnt:
            want = (w * full_sel * np.abs(F) ** 2).sum() / S ** 2 / want_cnt
            assert P[b] == pytest.approx(want, rel=1e-12)
    del last


def test_spectrum_linear_bins_merge_shells():
    N = 16
    f = rng.standard_normal((N, N))
    P1, c1, _ = o.spectrum(f, 1.0)
    P2, c2, k2 = o.spectrum(f, 1.0, nbins=4, spb=2)
    for j in range(4):
        cc = c1[2 * j] + c1[2 * j + 1]
        assert c2[j] == cc
        assert P2[j] == pytest.approx((P1[2 * j] * c1[2 * j] + P1[2 * j + 1] * c1[2 * j + 1]) / cc, rel=1e-13)


# ---------------------------------------------------------------- schedule, ICs, RNG
@pytest.mark.parametrize("nsteps,nmeas,want", [
    (10, 20, [k * 9 // 19 for k in range(20)]), (50, 50, list(range(50))), (7, 1, [6]),
    (0, 5, []), (5, 0, []), (100, 3, [0, 49, 99])])
def test_meas_list(nsteps, nmeas, want):
    assert o.meas_list(nsteps, nmeas) == want


@pytest.mark.parametrize("dim,N", [(2, 4), (2, 8), (3, 4), (3, 8)])
def test_solitons_ic_matches_reference_helper(dim, N):
    """tests/test_ic.py:12-37 (psi values at the reference's indices)."""
    got = o.ic_solitons(dim, N, 1.0, rm.read_solitons(), bkg=1.0)
    want = rm.solitons(dim, N)
    assert np.abs(got - want).max() < 1e-14


def test_rng_matches_python_restatement():
    from qaxions_amd import rng as prng
    for seed, st, i in [(0, 0, 0), (7, 1, 12345), (2 ** 40 + 3, 17, 99)]:
        assert o.rng_u64(seed, st, i) == prng.u64(seed, st, i)
        assert o.rng_uniform(seed, st, i) == prng.uniform(seed, st, i)


@pytest.mark.parametrize("dim,N", [(2, 64), (3, 16)])
def test_grf_ic_properties(dim, N):
    L = 10.0
    psi = o.ic_grf(dim, N, L, 42, amp=0.1)
    rho = np.abs(psi) ** 2
    assert rho.mean() == pytest.approx(1.0, abs=1e-12)          # <delta> = 0
    assert (rho - 1).std() == pytest.approx(0.1, rel=1e-9)      # sigma normalisation
    D = np.fft.fftn(rho - 1)
    n = np.fft.fftfreq(N, 1.0 / N)
    m = sum(g ** 2 for g in np.meshgrid(*([n] * dim), indexing="ij"))
    assert np.abs(D[16 * m > N * N]).max() < 1e-9 * np.abs(D).max()  # cut at N/4
    # theta is low-passed too
    th = np.angle(psi)
    T = np.fft.fftn(np.exp(1j * th) * 0 + th)
    assert np.isfinite(T).all()
    # deterministic in the seed
    assert np.array_equal(psi, o.ic_grf(dim, N, L, 42, amp=0.1))
    assert not np.array_equal(psi, o.ic_grf(dim, N, L, 43, amp=0.1))


def test_blobs_ic_closed_form():
    tab = o.blob_table(32, 9, 20.0)
    assert tab.shape == (32, 6)
    assert (tab[:, 0] >= 1).all() and (tab[:, 0] <= 10).all()
    assert (tab[:, 1] >= 0.3).all() and (tab[:, 1] <= 1).all()
    one = tab[:1].copy()
    one[0, 2:5] = [5.0, 5.0, 5.0]
    one[0, 5] = 0.0
    psi = o.ic_blobs(3, 16, 20.0, one, bkg=0.1)
    c = 4  # lattice point x = 4 * 20/16 = 5.0
    assert psi[c, c, c].real == pytest.approx(math.sqrt(0.1) + math.sqrt(one[0, 0]), rel=1e-14)
    # minimum image: distance across the periodic boundary
    one[0, 2:5] = [0.0, 0.0, 0.0]
    psi = o.ic_blobs(3, 16, 20.0, one, bkg=0.1)
    assert psi[0, 0, 15] == pytest.approx(psi[0, 0, 1], rel=1e-14)
