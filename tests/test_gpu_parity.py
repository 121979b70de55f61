"""Parity of the HIP path (through the C ABI) with the FP64 oracle on the same
seeded inputs.  Tolerances: FP64 contexts 1e-12 relative, FP32 contexts as
stated per test (the FP32 bar: <= 1e-5 relative L2 after 100 steps)."""
import json
import math

import numpy as np
import pytest

import oracle as o
import refmath as rm

pytestmark = pytest.mark.gpu

rng = np.random.default_rng(7)


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def lattice(qx, dim, N, L=1.0, prec="fp64", **ph):
    print(f"[case] dim={dim} N={N} L={L} prec={prec}", flush=True)
    return qx.Lattice(dim, N, L, precision=prec, **ph)


TOL = {"fp64": 1e-12, "fp32": 2e-6}


def spectrum_errors(P, Po, N, spb):
    """Per-bin relative error max_b |P_b - Po_b| / Po_b over the bins whose
    shells all lie inside the GRF IC's spectral cut (|n| <= N/4), and the
    largest absolute error of the other bins relative to bin 0 (those hold
    only what the evolution moved there, or round-off at the IC)."""
    shells_hi = (np.arange(len(P)) + 1) * spb - 1
    cut = shells_hi <= N // 4
    rel = np.abs(P - Po)[cut] / Po[cut]
    out = np.abs(P - Po)[~cut].max() / Po[0] if (~cut).any() else 0.0
    return float(rel.max()), float(out)

FFT_CASES = [(2, n) for n in (2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096)] + \
            [(3, n) for n in (2, 4, 8, 16, 32, 64, 128)]


# ------------------------------------------------------------------ transforms
@pytest.mark.parametrize("dim,N", FFT_CASES)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_fft_c2c(qx, dim, N, prec):
    lat = lattice(qx, dim, N, prec=prec)
    x = rng.standard_normal((N,) * dim) + 1j * rng.standard_normal((N,) * dim)
    for sign in (-1, 1):
        want = o.fft_c2c(x, sign)
        got = lat.fft_c2c(x, sign)
        assert rel_l2(got, want) < TOL[prec] * max(1.0, math.log2(N)), (sign, rel_l2(got, want))


@pytest.mark.parametrize("dim,N", FFT_CASES)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_fft_r2c_c2r(qx, dim, N, prec):
    lat = lattice(qx, dim, N, prec=prec)
    x = rng.standard_normal((N,) * dim)
    H = lat.fft_r2c(x)
    Hw = o.fft_r2c(x)
    assert rel_l2(H, Hw) < TOL[prec] * max(1.0, math.log2(N))
    back = lat.fft_c2r(Hw)
    assert rel_l2(back, o.fft_c2r(Hw)) < TOL[prec] * max(1.0, math.log2(N))


# ------------------------------------------------------------------ operators
OP_CASES = [(2, 8, 1.0), (2, 64, 10.0), (2, 256, 10.0), (3, 8, 1.0), (3, 32, 20.0)]


@pytest.mark.parametrize("dim,N,L", OP_CASES)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_kick_drift_potential(qx, dim, N, L, prec):
    pref = 4 * math.pi if L == 1.0 else 1.0
    lat = lattice(qx, dim, N, L, prec, pref=pref)
    psi = o.ic_grf(dim, N, L, 3) if N >= 16 else o.ic_solitons(dim, N, L, rm.read_solitons())
    lat.psi = psi
    vmax = lat.update_potential()
    V, _, vmax_o = o.update_potential(psi, L, pref)
    assert rel_l2(lat.V, V) < 10 * TOL[prec]
    assert vmax == pytest.approx(vmax_o, rel=10 * TOL[prec])
    lat.kick(0.1)
    p1 = o.kick(psi, V, 0.1)
    assert rel_l2(lat.psi, p1) < TOL[prec]
    lat.drift(0.05)
    p2 = o.drift(p1, L, 0.05)
    assert rel_l2(lat.psi, p2) < 2 * TOL[prec] * math.log2(N)


@pytest.mark.parametrize("k0_mode,rho_bar_mode", [(1, 0), (0, 0), (1, 1)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_potential_conventions(qx, k0_mode, rho_bar_mode, prec):
    dim, N, L = 3, 16, 20.0
    tab = o.blob_table(4, 1, L)
    psi = o.ic_blobs(dim, N, L, tab, 0.1)  # <rho> != 1
    lat = lattice(qx, dim, N, L, prec, pref=1.0, rho_bar=1.0, k0_mode=k0_mode,
                  rho_bar_mode=rho_bar_mode)
    lat.psi = psi
    lat.update_potential()
    V, _, _ = o.update_potential(psi, L, 1.0, 1.0, k0_mode, rho_bar_mode)
    assert rel_l2(lat.V, V) < 10 * TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_poisson_half_and_compute_energy(qx, prec):
    dim, N, L = 2, 32, 10.0
    lat = lattice(qx, dim, N, L, prec, pref=1.0)
    vh = o.fft_r2c(rng.standard_normal((N, N)))
    got = lat.poisson_half(vh)
    assert rel_l2(got, o.poisson(vh, L, 1)) < TOL[prec]
    psi = o.ic_grf(dim, N, L, 9)
    lat.psi = psi
    r = lat.compute_energy()
    Vo, ro = o.compute_energy(psi)
    assert r == pytest.approx(ro, rel=10 * TOL[prec])
    assert rel_l2(lat.V, Vo) < TOL[prec]
    rr, mass = lat.density_stats()
    ro2, s = o.density_stats(psi)
    assert rr == pytest.approx(ro2, rel=10 * TOL[prec])
    assert mass == pytest.approx(s * (L / N) ** 2, rel=10 * TOL[prec])


@pytest.mark.parametrize("dim,N,L", [(2, 64, 10.0), (3, 16, 20.0)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_energy(qx, dim, N, L, prec):
    lat = lattice(qx, dim, N, L, prec, pref=1.0)
    psi = o.ic_grf(dim, N, L, 21)
    lat.psi = psi
    lat.update_potential()
    V, _, _ = o.update_potential(psi, L, 1.0)
    m, k, p = lat.energy()
    mo, ko, po = o.mass_energy(psi, V, L)
    assert m == pytest.approx(mo, rel=10 * TOL[prec])
    assert k == pytest.approx(ko, rel=10 * TOL[prec])
    assert p == pytest.approx(po, rel=1e3 * TOL[prec], abs=1e-9)


@pytest.mark.parametrize("dim,N,nbins,spb", [(2, 16, 9, 1), (2, 256, 64, 2), (3, 16, 9, 1),
                                             (3, 32, 8, 2), (2, 10, 6, 1), (2, 4, 3, 1), (3, 4, 3, 1),
                                             (3, 64, 33, 1), (3, 128, 21, 3), (2, 1024, 100, 5),
                                             (2, 4096, 2049, 1), (3, 8, 5, 1)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_spectrum(qx, dim, N, nbins, spb, prec):
    if N & (N - 1):
        with pytest.raises(qx.QxError):
            qx.Lattice(dim, N, 1.0, precision=prec)
        return
    L = 10.0
    lat = lattice(qx, dim, N, L, prec)
    psi = o.ic_grf(dim, N, L, 5)
    lat.psi = psi
    P, c, k = lat.spectrum(source="rho", nbins=nbins, spb=spb)
    Po, co, ko = o.spectrum(np.abs(psi) ** 2, L, nbins, spb)
    assert np.array_equal(c, co)
    assert np.allclose(k, ko, rtol=1e-15)
    # bins inside the IC's spectral cut: per-bin relative error (FP32
    # measured <= 1.4e-7 at 512^3, 1.8e-6 at 2D 4096^2 in single-shell bins:
    # the round-off of log2(S) stages against the mean's share of |rho|^2);
    # the bins beyond it hold round-off only (|psi|^2 of this IC is
    # band-limited), compared against bin 0
    rtol, otol = (1e-12, 1e-20) if prec == "fp64" else (max(1e-6, 1e-7 * math.log2(N ** dim)), 1e-12)
    rel, out = spectrum_errors(P, Po, N, spb)
    print(f"P(k) {prec}: max rel err in the cut {rel:.2e}, outside {out:.2e}")
    assert rel <= rtol and out <= otol
    # the reference's choice: spectrum of V as it stands after computeEnergy
    lat.compute_energy()
    P2, c2, _ = lat.spectrum(source="V", nbins=nbins, spb=spb)
    assert np.array_equal(c2, co)
    rel, out = spectrum_errors(P2, Po, N, spb)
    assert rel <= rtol and out <= otol


# ------------------------------------------------------------------ reference tests
@pytest.mark.parametrize("n,N,dt,idx", rm.PARAMS)
@pytest.mark.parametrize("fp32", [False, True])
def test_reference_fieldtest_module(n, N, dt, idx, fp32, in_golden_dir):
    """tests/test_field.py:12-129 through the pybind11 module, same tolerance."""
    import qaxions_amd.fieldtest as field
    g = np.load("golden_field.npz")
    assert field.test_kick(n, N, dt, idx, fp32=fp32) == pytest.approx(
        g[f"kick_{n}_{N}_{dt}_{idx}"], abs=1e-5, rel=1e-4)
    assert field.test_drift(n, N, dt, idx, fp32=fp32) == pytest.approx(
        g[f"drift_{n}_{N}_{dt}_{idx}"], abs=1e-5, rel=1e-4)
    assert field.test_kick_drift_step(n, N, dt, idx, fp32=fp32) == pytest.approx(
        g[f"kds_{n}_{N}_{dt}_{idx}"], abs=1e-5, rel=1e-4)
    if not fp32:  # FP64 device state: parity at the oracle's level
        assert field.test_kick_drift_step(n, N, dt, idx) == pytest.approx(
            g[f"kds_{n}_{N}_{dt}_{idx}"], abs=1e-12)


@pytest.mark.parametrize("N", [4, 8])
@pytest.mark.parametrize("idx", [1, 9])
def test_reference_ictest_module(N, idx, in_golden_dir):
    """tests/test_ic.py:11-37 through the pybind11 module (the reference's
    tolerance, and the FP64 host IC to 1e-14)."""
    import qaxions_amd.ictest as ic
    for dim, f in ((2, ic.test_solitons), (3, ic.test_solitons_3D)):
        want = rm.solitons(dim, N).flat[idx].real
        got = f(N, idx)
        assert got == pytest.approx(want, abs=1e-5, rel=1e-4)
        assert abs(got - want) < 1e-14


@pytest.mark.parametrize("n,N", [(2, 8), (2, 16), (3, 8)])
@pytest.mark.parametrize("fp32", [False, True])
def test_field_vhat_mirror(n, N, fp32, in_golden_dir):
    """Field::Vhat() (src/field.h:45) through the reference's sequences:
    after updatePotential the Poisson-multiplied spectrum of the density
    (src/field.cxx:300-321); after computeEnergy and Spectrum::compute,
    r2c of V = |psi|^2 (src/spectrum/spectrum.cxx:48); Field::Poisson()
    then multiplies that spectrum in place (src/field.cxx:230-291)."""
    import qaxions_amd.fieldtest as field
    psi, vh_pot, vh_spec, vh_pois, V = field.vhat_sequence(n, N, fp32=fp32)
    shape, half = (N,) * n, (N,) * (n - 1) + (N // 2 + 1,)
    psi = psi.reshape(shape)
    L, pref = 1.0, 4 * math.pi   # defaults(): Lbox = 1, norm = 1 (STATIC)
    Vo, Vho, _ = o.update_potential(psi, L, pref)
    # FP32: V's constant -pref(<rho> - 1) (the reference's k = 0 rule) is a
    # cancellation and carries ~1e-6 absolute
    tol = 1e-12 if not fp32 else 1e-5
    assert rel_l2(V.reshape(shape), Vo) < tol
    assert rel_l2(vh_pot.reshape(half), Vho) < tol
    tol = 1e-12 if not fp32 else 2e-6
    rho_hat = o.fft_r2c(np.abs(psi) ** 2)
    assert rel_l2(vh_spec.reshape(half), rho_hat) < tol
    assert rel_l2(vh_pois.reshape(half), o.poisson(rho_hat, L, 1)) < tol


@pytest.mark.parametrize("n,N,idx", rm.PARAMS_POT)
def test_reference_potential_module(n, N, idx, in_golden_dir):
    import qaxions_amd.fieldtest as field
    g = np.load("golden_field.npz")
    assert field.test_potential(n, N, idx) == pytest.approx(g[f"potential_{n}_{N}_{idx}"], abs=1e-12)
    assert field.test_potential(n, N, idx, fp32=True) == pytest.approx(
        g[f"potential_{n}_{N}_{idx}"], abs=1e-5, rel=1e-4)


# ------------------------------------------------------------------ step compositions
@pytest.mark.parametrize("dim,N,L", [(2, 8, 1.0), (2, 32, 10.0), (3, 8, 1.0), (3, 16, 20.0),
                                     (2, 512, 10.0), (3, 64, 20.0)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_fused_leapfrog_matches_oracle(qx, dim, N, L, prec):
    ds, ns = 2e-3, 7
    lat = lattice(qx, dim, N, L, prec, pref=1.0)
    psi = o.ic_grf(dim, N, L, 13)
    lat.psi = psi
    lat.update_potential()
    V, _, _ = o.update_potential(psi, L, 1.0)
    lat.leapfrog(ds, ns)
    want, Vw = o.leapfrog(psi, V, L, ds, ns, 1.0)
    assert rel_l2(lat.psi, want) < (1e-12 if prec == "fp64" else 5e-6)
    assert rel_l2(lat.V, Vw) < (1e-11 if prec == "fp64" else 5e-5)
    assert lat.vmax == pytest.approx(np.abs(Vw).max(), rel=(1e-12 if prec == "fp64" else 1e-5))


@pytest.mark.parametrize("dim,N,L,ns", [(2, 2, 1.0, 5), (2, 4, 1.0, 5), (3, 2, 1.0, 5), (3, 4, 1.0, 5),
                                        (2, 2048, 40.0, 2), (2, 4096, 40.0, 2)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_fused_leapfrog_extreme_sizes(qx, dim, N, L, ns, prec):
    """The smallest lattices (single-stage transforms, fewer rows than a
    row block) and the largest 2D ones, against the oracle."""
    ds = 1e-3
    lat = lattice(qx, dim, N, L, prec, pref=1.0)
    psi = o.ic_grf(dim, N, L, 31) if N >= 16 else (
        rng.standard_normal((N,) * dim) + 1j * rng.standard_normal((N,) * dim) + 2.0)
    lat.psi = psi
    lat.update_potential()
    V, _, _ = o.update_potential(psi, L, 1.0)
    lat.leapfrog(ds, ns)
    want, Vw = o.leapfrog(psi, V, L, ds, ns, 1.0)
    assert rel_l2(lat.psi, want) < (1e-12 if prec == "fp64" else 5e-6)
    assert rel_l2(lat.V, Vw) < (1e-11 if prec == "fp64" else 5e-5)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_leapfrog_segments_and_propagate(qx, prec):
    dim, N, L, ds = 2, 64, 10.0, 1e-3
    psi = o.ic_grf(dim, N, L, 17)
    a = lattice(qx, dim, N, L, prec, pref=1.0)
    a.psi = psi
    a.update_potential()
    a.leapfrog(ds, 6)
    b = lattice(qx, dim, N, L, prec, pref=1.0)
    b.psi = psi
    b.update_potential()
    b.leapfrog(ds, 2, True, False)
    b.leapfrog(ds, 4, False, True)
    assert rel_l2(b.psi, a.psi) < (1e-13 if prec == "fp64" else 1e-6)
    c = lattice(qx, dim, N, L, prec, pref=1.0)
    c.psi = psi
    c.update_potential()
    for _ in range(6):
        c.propagate(ds)
    assert rel_l2(c.psi, a.psi) < (1e-13 if prec == "fp64" else 2e-6)
    # zero steps: K(ds/2) K(ds/2)
    d = lattice(qx, dim, N, L, prec, pref=1.0)
    d.psi = psi
    d.update_potential()
    V = d.V
    d.leapfrog(ds, 0)
    assert rel_l2(d.psi, o.kick(o.kick(psi, V, ds / 2), V, ds / 2)) < TOL[prec]


MRE_RHO_BAR = 1.5 * (2 * math.pi) ** 2 * 1.27 ** 2   # src/field.cxx:86


@pytest.mark.parametrize("dim,N,L,steps", [(2, 64, 10.0, 100), (3, 8, 1.0, 40), (3, 32, 20.0, 40)])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_mre_schedule_matches_oracle(qx, dim, N, L, steps, prec, tol):
    """The MRE run (src/field.cxx:82-89, 349-361; a_i = 0.1, dt = 1e-4): ds
    and pref = a change every step.  More steps than one table upload holds
    (32), untiled (8^3) and tiled (32^3) 3D."""
    dd, kk, pp, _, _ = o.mre_schedule(0.1, 1e-4, steps)
    psi = o.ic_grf(dim, N, L, 21)
    V, _, _ = o.update_potential(psi, L, 0.1, MRE_RHO_BAR)
    want, Vw = o.leapfrog_schedule(psi, V, L, dd, kk, pp, rho_bar=MRE_RHO_BAR)
    lat = lattice(qx, dim, N, L, prec, pref=0.1, rho_bar=MRE_RHO_BAR)
    lat.psi = psi
    lat.update_potential()
    lat.leapfrog_schedule(dd, kk, pp)
    err = rel_l2(lat.psi, want)
    print(f"psi rel L2 after {steps} MRE steps ({prec}): {err:.3e}")
    assert err <= tol
    assert rel_l2(lat.V, Vw) <= 10 * tol
    assert lat.vmax == pytest.approx(np.abs(Vw).max(), rel=10 * tol)


@pytest.mark.parametrize("dim,N,prec,ns", [(2, 64, "fp64", 170), (3, 32, "fp32", 100), (3, 8, "fp32", 20),
                                         (2, 1024, "fp32", 12)])
def test_constant_schedule_is_leapfrog_bitwise(qx, dim, N, prec, ns):
    """Constant per-step values run the same kernels with the same tables as
    qx_leapfrog (graph-captured there for these sizes).  170 steps take six
    chunks of uploaded tables through both buffers; a second run continues
    from the first, and a third runs with new physics."""
    L, ds = 10.0, 1e-3
    psi = o.ic_grf(dim, N, L, 22)
    a = lattice(qx, dim, N, L, prec, pref=1.0)
    a.psi = psi
    a.update_potential()
    b = lattice(qx, dim, N, L, prec, pref=1.0)
    b.psi = psi
    b.update_potential()
    for run in range(3):
        if run == 2:
            for lat in (a, b):
                lat.set_physics(2.0, rho_bar=1.0, k0_mode=0)
                lat.update_potential()
        a.leapfrog(ds, ns)
        b.leapfrog_schedule([ds] * ns, [ds / 2] + [ds] * (ns - 1) + [ds / 2], [a.get_physics()["pref"]] * ns)
        assert np.array_equal(a.psi, b.psi), run
        assert np.array_equal(a.V, b.V), run
    with pytest.raises(qx.QxError):
        b.leapfrog_schedule([], [ds], [])
    with pytest.raises(qx.QxError):
        b.leapfrog_schedule([ds], [ds, float("nan")], [1.0])


@pytest.mark.parametrize("n,N,fp32", [(2, 32, False), (3, 16, False), (3, 16, True)])
def test_field_mre_leapfrog_fused_equals_operators(n, N, fp32, in_golden_dir):
    """Field::leapfrog in an MRE run (one fused call) against the reference's
    operator-by-operator loop on the same Field; a, ds and s against the
    oracle's unrolled schedule."""
    import qaxions_amd.fieldtest as field
    print(f"[case] mre n={n} N={N} fp32={fp32}", flush=True)
    steps = 36
    pf, af, dsf, sf = field.mre_leapfrog(n, N, steps, fused=True, fp32=fp32)
    po, ao, dso, so = field.mre_leapfrog(n, N, steps, fused=False, fp32=fp32)
    _, kk, pp, a, s = o.mre_schedule(0.1, 1e-4, steps)
    assert af == ao == a and sf == so == s
    assert dsf == dso and dsf == pytest.approx(2 * kk[-1], rel=1e-15)
    assert rel_l2(pf, po) < (1e-12 if not fp32 else 1e-5)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_plain_method_256sq_100_steps(qx, prec, tol):
    """2D 256^2, L = 10, dt = 1e-3, 100 KDK steps, GRF IC: psi relative L2
    against the FP64 oracle (<= 1e-5 FP32, <= 1e-12 FP64), mass drift and
    the 64-bin P(k) at the end."""
    dim, N, L, dt, steps = 2, 256, 10.0, 1e-3, 100
    psi0 = o.ic_grf(dim, N, L, 0)
    V0, _, _ = o.update_potential(psi0, L, 1.0)
    want, _ = o.leapfrog(psi0, V0, L, dt, steps, 1.0)
    lat = lattice(qx, dim, N, L, prec, pref=1.0)
    lat.psi = psi0
    lat.update_potential()
    lat.leapfrog(dt, steps)
    got = lat.psi
    err = rel_l2(got, want)
    print(f"psi rel L2 after {steps} steps ({prec}): {err:.3e}")
    assert err <= tol
    m0 = (np.abs(psi0) ** 2).sum()
    # FP32: the drift's transforms take two-word twiddles, so 100 steps of
    # single-precision passes keep the mass to 3.4e-8 (one-word twiddles:
    # ~1e-6, growing with every step); what is left is mostly the FP32
    # rounding of the per-axis drift factors, the same every step
    assert abs((np.abs(got) ** 2).sum() - m0) / m0 < (1e-13 if prec == "fp64" else 1e-7)
    # 64-bin P(k) of |psi|^2 at the end against the oracle's: per-bin
    # relative error inside the IC's cut (FP32 measured 8.3e-7: the state's
    # own error, the spectrum of one state agrees to 1e-7)
    P, c, _ = lat.spectrum(source="rho", nbins=64, spb=2)
    Po, co, _ = o.spectrum(np.abs(want) ** 2, L, 64, 2)
    assert np.array_equal(c, co)
    rel, out = spectrum_errors(P, Po, N, 2)
    print(f"P(k) after {steps} steps ({prec}): max rel err in the cut {rel:.2e}, outside {out:.2e}")
    assert rel <= (1e-12 if prec == "fp64" else 5e-6)
    assert out <= (1e-16 if prec == "fp64" else 1e-10)


@pytest.mark.parametrize("N", [32, 64])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_config1_blobs_vs_oracle(qx, N, prec, tol):
    """The 3D soliton-blob workload (32 seeded Gaussian blobs on background
    0.1, L = 20, dt = 5e-4, 200 KDK steps, mean density subtracted) at 32^3
    and 64^3, where the FP64 oracle runs in seconds: psi against the oracle
    (FP32 measured 2.7e-6 and 2.9e-6), and the oracle's own mass and energy
    conservation reproduced."""
    L, dt, steps = 20.0, 5e-4, 200
    tab = o.blob_table(32, 7, L)
    psi0 = o.ic_blobs(3, N, L, tab, 0.1)
    V0, _, _ = o.update_potential(psi0, L, 1.0, 1.0, 0, 0)
    want, Vw = o.leapfrog(psi0, V0, L, dt, steps, 1.0, 1.0, 0, 0)
    lat = qx.Lattice(3, N, L, prec, pref=1.0, rho_bar=1.0, k0_mode=0)
    lat.psi = psi0
    lat.update_potential()
    e0 = lat.energy()
    lat.leapfrog(dt, steps)
    err = rel_l2(lat.psi, want)
    print(f"psi rel L2 after {steps} steps ({prec}): {err:.3e}")
    assert err <= tol
    assert rel_l2(lat.V, Vw) <= (1e-12 if prec == "fp64" else 5e-6)
    e1 = lat.energy()
    assert abs(e1[0] - e0[0]) / e0[0] < (1e-13 if prec == "fp64" else 1e-7)


def test_config1_256cube_fp32_against_fp64(qx):
    """The same workload at its full 256^3 size: the FP32 run against the
    FP64 run (pinned to the oracle above) after 100 and 200 steps, mass and
    energy drift over 200 steps.  FP64 energy changes by the method's own
    3e-9; FP32 adds round-off."""
    N, L, dt, steps = 256, 20.0, 5e-4, 200
    tab = o.blob_table(32, 7, L)
    out, mid, drift = {}, {}, {}
    for prec in ("fp64", "fp32"):
        print(f"[case] 256^3 blobs {prec}", flush=True)
        lat = qx.Lattice(3, N, L, prec, pref=1.0, rho_bar=1.0, k0_mode=0)
        lat.ic_blobs(tab, 0.1)
        lat.update_potential()
        e0 = lat.energy()
        lat.leapfrog(dt, steps // 2, first_half=True, last_half=False)
        mid[prec] = lat.psi
        lat.leapfrog(dt, steps - steps // 2, first_half=False, last_half=True)
        e1 = lat.energy()
        out[prec] = lat.psi
        drift[prec] = (abs(e1[0] - e0[0]) / e0[0], abs((e1[1] + e1[2]) - (e0[1] + e0[2])) / abs(e0[1] + e0[2]))
        lat.close()
    err_mid = rel_l2(mid["fp32"], mid["fp64"])
    err = rel_l2(out["fp32"], out["fp64"])
    print(f"256^3 fp32 vs fp64 after {steps // 2} steps: {err_mid:.3e}, after {steps}: {err:.3e}; drifts {drift}")
    assert err_mid < 1e-5 and err < 1e-5
    assert drift["fp64"][0] < 1e-12 and drift["fp64"][1] < 1e-8
    assert drift["fp32"][0] < 1e-7 and drift["fp32"][1] < 1e-4


def test_headline_512cube_fp32_against_fp64_100_steps(qx):
    """The headline workload at full size (3D 512^3, L = 40, dt = 2.5e-4,
    GRF seed 0): after 100 steps the FP32 state is within 1e-5 relative L2
    of the FP64 run, which matches the FP64 oracle to 1e-12 (the oracle
    comparison at this size is test_headline_512cube_against_oracle); mass
    drift 1e-13 (FP64) and 1e-8 (FP32)."""
    N, L, dt, steps = 512, 40.0, 2.5e-4, 100
    out, drift = {}, {}
    for prec in ("fp64", "fp32"):
        print(f"[case] 512^3 {prec} {steps} steps", flush=True)
        lat = qx.Lattice(3, N, L, prec, pref=1.0, rho_bar=1.0, k0_mode=0)
        lat.ic_grf(0, 0.1)
        lat.update_potential()
        _, m0 = lat.density_stats()
        lat.leapfrog(dt, steps)
        _, m1 = lat.density_stats()
        out[prec] = lat.psi
        drift[prec] = abs(m1 - m0) / m0
        lat.close()
    err = rel_l2(out["fp32"], out["fp64"])
    print(f"512^3 fp32 vs fp64 after {steps} steps: {err:.3e}; mass drift {drift}")
    assert err <= 1e-5
    assert drift["fp64"] < 1e-13 and drift["fp32"] < 1e-8


@pytest.mark.parametrize("N,steps", [(128, 20), (512, 8), (1024, 2)])
def test_headline_lattice_against_oracle(qx, N, steps):
    """The headline workload's lattice (3D, L = 40, dt = 2.5e-4, GRF seed 0,
    k = 0 zeroed) stepped on the GPU and by the FP64 oracle from the same
    state: at 512^3 this runs the paths no smaller lattice reaches (16-column
    tiles, streaming instantiations, the 512-point Poisson pass), at 1024^3
    the 1024-point ones (one-word twiddles, 16-column tiles under 8-column
    strided blocks, 1024-point row kernels).  FP64 within 1e-12 and FP32
    within 1e-5 relative L2 (FP32 measured 1.4e-6 after 50 steps against
    FP64 at 512^3); the 128-bin P(k) of |psi|^2 at the end per bin."""
    L, dt = 40.0, 2.5e-4
    print(f"[case] 3D {N}^3 oracle IC and {steps} steps", flush=True)
    psi0 = o.ic_grf(3, N, L, 0)
    V0, _, _ = o.update_potential(psi0, L, 1.0, 1.0, 0, 0)
    want, Vw = o.leapfrog(psi0, V0, L, dt, steps, 1.0, 1.0, 0, 0)
    nb = min(128, N // 2)
    Po, co, _ = o.spectrum(np.abs(want) ** 2, L, nb, 2)
    errs = {}
    for prec in ("fp64", "fp32"):
        print(f"[case] 3D {N}^3 {prec} {steps} steps", flush=True)
        lat = qx.Lattice(3, N, L, prec, pref=1.0, rho_bar=1.0, k0_mode=0)
        lat.psi = psi0
        lat.update_potential()
        lat.leapfrog(dt, steps)
        P, c, _ = lat.spectrum(source="rho", nbins=nb, spb=2)
        assert np.array_equal(c, co)
        errs[prec] = (rel_l2(lat.psi, want), rel_l2(lat.V, Vw)) + spectrum_errors(P, Po, N, 2)
        lat.close()
    print(f"{N}^3 after {steps} steps (psi, V, P(k) in the cut, outside): {errs}")
    e64, e32 = errs["fp64"], errs["fp32"]
    assert e64[0] <= 1e-12 and e64[1] <= 1e-12 and e64[2] <= 1e-12
    assert e32[0] <= 1e-5 and e32[1] <= 1e-5 and e32[2] <= 1e-5


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_streaming_accesses_bitwise(qx, prec):
    """Streaming (non-temporal) loads and stores change the cache policy only:
    3 fused steps at 512^3 with them and without give the same bits.  They are
    on by default at this size, and absent below N = 512."""
    N, dt = 512, 2.5e-4
    out = []
    for on in (True, False):
        print(f"[case] 512^3 {prec} streaming={on}", flush=True)
        lat = qx.Lattice(3, N, 40.0, prec, pref=1.0, rho_bar=1.0, k0_mode=0)
        assert lat.set_stream_io(True) is True  # the default here
        assert lat.set_stream_io(on) is on
        lat.ic_grf(0, 0.1)
        lat.update_potential()
        lat.leapfrog(dt, 3)
        out.append((lat.psi, lat.V))
        lat.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    small = qx.Lattice(3, 64, 20.0, prec)
    assert small.set_stream_io(True) is False
    small.close()


def test_1024cube_fp32_against_fp64(qx):
    """3D 1024^3, the largest lattice: FP32 (16-column tiles, two-wavefront
    lines) against FP64 (4-column tiles) from the same GRF IC, 4 steps."""
    N, L, dt, steps = 1024, 40.0, 2.5e-4, 4
    out, drift = {}, {}
    for prec in ("fp64", "fp32"):
        print(f"[case] 1024^3 {prec} {steps} steps", flush=True)
        lat = qx.Lattice(3, N, L, prec, pref=1.0, rho_bar=1.0, k0_mode=0)
        lat.ic_grf(0, 0.1)
        lat.update_potential()
        _, m0 = lat.density_stats()
        lat.leapfrog(dt, steps)
        _, m1 = lat.density_stats()
        out[prec] = lat.psi
        drift[prec] = abs(m1 - m0) / m0
        lat.close()
    err = rel_l2(out["fp32"], out["fp64"])
    del out
    print(f"1024^3 fp32 vs fp64 after {steps} steps: {err:.3e}; mass drift {drift}")
    assert err <= 2e-6
    assert drift["fp64"] < 1e-13 and drift["fp32"] < 1e-6


def test_dithered_long_lines_fp32_against_fp64(qx):
    """FP32 lines of 512 points and more take dithered one-word twiddle
    tables in the drift (one table per drift, picked by the lattice's drift
    counter): a rounded table's error would come back every step and drift
    the mass and phase linearly; over the dithered tables it stays bounded.
    2D 2048^2 GRF, 400 steps, FP32 against FP64 from the same IC."""
    N, L, dt, steps = 2048, 40.0, 2.5e-4, 400
    out, drift = {}, {}
    for prec in ("fp64", "fp32"):
        print(f"[case] 2D {N}^2 {prec} {steps} steps", flush=True)
        lat = qx.Lattice(2, N, L, prec, pref=1.0, rho_bar=1.0, k0_mode=0)
        lat.ic_grf(4, 0.1)
        lat.update_potential()
        _, m0 = lat.density_stats()
        lat.leapfrog(dt, steps // 2)
        lat.leapfrog(dt, steps // 2)
        _, m1 = lat.density_stats()
        out[prec] = lat.psi
        drift[prec] = abs(m1 - m0) / m0
        lat.close()
    err = rel_l2(out["fp32"], out["fp64"])
    print(f"2D {N}^2 fp32 vs fp64 after {steps} steps: {err:.3e}; mass drift {drift}")
    assert err <= 1e-5
    assert drift["fp64"] < 1e-13 and drift["fp32"] < 1e-7


# ------------------------------------------------------------------ initial conditions
@pytest.mark.parametrize("dim,N", [(2, 64), (3, 16)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_device_ics(qx, dim, N, prec):
    L = 10.0
    lat = lattice(qx, dim, N, L, prec)
    lat.ic_grf(42, 0.1)
    assert rel_l2(lat.psi, o.ic_grf(dim, N, L, 42, 0.1)) < (1e-12 if prec == "fp64" else 5e-6)
    tab = o.blob_table(8, 3, L)
    lat.ic_blobs(tab, 0.1)
    assert rel_l2(lat.psi, o.ic_blobs(dim, N, L, tab, 0.1)) < (1e-14 if prec == "fp64" else 1e-6)
    sol = rm.read_solitons()
    lat.ic_solitons(sol, 1.0)
    assert rel_l2(lat.psi, o.ic_solitons(dim, N, L, sol, 1.0)) < (1e-14 if prec == "fp64" else 1e-6)


# ------------------------------------------------------------------ larger lattices
@pytest.mark.parametrize("dim,N", [(3, 256), (2, 2048), (3, 1024)])
def test_large_lattice_invariants(qx, dim, N):
    """Size-independent properties at large N: round trip of the transform,
    mass conservation through fused steps, determinism."""
    L = 40.0
    lat = lattice(qx, dim, N, L, "fp32", pref=1.0)
    lat.ic_grf(1, 0.1)
    r0, m0 = lat.density_stats()
    lat.update_potential()
    lat.leapfrog(2.5e-4, 5)
    r1, m1 = lat.density_stats()
    assert abs(m1 - m0) / m0 < 1e-5
    e0 = lat.energy()
    lat.leapfrog(2.5e-4, 5)
    e1 = lat.energy()
    assert abs((e1[1] + e1[2]) - (e0[1] + e0[2])) / abs(e0[1] + e0[2]) < 1e-3


# ------------------------------------------------------------------ ensembles
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("graph_steps", [0, 3])
def test_ensemble_members_match_solo_runs(qx, prec, graph_steps):
    """Each member of a batched ensemble of uneven sizes evolves exactly as
    the same lattice run alone (same kernels, same constants: bitwise)."""
    Ns = [16, 64, 16, 32, 64, 16]
    Ls = [5.0, 12.0, 40.0, 7.5, 20.0, 9.0]
    seeds = [11, 12, 13, 14, 15, 16]
    dx0 = 10.0 / 256
    ds = [1e-3 * (L / N / dx0) ** 2 for N, L in zip(Ns, Ls)]
    print(f"[case] ensemble prec={prec} graph_steps={graph_steps}", flush=True)
    ens = qx.Ensemble(Ns, Ls, precision=prec, pref=1.0, graph_steps=graph_steps)
    ens.ic_grf(seeds)
    ens.update_potential()
    ens.leapfrog(ds, 9)
    m_after = ens.mass()
    for i, (N, L) in enumerate(zip(Ns, Ls)):
        solo = qx.Lattice(2, N, L, precision=prec, pref=1.0)
        solo.ic_grf(seeds[i])
        solo.update_potential()
        solo.leapfrog(ds[i], 9)
        got = ens.psi(i)
        want = solo.psi
        assert np.array_equal(got, want), (i, rel_l2(got, want))
        assert m_after[i] == pytest.approx(solo.density_stats()[1], rel=1e-12)
    # and against the FP64 oracle for one member
    psi0 = o.ic_grf(2, Ns[1], Ls[1], seeds[1])
    V0, _, _ = o.update_potential(psi0, Ls[1], 1.0)
    want, _ = o.leapfrog(psi0, V0, Ls[1], ds[1], 9, 1.0)
    assert rel_l2(ens.psi(1), want) < (1e-12 if prec == "fp64" else 5e-6)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_ensemble_config_size_classes(qx, prec):
    """The batched ensemble's four size classes (128^2, 512^2, 1024^2,
    2048^2, two members each, own seeds and boxes L in [5, 40], dt scaled to
    dx^2), graph-captured steps with the tuned tile widths: every member
    bitwise equal to its solo run, and the 512^2 member against the FP64
    oracle."""
    Ns = [128, 512, 1024, 2048, 128, 512, 1024, 2048]
    Ls = [5.0, 17.5, 40.0, 23.0, 31.0, 8.0, 12.5, 36.0]
    seeds = [40, 41, 42, 43, 44, 45, 46, 47]
    dx0 = 10.0 / 256
    ds = [1e-3 * (L / N / dx0) ** 2 for N, L in zip(Ns, Ls)]
    steps = 20
    print(f"[case] ensemble size classes prec={prec}", flush=True)
    ens = qx.Ensemble(Ns, Ls, precision=prec, pref=1.0, graph_steps=8)
    ens.ic_grf(seeds)
    ens.update_potential()
    ens.tune(ds, reps=2)
    ens.leapfrog(ds, steps)
    for i, (N, L) in enumerate(zip(Ns, Ls)):
        solo = qx.Lattice(2, N, L, precision=prec, pref=1.0)
        solo.ic_grf(seeds[i])
        solo.update_potential()
        solo.leapfrog(ds[i], steps)
        got, want = ens.psi(i), solo.psi
        assert np.array_equal(got, want), (i, N, rel_l2(got, want))
        solo.close()
    i = 1
    psi0 = o.ic_grf(2, Ns[i], Ls[i], seeds[i])
    V0, _, _ = o.update_potential(psi0, Ls[i], 1.0)
    want, _ = o.leapfrog(psi0, V0, Ls[i], ds[i], steps, 1.0)
    err = rel_l2(ens.psi(i), want)
    print(f"512^2 member after {steps} steps against the oracle ({prec}): {err:.2e}")
    assert err < (1e-12 if prec == "fp64" else 5e-6)


def test_ensemble_96_members(qx):
    """The batched ensemble at its full make-up: 32 x 128^2, 32 x 512^2,
    24 x 1024^2 and 8 x 2048^2, seeds 0-95, L ~ U[5, 40], dt scaled to dx^2,
    tuned, 16 graph-captured steps (two replays of 8).  Every member keeps its
    mass (FP32), the first and last member of each size class equal their solo
    runs bit for bit, and the first 128^2 member matches the FP64 oracle."""
    from qaxions_amd.rng import uniform
    Ns = [128] * 32 + [512] * 32 + [1024] * 24 + [2048] * 8
    Ls = [5.0 + 35.0 * uniform(2024, 99, i) for i in range(len(Ns))]
    dx0 = 10.0 / 256
    ds = [1e-3 * ((L / N) / dx0) ** 2 for N, L in zip(Ns, Ls)]
    steps = 16
    print("[case] ensemble 96 members fp32", flush=True)
    ens = qx.Ensemble(Ns, Ls, precision="fp32", pref=1.0, graph_steps=8)
    ens.ic_grf(list(range(len(Ns))))
    ens.update_potential()
    ens.tune(ds, reps=2)
    m0 = ens.mass()
    ens.leapfrog(ds, steps)
    drift = np.abs(ens.mass() / m0 - 1.0)
    print(f"96 members after {steps} steps: worst relative mass drift {drift.max():.2e}")
    assert drift.max() < 1e-5
    for i in (0, 31, 32, 63, 64, 87, 88, 95):
        solo = qx.Lattice(2, Ns[i], Ls[i], precision="fp32", pref=1.0)
        solo.ic_grf(i)
        solo.update_potential()
        solo.leapfrog(ds[i], steps)
        assert np.array_equal(ens.psi(i), solo.psi), (i, Ns[i])
        solo.close()
    psi0 = o.ic_grf(2, Ns[0], Ls[0], 0)
    V0, _, _ = o.update_potential(psi0, Ls[0], 1.0)
    want, _ = o.leapfrog(psi0, V0, Ls[0], ds[0], steps, 1.0)
    err = rel_l2(ens.psi(0), want)
    print(f"member 0 (128^2) against the oracle: {err:.2e}")
    assert err < 5e-6
    ens.close()


def test_ensemble_tune_cache(qx, tmp_path):
    """The ensemble tuner's cache: the first tune stores {N: variant} under
    the class make-up, a second ensemble of that make-up reinstalls the
    stored choice without timing (a forged entry comes back as stored), and
    stepping after either equals an untuned ensemble bit for bit."""
    Ns, Ls = [64, 128, 64, 128, 64], [10.0, 20.0, 12.0, 25.0, 8.0]
    ds = [1e-3] * len(Ns)
    cache = str(tmp_path / "ens_tiles.json")
    print("[case] ensemble tune cache", flush=True)

    def run(tune):
        e = qx.Ensemble(Ns, Ls, precision="fp32", pref=1.0, graph_steps=4)
        e.ic_grf(list(range(7, 7 + len(Ns))))
        e.update_potential()
        got = tune(e) if tune else None
        e.leapfrog(ds, 12)
        return got, [e.psi(i) for i in range(len(Ns))]

    _, plain = run(None)
    first, a = run(lambda e: e.tune(ds, reps=2, cache=cache))
    assert set(first) == {64, 128} and set(first.values()) <= {0, 1}
    with open(cache) as f:
        db = json.load(f)
    key = "ens-fp32-64x3,128x2"
    assert db[key] == {str(k): v for k, v in first.items()}
    forged = {64: 1 - first[64], 128: 1 - first[128]}
    db[key] = {str(k): v for k, v in forged.items()}
    with open(cache, "w") as f:
        json.dump(db, f)
    second, b = run(lambda e: e.tune(ds, reps=2, cache=cache))
    assert second == forged
    for x, y, z in zip(plain, a, b):
        assert np.array_equal(x, y) and np.array_equal(x, z)
    with pytest.raises(Exception):
        e = qx.Ensemble(Ns, Ls, precision="fp32")
        e.set_tile_variants({64: 2, 128: 0})


def test_pass_timing_and_tune_keep_state(qx):
    """Per-pass timing (inside consecutive steps on scratch copies) and the
    tile-variant tuner leave the state untouched, time every pass, and the
    chosen variants do not change results (at 2D 1024^2 also through the
    drift counter that picks the dithered tables, which timing restores)."""
    for dim, N in ((3, 32), (2, 64), (2, 1024)):
        print(f"[case] pass timing dim={dim} N={N}", flush=True)
        a = qx.Lattice(dim, N, 10.0, "fp32", pref=1.0)
        b = qx.Lattice(dim, N, 10.0, "fp32", pref=1.0)
        for lat in (a, b):
            lat.ic_grf(5, 0.1)
            lat.update_potential()
        psi0, V0 = a.psi, a.V
        names, ms, by = a.time_step_passes(1e-3, reps=3)
        assert len(names) == (7 if dim == 3 else 4) and len(set(names)) == len(names)
        assert np.all(ms > 0) and np.all(by > 0)
        v = a.tune(1e-3, reps=2)
        assert len(v) == 8 and set(v) <= {0, 1}
        assert np.array_equal(a.psi, psi0) and np.array_equal(a.V, V0)
        a.leapfrog(1e-3, 4)
        b.leapfrog(1e-3, 4)
        assert np.array_equal(a.psi, b.psi)
        a.close()
        b.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_ensemble_tune_keeps_state_and_results(qx, prec):
    """The ensemble autotuner reports one tile variant per size class, leaves
    every member's state as it was, and the tile variant does not change
    results (each column's transform is the same whatever the tile width)."""
    Ns = [16, 64, 16, 32]
    Ls = [5.0, 12.0, 40.0, 7.5]
    ds = [1e-3 * (L / N / (10.0 / 256)) ** 2 for N, L in zip(Ns, Ls)]
    a = qx.Ensemble(Ns, Ls, precision=prec, pref=1.0, graph_steps=0)
    b = qx.Ensemble(Ns, Ls, precision=prec, pref=1.0, graph_steps=0)
    for e in (a, b):
        e.ic_grf([21, 22, 23, 24])
        e.update_potential()
    before = [a.psi(i) for i in range(len(Ns))]
    v = a.tune(ds, reps=2)
    assert sorted(v) == sorted(set(Ns)) and set(v.values()) <= {0, 1}
    for i in range(len(Ns)):
        assert np.array_equal(a.psi(i), before[i])
    a.leapfrog(ds, 5)
    b.leapfrog(ds, 5)
    for i in range(len(Ns)):
        assert np.array_equal(a.psi(i), b.psi(i))
