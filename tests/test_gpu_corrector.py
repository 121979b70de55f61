"""The corrector's HIP forward and backward kernels against the FP64 oracle
(oracle/corrector.py, with the device's BF16 rounding points), and what the
trained network gains over coarse stepping without it."""
import numpy as np
import pytest

from oracle import corrector as cn

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("H,W", [(32, 32), (16, 48), (12, 12), (8, 20)])
def test_forward_matches_oracle(qx, H, W):
    """W % 16 == 0 runs the MFMA path for the 32 -> 32 layers, else the vector path."""
    print(f"[case] corrector forward {H}x{W}", flush=True)
    rng = np.random.default_rng(H * 100 + W)
    theta = cn.init_params(5, last_scale=0.5)
    x = rng.standard_normal((H, W, 3))
    net = qx.Corrector(H, W, params=theta.astype(np.float32))
    y = net.forward(x.astype(np.float32)).astype(np.float64)
    theta32 = theta.astype(np.float32).astype(np.float64)
    x32 = x.astype(np.float32).astype(np.float64)
    want = x32[..., :2] + cn.forward(theta32, x32, bf16=True)
    # compare the network part (y - psi), where rounding points decide agreement
    assert rel(y - x32[..., :2], want - x32[..., :2]) < 1e-2


@pytest.mark.parametrize("H,W", [(16, 16), (12, 20)])
def test_gradient_matches_oracle(qx, H, W):
    print(f"[case] corrector gradient {H}x{W}", flush=True)
    rng = np.random.default_rng(7)
    theta = cn.init_params(9, last_scale=0.5).astype(np.float32)
    x = rng.standard_normal((H, W, 3)).astype(np.float32)
    t = (x[..., :2] + 0.1 * rng.standard_normal((H, W, 2))).astype(np.float32)
    net = qx.Corrector(H, W, params=theta)
    net.output_scale = 0.25
    loss, g = net.loss_grad(x, t)
    lo, go = cn.loss_and_grad(theta.astype(np.float64), x.astype(np.float64), t.astype(np.float64), bf16=True,
                              scale=0.25)
    assert loss == pytest.approx(lo, rel=1e-3)
    assert rel(g, go) < 2e-2
    # every layer's weights and biases individually
    o = 0
    for (ws, bs) in cn.param_shapes():
        nw, nb = int(np.prod(ws)), int(np.prod(bs))
        assert rel(g[o:o + nw], go[o:o + nw]) < 5e-2
        assert rel(g[o + nw:o + nw + nb], go[o + nw:o + nw + nb]) < 5e-2
        o += nw + nb


def test_adam_step_lowers_loss(qx):
    rng = np.random.default_rng(3)
    H = W = 16
    x = rng.standard_normal((H, W, 3)).astype(np.float32)
    t = (x[..., :2] + 0.05 * np.roll(x[..., :2], 1, axis=0) + 0.05 * x[..., 2:]).astype(np.float32)
    net = qx.Corrector(H, W, params=cn.init_params(1).astype(np.float32))
    l0, _ = net.loss_grad(x, t)
    for _ in range(30):
        net.loss_grad(x, t)
        qx.lib().qx_net_adam_step(net._h, 1e-3, 0.9, 0.999, 1e-8)
    l1, _ = net.loss_grad(x, t)
    assert l1 < 0.5 * l0


@pytest.mark.parametrize("N", [64, 1024])
def test_forced_fallback_equals_fine_steps(qx, N):
    """mass_tol < 0 forces every chunk back to fine stepping: the result is
    exactly k*n fine steps (the failure path, forced from the host).  At
    N = 1024 the drift's dithered tables follow the lattice's drift counter,
    which a redone chunk restores with the state."""
    from qaxions_amd.learned import init_params
    L, dt, k = 10.0, 1e-3, 4
    print(f"[case] forced fallback N={N}", flush=True)
    net = qx.Corrector(N, N, params=init_params())
    a = qx.Lattice(2, N, L, "fp32", pref=1.0)
    a.ic_grf(5)
    a.update_potential()
    b = qx.Lattice(2, N, L, "fp32", pref=1.0)
    b.copy_from(a)
    fb = net.coarse_steps(a, dt, k, 6, check_every=3, mass_tol=-1.0)
    assert fb == 2
    b.leapfrog(dt, k * 3)
    b.leapfrog(dt, k * 3)
    assert np.array_equal(a.psi, b.psi)


def test_corrector_fits_a_solver_pair(qx):
    """Trained on one solver pair ((psi, V) after a coarse step of 4 dt -> psi
    after 4 fine steps) the corrector removes most of the coarse step's
    error: lattice input, scales, forward, gradient and Adam work end to end.
    pref = 100 puts the splitting error well above FP32 round-off."""
    from qaxions_amd.learned import init_params
    N, L, dt, k, pref = 64, 10.0, 2e-3, 4, 100.0
    traj = qx.Lattice(2, N, L, "fp32", pref=pref)
    traj.ic_grf(1000)
    traj.update_potential()
    coarse = qx.Lattice(2, N, L, "fp32", pref=pref)
    coarse.copy_from(traj)
    coarse.leapfrog(k * dt, 1)
    fine = qx.Lattice(2, N, L, "fp32", pref=pref)
    fine.copy_from(traj)
    fine.leapfrog(dt, k)
    d = fine.psi - coarse.psi
    net = qx.Corrector(N, N, params=init_params())
    net.output_scale = float(np.sqrt(np.mean(np.abs(d) ** 2) / 2.0))
    net.input_scale = float(1.0 / np.std(coarse.V))
    base = net.pair_mse(coarse, fine)
    assert base == pytest.approx(np.mean(np.abs(d) ** 2) / 2.0, rel=1e-4)
    rel = [net.train_step_lattices(coarse, fine, 1e-2) / base for _ in range(1200)]
    print(f"relative loss: first {rel[0]:.3f}, last {np.mean(rel[-100:]):.4f}")
    assert rel[0] == pytest.approx(1.0, abs=0.05)   # starts near the identity
    assert np.mean(rel[-100:]) < 0.1


def test_zero_network_is_plain_coarse_stepping(qx):
    """With the last layer zero the corrector is the identity, so corrected
    coarse stepping is coarse stepping with a potential update after each
    step: equal to plain coarse steps to FP32 round-off."""
    from qaxions_amd.learned import init_params
    N, L, dt, k, n, pref = 64, 10.0, 2e-3, 4, 6, 100.0
    net = qx.Corrector(N, N, params=init_params(last_scale=0.0))
    a = qx.Lattice(2, N, L, "fp32", pref=pref)
    a.ic_grf(7)
    a.update_potential()
    b = qx.Lattice(2, N, L, "fp32", pref=pref)
    b.copy_from(a)
    assert net.coarse_steps(a, dt, k, n, check_every=3, mass_tol=1e-3) == 0
    for _ in range(n):
        b.leapfrog(k * dt, 1)
    assert rel(a.psi, b.psi) < 1e-5


@pytest.mark.xfail(strict=False, reason=(
    "measured at 512^2, pref = 100, dt = 1e-3, k = 4, 1000 training pairs: after 10 coarse steps "
    "against the dt/4 reference, corrected 1.0e-3, plain coarse 8.8e-4, plain fine 7.8e-5 (gain "
    "over coarse 0.87); 10 corrected coarse steps take 2.39 ms against 1.50 ms for the 40 fine "
    "steps they replace (speed-up at equal error 0.16)"))
def test_trained_corrector_beats_plain_coarse(qx):
    """The A/B of learned coarse stepping at the config's 512^2: a corrector
    trained on solver pairs (seeds disjoint from the evaluation) against
    plain coarse and plain fine stepping, errors against a dt/4 reference
    inside the predictable horizon (10 coarse steps), with times.  Passes if
    the network lowers the coarse error and the corrected steps beat plain
    stepping by 2x at equal error."""
    from qaxions_amd import learned
    N, L, dt, k, pref = 512, 10.0, 1e-3, 4, 100.0
    print("[case] corrector A/B 512^2", flush=True)
    net, _, relative = learned.train(N, L, dt, k, seeds=range(1000, 1004), steps_per_seed=250, lr=1e-3,
                                     lr_final=1e-4, pref=pref)
    res = learned.evaluate(net, N, L, dt, k, seed=0, n_coarse=10, check_every=10, mass_tol=1e-6,
                           pref=pref, timing_reps=2, warm_steps=200)
    e = {m: res[m]["rel_l2_vs_reference"] for m in ("fine", "coarse", "learned")}
    t = {m: res[m]["seconds"] for m in ("fine", "coarse", "learned")}
    print(f"A/B 512^2: training relative loss {relative[:100].mean():.3f} -> {relative[-100:].mean():.3f}; "
          f"errors vs dt/4 reference {e}; seconds {t}; gain over coarse {res['error_gain_learned_vs_coarse']:.2f}; "
          f"speed-up at equal error {res['speedup_learned_vs_plain_at_equal_error']:.2f}")
    assert res["learned"]["fallback_chunks"] == 0   # the scaled correction keeps the mass
    assert e["learned"] < e["coarse"]
    assert res["speedup_learned_vs_plain_at_equal_error"] >= 2.0


def test_correction_keeps_the_mass(qx):
    """The corrected state is scaled back to its mass before the correction:
    a network far from the identity changes psi, not sum |psi|^2."""
    from qaxions_amd.learned import init_params
    N, L = 64, 10.0
    net = qx.Corrector(N, N, params=init_params(last_scale=0.5))
    net.output_scale = 0.05
    lat = qx.Lattice(2, N, L, "fp32", pref=1.0)
    lat.ic_grf(3)
    lat.update_potential()
    p0 = lat.psi
    net.correct(lat)
    p1 = lat.psi
    m0, m1 = np.sum(np.abs(p0) ** 2), np.sum(np.abs(p1) ** 2)
    assert rel(p1, p0) > 1e-3
    assert abs(m1 - m0) / m0 < 1e-6


def test_energy_gate_falls_back(qx):
    """With an energy tolerance below what a far-from-identity correction
    changes, every chunk falls back to fine steps (forced from the host)."""
    from qaxions_amd.learned import init_params
    N, L, dt, k = 64, 10.0, 1e-3, 4
    net = qx.Corrector(N, N, params=init_params(last_scale=0.5))
    net.output_scale = 0.05
    net.set_energy_tol(1e-9)
    a = qx.Lattice(2, N, L, "fp32", pref=1.0)
    a.ic_grf(5)
    a.update_potential()
    b = qx.Lattice(2, N, L, "fp32", pref=1.0)
    b.copy_from(a)
    assert net.coarse_steps(a, dt, k, 6, check_every=3, mass_tol=1.0) == 2
    b.leapfrog(dt, k * 3)
    b.leapfrog(dt, k * 3)
    assert np.array_equal(a.psi, b.psi)


def test_zero_input_scale_is_a_psi_only_network(qx):
    """Input scale 0 feeds the V channel zeros: the corrected state no longer
    depends on V (the config's 2-channel network), while with a positive
    scale it does."""
    from qaxions_amd.learned import init_params
    N, L = 64, 10.0
    out = {}
    for a in (0.0, 1.0):
        for vs in (1.0, 3.0):
            net = qx.Corrector(N, N, params=init_params(last_scale=0.5))
            net.output_scale = 0.05
            net.input_scale = a
            lat = qx.Lattice(2, N, L, "fp32", pref=1.0)
            lat.ic_grf(7)
            lat.update_potential()
            lat.V = lat.V * vs
            net.correct(lat)
            out[(a, vs)] = lat.psi
            lat.close()
    assert np.array_equal(out[(0.0, 1.0)], out[(0.0, 3.0)])
    assert rel(out[(1.0, 1.0)], out[(1.0, 3.0)]) > 1e-4
