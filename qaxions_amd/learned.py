"""Learned coarse stepping: training pairs from the solver, GPU training of the
corrector and its evaluation against plain fine and coarse stepping.

A pair is ((psi, V) after one coarse KDK step of k*dt, psi after k fine KDK
steps of dt) from the same state; states are taken along fine trajectories
that start from Gaussian-random-field ICs.  Everything runs on the device;
only scalars come back to the host.
"""
from __future__ import annotations

import time

import numpy as np

from . import Corrector, Lattice


def init_params(seed=0, last_scale=1e-3):
    """He-style initial parameters; the last layer starts near zero, so the
    corrector starts close to the identity on psi."""
    rng = np.random.default_rng(seed)
    chans = (3, 32, 32, 32, 2)
    out = []
    for i, (ci, co) in enumerate(zip(chans[:-1], chans[1:])):
        w = rng.standard_normal((co, ci, 3, 3)) * np.sqrt(2.0 / (ci * 9))
        if i == 3:
            w *= last_scale
        out += [w.ravel(), np.zeros(co)]
    return np.concatenate(out).astype(np.float32)


def train(N, L, dt, k, seeds, steps_per_seed, lr=1e-3, lr_final=None, net=None, log_every=0,
          pref=1.0, potential=True):
    """Train a corrector on len(seeds) * steps_per_seed pairs.  Returns
    (net, losses, relative) where relative[i] = loss / mean((coarse - fine)^2)
    of pair i: the fraction of the coarse step's error the network leaves.
    potential=False feeds the V channel zeros (input scale 0): the network
    then sees psi alone, as a 2-channel network would."""
    if net is None:
        net = Corrector(N, N, params=init_params())
    traj = Lattice(2, N, L, "fp32", pref=pref)
    coarse = Lattice(2, N, L, "fp32", pref=pref)
    fine = Lattice(2, N, L, "fp32", pref=pref)
    total = len(seeds) * steps_per_seed
    losses, rel = [], []
    it = 0
    scale_set = net.output_scale != 1.0
    for seed in seeds:
        traj.ic_grf(int(seed))
        traj.update_potential()
        for _ in range(steps_per_seed):
            coarse.copy_from(traj)
            coarse.leapfrog(k * dt, 1)
            fine.copy_from(traj)
            fine.leapfrog(dt, k)
            if not scale_set:
                # the network predicts the correction in units of its typical
                # size and sees V in units of its spread
                d = fine.psi - coarse.psi
                net.output_scale = float(np.sqrt(np.mean(np.abs(d) ** 2) / 2.0))
                v = coarse.V
                net.input_scale = float(1.0 / max(np.std(v), 1e-30)) if potential else 0.0
                scale_set = True
            lr_it = lr if lr_final is None else lr * (lr_final / lr) ** (it / max(total - 1, 1))
            base = net.pair_mse(coarse, fine)
            losses.append(net.train_step_lattices(coarse, fine, lr_it))
            rel.append(losses[-1] / base if base > 0 else 1.0)
            traj.copy_from(fine)
            it += 1
            if log_every and it % log_every == 0:
                print(f"[train] {it}/{total} loss {np.mean(losses[-log_every:]):.3e} "
                      f"relative {np.mean(rel[-log_every:]):.3f}", flush=True)
    for x in (traj, coarse, fine):
        x.close()
    return net, np.array(losses), np.array(rel)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def evaluate(net, N, L, dt, k, seed, n_coarse, check_every=10, mass_tol=1e-6, ref_refine=4,
             pref=1.0, timing_reps=1, warm_steps=0):
    """Errors after n_coarse coarse intervals (time n_coarse*k*dt) against a
    reference run with dt/ref_refine, for plain fine stepping (dt), plain
    coarse stepping (k dt) and corrected coarse stepping; plus device wall
    times of the fine and corrected runs.  The runs start after warm_steps
    fine steps from the IC (a GRF IC is real at t = 0, unlike the states the
    pairs are drawn from)."""
    lat0 = Lattice(2, N, L, "fp32", pref=pref)
    lat0.ic_grf(int(seed))
    lat0.update_potential()
    if warm_steps:
        lat0.leapfrog(dt, int(warm_steps))
    runs = {}

    def run(name, fn):
        lat = Lattice(2, N, L, "fp32", pref=pref)
        best = None
        for _ in range(timing_reps):
            lat.copy_from(lat0)
            lat.synchronize()
            t0 = time.perf_counter()
            extra = fn(lat)
            lat.synchronize()
            t = time.perf_counter() - t0
            best = t if best is None else min(best, t)
        runs[name] = (lat.psi, best, extra)
        lat.close()

    run("reference", lambda lat: lat.leapfrog(dt / ref_refine, n_coarse * k * ref_refine))
    run("fine", lambda lat: lat.leapfrog(dt, n_coarse * k))
    run("coarse", lambda lat: lat.leapfrog(k * dt, n_coarse))
    run("learned", lambda lat: net.coarse_steps(lat, dt, k, n_coarse, check_every, mass_tol))
    ref = runs["reference"][0]
    out = {name: {"rel_l2_vs_reference": _rel(p, ref), "seconds": t}
           for name, (p, t, _) in runs.items() if name != "reference"}
    out["learned"]["fallback_chunks"] = runs["learned"][2]
    out["learned"]["rel_l2_vs_fine"] = _rel(runs["learned"][0], runs["fine"][0])
    out["coarse"]["rel_l2_vs_fine"] = _rel(runs["coarse"][0], runs["fine"][0])
    out["speedup_learned_vs_fine"] = runs["fine"][1] / runs["learned"][1]
    out.update(equal_error(out, dt, k))
    out["horizon"] = {"n_coarse": n_coarse, "time": n_coarse * k * dt}
    lat0.close()
    return out


def equal_error(res, dt, k):
    """Speed-up of corrected coarse stepping over plain stepping at the same
    error.  Plain KDK stepping's error against the reference grows as h^2:
    e(h) = C h^2, with C fitted to the plain fine (h = dt) and coarse
    (h = k dt) errors.  Plain stepping reaches the corrected run's error at
    h_eq = sqrt(e_learned / C), at a cost of (fine run's time) * dt / h_eq;
    the speed-up is that cost over the corrected run's time.  The gain the
    network itself brings is e_coarse / e_learned (1: none)."""
    ef, ec = res["fine"]["rel_l2_vs_reference"], res["coarse"]["rel_l2_vs_reference"]
    el = res["learned"]["rel_l2_vs_reference"]
    C = (ef / dt ** 2 + ec / (k * dt) ** 2) / 2.0
    h_eq = float(np.sqrt(el / C)) if C > 0 else float("nan")
    t_plain = res["fine"]["seconds"] * dt / h_eq
    return {"error_gain_learned_vs_coarse": ec / el if el > 0 else float("inf"),
            "order_check_coarse_over_fine": ec / ef if ef > 0 else float("nan"),
            "h_equal_error": h_eq,
            "speedup_learned_vs_plain_at_equal_error": t_plain / res["learned"]["seconds"]}
