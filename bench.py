#!/usr/bin/env python3
"""Throughput of the fused kick-drift-kick Schroedinger-Poisson step on one MI355X.

Workload (default): 3D lattice 512^3, L = 40, complex64 state with FP64
phase reduction and sums, dt = 2.5e-4, seeded Gaussian-random-field IC
(P(k) ~ k^-2 cut at k_Ny/2), code units pref = 1, rho_bar = 1, k = 0 mode
zeroed.  A step is one merged-kick KDK step of the reference time loop
(src/main.cxx:65-85): K, D (forward FFT, e^{-i dt k^2/2}, inverse FFT),
P (|psi|^2, r2c, -1/k^2, c2r, max|V|).

Prints ONE JSON line: site-updates/s (value), ms per step, the roofline of
the dominant pass and of the whole step, and the FP64 CPU oracle timed on
the host cores on a bounded sample of the same workload.
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X


def algorithmic_bytes_per_site(N, c=8, r=4):
    """Each operator of one merged-kick step reads its inputs and writes its
    outputs once: B = (9c + 5r) S + 4c H, H = S (N/2+1)/N."""
    return (9 * c + 5 * r) + 4 * c * (N // 2 + 1) / N


class HipEvents:
    """hipEvent timing on a given stream (the library's stream)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        for n in ("hipEventCreate", "hipEventRecord", "hipEventSynchronize",
                  "hipEventElapsedTime", "hipEventDestroy"):
            getattr(self.h, n).restype = ctypes.c_int

    def event(self):
        e = ctypes.c_void_p()
        assert self.h.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self, e0, e1):
        assert self.h.hipEventSynchronize(e1) == 0
        t = ctypes.c_float()
        assert self.h.hipEventElapsedTime(ctypes.byref(t), e0, e1) == 0
        return t.value


def measured_traffic(pass_name, path=os.path.join(ROOT, "profiles", "kdk512_traffic.json")):
    """HBM bytes per launch of a pass of the 512^3 FP32 step, from the
    committed FETCH_SIZE / WRITE_SIZE counter runs (scaled by the copy
    calibration recorded in the same file); None when not measured."""
    try:
        with open(path) as f:
            rec = json.load(f)["passes"][pass_name]
        return rec["fetch_bytes"] + rec["write_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def host_cpu():
    """CPU model and physical core count of the host (lscpu), and the
    threads the CPU leg may use: the physical cores, capped by the CPU share
    this process is given (OMP_NUM_THREADS where set, else its affinity)."""
    info = {"cpu_model": None, "physical_cores": None}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        info["cpu_model"] = kv.get("Model name")
        cps, sockets = int(kv.get("Core(s) per socket", "0")), int(kv.get("Socket(s)", "0"))
        if cps and sockets:
            info["physical_cores"] = cps * sockets
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    share = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        share = min(share, int(os.environ["OMP_NUM_THREADS"]))
    info["cpu_share"] = share
    info["threads"] = min(share, info["physical_cores"] or share)
    return info


def spectrum_errors(P, Po, N, spb):
    """max_b |P_b - Po_b| / Po_b over the bins whose shells lie inside the
    GRF IC's spectral cut (|n| <= N/4)"""
    cut = (np.arange(len(P)) + 1) * spb - 1 <= N // 4
    return float((np.abs(P - Po)[cut] / Po[cut]).max())


def oracle_leg(qx, N, L, dt, warm, timed, threads):
    """The FP64 CPU oracle (oracle/qx_oracle.c, OpenMP, threads bound to
    cores) on the headline lattice from the oracle's own GRF IC: `warm`
    merged-kick steps, then `timed` steps each timed alone (their median is
    the baseline).  The same steps from the same IC on the GPU, in FP32 and
    FP64, are then compared with it: psi relative L2 and the per-bin P(k)."""
    import oracle
    oracle.set_threads(threads)
    dim = 3
    psi0 = oracle.ic_grf(dim, N, L, 0)
    V, Vh, _ = oracle.update_potential(psi0, L, 1.0, 1.0, 0, 0)
    psi = psi0.copy()
    times = []
    for i in range(warm + timed):
        t0 = time.perf_counter()
        oracle.leapfrog_inplace(psi, V, Vh, L, dt, 1, 1.0, 1.0, 0, 0, first_half=(i == 0), last_half=0)
        if i >= warm:
            times.append(time.perf_counter() - t0)
    steps = warm + timed
    nb = max(1, (N // 2) // 2)
    Po, _, _ = oracle.spectrum(np.abs(psi) ** 2, L, nb, 2)
    acc = {"steps": steps, "pk_bins": nb}
    for prec in ("fp32", "fp64"):
        lat = qx.Lattice(dim, N, L, precision=prec, pref=1.0, rho_bar=1.0, k0_mode=0)
        lat.psi = psi0
        lat.update_potential()
        lat.leapfrog(dt, steps, first_half=True, last_half=False)
        got = lat.psi
        acc[f"{prec}_rel_l2"] = float(np.linalg.norm(got - psi) / np.linalg.norm(psi))
        P, _, _ = lat.spectrum(source="rho", nbins=nb, spb=2)
        acc[f"{prec}_pk_max_rel_err_in_cut"] = spectrum_errors(P, Po, N, 2)
        del got
        lat.close()
    med = float(np.median(times))
    return N ** dim / med, med, times, acc


def fp32_vs_fp64(qx, N, dim, L, dt, steps):
    """The headline workload stepped `steps` times from the same seeded IC in
    an FP32 and an FP64 context: relative L2 of psi and each run's mass
    drift.  The FP64 device path matches the FP64 oracle to 1e-12 at the
    sizes the oracle can run, so it stands in for it at this size."""
    import numpy as np
    out = {}
    psis = {}
    for prec in ("fp64", "fp32"):
        lat = qx.Lattice(dim, N, L, precision=prec, pref=1.0, rho_bar=1.0, k0_mode=0)
        lat.ic_grf(seed=0, amp=0.1)
        lat.update_potential()
        _, m0 = lat.density_stats()
        lat.leapfrog(dt, steps)
        _, m1 = lat.density_stats()
        psis[prec] = lat.psi
        out[f"mass_drift_{prec}"] = abs(m1 - m0) / m0
        lat.close()
    a, b = psis["fp32"], psis["fp64"]
    out["rel_l2"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    out["steps"] = steps
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--N", type=int, default=512)
    ap.add_argument("--dim", type=int, default=3)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--cpu-warmup", type=int, default=2)
    ap.add_argument("--cpu-steps", type=int, default=10,
                    help="oracle steps timed one by one after the warm-up (median reported)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the FP32-against-FP64 run of the same workload")
    ap.add_argument("--pass-reps", type=int, default=10)
    ap.add_argument("--no-tune", action="store_true", help="skip the tile-variant autotuner")
    ap.add_argument("--workload", default="kdk", choices=["kdk", "ensemble", "learned", "copy"],
                    help="kdk: the headline 512^3 step (default); ensemble: 96 uneven 2D "
                         "lattices; learned: corrector training and evaluation at 512^2; "
                         "copy: device-to-device state copies of known size (the reference "
                         "dispatch for scaling HBM traffic counters)")
    ap.add_argument("--graph-steps", type=int, default=16,
                    help="ensemble: steps per captured graph (0: every pass launched directly)")
    ap.add_argument("--compare-solo", action="store_true",
                    help="ensemble: also time each member stepped alone in its own context")
    ap.add_argument("--train-pairs", type=int, default=2000)
    ap.add_argument("--eval-steps", type=int, default=10)
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU (a single MI355X); the multi-GPU slab decomposition "
              "with all-to-all transposes is designed, not built.", file=sys.stderr)
        sys.exit(2)

    # the CPU leg's OpenMP threads bound to cores (read when the oracle's
    # library loads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import qaxions_amd as qx

    if args.workload == "ensemble":
        return bench_ensemble(qx, args)
    if args.workload == "learned":
        return bench_learned(qx, args)
    if args.workload == "copy":
        return bench_copy(qx, args)

    N, dim, L, dt = args.N, args.dim, 40.0, 2.5e-4
    S = N ** dim
    print(f"[bench] {dim}D {N}^{dim} {args.precision} L={L} dt={dt}", file=sys.stderr, flush=True)
    lat = qx.Lattice(dim, N, L, precision=args.precision, pref=1.0, rho_bar=1.0, k0_mode=0)
    lat.ic_grf(seed=0, amp=0.1)
    lat.update_potential()
    variants = lat.tune(dt, reps=5) if not args.no_tune else None
    r0, m0 = lat.density_stats()
    ev = HipEvents()
    stream = lat.stream

    # warmup: a segment left open (continues with K(ds))
    lat.leapfrog(dt, max(args.warmup, 1), first_half=True, last_half=False)
    lat.synchronize()
    e0, e1 = ev.event(), ev.event()
    w0 = time.perf_counter()
    ev.record(e0, stream)
    lat.leapfrog(dt, args.steps, first_half=False, last_half=True)
    ev.record(e1, stream)
    ms_total = ev.elapsed_ms(e0, e1)
    lat.synchronize()
    wall = time.perf_counter() - w0
    ms_step = ms_total / args.steps
    value = S / (ms_step * 1e-3)
    r1, m1 = lat.density_stats()
    print(f"[bench] {ms_step:.4f} ms/step (events), wall {wall*1e3/args.steps:.4f} ms/step, "
          f"mass drift {abs(m1-m0)/m0:.2e}", file=sys.stderr, flush=True)
    # four more segments of the same length, untimed by the contract's
    # clock: the spread and the median of five
    seg = [ms_step]
    for _ in range(4):
        lat.leapfrog(dt, 1, first_half=True, last_half=False)
        ev.record(e0, stream)
        lat.leapfrog(dt, args.steps, first_half=False, last_half=True)
        ev.record(e1, stream)
        seg.append(ev.elapsed_ms(e0, e1) / args.steps)
    seg_sorted = sorted(seg)
    print(f"[bench] five segments: {' '.join(f'{x:.4f}' for x in seg)} ms/step, median {seg_sorted[2]:.4f}",
          file=sys.stderr, flush=True)

    # the measurement between steps, timed on its own: P(k) of |psi|^2 in
    # linear bins of `spb` shells (128 bins at 512^3), every 10 steps
    nbins = max(1, (N // 2) // 2)
    lat.spectrum(source="rho", nbins=nbins, spb=2)
    reps_pk = 5
    t0 = time.perf_counter()
    for _ in range(reps_pk):
        lat.spectrum(source="rho", nbins=nbins, spb=2)
    ms_pk = (time.perf_counter() - t0) * 1e3 / reps_pk
    print(f"[bench] P(k) in {nbins} bins: {ms_pk:.3f} ms (wall, incl. readback); every 10 steps adds "
          f"{100 * ms_pk / (10 * ms_step):.1f}%", file=sys.stderr, flush=True)

    # per-pass live timing: the dominant pass and its roofline
    names, pms, pbytes = lat.time_step_passes(dt, reps=args.pass_reps)
    dom = max(range(len(names)), key=lambda i: pms[i])
    for n_, t_, b_ in zip(names, pms, pbytes):
        print(f"[bench] pass {n_:>18s}: {t_*1e3:8.1f} us  {b_/t_/1e6:8.1f} GB/s",
              file=sys.stderr, flush=True)
    csz = 8 if args.precision == "fp32" else 16
    B_site = algorithmic_bytes_per_site(N, c=csz, r=csz // 2)
    step_achieved = B_site * S / (ms_step * 1e-3)

    out = {
        "metric": (f"site-updates/s per KDK step at {N}^{dim} (FP32 state, FP64 phase reduction and sums)"
                   if args.precision == "fp32" else f"site-updates/s per KDK step at {N}^{dim} (FP64)"),
        "value": value,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic: seeded Gaussian random field IC, P(k)~k^-2 cut at k_Ny/2",
        "config": {"workload": f"KDK Schroedinger-Poisson step, {dim}D {N}^{dim}, L={L}, dt={dt}, "
                               f"pref=1, rho_bar=1, k=0 zeroed, GRF seed 0",
                   "lattice": [N] * dim, "precision": args.precision, "parallelism": "single GPU",
                   "tile_variants": variants},
        "roofline": {"bound": "hbm", "kernel": names[dom],
                     "achieved": pbytes[dom] / (pms[dom] * 1e-3) / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": pbytes[dom] / (pms[dom] * 1e-3) / HBM_PEAK,
                     "traffic": measured_traffic(names[dom]) if (N, dim, args.precision) == (512, 3, "fp32") else None,
                     "bytes_per_launch": pbytes[dom], "ms_per_launch": pms[dom]},
        "step_roofline": {"bound": "hbm", "algorithmic_bytes_per_site": B_site,
                          "achieved": step_achieved / 1e9, "peak": HBM_PEAK / 1e9,
                          "unit": "GB/s", "frac": step_achieved / HBM_PEAK,
                          "physical_bytes_per_site": sum(pbytes) / S,
                          "physical_frac": sum(pbytes) / (ms_step * 1e-3) / HBM_PEAK},
        "passes": {n_: {"us": t_ * 1e3, "GBps": b_ / t_ / 1e6} for n_, t_, b_ in zip(names, pms, pbytes)},
        "spectrum": {"nbins": nbins, "shells_per_bin": 2, "ms": ms_pk,
                     "overhead_every_10_steps": ms_pk / (10 * ms_step)},
        "mass_drift": abs(m1 - m0) / m0,
        "wall_ms_per_step": wall * 1e3 / args.steps,
        "segments_ms_per_step": seg, "median_ms_per_step": seg_sorted[2],
    }
    if not args.no_accuracy and args.precision == "fp32":
        print(f"[bench] accuracy: {args.steps} steps in FP32 and FP64 from the same IC",
              file=sys.stderr, flush=True)
        acc = fp32_vs_fp64(qx, N, dim, L, dt, args.steps)
        print(f"[bench] FP32 vs FP64 psi rel L2 {acc['rel_l2']:.3e}; mass drift FP32 "
              f"{acc['mass_drift_fp32']:.2e}, FP64 {acc['mass_drift_fp64']:.2e}", file=sys.stderr, flush=True)
        out["fp32_vs_fp64"] = acc
    if not args.no_cpu_baseline and dim == 3:
        cpu = host_cpu()
        print(f"[bench] cpu baseline: FP64 oracle, {args.cpu_warmup} + {args.cpu_steps} steps on "
              f"{cpu['threads']} threads ({cpu['cpu_model']}, {cpu['physical_cores']} physical cores, "
              f"share {cpu['cpu_share']})", file=sys.stderr, flush=True)
        v, med, times, acc = oracle_leg(qx, N, L, dt, args.cpu_warmup, args.cpu_steps, cpu["threads"])
        print(f"[bench] oracle {med:.3f} s/step (median of {len(times)}); after {acc['steps']} steps "
              f"FP32 vs oracle {acc['fp32_rel_l2']:.3e}, FP64 vs oracle {acc['fp64_rel_l2']:.3e}, "
              f"P(k) per bin FP32 {acc['fp32_pk_max_rel_err_in_cut']:.2e}", file=sys.stderr, flush=True)
        out["cpu_baseline"] = {"value": v, "unit": "site-updates/s", "cores": cpu["threads"], "kind": "port",
                               "cpu_model": cpu["cpu_model"], "physical_cores": cpu["physical_cores"],
                               "cpu_share": cpu["cpu_share"], "omp": "OMP_PROC_BIND=close OMP_PLACES=cores",
                               "fft": "project radix-2 (FFTW absent)",
                               "sample": f"median of {len(times)} merged-kick KDK steps, each timed alone after "
                                         f"{args.cpu_warmup} warm-up steps, of the FP64 C oracle on the same "
                                         f"{N}^{dim} lattice and IC",
                               "step_s": times}
        out["fp_vs_oracle"] = acc
    print(json.dumps(out), flush=True)


def bench_copy(qx, args):
    """Copies of one 512^3 FP32 state into another: per copy three runtime
    copy dispatches, psi (2^27 complex64 = 1 GiB read + 1 GiB written), V
    (0.5 GiB) and vt (2^26 + 2^18 complex64).  A counter's reading on the psi
    dispatch over 2 x 1 GiB is the factor that scales it to bytes."""
    N = args.N
    a = qx.Lattice(3, N, 40.0, precision="fp32")
    b = qx.Lattice(3, N, 40.0, precision="fp32")
    a.ic_grf(seed=0, amp=0.1)
    reps = 3
    ev = HipEvents()
    e0, e1 = ev.event(), ev.event()
    b.copy_from(a)
    ev.record(e0, b.stream)
    for _ in range(reps):
        b.copy_from(a)
    ev.record(e1, b.stream)
    ms = ev.elapsed_ms(e0, e1) / reps
    S = N ** 3
    moved = 2 * (8 * S + 4 * S + 8 * (S // N) * (N // 2 + 1))
    print(json.dumps({"metric": "device-to-device state copy", "value": moved / (ms * 1e-3) / 1e9,
                      "unit": "GB/s", "ms_per_copy": ms, "bytes_per_copy": moved,
                      "psi_dispatch_bytes": {"read": 8 * S, "written": 8 * S}}), flush=True)


def bench_ensemble(qx, args):
    """96 independent 2D realizations: 32 x 128^2, 32 x 512^2, 24 x 1024^2,
    8 x 2048^2, seeds 0-95, L ~ U[5, 40] (keyed RNG), dt = 1e-3 (dx/dx0)^2
    with dx0 = 10/256, hipGraph-captured KDK steps."""
    from qaxions_amd.rng import uniform
    Ns = [128] * 32 + [512] * 32 + [1024] * 24 + [2048] * 8
    Ls = [5.0 + 35.0 * uniform(2024, 99, i) for i in range(len(Ns))]
    dx0 = 10.0 / 256
    ds = [1e-3 * ((L / N) / dx0) ** 2 for N, L in zip(Ns, Ls)]
    steps = args.steps if args.steps != 50 else 500
    ens = qx.Ensemble(Ns, Ls, precision="fp32", pref=1.0, graph_steps=args.graph_steps)
    ens.ic_grf(list(range(len(Ns))))
    ens.update_potential()
    variants = ens.tune(ds, reps=3) if not args.no_tune else None
    m0 = ens.mass()
    ens.leapfrog(ds, max(args.warmup, 1))
    ev = HipEvents()
    e0, e1 = ev.event(), ev.event()
    ev.record(e0, ens.stream)
    ens.leapfrog(ds, steps)
    ev.record(e1, ens.stream)
    ms = ev.elapsed_ms(e0, e1)
    m1 = ens.mass()
    S = ens.sites
    solo = None
    if args.compare_solo:
        # the same members one lattice at a time, each in its own context
        # (its own launches; graph replays where the context uses them)
        tot = 0.0
        for i, (N, L) in enumerate(zip(Ns, Ls)):
            lat = qx.Lattice(2, N, L, "fp32", pref=1.0)
            lat.ic_grf(i)
            lat.update_potential()
            lat.leapfrog(ds[i], max(args.warmup, 1))
            f0, f1 = ev.event(), ev.event()
            ev.record(f0, lat.stream)
            lat.leapfrog(ds[i], steps)
            ev.record(f1, lat.stream)
            tot += ev.elapsed_ms(f0, f1)
            lat.close()
        solo = {"ms_per_step": tot / steps, "value": S * steps / (tot * 1e-3),
                "speedup_of_ensemble": tot / ms}
        print(f"[bench] members alone: {tot / steps:.4f} ms/step ({S * steps / (tot * 1e-3):.3e} "
              f"site-updates/s); the ensemble is {tot / ms:.2f}x that", file=sys.stderr, flush=True)
    c, r = 8, 4
    B = sum(((9 * c + 5 * r) + 4 * c * (N // 2 + 1) / N) * N * N for N in Ns)
    out = {"metric": "ensemble site-updates/s (96 uneven 2D lattices, FP32)",
           "value": S * steps / (ms * 1e-3), "unit": "site-updates/s", "n_gpus": 1, "steps": steps,
           "ms_per_step": ms / steps, "higher_is_better": True, "dtype": "f32",
           "data": "synthetic: per-member GRF ICs, seeds 0-95",
           "config": {"workload": "32x128^2 + 32x512^2 + 24x1024^2 + 8x2048^2, L~U[5,40], "
                                  f"dt=1e-3 (dx/dx0)^2, hipGraph {args.graph_steps} steps/replay",
                      "tile_variants": {str(k): v for k, v in variants.items()} if variants else None},
           "step_roofline": {"bound": "hbm", "algorithmic_bytes": B,
                             "achieved": B / (ms / steps * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9,
                             "unit": "GB/s", "frac": B / (ms / steps * 1e-3) / HBM_PEAK},
           "max_rel_mass_drift": float(np.max(np.abs(m1 - m0) / m0)),
           "members_alone": solo}
    print(json.dumps(out), flush=True)


def bench_learned(qx, args):
    """Learned coarse stepping at 512^2 (coarse 4 dt against 4 fine dt): a
    corrector trained on solver pairs per regime, evaluated inside the
    predictable horizon against a dt/4 reference, with the speed-up over
    plain stepping at equal error.  Regimes: the config's physics (pref = 1,
    dt = 1e-3), a larger step at pref = 1 (dt = 1e-2), strong coupling
    (pref = 100, dt = 1e-3)."""
    from qaxions_amd import learned
    N, L, k = 512, 10.0, 4
    regimes = [("pref1_dt1e-3", 1.0, 1e-3), ("pref1_dt1e-2", 1.0, 1e-2), ("pref100_dt1e-3", 100.0, 1e-3)]
    per_seed = 500
    nseeds = max(1, args.train_pairs // per_seed)
    evals, best = {}, None
    for name, pref, dt in regimes:
        t0 = time.perf_counter()
        net, losses, relative = learned.train(N, L, dt, k, seeds=range(1000, 1000 + nseeds),
                                              steps_per_seed=min(per_seed, args.train_pairs), lr=1e-3,
                                              lr_final=1e-4, pref=pref)
        t_train = time.perf_counter() - t0
        res = learned.evaluate(net, N, L, dt, k, seed=0, n_coarse=args.eval_steps, check_every=10,
                               mass_tol=1e-6, timing_reps=2, pref=pref, warm_steps=200)
        res["train"] = {"pairs": int(len(losses)), "seconds": t_train,
                        "relative_loss_first_200": float(relative[:200].mean()),
                        "relative_loss_last_200": float(relative[-200:].mean())}
        evals[name] = res
        sp = res["speedup_learned_vs_plain_at_equal_error"]
        best = sp if best is None else max(best, sp)
        print(f"[bench] learned {name}: errors vs dt/4 reference fine {res['fine']['rel_l2_vs_reference']:.2e} "
              f"coarse {res['coarse']['rel_l2_vs_reference']:.2e} corrected {res['learned']['rel_l2_vs_reference']:.2e}; "
              f"speed-up at equal error {sp:.2f}", file=sys.stderr, flush=True)
    out = {"metric": "learned coarse stepping at 512^2: speed-up over plain stepping at equal error (best regime)",
           "value": best, "unit": "x", "n_gpus": 1, "higher_is_better": True, "eval": evals,
           "dtype": "f32 (bf16 network weights, fp32 accumulation)", "data": "synthetic: seeded GRF ICs",
           "config": {"workload": f"2D {N}^2, L={L}, coarse {k} dt, {args.eval_steps} coarse steps, "
                                  f"mass tol 1e-6 per 10-step chunk, dt/4 reference"}}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
