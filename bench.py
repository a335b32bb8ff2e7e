"""Throughput of GALIC's equilibrium hot path on one MI355X.

Workload: the precision-form model -- a beta=0.3 Hernquist halo (M=100,
a=30 kpc) plus an exponential sech^2 disk (Md=4, Rd=3.5, z0=0.35 kpc),
N=1e6 (80% halo, 20% disk), 256^2 force tables, 64^2 response grids.

A step is one optimisation iteration of the batched FP32 path: every particle
gets K=8 trial velocities (Gaussian perturbations of 10% of its target
dispersions, drawn on the device), its current orbit and the K trial orbits
are integrated and binned, the merit change of each trial is evaluated, the
best improving trial is taken, and the grids are committed and re-fitted.
orbits/s counts every integrated and binned orbit (current + trials).

Beside it the line reports the plain FP64 path on the same model with the
reference's method (one fresh trial per particle, reps of 1e5), the relative
difference of the final chi^2 (Sall) between the FP32 and FP64 paths run on
identical trials, the roofline of the FP32 orbit kernel, and the CPU oracle
timed on the host's cores on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from galic_amd import _abi, model  # noqa: E402
from galic_amd.runtime import GalicGPU  # noqa: E402

F_STEP = 376          # algorithmic FP operations per leapfrog step (orbit_response.c:211-458 + lookups)
PEAK_FP32 = 157.3     # TFLOP/s, MI355X published FP32 rate
PEAK_HBM = 8000.0     # GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fp32_iteration(g, perm, k, it, sigma):
    acc = g.fp32_rep(perm, k=k, sigma_frac=sigma, seed=1000 + it)
    return int((acc > 0).sum())


def cpu_baseline(cfg, seconds, nthreads):
    """The oracle's optimize() (orbit pair + two merits + write-back, main.c:365-527),
    rep after rep of 1000 trials until about `seconds` of CPU work.  The oracle holds a
    20000-particle realisation of the same model, whose initial response pass gives
    the global state the merits are evaluated on; per-trial work is that of the full
    model (same tables, grids and orbit lengths)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the checker, used here only as the CPU baseline
    ws = model.make_workload(cfg, n=20000, target_samples=200000)
    o = oracle.Oracle(ws)
    o.calc_all_response_fields(nthreads)
    perm = model.permutation(ws, 3)
    ntr = 0
    dt = 0.0
    rep = 0
    while dt < seconds and rep < 400:
        idx = perm[(rep * 1000) % ws.n:(rep * 1000) % ws.n + 1000]
        v = model.trial_velocities(ws, idx, 77, rep)
        t0 = time.perf_counter()
        o.rep(idx, v, nthreads)
        dt += time.perf_counter() - t0
        ntr += len(idx)
        rep += 1
    o.close()
    return 2 * ntr / dt, dt, ntr


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1000000)
    ap.add_argument("--k", type=int, default=8)
    ap.add_argument("--sigma", type=float, default=0.1)
    ap.add_argument("--fp64-reps", type=int, default=10, help="reps of the FP64 iteration (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline budget in seconds (0: skip)")
    ap.add_argument("--target-samples", type=int, default=4000000)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              "(the path shards by particles, see the design notes, but no multi-GPU run is made).",
              file=sys.stderr)
        sys.exit(2)

    t0 = time.perf_counter()
    cfg = model.config("halo-disk-1e6")
    w = model.make_workload(cfg, n=args.n, target_samples=args.target_samples)
    log(f"[bench] workload {cfg.name} n={w.n} built in {time.perf_counter() - t0:.1f}s")
    perm = model.permutation(w, 1)

    # ---------------------------------------------------------- FP32 batched
    g = GalicGPU(w)
    g.fp32_calc_all_response_fields()
    variant, bpc = g.fp32_autotune(min(w.n, 524288))
    log(f"[bench] fp32 initial pass done; autotuned orbit kernel: flush variant {variant}, {bpc} workgroups/CU")
    for it in range(args.warmup):
        fp32_iteration(g, perm, args.k, it, args.sigma)
        log(f"[bench] warmup {it} done")
    g.reset_timings()
    t0 = time.perf_counter()
    accepted = 0
    for it in range(args.steps):
        accepted += fp32_iteration(g, perm, args.k, args.warmup + it, args.sigma)
        log(f"[bench] step {it} done")
    wall = time.perf_counter() - t0
    tm = g.timings()
    orbits = args.steps * w.n * (args.k + 1)
    sall32 = {t: float(g.fit(t)[7]) for t in w.types}
    g.close()
    value = orbits / wall
    steps_total = tm["steps"]
    achieved = F_STEP * steps_total / (tm["orbit_ms"] * 1e-3) / 1e12
    log(f"[bench] fp32: {value:.4g} orbits/s, {steps_total / wall:.4g} steps/s, orbit kernel {tm['orbit_ms']:.1f} ms "
        f"of {wall * 1e3:.1f} ms, {achieved:.2f} TFLOP/s algorithmic")

    # ------------------------------------------------------ FP64 plain path
    fp64 = None
    chi2 = None
    if args.fp64_reps > 0:
        ntr = w.n // args.fp64_reps
        g64 = GalicGPU(w)
        g64.calc_all_response_fields()
        g32 = GalicGPU(w)
        g32.fp32_calc_all_response_fields()
        reps = [perm[r * ntr:(r + 1) * ntr] for r in range(args.fp64_reps)]
        trials = [model.trial_velocities(w, idx, 5, r) for r, idx in enumerate(reps)]
        init = {t: {"fp64": float(g64.fit(t)[7]), "fp32": float(g32.fit(t)[7])} for t in w.types}
        g64.reset_timings()
        t1 = time.perf_counter()
        acc64 = [g64.rep(idx, v)[0] for idx, v in zip(reps, trials)]
        wall64 = time.perf_counter() - t1
        tm64 = g64.timings()
        agree = []
        for (idx, v), a64 in zip(zip(reps, trials), acc64):
            a32 = g32.fp32_rep(idx, k=1, vtry=v)
            agree.append(float(np.mean((a32 > 0) == (a64 > 0))))
        log(f"[bench] chi2 check: initial Sall {init}, decision agreement per rep {agree}")
        chi2 = {t: {"fp64": float(g64.fit(t)[7]), "fp32": float(g32.fit(t)[7])} for t in w.types}
        worst = max(abs(c["fp32"] - c["fp64"]) / abs(c["fp64"]) for c in chi2.values())
        fp64 = dict(orbits_per_s=2 * ntr * args.fp64_reps / wall64, ms_per_iteration=wall64 * 1e3,
                    steps_per_s=tm64["steps"] / wall64, orbit_kernel_ms=tm64["orbit_ms"])
        g64.close()
        g32.close()
        log(f"[bench] fp64: {fp64['orbits_per_s']:.4g} orbits/s; chi2 rel diff fp32 vs fp64 {worst:.2e}")
        chi2 = dict(per_type=chi2, max_rel_diff=worst, initial=init, decision_agreement=agree)

    # ------------------------------------------------------------- CPU
    cpu = None
    if args.cpu_seconds > 0:
        nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        rate, dt, ntr = cpu_baseline(cfg, args.cpu_seconds, nthreads)
        cpu = {"value": rate, "unit": "orbits/s", "cores": nthreads, "kind": "port",
               "sample": f"{ntr} trials in reps of 1000 (2 orbits + 2 merit evaluations each, the reference's "
                         f"optimize()) on a 20000-particle realisation of this model, {nthreads} OpenMP threads, "
                         f"{dt:.1f} s"}
        log(f"[bench] cpu: {rate:.4g} orbits/s on {nthreads} threads ({ntr} trials, {dt:.1f} s)")

    # ---------------------------------------------------- HBM calibration
    copy_ms = np.zeros(1)
    nbytes = 1 << 30
    _abi.check(_abi.lib().galic_copy_probe(0, nbytes, 5, _abi.dptr(copy_ms)))
    copy_gbs = 2 * nbytes / (copy_ms[0] * 1e-3) / 1e9

    line = {
        "metric": "orbits/s integrated+binned: batched FP32 vs plain FP64 GPU at equal final chi^2",
        "value": value,
        "unit": "orbits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "precision form: beta=0.3 Hernquist halo + exponential sech^2 disk",
                   "n_particles": w.n, "halo_disk_split": [int(w.params.NType[1]), int(w.params.NType[2])],
                   "force_table": [w.params.FG_Nbin, w.params.FG_Nbin], "response_grid": [1 << w.params.DG_MaxLevel] * 2,
                   "trials_per_particle": args.k, "trial_sigma_frac": args.sigma,
                   "orbits_per_step": w.n * (args.k + 1), "torbit_fac": cfg.torbit_fac,
                   "autotuned": {"slot_flush": ["load-add-store", "atomic"][variant], "workgroups_per_cu": bpc}},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32, "traffic": None,
                     "kernel": "k_orbit32", "note": "compute-bound on the FP32 vector ALU (MFMA unused); "
                     f"achieved = {F_STEP} algorithmic flop/step x leapfrog steps / HIP-event time of the kernel"},
        "steps_per_s": steps_total / wall,
        "orbit_kernel_ms_per_step": tm["orbit_ms"] / args.steps,
        "accept_frac": accepted / (args.steps * w.n),
        "final_sall": sall32,
        "fp64_plain": fp64,
        "fp32_vs_fp64_speedup": (value / fp64["orbits_per_s"]) if fp64 else None,
        "chi2_check": chi2,
        "cpu_baseline": cpu,
        "copy_probe_GBs": copy_gbs,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
