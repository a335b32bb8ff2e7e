"""Throughput of GALIC's equilibrium hot path on one MI355X.

Workload: the precision-form model -- a beta=0.3 Hernquist halo (M=100,
a=30 kpc) plus an exponential sech^2 disk (Md=4, Rd=3.5, z0=0.35 kpc),
N=1e6 (80% halo, 20% disk), 256^2 force tables, 64^2 response grids.

A step is one optimisation iteration of the batched FP32 path over all N
particles in reps of 1e4 (each rep evaluated against the grids at its start and
committed before the next, main.c:93-100): every particle gets K=8 trial
velocities drawn on the device -- one fresh draw from optimize()'s own trial
distribution (main.c:383-435) and 7 Gaussian perturbations of 10% of its target
dispersions -- its current orbit and the K trial orbits are integrated and
binned, the merit change of each trial is evaluated, the best improving trial
is taken, and each rep's deltas are committed and re-fitted.  orbits/s counts
every integrated and binned orbit (current + trials).

Beside it the line reports the plain FP64 path on the same model with the
reference's method (one fresh trial per particle, reps of 1e4), the final
chi^2 (Sall) of the FP32 and FP64 paths run on identical trials, the chi^2
reached by each method after the same number of iterations, the roofline of
the FP32 orbit kernel, and the CPU oracle timed on the host's cores on a
bounded sample.
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
from galic_amd.runtime import (TREE_FP32, TREE_PLUMMER, TREE_QUADRUPOLE, GalicGPU, direct_forces,  # noqa: E402
                               tree_fill, tree_forces, tree_last_timings)

F_STEP = 376          # algorithmic FP operations per leapfrog step (orbit_response.c:211-458 + lookups)
# HBM bytes per leapfrog step of k_orbit32_log in the iteration's launches, from the FETCH_SIZE and
# WRITE_SIZE counter runs calibrated against a plain copy (profiles/k_orbit32_log_halo-disk-1e6.txt)
HBM_BYTES_PER_STEP = 22.8
PEAK_FP32 = 157.3     # TFLOP/s, MI355X published FP32 rate
PEAK_HBM = 8000.0     # GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fp32_iteration(g, perm, k, it, sigma, rep):
    acc = g.fp32_iterate(perm, k=k, rep_size=rep, sigma_frac=sigma, seed=1000 + it)
    return int((acc > 0).sum())


def fp64_trials(w, perm, rep, it):
    """Fresh trial velocities of one iteration of the plain method (main.c:407-435), per rep."""
    return [(perm[a:a + rep], model.trial_velocities(w, perm[a:a + rep], 5 + it, r))
            for r, a in enumerate(range(0, len(perm), rep))]


def fp64_iteration(g, trials):
    return np.concatenate([g.rep(idx, v)[0] for idx, v in trials])


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


def tree_sources(p, n, seed):
    """The composite's mass (halo 75%, disk 20%, bulge 5% of the particles) sampled
    from its profiles, as forcedensitygrid_calculate's sources (grid.c:659-678)."""
    ids = np.arange(n, dtype=np.uint64)
    nd, nb = n // 5, n // 20
    nh = n - nd - nb
    pos = np.concatenate([model.sample_halo(p, seed, 11, ids[:nh]), model.sample_disk(p, seed, 12, ids[nh:nh + nd]),
                          model.sample_halo(p, seed, 13, ids[nh + nd:], a=p.Bulge_A)])
    mass = np.concatenate([np.full(nh, p.Halo_Mass / nh), np.full(nd, p.Disk_Mass / nd), np.full(nb, p.Bulge_Mass / nb)])
    return np.ascontiguousarray(pos), mass


def tree_leg(n, nbin, ncheck):
    """Force-field fill: Barnes-Hut table fill over n composite sources (seed 3), theta = 0.4,
    FG_Nbin^2 x 8 sections of targets.  Two walks: the force-field configuration (FP32 walk,
    FP64 quadrupole moments, Plummer softening 0.05) and the reference's kernel (monopole,
    spline softening) in FP32; force errors against FP64 direct summation with the same
    softening on ncheck random table nodes."""
    cfg = model.config("composite-4e6", fg_nbin=nbin, theta=0.4)
    p = model.make_params(cfg)
    t0 = time.perf_counter()
    spos, smass = tree_sources(p, n, 3)
    log(f"[bench] tree: {n} sources sampled in {time.perf_counter() - t0:.1f}s")
    ntarget = 8 * nbin * nbin
    # error check on random table nodes (section, z row, R column)
    rng = np.random.default_rng(3)
    s, k, j = rng.integers(0, 8, ncheck), rng.integers(0, nbin, ncheck), rng.integers(0, nbin, ncheck)
    r1 = p.FG_Rmin * (p.FG_Fac ** j - 1.0)
    z1 = p.FG_Rmin * (p.FG_Fac ** k - 1.0)
    phi = 2 * np.pi / 8 * s
    tpos = np.stack([r1 * np.cos(phi), r1 * np.sin(phi), z1], axis=1)
    out = dict(sources=n, targets=ntarget, fg_nbin=nbin, theta=0.4, softening=cfg.softening)
    for name, mode in (("quadrupole_plummer_fp32", TREE_FP32 | TREE_QUADRUPOLE | TREE_PLUMMER),
                       ("reference_kernel_fp32", TREE_FP32)):
        t0 = time.perf_counter()
        tree_fill(p, spos, smass, precision=mode)
        wall = time.perf_counter() - t0
        tm = tree_last_timings()
        acc32, pot32, _ = tree_forces(p, spos, smass, tpos, precision=mode)
        accd, potd = direct_forces(p, spos, smass, tpos, plummer=bool(mode & TREE_PLUMMER))
        err = np.linalg.norm(acc32 - accd, axis=1) / np.linalg.norm(accd, axis=1)
        perr = np.abs(pot32 / potd - 1)
        out[name] = dict(wall_ms=wall * 1e3, build_ms=tm["build_ms"], walk_ms=tm["walk_ms"],
                         interactions=tm["interactions"], interactions_per_s=tm["interactions"] / (tm["walk_ms"] * 1e-3),
                         force_err_vs_direct_fp64=dict(nodes=ncheck, median=float(np.median(err)),
                                                       p99=float(np.quantile(err, 0.99)), max=float(err.max())),
                         pot_err_vs_direct_fp64=dict(median=float(np.median(perr)), max=float(perr.max())))
        log(f"[bench] tree fill {name}: {wall * 1e3:.0f} ms wall (build {tm['build_ms']:.0f} ms, walk "
            f"{tm['walk_ms']:.0f} ms, {tm['interactions']:.3g} interactions), force error vs direct median "
            f"{np.median(err):.2e}")
    return out


def target_leg(cfg, n, ncpu):
    """Sampled targets of the bench model (grid.c:536-617, 941-1076), as the reference builds them:
    n fresh points per component for the density map and another n for the five energy-grid maps.
    CPU baseline: the same kind of points drawn by the host sampler and binned by the oracle
    (one thread), ncpu points per map."""
    w = model.make_workload(cfg, n=2000, target_samples=20000)   # parameters and tables only
    p = w.params
    g = GalicGPU(w)
    per_type = {}
    for t in w.types:
        beta, tabs = model.disp_tables(p, cfg, t)
        g.set_disp_tables(t, beta, **tabs)
        g.sample_targets(t, 1 << 20, seed=2)   # warm-up
        t0 = time.perf_counter()
        _, _, nd, ms_d = g.sample_targets(t, n, seed=1, stream=300 + t, energy=False)
        _, _, ne, ms_e = g.sample_targets(t, n, seed=1, stream=400 + t, density=False)
        wall = time.perf_counter() - t0
        per_type[t] = dict(points=2 * n, binned=nd + ne, density_kernel_ms=ms_d, energy_kernel_ms=ms_e,
                           wall_ms=wall * 1e3, points_per_s=2 * n / ((ms_d + ms_e) * 1e-3))
        log(f"[bench] targets type {t}: 2 x {n:.3g} points, kernels {ms_d:.1f} + {ms_e:.1f} ms, wall {wall * 1e3:.0f} ms")
    g.close()
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the checker, used here only as the CPU baseline
    o = oracle.Oracle(w, set_particles=False)
    cpu_s = 0.0
    for t in w.types:
        beta, tabs = model.disp_tables(p, cfg, t)
        o.set_disp_tables(t, beta, **tabs)
        ids = np.arange(ncpu, dtype=np.uint64)
        t0 = time.perf_counter()
        pos = model.sample_type(p, t, 1, 300 + t, ids)
        o.bin_targets(t, pos, energy=False)
        pos = model.sample_type(p, t, 1, 400 + t, ids)
        o.bin_targets(t, pos, density=False)
        cpu_s += time.perf_counter() - t0
    o.close()
    cpu_rate = 2 * ncpu * len(w.types) / cpu_s
    gpu_rate = sum(v["points"] for v in per_type.values()) / sum(
        (v["density_kernel_ms"] + v["energy_kernel_ms"]) * 1e-3 for v in per_type.values())
    log(f"[bench] targets: GPU {gpu_rate:.4g} points/s, CPU {cpu_rate:.4g} points/s (1 thread)")
    return dict(points_per_map=n, per_type=per_type, points_per_s=gpu_rate,
                cpu=dict(points_per_s=cpu_rate, threads=1, kind="port",
                         sample=f"{ncpu} points per map and type: host sampler (numpy) + oracle binning"),
                gpu_over_cpu=gpu_rate / cpu_rate)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1000000)
    ap.add_argument("--k", type=int, default=8)
    ap.add_argument("--sigma", type=float, default=0.1)
    ap.add_argument("--fresh", type=int, default=1,
                    help="of the K trials, how many are fresh draws from optimize()'s distribution")
    ap.add_argument("--rep", type=int, default=10000, help="trials per rep (both paths)")
    ap.add_argument("--fp64-iters", type=int, default=-1,
                    help="iterations of the plain FP64 path (default warmup+steps; 0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline budget in seconds (0: skip)")
    ap.add_argument("--target-samples", type=int, default=4000000)
    ap.add_argument("--precision-sample", type=int, default=2000,
                    help="orbits integrated by both paths for the FP32 error figures (0: skip)")
    ap.add_argument("--tree-n", type=int, default=10000000, help="sources of the table-fill leg (0: skip)")
    ap.add_argument("--tree-nbin", type=int, default=512)
    ap.add_argument("--target-points", type=int, default=100000000,
                    help="points per map and component of the sampled-target leg (0: skip)")
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
    g.fp32_set_fresh_trials(args.fresh)
    hist32 = [{t: float(g.fit(t)[7]) for t in w.types}]
    variant, bpc = g.fp32_autotune(min(w.n, 524288))
    log(f"[bench] fp32 initial pass done; autotuned orbit kernel: {bpc} workgroups/CU")
    for it in range(args.warmup):
        fp32_iteration(g, perm, args.k, it, args.sigma, args.rep)
        hist32.append({t: float(g.fit(t)[7]) for t in w.types})
        log(f"[bench] warmup {it} done")
    g.reset_timings()
    t0 = time.perf_counter()
    accepted = 0
    for it in range(args.steps):
        accepted += fp32_iteration(g, perm, args.k, args.warmup + it, args.sigma, args.rep)
        hist32.append({t: float(g.fit(t)[7]) for t in w.types})
        log(f"[bench] step {it} done")
    wall = time.perf_counter() - t0
    tm = g.timings()
    orbits = args.steps * w.n * (args.k + 1)
    g.close()
    value = orbits / wall
    steps_total = tm["steps"]
    achieved = F_STEP * steps_total / (tm["orbit_ms"] * 1e-3) / 1e12
    breakdown = dict(orbit_kernel=tm["orbit_ms"] / args.steps, merit_kernel=tm["eval_ms"] / args.steps,
                     fit=tm["fit_ms"] / args.steps,
                     other=(wall * 1e3 - tm["orbit_ms"] - tm["eval_ms"] - tm["fit_ms"]) / args.steps)
    log(f"[bench] fp32: {value:.4g} orbits/s, {steps_total / wall:.4g} steps/s, orbit kernel {tm['orbit_ms']:.1f} ms "
        f"of {wall * 1e3:.1f} ms, {achieved:.2f} TFLOP/s algorithmic; ms per step {breakdown}, "
        f"log relaunches {tm['log_relaunches']:.0f}")

    # ------------------------------------------------------ FP64 plain path
    fp64 = None
    chi2 = None
    iters64 = args.warmup + args.steps if args.fp64_iters < 0 else args.fp64_iters
    if iters64 > 0:
        trials = [fp64_trials(w, perm, args.rep, it) for it in range(iters64)]
        g64 = GalicGPU(w)
        g64.calc_all_response_fields()
        hist64 = [{t: float(g64.fit(t)[7]) for t in w.types}]
        g64.reset_timings()
        wall64 = 0.0
        for it in range(iters64):
            t1 = time.perf_counter()
            fp64_iteration(g64, trials[it])
            wall64 += time.perf_counter() - t1
            hist64.append({t: float(g64.fit(t)[7]) for t in w.types})
            log(f"[bench] fp64 iteration {it} done")
        del trials
        tm64 = g64.timings()
        g64.close()
        fp64 = dict(orbits_per_s=2 * w.n * iters64 / wall64, ms_per_iteration=wall64 / iters64 * 1e3,
                    steps_per_s=tm64["steps"] / wall64, orbit_kernel_ms=tm64["orbit_ms"], iterations=iters64)
        log(f"[bench] fp64: {fp64['orbits_per_s']:.4g} orbits/s")

        # identical trials on both paths (K=1): the precision of the FP32 path alone
        g64 = GalicGPU(w)
        g64.calc_all_response_fields()
        g32 = GalicGPU(w)
        g32.fp32_calc_all_response_fields()
        init = {t: {"fp64": float(g64.fit(t)[7]), "fp32": float(g32.fit(t)[7])} for t in w.types}
        agree, drift = [], []
        for r, a in enumerate(range(0, w.n, args.rep)):
            idx = perm[a:a + args.rep]
            v = model.trial_velocities(w, idx, 5, r)
            a64 = g64.rep(idx, v)[0]
            a32 = g32.fp32_rep(idx, k=1, vtry=v)
            agree.append(round(float(np.mean((a32 > 0) == (a64 > 0))), 5))
            drift.append({t: float(g32.fit(t)[7] / g64.fit(t)[7] - 1) for t in w.types})
        same = {t: {"fp64": float(g64.fit(t)[7]), "fp32": float(g32.fit(t)[7])} for t in w.types}
        g64.close()
        g32.close()
        worst = max(abs(c["fp32"] - c["fp64"]) / abs(c["fp64"]) for c in same.values())
        every = max(1, len(agree) // 10)
        agree, drift = agree[every - 1::every], drift[every - 1::every]
        log(f"[bench] identical trials: initial Sall {init}, final {same}, rel diff {worst:.2e}, "
            f"decision agreement every {every} reps {agree}, Sall fp32/fp64 - 1 every {every} reps {drift}")
        n_cmp = min(len(hist32), len(hist64)) - 1
        # time to reach the plain path's final chi^2: iterations of the batched path
        # until every type's Sall is at or below the plain path's after iters64
        goal = hist64[-1]
        reach = next((i for i in range(1, len(hist32)) if all(hist32[i][t] <= goal[t] for t in w.types)), None)
        ms32 = wall / args.steps * 1e3
        to_chi2 = dict(plain_fp64_iterations=iters64, plain_fp64_ms=fp64["ms_per_iteration"] * iters64,
                       batched_fp32_iterations=reach, batched_fp32_ms=(reach * ms32) if reach else None,
                       speedup=(fp64["ms_per_iteration"] * iters64 / (reach * ms32)) if reach else None)
        # equal final chi^2: after the same iterations the batched path's Sall is within
        # 1e-3 relative of the plain path's, or below it, for every type
        equal = {t: hist32[n_cmp][t] <= hist64[n_cmp][t] * (1 + 1e-3) for t in w.types}
        chi2 = dict(equal_final_chi2=all(equal.values()),
                    identical_trials=dict(per_type=same, max_rel_diff=worst, initial=init, decision_agreement=agree,
                                          rel_diff_per_rep=drift, sampled_every_reps=every, iterations=1,
                                          rep=args.rep),
                    methods=dict(iterations=n_cmp, batched_fp32=hist32[n_cmp], plain_fp64=hist64[n_cmp],
                                 history_batched_fp32=hist32, history_plain_fp64=hist64),
                    time_to_plain_final_chi2=to_chi2)
        log(f"[bench] time to the plain path's final chi^2: {to_chi2}")

    # ------------------------------------------- FP32 orbits against FP64
    precision = None
    if args.precision_sample > 0:
        g = GalicGPU(w)
        sel = np.random.default_rng(7).choice(w.n, args.precision_sample, replace=False)
        errs, d32, d64, dsteps = [], [], [], []
        for a in range(0, len(sel), 500):
            s = sel[a:a + 500]
            orb = (w.pos[s], w.vel[s], w.mass[s], w.tint[s], w.ptype[s])
            r64 = g.orbit_responses(*orb)
            r32 = g.orbit_responses(*orb, precision=32)
            errs.append(np.abs(r32["fields"] - r64["fields"]).sum(axis=(1, 2)) / np.abs(r64["fields"]).sum(axis=(1, 2)))
            d32.append(np.abs(r32["e1"] / r32["e0"] - 1))
            d64.append(np.abs(r64["e1"] / r64["e0"] - 1))
            dsteps.append(np.abs(r32["steps"] - r64["steps"]) / r64["steps"])
        g.close()
        errs, d32, d64, dsteps = map(np.concatenate, (errs, d32, d64, dsteps))
        precision = dict(orbits=len(sel), response_l1_rel_err=dict(median=float(np.median(errs)),
                                                                  p99=float(np.quantile(errs, 0.99))),
                         energy_drift=dict(fp32_median=float(np.median(d32)), fp64_median=float(np.median(d64)),
                                           fp32_p99=float(np.quantile(d32, 0.99)),
                                           fp64_p99=float(np.quantile(d64, 0.99))),
                         step_count_rel_diff_median=float(np.median(dsteps)))
        log(f"[bench] fp32 vs fp64 orbits: {precision}")

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

    # ------------------------------------------------------ table fill
    tree = tree_leg(args.tree_n, args.tree_nbin, 4096) if args.tree_n > 0 else None
    # ------------------------------------------------- sampled targets
    targets = target_leg(cfg, args.target_points, 1000000) if args.target_points > 0 else None

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
                   "trials_per_particle": args.k, "fresh_trials": args.fresh, "trial_sigma_frac": args.sigma,
                   "orbits_per_step": w.n * (args.k + 1), "rep": args.rep, "torbit_fac": cfg.torbit_fac,
                   "autotuned": {"workgroups_per_cu": bpc}},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32,
                     "traffic": HBM_BYTES_PER_STEP * steps_total / max(1, tm["orbit_launches"]),
                     "kernel": "k_orbit32_log",
                     "note": f"roof: the FP32 vector ALU (no step is a contraction, MFMA unused); achieved = {F_STEP} "
                             "algorithmic flop/step x leapfrog steps / HIP-event time of the kernel; traffic = HBM "
                             f"bytes per launch ({HBM_BYTES_PER_STEP} B/step from calibrated FETCH_SIZE+WRITE_SIZE, "
                             "almost all of it response-log writes); waves spend most cycles waiting on memory "
                             "(SQ_WAIT_ANY), not on the ALU"},
        "steps_per_s": steps_total / wall,
        "ms_per_step_breakdown": breakdown,
        "accept_frac": accepted / (args.steps * w.n),
        "fp64_plain": fp64,
        "fp32_vs_fp64_speedup": (value / fp64["orbits_per_s"]) if fp64 else None,
        "chi2_check": chi2,
        "fp32_orbit_precision": precision,
        "cpu_baseline": cpu,
        "tree_fill": tree,
        "target_fill": targets,
        "copy_probe_GBs": copy_gbs,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
