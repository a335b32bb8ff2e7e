"""Throughput of GALIC's equilibrium hot path on one MI355X: the full v1.1 model.

Workload: a Hernquist halo (M=100, a=30 kpc, isotropic) + an exponential sech^2
disk (Md=4, Rd=3.5, z0=0.35 kpc, velocity structure 3 with sigma_R = sigma_z) +
a Hernquist bulge (Mb=1, ab=0.7 kpc, isotropic), N=2e7 particles (75% halo, 20%
disk, 5% bulge), 512^2 force tables filled by the GPU tree from 1e7 sampled
sources (FP32 walk, quadrupole moments, Plummer softening 0.05 kpc) and filled
again at the start of every iteration, 512^2 response grids (level 9), energy
targets from 1e8 device-sampled points per component.

A step is a block of reps of the batched FP32 path (reps of 1e4 particles, each
evaluated against the grids at its start and committed before the next,
main.c:93-100): every particle gets K=8 trial velocities drawn on the device --
one fresh draw from optimize()'s own distribution (main.c:383-435) and 7
Gaussian perturbations of 10% of its target dispersions -- its current orbit and
the K trial orbits are integrated and binned, the merit change of each trial is
evaluated, the best improving trial is taken, and each rep is committed and
re-fitted.  A run of steps covering all N particles is one iteration; the table
fill at the start of each iteration is inside the timed region.  orbits/s counts
every integrated and binned orbit (current + trials).

Beside it the line reports the plain FP64 path on the same model with the
reference's method (one fresh trial per particle, reps of 1e4), the chi^2 (Sall)
reached by both methods on a smaller realisation of the same model and the time
the batched path takes to reach the plain path's final chi^2, the FP32 and FP64
paths on identical trials, the roofline of the dominant kernel, and the CPU
oracle timed on the host's cores on a bounded sample.
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
from galic_amd.runtime import (TREE_FP32, TREE_GROUP, TREE_PLUMMER, TREE_QUADRUPOLE, GalicGPU, direct_forces,  # noqa: E402
                               tree_fill, tree_forces, tree_last_timings)

F_STEP = 376          # algorithmic FP operations per leapfrog step (orbit_response.c:211-458 + lookups)
PEAK_FP32 = 157.3     # TFLOP/s, MI355X published FP32 vector rate
PEAK_HBM = 8000.0     # GB/s
TREE_MODE_TARGET = TREE_FP32 | TREE_QUADRUPOLE | TREE_PLUMMER   # the force-field configuration's walk
TREE_MODE = TREE_MODE_TARGET | TREE_GROUP   # ... with the targets' groups walking together
# HBM bytes per leapfrog step of k_orbit32_log in the full model's timed launches, from the FETCH_SIZE
# and WRITE_SIZE counter runs calibrated against a plain copy (profiles/k_orbit32_log_composite-2e7.txt:
# 3.0 B fetched + 38.1 B written per step)
HBM_BYTES_PER_STEP = 41.1


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def full_model(n, tree_n):
    """The full v1.1 model: workload arrays on the host, every component sampled into the table."""
    cfg = model.config("composite-2e7", n=n, halo_in_table=True)
    w = model.make_workload(cfg, n=n, target_samples=200000)
    p = w.params
    nh, nd = int(0.75 * tree_n), int(0.2 * tree_n)
    p.SampleForceNhalo, p.SampleForceNdisk, p.SampleForceNbulge = nh, nd, tree_n - nh - nd
    return cfg, w


def setup(w, cfg, target_points, targets=None):
    """A context of the model: energy targets sampled on the device (or given), tree sources
    sampled on the device (seed 3) and the tables filled from them."""
    g = GalicGPU(w)
    if targets is None:
        targets = {}
        for t in w.types:
            beta, tabs = model.disp_tables(w.params, cfg, t)
            g.set_disp_tables(t, beta, **tabs)
            _, eg, _, _ = g.sample_targets(t, target_points, seed=1, stream=400 + t, density=False)
            targets[t] = [w.targets[t][0]] + [eg[f] for f in range(5)]
    for t, st in targets.items():
        g.set_targets(t, st)
    g.sample_tree_sources(3)
    g.retabulate(TREE_MODE)
    return g, targets


def fp64_trials(w, idx, seed, rep):
    """Fresh trial velocities of the plain method (main.c:407-435)."""
    return model.trial_velocities(w, idx, seed, rep)


def cpu_baseline(seconds, nthreads, n):
    """The oracle's optimize() (orbit pair + two merits + write-back, main.c:365-527) on an
    n-particle realisation of the full model (same grids and tables), rep after rep of 200
    trials until about `seconds` of CPU work; the initial response pass gives the state the
    merits are evaluated on."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the checker, used here only as the CPU baseline
    cfg = model.config("composite-2e7")
    ws = model.make_workload(cfg, n=n, target_samples=200000)
    o = oracle.Oracle(ws)
    t0 = time.perf_counter()
    o.calc_all_response_fields(nthreads)
    t_init = time.perf_counter() - t0
    perm = model.permutation(ws, 3)
    ntr, dt, rep = 0, 0.0, 0
    while dt < seconds and rep < 1000:
        a = (rep * 200) % ws.n
        idx = perm[a:a + 200]
        v = model.trial_velocities(ws, idx, 77, rep)
        t0 = time.perf_counter()
        o.rep(idx, v, nthreads)
        dt += time.perf_counter() - t0
        ntr += len(idx)
        rep += 1
    o.close()
    return 2 * ntr / dt, dt, ntr, t_init


def tree_sources(p, n, seed):
    """The composite's mass (halo 75%, disk 20%, bulge 5% of the particles) sampled from its
    profiles, as forcedensitygrid_calculate's sources (grid.c:659-678)."""
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
    rng = np.random.default_rng(3)
    s, k, j = rng.integers(0, 8, ncheck), rng.integers(0, nbin, ncheck), rng.integers(0, nbin, ncheck)
    r1 = p.FG_Rmin * (p.FG_Fac ** j - 1.0)
    z1 = p.FG_Rmin * (p.FG_Fac ** k - 1.0)
    phi = 2 * np.pi / 8 * s
    tpos = np.stack([r1 * np.cos(phi), r1 * np.sin(phi), z1], axis=1)
    # the group walk's error on the same nodes, each walked inside its own 8 x 8 block of its
    # section, as the fill groups them
    kk, jj = np.meshgrid(np.arange(8), np.arange(8), indexing="ij")
    bk = (k // 8 * 8)[:, None] + kk.ravel()[None, :]
    bj = (j // 8 * 8)[:, None] + jj.ravel()[None, :]
    br = p.FG_Rmin * (p.FG_Fac ** bj - 1.0)
    bz = p.FG_Rmin * (p.FG_Fac ** bk - 1.0)
    bpos = np.stack([br * np.cos(phi)[:, None], br * np.sin(phi)[:, None], bz], axis=2).reshape(-1, 3)
    lane = np.arange(ncheck) * 64 + (k % 8) * 8 + (j % 8)
    out = dict(sources=n, targets=ntarget, fg_nbin=nbin, theta=0.4, softening=cfg.softening)
    for name, mode in (("quadrupole_plummer_fp32", TREE_MODE_TARGET), ("reference_kernel_fp32", TREE_FP32),
                       ("quadrupole_plummer_fp32_group", TREE_MODE_TARGET | TREE_GROUP)):
        t0 = time.perf_counter()
        tree_fill(p, spos, smass, precision=mode)
        wall = time.perf_counter() - t0
        tm = tree_last_timings()
        if mode & TREE_GROUP:
            acc32, pot32, _ = tree_forces(p, spos, smass, bpos, precision=mode)
            acc32, pot32 = acc32[lane], pot32[lane]
        else:
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


def chi2_leg(n, iters64, k, sigma, fresh, rep, tree_n, target_points):
    """Both methods on an n-particle realisation of the full model (same grids, tables and
    targets): the plain FP64 path for iters64 iterations of one fresh trial per particle, the
    batched FP32 path iteration by iteration until its Sall reaches the plain path's final
    Sall for every type (at most 3 iters64 iterations); and one iteration of both paths on
    identical trials (K=1), the precision of the FP32 path alone."""
    cfg, w = full_model(n, tree_n)
    perm = model.permutation(w, 1)
    g64, targets = setup(w, cfg, target_points)
    t0 = time.perf_counter()
    g64.calc_all_response_fields()
    hist64 = [{t: float(g64.fit(t)[7]) for t in w.types}]
    wall64 = 0.0
    for it in range(iters64):
        trials = [(perm[a:a + rep], fp64_trials(w, perm[a:a + rep], 5 + it, r))
                  for r, a in enumerate(range(0, w.n, rep))]
        g64.retabulate(TREE_MODE)
        t1 = time.perf_counter()
        for idx, v in trials:
            g64.rep(idx, v)
        wall64 += time.perf_counter() - t1
        hist64.append({t: float(g64.fit(t)[7]) for t in w.types})
        log(f"[bench] chi2 leg: fp64 iteration {it} done ({time.perf_counter() - t0:.0f}s)")
    g64.close()
    g32, _ = setup(w, cfg, target_points, targets)
    g32.fp32_calc_all_response_fields()
    g32.fp32_set_fresh_trials(fresh)
    hist32 = [{t: float(g32.fit(t)[7]) for t in w.types}]
    goal = hist64[-1]
    wall32, reach = 0.0, None
    for it in range(3 * iters64):
        t1 = time.perf_counter()
        g32.retabulate(TREE_MODE)
        g32.fp32_iterate(perm, k=k, rep_size=rep, sigma_frac=sigma, seed=1000 + it)
        wall32 += time.perf_counter() - t1
        hist32.append({t: float(g32.fit(t)[7]) for t in w.types})
        if all(hist32[-1][t] <= goal[t] for t in w.types):
            reach = it + 1
            break
    g32.close()
    log(f"[bench] chi2 leg: fp64 {hist64[-1]} after {iters64} iterations ({wall64:.1f}s), fp32 {hist32[-1]} after "
        f"{len(hist32) - 1} ({wall32:.1f}s)")
    # identical trials (K=1) over one iteration
    g64, _ = setup(w, cfg, target_points, targets)
    g64.calc_all_response_fields()
    g32, _ = setup(w, cfg, target_points, targets)
    g32.fp32_calc_all_response_fields()
    agree, drift = [], []
    for r, a in enumerate(range(0, w.n, rep)):
        idx = perm[a:a + rep]
        v = fp64_trials(w, idx, 5, r)
        a64 = g64.rep(idx, v)[0]
        a32 = g32.fp32_rep(idx, k=1, vtry=v)
        agree.append(round(float(np.mean((a32 > 0) == (a64 > 0))), 5))
        drift.append({t: float(g32.fit(t)[7] / g64.fit(t)[7] - 1) for t in w.types})
    same = {t: {"fp64": float(g64.fit(t)[7]), "fp32": float(g32.fit(t)[7])} for t in w.types}
    g64.close()
    g32.close()
    worst = max(abs(c["fp32"] - c["fp64"]) / abs(c["fp64"]) for c in same.values())
    final_rel = {t: (hist32[-1][t] / goal[t] - 1) for t in w.types}
    return dict(
        n_particles=w.n, rep=rep,
        plain_fp64=dict(iterations=iters64, seconds=wall64, history=hist64),
        batched_fp32=dict(iterations=len(hist32) - 1, seconds=wall32, history=hist32, k=k, fresh=fresh,
                          sigma_frac=sigma),
        time_to_plain_final_chi2=dict(batched_fp32_iterations=reach, batched_fp32_s=wall32 if reach else None,
                                      plain_fp64_s=wall64, speedup=(wall64 / wall32) if reach else None,
                                      batched_final_over_plain_final_minus_1=final_rel),
        identical_trials=dict(per_type=same, max_rel_diff=worst, decision_agreement=agree[::max(1, len(agree) // 10)],
                              rel_diff_per_rep=drift[::max(1, len(drift) // 10)], iterations=1))


def control_leg(n, rep, reps):
    """Sensitivity of the optimisation to a perturbation of the FP64 path itself, beside the FP32
    path's difference: the precision-form model (beta=0.3 halo + disk, 64^2 grids) with n
    particles; runs on identical K=1 trials for `reps` reps of `rep`: the FP64 path, the FP64
    path whose initial velocities differ by one ulp, the FP64 path whose initial velocities are
    moved by 2^-24 of their size (FP32's rounding, random sign per component), and the FP32 path.
    Reported per rep: decision agreement with the first run and Sall / Sall(first run) - 1 per
    type."""
    cfg = model.config("halo-disk-1e6")
    w = model.make_workload(cfg, n=n, target_samples=4000000)
    perm = model.permutation(w, 1)
    runs = {}
    sgn = np.where(np.random.default_rng(11).random(w.vel.shape) < 0.5, -1.0, 1.0)
    for name in ("fp64", "fp64_ulp", "fp64_eps32", "fp32"):
        g = GalicGPU(w)
        if name == "fp64_ulp":
            g.set_particles(w.pos, np.nextafter(w.vel, np.inf))
        if name == "fp64_eps32":
            g.set_particles(w.pos, w.vel * (1.0 + sgn * 2.0 ** -24))
        if name == "fp32":
            g.fp32_calc_all_response_fields()
        else:
            g.calc_all_response_fields()
        acc, sall = [], []
        for r in range(reps):
            idx = perm[(r * rep) % w.n:(r * rep) % w.n + rep]
            v = model.trial_velocities(w, idx, 5, r)
            a = g.fp32_rep(idx, k=1, vtry=v) if name == "fp32" else g.rep(idx, v)[0]
            acc.append(a > 0)
            sall.append({t: float(g.fit(t)[7]) for t in w.types})
        g.close()
        runs[name] = (acc, sall)
        log(f"[bench] control: {name} done")
    ref_acc, ref_sall = runs["fp64"]
    out = dict(n_particles=w.n, rep=rep, reps=reps)
    every = max(1, reps // 10)
    for name in ("fp64_ulp", "fp64_eps32", "fp32"):
        acc, sall = runs[name]
        agree = [float(np.mean(a == b)) for a, b in zip(acc, ref_acc)]
        rel = [{t: s[t] / r[t] - 1 for t in w.types} for s, r in zip(sall, ref_sall)]
        out[name] = dict(decision_agreement=agree[every - 1::every], sall_rel_diff=rel[every - 1::every],
                         final_sall_rel_diff=rel[-1],
                         max_abs_sall_rel_diff=max(abs(v) for d in rel for v in d.values()))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=12, help="timed steps (blocks of reps)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=20000000)
    ap.add_argument("--k", type=int, default=8)
    ap.add_argument("--sigma", type=float, default=0.1)
    ap.add_argument("--fresh", type=int, default=1,
                    help="of the K trials, how many are fresh draws from optimize()'s distribution")
    ap.add_argument("--rep", type=int, default=10000, help="trials per rep (both paths)")
    ap.add_argument("--steps-per-iteration", type=int, default=80,
                    help="steps that make one iteration over all particles (the table fill opens each)")
    ap.add_argument("--tree-n", type=int, default=10000000, help="sources of the table fill")
    ap.add_argument("--target-points", type=int, default=100000000, help="points per component for the targets")
    ap.add_argument("--fp64-reps", type=int, default=3, help="reps of the plain FP64 path timed (0: skip)")
    ap.add_argument("--fp64-n", type=int, default=2000000,
                    help="particles of the model realisation the plain FP64 reps run on (a rep's cost does not "
                         "depend on N: 1e4 trials against grids of the same size)")
    ap.add_argument("--chi2-n", type=int, default=100000, help="particles of the chi^2 comparison (0: skip)")
    ap.add_argument("--chi2-iters", type=int, default=2, help="plain FP64 iterations of the chi^2 comparison")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline budget in seconds (0: skip)")
    ap.add_argument("--cpu-n", type=int, default=3000, help="particles of the CPU baseline's realisation")
    ap.add_argument("--precision-sample", type=int, default=2000,
                    help="orbits integrated by both paths for the FP32 error figures (0: skip)")
    ap.add_argument("--c4-tree-n", type=int, default=10000000, help="sources of the table-fill leg (0: skip)")
    ap.add_argument("--overlap", type=int, default=-1, choices=(-1, 0, 1),
                    help="integrate the next group's orbits while a group is judged (1), never (0), "
                         "or when a step needs more than one orbit launch (-1)")
    ap.add_argument("--orbit-wg-per-cu", type=int, default=0, help="orbit kernel workgroups per CU (0: autotuned)")
    ap.add_argument("--merit-wg-per-cu", type=int, default=0, help="merit kernel workgroups per CU (0: as many as fit)")
    ap.add_argument("--tune-schedule", type=int, default=1, choices=(0, 1),
                    help="tune the CU sharing of overlapped orbit launches and judging in the first warm-up step")
    ap.add_argument("--graph", type=int, default=1, choices=(0, 1),
                    help="capture the reps of each orbit launch's group in one hipGraph (1) or launch them "
                         "rep by rep (0)")
    ap.add_argument("--control-n", type=int, default=0,
                    help="particles of the FP64 perturbation control on the precision-form model (0: skip)")
    ap.add_argument("--control-reps", type=int, default=100)
    ap.add_argument("--only-control", action="store_true", help="run the control leg alone and print its JSON")
    ap.add_argument("--only-tree", action="store_true", help="run the table-fill leg alone and print its JSON")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              "(the path shards by particles, see the design notes, but no multi-GPU run is made).",
              file=sys.stderr)
        sys.exit(2)

    if args.only_tree:
        print(json.dumps(tree_leg(args.c4_tree_n or 10000000, 512, 4096)), flush=True)
        return

    if args.only_control:
        print(json.dumps(control_leg(args.control_n or 1000000, args.rep, args.control_reps)), flush=True)
        return

    t0 = time.perf_counter()
    cfg, w = full_model(args.n, args.tree_n)
    log(f"[bench] workload {cfg.name} n={w.n} built in {time.perf_counter() - t0:.1f}s")
    perm = model.permutation(w, 1)
    spi = args.steps_per_iteration
    bounds = np.linspace(0, w.n, spi + 1).astype(np.int64)
    # a step covers whole reps
    bounds = (bounds // args.rep) * args.rep
    bounds[-1] = w.n

    # ---------------------------------------------------------- FP32 batched
    t0 = time.perf_counter()
    g, targets = setup(w, cfg, args.target_points)
    tm0 = tree_last_timings()
    log(f"[bench] targets and table fill ready in {time.perf_counter() - t0:.1f}s (tree walk {tm0['walk_ms']:.0f} ms)")
    t0 = time.perf_counter()
    g.fp32_calc_all_response_fields()
    init_s = time.perf_counter() - t0
    ti = g.timings()
    log(f"[bench] fp32 initial pass {init_s:.1f}s (orbit kernel {ti['orbit_ms']:.0f} ms)")
    g.fp32_set_fresh_trials(args.fresh)
    g.fp32_set_schedule(0, args.overlap, args.graph)
    variant, bpc = g.fp32_autotune(min(w.n, 524288))
    log(f"[bench] autotuned orbit kernel: {bpc} workgroups/CU")
    if args.orbit_wg_per_cu:
        g.fp32_set_launch(0, args.orbit_wg_per_cu)
        bpc = args.orbit_wg_per_cu
    g.fp32_set_merit_launch(args.merit_wg_per_cu)
    sched = None

    def step(s, seed):
        j = s % spi
        retab_ms = 0.0
        if j == 0:
            t1 = time.perf_counter()
            g.retabulate(TREE_MODE)
            retab_ms = (time.perf_counter() - t1) * 1e3
        idx = perm[bounds[j]:bounds[j + 1]]
        acc = g.fp32_iterate(idx, k=args.k, rep_size=args.rep, sigma_frac=args.sigma, seed=seed)
        return len(idx), int((acc > 0).sum()), retab_ms

    w0 = 0
    if args.tune_schedule and args.warmup > 0:
        # the warm-up steps (real steps of the optimisation, as any other, run as one call) also
        # pick the sharing of the CUs between the orbit launches and the judging they overlap
        w0 = min(args.warmup, spi)
        idx = perm[bounds[0]:bounds[w0]]
        g.retabulate(TREE_MODE)
        _, ob, mb = g.fp32_autotune_schedule(idx, k=args.k, rep_size=args.rep, sigma_frac=args.sigma, seed=500)
        sched = {"orbit_wg_per_cu_beside_judging": ob, "merit_wg_per_cu": mb or "all"}
        log(f"[bench] schedule tuned over {w0} warm-up steps: {sched}")
    for s in range(w0, args.warmup):
        step(s, 500 + s)
        log(f"[bench] warmup {s} done")
    g.reset_timings()
    hist = [{t: float(g.fit(t)[7]) for t in w.types}]
    t0 = time.perf_counter()
    nparticles = accepted = 0
    retab_ms = 0.0
    for s in range(args.steps):
        n1, a1, r1 = step(s, 1000 + s)
        nparticles += n1
        accepted += a1
        retab_ms += r1
        log(f"[bench] step {s} done ({n1} particles, {time.perf_counter() - t0:.1f}s)")
    wall = time.perf_counter() - t0
    hist.append({t: float(g.fit(t)[7]) for t in w.types})
    tm = g.timings()
    orbits = nparticles * (args.k + 1)
    value = orbits / wall
    steps_total = tm["steps"]
    achieved = F_STEP * steps_total / (tm["orbit_ms"] * 1e-3) / 1e12
    # the orbit launches of the next group run beside the judging of a group: the parts can add
    # up to more than the wall time, by what the overlap hid
    rest = (wall * 1e3 - tm["orbit_ms"] - tm["eval_ms"] - tm["fit_ms"] - retab_ms) / args.steps
    breakdown = dict(orbit_kernel=tm["orbit_ms"] / args.steps, merit_kernel=tm["eval_ms"] / args.steps,
                     fit=tm["fit_ms"] / args.steps, table_fill=retab_ms / args.steps,
                     other=max(rest, 0.0), hidden_by_overlap=max(-rest, 0.0))
    log(f"[bench] fp32: {value:.4g} orbits/s, {steps_total / wall:.4g} steps/s, orbit kernel {tm['orbit_ms']:.1f} ms "
        f"of {wall * 1e3:.1f} ms, {achieved:.2f} TFLOP/s algorithmic; ms per step {breakdown}, "
        f"log relaunches {tm['log_relaunches']:.0f}")

    # ------------------------------------------- FP32 orbits against FP64
    precision = None
    if args.precision_sample > 0:
        sel = np.random.default_rng(7).choice(w.n, args.precision_sample, replace=False)
        errs, d32, d64, dsteps = [], [], [], []
        for a in range(0, len(sel), 250):
            s_ = sel[a:a + 250]
            orb = (w.pos[s_], w.vel[s_], w.mass[s_], w.tint[s_], w.ptype[s_])
            r64 = g.orbit_responses(*orb)
            r32 = g.orbit_responses(*orb, precision=32)
            errs.append(np.abs(r32["fields"] - r64["fields"]).sum(axis=(1, 2)) / np.abs(r64["fields"]).sum(axis=(1, 2)))
            d32.append(np.abs(r32["e1"] / r32["e0"] - 1))
            d64.append(np.abs(r64["e1"] / r64["e0"] - 1))
            dsteps.append(np.abs(r32["steps"] - r64["steps"]) / r64["steps"])
        errs, d32, d64, dsteps = map(np.concatenate, (errs, d32, d64, dsteps))
        precision = dict(orbits=len(sel), response_l1_rel_err=dict(median=float(np.median(errs)),
                                                                  p99=float(np.quantile(errs, 0.99))),
                         energy_drift=dict(fp32_median=float(np.median(d32)), fp64_median=float(np.median(d64)),
                                           fp32_p99=float(np.quantile(d32, 0.99)),
                                           fp64_p99=float(np.quantile(d64, 0.99))),
                         step_count_rel_diff_median=float(np.median(dsteps)))
        log(f"[bench] fp32 vs fp64 orbits: {precision}")
    g.close()

    # ------------------------------------------------------ FP64 plain path
    n_total = w.n
    split = [int(w.params.NType[t]) for t in (1, 2, 3)]
    fp64 = None
    if args.fp64_reps > 0:
        if args.fp64_n != w.n:
            del w
            cfg, w = full_model(args.fp64_n, args.tree_n)
            perm = model.permutation(w, 1)
            targets = None
        g64, _ = setup(w, cfg, args.target_points, targets)
        t0 = time.perf_counter()
        g64.calc_all_response_fields()
        init64 = time.perf_counter() - t0
        g64.reset_timings()
        wall64 = 0.0
        ntr = 0
        for r in range(args.fp64_reps):
            idx = perm[r * args.rep:(r + 1) * args.rep]
            v = fp64_trials(w, idx, 5, r)
            t1 = time.perf_counter()
            g64.rep(idx, v)
            wall64 += time.perf_counter() - t1
            ntr += len(idx)
        tm64 = g64.timings()
        g64.close()
        fp64 = dict(orbits_per_s=2 * ntr / wall64, ms_per_rep=wall64 / args.fp64_reps * 1e3, reps=args.fp64_reps,
                    steps_per_s=tm64["steps"] / wall64, orbit_kernel_ms=tm64["orbit_ms"], initial_pass_s=init64,
                    n_particles=w.n, method="one fresh trial per particle (main.c:383-435), reps of 1e4")
        log(f"[bench] fp64: {fp64['orbits_per_s']:.4g} orbits/s over {args.fp64_reps} reps, initial pass {init64:.1f}s")
    del w

    # --------------------------------------------- chi^2 of both methods
    chi2 = None
    if args.chi2_n > 0:
        chi2 = chi2_leg(args.chi2_n, args.chi2_iters, args.k, args.sigma, args.fresh, args.rep,
                        min(args.tree_n, 2000000), args.target_points // 10)
        log(f"[bench] chi2: {chi2['time_to_plain_final_chi2']}")

    control = control_leg(args.control_n, args.rep, args.control_reps) if args.control_n > 0 else None

    # ------------------------------------------------------------- CPU
    cpu = None
    if args.cpu_seconds > 0:
        nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        rate, dt, ntr, tinit = cpu_baseline(args.cpu_seconds, nthreads, args.cpu_n)
        cpu = {"value": rate, "unit": "orbits/s", "cores": nthreads, "kind": "port",
               "sample": f"{ntr} trials in reps of 200 (2 orbits + 2 merit evaluations each, the reference's "
                         f"optimize()) on a {args.cpu_n}-particle realisation of the full model (512^2 grids and "
                         f"tables), {nthreads} OpenMP threads, {dt:.1f} s (initial pass {tinit:.1f} s not counted)"}
        log(f"[bench] cpu: {rate:.4g} orbits/s on {nthreads} threads ({ntr} trials, {dt:.1f} s)")

    # ------------------------------------------------------ table fill
    tree = tree_leg(args.c4_tree_n, 512, 4096) if args.c4_tree_n > 0 else None

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
        "config": {"workload": "full v1.1 model: isotropic Hernquist halo + structure-3 exponential sech^2 disk + "
                               "isotropic Hernquist bulge, tables re-filled by the GPU tree every iteration",
                   "n_particles": n_total, "type_split": split,
                   "force_table": [512, 512], "response_grid": [512, 512],
                   "tree_sources": args.tree_n, "target_points_per_component": args.target_points,
                   "trials_per_particle": args.k, "fresh_trials": args.fresh, "trial_sigma_frac": args.sigma,
                   "rep": args.rep, "particles_per_step": int(nparticles / args.steps),
                   "steps_per_iteration": spi, "orbits_per_step": int(orbits / args.steps),
                   "torbit_fac": cfg.torbit_fac, "autotuned": {"workgroups_per_cu": bpc, "schedule": sched},
                   "overlap": args.overlap, "graph": args.graph,
                   "graph_launches_per_step": tm["graph_launches"] / args.steps},
        "roofline": {"bound": "valu_fp32", "achieved": achieved, "peak": PEAK_FP32, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32,
                     "traffic": HBM_BYTES_PER_STEP * steps_total / max(1, tm["orbit_launches"]),
                     "kernel": "k_orbit32_log",
                     "note": f"roof: the FP32 vector ALU (no step is a contraction, MFMA unused); achieved = {F_STEP} "
                             "algorithmic flop/step x leapfrog steps / HIP-event time of the kernel; traffic = HBM "
                             f"bytes per launch ({HBM_BYTES_PER_STEP} B/step from calibrated FETCH_SIZE+WRITE_SIZE "
                             "of the profile: the response-log writes and the table lookups that miss L2)"},
        "steps_per_s": steps_total / wall,
        "timed_orbit_launches": int(tm["orbit_launches"]), "timed_leapfrog_steps": int(steps_total),
        "ms_per_step_breakdown": breakdown,
        "fp32_initial_pass_s": init_s,
        "accept_frac": accepted / max(1, nparticles),
        "sall_before_after_timed_steps": hist,
        "fp64_plain": fp64,
        "fp32_vs_fp64_speedup": (value / fp64["orbits_per_s"]) if fp64 else None,
        "chi2_check": chi2,
        "perturbation_control": control,
        "fp32_orbit_precision": precision,
        "cpu_baseline": cpu,
        "gpu_over_cpu": (value / cpu["value"]) if cpu else None,
        "tree_fill": tree,
        "copy_probe_GBs": copy_gbs,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
