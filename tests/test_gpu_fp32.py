"""The batched FP32 path: FP32 orbits, fixed-point grids, FP64 merit.

Its bar is the FP64 path, not the oracle bit for bit:
  * response grids of the initial pass within 1e-3 (normwise) of the FP64 path;
  * every fixed-point grid bit-identical between two runs (integer atomics);
  * with the same K=1 trials the accept decisions agree except for trials whose
    merit change is within the FP32 noise, and the Sall history stays within
    1e-3 relative of the FP64 path.
"""
import numpy as np
import pytest

from galic_amd import model
from galic_amd.runtime import GalicGPU

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def w():
    cfg = model.config("halo-disk-1e6", max_steps=2000000)
    return model.make_workload(cfg, n=6000, target_samples=200000)


def normwise(a, b):
    return np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300)


def test_fp32_initial_pass_close_to_fp64(w):
    print("[gpu] fp32 initial pass", flush=True)
    g64 = GalicGPU(w)
    g64.calc_all_response_fields()
    g32 = GalicGPU(w)
    g32.fp32_calc_all_response_fields()
    for t in w.types:
        for which in range(10):
            assert normwise(g32.stack(t, which), g64.stack(t, which)) < 1e-3, (t, which)
        np.testing.assert_allclose(g32.fit(t)[:8], g64.fit(t)[:8], rtol=2e-3)
    g64.close()
    g32.close()


def test_fp32_orbits_against_fp64(w):
    """Per-orbit error of the FP32 path against the FP64 path (the oracle's
    double): the integration time t equals Tint up to the FP32 rounding of the
    last, clipped step (time is summed in FP64), step counts agree
    to 1%, response fields agree normwise to 1e-3 for the median orbit, and the
    energy drift stays of the order of the FP64 path's own."""
    print("[gpu] fp32 orbits vs fp64", flush=True)
    g = GalicGPU(w)
    sel = np.arange(0, w.n, 3)[:1500]
    args = (w.pos[sel], w.vel[sel], w.mass[sel], w.tint[sel], w.ptype[sel])
    r64 = g.orbit_responses(*args)
    r32 = g.orbit_responses(*args, precision=32)
    g.close()
    np.testing.assert_allclose(r32["t"], w.tint[sel], rtol=1e-8)
    assert np.all(np.abs(r32["steps"] - r64["steps"]) <= 0.01 * r64["steps"] + 2)
    f32, f64 = r32["fields"], r64["fields"]
    err = np.abs(f32 - f64).sum(axis=(1, 2)) / np.abs(f64).sum(axis=(1, 2))
    drift32 = np.abs(r32["e1"] / r32["e0"] - 1)
    drift64 = np.abs(r64["e1"] / r64["e0"] - 1)
    print(f"response L1 error: median {np.median(err):.2e} p99 {np.percentile(err, 99):.2e}; energy drift median "
          f"fp32 {np.median(drift32):.2e} fp64 {np.median(drift64):.2e}")
    assert np.median(err) < 1e-3 and np.percentile(err, 99) < 5e-2
    assert np.median(drift32) < 3 * np.median(drift64) + 1e-6


def test_fp32_grids_are_bit_deterministic(w):
    print("[gpu] fp32 determinism", flush=True)
    perm = model.permutation(w, 1)
    runs = []
    for _ in range(2):
        g = GalicGPU(w)
        g.fp32_calc_all_response_fields()
        g.fp32_rep(perm[: w.n // 2], k=4, sigma_frac=0.1, seed=11)
        runs.append([g.stack(t, wh) for t in w.types for wh in range(10)] + [g.particles()[0]])
        g.close()
    for a, b in zip(*runs):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("launch", [(0, 1), (0, 2), (0, 8)])
def test_fp32_launch_variants_give_identical_bits(w, launch):
    """The persistent grid size changes the schedule only: grids, velocities and
    accept flags are bit-identical to the reference launch (4 workgroups per CU)."""
    print(f"[gpu] fp32 launch variant {launch}", flush=True)
    perm = model.permutation(w, 1)
    runs = []
    for variant, bpc in [(0, 4), launch]:
        g = GalicGPU(w)
        g.fp32_calc_all_response_fields()
        g.fp32_set_launch(variant, bpc)
        acc = g.fp32_rep(perm[: w.n // 2], k=4, sigma_frac=0.1, seed=11)
        runs.append([acc] + [g.stack(t, wh) for t in w.types for wh in range(10)] + [g.particles()[0]])
        g.close()
    for a, b in zip(*runs):
        np.testing.assert_array_equal(a, b)


def test_fp32_iterate_equals_rep_by_rep(w):
    """galic_fp32_iterate integrates the orbits of several reps in one launch and
    then evaluates and commits the reps in turn: bit-identical to one
    galic_fp32_rep per rep with the same seed."""
    print("[gpu] fp32 iterate vs rep by rep", flush=True)
    perm = model.permutation(w, 2)
    rep = 700
    out = []
    for mode in ("iterate", "reps"):
        g = GalicGPU(w)
        g.fp32_calc_all_response_fields()
        if mode == "iterate":
            acc = g.fp32_iterate(perm, k=3, rep_size=rep, sigma_frac=0.1, seed=5)
        else:
            acc = np.concatenate([g.fp32_rep(perm[a:a + rep], k=3, sigma_frac=0.1, seed=5)
                                  for a in range(0, len(perm), rep)])
        out.append([acc] + [g.stack(t, wh) for t in w.types for wh in range(10)] + [g.particles()[0]]
                   + [g.fit(t) for t in w.types])
        g.close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("group,overlap,graph", [(700, 1, 1), (1000, 1, 1), (1000, 0, 1), (0, 0, 1), (1000, 1, 0)])
def test_fp32_schedule_changes_nothing(w, group, overlap, graph):
    """Orbit launches of a few reps each, integrated on a second stream while
    the previous group is judged (overlap 1) or one after the other (0), with
    each group's reps captured in one hipGraph (graph 1) or not, give the same
    bits as one launch for the whole iteration with its reps launched one by one."""
    print(f"[gpu] fp32 schedule group={group} overlap={overlap} graph={graph}", flush=True)
    perm = model.permutation(w, 2)
    out = []
    for sched in ((0, 0, 0), (group, overlap, graph)):
        g = GalicGPU(w)
        g.fp32_calc_all_response_fields()
        g.fp32_set_schedule(*sched)
        acc = g.fp32_iterate(perm, k=3, rep_size=700, sigma_frac=0.1, seed=5)
        out.append([acc] + [g.stack(t, wh) for t in w.types for wh in range(10)] + [g.particles()[0]]
                   + [g.fit(t) for t in w.types])
        launches = g.timings()["graph_launches"]
        assert launches == 0 if sched[2] == 0 else launches >= 1
        g.close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)


def test_fp32_schedule_tuning_changes_nothing(w):
    """The schedule tuner runs the iteration's reps in three runs, each with
    another CU sharing, and gives the bits of one plain iteration."""
    print("[gpu] fp32 schedule tuning", flush=True)
    perm = model.permutation(w, 2)
    out = []
    for tune in (False, True):
        g = GalicGPU(w)
        g.fp32_calc_all_response_fields()
        if tune:
            g.fp32_autotune(2000)
            g.fp32_set_schedule(1000, 1, 1)
            acc, ob, mb = g.fp32_autotune_schedule(perm, k=3, rep_size=700, sigma_frac=0.1, seed=5)
            assert ob >= 1 and mb in (0, 1)
        else:
            acc = g.fp32_iterate(perm, k=3, rep_size=700, sigma_frac=0.1, seed=5)
        out.append([acc] + [g.stack(t, wh) for t in w.types for wh in range(10)] + [g.particles()[0]]
                   + [g.fit(t) for t in w.types])
        g.close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)


def test_fp32_history_tracks_fp64(w):
    """Same trials (K=1) on both paths for several reps."""
    print("[gpu] fp32 vs fp64 history", flush=True)
    g64 = GalicGPU(w)
    g64.calc_all_response_fields()
    g32 = GalicGPU(w)
    g32.fp32_calc_all_response_fields()
    perm = model.permutation(w, 1)
    ntr = w.n // 3
    agree = []
    for rep in range(6):
        idx = perm[(rep % 3) * ntr:((rep % 3) + 1) * ntr]
        # trials drawn from the FP64 path's current velocities
        v64, _, _ = g64.particles()
        w2 = w
        v = model.trial_velocities(w2, idx, 5, rep)
        a64, dF = g64.rep(idx, v)
        a32 = g32.fp32_rep(idx, k=1, vtry=v)
        agree.append(np.mean((a32 > 0) == (a64 > 0)))
    print("agreement per rep", agree)
    assert min(agree) > 0.97
    for t in w.types:
        s64 = g64.fit(t)[7]
        s32 = g32.fit(t)[7]
        assert abs(s32 - s64) / s64 < 1e-3, (t, s32, s64)
    g64.close()
    g32.close()


def test_device_trial_draws_match_host_philox(w):
    """Trials drawn on the device (k_draw_trials) equal the same Philox draws on
    the host: every accepted velocity is v + sigma_frac * (g0 s_r e_r + g1 s_t e_t
    + g2 s_p e_p) for a perturbation trial k, and v_str e_p + g0 s_r e_r + g1 s_t
    e_t + g2 s_p e_p for the fresh trial (basis of main.c:383-405)."""
    from galic_amd.rng import gaussians
    print("[gpu] device trial draws", flush=True)
    g = GalicGPU(w)
    g.fp32_calc_all_response_fields()
    g.fp32_set_fresh_trials(1)
    v0 = g.particles()[0].copy()
    idx = model.permutation(w, 1)[:3000]
    K, sig, seed = 4, 0.1, 77
    acc = g.fp32_rep(idx, k=K, sigma_frac=sig, seed=seed)
    v1 = g.particles()[0]
    g.close()
    sel = np.nonzero(acc > 0)[0]
    assert len(sel) > 100
    p = w.params
    kinds = set()
    for i in sel[:1500]:
        pi = idx[i]
        k = acc[i] - 1
        x = w.pos[pi]
        tovs = p.TypeOfVelocityStructure[w.ptype[pi]]
        R, T, Q = model.local_basis(x[None, :], tovs == 2)
        # trial 0 is a fresh draw around (0, 0, v_str) with the full dispersions
        # (main.c:383-435), the others perturb the current velocity by sig * sigma
        s = (1.0 if k == 0 else sig) * np.sqrt(w.disp[pi, :3])
        centre = w.vstr[pi] * Q[0] if k == 0 else v0[pi]
        lim = p.MaxVelInUnitsVesc * w.vesc[pi]
        for attempt in range(64):
            gg = gaussians(seed, 3000, np.array([pi], dtype=np.uint64), attempt, k)[:, 0]
            t = centre + gg[0] * s[0] * R[0] + gg[1] * s[1] * T[0] + gg[2] * s[2] * Q[0]
            if np.linalg.norm(t) < lim:
                break
        np.testing.assert_allclose(v1[pi], t, rtol=1e-10, atol=1e-9)
        kinds.add(k == 0)
    assert kinds == {True, False}


def test_fp32_level7_grids():
    """128 x 128 response grids (the level of the reference's parameter files):
    the merit kernel sums both logs into one LDS grid.  Reps in one launch equal
    rep-by-rep bit for bit, and decisions follow the FP64 path."""
    print("[gpu] fp32 level-7 grids", flush=True)
    cfg = model.config("halo-disk-1e6", level=7, max_steps=2000000)
    w7 = model.make_workload(cfg, n=3000, target_samples=200000)
    perm = model.permutation(w7, 4)
    out = []
    for mode in ("iterate", "reps"):
        g = GalicGPU(w7)
        g.fp32_calc_all_response_fields()
        if mode == "iterate":
            acc = g.fp32_iterate(perm, k=3, rep_size=500, sigma_frac=0.1, seed=9)
        else:
            acc = np.concatenate([g.fp32_rep(perm[a:a + 500], k=3, sigma_frac=0.1, seed=9)
                                  for a in range(0, len(perm), 500)])
        out.append([acc] + [g.stack(t, wh) for t in w7.types for wh in range(10)] + [g.particles()[0]])
        g.close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)
    g64 = GalicGPU(w7)
    g64.calc_all_response_fields()
    g32 = GalicGPU(w7)
    g32.fp32_calc_all_response_fields()
    idx = perm[:1000]
    v = model.trial_velocities(w7, idx, 5, 0)
    a64 = g64.rep(idx, v)[0]
    a32 = g32.fp32_rep(idx, k=1, vtry=v)
    g64.close()
    g32.close()
    assert np.mean((a32 > 0) == (a64 > 0)) > 0.97


@pytest.mark.parametrize("level", [8, 9])
def test_fp32_grids_above_level7(level):
    """256^2 and 512^2 response grids: the merit kernel keeps its two grids in HBM
    scratch.  Reps in one launch equal rep-by-rep bit for bit, and K=1 decisions
    follow the FP64 path (itself checked against the oracle at level 9)."""
    print(f"[gpu] fp32 level-{level} grids", flush=True)
    cfg = model.config("halo-disk-1e6", level=level, max_steps=2000000)
    wl = model.make_workload(cfg, n=2000, target_samples=200000)
    perm = model.permutation(wl, 4)
    out = []
    for mode in ("iterate", "reps"):
        g = GalicGPU(wl)
        g.fp32_calc_all_response_fields()
        if mode == "iterate":
            acc = g.fp32_iterate(perm, k=3, rep_size=500, sigma_frac=0.1, seed=9)
        else:
            acc = np.concatenate([g.fp32_rep(perm[a:a + 500], k=3, sigma_frac=0.1, seed=9)
                                  for a in range(0, len(perm), 500)])
        out.append([acc] + [g.stack(t, wh) for t in wl.types for wh in range(10)] + [g.particles()[0]])
        g.close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)
    assert (out[0][0] > 0).mean() > 0.2
    g64 = GalicGPU(wl)
    g64.calc_all_response_fields()
    g32 = GalicGPU(wl)
    g32.fp32_calc_all_response_fields()
    idx = perm[:800]
    v = model.trial_velocities(wl, idx, 5, 0)
    a64 = g64.rep(idx, v)[0]
    a32 = g32.fp32_rep(idx, k=1, vtry=v)
    s64 = [g64.fit(t)[7] for t in wl.types]
    s32 = [g32.fit(t)[7] for t in wl.types]
    g64.close()
    g32.close()
    assert np.mean((a32 > 0) == (a64 > 0)) > 0.97
    np.testing.assert_allclose(s32, s64, rtol=2e-3)


def composite_workload(n=3000, **kw):
    cfg = model.config("composite-4e6", max_steps=2000000, **kw)
    return model.make_workload(cfg, n=n, target_samples=200000)


def k_trials(wl, idx, K, seed):
    """K trial velocities per particle: one fresh draw of optimize()'s distribution
    (main.c:383-435) and K-1 Gaussian perturbations of 10% of the target dispersions
    in the local basis, all below MaxVelInUnitsVesc * Vesc; shape (n, K, 3)."""
    from galic_amd.rng import gaussians
    v = np.zeros((len(idx), K, 3))
    v[:, 0] = model.trial_velocities(wl, idx, seed, 0)
    p = wl.params
    tovs = np.array([p.TypeOfVelocityStructure[t] for t in range(4)])[wl.ptype[idx]]
    R, T, Q = model.local_basis(wl.pos[idx], False)
    Rc, Tc, Qc = model.local_basis(wl.pos[idx], True)
    cyl = (tovs == 2)[:, None]
    R, T, Q = np.where(cyl, Rc, R), np.where(cyl, Tc, T), np.where(cyl, Qc, Q)
    s = 0.1 * np.sqrt(wl.disp[idx, :3])
    lim = p.MaxVelInUnitsVesc * wl.vesc[idx]
    for k in range(1, K):
        g = gaussians(seed, 500 + k, idx.astype(np.uint64), 0)
        t = wl.vel[idx] + g[0][:, None] * s[:, :1] * R + g[1][:, None] * s[:, 1:2] * T + g[2][:, None] * s[:, 2:3] * Q
        ok = np.linalg.norm(t, axis=1) < lim
        v[:, k] = np.where(ok[:, None], t, wl.vel[idx])
    return v


@pytest.mark.parametrize("form", [2, 4])
def test_fp32_best_of_k_equals_k_single_trials(form):
    """The fused K-trial launch (K=8) evaluates each trial as a launch of that trial alone
    would, and the particle takes the trial of the smallest negative merit change
    (main.c:464-465 applied to the best of K).  Under the int64 global fixed point (form 2) the
    per-trial merit changes are bit-identical to K=1 launches from the same state; under form 4
    a particle's fixed point follows the largest total of the orbits evaluated with it, so they
    agree to that rounding (median 1e-6, at most 1e-4 of the particle's largest |change|: a
    record's value is rounded to 2^-31 of the largest total, and a merit change is a
    difference of sums over thousands of nodes)."""
    print(f"[gpu] fp32 best of K=8 vs single trials, merit form {form}", flush=True)
    wl = composite_workload()
    idx = model.permutation(wl, 3)[:1000]
    K = 8
    V = k_trials(wl, idx, K, 21)
    g = GalicGPU(wl)
    g.fp32_calc_all_response_fields()
    g.fp32_set_merit_form(form)
    acc = g.fp32_rep(idx, k=K, vtry=V.reshape(-1, 3))
    dfk = g.fp32_trial_merits(len(idx), K)
    g.close()
    best = np.where(dfk.min(axis=1) < 0, dfk.argmin(axis=1) + 1, 0)
    np.testing.assert_array_equal(acc, best)
    assert (acc > 0).mean() > 0.2 and len(set(acc[acc > 0])) == K
    scale = np.abs(dfk).max(axis=1)
    for k in (0, 3, 7):
        g1 = GalicGPU(wl)
        g1.fp32_calc_all_response_fields()
        g1.fp32_set_merit_form(form)
        g1.fp32_rep(idx, k=1, vtry=np.ascontiguousarray(V[:, k]))
        d1 = g1.fp32_trial_merits(len(idx), 1)[:, 0]
        g1.close()
        if form == 2:
            np.testing.assert_array_equal(d1, dfk[:, k])
        else:
            err = np.abs(d1 - dfk[:, k]) / scale
            print(f"trial {k}: largest difference {err.max():.2e}, median {np.median(err):.2e} of the "
                  f"particle's largest |change|")
            assert err.max() < 1e-4 and np.median(err) < 1e-6


def test_fp32_best_of_k_matches_oracle(oracle_mod):
    """Every trial of a K=8 launch against the oracle's merit change of that trial alone
    (eval_trials, main.c:365-527, FP64 orbits on the oracle's own state): the FP32 path's
    per-trial merit changes follow the oracle's to the FP32 orbit error, and the chosen trial
    is the oracle's best wherever the oracle's best and second best differ by more than that
    noise."""
    print("[gpu] fp32 best of K=8 vs oracle", flush=True)
    wl = composite_workload()
    idx = model.permutation(wl, 3)[:600]
    K = 8
    V = k_trials(wl, idx, K, 22)
    g = GalicGPU(wl)
    g.fp32_calc_all_response_fields()
    acc = g.fp32_rep(idx, k=K, vtry=V.reshape(-1, 3))
    dfk = g.fp32_trial_merits(len(idx), K)
    g.close()
    o = oracle_mod.Oracle(wl)
    o.calc_all_response_fields(8)
    dfo = np.zeros_like(dfk)
    for k in range(K):
        fb, ft = o.eval_trials(idx, np.ascontiguousarray(V[:, k]), 8)
        dfo[:, k] = ft - fb
    o.close()
    scale = np.abs(dfo).max(axis=1, keepdims=True)
    rel = np.abs(dfk - dfo) / np.maximum(np.abs(dfo), 1e-3 * scale)
    print(f"per-trial merit change vs oracle: median rel {np.median(rel):.2e}, p90 {np.quantile(rel, 0.9):.2e}")
    assert np.median(rel) < 2e-2
    best_o = np.where(dfo.min(axis=1) < 0, dfo.argmin(axis=1) + 1, 0)
    srt = np.sort(dfo, axis=1)
    margin = np.where(srt[:, 0] < 0, (srt[:, 1] - srt[:, 0]) / np.abs(srt[:, 0]), np.abs(srt[:, 0]) / scale[:, 0])
    clear = margin > 0.2
    agree = (acc == best_o)[clear].mean()
    print(f"chosen trial equals the oracle's on {agree:.4f} of {clear.sum()} clear cases; "
          f"all cases {(acc == best_o).mean():.4f}")
    assert clear.sum() > 200 and agree > 0.97


def run_forms(wl, forms, K, n1, rep, seed):
    """One rep of n1 particles, then reps of `rep` over the rest, under each merit form."""
    perm = model.permutation(wl, 6)
    out = []
    for form in forms:
        g = GalicGPU(wl)
        g.fp32_calc_all_response_fields()
        g.fp32_set_merit_form(form)
        acc = g.fp32_rep(perm[:n1], k=K, sigma_frac=0.1, seed=seed)
        dfk = g.fp32_trial_merits(n1, K)
        acc2 = g.fp32_iterate(perm[n1:], k=K, rep_size=rep, sigma_frac=0.1, seed=seed + 1)
        out.append(dict(acc=np.concatenate([acc, acc2]), dfk=dfk,
                        grids=[g.stack(t, wh) for t in wl.types for wh in range(10)],
                        vel=g.particles()[0], fit=[g.fit(t) for t in wl.types]))
        g.close()
    return out


def assert_same_run(a, b):
    np.testing.assert_array_equal(a["acc"], b["acc"])
    for x, y in zip(a["grids"], b["grids"]):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(a["vel"], b["vel"])
    scale = np.abs(b["dfk"]).max()
    np.testing.assert_allclose(a["dfk"], b["dfk"], rtol=1e-11, atol=1e-13 * scale)
    assert (a["acc"] > 0).mean() > 0.2


@pytest.mark.parametrize("level,form", [(6, 1), (7, 1), (6, 3), (7, 3)])
def test_fp32_tile_merit_equals_lds_merit(level, form):
    """The tile form of the merit (the records of each log listed by 32 x 32-cell tile, each
    tile summed in LDS) against the whole-grid LDS kernels on 64^2 and 128^2 grids: the same
    decisions, bit-identical grids and velocities, per-trial merit changes equal up to the
    order of the FP64 sums."""
    print(f"[gpu] fp32 tile merit (form {form}) vs LDS merit, level {level}", flush=True)
    wl = composite_workload(n=2500, level=level)
    tiles, lds = run_forms(wl, (form, 2), K=6, n1=1000, rep=500, seed=13)
    assert_same_run(tiles, lds)


@pytest.mark.parametrize("level,ref", [(6, 2), (7, 2), (9, 1)])
def test_fp32_t16_merit_matches_int64_forms(level, ref):
    """Form 4 (16 x 16 tiles in each particle's own int32 fixed point) against a form of the
    int64 global fixed point (the whole-grid LDS form on 64^2 and 128^2 grids, the 32 x 32 tile
    form on the full model's 512^2 grids) on one rep from the same state: per-trial merit
    changes equal to the rounding of the fixed points (median below 1e-6, at most 1e-4 of the
    particle's largest |change|), the same choice wherever the best trial is separated from the
    runner-up and from zero by more than 1e-4, and -- the deltas being the same int64 sums --
    bit-identical grids and velocities when every choice agrees."""
    print(f"[gpu] fp32 16x16 tile merit vs form {ref}, level {level}", flush=True)
    if level == 9:
        wl = model.make_workload(model.config("composite-2e7", max_steps=2000000), n=1500, target_samples=200000)
        K, n1 = 4, 600
    else:
        wl = composite_workload(n=2500, level=level)
        K, n1 = 6, 1000
    perm = model.permutation(wl, 6)
    out = []
    for form in (4, ref):
        g = GalicGPU(wl)
        g.fp32_calc_all_response_fields()
        g.fp32_set_merit_form(form)
        acc = g.fp32_rep(perm[:n1], k=K, sigma_frac=0.1, seed=13)
        out.append(dict(acc=acc, dfk=g.fp32_trial_merits(n1, K),
                        grids=[g.stack(t, wh) for t in wl.types for wh in range(10)], vel=g.particles()[0]))
        g.close()
    a, b = out
    scale = np.abs(b["dfk"]).max(axis=1, keepdims=True)
    err = np.abs(a["dfk"] - b["dfk"]) / scale
    print(f"per-trial merit change: largest difference {err.max():.2e}, median {np.median(err):.2e} "
          f"of the particle's largest |change|")
    assert err.max() < 1e-4 and np.median(err) < 1e-6
    d = b["dfk"]
    srt = np.sort(d, axis=1)
    margin = np.minimum(np.abs(srt[:, 0]), srt[:, 1] - srt[:, 0]) / scale[:, 0]
    clear = margin > 1e-4
    print(f"choices equal on {(a['acc'] == b['acc']).mean():.4f}; clear cases {clear.mean():.4f}")
    assert clear.mean() > 0.98
    np.testing.assert_array_equal(a["acc"][clear], b["acc"][clear])
    assert (a["acc"] > 0).mean() > 0.2
    if (a["acc"] == b["acc"]).all():
        for x, y in zip(a["grids"], b["grids"]):
            np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(a["vel"], b["vel"])


@pytest.mark.parametrize("level,form", [(7, 2), (7, 4), (9, 4)])
def test_fp32_loc_grids_equal_fresh_initial_pass(level, form):
    """The FP32 path's orbit-averaged Loc grids are always the sum over the particles of their
    current orbits' logged records, each record rounded into the fixed point on its own: after
    reps of taken trials (each taken orbit added and the orbit it replaces taken away, through
    the region commit), an initial pass from the particles' current velocities on a fresh
    context gives the same bits (main.c:503-513 keeps Loc = the sum of the current responses)."""
    print(f"[gpu] fp32 Loc grids after reps vs a fresh initial pass, level {level}, merit form {form}", flush=True)
    if level == 9:
        wl = model.make_workload(model.config("composite-2e7", max_steps=2000000), n=1500, target_samples=200000)
        K, rep = 4, 300
    else:
        wl = composite_workload(n=2500, level=level)
        K, rep = 6, 500
    perm = model.permutation(wl, 6)
    g = GalicGPU(wl)
    g.fp32_calc_all_response_fields()
    g.fp32_set_merit_form(form)
    acc = g.fp32_iterate(perm, k=K, rep_size=rep, sigma_frac=0.1, seed=31)
    vel = g.particles()[0]
    grids = [g.stack(t, wh) for t in wl.types for wh in (0, 6, 7, 8, 9)]
    g.close()
    assert (acc > 0).mean() > 0.2
    assert not np.array_equal(vel, wl.vel)
    g2 = GalicGPU(wl)
    g2.set_particles(wl.pos, vel)
    g2.fp32_calc_all_response_fields()
    fresh = [g2.stack(t, wh) for t in wl.types for wh in (0, 6, 7, 8, 9)]
    g2.close()
    for x, y in zip(grids, fresh):
        np.testing.assert_array_equal(x, y)


def test_fp32_tile_merit_equals_hbm_form_level9():
    """On 512^2 grids (the full v1.1 model's form: isotropic halo and bulge, structure-3 disk,
    512^2 force tables) the tile form against the form that keeps two int64 grids per
    workgroup in HBM."""
    print("[gpu] fp32 tile merit vs HBM form, level 9", flush=True)
    wl = model.make_workload(model.config("composite-2e7", max_steps=2000000), n=1500, target_samples=200000)
    tiles, hbm, sparse = run_forms(wl, (1, 2, 3), K=4, n1=600, rep=300, seed=17)
    assert_same_run(tiles, hbm)
    assert_same_run(sparse, hbm)


def test_fp32_level7_many_trials_run():
    """On 128^2 grids one whole-grid LDS grid plus the block tables of K=24 trials exceed the
    LDS of a workgroup; the automatic choice then takes a form that fits (the HBM-scratch
    form: the tile grids of 26 logs do not fit either) instead of failing, and reps in one
    launch equal reps one by one, bit for bit."""
    print("[gpu] fp32 level 7 with K=24", flush=True)
    wl = composite_workload(n=1200, level=7)
    perm = model.permutation(wl, 6)
    out = []
    for mode in ("iterate", "reps"):
        g = GalicGPU(wl)
        g.fp32_calc_all_response_fields()
        if mode == "iterate":
            acc = g.fp32_iterate(perm, k=24, rep_size=400, sigma_frac=0.1, seed=19)
        else:
            acc = np.concatenate([g.fp32_rep(perm[a:a + 400], k=24, sigma_frac=0.1, seed=19)
                                  for a in range(0, len(perm), 400)])
        out.append([acc] + [g.stack(t, wh) for t in wl.types for wh in range(10)] + [g.particles()[0]])
        g.close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)
    assert (out[0][0] > 0).mean() > 0.2
