"""Barnes-Hut table fill on the GPU against the oracle tree.

The FP64 walk builds the same tree as the reference's insertion build
(forcetree.c:369-495) and walks it in the same order, so it matches the oracle
tree to rounding (1e-10 relative).  The FP32 walk (FP64 moments) is compared
with FP64 direct summation on a random subset of targets.
"""
import ctypes

import numpy as np
import pytest

from galic_amd import model
from galic_amd.runtime import (TREE_FP32, TREE_GROUP, TREE_PLUMMER, TREE_QUADRUPOLE, direct_forces, tree_fill,
                               tree_forces,
                                tree_last_timings)

pytestmark = pytest.mark.gpu


def sources(n, seed=3):
    cfg = model.config("composite-4e6")
    p = model.make_params(cfg)
    ids = np.arange(n, dtype=np.uint64)
    nd = n // 5
    nb = n // 20
    pos = np.concatenate([model.sample_halo(p, seed, 11, ids[:n - nd - nb]),
                          model.sample_disk(p, seed, 12, ids[n - nd - nb:n - nb]),
                          model.sample_halo(p, seed, 13, ids[n - nb:], a=p.Bulge_A)])
    mass = np.concatenate([np.full(n - nd - nb, p.Halo_Mass / (n - nd - nb)), np.full(nd, p.Disk_Mass / nd),
                           np.full(nb, p.Bulge_Mass / nb)])
    return p, np.ascontiguousarray(pos), mass


def oracle_forces(oracle_mod, p, spos, smass, tpos):
    P = ctypes.POINTER(ctypes.c_double)
    acc = np.zeros((len(tpos), 3))
    pot = np.zeros(len(tpos))
    n = oracle_mod.lib().or_tree_forces(ctypes.byref(p), len(spos), oracle_mod.d(spos), oracle_mod.d(smass),
                                        len(tpos), oracle_mod.d(tpos), acc.ctypes.data_as(P),
                                        pot.ctypes.data_as(P), 8)
    return acc, pot, n


def test_tree_fp64_matches_oracle_tree(oracle_mod):
    print("[gpu] tree fp64 vs oracle tree", flush=True)
    p, spos, smass = sources(20000)
    tpos = np.random.default_rng(4).normal(size=(3000, 3)) * 8
    acc_o, pot_o, n_o = oracle_forces(oracle_mod, p, spos, smass, tpos)
    acc_g, pot_g, n_g = tree_forces(p, spos, smass, tpos, precision=0)
    assert n_g == n_o
    np.testing.assert_allclose(acc_g, acc_o, rtol=1e-10, atol=1e-12 * np.abs(acc_o).max())
    np.testing.assert_allclose(pot_g, pot_o, rtol=1e-10)


def test_tree_theta_zero_equals_direct_sum(oracle_mod):
    p, spos, smass = sources(4000)
    p.ErrTolTheta = 1e-8
    tpos = np.random.default_rng(5).normal(size=(500, 3)) * 5
    P = ctypes.POINTER(ctypes.c_double)
    acc_d = np.zeros((500, 3))
    pot_d = np.zeros(500)
    oracle_mod.lib().or_direct_forces(ctypes.byref(p), len(spos), oracle_mod.d(spos), oracle_mod.d(smass), 500,
                                      oracle_mod.d(tpos), acc_d.ctypes.data_as(P), pot_d.ctypes.data_as(P), 8)
    acc_g, pot_g, n = tree_forces(p, spos, smass, tpos, precision=0)
    assert n == 500 * 4000
    np.testing.assert_allclose(acc_g, acc_d, rtol=1e-11, atol=1e-13 * np.abs(acc_d).max())
    np.testing.assert_allclose(pot_g, pot_d, rtol=1e-11)


def test_tree_fp32_walk_accuracy(oracle_mod):
    print("[gpu] tree fp32 walk", flush=True)
    p, spos, smass = sources(100000)
    tpos = np.random.default_rng(6).normal(size=(4096, 3)) * 6
    acc64, pot64, _ = tree_forces(p, spos, smass, tpos, precision=0)
    acc32, pot32, _ = tree_forces(p, spos, smass, tpos, precision=1)
    err = np.linalg.norm(acc32 - acc64, axis=1) / np.linalg.norm(acc64, axis=1)
    assert np.median(err) < 1e-5 and err.max() < 1e-3
    np.testing.assert_allclose(pot32, pot64, rtol=1e-4)
    # against FP64 direct summation: the theta = 0.4 monopole error dominates
    P = ctypes.POINTER(ctypes.c_double)
    acc_d = np.zeros((4096, 3))
    pot_d = np.zeros(4096)
    oracle_mod.lib().or_direct_forces(ctypes.byref(p), len(spos), oracle_mod.d(spos), oracle_mod.d(smass), 4096,
                                      oracle_mod.d(tpos), acc_d.ctypes.data_as(P), pot_d.ctypes.data_as(P), 8)
    e = np.linalg.norm(acc32 - acc_d, axis=1) / np.linalg.norm(acc_d, axis=1)
    print("fp32 walk vs direct: median", np.median(e), "p99", np.quantile(e, 0.99))
    assert np.median(e) < 5e-3


def test_table_fill_matches_oracle(oracle_mod):
    print("[gpu] table fill", flush=True)
    cfg = model.config("halo-disk-1e6", fg_nbin=32)
    p = model.make_params(cfg)
    p2, spos, smass = sources(20000)
    P = ctypes.POINTER(ctypes.c_double)
    nb2 = p.FG_Nbin ** 2
    po, ro, zo = np.zeros(nb2), np.zeros(nb2), np.zeros(nb2)
    oracle_mod.lib().or_table_fill(ctypes.byref(p), len(spos), oracle_mod.d(spos), oracle_mod.d(smass),
                                   po.ctypes.data_as(P), ro.ctypes.data_as(P), zo.ctypes.data_as(P), 8)
    pg, rg, zg, n = tree_fill(p, spos, smass, precision=0)
    np.testing.assert_allclose(pg, po, rtol=1e-10)
    np.testing.assert_allclose(rg, ro, rtol=1e-9, atol=1e-11 * np.abs(ro).max())
    np.testing.assert_allclose(zg, zo, rtol=1e-9, atol=1e-11 * np.abs(zo).max())


def test_direct_forces_match_oracle(oracle_mod):
    """galic_direct_forces (the FP64 check of the tree at full size) against the
    oracle's direct summation with the same spline kernel."""
    print("[gpu] direct summation", flush=True)
    p, spos, smass = sources(30000)
    tpos = np.random.default_rng(7).normal(size=(700, 3)) * 4
    P = ctypes.POINTER(ctypes.c_double)
    acc_d = np.zeros((700, 3))
    pot_d = np.zeros(700)
    oracle_mod.lib().or_direct_forces(ctypes.byref(p), len(spos), oracle_mod.d(spos), oracle_mod.d(smass), 700,
                                      oracle_mod.d(tpos), acc_d.ctypes.data_as(P), pot_d.ctypes.data_as(P), 8)
    acc_g, pot_g = direct_forces(p, spos, smass, tpos)
    np.testing.assert_allclose(acc_g, acc_d, rtol=1e-11, atol=1e-13 * np.abs(acc_d).max())
    np.testing.assert_allclose(pot_g, pot_d, rtol=1e-11)
    tree_forces(p, spos, smass, tpos, precision=1)
    tm = tree_last_timings()
    assert tm["interactions"] > 0 and tm["walk_ms"] > 0 and tm["build_ms"] > 0


def plummer_direct_np(p, spos, smass, tpos):
    """Plummer-softened direct summation in numpy (potential -G m / sqrt(r^2 + eps^2))."""
    eps = p.ForceSoftening / 2.8
    d = spos[None, :, :] - tpos[:, None, :]
    s2 = (d * d).sum(-1) + eps * eps
    inv = 1 / np.sqrt(s2)
    acc = p.G * (smass[None, :, None] * d * inv[..., None] ** 3).sum(1)
    pot = -p.G * (smass[None, :] * inv).sum(1)
    return acc, pot


def test_plummer_direct_sum_matches_numpy():
    p, spos, smass = sources(800)
    tpos = np.random.default_rng(8).normal(size=(60, 3)) * 3
    acc_n, pot_n = plummer_direct_np(p, spos, smass, tpos)
    acc_g, pot_g = direct_forces(p, spos, smass, tpos, plummer=True)
    np.testing.assert_allclose(acc_g, acc_n, rtol=1e-12, atol=1e-14 * np.abs(acc_n).max())
    np.testing.assert_allclose(pot_g, pot_n, rtol=1e-12)
    # theta -> 0 opens every node: the Plummer walk equals the direct sum
    p.ErrTolTheta = 1e-8
    for mode in (TREE_PLUMMER, TREE_PLUMMER | TREE_QUADRUPOLE):
        acc_t, pot_t, _ = tree_forces(p, spos, smass, tpos, precision=mode)
        np.testing.assert_allclose(acc_t, acc_n, rtol=1e-11, atol=1e-13 * np.abs(acc_n).max())


def test_quadrupole_of_a_distant_cluster():
    """A compact cluster seen from afar: the monopole error falls as (s/D)^2, the
    quadrupole error as (s/D)^3, so at s/D ~ 1/50 the quadrupole gains > 10x."""
    p = model.make_params(model.config("composite-4e6"))
    p.ForceSoftening = 1e-3
    p.ErrTolTheta = 0.7
    g = np.random.default_rng(9)
    spos = g.normal(size=(400, 3)) * np.array([1.0, 0.5, 0.25])   # an anisotropic cluster: large quadrupole
    smass = g.uniform(0.5, 1.5, 400)
    u = g.normal(size=(200, 3))
    tpos = 60 * u / np.linalg.norm(u, axis=1, keepdims=True)
    acc_d, pot_d = direct_forces(p, spos, smass, tpos)
    errs = {}
    for mode in (0, TREE_QUADRUPOLE):
        acc_t, pot_t, n = tree_forces(p, spos, smass, tpos, precision=mode)
        assert n < 400 * 200   # the walk takes nodes, not every particle
        errs[mode] = np.median(np.linalg.norm(acc_t - acc_d, axis=1) / np.linalg.norm(acc_d, axis=1))
    print("distant cluster: monopole", errs[0], "quadrupole", errs[TREE_QUADRUPOLE])
    assert errs[TREE_QUADRUPOLE] < errs[0] / 10


def test_quadrupole_walks_on_the_composite():
    """theta = 0.4 over 2e5 composite sources: quadrupole moments cut the force error
    against direct summation, FP64 and FP32 walks agree, for both softenings."""
    print("[gpu] quadrupole walks", flush=True)
    p, spos, smass = sources(200000)
    p.ErrTolTheta = 0.4
    tpos = np.random.default_rng(10).normal(size=(2048, 3)) * 6
    for soft in (0, TREE_PLUMMER):
        acc_d, pot_d = direct_forces(p, spos, smass, tpos, plummer=bool(soft))
        med = {}
        for mode in (soft, soft | TREE_QUADRUPOLE, soft | TREE_QUADRUPOLE | TREE_FP32):
            acc_t, pot_t, _ = tree_forces(p, spos, smass, tpos, precision=mode)
            e = np.linalg.norm(acc_t - acc_d, axis=1) / np.linalg.norm(acc_d, axis=1)
            med[mode] = np.median(e)
        print("softening", soft, "median force errors", med)
        assert med[soft | TREE_QUADRUPOLE] < med[soft] / 3
        assert med[soft | TREE_QUADRUPOLE | TREE_FP32] < 2 * med[soft | TREE_QUADRUPOLE] + 1e-5


@pytest.mark.parametrize("mode", [TREE_FP32, TREE_FP32 | TREE_QUADRUPOLE | TREE_PLUMMER])
def test_group_walk_no_less_accurate(mode):
    """The group walk (64 targets share one interaction list, a node accepted only where the
    criterion holds for the group's bounding box) against FP64 direct summation: every target's
    median and 99th-percentile errors no larger than the per-target FP32 walk's, up to FP32
    rounding, with at least as many interactions; a ragged last group of widely spread targets
    is among them."""
    print(f"[gpu] group walk mode {mode}", flush=True)
    p, spos, smass = sources(200000)
    p.ErrTolTheta = 0.4
    rng = np.random.default_rng(12)
    centres = rng.normal(size=(40, 3)) * 6
    tpos = (centres[:, None, :] + rng.normal(size=(40, 64, 3)) * 0.05).reshape(-1, 3)
    tpos = np.concatenate([tpos, rng.normal(size=(100, 3)) * 6])  # a last, ragged and spread group
    acc_d, pot_d = direct_forces(p, spos, smass, tpos, plummer=bool(mode & TREE_PLUMMER))
    acc_w, pot_w, n_w = tree_forces(p, spos, smass, tpos, precision=mode)
    acc_g, pot_g, n_g = tree_forces(p, spos, smass, tpos, precision=mode | TREE_GROUP)
    e_w = np.linalg.norm(acc_w - acc_d, axis=1) / np.linalg.norm(acc_d, axis=1)
    e_g = np.linalg.norm(acc_g - acc_d, axis=1) / np.linalg.norm(acc_d, axis=1)
    print("median errors: per target", np.median(e_w), "group", np.median(e_g), "interactions", n_w, n_g)
    assert n_g >= n_w
    assert np.median(e_g) <= np.median(e_w) * 1.05 + 1e-6
    assert np.quantile(e_g, 0.99) <= np.quantile(e_w, 0.99) * 1.05 + 1e-6
    p_w, p_g = np.abs(pot_w / pot_d - 1), np.abs(pot_g / pot_d - 1)
    assert np.median(p_g) <= np.median(p_w) * 1.05 + 1e-7
    assert p_g.max() <= p_w.max() * 1.5 + 1e-6


# ----------------------------------------------- re-tabulation and drop-ins --

def table_model(nbin=32, **kw):
    cfg = model.config("composite-4e6", fg_nbin=nbin, max_steps=2000000, **kw)
    return model.make_workload(cfg, n=1200, target_samples=50000)


@pytest.mark.parametrize("mode", [0, TREE_FP32 | TREE_QUADRUPOLE | TREE_PLUMMER,
                                  TREE_FP32 | TREE_QUADRUPOLE | TREE_PLUMMER | TREE_GROUP])
def test_retabulate_equals_tree_fill(mode):
    """galic_retabulate (the table fill on the device from the context's sources, installed in
    place) builds the same targets and domain as galic_tree_fill: identical tables; and the
    orbits integrated on them equal those after galic_set_force_tables of the same tables."""
    from galic_amd.runtime import GalicGPU
    print(f"[gpu] retabulate mode {mode}", flush=True)
    w = table_model()
    p, spos, smass = sources(30000)
    pt, rt, zt, nt = tree_fill(w.params, spos, smass, precision=mode)
    g = GalicGPU(w)
    g.set_tree_sources(spos, smass)
    n = g.retabulate(mode)
    pg, rg, zg = g.force_tables()
    assert n == nt
    np.testing.assert_array_equal(pg, pt)
    np.testing.assert_array_equal(rg, rt)
    np.testing.assert_array_equal(zg, zt)
    sel = np.arange(0, w.n, 40)
    args = (w.pos[sel], w.vel[sel], w.mass[sel], w.tint[sel], w.ptype[sel])
    a = g.orbit_responses(*args)
    a32 = g.orbit_responses(*args, precision=32)
    g.close()
    g2 = GalicGPU(w)
    import ctypes as ct
    from galic_amd._abi import check, dptr
    check(g2.lib.galic_set_force_tables(g2.h, dptr(pt), dptr(rt), dptr(zt)))
    b = g2.orbit_responses(*args)
    b32 = g2.orbit_responses(*args, precision=32)
    g2.close()
    for x, y in ((a, b), (a32, b32)):
        np.testing.assert_array_equal(x["fields"], y["fields"])
        np.testing.assert_array_equal(x["steps"], y["steps"])


def test_forcedensitygrid_calculate_dropin(oracle_mod):
    """forcedensitygrid_calculate() on the bound context (grid.c:621-757): SampleForceN points of
    each sampled component drawn on the device (the samplers of halo.c, disk.c, bulge.c under
    Philox keys (42, 200 + type)), the reference's walk, the tables installed; against the
    oracle's table fill from the same points drawn by the oracle's samplers."""
    from galic_amd.rng import uniforms
    from galic_amd.runtime import GalicGPU
    print("[gpu] forcedensitygrid_calculate", flush=True)
    w = table_model()
    p = w.params
    p.SampleForceNhalo, p.SampleForceNdisk, p.SampleForceNbulge = 12000, 4000, 1000
    g = GalicGPU(w)
    g.bind()
    fn = g.lib.forcedensitygrid_calculate
    fn.restype = None
    fn.argtypes = []
    fn()
    pg, rg, zg = g.force_tables()
    g.close()
    pts, ms = [], []
    for t, n in ((1, 12000), (2, 4000), (3, 1000)):
        pts.append(oracle_mod.sample_points(p, t, np.arange(n), uniforms, 42, 200 + t))
        ms.append(np.full(n, p.MType[t] / n))
    spos, smass = np.ascontiguousarray(np.concatenate(pts)), np.concatenate(ms)
    P = ctypes.POINTER(ctypes.c_double)
    nb2 = p.FG_Nbin ** 2
    po, ro, zo = np.zeros(nb2), np.zeros(nb2), np.zeros(nb2)
    oracle_mod.lib().or_table_fill(ctypes.byref(p), len(spos), oracle_mod.d(spos), oracle_mod.d(smass),
                                   po.ctypes.data_as(P), ro.ctypes.data_as(P), zo.ctypes.data_as(P), 8)
    np.testing.assert_allclose(pg, po, rtol=1e-10)
    np.testing.assert_allclose(rg, ro, rtol=1e-9, atol=1e-11 * np.abs(ro).max())
    np.testing.assert_allclose(zg, zo, rtol=1e-9, atol=1e-11 * np.abs(zo).max())


def test_forcetree_dropins_match_oracle(oracle_mod):
    """The forcetree module's entry points on a bound P[] (sources of types 1-3 and Type-5
    targets): gravity() against the oracle's tree on the same sources and targets; then
    force_treebuild, force_update_node_recursive from the root and force_treeevaluate target by
    target give gravity()'s values without the factor G and the same interaction count."""
    from galic_amd.runtime import GalicGPU
    print("[gpu] forcetree drop-ins", flush=True)
    w = table_model()
    p, spos, smass = sources(20000)
    rng = np.random.default_rng(8)
    tpos = rng.normal(size=(600, 3)) * 8
    n = len(spos) + len(tpos)
    pos = np.ascontiguousarray(np.concatenate([spos[:7000], tpos[:300], spos[7000:], tpos[300:]]))
    mass = np.concatenate([smass[:7000], np.zeros(300), smass[7000:], np.zeros(300)])
    typ = np.concatenate([np.ones(7000), np.full(300, 5), np.full(len(spos) - 7000, 2), np.full(300, 5)]).astype(np.int32)
    acc = np.zeros((n, 3))
    pot = np.zeros(n)
    g = GalicGPU(w)
    g.bind()
    lib = g.lib
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    assert lib.galic_bind_tree_particles(n, pos.ctypes.data_as(dp), mass.ctypes.data_as(dp), typ.ctypes.data_as(ip),
                                         acc.ctypes.data_as(dp), pot.ctypes.data_as(dp)) == 0
    lib.gravity.restype = None
    lib.gravity.argtypes = []
    lib.gravity()
    is_t = typ == 5
    acc_o, pot_o, n_o = oracle_forces(oracle_mod, w.params, np.ascontiguousarray(pos[~is_t]), mass[~is_t],
                                      np.ascontiguousarray(pos[is_t]))
    np.testing.assert_allclose(acc[is_t], acc_o, rtol=1e-10, atol=1e-12 * np.abs(acc_o).max())
    np.testing.assert_allclose(pot[is_t], pot_o, rtol=1e-10)
    assert (acc[~is_t] == 0).all()
    acc_g, pot_g = acc[is_t].copy(), pot[is_t].copy()
    acc[:] = 0
    pot[:] = 0
    lib.force_treebuild.restype = ctypes.c_int
    lib.force_treebuild.argtypes = [ctypes.c_int, ctypes.c_int]
    nodes = lib.force_treebuild(n, 1)
    assert 0 < nodes < 2 * len(spos)
    last = ctypes.c_int(-7)
    lib.force_update_node_recursive.restype = None
    lib.force_update_node_recursive.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip]
    lib.force_update_node_recursive(n, -1, -1, ctypes.byref(last))
    assert 0 <= last.value < n and typ[last.value] != 5
    lib.force_treeevaluate.restype = ctypes.c_int
    lib.force_treeevaluate.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    total = 0
    for i in range(int(is_t.sum())):
        total += lib.force_treeevaluate(i, 0, 0)
    assert total == n_o
    G = w.params.G
    np.testing.assert_array_equal(acc[is_t] * G, acc_g)
    np.testing.assert_array_equal(pot[is_t] * G, pot_g)
    g.close()
