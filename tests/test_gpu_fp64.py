"""Parity of the plain FP64 GPU path with the CPU oracle (same inputs).

Tolerance: 1e-10 relative.  Grids are compared cell by cell against
1e-10 * (|oracle| + 1e-6 * max|oracle|): the floor covers cells that are
differences of much larger contributions (a delta of near-equal orbits), where
a single rounding of the larger terms exceeds 1e-10 of the small result.
Integer outputs (orbit counts, step counts, accept flags) must be equal.
"""
import numpy as np
import pytest

from galic_amd import model
from galic_amd.runtime import GalicGPU

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def grid_close(a, b):
    """Grids: normwise (L1) relative error <= 1e-10, and no cell off by more
    than 1e-8 of |cell| + 1e-6 max|grid|.  Sparsely populated cells are sums of
    few orbits that barely touch them; ulp-level differences between the device
    and host log() move those orbits by ~1e-13 and such cells by up to ~1e-9."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    nrm = np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300)
    cell_ok, cell_worst = close(a, b, rtol=1e-8)
    return (nrm <= RTOL and cell_ok), (nrm, cell_worst)


def close(a, b, rtol=RTOL, floor=1e-6):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    scale = np.abs(b) + floor * (np.abs(b).max() if b.size else 0.0)
    bad = np.abs(a - b) > rtol * scale
    return not bad.any(), (np.abs(a - b) / np.where(scale > 0, scale, 1)).max() if b.size else 0.0


def workload(name, n, **kw):
    kw.setdefault("max_steps", 2000000)
    cfg = model.config(name, **kw)
    return model.make_workload(cfg, n=n, target_samples=max(100000, 20 * n))


@pytest.fixture(scope="module")
def halo():
    return workload("halo-1e5", 3000)


@pytest.fixture(scope="module")
def halodisk():
    return workload("halo-disk-1e6", 3000)


@pytest.fixture(scope="module")
def halodisk9():
    """512 x 512 response grids (level 9): the merit pyramid's coarse levels no longer fit in LDS."""
    return workload("halo-disk-1e6", 1200, level=9)


@pytest.fixture(scope="module")
def composite():
    """Halo (beta 0.3, structure 1) + disk (structure 3) + Hernquist bulge (structure 0): every
    analytic component of the force (halo.c:63-84, bulge.c:64-85) and all three types."""
    return workload("composite-4e6", 3000)


@pytest.fixture(scope="module")
def composite9():
    """The full v1.1 model's form (isotropic halo and bulge, structure-3 disk) on 512^2 force
    tables and 512^2 response grids."""
    return workload("composite-2e7", 1500)


@pytest.mark.parametrize("wname", ["halo", "halodisk", "composite"])
def test_orbit_responses_match_oracle(wname, request, oracle_mod):
    w = request.getfixturevalue(wname)
    print(f"[gpu] orbit responses {wname} n=64", flush=True)
    o = oracle_mod.Oracle(w, set_particles=False)
    g = GalicGPU(w)
    sel = np.arange(0, w.n, max(1, w.n // 64))[:64]
    r = g.orbit_responses(w.pos[sel], w.vel[sel], w.mass[sel], w.tint[sel], w.ptype[sel])
    nbad = 0
    for k, i in enumerate(sel):
        ro = o.orbit(w.pos[i], w.vel[i], w.mass[i], w.tint[i], w.ptype[i])
        assert r["steps"][k] == ro["steps"], (i, r["steps"][k], ro["steps"])
        assert r["orbits"][k] == ro["orbits"]
        assert r["t"][k] == pytest.approx(ro["t"], rel=1e-14)
        # a single orbit is compared normwise per field: ulp-level differences of
        # the device and host log() move the trajectory by ~1e-13 of its size,
        # which changes cells the orbit barely touches by more than 1e-10 of
        # their own (tiny) value; the grids below, sums over many orbits, are
        # compared cell by cell.
        a, b = r["fields"][k], ro["fields"]
        rel = np.abs(a - b).sum(axis=1) / np.maximum(np.abs(b).sum(axis=1), 1e-300)
        if not (rel <= RTOL).all():
            nbad += 1
            print("orbit", i, "normwise rel", rel)
        assert r["e1"][k] == pytest.approx(ro["e1"], rel=1e-10)
    assert nbad == 0
    g.close()


@pytest.mark.parametrize("wname", ["halo", "halodisk", "composite", "composite9"])
def test_field_lookup_matches_oracle(wname, request, oracle_mod):
    w = request.getfixturevalue(wname)
    o = oracle_mod.Oracle(w, set_particles=False)
    g = GalicGPU(w)
    acc, pot = g.field_at(w.pos[:500])
    for i in range(0, 500, 7):
        np.testing.assert_allclose(acc[i], o.accel(w.pos[i]), rtol=1e-13, atol=1e-13 * np.abs(acc[i]).max())
        assert pot[i] == pytest.approx(o.potential(w.pos[i]), rel=1e-13)
    g.close()


@pytest.mark.parametrize("wname", ["halo", "halodisk", "halodisk9", "composite", "composite9"])
def test_optimiser_history_matches_oracle(wname, request, oracle_mod):
    """calc_all_response_fields, then reps of trials: grids, accept flags and
    the Sall history agree with the oracle."""
    w = request.getfixturevalue(wname)
    print(f"[gpu] optimiser history {wname}", flush=True)
    o = oracle_mod.Oracle(w)
    g = GalicGPU(w)
    o.calc_all_response_fields(8)
    g.calc_all_response_fields()
    types = w.types
    for t in types:
        for which in range(10):
            ok, worst = grid_close(g.stack(t, which), o.stack(t, which))
            print("initial", t, which, worst)
            assert ok, (t, which, worst)
        np.testing.assert_allclose(g.fit(t), o.fit(t), rtol=RTOL)
    perm = model.permutation(w, 1)
    ntr = w.n // 4
    for rep in range(3):
        idx = perm[rep * ntr:(rep + 1) * ntr]
        v = model.trial_velocities(w, idx, 1, rep)
        acc_o, dF_o = o.rep(idx, v, 8)
        acc_g, dF_g = g.rep(idx, v)
        flips = np.nonzero(acc_o != acc_g)[0]
        assert len(flips) == 0, (rep, flips[:5], dF_o[flips[:5]], dF_g[flips[:5]])
        scale = np.abs(dF_o).max()
        np.testing.assert_allclose(dF_g, dF_o, rtol=1e-8, atol=1e-9 * scale)
        for t in types:
            for which in range(10):
                ok, worst = grid_close(g.stack(t, which), o.stack(t, which))
                assert ok, (rep, t, which, worst)
            np.testing.assert_allclose(g.fit(t), o.fit(t), rtol=RTOL)
    vel_g, orb_g, comp_g = g.particles()
    vel_o, orb_o, comp_o = o.particles()
    np.testing.assert_array_equal(vel_g, vel_o)
    np.testing.assert_array_equal(orb_g, orb_o)
    np.testing.assert_allclose(comp_g, comp_o, rtol=RTOL, atol=1e-12 * np.abs(comp_o).max())
    g.close()


def test_dropin_orbit_signature(halo, oracle_mod):
    """produce_orbit_response_field_mod with the reference signature (orbit_response.c:155)."""
    import ctypes
    w = halo
    g = GalicGPU(w)
    g.bind()
    lib = g.lib
    ng = g.ngrid
    bufs = [np.zeros(ng) for _ in range(5)]
    dp = ctypes.POINTER(ctypes.c_double)
    fn = lib.produce_orbit_response_field_mod
    fn.restype = ctypes.c_double
    fn.argtypes = [dp, dp, ctypes.c_int, dp, dp, dp, dp, dp, ctypes.c_double, ctypes.c_double,
                   ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    norb = ctypes.c_int(0)
    pos = np.ascontiguousarray(w.pos[5])
    vel = np.ascontiguousarray(w.vel[5])
    t = fn(pos.ctypes.data_as(dp), vel.ctypes.data_as(dp), 5, *[b.ctypes.data_as(dp) for b in bufs],
           w.mass[5], w.tint[5], ctypes.byref(norb), 1)
    o = oracle_mod.Oracle(w, set_particles=False)
    ro = o.orbit(w.pos[5], w.vel[5], w.mass[5], w.tint[5], 1)
    assert t == pytest.approx(ro["t"], rel=1e-14)
    assert norb.value == ro["orbits"]
    ok, worst = close(np.stack(bufs), ro["fields"])
    assert ok, worst
    g.close()


def test_dropin_plain_orbit_signature(composite, oracle_mod):
    """produce_orbit_response_field (orbit_response.c:33-151, the plain variant: the mass field
    only) with the reference signature, for particles of each type, against the oracle's plain
    orbit."""
    import ctypes
    w = composite
    g = GalicGPU(w)
    g.bind()
    fn = g.lib.produce_orbit_response_field
    dp = ctypes.POINTER(ctypes.c_double)
    fn.restype = ctypes.c_double
    fn.argtypes = [dp, dp, ctypes.c_int, dp, ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_int)]
    o = oracle_mod.Oracle(w, set_particles=False)
    L = oracle_mod.lib()
    for t in w.types:
        for i in np.nonzero(w.ptype == t)[0][:3]:
            pos = np.ascontiguousarray(w.pos[i])
            vel = np.ascontiguousarray(w.vel[i])
            mf = np.zeros(g.ngrid)
            norb = ctypes.c_int(0)
            tg = fn(pos.ctypes.data_as(dp), vel.ctypes.data_as(dp), int(i), mf.ctypes.data_as(dp), w.mass[i],
                    w.tint[i], ctypes.byref(norb))
            mo = np.zeros(g.ngrid)
            no = ctypes.c_int(0)
            so = ctypes.c_longlong(0)
            to = L.or_orbit_response_plain(o.h, oracle_mod.d(pos), oracle_mod.d(vel), mo.ctypes.data_as(dp),
                                           w.mass[i], w.tint[i], ctypes.byref(no), ctypes.byref(so))
            assert tg == pytest.approx(to, rel=1e-14), (t, i)
            assert norb.value == no.value
            assert np.abs(mf - mo).sum() <= RTOL * np.abs(mo).sum(), (t, i)
            assert mf.sum() == pytest.approx(w.mass[i], rel=1e-12)
    o.close()
    g.close()


def test_dropin_lookups_and_stack_differences(halodisk, oracle_mod):
    """The remaining reference-signature entry points (grid.c:15-55, 494-534, 896-938,
    1086-1151; orbit_response.c:16-26) on the bound context equal the oracle's restatements."""
    import ctypes
    w = halodisk
    p = w.params
    g = GalicGPU(w)
    g.bind()
    lib = g.lib
    o = oracle_mod.Oracle(w, set_particles=False)
    L = oracle_mod.lib()
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    rng = np.random.default_rng(12)
    pts = np.concatenate([rng.normal(size=(300, 3)) * 10, rng.normal(size=(50, 3)) * 300, np.zeros((1, 3))])
    for name in ("forcegrid_get_cell", "densitygrid_get_cell", "energygrid_get_cell"):
        f = getattr(lib, name)
        f.restype = None
        f.argtypes = [dp, ip, ip, dp, dp]
        ref = getattr(L, "or_" + name)
        for x in pts:
            x = np.ascontiguousarray(x)
            a = [ctypes.c_int(), ctypes.c_int(), ctypes.c_double(), ctypes.c_double()]
            b = [ctypes.c_int(), ctypes.c_int(), ctypes.c_double(), ctypes.c_double()]
            f(x.ctypes.data_as(dp), *[ctypes.byref(v) for v in a])
            ref(o.h, oracle_mod.d(x), *[ctypes.byref(v) for v in b])
            assert [v.value for v in a] == [v.value for v in b], (name, x)
    ts = lib.get_timestep
    ts.restype = ctypes.c_double
    ts.argtypes = [dp, dp, dp, ctypes.c_int]
    for k in range(50):
        x, v, a = [np.ascontiguousarray(rng.normal(size=3) * s) for s in (5, 100, 1000)]
        icell = int(rng.integers(0, g.ngrid))
        assert ts(x.ctypes.data_as(dp), v.ctypes.data_as(dp), a.ctypes.data_as(dp), icell) == \
            L.or_get_timestep(o.h, oracle_mod.d(x), oracle_mod.d(v), oracle_mod.d(a), icell)
    ns = g.nstack
    Lv = p.DG_MaxLevel
    d1, d2, r1, r2 = [np.ascontiguousarray(rng.uniform(0, 1, ns)) for _ in range(4)]
    r1[rng.random(ns) < 0.1] = 0
    dist = np.ascontiguousarray(rng.uniform(0.1, 50, ns))
    for name, flags in (("calc_stack_difference", (0, 1)), ("calc_stack_difference_mod", (0, 1, 2))):
        f = getattr(lib, name)
        f.restype = ctypes.c_double
        f.argtypes = [dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_double, dp,
                      ctypes.c_int]
        ref = getattr(L, "or_" + name)
        for flag in flags:
            for thresh in (0.2, 0.6):
                args = [d1.ctypes.data_as(dp), d2.ctypes.data_as(dp), 0, 0, 0, Lv, r1.ctypes.data_as(dp),
                        r2.ctypes.data_as(dp), thresh, dist.ctypes.data_as(dp), flag]
                a = f(*args)
                b = ref(o.h, oracle_mod.d(d1), oracle_mod.d(d2), 0, 0, 0, Lv, oracle_mod.d(r1), oracle_mod.d(r2),
                        thresh, oracle_mod.d(dist), flag)
                assert a == b and a > 0, (name, flag, thresh)
    o.close()
    g.close()
