"""Pinning of the oracle's sampled-target functions (grid.c:536-617, 941-1076).

The samplers are checked against the closed forms they invert, the dispersion
lookup against its tables at the nodes and in between, and the binning against
invariants (mass, linearity, a point at a cell centre) and against the analytic
density target of the same model, to the sampling noise.
"""
import ctypes

import numpy as np
import pytest

from galic_amd import model, rng


def wl(name, n=2000, **kw):
    return model.make_workload(model.config(name, **kw), n=n, target_samples=20000)


def draw(oracle_mod, p, t, u, stretch=1.0):
    pos = np.zeros(3)
    u = np.ascontiguousarray(u, dtype=np.float64)
    ok = oracle_mod.lib().or_sample_point(ctypes.byref(p), t, stretch, oracle_mod.d(u), oracle_mod.d(pos))
    return ok, pos


def test_halo_sampler_inverts_hernquist_mass(oracle_mod):
    """r = a(q + sqrt q)/(1 - q) solves M(<r)/M = (r/(r+a))^2 = q; phi = 2 pi u1, cos theta = 2 u2 - 1."""
    p = wl("halo-1e5").params
    a = p.Halo_A
    g = np.random.default_rng(3)
    kept = 0
    for _ in range(400):
        u = g.random(4)
        ok, x = draw(oracle_mod, p, 1, u)
        r = np.sqrt((x * x).sum())
        if not ok:
            assert r > p.Rmax
            continue
        kept += 1
        assert (r / (r + a)) ** 2 == pytest.approx(u[0], rel=1e-12)
        assert x[2] / r == pytest.approx(2 * u[2] - 1, abs=1e-12)
        assert np.mod(np.arctan2(x[1], x[0]), 2 * np.pi) == pytest.approx(2 * np.pi * u[1], abs=1e-9)
    assert kept > 300
    # the stretch divides z, and the halo tests the stretched radius
    ok, x = draw(oracle_mod, p, 1, [0.3, 0.2, 0.9, 0.0], stretch=2.0)
    ok1, x1 = draw(oracle_mod, p, 1, [0.3, 0.2, 0.9, 0.0])
    assert ok and ok1 and x[2] == pytest.approx(x1[2] / 2, rel=1e-15) and x[0] == x1[0]


def test_disk_sampler_inverts_exponential_and_sech2(oracle_mod):
    """(1 + R/h) exp(-R/h) = 1 - q (exponential disk mass) and tanh(z/z0) = 2 u0 - 1 (sech^2 layer)."""
    p = wl("halo-disk-1e6", n=2000).params
    g = np.random.default_rng(4)
    for _ in range(300):
        u = g.random(4)
        ok, x = draw(oracle_mod, p, 2, u)
        R = np.hypot(x[0], x[1]) / p.Disk_H
        if ok:
            assert (1 + R) * np.exp(-R) == pytest.approx(1 - u[1], abs=1e-12)
            assert np.tanh(x[2] / p.Disk_Z0) == pytest.approx(2 * u[0] - 1, abs=1e-12)
            assert np.mod(np.arctan2(x[1], x[0]), 2 * np.pi) == pytest.approx(2 * np.pi * u[2], abs=1e-9)
        else:
            assert (x * x).sum() > p.Rmax ** 2


def test_bulge_sampler_rejects_on_radius_first(oracle_mod):
    """bulge.c:17-38 redraws q alone while r > Rmax, then draws the direction; z / stretch."""
    p = wl("composite-4e6", n=2000).params
    a = p.Bulge_A
    q_out = (p.Rmax / (p.Rmax + a)) ** 2 * 1.0001   # just outside Rmax
    ok, _ = draw(oracle_mod, p, 3, [min(q_out, 0.999999), 0.5, 0.5, 0.5])
    assert not ok
    ok, x = draw(oracle_mod, p, 3, [0.25, 0.1, 0.7, 0.0], stretch=1.5)
    r = a * (0.25 + 0.5) / 0.75
    assert ok and x[2] == pytest.approx(r * (2 * 0.7 - 1) / 1.5, rel=1e-13)


def tables_of(w, t):
    beta, tabs = model.disp_tables(w.params, w.cfg, t)
    return beta, tabs


def test_disp_rtp_at_nodes_and_between(oracle_mod):
    """get_disp_rtp returns the table at a node, the linear (log-index) mean between nodes, and the
    structure's combinations: (1 - beta) for structure 1, the slope-dependent beta, v_str^2 in q."""
    w = wl("halo-disk-1e6", n=2000)
    p = w.params
    o = oracle_mod.Oracle(w, set_particles=False)
    nodes = model.fg_nodes(p)
    # spherical halo, structure 1, beta 0.3
    beta, tabs = tables_of(w, 1)
    o.set_disp_tables(1, beta, **tabs)
    T = tabs["disp_r"]
    for j in (3, 40, 100):
        d = o.disp_rtp([0.0, nodes[j] * 0.6, nodes[j] * 0.8], 1)
        assert d[0] == pytest.approx(T[j], rel=1e-10)
        assert d[1] == pytest.approx((1 - beta) * T[j], rel=1e-10)
        assert d[2] == d[1] and d[3] == d[1]
        rmid = p.FG_Rmin * (p.FG_Fac ** (j + 0.5) - 1)
        d = o.disp_rtp([rmid, 0.0, 0.0], 1)
        assert d[0] == pytest.approx(0.5 * (T[j] + T[j + 1]), rel=1e-9)
    # slope-dependent beta (parameter >= 1): beta = -0.15 - 0.2 dlog rho/dlog r of the Hernquist profile
    o.set_disp_tables(1, 1.0, **tabs)
    r = nodes[50]
    d = o.disp_rtp([r, 0.0, 0.0], 1)
    b = -0.15 - 0.20 * (-1.0 - 3.0 * r / (r + p.Halo_A))
    assert d[1] == pytest.approx((1 - b) * d[0], rel=1e-13)
    # the disk: structure 3 (tilted) reads tilt_r, tilt_theta, disp_phi and adds v_str^2 in q
    beta, tabs = tables_of(w, 2)
    assert p.TypeOfVelocityStructure[2] == 3
    o.set_disp_tables(2, beta, **tabs)
    nb = p.FG_Nbin
    for (k, j) in ((2, 10), (20, 60)):
        pos = [nodes[j], 0.0, nodes[k]]
        d = o.disp_rtp(pos, 2)
        vs = w.vstream[2][k * nb + j]
        assert d[0] == pytest.approx(tabs["tilt_r"][k * nb + j], rel=1e-9)
        assert d[1] == pytest.approx(tabs["tilt_theta"][k * nb + j], rel=1e-9)
        assert d[2] == pytest.approx(tabs["disp_phi"][k * nb + j], rel=1e-9)
        assert d[3] == pytest.approx(d[2] + vs * vs, rel=1e-9)
    o.close()


def test_binning_invariants(oracle_mod):
    """Each point adds CIC weights that sum to one, so every map's total is MType (times the
    dispersion for a constant table); a point at a cell centre lands in that cell alone."""
    w = wl("halo-1e5")
    p = w.params
    o = oracle_mod.Oracle(w, set_particles=False)
    c = 1234.5
    o.set_disp_tables(1, 0.0, disp_r=np.full(p.FG_Nbin, c))
    pos = model.sample_type(p, 1, 7, 101, np.arange(5000, dtype=np.uint64))
    dg, eg, n = o.bin_targets(1, pos)
    keep = np.sqrt(pos[:, 0] ** 2 + pos[:, 2] ** 2) < p.Rmax
    assert n == keep.sum()
    L = p.DG_MaxLevel
    off = model.stack_offset(L, 0, 0)
    assert dg[0] == pytest.approx(p.MType[1], rel=1e-12)
    assert dg[off:].sum() == pytest.approx(p.MType[1], rel=1e-12)
    assert eg[0][0] == pytest.approx(p.MType[1], rel=1e-12)
    for f in range(1, 5):   # a constant table: every dispersion map is c times the mass map
        np.testing.assert_allclose(eg[f], c * eg[0], rtol=1e-12)
    # one point at the centre of density cell (iz, iR) = (5, 9)
    nb = 1 << L
    Rc = p.DG_Rmin * (p.DG_Fac ** (9 + 0.5) - 1)
    zc = p.DG_Rmin * (p.DG_Fac ** (5 + 0.5) - 1)
    dg1, _, n1 = o.bin_targets(1, np.array([[Rc, 0.0, zc]]), energy=False)
    fine = dg1[off:].reshape(nb, nb)
    assert n1 == 1 and fine[5, 9] == pytest.approx(p.MType[1], rel=1e-9)
    assert abs(fine.sum() - fine[5, 9]) < 1e-9 * p.MType[1]
    o.close()


@pytest.mark.slow
def test_sampled_density_target_matches_analytic(oracle_mod):
    """The sampled DGs_MassTarget of the halo has the shape of the analytic vol*rho target of the same
    profile (the reference's other method, grid.c:830-877) within the sampling noise of each populated
    cell.  The two differ by one factor: the analytic target spreads MType over the whole (R, |z|)
    square, corners outside the truncation sphere included, the sampled one over the sphere."""
    w = wl("halo-1e5")
    p = w.params
    o = oracle_mod.Oracle(w, set_particles=False)
    n = 400000
    pos = model.sample_type(p, 1, 11, 101, np.arange(n, dtype=np.uint64))
    dg, _, cnt = o.bin_targets(1, pos, energy=False)
    ref = model.density_target(p, 1)
    L = p.DG_MaxLevel
    # compare at level L-2 (4x4 cells), where the CIC spread and the cell-centre rule of the
    # analytic target differ least; cells with >= 2000 expected points
    lv = L - 2
    a = slice(model.stack_offset(lv, 0, 0), model.stack_offset(lv + 1, 0, 0))
    expect_pts = ref[a] / p.MType[1] * cnt
    sel = expect_pts > 2000
    assert sel.sum() > 20
    ratio = dg[a][sel] / ref[a][sel]
    scale = np.median(ratio)
    assert 1.0 < scale < 1.15
    noise = 1 / np.sqrt(expect_pts[sel])
    assert np.median(np.abs(ratio / scale - 1) / noise) < 3
    o.close()


def test_oracle_points_follow_host_draws(oracle_mod):
    """sample_points with rng.uniforms gives the model's own sampler's points (same Philox rounds)."""
    w = wl("halo-disk-1e6", n=2000)
    p = w.params
    ids = np.arange(300, dtype=np.uint64)
    for t in (1, 2):
        a = oracle_mod.sample_points(p, t, ids, rng.uniforms, 5, 100 + t)
        b = model.sample_type(p, t, 5, 100 + t, ids)
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12 * p.Rmax)
