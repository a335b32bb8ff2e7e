"""Pinning of the CPU oracle (oracle/galic_oracle.c) without the reference.

The reference cannot be compiled in this image (it needs GSL and a generated
config header), and it ships no tests or fixtures.  Each oracle function is
therefore checked against something its mistakes would break: a closed form,
an invariant, or a brute-force evaluation by a different algorithm.
"""
import ctypes

import numpy as np
import pytest

from galic_amd import _abi, model


def small(name="halo-1e5", n=400, **kw):
    kw.setdefault("max_steps", 2000000)
    return model.make_workload(model.config(name, **kw), n=n, target_samples=50000)


# ------------------------------------------------------------------ geometry --

def test_grid_geometry_closed_form(oracle_mod):
    """Rmin*(Fac^Nbin - 1) = Rmax and Rmin*(Fac - 1) = Rin (grid.c:347-374)."""
    p = model.make_params(model.config("halo-1e5"))
    q = p.copy()
    assert oracle_mod.lib().or_grid_create(ctypes.byref(q), model.config("halo-1e5").outer_frac, 1e-7) == 0
    a = p.Halo_A
    f = 1e-7
    rin = a * (f + np.sqrt(f)) / (1 - f)
    assert q.Rmax == pytest.approx(30 * a, rel=1e-14)   # truncation at 30 a
    for nb, rmin, fac in [(1 << p.DG_MaxLevel, q.DG_Rmin, q.DG_Fac), (p.FG_Nbin, q.FG_Rmin, q.FG_Fac)]:
        assert rmin * (fac ** nb - 1) == pytest.approx(q.Rmax, rel=1e-6)
        assert rmin * (fac - 1) == pytest.approx(rin, rel=1e-12)
    # the product's set-up routine gives bit-identical geometry
    for k in ("Rmax", "DG_Rmin", "DG_Fac", "FG_Rmin", "FG_Fac", "EG_Rmin", "EG_Fac"):
        assert getattr(p, k) == getattr(q, k), k


def test_density_geometry_closed_form(oracle_mod):
    w = small()
    o = oracle_mod.Oracle(w, set_particles=False)
    p = w.params
    nb = 1 << p.DG_MaxLevel
    edges = p.DG_Rmin * (p.DG_Fac ** np.arange(nb + 1) - 1.0)
    width = np.diff(edges)
    cs = o.cellsize().reshape(nb, nb)
    np.testing.assert_allclose(cs, np.maximum(width[None, :], width[:, None]), rtol=1e-12)
    # finest-level distance = |(R_c, z_c)| of the cell centres (grid.c:388-416)
    ctr = 0.5 * (edges[:-1] + edges[1:])
    dist = o.distance()
    fine = dist[model.stack_offset(p.DG_MaxLevel, 0, 0):].reshape(nb, nb)
    np.testing.assert_allclose(fine, np.hypot(ctr[None, :], ctr[:, None]), rtol=1e-13)
    # coarse level: geometric mean of the children's R and z (mean of logs)
    L = p.DG_MaxLevel
    lr = np.log(ctr)
    i, j = 3, 5
    kids = [(2 * i + a, 2 * j + b) for a in (0, 1) for b in (0, 1)]
    Rg = np.exp(np.mean([lr[jj] for _, jj in kids]))
    zg = np.exp(np.mean([lr[ii] for ii, _ in kids]))
    assert dist[model.stack_offset(L - 1, i, j)] == pytest.approx(np.hypot(Rg, zg), rel=1e-13)
    # product geometry equals the oracle's bit for bit
    cs2 = np.zeros(nb * nb)
    d2 = np.zeros(len(dist))
    _abi.check(_abi.lib().galic_density_geometry(ctypes.byref(p), _abi.dptr(cs2), None, _abi.dptr(d2)))
    np.testing.assert_array_equal(cs2, o.cellsize())
    np.testing.assert_array_equal(d2, dist)


# ------------------------------------------------------------------- lookups --

def test_halo_field_is_hernquist(oracle_mod):
    """Analytic mode: a = -grad Phi of Hernquist (with the 1e-6 a softening of halo.c:79)."""
    w = small()
    o = oracle_mod.Oracle(w, set_particles=False)
    p = w.params
    rng = np.random.default_rng(0)
    for _ in range(20):
        x = rng.normal(size=3) * 20
        r = np.linalg.norm(x)
        assert o.potential(x) == pytest.approx(-p.G * p.Halo_Mass / (r + p.Halo_A), rel=1e-14)
        h = 1e-4 * r
        grad = np.array([(o.potential(x + h * e) - o.potential(x - h * e)) / (2 * h) for e in np.eye(3)])
        # the force softens r -> r + 1e-6 a in one factor: -grad Phi * r/(r + 1e-6 a)
        np.testing.assert_allclose(o.accel(x), -grad * r / (r + 1e-6 * p.Halo_A), rtol=1e-7)


def test_table_interpolation_is_exact_for_bilinear_fields(oracle_mod):
    """Bilinear in the log index: a table T[k][j] = a + b j + c k + d j k is
    reproduced exactly at any point (grid.c:130-146)."""
    w = small(halo_in_table=True)
    p = w.params
    nb = p.FG_Nbin
    k, j = np.meshgrid(np.arange(nb), np.arange(nb), indexing="ij")
    T = (1.5 + 0.25 * j - 0.5 * k + 0.01 * j * k).ravel()
    w.dpotdr = np.ascontiguousarray(T)
    w.dpotdz = np.ascontiguousarray(2 * T)
    w.pot = np.ascontiguousarray(-T)
    p.Halo_Mass_saved = None
    o = oracle_mod.Oracle(w, set_particles=False)
    rng = np.random.default_rng(1)
    pts = rng.normal(size=(50, 3)) * 10
    for x in pts:
        R = np.hypot(x[0], x[1])
        jd = np.log(R / p.FG_Rmin + 1) / np.log(p.FG_Fac)
        kd = np.log(abs(x[2]) / p.FG_Rmin + 1) / np.log(p.FG_Fac)
        expect = 1.5 + 0.25 * jd - 0.5 * kd + 0.01 * jd * kd
        acc = o.accel(x)
        # halo is tabulated (SampleForceNhalo>0): no analytic term is added
        assert np.hypot(acc[0], acc[1]) == pytest.approx(abs(expect), rel=1e-12)
        assert acc[2] == pytest.approx(np.sign(x[2]) * 2 * expect, rel=1e-12)
        assert o.potential(x) == pytest.approx(-expect, rel=1e-12)


def test_cell_finders_brute_force(oracle_mod):
    """densitygrid_get_cell == search over the cell-centre grid (grid.c:494-534)."""
    w = small()
    o = oracle_mod.Oracle(w, set_particles=False)
    p = w.params
    nb = 1 << p.DG_MaxLevel
    lib = oracle_mod.lib()
    edges = p.DG_Rmin * (p.DG_Fac ** np.arange(nb + 1) - 1.0)
    # in log-index space the cell centre of bin j sits at j + 0.5
    rng = np.random.default_rng(2)
    for x in rng.normal(size=(200, 3)) * np.exp(rng.uniform(-3, 5, size=(200, 1))):
        iR, iz = ctypes.c_int(), ctypes.c_int()
        fR, fz = ctypes.c_double(), ctypes.c_double()
        lib.or_densitygrid_get_cell(o.h, oracle_mod.d(x), ctypes.byref(iR), ctypes.byref(iz), ctypes.byref(fR),
                                    ctypes.byref(fz))
        for val, ib, fb in ((np.hypot(x[0], x[1]), iR.value, fR.value), (abs(x[2]), iz.value, fz.value)):
            # continuous index u with edges[u] = val; centre of cell j at u = j + 0.5
            jj = np.searchsorted(edges, val) - 1
            u = jj + (np.log(val / p.DG_Rmin + 1) - np.log(edges[jj] / p.DG_Rmin + 1)) / np.log(p.DG_Fac) \
                if val > 0 else 0.0
            b = max(u - 0.5, 0.0)
            if b >= nb - 1:
                assert ib == nb - 2 and fb == 1.0
            else:
                assert ib == int(np.floor(b))
                assert fb == pytest.approx(b - ib, abs=1e-9)


# --------------------------------------------------------------------- orbits --

def circular_orbit(p, r, polar=False):
    vc = np.sqrt(p.G * p.Halo_Mass * r * r / ((r + 1e-6 * p.Halo_A) * (r + p.Halo_A) ** 2))
    if polar:
        return np.array([r, 0.0, 0.0]), np.array([0.0, 0.0, vc]), vc
    return np.array([r, 0.0, 0.0]), np.array([0.0, vc, 0.0]), vc


def test_circular_orbit_binning(oracle_mod):
    """Circular orbit in the plane z=0: all time sits at R=r, so the response is
    the CIC weights of r in row iz=0; vq2 = vc^2, vr2 = vt2 = 0; and the step
    count is ceil(T/dt) for the constant timestep (orbit_response.c:16-26, 276-319)."""
    w = small()
    o = oracle_mod.Oracle(w, set_particles=False)
    p = w.params
    nb = 1 << p.DG_MaxLevel
    r = 17.0
    pos, vel, vc = circular_orbit(p, r)
    T = 3 * 2 * np.pi * r / vc
    res = o.orbit(pos, vel, 0.5, T, 1)
    f = res["fields"].reshape(5, nb, nb)
    br = np.log(r / p.DG_Rmin + 1) / np.log(p.DG_Fac) - 0.5
    i = int(br)
    fr = br - i
    assert res["t"] == T
    assert f[0].sum() == pytest.approx(0.5, rel=1e-13)
    # the leapfrog keeps R within ~1e-5 of r; the CIC weights follow to ~1e-4
    assert f[0][0, i] == pytest.approx(0.5 * (1 - fr), rel=1e-3)
    assert f[0][0, i + 1] == pytest.approx(0.5 * fr, rel=1e-3)
    assert np.abs(f[0][1:]).max() == 0
    assert f[3].sum() == pytest.approx(0.5 * vc * vc, rel=1e-4)    # q field: vphi^2
    assert f[4].sum() == pytest.approx(0.5 * vc * vc, rel=1e-4)    # p field: (vphi - 0)^2
    assert f[1].sum() < 1e-8 * f[3].sum()                          # no radial motion
    assert f[2].sum() == 0                                         # no polar motion
    a = p.G * p.Halo_Mass * r / ((r + 1e-6 * p.Halo_A) * (r + p.Halo_A) ** 2)
    cell = (w.params.DG_Rmin * (p.DG_Fac ** (i + 1) - 1) - w.params.DG_Rmin * (p.DG_Fac ** i - 1))
    dt = min(p.TimeStepFactorOrbit * p.V200 / a, p.TimeStepFactorCellCross * max(cell, o.cellsize()[i]) / vc)
    assert abs(res["steps"] - np.ceil(T / dt)) <= 1
    assert abs(res["e1"] - res["e0"]) < 1e-6 * abs(res["e0"])
    # 3 revolutions of a circular orbit: radsign stays ~0, no pericentre is counted reliably;
    # the polar orbit puts all motion into the t field
    pos, vel, vc = circular_orbit(p, r, polar=True)
    res = o.orbit(pos, vel, 0.5, T, 1)
    f = res["fields"]
    assert f[2].sum() == pytest.approx(0.5 * vc * vc, rel=1e-4)
    assert f[3].sum() < 1e-12 * f[2].sum()


def test_pericentre_count_matches_radial_period(oracle_mod):
    """Eccentric orbit: number of pericentre passages = floor(T / T_r) (+-1 for
    the start phase), with T_r from the radial-period quadrature."""
    w = small()
    o = oracle_mod.Oracle(w, set_particles=False)
    p = w.params
    r0 = 20.0
    pos = np.array([r0, 0, 0.0])
    vel = np.array([0.0, 60.0, 0.0])        # starts at apocentre
    E = 0.5 * 60.0 ** 2 - p.G * p.Halo_Mass / (r0 + p.Halo_A)
    Lz = r0 * 60.0
    rr = np.linspace(0.01, r0, 400001)
    kin = 2 * (E + p.G * p.Halo_Mass / (rr + p.Halo_A)) - Lz * Lz / rr ** 2
    ok = kin > 0
    rp = rr[ok][0]
    # T_r = 2 int dr / sqrt(kin): substitute r = rp + (ra-rp) sin^2 to remove the end singularities
    th = np.linspace(0, np.pi / 2, 200001)[1:-1]
    rs = rp + (r0 - rp) * np.sin(th) ** 2
    ks = 2 * (E + p.G * p.Halo_Mass / (rs + p.Halo_A)) - Lz * Lz / rs ** 2
    Tr = 2 * np.trapezoid(2 * (r0 - rp) * np.sin(th) * np.cos(th) / np.sqrt(np.maximum(ks, 1e-300)), th)
    T = 7.3 * Tr
    res = o.orbit(pos, vel, 1.0, T, 1)
    assert abs(res["orbits"] - 7) <= 1
    assert abs(res["e1"] - res["e0"]) < 1e-4 * abs(res["e0"])


def test_response_converges_at_second_order(oracle_mod):
    """Halving both timestep factors shrinks the change of the time-averaged mass
    response by ~4x: KDK leapfrog with trapezoid time weights is second order.
    (The end energy is not used: the softened force of halo.c:79 is not exactly
    -grad of halo.c:71, which leaves an O(1e-6) energy drift at any timestep.)"""
    fs = []
    for f in (1.0, 0.5, 0.25):
        w = small(n=50, timestep_orbit=0.01 * f, timestep_cell=0.25 * f)
        o = oracle_mod.Oracle(w, set_particles=False)
        fs.append(o.orbit(np.array([15.0, 0, 3.0]), np.array([0, 90.0, 40.0]), 1.0, 3.0, 1)["fields"][0])
    d1 = np.abs(fs[0] - fs[1]).sum()
    d2 = np.abs(fs[1] - fs[2]).sum()
    assert d1 / d2 > 2.8


# --------------------------------------------------------------------- stacks --

def brute_stack(data, L):
    out = np.zeros(model.nstack(L))
    off = model.stack_offset(L, 0, 0)
    fine = data.reshape(1 << L, 1 << L)
    for l in range(L + 1):
        f = 1 << (L - l)
        n = 1 << l
        blk = fine.reshape(n, f, n, f).sum(axis=(1, 3))
        out[model.stack_offset(l, 0, 0): model.stack_offset(l, 0, 0) + n * n] = blk.ravel()
    return out, off


def test_smooth_stack_brute_force(oracle_mod):
    L = 4
    rng = np.random.default_rng(3)
    fine = rng.random(1 << (2 * L))
    st = np.zeros(model.nstack(L))
    st[model.stack_offset(L, 0, 0):] = fine
    oracle_mod.lib().or_smooth_stack(oracle_mod.d(st), L)
    # or_smooth_stack works in place only through a real pointer
    buf = np.ascontiguousarray(st)
    oracle_mod.lib().or_smooth_stack(buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), L)
    expect, _ = brute_stack(fine, L)
    np.testing.assert_allclose(buf, expect, rtol=1e-13)
    assert model.stack_offset(2, 1, 3) == 1 + 4 + 1 * 4 + 3


def cut_mask(ref, thresh, L):
    """Nodes where the descent stops, found level by level (not recursively):
    a node is a stop node if every ancestor has ref >= thresh and the node is at
    level L or has ref < thresh."""
    mask = np.zeros(len(ref), bool)
    alive = np.array([True])
    for l in range(L + 1):
        n = 1 << l
        o = model.stack_offset(l, 0, 0)
        r = ref[o:o + n * n].reshape(n, n)
        stop = alive & ((r < thresh) | (l == L))
        mask[o:o + n * n] = stop.ravel()
        cont = alive & ~stop
        alive = np.repeat(np.repeat(cont, 2, 0), 2, 1)
    return mask


@pytest.mark.parametrize("flag", [0, 1, 2])
def test_stack_difference_brute_force(flag, oracle_mod):
    w = small()
    o = oracle_mod.Oracle(w, set_particles=False)
    p = w.params
    L = 3
    rng = np.random.default_rng(10 + flag)
    ns = model.nstack(L)

    def mk():
        s = np.zeros(ns)
        s[model.stack_offset(L, 0, 0):] = rng.random(1 << (2 * L)) * (rng.random(1 << (2 * L)) > 0.2)
        return model.smooth_stack(s, L)
    d1, d2, r1, r2, dist = mk(), mk(), mk(), mk(), mk() + 0.5
    thresh = 0.9
    lib = oracle_mod.lib()
    fn = lib.or_calc_stack_difference_mod if flag == 2 else lib.or_calc_stack_difference
    D = oracle_mod.d
    got = fn(o.h, D(d1), D(d2), 0, 0, 0, L, D(r1), D(r2), thresh, D(dist), flag)
    m = cut_mask(r1, thresh, L)
    lim = np.minimum(dist, p.R200 / 20.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        if flag == 0:
            terms = np.abs(d1 - d2) / lim
        elif flag == 1:
            ok = (r1 > 0) & (r2 > 0)
            terms = np.where(ok, np.abs(d1 / r1 - d2 / r2) / np.maximum(d2 / r2, p.LowerDispLimit), 0.0)
        else:
            terms = np.where(r2 > 0, np.abs(d1 - d2) * (r1 / r2) / lim, 0.0)
    assert got == pytest.approx(terms[m].sum(), rel=1e-12)
    # the stop nodes partition the fine grid
    L2 = L
    cover = np.zeros((1 << L2, 1 << L2), int)
    for l in range(L2 + 1):
        n = 1 << l
        o2 = model.stack_offset(l, 0, 0)
        for idx in np.nonzero(m[o2:o2 + n * n])[0]:
            i, j = divmod(idx, n)
            f = 1 << (L2 - l)
            cover[i * f:(i + 1) * f, j * f:(j + 1) * f] += 1
    assert (cover == 1).all()


# --------------------------------------------------------------------- merit --

def test_merit_identity_and_global_fit(oracle_mod):
    """F_try == F_base when the trial is the current velocity; S of the global
    fit equals the cut-mask sum computed independently (main.c:294-299)."""
    w = small(n=600)
    o = oracle_mod.Oracle(w)
    o.calc_all_response_fields(4)
    idx = np.arange(0, 600, 37).astype(np.int32)
    fb, ft = o.eval_trials(idx, w.vel[idx], 4)
    np.testing.assert_allclose(ft, fb, rtol=1e-15, atol=0)
    p = w.params
    L = p.DG_MaxLevel
    R = o.stack(1, 0)
    T = o.stack(1, 10)
    thr = p.MinParticlesPerBinForDensityMeasurement * p.MType[1] / p.NType[1]
    m = cut_mask(T, thr, L)
    dist = o.distance()
    S = (np.abs(R - T) / np.minimum(dist, p.R200 / 20))[m].sum() / T[0]
    assert o.fit(1)[0] == pytest.approx(S, rel=1e-12)
    # the mass response is the sum of every particle's orbit-averaged mass
    off = model.stack_offset(L, 0, 0)
    assert R[off:].sum() == pytest.approx(p.MType[1], rel=1e-12)


def test_rep_accepts_only_improvements(oracle_mod):
    w = small(n=600)
    o = oracle_mod.Oracle(w)
    o.calc_all_response_fields(4)
    perm = model.permutation(w, 1)
    idx = perm[:150]
    v = model.trial_velocities(w, idx, 1, 0)
    fb, ft = o.eval_trials(idx, v, 4)
    acc, dF = o.rep(idx, v, 4)
    np.testing.assert_array_equal(acc, (ft < fb).astype(np.int32))
    np.testing.assert_allclose(dF, ft - fb, rtol=0, atol=0)
    vel, _, _ = o.particles()
    np.testing.assert_array_equal(vel[idx[acc == 1]], v[acc == 1])
    np.testing.assert_array_equal(vel[idx[acc == 0]], w.vel[idx[acc == 0]])


# ---------------------------------------------------------------------- tree --

def plummer_sources(n, seed=5):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 3)) * np.array([5, 5, 1])
    m = np.full(n, 1.0 / n)
    return x, m


def test_tree_theta_zero_is_direct_sum(oracle_mod):
    p = model.make_params(model.config("halo-1e5"))
    p.ErrTolTheta = 1e-6
    x, m = plummer_sources(300)
    t = np.random.default_rng(6).normal(size=(40, 3)) * 4
    lib = oracle_mod.lib()
    D = oracle_mod.d
    acc = np.zeros((40, 3))
    pot = np.zeros(40)
    acc2 = np.zeros((40, 3))
    pot2 = np.zeros(40)
    lib.or_tree_forces(ctypes.byref(p), 300, D(x), D(m), 40, D(t), acc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                       pot.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1)
    lib.or_direct_forces(ctypes.byref(p), 300, D(x), D(m), 40, D(t),
                         acc2.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                         pot2.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1)
    np.testing.assert_allclose(acc, acc2, rtol=1e-12, atol=1e-14 * np.abs(acc2).max())
    np.testing.assert_allclose(pot, pot2, rtol=1e-12)
    # Newtonian beyond the softening: a single unit mass gives -G/r^2 and -G/r
    one = np.array([[0.0, 0, 0]])
    tt = np.array([[3.0, 4.0, 0.0]])
    a1 = np.zeros((1, 3))
    p1 = np.zeros(1)
    lib.or_direct_forces(ctypes.byref(p), 1, D(one), D(np.ones(1)), 1, D(tt),
                         a1.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                         p1.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1)
    np.testing.assert_allclose(a1[0], -p.G * np.array([3, 4, 0]) / 125.0, rtol=1e-13)
    assert p1[0] == pytest.approx(-p.G / 5.0, rel=1e-13)


def test_softening_kernel_consistency(oracle_mod):
    """Inside h the spline force is -dPhi/dr of the spline potential, and both join
    the Newtonian forms at r = h (forcetree_walk.c:160-183)."""
    p = model.make_params(model.config("halo-1e5"))
    h = p.ForceSoftening
    lib = oracle_mod.lib()
    D = oracle_mod.d
    one = np.zeros((1, 3))

    def fp(r):
        a = np.zeros((1, 3))
        pt = np.zeros(1)
        lib.or_direct_forces(ctypes.byref(p), 1, D(one), D(np.ones(1)), 1, D(np.array([[r, 0, 0]])),
                             a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             pt.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1)
        return a[0, 0], pt[0]
    for r in (0.2 * h, 0.45 * h, 0.7 * h, 0.95 * h):
        e = 1e-6 * h
        dphi = (fp(r + e)[1] - fp(r - e)[1]) / (2 * e)
        assert fp(r)[0] == pytest.approx(-dphi, rel=1e-5)   # a = -dPhi/dr, pointing inwards
    assert fp(h * (1 - 1e-12))[1] == pytest.approx(-p.G / h, rel=1e-9)
    assert fp(h * (1 - 1e-12))[0] == pytest.approx(-p.G / h ** 2, rel=1e-8)


def test_tree_monopole_accuracy(oracle_mod):
    """theta = 0.4 monopole tree: relative force error small against direct summation."""
    p = model.make_params(model.config("halo-1e5"))
    p.ErrTolTheta = 0.4
    x, m = plummer_sources(3000)
    t = np.random.default_rng(7).normal(size=(64, 3)) * 6
    lib = oracle_mod.lib()
    D = oracle_mod.d
    P = ctypes.POINTER(ctypes.c_double)
    acc = np.zeros((64, 3))
    pot = np.zeros(64)
    acc2 = np.zeros((64, 3))
    pot2 = np.zeros(64)
    nint = lib.or_tree_forces(ctypes.byref(p), 3000, D(x), D(m), 64, D(t), acc.ctypes.data_as(P),
                              pot.ctypes.data_as(P), 1)
    lib.or_direct_forces(ctypes.byref(p), 3000, D(x), D(m), 64, D(t), acc2.ctypes.data_as(P),
                         pot2.ctypes.data_as(P), 1)
    err = np.linalg.norm(acc - acc2, axis=1) / np.linalg.norm(acc2, axis=1)
    assert np.median(err) < 5e-3
    assert nint < 64 * 3000   # fewer interactions than direct summation


# ------------------------------------------------- first fit and the F_base terms --

def composite_small(n=1500, **kw):
    """Halo (beta model, structure 1) + disk (structure 3) + bulge (structure 0)."""
    kw.setdefault("max_steps", 2000000)
    return model.make_workload(model.config("composite-4e6", **kw), n=n, target_samples=60000)


def test_first_fit_sall_is_twice_s(oracle_mod):
    """At the first calc_global_fit, Srelfac = S / (sum Sdisp / count) (main.c:350-353), so the
    logged Sall = S + sum(Srelfac * Sdisp) / count (main.c:1270) is exactly 2 S for every type,
    whatever the velocity structure.  A dropped Srelfac factor, a wrong count or a misplaced
    scaling of the Sdisp terms breaks it."""
    w = composite_small()
    o = oracle_mod.Oracle(w)
    o.calc_all_response_fields(4)
    for t in w.types:
        f = o.fit(t)
        S, Srelfac, count, Sall = f[0], f[5], f[6], f[7]
        assert S > 0 and Srelfac > 0
        assert count == (3.0 if w.params.TypeOfVelocityStructure[t] in (0, 1) else 4.0)
        assert Sall == pytest.approx(2 * S, rel=1e-14, abs=0), (t, Sall, S)
        # the scaled dispersion terms average to S
        assert (f[1] + f[2] + f[3] + f[4]) / count == pytest.approx(S, rel=1e-14)


def test_fbase_instantaneous_terms_equal_global_fit(oracle_mod):
    """F_base of an unchanged velocity, with the orbit-averaged S weights set to zero, is the
    density term plus Srelfac/count times the instantaneous terms of the current state
    (main.c:1134-1216).  Those are the global fit's unscaled Sdisp values (main.c:327-346):
    two code paths of the oracle (the 4-cell update of a zero difference, re-smoothed, against
    the global stacks) must agree.  The V_q rules differ: the merit drops q for structure 0
    or for the disk (main.c:1195), the fit for structures 0 and 1 (main.c:344-346), so the halo's
    (structure 1) q term is evaluated separately from the global stacks."""
    w = composite_small()
    p = w.params
    for t in range(4):
        p.fac_value_rs[t] = p.fac_value_ts[t] = p.fac_value_qs[t] = p.fac_value_ps[t] = 0.0
    o = oracle_mod.Oracle(w)
    o.calc_all_response_fields(4)
    lib = oracle_mod.lib()
    D = oracle_mod.d
    L = p.DG_MaxLevel
    for t in w.types:
        idx = np.nonzero(w.ptype == t)[0][:5].astype(np.int32)
        fb, ft = o.eval_trials(idx, w.vel[idx], 2)
        np.testing.assert_array_equal(fb, ft)
        f = o.fit(t)
        S, Srelfac, count = f[0], f[5], f[6]
        dens = S * o.stack(t, 10)[0] / p.MType[t]
        inst = (f[1] + f[2] + f[3]) / count   # scaled r, t, p terms
        tovs = p.TypeOfVelocityStructure[t]
        if tovs == 1:
            # q of the halo: not in the fit, in the merit
            th = p.MinParticlesPerBinForDispersionMeasurement * p.MType[t] / p.NType[t]
            Iq, Eq, MR, MT = o.stack(t, 5), o.stack(t, 15), o.stack(t, 1), o.stack(t, 11)
            vq = lib.or_calc_stack_difference(o.h, D(Iq), D(Eq), 0, 0, 0, L, D(MR), D(MT), th, None, 1)
            assert vq > 0
            inst += Srelfac * vq / count
        want = dens + inst
        np.testing.assert_allclose(fb, want, rtol=1e-12)
        if tovs == 3:
            # the fit counts the disk's q term, the merit does not: the q term is not zero here
            assert f[4] > 0
