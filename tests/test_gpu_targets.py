"""Sampled target maps on the GPU (grid.c:536-617, 941-1076) against the oracle.

The device draws each sample's points from the Philox numbers of its id, so its
points equal the oracle's restatement of the reference's samplers driven by the
same numbers (to the last bits of the device's log/sqrt/trig).  Binned by the
oracle, those points give the device's maps to the fixed-point resolution of
the device sums (2^-32 of a point's weight).
"""
import numpy as np
import pytest

from galic_amd import _abi, model, rng
from galic_amd.runtime import GalicGPU

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def w():
    return model.make_workload(model.config("composite-4e6"), n=4000, target_samples=20000)


def setup_tables(w, g, o):
    for t in w.types:
        beta, tabs = model.disp_tables(w.params, w.cfg, t)
        g.set_disp_tables(t, beta, **tabs)
        o.set_disp_tables(t, beta, **tabs)


def test_device_points_equal_host_draws(w, oracle_mod):
    p = w.params
    g = GalicGPU(w)
    try:
        n = 3000
        for t in w.types:
            print(f"type {t}: {n} samples", flush=True)
            _, _, _, _, pts = g.sample_targets(t, n, seed=9, stream=200 + t, energy=False, points=True)
            ref = oracle_mod.sample_points(p, t, np.arange(n, dtype=np.uint64), rng.uniforms, 9, 200 + t)
            np.testing.assert_allclose(pts, ref, rtol=1e-12, atol=1e-13 * p.Rmax)
    finally:
        g.close()


@pytest.mark.parametrize("name,kw", [("composite-4e6", {}), ("halo-disk-1e6", {"disk_tovs": 2}),
                                     ("composite-4e6", {"level": 5}), ("composite-4e6", {"level": 7})])
def test_device_targets_match_oracle_binning(name, kw, oracle_mod):
    """Velocity structures 1 and 0 (halo, bulge), 3 and 2 (disk); at levels 5, 6 and 7, whose maps
    the sampler sums in three layouts (in LDS the density map, then all five energy maps in one
    pass, at 32^2; the density map, then the energy maps in groups, at 64^2; atomics straight to
    memory at 128^2)."""
    w = model.make_workload(model.config(name, **kw), n=4000, target_samples=20000)
    p = w.params
    g = GalicGPU(w)
    o = oracle_mod.Oracle(w, set_particles=False)
    try:
        setup_tables(w, g, o)
        n = 20000
        for t in w.types:
            print(f"type {t} (velocity structure {p.TypeOfVelocityStructure[t]}): {n} samples", flush=True)
            dg, eg, nb, ms, pts = g.sample_targets(t, n, seed=4, points=True)
            rdg, reg, rn = o.bin_targets(t, pts)
            assert nb == rn
            tol = 1e-9 * p.MType[t]
            np.testing.assert_allclose(dg, rdg, rtol=1e-9, atol=tol)
            scale = np.abs(reg).max(axis=1, keepdims=True)
            assert np.all(np.abs(eg - reg) <= 1e-9 * scale)
    finally:
        o.close()
        g.close()


def test_targets_bit_identical_run_to_run(w):
    g = GalicGPU(w)
    try:
        beta, tabs = model.disp_tables(w.params, w.cfg, 2)
        g.set_disp_tables(2, beta, **tabs)
        a = g.sample_targets(2, 200000, seed=3)
        b = g.sample_targets(2, 200000, seed=3)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
    finally:
        g.close()


def test_large_sample_mass_and_shape(w):
    """At 2e7 points: every map's root holds MType, the density-grid and energy-grid mass maps agree
    (the grids coincide), and the halo's map has the shape of the analytic target to the noise."""
    p = w.params
    g = GalicGPU(w)
    try:
        beta, tabs = model.disp_tables(p, w.cfg, 1)
        g.set_disp_tables(1, beta, **tabs)
        dg, eg, nb, ms = g.sample_targets(1, 20_000_000, seed=5)
        print(f"2e7 halo samples in {ms:.1f} ms of kernel time", flush=True)
        assert dg[0] == pytest.approx(p.MType[1], rel=1e-12)
        assert eg[0][0] == pytest.approx(p.MType[1], rel=1e-12)
        if p.EG_Rmin == p.DG_Rmin and p.EG_Fac == p.DG_Fac:
            assert np.array_equal(dg, eg[0])
        ref = model.density_target(p, 1)
        L = p.DG_MaxLevel
        a = slice(model.stack_offset(L, 0, 0), None)
        expect = ref[a] / p.MType[1] * nb
        sel = expect > 5000
        ratio = dg[a][sel] / ref[a][sel]
        s = np.median(ratio)
        assert 1.0 < s < 1.15
        # the fine cells also carry the CIC spread against the analytic cell-centre rule: a few percent
        assert np.median(np.abs(ratio / s - 1)) < 0.05
    finally:
        g.close()


def test_sampling_errors(w):
    g = GalicGPU(w)
    try:
        for bad in (0, (1 << 30) + 1):
            with pytest.raises(_abi.GalicError):
                g.sample_targets(1, bad)
        with pytest.raises(_abi.GalicError):   # dispersion targets need the tables first
            g.sample_targets(1, 1000, density=False)
    finally:
        g.close()
