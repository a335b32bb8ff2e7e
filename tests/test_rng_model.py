"""The synthetic-input side: the counter-based RNG against the published
Philox4x32-10 known-answer vectors (Random123, Salmon et al. 2011), and the
samplers against the analytic profiles they invert."""
import numpy as np
import pytest

from galic_amd import model
from galic_amd.rng import gaussians, philox4x32, uniforms


@pytest.mark.parametrize("ctr,key,expect", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox_known_answers(ctr, key, expect):
    out = philox4x32(*[[c] for c in ctr], *key)
    assert tuple(int(o[0]) for o in out) == expect


def test_uniform_and_gaussian_moments():
    u = uniforms(1, 2, np.arange(200000, dtype=np.uint64))
    assert (u > 0).all() and (u < 1).all()
    assert abs(u.mean() - 0.5) < 2e-3
    g = gaussians(1, 3, np.arange(200000, dtype=np.uint64))
    assert abs(g.mean()) < 5e-3 and abs(g.std() - 1) < 5e-3


def test_halo_sampler_follows_hernquist_mass():
    p = model.make_params(model.config("halo-1e5"))
    x = model.sample_halo(p, 1, 1, np.arange(100000, dtype=np.uint64))
    r = np.sqrt((x * x).sum(1))
    assert r.max() <= p.Rmax
    a = p.Halo_A
    # truncated Hernquist: P(<r) = (r/(r+a))^2 / (Rmax/(Rmax+a))^2
    for rr in (5.0, 30.0, 120.0):
        expect = (rr / (rr + a)) ** 2 / (p.Rmax / (p.Rmax + a)) ** 2
        assert abs((r < rr).mean() - expect) < 5e-3
    # isotropic: <cos theta> = 0, <cos^2 theta> = 1/3
    c = x[:, 2] / r
    assert abs(c.mean()) < 1e-2 and abs((c * c).mean() - 1 / 3) < 1e-2


def test_disk_sampler_follows_exponential_sech2():
    p = model.make_params(model.config("halo-disk-1e6"))
    x = model.sample_disk(p, 1, 2, np.arange(100000, dtype=np.uint64))
    R = np.hypot(x[:, 0], x[:, 1])
    h = p.Disk_H
    for rr in (2.0, 5.0, 10.0):
        assert abs((R < rr).mean() - (1 - (1 + rr / h) * np.exp(-rr / h))) < 5e-3
    # sech^2(z/z0): P(|z| < z) = tanh(z/z0)
    z0 = p.Disk_Z0
    assert abs((np.abs(x[:, 2]) < z0).mean() - np.tanh(1.0)) < 5e-3


def test_tint_and_vesc():
    w = model.make_workload(model.config("halo-1e5"), n=2000, target_samples=20000)
    p = w.params
    r = np.sqrt((w.pos ** 2).sum(1))
    pot = -p.G * p.Halo_Mass / (r + p.Halo_A)
    np.testing.assert_allclose(w.vesc, np.sqrt(-2 * pot), rtol=1e-12)
    vc = np.sqrt(p.G * p.Halo_Mass * r * r / ((r + 1e-6 * p.Halo_A) * (r + p.Halo_A) ** 2))
    np.testing.assert_allclose(w.tint, 4.0 * 2 * np.pi * r / vc, rtol=1e-10)   # set_particles.c:87
    assert (np.sqrt((w.vel ** 2).sum(1)) < 0.9999 * w.vesc).all()


def test_targets_normalised_and_stale_coarse():
    w = model.make_workload(model.config("halo-disk-1e6"), n=3000, target_samples=50000)
    p = w.params
    L = p.DG_MaxLevel
    off = model.stack_offset(L, 0, 0)
    for t in w.types:
        Tm = w.targets[t][0]
        assert Tm[off:].sum() == pytest.approx(p.MType[t], rel=1e-12)   # grid.c:874-877
        MT = w.targets[t][1]
        assert MT[0] == pytest.approx(MT[off:].sum(), rel=1e-12)        # smoothed
