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
