"""galic_field_at_static (the lookups of galic_field_at with static device staging, for callers
that cannot allocate) against galic_field_at through the C ABI: bit-identical accelerations and
potentials on two calls in a row, the second with fewer points, and a clear error beyond the
static staging's size."""
import ctypes

import numpy as np
import pytest

from galic_amd import _abi, model
from galic_amd.runtime import GalicGPU

pytestmark = pytest.mark.gpu


def lookup(lib, h, name, pos):
    n = len(pos)
    acc = np.zeros((n, 3))
    pot = np.zeros(n)
    rc = getattr(lib, name)(h, n, _abi.dptr(np.ascontiguousarray(pos)), _abi.dptr(acc), _abi.dptr(pot))
    _abi.check(rc)
    return acc, pot


def test_field_at_static_equals_field_at():
    print("[gpu] galic_field_at_static vs galic_field_at", flush=True)
    w = model.make_workload(model.config("composite-4e6", max_steps=2000000), n=3000, target_samples=50000)
    g = GalicGPU(w)
    lib = _abi.lib()
    rng = np.random.default_rng(4)
    for n in (20000, 700):
        pos = np.concatenate([w.pos[rng.integers(0, w.n, n // 2)], rng.normal(size=(n - n // 2, 3)) * 40])
        a_s, p_s = lookup(lib, g.h, "galic_field_at_static", pos)
        a_r, p_r = lookup(lib, g.h, "galic_field_at", pos)
        np.testing.assert_array_equal(a_s, a_r)
        np.testing.assert_array_equal(p_s, p_r)
        assert np.isfinite(a_s).all() and (p_s < 0).all()
    big = np.zeros((65537, 3))
    rc = lib.galic_field_at_static(g.h, len(big), _abi.dptr(big), _abi.dptr(np.zeros((len(big), 3))),
                                   _abi.dptr(np.zeros(len(big))))
    assert rc == -1 and b"exceed" in lib.galic_last_error()
    g.close()
