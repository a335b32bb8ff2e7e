"""The reference's output files written from the device state, read back by
independent readers written from the formats (main.c:1275-1283; io.c:36-457;
io.c:460-1240 with allvars.h:998-1031), and compared with the state they came from."""
import struct

import numpy as np
import pytest

from galic_amd import model
from galic_amd.runtime import GalicGPU

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def state():
    w = model.make_workload(model.config("halo-disk-1e6"), n=3000, target_samples=20000)
    g = GalicGPU(w)
    g.calc_all_response_fields()
    perm = model.permutation(w, 1)[:500]
    acc, _ = g.rep(perm, model.trial_velocities(w, perm, 5, 0))
    yield w, g, perm, acc
    g.close()


def test_fit_lines_are_the_reference_format(state, tmp_path):
    w, g, perm, acc = state
    vel, _, _ = g.particles()
    for t in w.types:
        path = tmp_path / f"fit_{t}.txt"
        sel = w.ptype[perm] == t
        tries, changes = int(sel.sum()), int(acc[sel].sum())
        g.write_fit_line(t, path, 500 / w.n, tries, changes)
        g.write_fit_line(t, path, 500 / w.n, tries, changes)
        lines = path.read_text().splitlines()
        assert len(lines) == 2 and lines[0] == lines[1]
        f = g.fit(t)
        m = w.ptype == t
        vavg = np.sqrt((vel[m] ** 2).sum() / m.sum())
        S, sr, st, sp, sq, relc = f[0], f[1], f[2], f[3], f[4], f[6]
        expect = "%10.7g  %10.7g  %10.7g     %10.7g %10.7g %10.7g %10.7g     %10.7g    %10.7g" % (
            500 / w.n, S, vavg, sr, st, sp, sq, S + (sr + st + sp + sq) / relc, changes / (tries + 1e-60))
        assert lines[0] == expect
        assert float(lines[0].split()[7]) == pytest.approx(f[7], rel=1e-6)   # the logged Sall


def smoothed(din, L, ref, thresh):
    """calc_smoothed_stack (grid.c:1217-1244), recursion as written there."""
    nb = 1 << L
    out = np.zeros((nb, nb))

    def ev(l, i, j):
        if l == 0:
            return din[model.stack_offset(0, i, j)] / (1 << (L - l)) ** 2
        if ref[model.stack_offset(l - 1, i // 2, j // 2)] > thresh:
            return din[model.stack_offset(l, i, j)] / (1 << (L - l)) ** 2
        return ev(l - 1, i // 2, j // 2)

    for i in range(nb):
        for j in range(nb):
            out[i, j] = ev(L, i, j)
    return out


def test_densfield_file(state, tmp_path):
    w, g, _, _ = state
    p = w.params
    L = p.DG_MaxLevel
    nb = 1 << L
    for t in w.types:
        path = tmp_path / f"densfield_{t}_000.dat"
        g.write_densfield(t, path)
        buf = path.read_bytes()
        pos = 0

        def rint():
            nonlocal pos
            v = struct.unpack_from("<i", buf, pos)[0]
            pos += 4
            return v

        def rdoubles(k):
            nonlocal pos
            a = np.frombuffer(buf, dtype="<f8", count=k, offset=pos)
            pos += 8 * k
            return a

        assert rint() == L
        dmt = g.stack(t, 10)
        for lev in range(L, -1, -1):
            n = rint()
            assert n == 1 << lev
            np.testing.assert_array_equal(rdoubles(n * n), dmt[model.stack_offset(lev, 0, 0):][:n * n])
        thr_d = p.MinParticlesPerBinForDensityMeasurement * p.MType[t] / p.NType[t]
        thr_s = p.MinParticlesPerBinForDispersionMeasurement * p.MType[t] / p.NType[t]
        dmr = g.stack(t, 0)
        for din in (dmt, dmr):
            assert rint() == nb
            np.testing.assert_allclose(rdoubles(nb * nb), smoothed(din, L, dmt, thr_d).ravel(), rtol=1e-15)
        assert rint() == nb
        np.testing.assert_array_equal(rdoubles(nb * nb), dmr[model.stack_offset(L, 0, 0):])
        for rmin, fac in ((p.DG_Rmin, p.DG_Fac), (p.EG_Rmin, p.EG_Fac)):
            k = np.arange(nb)
            c = rmin * 0.5 * (fac ** k + fac ** (k + 1) - 2.0)
            assert rint() == nb
            np.testing.assert_allclose(rdoubles(nb * nb), np.tile(c, nb), rtol=1e-14)
            assert rint() == nb
            np.testing.assert_allclose(rdoubles(nb * nb), np.repeat(c, nb), rtol=1e-14)
        emr = g.stack(t, 1)
        for tgt, resp in ((11, 1), (12, 2), (13, 3), (14, 4), (15, 5)):
            for which in (tgt, resp):
                assert rint() == nb
                np.testing.assert_allclose(rdoubles(nb * nb), smoothed(g.stack(t, which), L, emr, thr_s).ravel(),
                                           rtol=1e-15)
        assert pos == len(buf)


HEADER = struct.Struct("<6i6d2d2i6I2i4d2i6I2iif2i40x")


@pytest.mark.parametrize("double_precision", [False, True])
def test_gadget_snapshot(state, tmp_path, double_precision):
    w, g, _, _ = state
    path = tmp_path / "snap_000"
    vt = np.random.default_rng(1).normal(size=(w.n, 3))
    g.write_snapshot(path, double_precision=double_precision, vel_theo=vt)
    buf = path.read_bytes()
    assert HEADER.size == 256
    assert struct.unpack_from("<i", buf, 0)[0] == 256 and struct.unpack_from("<i", buf, 260)[0] == 256
    h = HEADER.unpack_from(buf, 4)
    npart = np.array(h[0:6])
    assert h[6:12] == (0.0,) * 6                     # MassTable: masses stored per particle
    assert list(npart) == [int((w.ptype == t).sum()) for t in range(6)]
    assert h[-1] == 0 and h[-5] == int(double_precision)   # flag_doubleprecision before the lpt fields
    order = np.concatenate([np.flatnonzero(w.ptype == t) for t in range(6)])
    fl = "<f8" if double_precision else "<f4"
    fsz = 8 if double_precision else 4
    pos = 264
    n = w.n
    vel, _, _ = g.particles()

    def block(nbytes, dtype, count):
        nonlocal pos
        assert struct.unpack_from("<i", buf, pos)[0] == nbytes
        a = np.frombuffer(buf, dtype=dtype, count=count, offset=pos + 4)
        assert struct.unpack_from("<i", buf, pos + 4 + nbytes)[0] == nbytes
        pos += nbytes + 8
        return a

    cast = (lambda a: a) if double_precision else (lambda a: a.astype(np.float32))
    np.testing.assert_array_equal(block(3 * fsz * n, fl, 3 * n).reshape(n, 3), cast(w.pos[order]))
    np.testing.assert_array_equal(block(3 * fsz * n, fl, 3 * n).reshape(n, 3), cast(vel[order]))
    np.testing.assert_array_equal(block(4 * n, "<u4", n), order + 1)
    np.testing.assert_array_equal(block(fsz * n, fl, n), cast(w.mass[order]))
    np.testing.assert_array_equal(block(3 * fsz * n, fl, 3 * n).reshape(n, 3), cast(vt[order]))
    assert pos == len(buf)
