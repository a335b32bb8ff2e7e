"""Host driver over the C ABI: one model resident on one GPU.

Mirrors the reference optimiser's control flow (main.c:83-110): an initial
response pass, then per iteration a number of reps, each a batch of
independent trials followed by a commit and a global fit.  All arithmetic runs
in libgalic_hip.so on the GPU; this module only moves arrays across the ABI.
"""
import ctypes
import os

import numpy as np

from . import _abi
from ._abi import check, dptr, iptr, lptr


class GalicGPU:
    """A model on one GPU (the reference's globals, device resident)."""

    def __init__(self, w, device=0):
        self.w = w
        self.lib = _abi.lib()
        self.h = ctypes.c_void_p()
        check(self.lib.galic_create(ctypes.byref(w.params), device, ctypes.byref(self.h)))
        self.L = w.params.DG_MaxLevel
        self.ngrid = 1 << (2 * self.L)
        self.nstack = ((1 << (2 + 2 * self.L)) - 1) // 3
        check(self.lib.galic_set_force_tables(self.h, dptr(w.pot), dptr(w.dpotdr), dptr(w.dpotdz)))
        for t in (1, 2, 3):
            check(self.lib.galic_set_vstream(self.h, t, dptr(np.ascontiguousarray(w.vstream[t]))))
        for t, st in w.targets.items():
            st = [np.ascontiguousarray(s, dtype=np.float64) for s in st]
            check(self.lib.galic_set_targets(self.h, t, *[dptr(s) for s in st]))
        self.set_particles(w.pos, w.vel)
        self.set_dispersions(w.disp)

    def set_targets(self, t, stacks):
        """DGs_MassTarget, EGs_MassTarget, EGs_EgyTarget_{r,t,p,q} of type t (6 stacks)."""
        st = [np.ascontiguousarray(s, dtype=np.float64) for s in stacks]
        check(self.lib.galic_set_targets(self.h, t, *[dptr(s) for s in st]))

    def set_dispersions(self, disp):
        disp = np.ascontiguousarray(disp, dtype=np.float64)
        check(self.lib.galic_set_dispersions(self.h, len(disp), dptr(disp)))

    def write_fit_line(self, t, path, opt_frac, tries, changes):
        """Appends one fit_<type>.txt line (main.c:1221-1288)."""
        check(self.lib.galic_write_fit_line(self.h, t, os.fsencode(path), float(opt_frac), int(tries), int(changes)))

    def write_densfield(self, t, path):
        """densfield_<type>_<num>.dat (io.c:36-457)."""
        check(self.lib.galic_write_densfield(self.h, t, os.fsencode(path)))

    def write_snapshot(self, path, double_precision=False, vel_theo=None):
        """One-file Gadget format-1 snapshot (io.c:460-1240)."""
        vt = None if vel_theo is None else np.ascontiguousarray(vel_theo, dtype=np.float64)
        check(self.lib.galic_write_snapshot(self.h, os.fsencode(path), int(bool(double_precision)), dptr(vt)))

    def set_disp_tables(self, t, beta=0.0, disp_r=None, disp_z=None, disp_phi=None, tilt_r=None, tilt_theta=None):
        """Target dispersion tables of type t as get_disp_rtp reads them (disp_fields.c:66-198)."""
        tabs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel()
                for a in (disp_r, disp_z, disp_phi, tilt_r, tilt_theta)]
        check(self.lib.galic_set_disp_tables(self.h, t, float(beta), *[dptr(a) for a in tabs]))

    def sample_targets(self, t, nsamples, seed=1, stream=None, stretch=1.0, density=True, energy=True, points=False):
        """Sampled DGs_MassTarget and EGs targets of type t (grid.c:536-617, 941-1076).

        Returns (dgs_mass_target or None, [5, Nstack] energy targets or None, points binned, kernel ms), and the
        [nsamples, 3] sample points as a fifth item if points is set."""
        dg = np.zeros(self.nstack) if density else None
        eg = np.zeros((5, self.nstack)) if energy else None
        pts = np.zeros((int(nsamples), 3)) if points else None
        nb = np.zeros(1, dtype=np.int64)
        ms = ctypes.c_float(0)
        stream = 100 + t if stream is None else stream
        check(self.lib.galic_sample_targets(self.h, t, int(nsamples), int(seed) & 0xFFFFFFFF, int(stream) & 0xFFFFFFFF,
                                            float(stretch), dptr(dg), dptr(eg), dptr(pts), lptr(nb),
                                            ctypes.byref(ms)))
        if points:
            return dg, eg, int(nb[0]), float(ms.value), pts
        return dg, eg, int(nb[0]), float(ms.value)

    def set_particles(self, pos, vel):
        w = self.w
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        check(self.lib.galic_set_particles(self.h, len(w.mass), dptr(pos), dptr(vel), dptr(w.mass), dptr(w.tint),
                                           iptr(w.ptype), dptr(w.vesc)))

    def close(self):
        if self.h:
            self.lib.galic_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------- FP64 path --
    def calc_all_response_fields(self):
        check(self.lib.galic_calc_all_response_fields(self.h))

    def rep(self, idx, vtry):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        vtry = np.ascontiguousarray(vtry, dtype=np.float64)
        acc = np.zeros(len(idx), dtype=np.int32)
        dF = np.zeros(len(idx))
        check(self.lib.galic_rep(self.h, len(idx), iptr(idx), dptr(vtry), iptr(acc), dptr(dF)))
        return acc, dF

    # ------------------------------------------------------------- FP32 path --
    def fp32_calc_all_response_fields(self):
        check(self.lib.galic_fp32_calc_all_response_fields(self.h))

    def fp32_rep(self, idx, k=1, vtry=None, sigma_frac=0.1, seed=1):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        acc = np.zeros(len(idx), dtype=np.int32)
        vt = None if vtry is None else np.ascontiguousarray(vtry, dtype=np.float64)
        check(self.lib.galic_fp32_rep(self.h, len(idx), iptr(idx), k, dptr(vt), sigma_frac, seed, iptr(acc)))
        return acc

    def fp32_iterate(self, idx, k=1, rep_size=100000, vtry=None, sigma_frac=0.1, seed=1):
        """Trials in reps of rep_size, each rep committed before the next."""
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        acc = np.zeros(len(idx), dtype=np.int32)
        vt = None if vtry is None else np.ascontiguousarray(vtry, dtype=np.float64)
        check(self.lib.galic_fp32_iterate(self.h, len(idx), iptr(idx), k, int(rep_size), dptr(vt), sigma_frac, seed,
                                          iptr(acc)))
        return acc

    def fp32_autotune_schedule(self, idx, k=1, rep_size=100000, sigma_frac=0.1, seed=1):
        """fp32_iterate (device-drawn trials, the same results) in six runs of reps, two for each
        of three sharings of the CUs between orbit launches and judging; keeps the fastest.
        Returns (accept, orbit workgroups per CU beside the judging, merit workgroups per CU)."""
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        acc = np.zeros(len(idx), dtype=np.int32)
        o = np.zeros(1, dtype=np.int32)
        m = np.zeros(1, dtype=np.int32)
        check(self.lib.galic_fp32_autotune_schedule(self.h, len(idx), iptr(idx), k, int(rep_size), sigma_frac, seed,
                                                    iptr(acc), iptr(o), iptr(m)))
        return acc, int(o[0]), int(m[0])

    def fp32_trial_merits(self, ntrial, k):
        """F_try - F_base of every trial of the last fp32_rep / fp32_iterate call, shape (ntrial, k)."""
        out = np.zeros((ntrial, k))
        check(self.lib.galic_fp32_trial_merits(self.h, int(ntrial), int(k), dptr(out)))
        return out

    def fp32_autotune(self, nsample=65536):
        v = np.zeros(1, dtype=np.int32)
        b = np.zeros(1, dtype=np.int32)
        check(self.lib.galic_fp32_autotune(self.h, nsample, iptr(v), iptr(b)))
        return int(v[0]), int(b[0])

    def fp32_set_launch(self, variant, blocks_per_cu):
        check(self.lib.galic_fp32_set_launch(self.h, int(variant), int(blocks_per_cu)))

    def fp32_set_merit_form(self, form):
        """0 automatic (4 where it applies), 1 dense 32x32 LDS tiles at every level, 2 never tiles,
        3 sparse 32x32 LDS tiles at every level, 4 16x16 tiles in each particle's int32 fixed point."""
        check(self.lib.galic_fp32_set_merit_form(self.h, int(form)))

    def fp32_set_merit_launch(self, wg_per_cu=0):
        """At most wg_per_cu workgroups per CU for the tile merit kernels (0: as many as fit)."""
        check(self.lib.galic_fp32_set_merit_launch(self.h, int(wg_per_cu)))

    def fp32_set_schedule(self, max_group_trials=0, overlap=-1, graph=1):
        """Group the FP32 reps' orbit launches at most max_group_trials trials each (0: as many as
        the log arena holds); overlap 1 integrates the next group while a group is judged, 0 never,
        -1 when a call needs more than one launch; graph 1 captures a group's reps in one hipGraph
        (16x16-tile merit), 0 launches them rep by rep.  The results do not depend on the schedule."""
        check(self.lib.galic_fp32_set_schedule(self.h, int(max_group_trials), int(overlap), int(graph)))

    def fp32_set_fresh_trials(self, nfresh):
        """The first nfresh device trials per particle are fresh draws (main.c:383-435)."""
        check(self.lib.galic_fp32_set_fresh_trials(self.h, int(nfresh)))

    # ---------------------------------------------------------------- reads --
    def stack(self, t, which):
        out = np.zeros(self.nstack)
        check(self.lib.galic_get_stack(self.h, t, which, dptr(out)))
        return out

    def fit(self, t):
        out = np.zeros(8)
        check(self.lib.galic_get_fit(self.h, t, dptr(out)))
        return out

    def particles(self):
        n = len(self.w.mass)
        vel = np.zeros((n, 3))
        orb = np.zeros(n, dtype=np.int32)
        v = [np.zeros(n) for _ in range(4)]
        check(self.lib.galic_get_particles(self.h, dptr(vel), iptr(orb), *[dptr(x) for x in v]))
        return vel, orb, np.stack(v, -1)

    def orbit_responses(self, pos, vel, mass, tint, ptype, fields=True, precision=64):
        """One orbit per row: response fields [n][5][Ngrid], t, pericentre count,
        steps, E0, E1.  precision=32 integrates with the FP32 path."""
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        n = len(pos)
        mass = np.ascontiguousarray(np.broadcast_to(mass, (n,)), dtype=np.float64)
        tint = np.ascontiguousarray(np.broadcast_to(tint, (n,)), dtype=np.float64)
        ptype = np.ascontiguousarray(np.broadcast_to(ptype, (n,)), dtype=np.int32)
        f = np.zeros((n, 5, self.ngrid)) if fields else None
        t = np.zeros(n)
        orb = np.zeros(n, dtype=np.int32)
        steps = np.zeros(n, dtype=np.int64)
        e0 = np.zeros(n)
        e1 = np.zeros(n)
        fn = self.lib.galic_fp32_orbit_responses if precision == 32 else self.lib.galic_orbit_responses
        check(fn(self.h, n, dptr(pos), dptr(vel), dptr(mass), dptr(tint), iptr(ptype), dptr(f), dptr(t), iptr(orb),
                 lptr(steps), dptr(e0), dptr(e1)))
        return dict(fields=f, t=t, orbits=orb, steps=steps, e0=e0, e1=e1)

    def field_at(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        acc = np.zeros_like(pos)
        pot = np.zeros(len(pos))
        check(self.lib.galic_field_at(self.h, len(pos), dptr(pos), dptr(acc), dptr(pot)))
        return acc, pot

    def timings(self):
        out = np.zeros(11)
        check(self.lib.galic_timings(self.h, dptr(out)))
        keys = ["last_orbit_ms", "last_eval_ms", "last_fit_ms", "orbit_ms", "orbit_launches", "orbits", "steps",
                "eval_ms", "fit_ms", "log_relaunches", "graph_launches"]
        return dict(zip(keys, out.tolist()))

    def reset_timings(self):
        check(self.lib.galic_reset_timings(self.h))

    def bind(self):
        check(self.lib.galic_bind(self.h))

    # -------------------------------------------------------- table fill --
    def set_tree_sources(self, spos, smass):
        """Sources of the context's table fill (forcedensitygrid_calculate's samples)."""
        spos = np.ascontiguousarray(spos, dtype=np.float64).reshape(-1, 3)
        smass = np.ascontiguousarray(smass, dtype=np.float64)
        check(self.lib.galic_set_tree_sources(self.h, len(spos), dptr(spos), dptr(smass)))

    def sample_tree_sources(self, seed=42):
        check(self.lib.galic_sample_tree_sources(self.h, int(seed)))

    def retabulate(self, precision=0):
        """Re-fill the context's force tables from its tree sources on the GPU; returns the interactions."""
        n = np.zeros(1, dtype=np.int64)
        check(self.lib.galic_retabulate(self.h, int(precision), lptr(n)))
        return int(n[0])

    def force_tables(self):
        n2 = self.w.params.FG_Nbin ** 2
        pot, dr, dz = np.zeros(n2), np.zeros(n2), np.zeros(n2)
        check(self.lib.galic_get_force_tables(self.h, dptr(pot), dptr(dr), dptr(dz)))
        return pot, dr, dz


def tree_forces(params, spos, smass, tpos, precision=0, device=0):
    spos = np.ascontiguousarray(spos, dtype=np.float64).reshape(-1, 3)
    smass = np.ascontiguousarray(smass, dtype=np.float64)
    tpos = np.ascontiguousarray(tpos, dtype=np.float64).reshape(-1, 3)
    acc = np.zeros_like(tpos)
    pot = np.zeros(len(tpos))
    nint = np.zeros(1, dtype=np.int64)
    check(_abi.lib().galic_tree_forces(ctypes.byref(params), device, len(spos), dptr(spos), dptr(smass), len(tpos),
                                       dptr(tpos), dptr(acc), dptr(pot), precision, lptr(nint)))
    return acc, pot, int(nint[0])


def tree_fill(params, spos, smass, precision=0, device=0):
    spos = np.ascontiguousarray(spos, dtype=np.float64).reshape(-1, 3)
    smass = np.ascontiguousarray(smass, dtype=np.float64)
    n2 = params.FG_Nbin * params.FG_Nbin
    pot, dr, dz = np.zeros(n2), np.zeros(n2), np.zeros(n2)
    nint = np.zeros(1, dtype=np.int64)
    check(_abi.lib().galic_tree_fill(ctypes.byref(params), device, len(spos), dptr(spos), dptr(smass), dptr(pot),
                                     dptr(dr), dptr(dz), precision, lptr(nint)))
    return pot, dr, dz, int(nint[0])


# walk modes of the tree calls (include/galic_hip.h)
TREE_FP32, TREE_QUADRUPOLE, TREE_PLUMMER, TREE_GROUP = 1, 2, 4, 8


def direct_forces(params, spos, smass, tpos, device=0, plummer=False):
    """FP64 direct summation with the tree's interaction (the spline, or Plummer): (acc [n][3], pot [n])."""
    spos = np.ascontiguousarray(spos, dtype=np.float64).reshape(-1, 3)
    smass = np.ascontiguousarray(smass, dtype=np.float64)
    tpos = np.ascontiguousarray(tpos, dtype=np.float64).reshape(-1, 3)
    acc = np.zeros_like(tpos)
    pot = np.zeros(len(tpos))
    check(_abi.lib().galic_direct_forces_mode(ctypes.byref(params), device, len(spos), dptr(spos), dptr(smass),
                                              len(tpos), dptr(tpos), dptr(acc), dptr(pot),
                                              TREE_PLUMMER if plummer else 0))
    return acc, pot


def tree_last_timings():
    """Device ms of the last tree call's build and walk, and its interaction count."""
    out = np.zeros(3)
    check(_abi.lib().galic_tree_last_timings(dptr(out)))
    return dict(build_ms=out[0], walk_ms=out[1], interactions=out[2])
