"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (galic_oracle.c).

Used by tests/, the smoke check and the cpu_baseline leg of bench.py as the
checker; the product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libgalic_oracle.so")

_P = ctypes.POINTER
_dp = _P(ctypes.c_double)
_ip = _P(ctypes.c_int)
_lp = _P(ctypes.c_longlong)
_vp = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    from galic_amd._abi import GalicParams
    L = ctypes.CDLL(LIB)
    pp = _P(GalicParams)
    sig = {
        "or_solve_fac": (ctypes.c_double, [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]),
        "or_grid_create": (ctypes.c_int, [pp, ctypes.c_double, ctypes.c_double]),
        "or_model_new": (_vp, [pp]),
        "or_model_free": (None, [_vp]),
        "or_set_force_tables": (None, [_vp, _dp, _dp, _dp]),
        "or_set_vstream": (None, [_vp, ctypes.c_int, _dp]),
        "or_set_targets": (None, [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]),
        "or_set_particles": (None, [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, _ip, _dp]),
        "or_cellsize": (_dp, [_vp]),
        "or_distance": (_dp, [_vp]),
        "or_stack": (_dp, [_vp, ctypes.c_int, ctypes.c_int]),
        "or_fit_value": (ctypes.c_double, [_vp, ctypes.c_int, ctypes.c_int]),
        "or_get_particles": (None, [_vp, _dp, _ip, _dp, _dp, _dp, _dp]),
        "or_forcegrid_get_cell": (None, [_vp, _dp, _ip, _ip, _dp, _dp]),
        "or_forcegrid_get_potential": (ctypes.c_double, [_vp, _dp]),
        "or_forcegrid_get_acceleration": (None, [_vp, _dp, _dp]),
        "or_densitygrid_get_cell": (None, [_vp, _dp, _ip, _ip, _dp, _dp]),
        "or_energygrid_get_cell": (None, [_vp, _dp, _ip, _ip, _dp, _dp]),
        "or_get_vstream": (ctypes.c_double, [_vp, _dp, ctypes.c_int]),
        "or_get_timestep": (ctypes.c_double, [_vp, _dp, _dp, _dp, ctypes.c_int]),
        "or_orbit_response_mod": (ctypes.c_double, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_double,
                                                    ctypes.c_double, _ip, ctypes.c_int, _lp, _dp, _dp]),
        "or_orbit_response_plain": (ctypes.c_double, [_vp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double, _ip,
                                                      _lp]),
        "or_smooth_stack": (None, [_dp, ctypes.c_int]),
        "or_calc_stack_difference": (ctypes.c_double, [_vp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                       ctypes.c_int, _dp, _dp, ctypes.c_double, _dp, ctypes.c_int]),
        "or_calc_stack_difference_mod": (ctypes.c_double, [_vp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                           ctypes.c_int, _dp, _dp, ctypes.c_double, _dp,
                                                           ctypes.c_int]),
        "or_calc_disp_components": (None, [_vp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]),
        "or_calc_all_response_fields": (None, [_vp, ctypes.c_int]),
        "or_calc_global_fit": (None, [_vp]),
        "or_rep": (None, [_vp, ctypes.c_int, _ip, _dp, _ip, _dp, ctypes.c_int]),
        "or_eval_trials": (None, [_vp, ctypes.c_int, _ip, _dp, _dp, _dp, ctypes.c_int]),
        "or_tree_forces": (ctypes.c_longlong, [pp, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp, _dp,
                                               ctypes.c_int]),
        "or_direct_forces": (None, [pp, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]),
        "or_table_fill": (ctypes.c_longlong, [pp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, ctypes.c_int]),
        "or_sample_point": (ctypes.c_int, [pp, ctypes.c_int, ctypes.c_double, _dp, _dp]),
        "or_set_disp_tables": (None, [_vp, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, _dp, _dp]),
        "or_get_disp_rtp": (None, [_vp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp]),
        "or_bin_targets": (ctypes.c_longlong, [_vp, ctypes.c_int, ctypes.c_longlong, _dp, _dp, _dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def d(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_dp)


class Oracle:
    """One oracle model (the reference's globals) built from a galic_amd Workload."""

    def __init__(self, w, set_particles=True):
        L = lib()
        self.w = w
        self.p = w.params
        self.h = L.or_model_new(ctypes.byref(w.params))
        self.L = w.params.DG_MaxLevel
        self.nb = 1 << self.L
        self.ngrid = self.nb * self.nb
        self.nstack = ((1 << (2 + 2 * self.L)) - 1) // 3
        self._keep = []
        L.or_set_force_tables(self.h, d(w.pot), d(w.dpotdr), d(w.dpotdz))
        for t in (1, 2, 3):
            L.or_set_vstream(self.h, t, d(w.vstream[t]))
        for t, st in w.targets.items():
            st = [np.ascontiguousarray(s) for s in st]
            L.or_set_targets(self.h, t, *[d(s) for s in st])
        if set_particles:
            self.set_particles(w.pos, w.vel)

    def set_particles(self, pos, vel):
        w = self.w
        L = lib()
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        L.or_set_particles(self.h, len(w.mass), d(pos), d(vel), d(w.mass), d(w.tint),
                           w.ptype.ctypes.data_as(_ip), d(w.vesc))

    def close(self):
        if self.h:
            lib().or_model_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- hot path ---
    def orbit(self, pos, vel, mass, tint, ptype):
        ng = self.ngrid
        f = np.zeros((5, ng))
        norb = ctypes.c_int(0)
        steps = ctypes.c_longlong(0)
        e0 = ctypes.c_double(0)
        e1 = ctypes.c_double(0)
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        fr = [np.zeros(ng) for _ in range(5)]
        t = lib().or_orbit_response_mod(self.h, d(pos), d(vel), *[d(x) for x in fr], mass, tint,
                                        ctypes.byref(norb), int(ptype), ctypes.byref(steps), ctypes.byref(e0),
                                        ctypes.byref(e1))
        for k in range(5):
            f[k] = fr[k]
        return dict(t=t, fields=f, orbits=norb.value, steps=steps.value, e0=e0.value, e1=e1.value)

    def calc_all_response_fields(self, nthreads=0):
        lib().or_calc_all_response_fields(self.h, nthreads)

    def rep(self, idx, vtry, nthreads=0):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        vtry = np.ascontiguousarray(vtry, dtype=np.float64)
        acc = np.zeros(len(idx), dtype=np.int32)
        dF = np.zeros(len(idx))
        lib().or_rep(self.h, len(idx), idx.ctypes.data_as(_ip), d(vtry), acc.ctypes.data_as(_ip), d(dF), nthreads)
        return acc, dF

    def eval_trials(self, idx, vtry, nthreads=0):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        vtry = np.ascontiguousarray(vtry, dtype=np.float64)
        fb = np.zeros(len(idx))
        ft = np.zeros(len(idx))
        lib().or_eval_trials(self.h, len(idx), idx.ctypes.data_as(_ip), d(vtry), d(fb), d(ft), nthreads)
        return fb, ft

    def stack(self, t, which):
        ptr = lib().or_stack(self.h, t, which)
        return np.ctypeslib.as_array(ptr, shape=(self.nstack,)).copy()

    def fit(self, t):
        return np.array([lib().or_fit_value(self.h, t, k) for k in range(8)])

    def particles(self):
        n = len(self.w.mass)
        vel = np.zeros((n, 3))
        orb = np.zeros(n, dtype=np.int32)
        v = [np.zeros(n) for _ in range(4)]
        lib().or_get_particles(self.h, d(vel), orb.ctypes.data_as(_ip), *[d(x) for x in v])
        return vel, orb, np.stack(v, -1)

    def set_disp_tables(self, t, beta=0.0, disp_r=None, disp_z=None, disp_phi=None, tilt_r=None, tilt_theta=None):
        tabs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel()
                for a in (disp_r, disp_z, disp_phi, tilt_r, tilt_theta)]
        lib().or_set_disp_tables(self.h, t, float(beta), *[d(a) for a in tabs])

    def disp_rtp(self, pos, t):
        out = [ctypes.c_double(0) for _ in range(4)]
        lib().or_get_disp_rtp(self.h, d(np.asarray(pos, float)), int(t), *[ctypes.byref(o) for o in out])
        return np.array([o.value for o in out])

    def bin_targets(self, t, pos, density=True, energy=True):
        """DGs_MassTarget [Nstack] and the EGs targets [5, Nstack] of the given points, and the count binned."""
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        dg = np.zeros(self.nstack) if density else None
        eg = np.zeros((5, self.nstack)) if energy else None
        n = lib().or_bin_targets(self.h, int(t), len(pos), d(pos), d(dg), d(eg))
        return dg, eg, int(n)

    def cellsize(self):
        return np.ctypeslib.as_array(lib().or_cellsize(self.h), shape=(self.ngrid,)).copy()

    def distance(self):
        return np.ctypeslib.as_array(lib().or_distance(self.h), shape=(self.nstack,)).copy()

    def accel(self, pos):
        out = np.zeros(3)
        lib().or_forcegrid_get_acceleration(self.h, d(np.asarray(pos, float)), d(out))
        return out

    def potential(self, pos):
        return lib().or_forcegrid_get_potential(self.h, d(np.asarray(pos, float)))


def sample_points(params, t, ids, uniforms, seed, stream, stretch=1.0, max_attempts=1000):
    """Points of a component's fresh-coordinate draw (halo.c:17-40, disk.c:17-51, bulge.c:17-38) for the
    given sample ids: round a of id i uses uniforms(seed, stream, [i], a)[0:3] in the reference's order,
    until the reference would keep the point.  Returns [len(ids), 3] (rows of zeros if no round kept one)."""
    L = lib()
    ids = np.asarray(ids, dtype=np.uint64)
    out = np.zeros((len(ids), 3))
    todo = np.arange(len(ids))
    pos = np.zeros(3)
    for attempt in range(max_attempts):
        if len(todo) == 0:
            break
        u = np.ascontiguousarray(uniforms(seed, stream, ids[todo], attempt).T)
        keep = np.zeros(len(todo), dtype=bool)
        for k in range(len(todo)):
            if L.or_sample_point(ctypes.byref(params), int(t), float(stretch), d(u[k]), d(pos)):
                out[todo[k]] = pos
                keep[k] = True
        todo = todo[~keep]
    return out

