"""ctypes binding of the CPU oracle (oracle/liboracle.so) for tests and the CPU baseline.

The oracle is the checker: tests compare the product (libassist.so) against it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)


class OrcOpts(C.Structure):
    _fields_ = [("forces", C.c_int), ("gr_eih_sources", C.c_int), ("geocentric", C.c_int),
                ("alpha", C.c_double), ("nk", C.c_double), ("nm", C.c_double), ("nn", C.c_double), ("r0", C.c_double)]


class OrcIas15Cfg(C.Structure):
    _fields_ = [("epsilon", C.c_double), ("min_dt", C.c_double), ("dt", C.c_double),
                ("exact_finish", C.c_int), ("max_steps", C.c_long), ("pc_tol", C.c_double),
                ("pc_tol1", C.c_double), ("pc_aref", C.c_double)]


class OrcEphem(C.Structure):
    _fields_ = [("jd_ref", C.c_double), ("planets", C.c_void_p), ("asteroids", C.c_void_p), ("ascii", C.c_void_p),
                ("target_index", C.c_int * 11), ("emb_index", C.c_int)] + [
        (k, C.c_double) for k in ("AU", "EMRAT", "J2E", "J3E", "J4E", "J2SUN", "RE", "CLIGHT", "ASUN",
                                  "Re_eq", "Rs_eq", "c_AU_per_day", "c_squared", "over_c_squared")]


class OrcSys(C.Structure):
    _fields_ = [("n_real", C.c_int), ("n_var", C.c_int), ("var_tp", C.c_void_p), ("t", C.c_double),
                ("p", C.c_void_p), ("params", C.c_void_p), ("status", C.c_int)]


class OrcStats(C.Structure):
    _fields_ = [("steps", C.c_long), ("rejected", C.c_long), ("iterations", C.c_long), ("evals", C.c_long),
                ("pc_maxed", C.c_long), ("status", C.c_int), ("dt_last_done", C.c_double), ("dt_next", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_ephem_load.restype = C.POINTER(OrcEphem)
        L.orc_ephem_load.argtypes = [C.c_char_p, C.c_char_p, _ip]
        L.orc_ephem_free.argtypes = [C.POINTER(OrcEphem)]
        L.orc_all_ephem.argtypes = [C.POINTER(OrcEphem), C.c_int, C.c_double, _dp]
        L.orc_spk_eval_target.argtypes = [C.POINTER(OrcEphem), C.c_int, C.c_int, C.c_double, _dp]
        L.orc_time_bounds.argtypes = [C.POINTER(OrcEphem), _dp, _dp]
        L.orc_opts_default.argtypes = [C.POINTER(OrcOpts)]
        L.orc_ias15_cfg_default.argtypes = [C.POINTER(OrcIas15Cfg)]
        L.orc_ias15_tables.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_batch_forces.argtypes = [C.POINTER(OrcEphem), C.POINTER(OrcOpts), C.c_int, C.c_int, _dp, _dp, _dp, _dp, _ip]
        L.orc_batch_integrate.argtypes = [C.POINTER(OrcEphem), C.POINTER(OrcOpts), C.POINTER(OrcIas15Cfg), C.c_int, C.c_int,
                                          _dp, _dp, _dp, _dp, C.c_int, _ip, _lp, _lp, _lp, _dp]
        L.orc_integrate.argtypes = [C.POINTER(OrcEphem), C.POINTER(OrcOpts), C.POINTER(OrcIas15Cfg), C.POINTER(OrcSys),
                                    C.c_double, C.POINTER(OrcStats)]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


class Oracle:
    def __init__(self, planets, asteroids):
        err = C.c_int(0)
        self.e = lib().orc_ephem_load(planets.encode(), asteroids.encode() if asteroids else None, C.byref(err))
        if not self.e:
            raise RuntimeError("oracle ephemeris load failed: %d" % err.value)

    def __del__(self):
        if getattr(self, "e", None):
            lib().orc_ephem_free(self.e)
            self.e = None

    @property
    def ephem(self):
        return self.e.contents

    def all_ephem(self, body, t):
        out = np.zeros(10)
        f = lib().orc_all_ephem(self.e, body, t, _d(out))
        return f, out

    def spk_eval(self, which, idx, t):
        out = np.zeros(9)
        f = lib().orc_spk_eval_target(self.e, which, idx, t, _d(out))
        return f, out

    def time_bounds(self):
        a, b = C.c_double(), C.c_double()
        lib().orc_time_bounds(self.e, C.byref(a), C.byref(b))
        return a.value, b.value

    @staticmethod
    def opts(**kw):
        o = OrcOpts()
        lib().orc_opts_default(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    @staticmethod
    def cfg(**kw):
        c = OrcIas15Cfg()
        lib().orc_ias15_cfg_default(C.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        return c

    def forces(self, t, states, params=None, opts=None):
        """states (n, 1+n_var, 6) -> acc (n, 1+n_var, 3), status (n,)."""
        states = np.ascontiguousarray(states, dtype=np.float64)
        n, nv, _ = states.shape
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(t, dtype=np.float64), (n,)))
        acc = np.zeros((n, nv, 3))
        st = np.zeros(n, dtype=np.int32)
        o = opts or self.opts()
        pp = None
        if params is not None:
            params = np.ascontiguousarray(params, dtype=np.float64)
            pp = _d(params)
        lib().orc_batch_forces(self.e, C.byref(o), n, nv - 1, _d(t), _d(states), pp, _d(acc), st.ctypes.data_as(_ip))
        return acc, st

    def integrate(self, t0, t1, states, params=None, opts=None, cfg=None, nthreads=0):
        states = np.array(states, dtype=np.float64, copy=True, order="C")
        n, nv, _ = states.shape
        t0 = np.ascontiguousarray(np.broadcast_to(np.asarray(t0, dtype=np.float64), (n,)))
        t1 = np.ascontiguousarray(np.broadcast_to(np.asarray(t1, dtype=np.float64), (n,)))
        o = opts or self.opts()
        c = cfg or self.cfg()
        st = np.zeros(n, dtype=np.int32)
        steps = np.zeros(n, dtype=np.int64)
        evals = np.zeros(n, dtype=np.int64)
        rej = np.zeros(n, dtype=np.int64)
        tout = np.zeros(n)
        pp = None
        if params is not None:
            params = np.ascontiguousarray(params, dtype=np.float64)
            pp = _d(params)
        lib().orc_batch_integrate(self.e, C.byref(o), C.byref(c), n, nv - 1, _d(t0), _d(t1), _d(states), pp, nthreads,
                                  st.ctypes.data_as(_ip), steps.ctypes.data_as(_lp), evals.ctypes.data_as(_lp),
                                  rej.ctypes.data_as(_lp), _d(tout))
        return dict(states=states, status=st, steps=steps, evals=evals, rejected=rej, t=tout)


    def integrate_simulation(self, t0, t1, reals, opts=None, cfg=None):
        """One reference-style simulation: the real particles reals (N, 6) share one IAS15 step
        (no variational particles), from t0 to t1; returns the final states and the stats."""
        reals = np.asarray(reals, dtype=np.float64)
        n = reals.shape[0]
        p = np.zeros((n, 9))
        p[:, :6] = reals
        s = OrcSys(n_real=n, n_var=0, var_tp=None, t=float(t0), p=p.ctypes.data, params=None, status=0)
        st = OrcStats()
        c = cfg or self.cfg()
        lib().orc_integrate(self.e, C.byref(opts or self.opts()), C.byref(c), C.byref(s), float(t1), C.byref(st))
        return p[:, :6].copy(), st


def ias15_tables():
    h, rr, c, d = np.zeros(8), np.zeros(28), np.zeros(21), np.zeros(21)
    lib().orc_ias15_tables(_d(h), _d(rr), _d(c), _d(d))
    return h, rr, c, d
