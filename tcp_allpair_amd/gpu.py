"""ctypes binding of libtcp_gpu.so (include/tcp_gpu.h).

The library is the only compute path: importing this module without the built
library, or using it without a HIP device, raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .params import MIXED, NBIN, CParams, TcpParams

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtcp_gpu.so")

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

# exported symbols and their argument types (include/tcp_gpu.h)
SIGNATURES = {
    "tcp_create": [ctypes.POINTER(_vp), ctypes.POINTER(CParams), ctypes.c_int],
    "tcp_destroy": [_vp],
    "tcp_last_error": [_vp],
    "tcp_ewald_self": [_vp, _dp, _dp, _dp, _dp],
    "tcp_set_state": [_vp, _dp, _dp, _dp],
    "tcp_get_state": [_vp, _dp, _dp, _dp],
    "tcp_force_direct": [_vp, _dp],
    "tcp_force_direct_ext": [_vp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp],
    "tcp_force_recip": [_vp, _dp],
    "tcp_remove_rigidmtn": [_vp],
    "tcp_vverlet1": [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double],
    "tcp_vverlet2": [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                     ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _dp, _dp],
    "tcp_get_histograms": [_vp, _dp, _dp, ctypes.c_int],
    "tcp_md_steps": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                     ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _dp],
    "tcp_last_timings": [_vp, _dp, _dp],
    "tcp_vverlet2_kick": [_vp, ctypes.c_double, _dp, _dp],
    "tcp_vverlet2_scale": [_vp, ctypes.c_double, ctypes.c_double, _dp, _dp],
    "tcp_set_lambda": [_vp, ctypes.c_double],
    "tcp_energy_adjoint": [_vp, _dp, _dp, _dp, _dp, _dp, _dp],
    "tcp_force_direct_noalloc": [ctypes.POINTER(CParams), ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                 _vp, _vp, _vp],
    "tcp_batch_create": [ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(CParams), ctypes.c_int],
    "tcp_batch_destroy": [_vp],
    "tcp_batch_last_error": [_vp],
    "tcp_batch_set_state": [_vp, ctypes.c_int, _dp, _dp, _dp],
    "tcp_batch_get_state": [_vp, ctypes.c_int, _dp, _dp, _dp],
    "tcp_batch_get_histograms": [_vp, ctypes.c_int, _dp, _dp, ctypes.c_int],
    "tcp_batch_md_steps": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp,
                           ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _dp],
    "tcp_batch_set_shape": [_vp, ctypes.c_int],
    "tcp_batch_replica_shape": [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
    "tcp_batch_autotune": [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _dp],
    "tcp_batch_last_timings": [_vp, _dp, _dp],
}

_LIB = None


class TcpError(RuntimeError):
    pass


def load():
    """Load libtcp_gpu.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise TcpError(f"{LIB_PATH} is missing: build it with python -m tcp_allpair_amd.build")
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_char_p if name.endswith("_last_error") else ctypes.c_int
        _LIB = lib
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _aos(a, n):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != (n, 3):
        raise ValueError(f"expected shape ({n}, 3), got {a.shape}")
    return a


class Context:
    """One plasma resident on the current HIP device (tcp_create / tcp_destroy)."""

    def __init__(self, p: TcpParams, precision: int = MIXED):
        self.lib = load()
        self.p = p
        self.precision = precision
        self._c = CParams.from_params(p)
        h = _vp()
        rc = self.lib.tcp_create(ctypes.byref(h), ctypes.byref(self._c), int(precision))
        if rc != 0:
            raise TcpError(f"tcp_create: {self.lib.tcp_last_error(None).decode()} ({rc})")
        self.h = h
        self.seed = 12345

    def close(self):
        if getattr(self, "h", None):
            self.lib.tcp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _ck(self, rc, what):
        if rc != 0:
            raise TcpError(f"{what}: {self.lib.tcp_last_error(self.h).decode()} ({rc})")

    # -- reference entry points -------------------------------------------------
    def ewald_self(self, with_psi=False):
        a, rc2, uo = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        psi = np.zeros(int(np.prod(self.p.K))) if with_psi else None
        self._ck(self.lib.tcp_ewald_self(self.h, ctypes.byref(a), ctypes.byref(rc2),
                                         ctypes.byref(uo), _ptr(psi)), "tcp_ewald_self")
        if with_psi:
            return a.value, rc2.value, uo.value, psi.reshape(self.p.K[2], self.p.K[1], self.p.K[0])
        return a.value, rc2.value, uo.value

    def set_state(self, x=None, v=None, f=None):
        n = self.p.n
        arrs = [None if a is None else _aos(a, n) for a in (x, v, f)]
        self._ck(self.lib.tcp_set_state(self.h, *(_ptr(a) for a in arrs)), "tcp_set_state")

    def get_state(self):
        n = self.p.n
        x, v, f = (np.zeros((n, 3)) for _ in range(3))
        self._ck(self.lib.tcp_get_state(self.h, _ptr(x), _ptr(v), _ptr(f)), "tcp_get_state")
        return x, v, f

    def force_direct(self):
        ur = ctypes.c_double()
        self._ck(self.lib.tcp_force_direct(self.h, ctypes.byref(ur)), "tcp_force_direct")
        return ur.value

    def force_direct_ext(self, xg, ne_g, ni_g):
        xg = _aos(xg, ne_g + ni_g)
        gemit = np.zeros_like(xg)
        ur = ctypes.c_double()
        self._ck(self.lib.tcp_force_direct_ext(self.h, _ptr(xg), int(ne_g), int(ni_g), _ptr(gemit),
                                               ctypes.byref(ur)), "tcp_force_direct_ext")
        return gemit, ur.value

    def force_recip(self):
        um = ctypes.c_double()
        self._ck(self.lib.tcp_force_recip(self.h, ctypes.byref(um)), "tcp_force_recip")
        return um.value

    def remove_rigidmtn(self):
        self._ck(self.lib.tcp_remove_rigidmtn(self.h), "tcp_remove_rigidmtn")

    def vverlet1(self, thermo, dt, fric=None):
        fric = self.p.fric if fric is None else fric
        self._ck(self.lib.tcp_vverlet1(self.h, int(thermo), float(dt), float(fric)), "tcp_vverlet1")

    def vverlet2(self, thermo, dt, Te=None, Ti=None, tau=None, fric=None, seed=None):
        p = self.p
        a, b = ctypes.c_double(), ctypes.c_double()
        self._ck(self.lib.tcp_vverlet2(self.h, int(thermo), float(dt),
                                       float(p.Te if Te is None else Te),
                                       float(p.Ti if Ti is None else Ti),
                                       float(p.tau if tau is None else tau),
                                       float(p.fric if fric is None else fric),
                                       int(self.seed if seed is None else seed),
                                       ctypes.byref(a), ctypes.byref(b)), "tcp_vverlet2")
        return a.value, b.value

    def vverlet2_kick(self, dt):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._ck(self.lib.tcp_vverlet2_kick(self.h, float(dt), ctypes.byref(a), ctypes.byref(b)),
                 "tcp_vverlet2_kick")
        return a.value, b.value

    def vverlet2_scale(self, lam_e, lam_i):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._ck(self.lib.tcp_vverlet2_scale(self.h, float(lam_e), float(lam_i), ctypes.byref(a),
                                             ctypes.byref(b)), "tcp_vverlet2_scale")
        return a.value, b.value

    def histograms(self, reset=False):
        rdf = np.zeros(3 * NBIN)
        vdf = np.zeros(2 * NBIN)
        self._ck(self.lib.tcp_get_histograms(self.h, _ptr(rdf), _ptr(vdf), int(bool(reset))),
                 "tcp_get_histograms")
        return rdf.reshape(3, NBIN), vdf.reshape(2, NBIN)

    def md_steps(self, nsteps, thermo, dt, Te=None, Ti=None, tau=None, fric=None, seed=None,
                 trace=True):
        p = self.p
        tr = np.zeros((nsteps, 6)) if trace else None
        self._ck(self.lib.tcp_md_steps(self.h, int(nsteps), int(thermo), float(dt),
                                       float(p.Te if Te is None else Te),
                                       float(p.Ti if Ti is None else Ti),
                                       float(p.tau if tau is None else tau),
                                       float(p.fric if fric is None else fric),
                                       int(self.seed if seed is None else seed), _ptr(tr)),
                 "tcp_md_steps")
        return tr

    # -- extensions with no reference counterpart --------------------------------
    def set_lambda(self, lam):
        self._ck(self.lib.tcp_set_lambda(self.h, float(lam)), "tcp_set_lambda")
        self.p = self.p.replace(rs=float(lam))

    def energy_adjoint(self, q=None, with_dfdl=True):
        """(U, dU/dx, dU/dq, dU/dlambda, dF/dlambda or None) at the current positions."""
        n = self.p.n
        qa = None
        if q is not None:
            qa = np.ascontiguousarray(q, dtype=np.float64)
            if qa.shape != (n,):
                raise ValueError(f"expected {n} charges, got shape {qa.shape}")
        U, dl = ctypes.c_double(), ctypes.c_double()
        dudx = np.zeros((n, 3))
        dudq = np.zeros(n)
        dfdl = np.zeros((n, 3)) if with_dfdl else None
        self._ck(self.lib.tcp_energy_adjoint(self.h, _ptr(qa), ctypes.byref(U), _ptr(dudx), _ptr(dudq),
                                             ctypes.byref(dl), _ptr(dfdl)), "tcp_energy_adjoint")
        return U.value, dudx, dudq, dl.value, dfdl

    def timings(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._ck(self.lib.tcp_last_timings(self.h, ctypes.byref(a), ctypes.byref(b)),
                 "tcp_last_timings")
        return a.value, b.value


class Batch:
    """Independent replicas stepped together (tcp_batch_*): one persistent pair
    launch per step for all replicas, the step captured in a hipGraph."""

    NSHAPES = 6

    def __init__(self, params, precision: int = MIXED):
        self.lib = load()
        self.params = list(params)
        self.precision = precision
        arr = (CParams * len(self.params))(*[CParams.from_params(p) for p in self.params])
        h = _vp()
        rc = self.lib.tcp_batch_create(ctypes.byref(h), len(self.params), arr, int(precision))
        if rc != 0:
            raise TcpError(f"tcp_batch_create: {self.lib.tcp_batch_last_error(None).decode()} ({rc})")
        self.h = h
        self.seed = 12345

    @property
    def nrep(self):
        return len(self.params)

    def close(self):
        if getattr(self, "h", None):
            self.lib.tcp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _ck(self, rc, what):
        if rc != 0:
            raise TcpError(f"{what}: {self.lib.tcp_batch_last_error(self.h).decode()} ({rc})")

    def set_state(self, rep, x=None, v=None, f=None):
        n = self.params[rep].n if 0 <= rep < self.nrep else 0
        arrs = [None if a is None else _aos(a, n) for a in (x, v, f)]
        self._ck(self.lib.tcp_batch_set_state(self.h, int(rep), *(_ptr(a) for a in arrs)),
                 "tcp_batch_set_state")

    def get_state(self, rep):
        n = self.params[rep].n if 0 <= rep < self.nrep else 1
        x, v, f = (np.zeros((n, 3)) for _ in range(3))
        self._ck(self.lib.tcp_batch_get_state(self.h, int(rep), _ptr(x), _ptr(v), _ptr(f)),
                 "tcp_batch_get_state")
        return x, v, f

    def histograms(self, rep, reset=False):
        rdf = np.zeros(3 * NBIN)
        vdf = np.zeros(2 * NBIN)
        self._ck(self.lib.tcp_batch_get_histograms(self.h, int(rep), _ptr(rdf), _ptr(vdf),
                                                   int(bool(reset))), "tcp_batch_get_histograms")
        return rdf.reshape(3, NBIN), vdf.reshape(2, NBIN)

    def md_steps(self, nsteps, thermo, dt, Te=None, Ti=None, tau=1.0, fric=0.0, seed=None,
                 trace=True):
        te = None if Te is None else np.ascontiguousarray(Te, dtype=np.float64)
        ti = None if Ti is None else np.ascontiguousarray(Ti, dtype=np.float64)
        for a in (te, ti):
            if a is not None and a.shape != (self.nrep,):
                raise ValueError(f"expected {self.nrep} thermostat targets")
        tr = np.zeros((nsteps, self.nrep, 6)) if trace else None
        self._ck(self.lib.tcp_batch_md_steps(self.h, int(nsteps), int(thermo), float(dt), _ptr(te),
                                             _ptr(ti), float(tau), float(fric),
                                             int(self.seed if seed is None else seed), _ptr(tr)),
                 "tcp_batch_md_steps")
        return tr

    def set_shape(self, shape):
        self._ck(self.lib.tcp_batch_set_shape(self.h, int(shape)), "tcp_batch_set_shape")

    def replica_shape(self, rep):
        """Block shape of replica rep's pair work (tcp_batch_replica_shape)."""
        sh = ctypes.c_int()
        self._ck(self.lib.tcp_batch_replica_shape(self.h, int(rep), ctypes.byref(sh)),
                 "tcp_batch_replica_shape")
        return sh.value

    def autotune(self, reps=3):
        best = ctypes.c_int()
        ms = np.full(self.NSHAPES, np.nan)
        self._ck(self.lib.tcp_batch_autotune(self.h, int(reps), ctypes.byref(best), _ptr(ms)),
                 "tcp_batch_autotune")
        return best.value, ms

    def timings(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._ck(self.lib.tcp_batch_last_timings(self.h, ctypes.byref(a), ctypes.byref(b)),
                 "tcp_batch_last_timings")
        return a.value, b.value
