"""ctypes binding of libtcp_io.so (tcp_allpair_amd/host/tcp_io.h): the readers and
writers of the reference's text files shared with the standalone driver tcp_run.
The writers write into the current directory, as result.f90 does."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .params import NBIN

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtcp_io.so")
RUN_PATH = os.path.join(_HERE, "tcp_run")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class Prm(ctypes.Structure):
    """tcp_prm of tcp_io.h: the records of tcp.prm (main.f90:221-291)."""
    _fields_ = [
        ("tmax", ctypes.c_double), ("trest", ctypes.c_double), ("tdump", ctypes.c_double),
        ("tsamp", ctypes.c_double), ("dt", ctypes.c_double), ("box", ctypes.c_double * 3),
        ("Te", ctypes.c_double), ("Ti", ctypes.c_double), ("NTel", ctypes.c_int),
        ("NTion", ctypes.c_int), ("rc", ctypes.c_double), ("rs", ctypes.c_double),
        ("xm", ctypes.c_double * 2), ("q", ctypes.c_double * 2), ("thermo", ctypes.c_char * 8),
        ("tau", ctypes.c_double), ("alpha", ctypes.c_double), ("rest", ctypes.c_char * 8),
        ("trace", ctypes.c_char * 8), ("dx", ctypes.c_double * 3), ("K", ctypes.c_int * 3),
    ]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with python -m tcp_allpair_amd.build")
        _LIB = ctypes.CDLL(LIB_PATH)
    return _LIB


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def format_es(v: float, w: int, d: int) -> str:
    buf = ctypes.create_string_buffer(w + 1)
    lib().tcp_io_format_es(ctypes.c_double(v), int(w), int(d), buf)
    return buf.value.decode()


def read_prm(path) -> Prm:
    p = Prm()
    err = ctypes.create_string_buffer(256)
    rc = lib().tcp_io_read_prm(str(path).encode(), ctypes.byref(p), err)
    if rc:
        raise RuntimeError(err.value.decode())
    return p


def read_restart(path, n):
    x, v, f = (np.zeros((n, 3)) for _ in range(3))
    ids = np.zeros(n, dtype=np.int32)
    err = ctypes.create_string_buffer(256)
    rc = lib().tcp_io_read_restart(str(path).encode(), int(n), _p(x), _p(v), _p(f), _p(ids, _ip), err)
    if rc:
        raise RuntimeError(err.value.decode())
    return x, v, f, ids


def ener_open(path):
    if lib().tcp_io_ener_open(str(path).encode()):
        raise OSError(f"cannot write {path}")


def ener_row(path, t_fs, Tel, Tion, Epot, Etot):
    d = ctypes.c_double
    if lib().tcp_io_ener_row(str(path).encode(), d(t_fs), d(Tel), d(Tion), d(Epot), d(Etot)):
        raise OSError(f"cannot write {path}")


def write_dump(nfreq, ntel, ntion, ndump, box, dx, dv_e, dv_i, rdf, vdf):
    box = np.ascontiguousarray(box, dtype=np.float64)
    dx = np.ascontiguousarray(dx, dtype=np.float64)
    rdf = np.ascontiguousarray(rdf, dtype=np.float64).reshape(3 * NBIN)
    vdf = np.ascontiguousarray(vdf, dtype=np.float64).reshape(2 * NBIN)
    if lib().tcp_io_write_dump(int(nfreq), int(ntel), int(ntion), int(ndump), _p(box), _p(dx),
                               ctypes.c_double(dv_e), ctypes.c_double(dv_i), _p(rdf), _p(vdf)):
        raise OSError("cannot write the RDF/VDF files")


def write_rest(nfreq, ntel, ntion, x, v, f, ids):
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, v, f)]
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    if lib().tcp_io_write_rest(int(nfreq), int(ntel), int(ntion), *(_p(a) for a in arrs),
                               _p(ids, _ip)):
        raise OSError("cannot write the restart file")
