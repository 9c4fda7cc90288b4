"""Driver of _ref/ref_harness, the reference's own Fortran routines compiled from
their sources by oracle/Makefile.  Test infrastructure only: it writes the
harness's binary input, runs it and parses its binary output."""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

from tcp_allpair_amd.params import NBIN, TcpParams

_HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(_HERE, "_ref", "ref_harness")

MODE_FORCE, MODE_PSI, MODE_EXT, MODE_RIGID, MODE_VV1, MODE_VV2, MODE_TRAJ, MODE_TIME = range(1, 9)


def available() -> bool:
    return os.access(HARNESS, os.X_OK)


def _write_input(path, mode, p: TcpParams, x, v, f, dt=0.0, nsteps=0, thermo=0, nown=(0, 0)):
    n = p.n
    hdr = np.array([mode, p.ne, p.ni, *p.K, nsteps, thermo, nown[0], nown[1]], dtype="<i4")
    reals = np.array([*p.box, *p.q, *p.xm, p.rc, p.rs, *p.dx, p.Te, p.Ti, dt, p.tau, p.fric],
                     dtype="<f8")
    with open(path, "wb") as fh:
        fh.write(hdr.tobytes())
        fh.write(reals.tobytes())
        for a in (x, v, f):
            a = np.zeros((n, 3)) if a is None else np.asarray(a, dtype="<f8")
            fh.write(np.ascontiguousarray(a).tobytes())


class _Reader:
    def __init__(self, buf):
        self.b = buf
        self.o = 0

    def f8(self, count):
        a = np.frombuffer(self.b, dtype="<f8", count=count, offset=self.o).copy()
        self.o += 8 * count
        return a


def run(mode, p: TcpParams, x, v=None, f=None, dt=0.0, nsteps=0, thermo=0, nown=(0, 0), timeout=600):
    """Run the harness; returns a dict of its outputs."""
    if not available():
        raise RuntimeError("oracle/_ref/ref_harness is not built (make -C oracle ref)")
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        _write_input(fin, mode, p, x, v, f, dt, nsteps, thermo, nown)
        subprocess.run([HARNESS, fin, fout], check=True, timeout=timeout,
                       stdout=subprocess.DEVNULL)
        buf = open(fout, "rb").read()
    r = _Reader(buf)
    out = {}
    if mode == MODE_TIME:
        pairs, secs, ur = r.f8(3)
        return dict(pairs=pairs, seconds=secs, Ur=ur)
    npart = nown[0] + nown[1] if mode == MODE_EXT else p.n
    s = r.f8(7)
    out.update(alpha=s[0], rc2=s[1], Uo=s[2], Ur=s[3], Epot=s[4], mv2el=s[5], mv2ion=s[6])
    out["x"] = r.f8(3 * npart).reshape(npart, 3)
    out["v"] = r.f8(3 * npart).reshape(npart, 3)
    out["f"] = r.f8(3 * npart).reshape(npart, 3)
    out["rdf"] = r.f8(3 * NBIN).reshape(3, NBIN)
    out["vdf"] = r.f8(2 * NBIN).reshape(2, NBIN)
    if mode == MODE_PSI:
        out["psi"] = r.f8(int(np.prod(p.K))).reshape(p.K[2], p.K[1], p.K[0])
    elif mode == MODE_EXT:
        next_ = p.n - npart
        out["Ur_self"], out["Ur_ext"] = r.f8(2)
        out["gemit"] = r.f8(3 * next_).reshape(next_, 3)
        out["f_after_sum"] = r.f8(3 * npart).reshape(npart, 3)
    elif mode == MODE_VV1:
        out["gsend"] = r.f8(3 * p.n).reshape(p.n, 3)
    elif mode == MODE_TRAJ:
        out["trace"] = r.f8(4 * nsteps).reshape(nsteps, 4)
    return out


def writers(cwd, p: TcpParams, x, v, f, ids, ndump, nloop, nrest, tnow, mv2el, mv2ion, epot, um,
            rdf, vdf, timeout=120):
    """Mode 9: the reference's own DUMP, SAMPLING and Restart (result.f90) write
    their files into cwd for the given data."""
    if not available():
        raise RuntimeError("oracle/_ref/ref_harness is not built (make -C oracle ref)")
    fin = os.path.join(cwd, "writers_in.bin")
    _write_input(fin, 9, p, x, v, f)
    with open(fin, "ab") as fh:
        fh.write(np.array([ndump, nloop, nrest], dtype="<i4").tobytes())
        fh.write(np.asarray(ids, dtype="<i4").tobytes())
        fh.write(np.array([tnow, mv2el, mv2ion, epot, um], dtype="<f8").tobytes())
        fh.write(np.ascontiguousarray(np.asarray(rdf, dtype="<f8").reshape(3, NBIN)).tobytes())
        fh.write(np.ascontiguousarray(np.asarray(vdf, dtype="<f8").reshape(2, NBIN)).tobytes())
    subprocess.run([HARNESS, fin, os.path.join(cwd, "writers_out.bin")], check=True,
                   timeout=timeout, cwd=cwd, stdout=subprocess.DEVNULL)
    os.remove(fin)
    os.remove(os.path.join(cwd, "writers_out.bin"))
