"""Seeded synthetic plasmas and the reference's text formats.

`make_plasma` builds the workloads of the benchmark and the tests: electrons
then ions, positions uniform in [-L/2, L/2)^3 with L = (N/density)^(1/3),
velocities Maxwell-Boltzmann per species.  `write_tcp_prm` / `write_restart_xyz`
emit the positional input formats read by main.f90:221-294 and :383-440 so the
unchanged Fortran driver can start from exactly the same state (rest='ON');
`write_input_xyz` writes the positions-only file of a fresh start (rest='OFF',
main.f90:347-379).
"""
from __future__ import annotations

import math

import numpy as np

from .params import EPS, NDIST, RECIP_SPME, TPS, TcpParams


def make_plasma(ne: int, ni: int, *, density: float = 1.0, T: float = EPS,
                lam: float = 0.5, seed: int = 1, K=(32, 32, 32), order: int = 4,
                rc: float | None = None, alpha: float = 0.0, recip: int = RECIP_SPME,
                kmax: int = 0, nbins: int = NDIST, xm=(1.0, 100.0), q=(-1.0, 1.0)):
    """Return (params, x, v) with x, v float64 arrays of shape (ne+ni, 3)."""
    n = ne + ni
    L = (n / density) ** (1.0 / 3.0)
    rc = L / 2.0 if rc is None else rc
    dx = rc / nbins
    p = TcpParams(ne=ne, ni=ni, box=(L, L, L), q=tuple(q), xm=tuple(xm), rc=rc, rs=lam,
                  alpha=alpha, dx=(dx, dx, dx), Te=T, Ti=T, K=tuple(K), order=order,
                  kmax=kmax, recip=recip)
    rng = np.random.Generator(np.random.PCG64(seed))
    x = (rng.random((n, 3)) - 0.5) * L
    sig = np.empty((n, 1))
    sig[:ne] = math.sqrt(T / xm[0])
    sig[ne:] = math.sqrt(T / xm[1])
    v = rng.standard_normal((n, 3)) * sig
    return p, x, v


def write_tcp_prm(path, p: TcpParams, *, tmax_fs: float, trest_fs: float, tdump_fs: float,
                  tsamp_fs: float, dt_fs: float, thermo: str = "NOTEMP", rest: str = "ON"):
    """tcp.prm in the positional layout of main.f90:221-291 (one header line per block)."""
    g = lambda v: repr(float(v))
    lines = [
        "# tcp.prm",
        "# tmax trest tdump tsamp dt (fs)",
        " ".join(g(v) for v in (tmax_fs, trest_fs, tdump_fs, tsamp_fs, dt_fs)),
        "# box", " ".join(g(b) for b in p.box),
        "# Te Ti", g(p.Te), g(p.Ti),
        "# NTel NTion", str(p.ne), str(p.ni),
        "# rc", g(p.rc),
        "# rs", g(p.rs),
        "# xm", g(p.xm[0]), g(p.xm[1]),
        "# q", g(p.q[0]), g(p.q[1]),
        "# thermo", f"'{thermo:<6s}'", f"{g(p.tau)} {g(p.fric)}",
        "# rest trace", f"'{rest}' 'OFF'",
        "# dx_ee dx_ei dx_ii", " ".join(g(d) for d in p.dx),
        "# K", " ".join(str(k) for k in p.K),
    ]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def write_restart_xyz(path, p: TcpParams, x, v, f):
    """restart.xyz as read by main.f90:389-438: name x y z vx vy vz fx fy fz id."""
    with open(path, "w") as fh:
        fh.write(f"{p.n}\nframe = 0\n")
        for i in range(p.n):
            name = "Cu" if i < p.ne else "C"
            idn = i + 1 if i < p.ne else i - p.ne + 1
            vals = " ".join("%.17g" % t for t in (*x[i], *v[i], *f[i]))
            fh.write(f"{name} {vals} {idn}\n")


def write_input_xyz(path, p: TcpParams, x):
    """input.xyz as read by the fresh start (rest = 'OFF', main.f90:347-379):
    two header lines, then name x y z id per particle, electrons first."""
    with open(path, "w") as fh:
        fh.write(f"{p.n}\nframe = 0\n")
        for i in range(p.n):
            name = "Cu" if i < p.ne else "C"
            idn = i + 1 if i < p.ne else i - p.ne + 1
            fh.write(f"{name} {' '.join('%.17g' % t for t in x[i])} {idn}\n")


def read_restart_xyz(path):
    """Inverse of write_restart_xyz (also reads REST##.xyz written by result.f90:74-131)."""
    with open(path) as fh:
        n = int(fh.readline().split()[0])
        fh.readline()
        rows = [fh.readline().split() for _ in range(n)]
    a = np.array([[float(t) for t in r[1:10]] for r in rows])
    return a[:, 0:3].copy(), a[:, 3:6].copy(), a[:, 6:9].copy()


def fs_per_internal() -> float:
    return TPS
