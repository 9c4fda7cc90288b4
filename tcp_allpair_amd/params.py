"""System parameters of the two-component plasma, mirroring the reference's types.

The fields follow TYPE(PM) (data.f90:68-74) and the counts / bin widths of
TYPE(NM) (data.f90:97-103) that the force-and-integrate path reads.  The
constants are those of data.f90:16-24.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math

EPS = 14.399644154020082   # e^2/(4 pi eps0) in eV*A, data.f90:21
TPS = 10.180505306645389   # fs per internal time unit, data.f90:22
PI = 3.1415926535897931    # data.f90:23
SQRTPI = 1.7724538509055159  # data.f90:24
NDIST = 500                # histogram bins, data.f90:16 (arrays hold NDIST+1)
NBIN = NDIST + 1

# Thermostat codes; the reference selects by the param%thermo string (main.f90:88-153)
NOTEMP, BEREND, ISOKIN, STOCH = 0, 1, 2, 3
THERMO_NAMES = {"NOTEMP": NOTEMP, "BEREND": BEREND, "ISOKIN": ISOKIN, "STOCH": STOCH}

# Reciprocal-space method
RECIP_NONE, RECIP_SPME, RECIP_EWALD = 0, 1, 2

# Arithmetic of the pair kernel
FP64, MIXED = 0, 1

# Regularisation options (no reference counterpart; 0 = Deutsch e-i only)
KELBG, REG_EE = 1, 2


@dataclasses.dataclass
class TcpParams:
    """Parameters of one plasma; electrons are particles 0..ne-1, ions ne..ne+ni-1."""
    ne: int
    ni: int
    box: tuple = (1.0, 1.0, 1.0)          # param%box (A)
    q: tuple = (-1.0, 1.0)                # param%q (e)
    xm: tuple = (1.0, 100.0)              # param%xm (amu)
    rc: float = 0.5                       # param%rc (A)
    rs: float = 0.5                       # param%rs: lambda of the e-i term (A)
    alpha: float = 0.0                    # <= 0 -> 3.1/rc (force.f90:23)
    eps: float = EPS
    dx: tuple = (0.01, 0.01, 0.01)        # NS%dx_ee, dx_ei, dx_ii
    Te: float = EPS                       # param%Te (eV)
    Ti: float = EPS                       # param%Ti (eV)
    K: tuple = (32, 32, 32)               # param%K
    order: int = 4                        # B-spline order (4 = Force_SPME4)
    kmax: int = 0                         # direct Ewald k-sum extent (RECIP_EWALD)
    recip: int = RECIP_SPME
    tau: float = 1.0                      # param%tau (Berendsen)
    fric: float = 0.0                     # param%alpha (Langevin friction)
    flags: int = 0                        # KELBG | REG_EE (0: the reference's model)

    @property
    def n(self) -> int:
        return self.ne + self.ni

    @property
    def volume(self) -> float:
        return self.box[0] * self.box[1] * self.box[2]

    @property
    def dv(self) -> tuple:
        """VDF bin widths NS%dv_e, NS%dv_i (main.f90:286-287)."""
        return (self.Te * 4.0 / NDIST, self.Ti * 4.0 / NDIST)

    @property
    def ewald_alpha(self) -> float:
        return self.alpha if self.alpha > 0 else 3.1 / self.rc

    def replace(self, **kw) -> "TcpParams":
        return dataclasses.replace(self, **kw)


class CParams(ctypes.Structure):
    """C layout shared by include/tcp_gpu.h (tcp_params) and oracle/tcp_oracle.h."""
    _fields_ = [
        ("ne", ctypes.c_int), ("ni", ctypes.c_int),
        ("K", ctypes.c_int * 3), ("order", ctypes.c_int), ("kmax", ctypes.c_int),
        ("recip", ctypes.c_int), ("flags", ctypes.c_int),
        ("box", ctypes.c_double * 3), ("q", ctypes.c_double * 2), ("xm", ctypes.c_double * 2),
        ("rc", ctypes.c_double), ("rs", ctypes.c_double), ("alpha", ctypes.c_double),
        ("eps", ctypes.c_double),
        ("dx_ee", ctypes.c_double), ("dx_ei", ctypes.c_double), ("dx_ii", ctypes.c_double),
        ("dv_e", ctypes.c_double), ("dv_i", ctypes.c_double),
    ]

    @classmethod
    def from_params(cls, p: TcpParams) -> "CParams":
        c = cls()
        c.ne, c.ni = int(p.ne), int(p.ni)
        c.K[:] = [int(k) for k in p.K]
        c.order, c.kmax, c.recip, c.flags = int(p.order), int(p.kmax), int(p.recip), int(p.flags)
        c.box[:] = [float(b) for b in p.box]
        c.q[:] = [float(v) for v in p.q]
        c.xm[:] = [float(v) for v in p.xm]
        c.rc, c.rs, c.alpha, c.eps = float(p.rc), float(p.rs), float(p.alpha), float(p.eps)
        c.dx_ee, c.dx_ei, c.dx_ii = (float(v) for v in p.dx)
        c.dv_e, c.dv_i = p.dv
        return c


def reduced_dt(dt_star: float, eps: float = EPS) -> float:
    """Time step in internal units (tps fs) for dt* in units of tau* = tps/sqrt(eps) fs."""
    return dt_star / math.sqrt(eps)
