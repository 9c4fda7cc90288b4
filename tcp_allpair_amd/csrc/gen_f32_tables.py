"""Writes f32_tables.h, the constant table of the mixed-precision pair kernels:

kErfcxTab: the scaled complementary error function erfcx(x) = exp(x^2) erfc(x)
on a uniform grid of intervals of width 1/16 from x = 0, one row of four FP32
coefficients (16 bytes) per interval: on interval k, x = (k + t)/16 with t in
[0, 1), erfcx(x) is the cubic c0 + c1 t + c2 t^2 + c3 t^3 interpolating it at
the 4 Chebyshev nodes of [0, 1] (40-digit arithmetic, mpmath), c1..c3 rounded
to FP32 and c0 then set so that the error averages to zero over the interval
before it is rounded in turn: a bias common to many pairs would add up in the
energy sum, whose attractive and repulsive parts largely cancel.  The kernel
finds the row from one multiply and one conversion (u = 16 x, k = int(u),
t = fract(u)) and forms erfc(x) = exp(-x^2) erfcx(x)
with the Gaussian it computes anyway for the force.  erfcx is smooth and
slowly varying where erfc itself falls by orders of magnitude, so a cubic per
interval holds it to 1.2e-7 relative in FP32 (3.8e-8 RMS) on [0, 4.5], where
the rational-argument degree-7 polynomial it replaces reached 1.1e-6.  Rows
reach x = 74/16 = 4.625: mixed contexts keep alpha rc <= 4.5
(kMixedErfcMax), and the kernels take the row from r masked to 0 beyond rc.

Checked by tests/test_f32_tables.py (against mpmath, with the kernel's
mapping and FP32 Horner steps).

    python tcp_allpair_amd/csrc/gen_f32_tables.py      (rewrites f32_tables.h)
"""
from __future__ import annotations

import os

import mpmath as mp
import numpy as np

PER_UNIT = 16
ROWS = 74  # x < 74/16 = 4.625
DEG = 3
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "f32_tables.h")


def erfcx(x):
    return mp.exp(x * x) * mp.erfc(x)


def row(k, deg=DEG):
    """FP32 monomial coefficients in t of the cubic interpolating erfcx((k + t)/16)
    at the Chebyshev nodes of [0, 1], with c1..c3 rounded first and c0 then
    chosen so that the rounded cubic's error averages to zero over the
    interval (before c0's own rounding)."""
    n = deg + 1
    nodes = [(mp.cos(mp.pi * (j + mp.mpf(1) / 2) / n) + 1) / 2 for j in range(n)]
    vander = mp.matrix([[t ** i for i in range(n)] for t in nodes])
    vals = mp.matrix([erfcx((k + t) / PER_UNIT) for t in nodes])
    c = [mp.mpf(float(np.float32(float(v)))) for v in mp.lu_solve(vander, vals)]
    mean_f = mp.quad(lambda t: erfcx((k + t) / PER_UNIT), [0, 1])
    mean_p = sum(c[i] / (i + 1) for i in range(1, n))
    c[0] = mean_f - mean_p
    return [float(np.float32(float(v))) for v in c]


def main():
    mp.mp.dps = 40
    rows = np.array([row(k) for k in range(ROWS)], dtype=np.float32)
    lines = [
        "// Generated by gen_f32_tables.py: do not edit.",
        "// kErfcxTab: exp(x^2) erfc(x) on interval k: sum_j kErfcxTab[k][j] t^j,",
        f"// x = (k + t)/{PER_UNIT}, t in [0, 1); rows up to x = {ROWS}/{PER_UNIT}.",
        "#pragma once",
        "",
        "namespace tcp {",
        "",
        f"constexpr int kErfcxPerUnit = {PER_UNIT};",
        f"constexpr int kErfcxRows = {ROWS};",
        "",
        "__device__ __constant__ const float kErfcxTab[kErfcxRows][4] = {",
    ]
    for r in rows:
        lines.append("    {" + ", ".join(float(c).hex() for c in r) + "},")
    lines += ["};", "", "}  // namespace tcp", ""]
    with open(OUT, "w") as f:
        f.write("\n".join(lines))
    print(f"wrote {OUT}: {ROWS} rows of erfcx cubics")


if __name__ == "__main__":
    main()
