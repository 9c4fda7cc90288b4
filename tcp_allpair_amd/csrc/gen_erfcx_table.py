"""Writes erfcx_table.h: piecewise polynomials of the scaled complementary error
function erfcx(x) = exp(x^2) erfc(x) for the FP64 pair kernel, which forms
erfc(x) = exp(-x^2) erfcx(x) from the exp(-x^2) it needs anyway for the
Gaussian term of the Ewald force.

Intervals: width 1/8 on [0, 8) (64 intervals) and width 1/2 on [8, 28)
(40 intervals); on each, erfcx(m + h t) for t in [-1, 1] is the degree-9
polynomial interpolating erfcx at the 10 Chebyshev nodes, computed with
40-digit arithmetic (mpmath) and rounded to double.  Past x = 28,
exp(-x^2) underflows to zero, so the value there does not matter.  Maximum
relative error of the double Horner evaluation: 1.9e-16 (checked by
tests/test_erfcx_table.py against scipy.special.erfcx).

    python tcp_allpair_amd/csrc/gen_erfcx_table.py      (rewrites erfcx_table.h)
"""
from __future__ import annotations

import os

import mpmath as mp

DEG = 9
SEGMENTS = ((0, 8, 8), (8, 28, 2))  # (from, to, intervals per unit)
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "erfcx_table.h")


def intervals():
    out = []
    for a, b, per in SEGMENTS:
        for i in range((b - a) * per):
            out.append((mp.mpf(a) + mp.mpf(i) / per, mp.mpf(a) + mp.mpf(i + 1) / per))
    return out


def erfcx(x):
    return mp.exp(x * x) * mp.erfc(x)


def interval_poly(a, b, deg=DEG):
    """Monomial coefficients (in t = (x - m)/h) of the Chebyshev interpolant."""
    n = deg + 1
    m, h = (a + b) / 2, (b - a) / 2
    nodes = [mp.cos(mp.pi * (k + mp.mpf(1) / 2) / n) for k in range(n)]
    vals = [erfcx(m + h * t) for t in nodes]
    cheb = [2 * mp.fsum(vals[k] * mp.cos(mp.pi * j * (k + mp.mpf(1) / 2) / n) for k in range(n)) / n
            for j in range(n)]
    cheb[0] /= 2
    basis = [[mp.mpf(1)], [mp.mpf(0), mp.mpf(1)]]  # T_j as monomials
    for j in range(2, n):
        nxt = [mp.mpf(0)] * (j + 1)
        for i, v in enumerate(basis[-1]):
            nxt[i + 1] += 2 * v
        for i, v in enumerate(basis[-2]):
            nxt[i] -= v
        basis.append(nxt)
    mono = [mp.mpf(0)] * n
    for j in range(n):
        for i, v in enumerate(basis[j]):
            mono[i] += cheb[j] * v
    return [float(c) for c in mono]


def main():
    mp.mp.dps = 40
    ivs = intervals()
    rows = [interval_poly(a, b) for a, b in ivs]
    lines = [
        "// Generated by gen_erfcx_table.py: do not edit.",
        "// erfcx(x) = exp(x^2) erfc(x) on interval k: sum_j kErfcxTab[k][j] t^j,",
        "// t = (x - m_k) / h_k in [-1, 1]; widths 1/8 on [0, 8), 1/2 on [8, 28).",
        "#pragma once",
        "",
        "namespace tcp {",
        "",
        f"constexpr int kErfcxDeg = {DEG};",
        f"constexpr int kErfcxLo = {(SEGMENTS[0][1] - SEGMENTS[0][0]) * SEGMENTS[0][2]};  // intervals on [0, 8)",
        f"constexpr int kErfcxHi = {(SEGMENTS[1][1] - SEGMENTS[1][0]) * SEGMENTS[1][2]};  // intervals on [8, 28)",
        "constexpr int kErfcxN = kErfcxLo + kErfcxHi;",
        "",
        "__device__ __constant__ const double kErfcxTab[kErfcxN][kErfcxDeg + 1] = {",
    ]
    for r in rows:
        lines.append("    {" + ", ".join(c.hex() for c in r) + "},")
    lines += ["};", "", "}  // namespace tcp", ""]
    with open(OUT, "w") as f:
        f.write("\n".join(lines))
    print(f"wrote {OUT}: {len(rows)} intervals, degree {DEG}")


if __name__ == "__main__":
    main()
