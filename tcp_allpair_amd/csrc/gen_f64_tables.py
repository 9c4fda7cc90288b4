"""Writes f64_tables.h, the constant tables of the FP64 pair kernel:

kErfcTab: piecewise polynomials of the complementary error function erfc(x).
Intervals: width 1/32 on [0, 6) (192 intervals, one uniform grid so that the
kernel finds the interval from one multiply and one conversion).  On each, with
x = m + h t, t in [-1, 1], the degree-7 polynomial P(t) is fitted through its
derivative: P'(t) = h erfc'(m + h t) = -(2h/sqrt(pi)) exp(-(m + h t)^2) is the
degree-6 polynomial interpolating it at the 7 Chebyshev nodes, and
P(t) = erfc(m) + the integral of P' from 0 to t (40-digit arithmetic, mpmath,
then rounded to double).  Fitting the derivative keeps the Gaussian term, which
the kernel takes from P', as close as erfc itself: interpolating erfc at 8
nodes leaves P' twice as far off.  A last row of zeros stands for x >= 6,
where erfc(x) < 2.2e-17 and (2/sqrt(pi)) exp(-x^2) < 2.7e-16; the reference's
1 - DERF(x) is already exactly 0 from x = 5.93.  The kernel evaluates P and P'
by one Horner pass; erfc'(x) = -(2/sqrt(pi)) exp(-x^2) = (1/h) P'(t) = 64 P'(t),
so the Gaussian term of the Ewald force, (2x/sqrt(pi)) exp(-x^2) = -64 x P'(t),
and Kelbg's exp(-s^2) come from the same row.  Maximum absolute error of the
double Horner evaluation: erfc 1.1e-16 (about an ulp of values near 1; the
reference's 1 - DERF(x) is good to about the same), (2/sqrt(pi)) exp(-x^2)
6.7e-16.

kExp2Tab: 2^(j/256), j = 0..255, correctly rounded, for exp(y) = 2^n 2^(j/256)
e^r with y = (256 n + j) ln2/256 + r, |r| <= ln2/512, where e^r is the Taylor
polynomial of degree 4 (remainder below r^5/120 < 4e-17).

Both are checked by tests/test_f64_tables.py (against mpmath, scipy.special.erfc,
the exact derivative, and numpy's exp).

    python tcp_allpair_amd/csrc/gen_f64_tables.py      (rewrites f64_tables.h)
"""
from __future__ import annotations

import os

import mpmath as mp

DEG = 7
XMAX, PER_UNIT = 6, 32  # intervals of width 1/32 on [0, 6), then a row of zeros
EXP2_N, EXP2_BITS = 256, 8
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "f64_tables.h")


def intervals():
    return [(mp.mpf(i) / PER_UNIT, mp.mpf(i + 1) / PER_UNIT) for i in range(XMAX * PER_UNIT)]


def cheb_interp(f, n):
    """Monomial coefficients in t of the polynomial of degree n - 1 that
    interpolates f(t) at the n Chebyshev nodes of [-1, 1]."""
    nodes = [mp.cos(mp.pi * (k + mp.mpf(1) / 2) / n) for k in range(n)]
    vals = [f(t) for t in nodes]
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
    return mono


def interval_poly(a, b, deg=DEG):
    """Monomial coefficients (in t = (x - m)/h) of erfc on [a, b]: P(0) = erfc(m)
    and P' the Chebyshev interpolant of h erfc'(m + h t) of degree deg - 1."""
    m, h = (a + b) / 2, (b - a) / 2
    dp = cheb_interp(lambda t: -2 * h / mp.sqrt(mp.pi) * mp.exp(-(m + h * t) ** 2), deg)
    return [float(mp.erfc(m))] + [float(dp[i] / (i + 1)) for i in range(deg)]


def main():
    mp.mp.dps = 40
    rows = [interval_poly(a, b) for a, b in intervals()]
    rows.append([0.0] * (DEG + 1))
    exp2 = [float(mp.power(2, mp.mpf(j) / EXP2_N)) for j in range(EXP2_N)]
    lines = [
        "// Generated by gen_f64_tables.py: do not edit.",
        "// kErfcTab: erfc(x) on interval k < kErfcN: sum_j kErfcTab[k][j] t^j,",
        f"// t = {2 * PER_UNIT} x - (2k + 1) in [-1, 1], intervals of width 1/{PER_UNIT} on [0, {XMAX});",
        f"// row kErfcN is zero (x >= {XMAX}).",
        "// kExp2Tab: 2^(j/256), j = 0..255.",
        "#pragma once",
        "",
        "namespace tcp {",
        "",
        f"constexpr int kErfcDeg = {DEG};",
        f"constexpr int kErfcPerUnit = {PER_UNIT};",
        f"constexpr int kErfcN = {XMAX * PER_UNIT};  // intervals of width 1/{PER_UNIT} on [0, {XMAX})",
        f"constexpr double kErfcScale = {2 * PER_UNIT}.0;  // t + 2k + 1 = kErfcScale x",
        f"constexpr int kExp2Bits = {EXP2_BITS};",
        f"constexpr int kExp2N = {EXP2_N};  // 1 << kExp2Bits",
        "",
        "__device__ __constant__ const double kErfcTab[kErfcN + 1][kErfcDeg + 1] = {",
    ]
    for r in rows:
        lines.append("    {" + ", ".join(c.hex() for c in r) + "},")
    lines += ["};", "", "__device__ __constant__ const double kExp2Tab[kExp2N] = {"]
    for i in range(0, EXP2_N, 4):
        lines.append("    " + ", ".join(c.hex() for c in exp2[i:i + 4]) + ",")
    lines += ["};", "", "}  // namespace tcp", ""]
    with open(OUT, "w") as f:
        f.write("\n".join(lines))
    print(f"wrote {OUT}: {len(rows) - 1} erfc intervals of degree {DEG} and a zero row, "
          f"{EXP2_N} powers of 2")


if __name__ == "__main__":
    main()
