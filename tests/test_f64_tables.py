"""The constant tables of the FP64 pair kernel (tcp_allpair_amd/csrc/f64_tables.h).

erfcx: its piecewise polynomials, evaluated with the kernel's interval mapping
(erfcx_tab in kernels_pair.hip), against scipy.special.erfcx, and erfc =
exp(-x^2) erfcx against scipy.special.erfc over the range the Ewald and Kelbg
terms use.  exp: the powers 2^(j/64) against mpmath at 40 digits, and the
table-driven reduction of exp_tab (the same steps in numpy) against numpy's exp
over the arguments the kernel meets (-x^2 and -r/lambda, down to underflow).
"""
import os
import re

import mpmath as mp
import numpy as np
import pytest
from scipy import special

HEADER = os.path.join(os.path.dirname(__file__), "..", "tcp_allpair_amd", "csrc", "f64_tables.h")


def _text():
    return open(HEADER).read()


def load_erfcx():
    text = _text()
    deg = int(re.search(r"kErfcxDeg = (\d+)", text).group(1))
    n = int(re.search(r"kErfcxN = (\d+)", text).group(1))
    body = text[text.index("kErfcxTab"):text.index("kExp2Tab[kExp2N]")]
    rows = re.findall(r"\{([^{}]*)\},", body)
    tab = np.array([[float.fromhex(v.strip()) for v in r.split(",")] for r in rows])
    assert tab.shape == (n, deg + 1)
    return tab


def load_exp2():
    text = _text()
    n = int(re.search(r"kExp2N = (\d+)", text).group(1))
    body = text[text.index("kExp2Tab[kExp2N] = {"):]
    body = body[body.index("{") + 1:body.index("};")]
    tab = np.array([float.fromhex(v.strip()) for v in body.split(",") if v.strip()])
    assert tab.shape == (n,)
    return tab


def erfcx_tab(x, tab):
    """The kernel's mapping: intervals of width 1/8, the last one past 28."""
    x = np.asarray(x, dtype=np.float64)
    k = np.minimum((x * 8.0).astype(np.int64), len(tab) - 1)
    t = np.minimum(x * 16.0 - (2 * k + 1), 1.0)
    c = tab[k]
    p = c[:, -1].copy()
    for j in range(tab.shape[1] - 2, -1, -1):
        p = p * t + c[:, j]
    return p


def test_table_matches_scipy_erfcx():
    tab = load_erfcx()
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(0.0, 28.0, 200000), np.arange(0, 28 * 8 + 1) / 8.0,
                        np.arange(0, 28 * 8) / 8.0 + 1e-12, [1e-300, 1e-8, 0.5, 7.999999999999, 8.0]])
    got = erfcx_tab(x, tab)
    ref = special.erfcx(x)
    rel = np.abs(got - ref) / ref
    assert rel.max() <= 1e-15, (x[rel.argmax()], rel.max())
    assert abs(erfcx_tab([0.0], tab)[0] - 1.0) <= 2e-16


@pytest.mark.parametrize("xmax", [4.0, 8.0, 27.0])
def test_erfc_from_table(xmax):
    tab = load_erfcx()
    x = np.linspace(0.0, xmax, 100001)
    got = np.exp(-x * x) * erfcx_tab(x, tab)
    ref = special.erfc(x)
    ok = ref > 1e-300
    rel = np.abs(got[ok] - ref[ok]) / ref[ok]
    # exp(-x^2) of a rounded x^2 costs up to 2 x^2 ulp here; the kernel splits x^2 in hi + lo
    assert rel.max() <= 4e-16 * (1 + 2 * xmax * xmax)


def test_past_the_table_is_finite():
    tab = load_erfcx()
    x = np.array([28.0, 30.0, 1e3, 1e6])
    p = erfcx_tab(x, tab)
    assert np.all(np.isfinite(p)) and np.all(np.exp(-x * x) * p == 0.0)


def test_exp2_table_correctly_rounded():
    tab = load_exp2()
    mp.mp.dps = 40
    for j, v in enumerate(tab):
        exact = mp.power(2, mp.mpf(j) / len(tab))
        assert abs(mp.mpf(v) - exact) <= mp.mpf(2) ** -53 * exact, j
    assert tab[0] == 1.0


def exp_tab(y, tab):
    """The steps of exp_tab (kernels_pair.hip) in numpy.  The reduction's fused
    multiply-add y - k ln2/64 is formed in extended precision (as the device's
    fma forms it exactly); the polynomial without fma, a rounding or two looser."""
    k = np.rint(y * 92.332482616893656)
    hi = np.longdouble(float.fromhex("0x1.62e42fefa39efp-7"))
    r = (np.longdouble(y) - np.longdouble(k) * hi).astype(np.float64)
    r = r - k * float.fromhex("0x1.abc9e3b39803fp-62")
    ki = k.astype(np.int64)
    q = ((((r * (1.0 / 120.0) + 1.0 / 24.0) * r + 1.0 / 6.0) * r + 0.5) * r + 1.0) * r
    t = tab[ki & 63]
    return np.ldexp(t + t * q, (ki >> 6).astype(np.int32))


def test_exp_reduction_accuracy():
    tab = load_exp2()
    rng = np.random.default_rng(9)
    y = -np.concatenate([rng.uniform(0, 1, 100000), rng.uniform(0, 50, 100000),
                         rng.uniform(0, 700, 100000), np.arange(0, 700, 0.25)])
    got = exp_tab(y, tab)
    ref = np.exp(y)
    rel = np.abs(got - ref) / ref
    assert rel.max() <= 4e-16, (y[rel.argmax()], rel.max())
    deep = exp_tab(np.array([-745.2, -800.0, -1e4]), tab)
    assert np.all(deep == 0.0)
