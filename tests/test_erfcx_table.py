"""The erfcx table of the FP64 pair kernel (tcp_allpair_amd/csrc/erfcx_table.h):
its piecewise polynomials, evaluated with the kernel's interval mapping
(erfcx_tab in kernels_pair.hip), against scipy.special.erfcx, and
erfc = exp(-x^2) erfcx against scipy.special.erfc over the range the Ewald and
Kelbg terms use."""
import os
import re

import numpy as np
import pytest
from scipy import special

HEADER = os.path.join(os.path.dirname(__file__), "..", "tcp_allpair_amd", "csrc", "erfcx_table.h")


def load_table():
    text = open(HEADER).read()
    deg = int(re.search(r"kErfcxDeg = (\d+)", text).group(1))
    lo = int(re.search(r"kErfcxLo = (\d+)", text).group(1))
    hi = int(re.search(r"kErfcxHi = (\d+)", text).group(1))
    rows = re.findall(r"\{([^{}]*)\},", text)
    tab = np.array([[float.fromhex(v.strip()) for v in r.split(",")] for r in rows])
    assert tab.shape == (lo + hi, deg + 1)
    return tab, lo, hi


def erfcx_tab(x, tab, lo_n, hi_n):
    """The kernel's mapping: width 1/8 below 8, 1/2 from 8 to 28, clamped past 28."""
    x = np.asarray(x, dtype=np.float64)
    lo = x < 8.0
    u = np.where(lo, x * 16.0, (x - 8.0) * 4.0)
    k = np.minimum((u * 0.5).astype(np.int64), np.where(lo, lo_n - 1, hi_n - 1))
    t = np.minimum(u - (2 * k + 1), 1.0)
    c = tab[k + np.where(lo, 0, lo_n)]
    p = c[:, -1].copy()
    for j in range(tab.shape[1] - 2, -1, -1):
        p = p * t + c[:, j]
    return p


def test_table_matches_scipy_erfcx():
    tab, lo, hi = load_table()
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(0.0, 28.0, 200000), np.arange(0, 28 * 8 + 1) / 8.0,
                        np.arange(0, 28 * 8) / 8.0 + 1e-12, [1e-300, 1e-8, 0.5, 7.999999999999, 8.0]])
    got = erfcx_tab(x, tab, lo, hi)
    ref = special.erfcx(x)
    rel = np.abs(got - ref) / ref
    assert rel.max() <= 1e-15, (x[rel.argmax()], rel.max())
    assert abs(erfcx_tab([0.0], tab, lo, hi)[0] - 1.0) <= 2e-16


@pytest.mark.parametrize("xmax", [4.0, 8.0, 27.0])
def test_erfc_from_table(xmax):
    tab, lo, hi = load_table()
    x = np.linspace(0.0, xmax, 100001)
    got = np.exp(-x * x) * erfcx_tab(x, tab, lo, hi)
    ref = special.erfc(x)
    ok = ref > 1e-300
    rel = np.abs(got[ok] - ref[ok]) / ref[ok]
    # exp(-x^2) of a rounded x^2 costs up to 2 x^2 ulp here; the kernel splits x^2 in hi + lo
    assert rel.max() <= 4e-16 * (1 + 2 * xmax * xmax)


def test_past_the_table_is_finite():
    tab, lo, hi = load_table()
    x = np.array([28.0, 30.0, 1e3, 1e6])
    p = erfcx_tab(x, tab, lo, hi)
    assert np.all(np.isfinite(p)) and np.all(np.exp(-x * x) * p == 0.0)
