"""The constant tables of the FP64 pair kernel (tcp_allpair_amd/csrc/f64_tables.h).

erfc: its piecewise polynomials, evaluated with the kernel's interval mapping
and one Horner pass for P and dP/dt (erfc_tab in kernels_pair.hip), against
mpmath and scipy.special.erfc, and the derivative against its closed form
erfc'(x) = -(2/sqrt(pi)) exp(-x^2) (scipy/numpy), over the range the Ewald and
Kelbg terms use; past x = 6 both are zero.  exp: the powers 2^(j/256) against
mpmath at 40 digits, and the table-driven reduction of exp_tab (the same steps
in numpy) against numpy's exp over the arguments the kernel meets (-r/lambda,
down to underflow).
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


def _scale():
    return float(re.search(r"kErfcScale = ([0-9.]+)", _text()).group(1))


def _xmax():
    text = _text()
    return int(re.search(r"kErfcN = (\d+)", text).group(1)) / int(
        re.search(r"kErfcPerUnit = (\d+)", text).group(1))


def load_erfc():
    text = _text()
    deg = int(re.search(r"kErfcDeg = (\d+)", text).group(1))
    n = int(re.search(r"kErfcN = (\d+)", text).group(1))
    body = text[text.index("kErfcTab"):text.index("kExp2Tab[kExp2N]")]
    rows = re.findall(r"\{([^{}]*)\},", body)
    tab = np.array([[float.fromhex(v.strip()) for v in r.split(",")] for r in rows])
    assert tab.shape == (n + 1, deg + 1)
    return tab


def load_exp2():
    text = _text()
    n = int(re.search(r"kExp2N = (\d+)", text).group(1))
    body = text[text.index("kExp2Tab[kExp2N] = {"):]
    body = body[body.index("{") + 1:body.index("};")]
    tab = np.array([float.fromhex(v.strip()) for v in body.split(",") if v.strip()])
    assert tab.shape == (n,)
    return tab


def _fma(a, b, c):
    """a b + c rounded once to double, as the device's fma (the 64-bit-mantissa
    intermediate leaves an error far below an ulp of the result)."""
    return (np.longdouble(a) * np.longdouble(b) + np.longdouble(c)).astype(np.float64)


def erfc_tab(x, tab):
    """The kernel's mapping (xs = kErfcScale x, interval floor(xs/2), the zero
    row past the table) and its Horner pass with fused multiply-adds: returns
    erfc(x) and erfc'(x) = kErfcScale dP/dt."""
    scale = _scale()
    xs = np.asarray(x, dtype=np.float64) * scale
    ki = np.minimum(xs, 2.0 * (len(tab) - 1) + 1.0).astype(np.int64)
    t = xs - (ki | 1)
    c = tab[ki >> 1]
    p = c[:, -1].copy()
    d = p.copy()
    p = _fma(p, t, c[:, -2])
    for j in range(tab.shape[1] - 3, -1, -1):
        d = _fma(d, t, p)
        p = _fma(p, t, c[:, j])
    return p, scale * d


def _points(xmax):
    rng = np.random.default_rng(5)
    grid = np.arange(0, int(xmax * _scale() / 2) + 1) / (_scale() / 2)
    return np.concatenate([rng.uniform(0.0, xmax, 200000), grid, grid + 1e-12,
                           [1e-300, 1e-8, 0.5, xmax - 1e-12]])


def test_erfc_table_matches_mpmath():
    """Against erfc at 40 digits: absolute to about an ulp of values near 1
    (the reference's 1 - DERF(x) is good to about the same, and is 0 from
    x = 5.93 on), relative to 1.5e-15 where erfc > 1e-5 (x < 3.1, the Ewald
    range alpha r < 3.1 of alpha = 3.1/rc; there the reference's absolute
    error is already 1e-11 relative)."""
    tab = load_erfc()
    rng = np.random.default_rng(11)
    xmax = _xmax()
    grid = np.arange(0, int(xmax * _scale() / 2)) / (_scale() / 2)
    x = np.concatenate([rng.uniform(0.0, xmax, 3000), grid, grid + 1e-12, [1e-300, xmax - 1e-12]])
    p, _ = erfc_tab(x, tab)
    mp.mp.dps = 40
    ref = np.array([float(mp.erfc(v)) for v in x])
    err = np.abs(p - ref)
    assert err.max() <= 1.5e-16, (x[err.argmax()], err.max())
    big = ref > 1e-5
    assert (err[big] / ref[big]).max() <= 1.5e-15
    assert abs(erfc_tab([0.0], tab)[0][0] - 1.0) <= 2e-16


def test_erfc_table_matches_scipy():
    """A dense sweep against scipy.special.erfc, whose own error reaches
    3e-16 absolute near x = 0.9."""
    tab = load_erfc()
    x = _points(_xmax())
    p, _ = erfc_tab(x, tab)
    assert np.abs(p - special.erfc(x)).max() <= 5e-16


def test_erfc_derivative_is_the_gaussian():
    """(2/sqrt(pi)) exp(-x^2) from the derivative: within 8e-16 absolute (the
    values reach 1.13) and 1e-13 relative where it exceeds 1e-4."""
    tab = load_erfc()
    x = _points(_xmax())
    _, d = erfc_tab(x, tab)
    g = 2.0 / np.sqrt(np.pi) * np.exp(-x * x)
    assert np.abs(-d - g).max() <= 8e-16
    big = g > 1e-4
    assert (np.abs(-d - g)[big] / g[big]).max() <= 1e-13


@pytest.mark.parametrize("xmax", [3.1, 4.5, 6.0 - 1e-9])
def test_ewald_force_term(xmax):
    """num = erfc(x) + (2x/sqrt(pi)) exp(-x^2) as the kernel forms it, x P - xs dP/dt."""
    tab = load_erfc()
    x = np.linspace(0.0, xmax, 100001)
    p, d = erfc_tab(x, tab)
    num = _fma(-_scale() * x, d / _scale(), p)
    ref = special.erfc(x) + 2.0 * x / np.sqrt(np.pi) * np.exp(-x * x)
    assert np.abs(num - ref).max() <= 1e-15


def test_past_the_table_is_zero():
    tab = load_erfc()
    assert np.all(tab[-1] == 0.0)
    xmax = _xmax()
    x = np.array([xmax, xmax + 0.5, 30.0, 1e3, 1e6, 1e300])
    p, d = erfc_tab(x, tab)
    assert np.all(p == 0.0) and np.all(d == 0.0)
    # erfc and (2/sqrt(pi)) exp(-x^2) dropped there are below 2.2e-17 and
    # 2.7e-16 (the reference's 1 - DERF(x) is exactly 0 from x = 5.93)
    assert xmax >= 6.0
    assert special.erfc(6.0) < 2.2e-17 and 2.0 / np.sqrt(np.pi) * np.exp(-36.0) < 2.7e-16
    # and so is the Ewald force term erfc(x) + (2x/sqrt(pi)) exp(-x^2), below
    # 1.6e-15 of its value 1 at x = 0
    assert (special.erfc(x) + 2.0 * x / np.sqrt(np.pi) * np.exp(-x * x)).max() < 1.6e-15
    assert 1.0 - special.erf(5.93) == 0.0


def test_table_rows_continuous():
    """Neighbouring intervals agree at their common end (t = 1 of row k, t = -1
    of row k + 1) to an ulp or two: no row is misplaced or mis-scaled."""
    tab = load_erfc()[:-1]
    right = tab.sum(1)
    left = (tab * (-1.0) ** np.arange(tab.shape[1])).sum(1)
    assert np.abs(right[:-1] - left[1:]).max() <= 4e-16


def test_exp2_table_correctly_rounded():
    tab = load_exp2()
    mp.mp.dps = 40
    for j, v in enumerate(tab):
        exact = mp.power(2, mp.mpf(j) / len(tab))
        assert abs(mp.mpf(v) - exact) <= mp.mpf(2) ** -53 * exact, j
    assert tab[0] == 1.0


def exp_tab(y, tab):
    """The steps of exp_tab (kernels_pair.hip) in numpy, fused multiply-adds
    as _fma forms them."""
    k = np.rint(y * float.fromhex("0x1.71547652b82fep+8"))
    r = _fma(-k, float.fromhex("0x1.62e42fefa39efp-9"), y)
    r = _fma(-k, float.fromhex("0x1.abc9e3b39803fp-64"), r)
    ki = k.astype(np.int64)
    q = _fma(r, 1.0 / 24.0, 1.0 / 6.0)
    q = _fma(q, r, 0.5)
    q = _fma(q, r, 1.0)
    q = q * r
    t = tab[ki & (len(tab) - 1)]
    return np.ldexp(_fma(t, q, t), (ki >> 8).astype(np.int32))


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
