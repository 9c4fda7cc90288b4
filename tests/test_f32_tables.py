"""The constant table of the mixed-precision pair kernels
(tcp_allpair_amd/csrc/f32_tables.h): erfcx(x) = exp(x^2) erfc(x) by a cubic per
interval of width 1/16, evaluated with the kernel's steps in FP32 (u = 16 x,
row int(u), t = fract(u), Horner with fused multiply-adds), against mpmath at
40 digits over the range the mixed kernels use (alpha r < alpha rc <= 4.5).
"""
import os
import re

import mpmath as mp
import numpy as np

HEADER = os.path.join(os.path.dirname(__file__), "..", "tcp_allpair_amd", "csrc", "f32_tables.h")
RANGE = 4.5  # kMixedErfcMax: mixed contexts keep alpha rc at or below it


def load():
    text = open(HEADER).read()
    per = int(re.search(r"kErfcxPerUnit = (\d+)", text).group(1))
    rows = int(re.search(r"kErfcxRows = (\d+)", text).group(1))
    body = text[text.index("kErfcxTab"):]
    vals = re.findall(r"\{([^{}]*)\},", body)
    tab = np.array([[float.fromhex(v.strip()) for v in r.split(",")] for r in vals])
    assert tab.shape == (rows, 4)
    assert np.all(tab.astype(np.float32).astype(np.float64) == tab)  # exact FP32 values
    return per, tab.astype(np.float32)


def _fma32(a, b, c):
    """a b + c rounded once to FP32 (the FP64 product of two FP32 values is
    exact, and its sum with an FP32 value rounds to FP32 as one fma would,
    barring double-rounding ties far below the tolerances here)."""
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)


def erfcx_tab(x, per, tab):
    u = (np.asarray(x, dtype=np.float32) * np.float32(per)).astype(np.float32)
    k = u.astype(np.int64)
    t = (u - k.astype(np.float32)).astype(np.float32)
    c = tab[k]
    p = c[:, 3]
    for j in (2, 1, 0):
        p = _fma32(p, t, c[:, j])
    return p


def _ref(x):
    mp.mp.dps = 40
    return np.array([float(mp.exp(mp.mpf(float(v)) ** 2) * mp.erfc(mp.mpf(float(v)))) for v in x])


def test_erfcx_table_accuracy():
    """1.2e-7 relative at most and 4e-8 RMS on [0, 4.5] (the polynomial in
    x/(2 + x) it replaced: 1.1e-6 and 2.9e-7)."""
    per, tab = load()
    rng = np.random.default_rng(3)
    grid = np.arange(0, int(RANGE * per) + 1) / per
    x = np.concatenate([rng.uniform(0.0, RANGE, 6000), grid, grid + 1e-6, [0.0, 1e-30, RANGE]])
    x = x[x <= RANGE].astype(np.float32)
    rel = np.abs(erfcx_tab(x, per, tab) - _ref(x)) / _ref(x)
    assert rel.max() <= 1.3e-7, (x[rel.argmax()], rel.max())
    assert np.sqrt((rel ** 2).mean()) <= 4.5e-8


def test_rows_cover_the_mixed_range():
    """The last row reaches past alpha rc <= 4.5 with FP32 rounding of
    u = 16 alpha r to spare, and the top rows still hold erfcx there."""
    per, tab = load()
    assert len(tab) / per > RANGE * (1 + 1e-6) + 1.0 / per
    x = np.array([RANGE * (1 + 2e-7), (len(tab) - 1e-3) / per], dtype=np.float32)
    rel = np.abs(erfcx_tab(x, per, tab) - _ref(x)) / _ref(x)
    assert rel.max() <= 1.3e-7


def test_rows_continuous():
    """Neighbouring rows agree at their common end (t = 1 of row k, t = 0 of
    row k + 1) to 2e-7 relative: no row misplaced or mis-scaled."""
    _, tab = load()
    t64 = tab.astype(np.float64)
    right, left = t64.sum(1), t64[:, 0]
    assert (np.abs(right[:-1] - left[1:]) / left[1:]).max() <= 2e-7


def test_erfc_from_the_table():
    """erfc(x) = exp(-x^2) erfcx(x) as the kernel forms it (the Gaussian
    rounded to FP32), relative to 2.5e-7 where erfc > 1e-30."""
    per, tab = load()
    x = np.linspace(0.0, RANGE, 4001).astype(np.float32)
    g = np.exp(-(x.astype(np.float64) ** 2)).astype(np.float32)
    ec = (g * erfcx_tab(x, per, tab)).astype(np.float32)
    mp.mp.dps = 40
    ref = np.array([float(mp.erfc(mp.mpf(float(v)))) for v in x])
    assert (np.abs(ec - ref) / ref).max() <= 2.5e-7
