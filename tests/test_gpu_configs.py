"""The named workloads and the free Ewald splitting, through the C ABI, against the oracle.

  plain CPU form, 1000 e + 1000 i at density 1, alpha = 5/L, direct k-sum kmax = 8
  32768 particles at density 1 (L = 32 A), SPME 64^3 order 4, alpha rc = 3.5
  131072 particles at density 1 (L = 50.797 A), SPME 128^3 order 6: FP64 on a sample
      of 2048 rows against the whole system, and mixed against FP64
  VVerletStoch2's deterministic part (vverlet.f90:92-97)

Tolerances (stated per test):
  FP64 vs the oracle:    per-particle relative force error <= 1e-10, energies relative <= 1e-12,
                         RDF counts identical
  MIXED vs GPU FP64:     RMS relative force error <= 1e-6
"""
import sys

import numpy as np
import pytest

from oracle import cpu
from tcp_allpair_amd.gpu import Context
from tcp_allpair_amd.inputs import make_plasma
from tcp_allpair_amd.params import FP64, MIXED, RECIP_EWALD, RECIP_NONE, RECIP_SPME, STOCH, reduced_dt

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_err(f, ref, rms_ref=None):
    """|dF_i| / max(|F_i|, 1e-3 RMS|F|); rms_ref: the RMS of the whole system's forces
    when f, ref hold only a sample of its rows."""
    nref = np.linalg.norm(ref, axis=1)
    rms = np.sqrt(np.mean(nref ** 2)) if rms_ref is None else rms_ref
    return np.linalg.norm(f - ref, axis=1) / np.maximum(nref, 1e-3 * rms)


def rms_rel(f, ref):
    return np.sqrt(((f - ref) ** 2).sum() / (ref ** 2).sum())


def forces(p, x, precision):
    """Real-space + reciprocal forces (Force_direct_self, Force_SPME4 + SPME_SUM) and
    (Ur, Um, RDF) at positions x."""
    with Context(p, precision) as c:
        c.set_state(x=x, f=np.zeros_like(x))
        alpha = c.ewald_self()[0]
        ur = c.force_direct()
        um = c.force_recip()
        f = c.get_state()[2]
        rdf, _ = c.histograms()
    return f, ur, um, rdf, alpha


def test_free_alpha_kspace_c0():
    """1000 electrons + 1000 ions at reduced density 1 (L = 12.599 A), alpha = 5/L
    instead of the reference's 3.1/rc (force.f90:23), rc = L/2, direct k-sum with
    kmax = 8: FP64 against the oracle (real space + k-sum), mixed against FP64."""
    say("case free alpha c0 kspace")
    p, x, _ = make_plasma(1000, 1000, seed=2000, recip=RECIP_EWALD, kmax=8)
    p = p.replace(alpha=5.0 / p.box[0])
    fo, uro, rdfo = cpu.force_allpair(p, x)
    fko, umo = cpu.ewald_kspace(p, x)
    f, ur, um, rdf, alpha = forces(p, x, FP64)
    assert alpha == 5.0 / p.box[0]
    assert rel_err(f, fo + fko).max() <= 1e-10
    assert abs(ur - uro) <= 1e-12 * abs(uro)
    assert abs(um - umo) <= 1e-12 * abs(umo)
    np.testing.assert_array_equal(rdf, rdfo)
    fm, urm, umm, _, _ = forces(p, x, MIXED)
    say("mixed rms rel err", rms_rel(fm, f), "dUr", urm - ur)
    assert rms_rel(fm, f) <= 1e-6
    assert abs(umm - um) <= 1e-14 * abs(um)  # the reciprocal part is FP64 in both


def test_free_alpha_spme_arc35():
    """32768 particles (L = 32 A), rc = L/2, SPME 64^3 order 4 with alpha rc = 3.5
    (erfc(3.5) = 7.4e-7 at the cut-off): FP64 against the oracle, mixed against FP64."""
    say("case free alpha spme arc 3.5")
    p, x, _ = make_plasma(16384, 16384, seed=1001, K=(64, 64, 64), order=4)
    p = p.replace(alpha=3.5 / p.rc)
    fo, uro, rdfo = cpu.force_allpair(p, x)
    fro, umo = cpu.spme(p, x)
    f, ur, um, rdf, _ = forces(p, x, FP64)
    say("Ur", ur, "oracle", uro, "Um", um, "oracle", umo)
    assert rel_err(f, fo + fro).max() <= 1e-10
    assert abs(ur - uro) <= 1e-12 * abs(uro)
    assert abs(um - umo) <= 1e-12 * abs(umo)
    np.testing.assert_array_equal(rdf, rdfo)
    fm, _, _, _, _ = forces(p, x, MIXED)
    say("mixed rms rel err", rms_rel(fm, f))
    assert rms_rel(fm, f) <= 1e-6


@pytest.mark.parametrize("cubic", [True, False])
def test_mixed_at_the_alpha_rc_limit(cubic):
    """alpha rc = 4.5, the largest a mixed context takes: pairs near rc read the
    top rows of the mixed kernels' erfcx table (x up to 4.5).  3000 + 3000
    particles, rc = L/2 (the shortest edge's half in the box of aspect
    1 : 1.25 : 0.8), no reciprocal part: FP64 against the oracle, mixed against
    FP64 to 1e-6 RMS and Ur to 1e-7 of the sum of |qi qj|/r."""
    say("case alpha rc 4.5, cubic", cubic)
    p, x, _ = make_plasma(3000, 3000, seed=4500, recip=RECIP_NONE)
    if not cubic:
        box = np.asarray(p.box) * np.array([1.0, 1.25, 0.8])
        x = x * np.array([1.0, 1.25, 0.8])
        p = p.replace(box=tuple(box), rc=0.5 * box.min())
    p = p.replace(alpha=4.5 / p.rc)
    fo, uro, rdfo = cpu.force_allpair(p, x)
    with Context(p, FP64) as c:
        c.set_state(x=x, f=np.zeros_like(x))
        ur = c.force_direct()
        f = c.get_state()[2]
        rdf, _ = c.histograms()
    assert rel_err(f, fo).max() <= 1e-10
    assert abs(ur - uro) <= 1e-12 * abs(uro)
    np.testing.assert_array_equal(rdf, rdfo)
    with Context(p, MIXED) as c:
        c.set_state(x=x, f=np.zeros_like(x))
        urm = c.force_direct()
        fm = c.get_state()[2]
    # sum over pairs inside rc of |qi qj|/r, which the near-neutral Ur cancels
    L = np.asarray(p.box)
    scale = 0.0
    for i0 in range(0, p.n, 512):
        d = x[i0:i0 + 512, None, :] - x[None, :, :]
        d -= L * np.rint(d / L)
        r = np.sqrt((d * d).sum(-1))
        scale += 0.5 * (1.0 / r[(r < p.rc) & (r > 0)]).sum()
    say("mixed rms rel err", rms_rel(fm, f), "dUr", urm - ur, "scale", scale)
    assert rms_rel(fm, f) <= 1e-6
    assert abs(urm - ur) <= 1e-7 * scale


def test_fullsize_fp64_rows_vs_oracle():
    """65536 electrons + 65536 ions (L = 50.797 A), rc = L/2, SPME 128^3 order 6, the
    headline workload.  FP64: the forces of 2048 sampled particles against the whole
    system (oracle rows) plus the oracle's SPME forces, per particle to 1e-10, and Um to
    1e-12 (Ur needs every pair: the full oracle sum is left to the smaller systems).
    Mixed: RMS relative force error against this FP64 result <= 1e-6 over all particles."""
    say("case fullsize fp64 rows")
    p, x, _ = make_plasma(65536, 65536, seed=7, K=(128, 128, 128), order=6)
    rows = np.sort(np.random.default_rng(11).choice(p.n, 2048, replace=False))
    f, ur, um, rdf, _ = forces(p, x, FP64)
    say("gpu Ur", ur, "Um", um)
    fr, ur_rows = cpu.force_rows(p, x, rows)
    fro, umo = cpu.spme(p, x)
    say("oracle Um", umo)
    rms = np.sqrt(np.mean((f ** 2).sum(1)))
    err = rel_err(f[rows], fr + fro[rows], rms_ref=rms)
    say("rows: max per-particle rel err", err.max())
    assert err.max() <= 1e-10
    assert abs(um - umo) <= 1e-12 * abs(umo)
    fm, urm, umm, _, _ = forces(p, x, MIXED)
    e = rms_rel(fm, f)
    say("mixed vs fp64 rms rel err", e, "dUr/Ur", (urm - ur) / ur)
    assert e <= 1e-6
    assert abs(umm - um) <= 1e-14 * abs(um)  # the reciprocal part is FP64 in both


def test_stoch2_deterministic_part():
    """VVerletStoch2 (vverlet.f90:92-97): f += eta beta, v = (v + dt f/(2m))/(1 + fric dt/(2m)).
    Two contexts draw the same noise (same seed, same step counter); one starts from
    (v1, f1), the other from (0, 0).  The second gives eta beta in f exactly; the
    difference of the velocities must be (v1 + dt f1/(2m))/(1 + fric dt/(2m)), and the
    first run must equal the oracle's Stoch2 with that noise supplied."""
    say("case stoch2 deterministic part")
    p, x, v1 = make_plasma(3000, 2500, seed=9)
    p = p.replace(fric=0.3, recip=RECIP_SPME)
    dt = reduced_dt(0.005)
    rng = np.random.default_rng(17)
    f1 = rng.standard_normal(x.shape) * 40.0
    out = []
    for v0, f0 in ((v1, f1), (np.zeros_like(x), np.zeros_like(x))):
        with Context(p, FP64) as c:
            c.set_state(x=x, v=v0, f=f0)
            mv2 = c.vverlet2(STOCH, dt, seed=31)
            _, v, f = c.get_state()
        out.append((v, f, mv2))
    (va, fa, mv2a), (vb, fb, _) = out
    m = np.where(np.arange(p.n) < p.ne, p.xm[0], p.xm[1])[:, None]
    expect = (v1 + 0.5 * dt * f1 / m) / (1.0 + 0.5 * p.fric * dt / m)
    scale = np.abs(va) + np.abs(vb) + np.abs(expect)
    assert np.all(np.abs((va - vb) - expect) <= 1e-14 * scale)
    np.testing.assert_allclose(fa - f1, fb, rtol=0, atol=1e-14 * (np.abs(fa) + np.abs(f1)).max())
    assert np.any(fb != 0) and np.any(np.abs(expect) > np.abs(vb))
    # the oracle's Stoch2 with the same draws: eta = (eta beta) / beta per species
    beta = np.where(np.arange(p.n) < p.ne, np.sqrt(2.0 * p.fric * p.Te / dt),
                    np.sqrt(2.0 * p.fric * p.Ti / dt))[:, None]
    vo, fo, mv2el, mv2ion, _ = cpu.vverlet2(p, STOCH, dt, v1, f1, eta=fb / beta)
    np.testing.assert_allclose(va, vo, rtol=1e-14, atol=1e-14 * np.abs(vo).max())
    np.testing.assert_allclose(fa, fo, rtol=1e-14, atol=1e-14 * np.abs(fo).max())
    assert mv2a[0] == pytest.approx(mv2el, rel=1e-13)
    assert mv2a[1] == pytest.approx(mv2ion, rel=1e-13)
