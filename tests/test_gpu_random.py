"""The HIP path against the oracle on the randomised systems of
tests/randsys.py (the oracle equals the reference's own routines bit for bit
on them, tests/test_oracle_random.py).

Small systems (one or two tiles, one species absent in some) and larger ones
(up to 6000 particles: several blocks, species boundaries and system ends
inside tiles), in orthorhombic boxes of random aspect with random cut-off,
lambda and RDF bin widths:
  FP64 real-space forces vs the oracle: per-particle relative <= 1e-10 (floor
      1e-3 of the RMS force), Ur to 1e-12 of the sum of |qi qj|/r over the
      pairs inside rc (the terms whose signed sum is Ur), RDF counts identical;
  mixed vs FP64: RMS relative force error <= 1e-6, Ur to 1e-6 relative or
      1e-8 of that sum where Ur cancels (a system of one species, whose pair
      energies all have one sign, shows the FP32 path's bias of a few ulps,
      1-3e-7 of Ur, in full),
      with the pair planted at exactly rc moved 1e-4 rc inside: a pair within
      FP32 rounding of rc may fall on the other side of the cut-off than in
      FP64, a step of the truncated pair force (eps q^2 (erfc(3.1) + 6.2
      e^{-9.61}/sqrt(pi))/rc^2, 3.5e-3/rc^2 eV/A here) that the reference's
      own truncation makes, not an arithmetic error;
  ten FP64 steps of main.f90:87-153 without the reciprocal part under NOTEMP,
      BEREND and ISOKIN vs the oracle's steps: Ur, Epot, mv2 per step to 1e-10
      relative, positions to 1e-10.
"""
import sys

import numpy as np
import pytest

from oracle import cpu
from randsys import random_system
from tcp_allpair_amd.gpu import Context
from tcp_allpair_amd.params import BEREND, FP64, ISOKIN, MIXED, NBIN, NOTEMP, reduced_dt

pytestmark = pytest.mark.gpu

CASES = [(s, 48) for s in range(20)] + [(s, 3000) for s in range(40, 48)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_err(f, ref):
    """Per-particle relative error, the denominator floored at 1e-3 of the RMS
    force (exact zeros expected where no pair lies inside rc)."""
    nref = np.linalg.norm(ref, axis=1)
    floor = 1e-3 * np.sqrt(np.mean(nref ** 2))
    if floor == 0.0:
        return np.where(np.linalg.norm(f, axis=1) == 0.0, 0.0, np.inf)
    return np.linalg.norm(f - ref, axis=1) / np.maximum(nref, floor)


def energy_scale(p, x):
    """Sum of |qi qj|/r over the minimum-image pairs inside rc."""
    q = np.where(np.arange(p.n) < p.ne, p.q[0], p.q[1])
    box = np.array(p.box)
    s = 0.0
    for i in range(p.n - 1):
        d = x[i] - x[i + 1:]
        d -= box * np.round(d / box)
        r = np.sqrt((d ** 2).sum(1))
        m = r < p.rc
        s += np.abs(q[i] * q[i + 1:][m] / r[m]).sum()
    return s


@pytest.mark.parametrize("seed,nmax", CASES)
def test_random_forces_vs_oracle(seed, nmax):
    p, x, _ = random_system(seed, nmax)
    say("case", seed, nmax, "ne", p.ne, "ni", p.ni, "box", p.box, "rc", p.rc)
    f_ref, ur_ref, rdf_ref = cpu.force_direct_self(p, x)
    scale = energy_scale(p, x)

    def run(prec, xs):
        with Context(p, prec) as c:
            c.set_state(x=xs, f=np.zeros_like(xs))
            ur = c.force_direct()
            return c.get_state()[2], ur, c.histograms()[0]

    f64, ur64, rdf64 = run(FP64, x)
    assert rel_err(f64, f_ref).max() <= 1e-10
    assert abs(ur64 - ur_ref) <= 1e-12 * scale
    np.testing.assert_array_equal(rdf64, rdf_ref)
    xm = x.copy()
    if p.n >= 5:  # the pair at exactly rc (randsys) moved inside
        xm[4] = xm[3] + np.array([p.rc * (1.0 - 1e-4), 0.0, 0.0])
        xm[4, 0] -= p.box[0] * np.round(xm[4, 0] / p.box[0])
    f64, ur64, _ = run(FP64, xm)
    fmx, urmx, _ = run(MIXED, xm)
    if (f64 ** 2).sum() > 0:
        assert np.sqrt(((fmx - f64) ** 2).sum() / (f64 ** 2).sum()) <= 1e-6
    else:  # no pair inside rc
        assert not fmx.any()
    assert abs(urmx - ur64) <= max(1e-6 * abs(ur64), 1e-8 * scale)


@pytest.mark.parametrize("seed,nmax,thermo", [(20, 48, NOTEMP), (21, 48, BEREND), (22, 48, ISOKIN),
                                              (48, 1500, NOTEMP), (49, 1500, BEREND),
                                              (50, 1500, ISOKIN)])
def test_random_steps_vs_oracle(seed, nmax, thermo):
    p, x, v = random_system(seed, nmax)
    say("case", seed, nmax, thermo, "ne", p.ne, "ni", p.ni)
    dt = reduced_dt(0.005)
    f0 = cpu.force_direct_self(p, x)[0]
    xs, vs, fs = (np.ascontiguousarray(a, dtype=np.float64).copy() for a in (x, v, f0))
    rdf, vdf = np.zeros(3 * NBIN), np.zeros(2 * NBIN)
    ref = []
    for _ in range(10):
        r = cpu.step(p, thermo, dt, xs, vs, fs, rdf, vdf)
        ref.append([r["Ur"], r["Epot"], r["mv2el"], r["mv2ion"]])
    with Context(p, FP64) as c:
        c.set_state(x=x, v=v, f=f0)
        tr = c.md_steps(10, thermo, dt)
        xg = c.get_state()[0]
    np.testing.assert_allclose(tr[:, [0, 3, 4, 5]], np.array(ref), rtol=1e-10, atol=0)
    np.testing.assert_allclose(xg, xs, rtol=0, atol=1e-10)
