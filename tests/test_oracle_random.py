"""The oracle (C restatement) against the reference's own routines on
randomised systems, bit for bit.

The reference's force.f90 and vverlet.f90, compiled from their sources into
oracle/_ref/ref_harness by oracle/Makefile, run on seeded random systems of a
few dozen particles: random electron and ion counts (one species may be
absent), orthorhombic boxes of random aspect, a random cut-off up to half the
shortest box edge, random lambda and RDF bin widths (some so narrow that pairs
fall past bin Ndist+1), particles placed on the box faces and at the cut-off,
and short trajectories under each thermostat the harness can seed (NOTEMP,
BEREND, ISOKIN).  The restatement keeps the reference's evaluation order, so
forces, energies, RDF/VDF counts and trajectories must agree exactly.  Skipped
where the harness is not built (no reference sources).
"""
import numpy as np
import pytest

from oracle import cpu
from oracle import refharness as rh
from randsys import random_system
from tcp_allpair_amd.params import BEREND, ISOKIN, NBIN, NOTEMP, reduced_dt

pytestmark = pytest.mark.skipif(not rh.available(), reason="oracle/_ref/ref_harness not built")


SEEDS = list(range(20))


@pytest.mark.parametrize("seed", SEEDS)
def test_force_direct_self_random_bitwise(seed):
    p, x, v = random_system(seed)
    ref = rh.run(rh.MODE_FORCE, p, x, v)
    a, rc2, uo = cpu.ewald_self(p)
    assert (a, rc2, uo) == (ref["alpha"], ref["rc2"], ref["Uo"])
    f, ur, rdf = cpu.force_direct_self(p, x)
    np.testing.assert_array_equal(f, ref["f"])
    assert ur == ref["Ur"]
    np.testing.assert_array_equal(rdf, ref["rdf"])
    assert (ur + uo) * p.eps == ref["Epot"]


@pytest.mark.parametrize("seed,thermo", [(s, t) for s, t in zip(range(20, 29),
                                                                   [NOTEMP, BEREND, ISOKIN] * 3)])
def test_trajectory_random_bitwise(seed, thermo):
    """Ten steps of main.f90:87-153 without the reciprocal part from the
    harness's own step-0 forces."""
    p, x, v = random_system(seed)
    dt = reduced_dt(0.005)
    f0 = rh.run(rh.MODE_FORCE, p, x, v)["f"]
    ref = rh.run(rh.MODE_TRAJ, p, x, v, f0, dt=dt, nsteps=10, thermo=thermo)
    xs, vs, fs = (np.ascontiguousarray(a, dtype=np.float64).copy() for a in (x, v, f0))
    rdf, vdf = np.zeros(3 * NBIN), np.zeros(2 * NBIN)
    for s in range(10):
        r = cpu.step(p, thermo, dt, xs, vs, fs, rdf, vdf)
        got = np.array([r["Ur"], r["Epot"], r["mv2el"], r["mv2ion"]])
        np.testing.assert_array_equal(got, ref["trace"][s])
    np.testing.assert_array_equal(xs, ref["x"])
    np.testing.assert_array_equal(vs, ref["v"])
    np.testing.assert_array_equal(fs, ref["f"])
    np.testing.assert_array_equal(rdf.reshape(3, NBIN), ref["rdf"])
    np.testing.assert_array_equal(vdf.reshape(2, NBIN), ref["vdf"])
