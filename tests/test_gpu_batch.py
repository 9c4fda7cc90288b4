"""Batched replicas (tcp_batch_*): one persistent pair launch and a captured step for all.

Tolerances:
  FP64 batch vs each replica's own context:   bit for bit (state, trace, histograms)
  MIXED batch at its context's block shape:   bit for bit
  MIXED batch, every shape, vs FP64:          RMS relative force error <= 1e-6
"""
import sys

import numpy as np
import pytest

from tcp_allpair_amd.gpu import Batch, Context, TcpError
from tcp_allpair_amd.inputs import make_plasma
from tcp_allpair_amd.params import (BEREND, EPS, FP64, ISOKIN, KELBG, MIXED, NOTEMP, RECIP_EWALD,
                                    RECIP_NONE, RECIP_SPME, REG_EE, STOCH, reduced_dt)

pytestmark = pytest.mark.gpu

SIZES = [(150, 150), (260, 250), (389, 388), (512, 512), (1000, 1000)]
DENS = [0.5, 1.0, 1.5, 2.0, 1.0]
TEMP = [0.5, 1.0, 2.0, 4.0, 1.0]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def replicas(recip=RECIP_EWALD, flags=0, kmax=4):
    out = []
    for r, ((ne, ni), d, t) in enumerate(zip(SIZES, DENS, TEMP)):
        p, x, v = make_plasma(ne, ni, density=d, T=t * EPS, seed=100 + r, recip=recip, kmax=kmax)
        out.append((p.replace(flags=flags), x, v))
    return out


def standalone(p, x, v, precision, nsteps, thermo, dt, **kw):
    with Context(p, precision) as c:
        c.set_state(x=x, v=v, f=np.zeros_like(x))
        tr = c.md_steps(nsteps, thermo, dt, **kw)
        st = c.get_state()
        h = c.histograms()
    return tr, st, h


def run_batch(reps, precision, nsteps, thermo, dt, shape=None, **kw):
    with Batch([p for p, _, _ in reps], precision) as b:
        if shape is not None:
            b.set_shape(shape)
        for r, (p, x, v) in enumerate(reps):
            b.set_state(r, x=x, v=v, f=np.zeros_like(x))
        tr = b.md_steps(nsteps, thermo, dt, **kw)
        st = [b.get_state(r) for r in range(len(reps))]
        h = [b.histograms(r) for r in range(len(reps))]
    return tr, st, h


def assert_bitwise(reps, precision, nsteps, thermo, dt, shape=None, batch_kw=None, rtol=0.0, **kw):
    """rtol = 0: bit for bit; otherwise state and trace to rtol (histograms exactly)."""
    btr, bst, bh = run_batch(reps, precision, nsteps, thermo, dt, shape, **(batch_kw or kw))
    for r, (p, x, v) in enumerate(reps):
        skw = dict(kw)
        if batch_kw:
            skw.update({k: batch_kw[k] for k in ("tau", "fric", "seed") if k in batch_kw})
            skw.update(Te=batch_kw["Te"][r], Ti=batch_kw["Ti"][r])
        tr, st, h = standalone(p, x, v, precision, nsteps, thermo, dt, **skw)
        if rtol:
            np.testing.assert_allclose(btr[:, r, :], tr, rtol=rtol, err_msg=f"trace of replica {r}")
            for a, b_, name in zip(st, bst[r], "xvf"):
                np.testing.assert_allclose(b_, a, rtol=rtol, atol=rtol * np.abs(a).max(),
                                           err_msg=f"{name} of replica {r}")
        else:
            np.testing.assert_array_equal(btr[:, r, :], tr, err_msg=f"trace of replica {r}")
            for a, b_, name in zip(st, bst[r], "xvf"):
                np.testing.assert_array_equal(b_, a, err_msg=f"{name} of replica {r}")
        np.testing.assert_array_equal(bh[r][0], h[0])
        np.testing.assert_array_equal(bh[r][1], h[1])


def test_fp64_batch_nve_ewald_bitwise():
    say("case batch fp64 nve ewald")
    assert_bitwise(replicas(), FP64, 6, NOTEMP, reduced_dt(0.005))


def test_fp64_batch_thermostats_bitwise():
    reps = replicas()
    Te = np.array([t * EPS for t in TEMP]) * 1.1
    Ti = Te * 0.9
    for thermo in (BEREND, ISOKIN):
        say("case batch fp64 thermostat", thermo)
        assert_bitwise(reps, FP64, 4, thermo, reduced_dt(0.005), batch_kw=dict(Te=Te, Ti=Ti, tau=0.5))
    say("case batch fp64 stoch")
    assert_bitwise(reps, FP64, 4, STOCH, reduced_dt(0.005),
                   batch_kw=dict(Te=Te, Ti=Ti, fric=0.3, seed=77))


def test_fp64_batch_no_recip_regularised_bitwise():
    say("case batch fp64 recip none kelbg")
    assert_bitwise(replicas(RECIP_NONE, KELBG | REG_EE), FP64, 5, NOTEMP, reduced_dt(0.005))


def test_fp64_batch_two_shapes_bitwise():
    """FP64 replicas keep their own contexts' block shapes: the small ones 4 x 4
    tiles and a 40000-particle one 8 x 8, in two persistent launches per step;
    every replica matches its standalone run bit for bit."""
    say("case batch fp64 two shapes")
    reps = replicas(RECIP_NONE)[:3] + [make_plasma(20000, 20000, seed=121, recip=RECIP_NONE)]
    with Batch([p for p, _, _ in reps], FP64) as b:
        shapes = [b.replica_shape(r) for r in range(len(reps))]
    say("replica shapes", shapes)
    assert set(shapes) == {0, 3}  # both persistent launches take part
    assert_bitwise(reps, FP64, 2, NOTEMP, reduced_dt(0.005))


@pytest.mark.parametrize("shape,large", [(3, False), (0, True)])
def test_mixed_batch_matches_standalone(shape, large):
    """A single context runs the same kernel body on 4 x 4 tile blocks for
    systems as small as the five replicas (shape 3) and on 8 x 8 tile blocks for
    a 60000-particle one (shape 0); the batch at the context's shape matches it
    bit for bit (the waves of a block add their reactions in fixed turns)."""
    say("case batch mixed vs standalone, shape", shape)
    reps = [make_plasma(30000, 30000, seed=120, recip=RECIP_NONE)] if large else replicas()
    assert_bitwise(reps, MIXED, 5 if not large else 2, NOTEMP, reduced_dt(0.005), shape=shape)


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4, 5])
def test_mixed_batch_shapes_vs_fp64(shape):
    """Every block shape: within 1e-6 RMS of FP64, and two runs of the same
    batch identical bit for bit (each shape's waves add reactions in fixed
    turns or, on 4 x 4 tiles by 4 waves, through per-wave slots; the replicas
    of 300..2000 particles leave the last block column partial)."""
    say("case batch mixed shape", shape)
    reps = replicas(flags=KELBG)
    btr, bst, bh = run_batch(reps, MIXED, 2, NOTEMP, reduced_dt(0.005), shape=shape)
    ftr, fst, fh = run_batch(reps, FP64, 2, NOTEMP, reduced_dt(0.005))
    for r in range(len(reps)):
        f, f64 = bst[r][2], fst[r][2]
        assert np.sqrt(((f - f64) ** 2).sum() / (f64 ** 2).sum()) <= 1e-6
        assert abs(bh[r][0].sum() - fh[r][0].sum()) <= 1e-4 * fh[r][0].sum()
    btr2, bst2, bh2 = run_batch(reps, MIXED, 2, NOTEMP, reduced_dt(0.005), shape=shape)
    np.testing.assert_array_equal(btr2, btr)
    for r in range(len(reps)):
        for a, b in zip(bst2[r], bst[r]):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(bh2[r][0], bh[r][0])


def test_batch_workload_vs_oracle():
    """The batching workload itself (tcp_allpair_amd.batchbench.batch_workload): 64
    replicas, N_e log-uniform in 256..8192 with as many ions, densities U[0.5, 2], T*
    U[0.5, 4], lambda = 0.5, rc = L/2, direct k-sum kmax = 6.  One step with dt = 0
    leaves the positions where they are and the forces of main.f90:100-141 (real space
    + k-sum, mean force removed).  FP64: six replicas spanning the sizes against the
    oracle (force_allpair + ewald_kspace + remove_rigidmtn): per-particle force
    <= 1e-10, Ur and Um <= 1e-12, RDF identical.  Mixed (autotuned shape): every
    replica's forces within 1e-6 RMS of the FP64 batch's."""
    from oracle import cpu
    from tcp_allpair_amd.batchbench import batch_workload
    say("case batch workload 64 replicas vs oracle")
    reps = batch_workload()
    n = np.array([p.n for p, _, _ in reps])
    order = np.argsort(n)
    picks = [int(order[k]) for k in np.linspace(0, len(reps) - 1, 6).round().astype(int)]
    res = {}
    for prec in (FP64, MIXED):
        with Batch([p for p, _, _ in reps], prec) as b:
            for r, (p, x, v) in enumerate(reps):
                b.set_state(r, x=x, v=np.zeros_like(v), f=np.zeros_like(x))
            if prec == MIXED:
                best, _ = b.autotune(1)
                say("mixed shape", best)
            tr = b.md_steps(1, NOTEMP, 0.0)
            st = [b.get_state(r) for r in range(len(reps))]
            h = [b.histograms(r)[0] for r in range(len(reps))]
        res[prec] = (tr, st, h)
    tr, st, h = res[FP64]
    for r in picks:
        p, x, _ = reps[r]
        say("replica", r, "n", p.n, "L", p.box[0])
        np.testing.assert_array_equal(st[r][0], x)
        fo, uro, rdfo = cpu.force_allpair(p, x)
        fko, umo = cpu.ewald_kspace(p, x)
        fo = cpu.remove_rigidmtn(p, fo + fko)
        f = st[r][2]
        nref = np.linalg.norm(fo, axis=1)
        err = np.linalg.norm(f - fo, axis=1) / np.maximum(nref, 1e-3 * np.sqrt(np.mean(nref ** 2)))
        assert err.max() <= 1e-10, (r, err.max())
        assert abs(tr[0, r, 0] - uro) <= 1e-12 * abs(uro), r
        assert abs(tr[0, r, 2] - umo) <= 1e-12 * abs(umo), r
        np.testing.assert_array_equal(h[r], rdfo)
    _, stm, _ = res[MIXED]
    for r in range(len(reps)):
        f, f64 = stm[r][2], st[r][2]
        e = np.sqrt(((f - f64) ** 2).sum() / (f64 ** 2).sum())
        assert e <= 1e-6, (r, e)


def test_autotune_and_graph_reuse():
    say("case batch autotune")
    reps = replicas()
    dt = reduced_dt(0.005)
    with Batch([p for p, _, _ in reps], MIXED) as b:
        for r, (p, x, v) in enumerate(reps):
            b.set_state(r, x=x, v=v, f=np.zeros_like(x))
        best, ms = b.autotune(2)
        say("autotune", best, ms)
        assert 0 <= best < Batch.NSHAPES
        assert np.all(np.isfinite(ms)) and np.all(ms > 0) and ms[best] == ms.min()
        pair_ms, _ = b.timings()
        assert pair_ms == ms[best]
        # the autotuner's launches leave the RDF counts alone
        assert all(not b.histograms(r)[0].any() for r in range(len(reps)))
        b.set_shape(0)
        t1 = b.md_steps(3, NOTEMP, dt)
        t2 = b.md_steps(3, NOTEMP, dt)  # replays the captured step
        s1 = [b.get_state(r) for r in range(len(reps))]
    tr, st, _ = run_batch(reps, MIXED, 6, NOTEMP, dt, shape=0)
    # replays of the captured step equal a fresh run bit for bit
    np.testing.assert_array_equal(np.concatenate([t1, t2]), tr)
    for r in range(len(reps)):
        np.testing.assert_array_equal(s1[r][0], st[r][0])


# SPME replicas: grids from 16^3 to 32^3, one anisotropic
SPME_K = [(16, 16, 16), (32, 32, 32), (16, 32, 16), (32, 32, 32), (32, 16, 32)]


def spme_replicas(order=4, flags=0):
    out = []
    for r, ((ne, ni), d, t) in enumerate(zip(SIZES, DENS, TEMP)):
        p, x, v = make_plasma(ne, ni, density=d, T=t * EPS, seed=140 + r, recip=RECIP_SPME,
                              K=SPME_K[r], order=order)
        out.append((p.replace(flags=flags), x, v))
    return out


@pytest.mark.parametrize("order", [4, 6])
def test_fp64_batch_spme_bitwise(order):
    """SPME in a batch: every replica's binning, spreading, five FFT passes and
    gather run over the replica table with the single-context kernel bodies,
    so each replica equals its own context's tcp_md_steps bit for bit."""
    say("case batch spme fp64 order", order)
    assert_bitwise(spme_replicas(order), FP64, 4, NOTEMP, reduced_dt(0.005))


def test_mixed_batch_spme_matches_standalone():
    say("case batch spme mixed")
    assert_bitwise(spme_replicas(flags=KELBG), MIXED, 3, NOTEMP, reduced_dt(0.005), shape=3)


def test_batch_spme_vs_oracle():
    """One SPME batch step with dt = 0 (positions stay, forces of main.f90:100-141):
    every replica's FP64 forces against the oracle's real space + SPME + mean-force
    removal <= 1e-10 per particle, Ur and Um <= 1e-12."""
    from oracle import cpu
    say("case batch spme vs oracle")
    reps = spme_replicas(order=6)
    with Batch([p for p, _, _ in reps], FP64) as b:
        for r, (p, x, v) in enumerate(reps):
            b.set_state(r, x=x, v=np.zeros_like(v), f=np.zeros_like(x))
        tr = b.md_steps(1, NOTEMP, 0.0)
        st = [b.get_state(r) for r in range(len(reps))]
    for r, (p, x, _) in enumerate(reps):
        fo, uro, _ = cpu.force_allpair(p, x)
        fso, umo = cpu.spme(p, x)
        fo = cpu.remove_rigidmtn(p, fo + fso)
        f = st[r][2]
        nref = np.linalg.norm(fo, axis=1)
        err = np.linalg.norm(f - fo, axis=1) / np.maximum(nref, 1e-3 * np.sqrt(np.mean(nref ** 2)))
        assert err.max() <= 1e-10, (r, err.max())
        assert abs(tr[0, r, 0] - uro) <= 1e-12 * abs(uro), r
        assert abs(tr[0, r, 2] - umo) <= 1e-12 * abs(umo), r


def test_batch_spme_needs_binned_grids():
    """Batched SPME takes brick-binned grids only (every K a multiple of 8)."""
    say("case batch spme grid")
    p, _, _ = make_plasma(200, 200, K=(4, 16, 8), recip=RECIP_SPME)
    q, _, _ = make_plasma(300, 300, K=(16, 16, 16), recip=RECIP_SPME)
    with pytest.raises(TcpError):
        Batch([q, p], FP64)
    r, _, _ = make_plasma(300, 300, K=(16, 16, 16), order=6, recip=RECIP_SPME)
    with pytest.raises(TcpError):
        Batch([q, r], FP64)  # one order per batch


def test_batch_errors():
    say("case batch errors")
    reps = replicas()
    with Batch([p for p, _, _ in reps], FP64) as b:
        with pytest.raises(TcpError):
            b.get_state(len(reps))
        with pytest.raises(TcpError):
            b.md_steps(1, BEREND, 0.001)
        with pytest.raises(TcpError):
            b.set_shape(1)  # FP64 batches have one shape
