"""Batching workload: 64 independent replicas of uneven size in one step launch.

Replica r: electrons N_e drawn log-uniform from 256 to 8192 (as many ions),
reduced density uniform in [0.5, 2], T* uniform in [0.5, 4], lambda = 0.5 A,
rc = L/2 with alpha = 3.1/rc, direct k-space sum with kmax = 6, mixed
precision, NVE with dt = 0.002 tau*.  All draws come from one seeded generator.

    python -m tcp_allpair_amd.batchbench [--steps 10] [--warmup 2] [--spme]

--spme runs the same replicas with SPME (order 4, each on its own grid of
spacing <= 1 A) in place of the k-sum.

prints one JSON line: the autotuned block shape, the batch's aggregate
pair-interactions/s and steps/s, and the same replicas stepped one context
after another (each with its own tcp_md_steps call).
"""
from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np

from .gpu import Batch, Context
from .inputs import make_plasma
from .params import EPS, MIXED, NOTEMP, RECIP_EWALD, RECIP_SPME, reduced_dt


def spme_grid(L):
    """SPME grid edge for a box of edge L (A): the power of two (at least 16) that
    keeps the spacing at or below 1 A."""
    k = 16
    while k < L:
        k *= 2
    return k


def batch_workload(nrep=64, seed=4242, kmax=6, recip=RECIP_EWALD):
    """recip = RECIP_SPME replaces the k-sum by SPME at order 4 on each replica's
    own grid (spme_grid of its box), the same draws otherwise."""
    rng = np.random.Generator(np.random.PCG64(seed))
    reps = []
    for r in range(nrep):
        ne = int(round(math.exp(rng.uniform(math.log(256), math.log(8192)))))
        dens = float(rng.uniform(0.5, 2.0))
        T = float(rng.uniform(0.5, 4.0))
        if recip == RECIP_SPME:
            L = (2 * ne / dens) ** (1.0 / 3.0)
            K = spme_grid(L)
            p, x, v = make_plasma(ne, ne, density=dens, T=T * EPS, seed=seed + 1 + r,
                                  recip=RECIP_SPME, K=(K, K, K), order=4, lam=0.5)
        else:
            p, x, v = make_plasma(ne, ne, density=dens, T=T * EPS, seed=seed + 1 + r,
                                  recip=RECIP_EWALD, kmax=kmax, lam=0.5)
        reps.append((p, x, v))
    return reps


def initial_forces(b: Batch, reps):
    """Forces before the first kick: a zero-length start, i.e. one step with dt = 0."""
    for r, (p, x, v) in enumerate(reps):
        b.set_state(r, x=x, v=np.zeros_like(v), f=np.zeros_like(x))
    b.md_steps(1, NOTEMP, 0.0, trace=False)
    return [b.get_state(r)[2] for r in range(len(reps))]


def run(steps=10, warmup=2, nrep=64, seed=4242, precision=MIXED, recip=RECIP_EWALD):
    reps = batch_workload(nrep, seed, recip=recip)
    dt = reduced_dt(0.002)
    n = np.array([p.n for p, _, _ in reps], dtype=np.float64)
    pairs = float((n * (n - 1) / 2).sum())
    with Batch([p for p, _, _ in reps], precision) as b:
        f0 = initial_forces(b, reps)
        for r, (p, x, v) in enumerate(reps):
            b.set_state(r, x=x, v=v, f=f0[r])
        best, shape_ms = b.autotune(3)
        if warmup:
            b.md_steps(warmup, NOTEMP, dt, trace=False)
        t0 = time.perf_counter()
        b.md_steps(steps, NOTEMP, dt, trace=False)
        t1 = time.perf_counter()
        pair_ms, dev_ms = b.timings()
    batch_s = t1 - t0
    # the same replicas, one context after another
    sep_s = 0.0
    sep_pair_ms = 0.0
    for r, (p, x, v) in enumerate(reps):
        with Context(p, precision) as c:
            c.set_state(x=x, v=v, f=f0[r])
            if warmup:
                c.md_steps(warmup, NOTEMP, dt, trace=False)
            t0 = time.perf_counter()
            c.md_steps(steps, NOTEMP, dt, trace=False)
            sep_s += time.perf_counter() - t0
            sep_pair_ms += c.timings()[0]
    return {
        "workload": ("batch64_ewald_kmax6_mixed_nve" if recip == RECIP_EWALD
                     else "batch64_spme_order4_mixed_nve"), "replicas": nrep, "seed": seed,
        "n_total": int(n.sum()), "n_min": int(n.min()), "n_max": int(n.max()),
        "pairs_unordered_per_step": pairs, "steps": steps, "warmup": warmup,
        "shape": best, "shape_pair_ms": [float(v) for v in shape_ms],
        "batch_ms_per_step": batch_s / steps * 1e3,
        "batch_device_ms_per_step": dev_ms / steps,
        "batch_pair_ms": pair_ms,
        "batch_pair_interactions_per_s": pairs * steps / batch_s,
        "batch_steps_per_s_per_replica": steps / batch_s,
        "separate_ms_per_step": sep_s / steps * 1e3,
        "separate_pair_ms": sep_pair_ms,
        "separate_pair_interactions_per_s": pairs * steps / sep_s,
        "separate_steps_per_s_per_replica": steps / sep_s,
        "speedup": sep_s / batch_s,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--replicas", type=int, default=64)
    ap.add_argument("--spme", action="store_true",
                    help="SPME (order 4, grid spacing <= 1 A) instead of the k-sum")
    a = ap.parse_args()
    print(json.dumps(run(a.steps, a.warmup, a.replicas,
                         recip=RECIP_SPME if a.spme else RECIP_EWALD)), flush=True)


if __name__ == "__main__":
    main()
