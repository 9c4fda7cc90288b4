"""Batching workload: 64 independent replicas of uneven size in one step launch.

Replica r: electrons N_e drawn log-uniform from 256 to 8192 (as many ions),
reduced density uniform in [0.5, 2], T* uniform in [0.5, 4], lambda = 0.5 A,
rc = L/2 with alpha = 3.1/rc, direct k-space sum with kmax = 6, mixed
precision, NVE with dt = 0.002 tau*.  All draws come from one seeded generator.

    python -m tcp_allpair_amd.batchbench [--steps 10] [--warmup 2]

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
from .params import EPS, MIXED, NOTEMP, RECIP_EWALD, reduced_dt


def batch_workload(nrep=64, seed=4242, kmax=6):
    rng = np.random.Generator(np.random.PCG64(seed))
    reps = []
    for r in range(nrep):
        ne = int(round(math.exp(rng.uniform(math.log(256), math.log(8192)))))
        dens = float(rng.uniform(0.5, 2.0))
        T = float(rng.uniform(0.5, 4.0))
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


def run(steps=10, warmup=2, nrep=64, seed=4242, precision=MIXED):
    reps = batch_workload(nrep, seed)
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
        "workload": "batch64_ewald_kmax6_mixed_nve", "replicas": nrep, "seed": seed,
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
    a = ap.parse_args()
    print(json.dumps(run(a.steps, a.warmup, a.replicas)), flush=True)


if __name__ == "__main__":
    main()
