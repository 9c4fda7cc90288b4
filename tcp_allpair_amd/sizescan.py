"""Step time against system size: one context per size and precision.

Electrons and ions in equal numbers at reduced density 1, T* = 1,
lambda = 0.5 A, rc = L/2, no reciprocal part (the pair kernel and the
integrator alone), NVE with dt = 0.002 tau*, RDF/VDF every step.  Small mixed
systems run the 4 x 4 tile blocks of pair_mixed_n3_fine_kernel, larger ones
the 8 x 8 blocks.

    python -m tcp_allpair_amd.sizescan [--sizes 2000 8000 32768 131072] [--steps 50]

prints one JSON line per size and precision: ms per step (host clock around
tcp_md_steps) and the pair kernel's ms per launch (HIP events).
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from .gpu import Context
from .inputs import make_plasma
from .params import FP64, MIXED, NOTEMP, RECIP_NONE, reduced_dt


def run(n, precision, steps, warmup=5):
    p, x, v = make_plasma(n // 2, n - n // 2, seed=3, recip=RECIP_NONE)
    dt = reduced_dt(0.002)
    with Context(p, precision) as c:
        c.set_state(x=x, v=v, f=np.zeros_like(x))
        c.md_steps(warmup, NOTEMP, dt, trace=False)
        t0 = time.perf_counter()
        c.md_steps(steps, NOTEMP, dt, trace=False)
        ms = (time.perf_counter() - t0) / steps * 1e3
        pair_ms = c.timings()[0]
    return {"n": n, "precision": "mixed" if precision == MIXED else "fp64", "steps": steps,
            "ms_per_step": ms, "pair_kernel_ms": pair_ms,
            "pair_interactions_per_s": n * (n - 1) / 2 / (ms * 1e-3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", type=int, nargs="+", default=[2000, 8000, 32768, 131072])
    ap.add_argument("--steps", type=int, default=50)
    a = ap.parse_args()
    for n in a.sizes:
        for prec in (MIXED, FP64):
            print(json.dumps(run(n, prec, a.steps if n < 100000 or prec == MIXED else 5)), flush=True)


if __name__ == "__main__":
    main()
