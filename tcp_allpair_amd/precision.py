"""Precision workload: the mixed path against the FP64 path over a long NVE run.

65536 electrons + 65536 ions, reduced density 1, T* = 1, lambda = 0.5 A,
rc = L/2 with alpha = 3.1/rc, SPME 128^3 at order 6, NVE with dt = 0.002 tau*,
both precisions from the same seeded start and the same FP64 initial forces.

Two trajectories diverge chaotically, so forces and energies are compared on
identical positions: every `every` steps of the FP64 run, the mixed path
evaluates the full force (real space + SPME, mean force removed) and Ur on the
FP64 run's positions, against the FP64 forces and Ur there (Ur relative to
itself, a sum that a near-neutral plasma mostly cancels, and eps dUr relative
to E_tot).  Energy drift is each path's own: the slope of E_tot = KE +
eps (Ur + Uo) + Um per step over its run, divided by |E_tot(0)|, and the
largest |E_tot(t) - E_tot(0)| / |E_tot(0)|.

    python -m tcp_allpair_amd.precision [--steps 5000] [--every 1000]

prints progress to stderr and one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from .gpu import Context
from .inputs import make_plasma
from .params import FP64, MIXED, NOTEMP, reduced_dt


def say(*a):
    print("[precision]", *a, file=sys.stderr, flush=True)


def workload():
    return make_plasma(65536, 65536, density=1.0, seed=2024, K=(128, 128, 128), order=6)


def full_force(c, x):
    c.set_state(x=x, f=np.zeros_like(x))
    ur = c.force_direct()
    c.force_recip()
    c.remove_rigidmtn()
    return c.get_state()[2], ur


def force_errors(f, ref):
    """RMS relative error and the largest per-particle relative error, the
    per-particle denominator floored at 1e-3 of the RMS force."""
    d = np.linalg.norm(f - ref, axis=1)
    nref = np.linalg.norm(ref, axis=1)
    rms = math.sqrt((d ** 2).sum() / (nref ** 2).sum())
    floor = 1e-3 * math.sqrt(np.mean(nref ** 2))
    return rms, float((d / np.maximum(nref, floor)).max())


def etot(trace):
    """E_tot per step from tcp_md_steps rows (Ur, Uo, Um, Epot, mv2el, mv2ion)."""
    return 0.5 * (trace[:, 4] + trace[:, 5]) + trace[:, 3] + trace[:, 2]


def drift(e):
    steps = np.arange(len(e))
    return (float(np.polyfit(steps, e, 1)[0] / abs(e[0])),
            float(np.abs(e - e[0]).max() / abs(e[0])))


def run(steps=5000, every=1000):
    p, x, v = workload()
    dt = reduced_dt(0.002)
    t0 = time.perf_counter()
    with Context(p, FP64) as c64, Context(p, MIXED) as cmx:
        f0, _ = full_force(c64, x)
        checks = []
        trajectories = {}
        for name, c in (("fp64", c64), ("mixed", cmx)):
            c.set_state(x=x, v=v, f=f0)
            rows = []
            done = 0
            while done < steps:
                k = min(every, steps - done)
                rows.append(c.md_steps(k, NOTEMP, dt))
                done += k
                if name == "fp64":
                    # the mixed force and Ur on the FP64 run's positions; the FP64
                    # context keeps the forces of its last step at those positions
                    xs, vs, f64 = c.get_state()
                    ur64 = float(rows[-1][-1, 0])
                    e64 = float(etot(rows[-1])[-1])
                    fmx, urmx = full_force(cmx, xs)
                    rms, mx = force_errors(fmx, f64)
                    checks.append({"step": done, "force_rms_rel": rms, "force_max_rel": mx,
                                   "ur_rel": abs(urmx - ur64) / abs(ur64),
                                   "etot_rel": abs(p.eps * (urmx - ur64)) / abs(e64)})
                    say(f"fp64 step {done}: mixed force rms {rms:.3e} max {mx:.3e}")
                else:
                    say(f"mixed step {done}")
            trajectories[name] = etot(np.vstack(rows))
    out = {"workload": "tcp_n131072_spme128_order6_nve", "steps": steps, "dt_star": 0.002,
           "checks_on_fp64_positions": checks}
    for name, e in trajectories.items():
        slope, dev = drift(e)
        out[f"{name}_drift_per_step"] = slope
        out[f"{name}_max_rel_energy_deviation"] = dev
    out["wall_s"] = time.perf_counter() - t0
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--every", type=int, default=1000)
    a = ap.parse_args()
    print(json.dumps(run(a.steps, a.every)), flush=True)


if __name__ == "__main__":
    main()
