"""Benchmark of the force-and-integrate step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (the metric's configuration): 65536 electrons + 65536 ions, reduced
density 1 (L = 50.797 A), T* = 1, lambda = 0.5 A, Ewald alpha = 3.1/rc with
rc = L/2, SPME on a 128^3 grid at B-spline order 6, NVE velocity Verlet with
dt = 0.002 tau*, RDF and VDF accumulated every step as the reference does.
A step is main.f90:87-153 on the device: kick-drift, SPME, the all-pair real
space sum, force sum, mean-force removal, kick.  Inputs are synthetic (seeded)
and resident in HBM before the timed region; the timed region is K calls of
the step bracketed by a device synchronisation on both sides.

Prints one JSON line.  value = pair-interactions/s of the mixed-precision path
(N(N-1)/2 unordered pairs per step x steps/s); the FP64 path (with its pair
kernel's rate in the same flop count against the 78.6 TFLOP/s FP64 vector
peak), their ratio and the RMS relative force error of the mixed path against
FP64 are reported beside it.  roofline: the pair kernel's algorithmic FP32 flops (68.5 per
unordered pair: 62.5 per pair plus 6 for the Newton's-third-law reaction) over
its average launch time from HIP events on its stream, against the
157.3 TFLOP/s FP32 vector peak.  cpu_baseline: the reference's own
Force_direct_ext (compiled from source, oracle/_ref) timed on the host cores on
a bounded row slice of the same workload.  batch: the 64-replica batching
workload of tcp_allpair_amd/batchbench.py (one persistent pair launch and a
captured step for all replicas) against the same replicas stepped one
context after another.  dropin_host_buffers: the step driven as the Fortran
drop-in drives it, the host arrays coming back every step (the PCIe-inclusive
rate; never value).  nve_drift: the slope of the total energy per step over
1000 NVE steps of the mixed path divided by |E_tot(0)| (nve_energy_slope_timed_steps
is the same slope over the timed steps only, a fluctuation measure).
roofline_fp64: the FP64 pair kernel in the same flop count against the
78.6 TFLOP/s FP64 vector peak.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from tcp_allpair_amd.gpu import Context  # noqa: E402
from tcp_allpair_amd.inputs import make_plasma  # noqa: E402
from tcp_allpair_amd.params import FP64, MIXED, NOTEMP, reduced_dt  # noqa: E402

# Algorithmic FP32 flops: 62.5 per ordered pair (FMA = 2; exp, rsqrt, rcp, rint = 1;
# the e-i / e-e mix), and 62.5 + 6 per unordered pair for a kernel that uses
# Newton's third law (the reaction accumulate), which the mixed kernel does.
FLOP_PER_UNORDERED_PAIR_NEWTON = 62.5 + 6.0
FP32_VECTOR_PEAK_TFLOPS = 157.3
FP32_NONPACKED_PEAK_TFLOPS = 78.6  # one FP32 FMA per lane per cycle, no packed math
FP64_VECTOR_PEAK_TFLOPS = 78.6
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "pair_mixed_n131072_pmc.md")


def workload():
    p, x, v = make_plasma(65536, 65536, density=1.0, seed=2024, K=(128, 128, 128), order=6)
    return p, x, v


def initial_forces(p, x, precision):
    with Context(p, precision) as c:
        c.set_state(x=x, f=np.zeros_like(x))
        c.force_direct()
        c.force_recip()
        c.remove_rigidmtn()
        return c.get_state()[2]


def run(p, x, v, f0, precision, steps, warmup, dt):
    with Context(p, precision) as c:
        c.set_state(x=x, v=v, f=f0)
        if warmup:
            c.md_steps(warmup, NOTEMP, dt, trace=False)
        t0 = time.perf_counter()
        tr = c.md_steps(steps, NOTEMP, dt, trace=True)   # returns after a stream sync
        t1 = time.perf_counter()
        pair_ms, steps_ms = c.timings()
    return dict(wall_s=t1 - t0, dev_ms=steps_ms, pair_ms=pair_ms, trace=tr)


def dropin_run(p, x, v, f0, precision, steps, dt):
    """The step as the Fortran drop-in drives it with one direct rank
    (tcp_allpair_amd/fortran/tcp_force_gpu.f90), the host state (AoS x, v, f)
    and the histograms coming back every step: an array is uploaded only when
    the host copy changed (here it never does between steps) and the step's
    results stay on the device from VVerlet*1 to VVerlet*2.  Per step:
    vverlet1, force_direct, force_recip, remove_rigidmtn, vverlet2, then one
    download of 9N doubles (x, v, f) and one of the RDF and VDF counts.  The
    shim's own loops between the PT records and these arrays are not
    included."""
    import ctypes
    with Context(p, precision) as c:
        lib, h = c.lib, c.h
        hx, hv, hf = (np.ascontiguousarray(a, dtype=np.float64).copy() for a in (x, v, f0))
        px, pv, pf = (a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for a in (hx, hv, hf))
        rdf, vdf = np.zeros(3 * 501), np.zeros(2 * 501)
        prdf, pvdf = (a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for a in (rdf, vdf))
        ur, um, a, b = (ctypes.c_double() for _ in range(4))

        def ck(rc, what):
            if rc != 0:
                raise RuntimeError(f"{what}: {lib.tcp_last_error(h).decode()}")

        def step():
            ck(lib.tcp_vverlet1(h, NOTEMP, dt, 0.0), "vverlet1")
            ck(lib.tcp_force_direct(h, ctypes.byref(ur)), "force_direct")
            ck(lib.tcp_force_recip(h, ctypes.byref(um)), "force_recip")
            ck(lib.tcp_remove_rigidmtn(h), "remove_rigidmtn")
            ck(lib.tcp_vverlet2(h, NOTEMP, dt, p.Te, p.Ti, p.tau, 0.0, 1, ctypes.byref(a),
                                ctypes.byref(b)), "vverlet2")
            ck(lib.tcp_get_state(h, px, pv, pf), "get_state")
            ck(lib.tcp_get_histograms(h, prdf, pvdf, 1), "get_histograms")

        ck(lib.tcp_set_state(h, px, pv, pf), "set_state")
        step()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        return (time.perf_counter() - t0) / steps * 1e3


def cpu_baseline(p, x):
    """The reference's Force_direct_ext on host cores: each worker process times
    one call on its own 512-particle row slice against all N particles (after one
    warm call; a few seconds each).  At most 16 processes: the share of host
    cores that goes with one GPU of the box."""
    from oracle import refharness as rh
    host_cores = os.cpu_count() or 1
    cores = max(1, min(16, host_cores))
    n_own = 512
    if not rh.available():
        return None
    import tempfile
    procs = []
    with tempfile.TemporaryDirectory() as td:
        for w in range(cores):
            # slice w: electrons and ions starting at a per-worker offset
            e0 = (w * n_own // 2) % (p.ne - n_own // 2)
            i0 = (w * n_own // 2) % (p.ni - n_own // 2)
            own = np.concatenate([x[e0:e0 + n_own // 2], x[p.ne + i0:p.ne + i0 + n_own // 2]])
            rest_e = np.concatenate([x[:e0], x[e0 + n_own // 2:p.ne]])
            rest_i = np.concatenate([x[p.ne:p.ne + i0], x[p.ne + i0 + n_own // 2:]])
            xx = np.concatenate([own[:n_own // 2], rest_e, own[n_own // 2:], rest_i])
            fin = os.path.join(td, f"in{w}.bin")
            fout = os.path.join(td, f"out{w}.bin")
            rh._write_input(fin, rh.MODE_TIME, p, xx, None, None, nsteps=1,
                            nown=(n_own // 2, n_own // 2))
            procs.append((subprocess.Popen([rh.HARNESS, fin, fout], stdout=subprocess.DEVNULL,
                                           stderr=subprocess.DEVNULL), fout))
        rates, secs = [], []
        for pr, fout in procs:
            if pr.wait(timeout=300) != 0:
                return None
            pairs, sec, _ = np.fromfile(fout, dtype="<f8", count=3)
            rates.append(pairs / sec)
            secs.append(sec)
    rate = float(sum(rates))
    return dict(value=rate, unit="pair-interactions/s", cores=cores, kind="reference",
                host_cores=host_cores, per_process_rate=float(np.mean(rates)),
                seconds_per_process=float(np.mean(secs)),
                # one step's N(N-1)/2 pair interactions at this rate
                pair_seconds_per_step=p.n * (p.n - 1) / 2 / rate,
                sample=f"Force_direct_ext (force.f90:183) of a {n_own}-particle row slice against "
                       f"all {p.n} particles per process ({n_own * p.n} ordered pairs), {cores} "
                       "single-threaded processes in parallel (at most 16: the host-core share of "
                       "one GPU), one timed call each after one warm call; rate summed over "
                       "processes")


def nve_drift(p, x, v, f0, steps, dt):
    """Slope of the total energy per step over `steps` NVE steps of the mixed
    path, divided by |E_tot(0)| (least squares over every step)."""
    with Context(p, MIXED) as c:
        c.set_state(x=x, v=v, f=f0)
        tr = c.md_steps(steps, NOTEMP, dt, trace=True)
    etot = 0.5 * (tr[:, 4] + tr[:, 5]) + tr[:, 3] + tr[:, 2]
    slope = np.polyfit(np.arange(len(etot)), etot, 1)[0]
    return float(slope / abs(etot[0])), float(np.ptp(etot) / abs(etot[0]))


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the rocprofv3 --pmc summary (scaled
    FETCH_SIZE + WRITE_SIZE), or None."""
    if not os.path.exists(PROFILE_TRAFFIC):
        return None
    for line in open(PROFILE_TRAFFIC):
        cols = [c.strip() for c in line.split("|")]
        if len(cols) > 4 and cols[1] == kernel:
            return float(cols[4])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--fp64-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=10,
                    help="steps of the 64-replica batching workload (0 skips it)")
    ap.add_argument("--dropin-steps", type=int, default=5,
                    help="steps driven as the Fortran drop-in drives them (0 skips it)")
    ap.add_argument("--drift-steps", type=int, default=1000,
                    help="NVE steps of the mixed path for the energy drift (0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the multi-GPU design is "
              "documented but not built", file=sys.stderr)
        sys.exit(2)
    p, x, v = workload()
    dt = reduced_dt(0.002)
    N = p.n
    pairs_unordered = N * (N - 1) / 2
    print(f"[bench] N={N} K={p.K} order={p.order}: initial forces", file=sys.stderr, flush=True)
    f64_0 = initial_forces(p, x, FP64)
    fmx_0 = initial_forces(p, x, MIXED)
    err = math.sqrt(((fmx_0 - f64_0) ** 2).sum() / (f64_0 ** 2).sum())
    print(f"[bench] mixed vs fp64 rms rel force err {err:.3e}; timing mixed", file=sys.stderr, flush=True)
    mx = run(p, x, v, f64_0, MIXED, args.steps, args.warmup, dt)
    print(f"[bench] mixed {mx['wall_s'] / args.steps * 1e3:.2f} ms/step; timing fp64", file=sys.stderr,
          flush=True)
    fp = run(p, x, v, f64_0, FP64, args.fp64_steps, 2, dt)
    ms_step = mx["wall_s"] / args.steps * 1e3
    sps = args.steps / mx["wall_s"]
    sps64 = args.fp64_steps / fp["wall_s"]
    tr = mx["trace"]
    etot = 0.5 * (tr[:, 4] + tr[:, 5]) + tr[:, 3] + tr[:, 2]
    flops = N * (N - 1) / 2 * FLOP_PER_UNORDERED_PAIR_NEWTON
    achieved = flops / (mx["pair_ms"] * 1e-3) / 1e12
    achieved64 = flops / (fp["pair_ms"] * 1e-3) / 1e12
    drift = None
    if args.drift_steps > 1:
        print(f"[bench] nve drift over {args.drift_steps} mixed steps", file=sys.stderr, flush=True)
        d, excursion = nve_drift(p, x, v, f64_0, args.drift_steps, dt)
        drift = {"steps": args.drift_steps, "drift_per_step": d, "max_excursion": excursion}
    dropin = None
    if args.dropin_steps > 0:
        print("[bench] host-buffer steps (the drop-in's call pattern)", file=sys.stderr, flush=True)
        dropin = {"steps": args.dropin_steps,
                  "mixed_ms_per_step": dropin_run(p, x, v, f64_0, MIXED, args.dropin_steps, dt),
                  "fp64_ms_per_step": dropin_run(p, x, v, f64_0, FP64, args.dropin_steps, dt),
                  "transfers_per_step": "one download of 9N doubles (x, v, f) and one of the "
                                        "histograms; no uploads while the host arrays are "
                                        "unchanged"}
    batch = None
    if args.batch_steps > 0:
        print("[bench] batching workload (64 replicas)", file=sys.stderr, flush=True)
        from tcp_allpair_amd import batchbench
        batch = batchbench.run(args.batch_steps, 2)
    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (reference on host cores)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(p, x)
    out = {
        "metric": "pair-interactions/s + steps/s at N=131072, mixed vs FP64 GPU at force err<=1e-6",
        "value": pairs_unordered * sps,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "tcp_n131072_spme128_order6_nve", "N": N, "ne": p.ne, "ni": p.ni,
                   "box": p.box[0], "rc": p.rc, "alpha": p.ewald_alpha, "lambda": p.rs,
                   "K": list(p.K), "order": p.order, "dt_star": 0.002, "thermo": "NOTEMP",
                   "precision": "mixed: FP32 pair arithmetic, FP64 accumulation/SPME/integrator",
                   "rdf_vdf": True},
        "steps_per_s": sps,
        "pair_interactions_per_s_ordered": 2 * pairs_unordered * sps,
        "fp64": {"steps_per_s": sps64, "pair_interactions_per_s": pairs_unordered * sps64,
                 "ms_per_step": fp["wall_s"] / args.fp64_steps * 1e3,
                 "pair_kernel_ms": fp["pair_ms"],
                 "pair_kernel_tflops": flops / (fp["pair_ms"] * 1e-3) / 1e12,
                 "pair_kernel_frac_of_fp64_peak": flops / (fp["pair_ms"] * 1e-3) / 1e12
                 / FP64_VECTOR_PEAK_TFLOPS},
        "mixed_over_fp64_steps": sps / sps64,
        "force_err_rms_vs_fp64": err,
        "nve_energy_drift_per_step": None if drift is None else drift["drift_per_step"],
        "nve_drift": drift,
        "nve_energy_slope_timed_steps": float(np.polyfit(np.arange(len(etot)), etot, 1)[0]
                                              / abs(etot[0])) if len(etot) > 1 else None,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_VECTOR_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_VECTOR_PEAK_TFLOPS,
                     "traffic": profiled_traffic("pair_mixed_n3_kernel<true, false>"),
                     "kernel": "pair_mixed_n3_kernel<true, false>",
                     "kernel_ms": mx["pair_ms"], "flop_per_launch": flops,
                     # packed FP32 (v_pk_*_f32) is disabled for all device code, so
                     # the kernel issues non-packed FP32, whose ceiling is half the peak
                     "peak_nonpacked": FP32_NONPACKED_PEAK_TFLOPS,
                     "frac_nonpacked": achieved / FP32_NONPACKED_PEAK_TFLOPS},
        "roofline_fp64": {"bound": "valu", "achieved": achieved64, "peak": FP64_VECTOR_PEAK_TFLOPS,
                          "unit": "TFLOP/s", "frac": achieved64 / FP64_VECTOR_PEAK_TFLOPS,
                          "traffic": profiled_traffic("pair_f64_n3_kernel<true, false>"),
                          "kernel": "pair_f64_n3_kernel<true, false>",
                          "kernel_ms": fp["pair_ms"], "flop_per_launch": flops},
        "cpu_baseline": cpu,
        "dropin_host_buffers": dropin,
        "batch": batch,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
