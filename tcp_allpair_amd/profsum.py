"""Summaries of rocprofv3 CSV output for profiles/ (text, a few KB).

    python -m tcp_allpair_amd.profsum stats <kernel_stats.csv> <out.md> [note] [kernel_trace.csv steps]
    python -m tcp_allpair_amd.profsum batchstats <kernel_stats.csv> <out.md> [note]
    python -m tcp_allpair_amd.profsum pmc <fetch.csv> <write.csv> <copy_fetch.csv> <copy_write.csv> <out.md>
    python -m tcp_allpair_amd.profsum valu <counter_collection.csv> <out.md>

`stats` tabulates rocprofv3 --stats per kernel.  `pmc` scales FETCH_SIZE and
WRITE_SIZE (KiB per dispatch) by what the same counters read for copy_probe's
copy_kernel of known size (512 MiB in, 512 MiB out) and writes the HBM bytes
per launch of each kernel of the step; the line `hbm_bytes_per_launch:` holds
the pair kernel's figure that bench.py reports as roofline.traffic.  `valu`
tabulates the SQ instruction counters of the two pair kernels per launch and
per pair evaluation.
"""
from __future__ import annotations

import collections
import csv
import sys

PAIR_KERNEL = "pair_mixed_n3_kernel<true, false>"
FP64_PAIR_KERNEL = "pair_f64_n3_kernel<true, false>"
FP64_TIMED_STEPS = 2  # --fp64-steps of STATS_CMD
STATS_CMD = ("rocprofv3 --kernel-trace --stats --output-format csv -d gpurun_out/prof_stats -o bench "
             "-- python3 bench.py --steps 5 --warmup 1 --fp64-steps 2 --no-cpu-baseline --batch-steps 0 "
             "--dropin-steps 0")
BATCH_CMD = ("rocprofv3 --kernel-trace --stats --output-format csv -d gpurun_out/prof_batch -o batch "
             "-- python3 -m tcp_allpair_amd.batchbench --steps 5 --warmup 1")
PMC_CMDS = [
    "rocprofv3 --pmc FETCH_SIZE --output-format csv -d gpurun_out/pmc_fetch -o bench -- python3 bench.py "
    "--steps 2 --warmup 0 --fp64-steps 1 --batch-steps 0 --dropin-steps 0 --no-cpu-baseline",
    "rocprofv3 --pmc WRITE_SIZE --output-format csv -d gpurun_out/pmc_write -o bench -- python3 bench.py "
    "--steps 2 --warmup 0 --fp64-steps 1 --batch-steps 0 --dropin-steps 0 --no-cpu-baseline",
    "rocprofv3 --pmc FETCH_SIZE --output-format csv -d gpurun_out/pmc_copy_fetch -o copy -- "
    "tcp_allpair_amd/copy_probe 512 3",
    "rocprofv3 --pmc WRITE_SIZE --output-format csv -d gpurun_out/pmc_copy_write -o copy -- "
    "tcp_allpair_amd/copy_probe 512 3",
]
VALU_CMD = ("rocprofv3 --pmc SQ_INSTS_VALU SQ_INSTS_SALU SQ_WAVES SQ_BUSY_CYCLES SQ_WAVE_CYCLES "
            "SQ_INSTS_VALU_TRANS_F32 --output-format csv -d gpurun_out/pmc_valu -o bench -- python3 "
            "bench.py --steps 2 --warmup 0 --fp64-steps 1 --batch-steps 0 --dropin-steps 0 --no-cpu-baseline")
N_BENCH = 131072


def short(name: str) -> str:
    name = name.replace("tcp::(anonymous namespace)::", "").replace("void ", "")
    return name.split("(")[0]


def timed_average(trace: str, kernel: str, steps: int) -> float:
    """Average duration (ms) of the last `steps` dispatches of `kernel` in a
    rocprofv3 kernel trace: the launches of bench.py's timed region."""
    d = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
         for r in csv.DictReader(open(trace)) if short(r["Kernel_Name"]) == kernel]
    tail = d[-steps:]
    return sum(tail) / len(tail)


def stats(src: str, out: str, note: str = "", kind: str = "step", trace: str = "",
          steps: int = 0) -> None:
    rows = list(csv.DictReader(open(src)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    if kind == "batch":
        head = ["# Kernel time summary: 64-replica batch, mixed precision, direct k-sum kmax 6\n\n",
                "Workload: tcp_allpair_amd/batchbench.py (64 replicas, 526..16238 particles, 279616 in "
                "all); the autotuner times each of the 6 block shapes, then 1 warm-up and 5 timed "
                "batch steps; the same replicas then run one context after another (the "
                "single-context kernels below).\n\n", BATCH_CMD]
    else:
        head = ["# Kernel time summary: N = 131072 step, mixed and FP64 paths\n\n",
                "Workload: 65536 e + 65536 i, L = 50.797 A, SPME 128^3 order 6, RDF/VDF on; bench.py "
                "with 1 warm-up and 5 timed mixed steps, 2 FP64 steps, and one initial force pass per "
                "precision.\n\n", STATS_CMD]
    lines = head[:2] + ["Command (repository root on the GPU box, TMPDIR=/tmp):\n\n    " + head[2] + "\n\n",
                        "| kernel | calls | average ms | total ms | % |\n|---|---|---|---|---|\n"]
    for r in rows[:22]:
        lines.append("| %s | %s | %.4f | %.2f | %.2f |\n" % (
            short(r["Name"]), r["Calls"], float(r["AverageNs"]) / 1e6,
            float(r["TotalDurationNs"]) / 1e6, float(r["Percentage"])))
    if trace and steps:
        lines.append("\n%s over the %d timed steps (the last %d dispatches of the kernel trace; the "
                     "earlier ones run before the clocks settle): %.4f ms per launch.\n"
                     % (PAIR_KERNEL, steps, steps, timed_average(trace, PAIR_KERNEL, steps)))
        lines.append("\n%s over the %d timed FP64 steps: %.4f ms per launch.\n"
                     % (FP64_PAIR_KERNEL, FP64_TIMED_STEPS,
                        timed_average(trace, FP64_PAIR_KERNEL, FP64_TIMED_STEPS)))
    if note:
        lines.append("\n" + note + "\n")
    open(out, "w").write("".join(lines))


def _agg(path: str) -> dict:
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        d[short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in d.items()}


def pmc(fetch: str, write: str, cfetch: str, cwrite: str, out: str) -> None:
    fe, wr, cf, cw = _agg(fetch), _agg(write), _agg(cfetch), _agg(cwrite)
    known = 512 * 1024  # KiB moved each way by copy_kernel
    ffac, wfac = known / cf["copy_kernel"], known / cw["copy_kernel"]
    lines = ["# HBM traffic per launch (rocprofv3 --pmc), N = 131072 step\n\n",
             "Commands (repository root on the GPU box, one counter per run):\n\n"]
    lines += ["    " + c + "\n" for c in PMC_CMDS]
    lines.append("\nCalibration: copy_kernel moves %d KiB in and %d KiB out per launch; FETCH_SIZE reads "
                 "%.1f KiB (scale x%.3f), WRITE_SIZE reads %.1f KiB (scale x%.3f).\n\n"
                 % (known, known, cf["copy_kernel"], ffac, cw["copy_kernel"], wfac))
    lines.append("| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | scaled HBM bytes per launch |\n"
                 "|---|---|---|---|\n")
    res = {}
    for k in sorted(fe):
        if any(s in k for s in ("pair_", "spread", "gather", "fft_pass", "combine", "bin_", "grid_")):
            tot = (fe[k] * ffac + wr.get(k, 0.0) * wfac) * 1024
            res[k] = tot
            lines.append("| %s | %.1f | %.1f | %.3e |\n" % (k, fe[k], wr.get(k, 0.0), tot))
    lines.append("\nhbm_bytes_per_launch: %.6e\n(%s; bench.py reports this line as roofline.traffic)\n"
                 % (res[PAIR_KERNEL], PAIR_KERNEL))
    open(out, "w").write("".join(lines))


def valu(src: str, out: str) -> None:
    d = collections.defaultdict(lambda: collections.defaultdict(list))
    dur = collections.defaultdict(list)
    for r in csv.DictReader(open(src)):
        k = short(r["Kernel_Name"])
        if k in (PAIR_KERNEL, FP64_PAIR_KERNEL):
            d[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
            dur[k].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6)
    steps = N_BENCH * (N_BENCH - 1) / 2 / 64  # wave-wide pair evaluations (64 lanes)
    lines = ["# VALU instruction counts of the pair kernels (rocprofv3 --pmc), N = 131072\n\n",
             "Command (repository root on the GPU box, TMPDIR=/tmp):\n\n    " + VALU_CMD + "\n\n",
             "Per launch (average over the launches of the run; the counters are summed over the "
             "chip).  Per pair: SQ_INSTS_VALU / (N(N-1)/2 / 64), the wave instructions one lane "
             "spends per unordered pair, reactions included.  Cycles per VALU instruction per "
             "SIMD: launch time x 2.4 GHz x 1024 SIMDs / SQ_INSTS_VALU (the launch times of this "
             "counter run, nominal clock).\n\n",
             "| kernel | SQ_INSTS_VALU | of which TRANS_F32 | VALU per pair | SQ_INSTS_SALU | launch ms "
             "| cycles per VALU per SIMD |\n|---|---|---|---|---|---|---|\n"]
    for k in (PAIR_KERNEL, FP64_PAIR_KERNEL):
        if k not in d:
            continue
        c = {n: sum(v) / len(v) for n, v in d[k].items()}
        ms = sorted(dur[k])[len(dur[k]) // 2]
        v = c.get("SQ_INSTS_VALU", 0.0)
        lines.append("| %s | %.4e | %.4e | %.1f | %.4e | %.3f | %.2f |\n" % (
            k, v, c.get("SQ_INSTS_VALU_TRANS_F32", 0.0), v / steps, c.get("SQ_INSTS_SALU", 0.0), ms,
            ms * 1e-3 * 2.4e9 * 1024 / v if v else 0.0))
    open(out, "w").write("".join(lines))


if __name__ == "__main__":
    if sys.argv[1] == "valu":
        valu(sys.argv[2], sys.argv[3])
    elif sys.argv[1] in ("stats", "batchstats"):
        stats(sys.argv[2], sys.argv[3], sys.argv[4] if len(sys.argv) > 4 else "",
              "batch" if sys.argv[1] == "batchstats" else "step",
              sys.argv[5] if len(sys.argv) > 5 else "", int(sys.argv[6]) if len(sys.argv) > 6 else 0)
    else:
        pmc(*sys.argv[2:7])
