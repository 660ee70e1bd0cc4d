"""Benchmark of the GARLI likelihood engine on one MI355X.

Workload (the metric's configuration): nucleotide GTR+I+G4, 128-taxon Yule
tree, branch lengths ~Exp(0.05), 100,000 simulated site patterns, pinv 0.2,
alpha ~U(0.2, 1.5) -- synthetic, seeded.  One step is a full-tree sweep: every
directional CLA (3*(N-2) = 378 CLA updates of all patterns, with rescaling)
followed by (lnL, d1, d2) on all 2N-3 = 253 edges, inputs resident in HBM.

value = CLA site-updates/s of the FP32-bulk form (378 * 100,000 per step).
Also reported: full-tree lnL evaluations/s with p50/p99 latency (126 CLA
updates + one edge integration per evaluation), the FP32-vs-FP64 |dlnL| and
derivative errors, the FP64 drop-in against the CPU oracle, the CLA kernel's
roofline fraction from HIP-event timing, and the CPU baseline.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--precision fp32|fp64]
                    [--schedule levels|fused] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "lnL evals/s and CLA site-updates/s at |ΔlnL|<1e-3 vs FP64; p99 latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--schedule", choices=["levels", "fused"], default="fused")
    ap.add_argument("--evals", type=int, default=60, help="timed full-tree lnL evaluations")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle legs")
    ap.add_argument("--patterns", type=int, default=100000)
    ap.add_argument("--taxa", type=int, default=128)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this engine has no multi-GPU path "
              "(pattern sharding across GPUs is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    from garli_amd import engine as E
    from garli_amd import sim

    t0 = time.time()
    ds = sim.gtr_gamma(n_tips=args.taxa, n_sites=None, n_patterns=args.patterns, mean_blen=0.05, alpha=None,
                       pinv=0.2, ambig_frac=0.0, seed=2)
    log("simulated %d patterns (%d sites) on %d taxa in %.1fs, alpha=%.3f" %
        (ds.n_patterns, int(ds.counts.sum()), ds.n_tips, time.time() - t0, ds.meta["alpha"]))
    E.init(0)
    N, P = ds.n_tips, ds.n_patterns
    n_cla = 3 * (N - 2)
    n_edges = 2 * N - 3
    prec = E.PREC_FP32 if args.precision == "fp32" else E.PREC_FP64

    inst = ds.make_instance(prec)
    tree = E.Tree(inst, ds.parent, ds.blen)
    tree.set_schedule(1 if args.schedule == "fused" else 0)

    for _ in range(args.warmup):
        tree.sweep()
    times = []
    for _ in range(args.steps):
        a = time.perf_counter()
        sw = tree.sweep()  # returns after the device work drained
        times.append(time.perf_counter() - a)
    ms_step = 1000.0 * float(np.mean(times))
    site_updates = n_cla * P / (ms_step / 1000.0)
    log("sweep: %.3f ms/step, %.3e CLA site-updates/s" % (ms_step, site_updates))

    # roofline of the dominant kernel (CLA update), HIP events on the instance stream
    inst.set_timing(True)
    for k in range(3):
        inst.kernel_stats(k, reset=True)
    for _ in range(max(3, args.steps // 2)):
        tree.sweep()
    inst.set_timing(False)
    n_l, ms_l, bytes_l = inst.kernel_stats(0)
    e_l, e_ms, e_bytes = inst.kernel_stats(1)
    p_l, p_ms, _ = inst.kernel_stats(2)
    avg_ms = ms_l / max(n_l, 1)
    achieved = bytes_l / max(ms_l, 1e-12) / 1e6  # GB/s
    log("CLA kernel: %d launches, avg %.3f ms, %.1f GB/s; edge kernel avg %.3f ms %.1f GB/s; pmat avg %.3f ms" %
        (n_l, avg_ms, achieved, e_ms / max(e_l, 1), e_bytes / max(e_ms, 1e-12) / 1e6, p_ms / max(p_l, 1)))

    # full-tree lnL evaluations (MakeAllNodesDirty + Score), latency distribution
    lat = []
    for i in range(args.evals + 3):
        tree.invalidate()
        a = time.perf_counter()
        lnl32 = tree.score()
        if i >= 3:
            lat.append(time.perf_counter() - a)
    lat = np.array(lat) * 1000.0
    evals_per_s = 1000.0 / float(np.mean(lat))
    log("lnL eval: %.3f ms mean, p50 %.3f, p99 %.3f -> %.1f evals/s" %
        (lat.mean(), np.percentile(lat, 50), np.percentile(lat, 99), evals_per_s))

    # accuracy: this form vs the FP64 drop-in on the GPU (same inputs)
    acc = {}
    if prec == E.PREC_FP32:
        inst64 = ds.make_instance(E.PREC_FP64)
        t64 = E.Tree(inst64, ds.parent, ds.blen)
        sw64 = t64.sweep()
        lnl64 = t64.score()
        acc["dlnl_vs_fp64"] = abs(lnl32 - lnl64)
        acc["max_edge_dlnl_vs_fp64"] = float(np.max(np.abs(sw[:, 0] - sw64[:, 0])))
        acc["max_d1_err_vs_fp64"] = float(np.max(np.abs(sw[:, 1] - sw64[:, 1]) / np.maximum(1.0, np.abs(sw64[:, 1]))))
        acc["max_d2_err_vs_fp64"] = float(np.max(np.abs(sw[:, 2] - sw64[:, 2]) / np.maximum(1.0, np.abs(sw64[:, 2]))))
        log("fp32 vs fp64: |dlnL| %.3e, max edge %.3e, d1 %.2e, d2 %.2e" %
            (acc["dlnl_vs_fp64"], acc["max_edge_dlnl_vs_fp64"], acc["max_d1_err_vs_fp64"], acc["max_d2_err_vs_fp64"]))
        inst64.close()
        ref_sw = sw64
    else:
        ref_sw = sw

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        # FP64 oracle (parity build) on the full workload: pins the GPU numbers
        o = O.Oracle(ds)
        a = time.time()
        osw = o.sweep()
        log("oracle parity sweep %.1fs" % (time.time() - a))
        acc["fp64_gpu_vs_oracle_max_rel_lnl"] = float(np.max(np.abs(ref_sw[:, 0] - osw[:, 0]) / np.abs(osw[:, 0])))
        acc["this_form_vs_oracle_max_dlnl"] = float(np.max(np.abs(sw[:, 0] - osw[:, 0])))
        # CPU baseline: the same sweep, oracle built with the reference's release flags, one core
        of = O.Oracle(ds, fast=True)
        a = time.time()
        of.sweep()
        dt = time.time() - a
        cpu = {"value": n_cla * P / dt, "unit": "CLA site-updates/s", "cores": 1, "kind": "port",
               "sample": "one full sweep of the same workload (378 CLA updates x %d patterns + %d edge "
                         "derivatives), FP64, -O3 -ffast-math, %.1fs" % (P, n_edges, dt)}
        log("cpu baseline %.3e site-updates/s (%.1fs)" % (cpu["value"], dt))

    out = {
        "metric": METRIC,
        "value": site_updates,
        "unit": "CLA site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == E.PREC_FP32 else "f64",
        "data": "synthetic (seeded GTR+I+G4 simulation on a Yule tree)",
        "config": {"workload": "nucleotide GTR+I+G4 full-tree sweep, %d taxa, %d patterns" % (N, P),
                   "taxa": N, "patterns": P, "states": 4, "rate_cats": 4, "pinv": 0.2,
                   "alpha": round(ds.meta["alpha"], 4), "form": "fp32-bulk" if prec == E.PREC_FP32 else "fp64",
                   "schedule": args.schedule, "cla_updates_per_step": n_cla, "edges_per_step": n_edges},
        "lnl_evals_per_s": evals_per_s,
        "lnl_eval_p50_ms": float(np.percentile(lat, 50)),
        "lnl_eval_p99_ms": float(np.percentile(lat, 99)),
        "accuracy": acc,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_cla_nuc<%s,4>" % ("float" if prec == E.PREC_FP32 else "double"),
                     "avg_launch_ms": avg_ms, "launches": n_l,
                     "bytes_per_site_update": 3 * 4 * 4 * (4 if prec == E.PREC_FP32 else 8) + 12},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
