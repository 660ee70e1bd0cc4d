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
PEAK_FP64_TFLOPS = 78.6  # MI355X published FP64 (vector and matrix)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(key):
    """HBM bytes per launch of a kernel on this workload from the committed
    rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, each scaled by what the
    counter reads for a copy kernel of known size), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "pmc_traffic.json")) as f:
            return json.load(f)[key]["traffic_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--schedule", choices=["levels", "fused"], default="levels")
    ap.add_argument("--evals", type=int, default=60, help="timed full-tree lnL evaluations")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle legs")
    ap.add_argument("--no-codon", action="store_true",
                    help="skip the codon served-vs-plain leg of the default workload")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the GA request-stream leg of the default workload")
    ap.add_argument("--patterns", type=int, default=100000)
    ap.add_argument("--taxa", type=int, default=128)
    ap.add_argument("--quick", action="store_true",
                    help="sweeps only (for profiler runs): no evals, accuracy or CPU legs")
    ap.add_argument("--calibrate", type=float, default=0.0,
                    help="also run a plain copy of this many GiB (counter calibration)")
    ap.add_argument("--leg", choices=["serve", "plain"], default=None,
                    help="aa / codon: run only the served or only the plain FP64 request leg (profiler runs)")
    ap.add_argument("--workload", choices=["gtr", "aa", "codon", "serve"], default="gtr",
                    help="gtr: the metric's configuration (default); aa / codon: the 20- and 61-state "
                         "configurations; serve: Poisson-arrival GA requests through the serving layer")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this engine has no multi-GPU path "
              "(pattern sharding across GPUs is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    from garli_amd import engine as E
    from garli_amd import sim

    if args.workload in ("aa", "codon"):
        print(json.dumps(bench_leg(args) if args.leg else bench_states(args)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(bench_serve(args)), flush=True)
        return

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
    n_heavy = inst.heavy_patterns()
    if n_heavy:
        log("%d heavily weighted patterns run in FP64 beside the FP32 bulk" % n_heavy)
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
    for k in range(6):
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
    h_l, h_ms, _ = inst.kernel_stats(3)
    he_l, he_ms, _ = inst.kernel_stats(4)
    if h_l:
        log("FP64 heavy part: CLA %d launches avg %.3f ms, edge avg %.3f ms (own stream)" %
            (h_l, h_ms / h_l, he_ms / max(he_l, 1)))

    copy = None
    if args.calibrate > 0:
        nbytes = int(args.calibrate * (1 << 30))
        ms = E.copy_calibrate(nbytes, 10)
        copy = {"bytes": nbytes, "ms": ms, "GB/s": 2 * nbytes / ms / 1e6}
        log("copy calibration: %d bytes in %.3f ms -> %.1f GB/s (read+write)" % (nbytes, ms, copy["GB/s"]))
    if args.quick:
        args.no_codon = args.no_serve = True
        print(json.dumps({"metric": METRIC, "value": site_updates, "unit": "CLA site-updates/s",
                          "ms_per_step": ms_step, "cla_avg_launch_ms": avg_ms, "cla_GBs": achieved,
                          "launches": n_l, "copy": copy}), flush=True)
        return

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
        n_sw = 4
        a = time.time()
        for _ in range(n_sw):
            of.sweep()
        dt = (time.time() - a) / n_sw
        one = n_cla * P / dt
        log("cpu baseline, one thread: %.3e site-updates/s (%.1fs per sweep)" % (one, dt))
        del o, of
        T = _cpu_threads()
        parts = _pattern_slices(O, ds, T)
        n_mt = 12
        wall = _threaded([lambda q=q: [q.sweep() for _ in range(n_mt)] for q in parts])
        cpu = {"value": n_mt * n_cla * P / wall, "unit": "CLA site-updates/s", "cores": len(parts), "kind": "port",
               "single_thread_value": one, "host_cores_allowed": _host_cores()[0],
               "host_cores_reported": _host_cores()[1],
               "sample": "%d full sweeps of the same workload (%d CLA updates x %d patterns + %d edge derivatives "
                         "each), FP64, patterns split over %d threads as the reference's OpenMP build splits "
                         "them, oracle built -O3 -ffast-math, %.1fs wall; one thread: %d sweeps, %.1fs" %
                         (n_mt, n_cla, P, n_edges, len(parts), wall, n_sw, dt * n_sw)}
        log("cpu baseline, %d threads: %.3e site-updates/s (%.1fs)" % (len(parts), cpu["value"], wall))

    traffic = None
    if (prec == E.PREC_FP32 and args.schedule == "levels" and (N, P) == (128, 100000)):
        traffic = measured_traffic("k_cla_nuc<float,4>|gtr 128x100000 levels")
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
                   "schedule": args.schedule, "cla_updates_per_step": n_cla, "edges_per_step": n_edges,
                   "fp64_heavy_patterns": n_heavy},
        "lnl_evals_per_s": evals_per_s,
        "lnl_eval_p50_ms": float(np.percentile(lat, 50)),
        "lnl_eval_p99_ms": float(np.percentile(lat, 99)),
        "accuracy": acc,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "algorithmic_bytes_per_launch": bytes_l / max(n_l, 1),
                     "kernel": "k_cla_nuc<%s,4>" % ("float" if prec == E.PREC_FP32 else "double"),
                     "avg_launch_ms": avg_ms, "launches": n_l,
                     "bytes_per_site_update": 3 * 4 * 4 * (4 if prec == E.PREC_FP32 else 8) + 12},
        "cpu_baseline": cpu,
    }
    if not args.no_codon:
        # the serving target's own configuration, timed in the same run: codon
        # GY94 M3, 32 taxa x 5,000 patterns, 64 GA offspring, FP32-bulk served
        # against the plain FP64 drop-in scoring each request on its own
        c = bench_states(args, workload="codon", cpu_leg=False)
        out["codon_served_evals_per_s"] = c["fp32_served_evals_per_s"]
        out["codon_fp64_plain_evals_per_s"] = c["fp64_plain_evals_per_s"]
        out["served_vs_fp64_plain"] = c["served_vs_fp64_plain"]
        out["codon_served_p50_ms"] = c["serve_stats"]["p50_ms"]
        out["codon_served_p99_ms"] = c["serve_stats"]["p99_ms"]
        out["codon_served_max_dlnl_vs_oracle"] = c["accuracy"]["serve_fp32_max_dlnl"]
        out["codon_roofline"] = c["roofline"]
        out["codon_config"] = c["config"]
    if not args.no_serve:
        # the GA request stream (Poisson arrivals over uneven nucleotide and
        # amino-acid instances), served FP32 bulk against the plain FP64 drop-in
        sv = bench_serve(args, cpu_leg=False)
        out["serve_requests_per_s"] = sv["value"]
        out["serve_fp64_plain_requests_per_s"] = sv["fp64_plain_requests_per_s"]
        out["serve_vs_fp64_plain"] = sv["served_vs_fp64_plain"]
        out["serve_p50_ms"] = sv["p50_ms"]
        out["serve_p99_ms"] = sv["p99_ms"]
        out["serve_max_dlnl_vs_fp64"] = sv["accuracy"]["served_fp32_vs_fp64_max_dlnl"]
        out["serve_config"] = sv["config"]
    print(json.dumps(out), flush=True)


def _cla_kinds(parent, n_tips):
    """(internal-internal, internal-tip, tip-tip) counts over all directional
    CLAs of the tree (tips are nodes 1..n_tips)."""
    nb = [[] for _ in range(len(parent))]
    for n, a in enumerate(parent):
        if a >= 0:
            nb[n].append(a)
            nb[a].append(n)
    kinds = [0, 0, 0]
    for a in range(len(parent)):
        if 1 <= a <= n_tips:
            continue
        for b in nb[a]:
            kinds[sum(1 for c in nb[a] if c != b and 1 <= c <= n_tips)] += 1
    return kinds


def _host_cores():
    """(cores this process may run on, cores the host reports)."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    return aff, os.cpu_count() or aff


def _cpu_threads():
    """Host threads for the CPU baseline: the cores this process may run on,
    at most 16 -- the host share that goes with one GPU on these machines
    (worker pools are sized to it whatever the host reports)."""
    return max(1, min(16, _host_cores()[0]))


def _pattern_slices(O, ds, T):
    """One FP64 oracle (reference release flags) per thread over a contiguous
    slice of the patterns: the reference's OpenMP split of its per-pattern
    loops (defs.h:29-43), with threads in place of an OpenMP team."""
    cuts = np.linspace(0, ds.n_patterns, T + 1).astype(int)
    return [O.Oracle(ds, fast=True, patterns=np.arange(cuts[k], cuts[k + 1])) for k in range(T)
            if cuts[k + 1] > cuts[k]]


def _threaded(fns):
    """Wall time of running every fn on its own thread (the oracle's ctypes
    calls release the GIL)."""
    import threading
    th = [threading.Thread(target=f) for f in fns]
    a = time.perf_counter()
    for t in th:
        t.start()
    for t in th:
        t.join()
    return time.perf_counter() - a


def _population(ds, n, seed):
    from garli_amd import sim
    rng = np.random.default_rng(seed)
    return [sim.perturb(ds.parent, ds.blen, rng, ds.n_tips, int(rng.integers(1, 4))) for _ in range(n)]


def _warm(inst):
    """One P(t) launch on a fresh instance's stream before timing (its first
    use costs a one-off set-up the steady state does not pay)."""
    inst.update_transition_matrices([0], [0.1])


def _eval_rates(ds, E, pop, prec):
    """Per-request plain drop-in: a fresh tree per request, Tree::Score."""
    inst = ds.make_instance(prec)
    _warm(inst)
    lat = []
    for p, b in pop:
        t = E.Tree(inst, p, b)
        a = time.perf_counter()
        t.score()
        lat.append(time.perf_counter() - a)
        t.close()
    inst.close()
    lat = np.array(lat) * 1000.0
    return 1000.0 / lat.mean(), lat


SERVE_REPEATS = 5  # cold-cache bursts timed; the median is reported


def _serve_rate(ds, E, pop, prec, slots):
    """The generation's offspring submitted in one call to a server over a
    fresh instance (cold cache); the median of SERVE_REPEATS such bursts."""
    runs = []
    for _ in range(SERVE_REPEATS):
        inst = ds.make_instance(prec, n_cla_slots=slots, n_pmat_slots=4 * slots)
        _warm(inst)
        srv = E.Server()
        a = time.perf_counter()
        ids = srv.submit_batch(inst, pop)
        res = list(srv.wait_all(ids))
        dt = time.perf_counter() - a
        st = srv.stats()
        st.update(srv.timing())
        st.update(srv.work())
        srv.close()
        inst.close()
        runs.append((len(pop) / dt, st, res))
    runs.sort(key=lambda x: x[0])
    rate, st, res = runs[len(runs) // 2]
    st["bursts_evals_per_s"] = [r[0] for r in runs]
    return rate, st, res


def _states_dataset(workload):
    from garli_amd import sim
    if workload == "aa":
        return sim.aa_gamma(n_tips=64, n_patterns=20000, seed=3), \
            "amino-acid 20-state G4 (random exchangeabilities), 64 taxa, 20000 patterns", 20
    return sim.codon_m3(n_tips=32, n_patterns=5000, seed=4), \
        "codon GY94 M3 (omega 0.1/1/3, kappa 2.5), 32 taxa, 5000 patterns", 61


def bench_leg(args):
    """One request leg of the aa / codon workload alone, for kernel traces:
    the 64 offspring served (FP32 bulk, SERVE_REPEATS cold-cache bursts) or
    scored one at a time by the plain FP64 drop-in."""
    from garli_amd import engine as E
    ds, name, _ = _states_dataset(args.workload)
    E.init(0)
    pop = _population(ds, 64, 11)
    if args.leg == "serve":
        rate, st, _ = _serve_rate(ds, E, pop, E.PREC_FP32, 3 * (ds.n_tips - 2) * 24)
        log("served %.1f evals/s" % rate)
        return {"workload": name, "leg": "serve", "evals_per_s": rate, "serve_stats": st}
    rate, lat = _eval_rates(ds, E, pop, E.PREC_FP64)
    log("plain fp64 %.1f evals/s" % rate)
    return {"workload": name, "leg": "plain", "evals_per_s": rate, "mean_ms": float(lat.mean())}


def bench_states(args, workload=None, cpu_leg=True):
    """AA (20 states, G4, 64 taxa, 20k patterns) or codon (GY94 M3, 32 taxa, 5k
    patterns): sweeps in both forms, lnL evaluations per request (FP64 plain
    drop-in) and batched through the serving layer (FP32 bulk).  Returns the
    record (printed by main for --workload aa|codon)."""
    from garli_amd import engine as E
    from garli_amd import sim
    from oracle import oracle as O
    workload = workload or args.workload
    ds, name, S = _states_dataset(workload)
    E.init(0)
    N, P = ds.n_tips, ds.n_patterns
    n_cla = 3 * (N - 2)
    res = {}
    sweeps = {}
    for prec, tag in ((E.PREC_FP64, "fp64"), (E.PREC_FP32, "fp32")):
        inst = ds.make_instance(prec)
        t = E.Tree(inst, ds.parent, ds.blen)
        t.set_schedule(0)
        for _ in range(args.warmup):
            t.sweep()
        inst.set_timing(True)
        inst.kernel_stats(0, reset=True)
        a = time.perf_counter()
        for _ in range(args.steps):
            sw = t.sweep()
        ms = 1000.0 * (time.perf_counter() - a) / args.steps
        inst.set_timing(False)
        n_l, ms_l, bytes_l = inst.kernel_stats(0)
        sweeps[tag] = sw
        res[tag] = {"ms_per_sweep": ms, "site_updates_per_s": n_cla * P / (ms / 1000.0),
                    "cla_kernel_GBs": bytes_l / max(ms_l, 1e-12) / 1e6, "cla_kernel_ms_per_sweep": ms_l / args.steps,
                    "cla_launches": n_l, "cla_algorithmic_bytes_per_launch": bytes_l / max(n_l, 1),
                    "cla_avg_launch_ms": ms_l / max(n_l, 1)}
        t.close()
        inst.close()
        log("%s sweep %.3f ms, CLA kernels %.1f GB/s" % (tag, ms, res[tag]["cla_kernel_GBs"]))
    if S == 61:
        # useful codon CLA flops per sweep: R(4S^2+S) per pattern for an
        # internal-internal update, R(2S^2+S) with one tip child, R S with two
        n_ii, n_it, n_tt = _cla_kinds(ds.parent, N)
        flops = ds.n_rates * P * (n_ii * (4 * S * S + S) + n_it * (2 * S * S + S) + n_tt * S)
        for tag in ("fp64", "fp32"):
            res[tag]["codon_mfma_TFLOPs"] = flops / (res[tag]["cla_kernel_ms_per_sweep"] / 1000.0) / 1e12
        ceiling = E.mfma_calibrate(0)
        ceiling44 = E.mfma_calibrate(1)
        log("codon CLA MFMA: %.1f TFLOP/s (fp32 form) of %.1f published, %.1f measured register-only ceiling "
            "of the 16x16x4 instruction the kernel uses (%.1f for 4x4x4_4b)"
            % (res["fp32"]["codon_mfma_TFLOPs"], PEAK_FP64_TFLOPS, ceiling, ceiling44))
        roof = {"bound": "mfma", "achieved": res["fp32"]["codon_mfma_TFLOPs"], "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": res["fp32"]["codon_mfma_TFLOPs"] / PEAK_FP64_TFLOPS,
                "traffic": measured_traffic("k_cla_codon<float,3>|codon 32x5000 levels"),
                "algorithmic_bytes_per_launch": res["fp32"]["cla_algorithmic_bytes_per_launch"],
                "avg_launch_ms": res["fp32"]["cla_avg_launch_ms"], "launches": res["fp32"]["cla_launches"],
                "measured_ceiling": ceiling, "frac_of_measured_ceiling": res["fp32"]["codon_mfma_TFLOPs"] / ceiling,
                "measured_ceiling_4x4x4_4b": ceiling44,
                "kernel": "k_cla_codon<float,3>", "useful_flops_per_sweep": flops}
    else:
        gbs = res["fp32"]["cla_kernel_GBs"]
        roof = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                "traffic": measured_traffic("k_cla_aa<float,4>|aa 64x20000 levels"),
                "algorithmic_bytes_per_launch": res["fp32"]["cla_algorithmic_bytes_per_launch"],
                "avg_launch_ms": res["fp32"]["cla_avg_launch_ms"], "launches": res["fp32"]["cla_launches"],
                "kernel": "k_cla_aa<float,4>"}
    if args.quick and workload == args.workload:
        copy = None
        if args.calibrate > 0:
            nbytes = int(args.calibrate * (1 << 30))
            cms = E.copy_calibrate(nbytes, 10)
            copy = {"bytes": nbytes, "ms": cms, "GB/s": 2 * nbytes / cms / 1e6}
        return {"workload": name, "sweeps": res, "roofline": roof, "copy": copy}
    pop = _population(ds, 64, 11)
    e64, lat64 = _eval_rates(ds, E, pop, E.PREC_FP64)
    e32, _ = _eval_rates(ds, E, pop, E.PREC_FP32)
    # HBM budget: CLA slots for the whole population's working sets in flight
    # (N - 2 per score request), 3 (N - 2) x 24 slots
    slots = 3 * (N - 2) * 24
    s32, st, served = _serve_rate(ds, E, pop, E.PREC_FP32, slots)
    o = O.Oracle(ds)
    ref = [o.score(parent=p, blen=b) for p, b in pop[:8]]
    cpu = None
    if cpu_leg:
        # CPU baseline: the plain FP64 path on the host, each request's full
        # score with its patterns split over the host threads
        T = _cpu_threads()
        parts = _pattern_slices(O, ds, T)
        n_cpu = 4
        wall = _threaded([lambda q=q: [q.score(parent=p, blen=b) for p, b in pop[:n_cpu]] for q in parts])
        cpu = {"value": n_cpu / wall, "unit": "lnL evals/s", "cores": len(parts), "kind": "port",
               "sample": "%d of the same requests scored in full (FP64, oracle built -O3 -ffast-math), patterns "
                         "split over %d threads as the reference's OpenMP build splits them, %.1fs wall" %
                         (n_cpu, len(parts), wall)}
        log("cpu baseline, %d threads: %.2f lnL evals/s" % (len(parts), cpu["value"]))
    acc = {"serve_fp32_max_dlnl": float(max(abs(x - y) for x, y in zip(served[:8], ref))),
           "sweep_fp32_vs_fp64_max_dlnl": float(np.max(np.abs(sweeps["fp32"][:, 0] - sweeps["fp64"][:, 0]))),
           "sweep_fp64_vs_oracle_max_rel": float(np.max(np.abs(sweeps["fp64"][:, 0] - o.sweep()[:, 0]) /
                                                        np.abs(sweeps["fp64"][:, 0])))}
    log("evals/s: fp64 plain %.1f, fp32 plain %.1f, fp32 served %.1f (%.1fx)" % (e64, e32, s32, s32 / e64))
    out = {"metric": METRIC, "value": s32, "unit": "lnL evals/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": res["fp32"]["ms_per_sweep"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded simulation)",
           "config": {"workload": name, "taxa": N, "patterns": P, "states": S, "rate_cats": ds.n_rates,
                      "requests": len(pop)},
           "fp64_plain_evals_per_s": e64, "fp32_plain_evals_per_s": e32, "fp32_served_evals_per_s": s32,
           "served_vs_fp64_plain": s32 / e64, "serve_stats": st, "sweeps": res, "accuracy": acc, "roofline": roof,
           "cpu_baseline": cpu,
           "serve_config": {"cla_slots": slots, "hbm_budget_GB": slots * ds.n_rates * ((P + 255) // 256 * 256) * (64 if S == 61 else S) * 4 / 1e9,
                            "submit": "one gle_serve_submit_batch call per burst", "bursts": SERVE_REPEATS}}
    return out


def bench_serve(args, cpu_leg=True):
    """The serving workload: a stream of 256 GA requests with Poisson arrivals
    over uneven instances (nucleotide and amino-acid, 16-512 taxa, 1k-200k
    patterns), each a parent tree plus 1-3 NNI/SPR moves; one in eight asks for
    a branch-length pass before its score (Individual::Mutate's
    optimisation, individual.cpp:127-222).  FP32 bulk served against the FP64
    plain drop-in running the same requests one at a time.  Returns the record
    (printed by main for --workload serve)."""
    import threading
    from garli_amd import engine as E
    from garli_amd import sim
    rng = np.random.default_rng(99)
    E.init(0)
    specs = [("gtr", 16, 1000), ("aa", 24, 2000), ("gtr", 64, 20000), ("aa", 48, 5000),
             ("gtr", 128, 50000), ("gtr", 512, 4000), ("aa", 96, 10000), ("gtr", 32, 200000)]
    dss = []
    t0 = time.time()
    for k, (kind, n, p) in enumerate(specs):
        if kind == "gtr":
            dss.append(sim.gtr_gamma(n_tips=n, n_sites=None, n_patterns=p, mean_blen=0.05, pinv=0.1,
                                     ambig_frac=0.0, seed=200 + k))
        else:
            dss.append(sim.aa_gamma(n_tips=n, n_patterns=p, seed=200 + k))
    log("simulated %d instances in %.1fs" % (len(dss), time.time() - t0))
    n_req = 256
    which = rng.integers(len(dss), size=n_req)
    reqs = []
    for k in range(n_req):
        ds = dss[which[k]]
        p, b = sim.perturb(ds.parent, ds.blen, rng, ds.n_tips, int(rng.integers(1, 4)))
        # branch-length passes on the instances of up to 128 taxa
        kind = E.REQ_OPTIMIZE if (k % 8 == 3 and ds.n_tips <= 128) else E.REQ_SCORE
        reqs.append((int(which[k]), p, b, kind))
    n_opt = sum(1 for r in reqs if r[3] == E.REQ_OPTIMIZE)
    # FP64 plain per-request drop-in
    insts64 = [ds.make_instance(E.PREC_FP64) for ds in dss]
    a = time.perf_counter()
    plain = []
    for w, p, b, kind in reqs:
        t = E.Tree(insts64[w], p, b)
        plain.append(t.score() if kind == E.REQ_SCORE else t.optimize_all_branches(0.01)[1])
        t.close()
    t64 = time.perf_counter() - a
    log("plain fp64: %d requests (%d branch passes) in %.2fs" % (n_req, n_opt, t64))
    # FP32 bulk served, Poisson arrivals at a rate above the plain path's capacity
    budget = 32 << 30
    insts = []
    for ds in dss:
        slot_bytes = ds.n_patterns * ds.n_rates * ds.n_states * 4
        slots = max(2 * (ds.n_tips - 2) + 4, min(8 * (ds.n_tips - 2), budget // len(dss) // slot_bytes))
        insts.append(ds.make_instance(E.PREC_FP32, n_cla_slots=int(slots), n_pmat_slots=int(8 * ds.n_tips)))
    srv = E.Server()
    rate = 4.0 * n_req / t64
    ids = [None] * n_req
    gaps = rng.exponential(1.0 / rate, size=n_req)

    def producer():
        for k, (w, p, b, kind) in enumerate(reqs):
            ids[k] = srv.submit(insts[w], p, b, kind=kind, precision=0.01)
            time.sleep(float(gaps[k]))

    a = time.perf_counter()
    th = threading.Thread(target=producer)
    th.start()
    th.join()
    served = [srv.wait(i, n_nodes=len(reqs[k][1])) for k, i in enumerate(ids)]
    t32 = time.perf_counter() - a
    st = srv.stats()
    tm = srv.timing()
    srv.close()
    # every served FP32 result against the FP64 drop-in on the same request:
    # scores directly, branch passes rescored at the served lengths in FP64
    dl, dopt = [], []
    for (w, p, b, kind), (lnl, _, bl), ref in zip(reqs, served, plain):
        if kind == E.REQ_SCORE:
            dl.append(abs(lnl - ref))
        else:
            t = E.Tree(insts64[w], p, bl)
            dl.append(abs(lnl - t.score()))
            dopt.append(lnl - ref)
            t.close()
    for i in insts64:
        i.close()
    acc = {"served_fp32_vs_fp64_max_dlnl": float(max(dl)), "served_fp32_vs_fp64_mean_dlnl": float(np.mean(dl)),
           "branch_pass_served_minus_fp64_pass": [float(min(dopt)), float(max(dopt))] if dopt else None}
    log("served fp32 vs fp64 plain: max |dlnL| %.2e over %d requests" % (max(dl), n_req))
    cpu = None
    if cpu_leg and not args.no_cpu:
        # CPU baseline: one score request of each instance in full by the FP64
        # oracle, patterns split over the host threads; the stream's score rate
        # is the inverse of the mean time over its mix of instances
        from oracle import oracle as O
        T = _cpu_threads()
        t_of = {}
        for w, p, b, kind in reqs:
            if w in t_of or kind != E.REQ_SCORE:
                continue
            parts = _pattern_slices(O, dss[w], T)
            t_of[w] = _threaded([lambda q=q, p=p, b=b: q.score(parent=p, blen=b) for q in parts])
        n_sc = sum(1 for r in reqs if r[3] == E.REQ_SCORE and r[0] in t_of)
        mean_t = sum(t_of[w] for w, _, _, kind in reqs if kind == E.REQ_SCORE and w in t_of) / max(n_sc, 1)
        cpu = {"value": 1.0 / mean_t, "unit": "lnL evals/s (score requests)", "cores": T, "kind": "port",
               "sample": "one score request of each of the %d instances in full (FP64, oracle built -O3 "
                         "-ffast-math), patterns split over %d threads as the reference's OpenMP build splits them, "
                         "%.2fs wall in all; weighted by the stream's mix of score requests (branch passes "
                         "not timed on the CPU)" % (len(t_of), T, sum(t_of.values()))}
        log("cpu baseline, %d threads: %.2f lnL evals/s" % (T, cpu["value"]))
    out = {"metric": METRIC, "value": n_req / t32, "unit": "requests/s", "n_gpus": 1, "steps": n_req,
           "warmup": 0, "ms_per_step": 1000.0 * t32 / n_req, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded simulations)",
           "config": {"workload": "GA request stream, Poisson arrivals, %d requests (%d with a branch-length pass) "
                      "over %d instances" % (n_req, n_opt, len(dss)),
                      "instances": [list(x) for x in specs], "arrival_rate_per_s": rate},
           "fp64_plain_requests_per_s": n_req / t64, "served_vs_fp64_plain": t64 / t32,
           "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_batch": st["mean_batch"],
           "serve_timing_ms": tm, "accuracy": acc, "cpu_baseline": cpu}
    log("serve: %.1f requests/s (plain fp64 %.1f), p50 %.2f ms p99 %.2f ms, mean batch %.1f" %
        (n_req / t32, n_req / t64, st["p50_ms"], st["p99_ms"], st["mean_batch"]))
    return out


if __name__ == "__main__":
    main()
