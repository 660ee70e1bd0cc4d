"""A GA request stream through the serving layer at the serving workload's
extremes: a 512-taxon and a 200,000-pattern nucleotide instance, amino-acid
instances, score and branch-optimisation requests, each a parent tree plus
1-3 SPR/NNI moves (Individual::Mutate, individual.cpp:127-222), arriving
with exponential gaps from a client thread while earlier ones run.

Every served FP32 result is checked against the FP64 drop-in (the tree
layer) on the same request: scores to |dlnL| < 1e-3; branch-optimisation
requests by rescoring the served branch lengths in FP64 (|dlnL| < 1e-3) and
against the drop-in's own Newton pass (both local optima, within 1e-2).  A
sample of scores is also checked against the CPU oracle."""
import threading
import time

import numpy as np
import pytest

from garli_amd import engine as E
from garli_amd import sim
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _instances():
    specs = [
        ("gtr", dict(n_tips=512, n_sites=None, n_patterns=2000, mean_blen=0.05, pinv=0.1, ambig_frac=0.0, seed=601)),
        ("gtr", dict(n_tips=32, n_sites=None, n_patterns=200000, mean_blen=0.05, pinv=0.1, ambig_frac=0.0, seed=602)),
        ("aa", dict(n_tips=48, n_patterns=3000, seed=603)),
        ("gtr", dict(n_tips=64, n_sites=None, n_patterns=10000, mean_blen=0.05, pinv=0.2, ambig_frac=0.02, seed=604)),
        ("aa", dict(n_tips=16, n_patterns=1000, seed=605)),
    ]
    return [sim.gtr_gamma(**kw) if k == "gtr" else sim.aa_gamma(**kw) for k, kw in specs]


def test_poisson_stream_at_the_extremes(gpu):
    dss = _instances()
    print("instances", [(d.n_tips, d.n_patterns, d.n_states) for d in dss], flush=True)
    rng = np.random.default_rng(77)
    reqs = []
    for k in range(40):
        w = k % len(dss)
        ds = dss[w]
        p, b = sim.perturb(ds.parent, ds.blen, rng, ds.n_tips, int(rng.integers(1, 4)))
        # branch-length passes on the amino-acid and 64-taxon instances
        kind = E.REQ_OPTIMIZE if (w in (2, 3, 4) and k % 3 == 0) else E.REQ_SCORE
        reqs.append((w, p, b, kind))
    assert {r[3] for r in reqs} == {E.REQ_SCORE, E.REQ_OPTIMIZE}
    insts = []
    for ds in dss:
        slots = max(4 * (ds.n_tips - 2) + 8, min(12 * (ds.n_tips - 2), (6 << 30) // (ds.n_patterns * ds.n_rates *
                                                                                      ds.n_states * 4)))
        insts.append(ds.make_instance(E.PREC_FP32, n_cla_slots=int(slots), n_pmat_slots=8 * ds.n_tips))
    srv = E.Server()
    ids = [None] * len(reqs)
    gaps = rng.exponential(0.002, size=len(reqs))

    def producer():
        for k, (w, p, b, kind) in enumerate(reqs):
            ids[k] = srv.submit(insts[w], p, b, kind=kind, precision=0.01)
            time.sleep(float(gaps[k]))

    th = threading.Thread(target=producer)
    th.start()
    th.join()
    served = [srv.wait(i, n_nodes=len(reqs[k][1])) for k, i in enumerate(ids)]
    st = srv.stats()
    srv.close()
    print("served", st, flush=True)
    # every request completes; each instance picks an arrival up as soon as
    # it has fewer than two parts in flight, so requests of different
    # instances overlap on the device even when a round of building holds a
    # single request (arrivals 2 ms apart over 8 instances seldom coincide):
    # the gathering of a burst is tested on its own
    assert st["completed"] == len(reqs) and st["mean_batch"] >= 1.0
    # the FP64 drop-in on every request
    insts64 = [ds.make_instance(E.PREC_FP64) for ds in dss]
    worst = 0.0
    for (w, p, b, kind), (lnl, _, bl) in zip(reqs, served):
        t = E.Tree(insts64[w], p, b)
        if kind == E.REQ_SCORE:
            ref = t.score()
            worst = max(worst, abs(lnl - ref))
            assert abs(lnl - ref) < 1e-3, (w, lnl, ref)
        else:
            t.set_topology(p, bl)  # the served branch lengths, scored in FP64
            at_served = t.score()
            assert abs(lnl - at_served) < 1e-3, (w, lnl, at_served)
            t2 = E.Tree(insts64[w], p, b)
            _, opt64 = t2.optimize_all_branches(0.01)
            assert abs(lnl - opt64) < 1e-2, (w, lnl, opt64)
            assert lnl > E.Tree(insts64[w], p, b).score()  # the pass improved the tree
            t2.close()
        t.close()
    print("max |dlnL| served fp32 vs fp64 drop-in: %.2e" % worst, flush=True)
    # a sample against the oracle
    for k in (0, 1, 2, 5, 6):
        w, p, b, kind = reqs[k]
        if kind != E.REQ_SCORE:
            continue
        ref = O.Oracle(dss[w]).score(parent=p, blen=b)
        assert abs(served[k][0] - ref) < 1e-3, (k, served[k][0], ref)


def test_tick_split_when_pmat_slots_run_out(gpu):
    """With only 2N - 3 transition-matrix slots one request fills them all, so
    a burst that gathered into one tick is split over several (the requests
    that do not fit wait, holding nothing); every result still equals the
    FP64 drop-in's and later requests succeed."""
    ds = sim.gtr_gamma(n_tips=18, n_sites=800, pinv=0.1, seed=606)
    rng = np.random.default_rng(8)
    pop = [sim.perturb(ds.parent, ds.blen, rng, ds.n_tips, int(rng.integers(1, 4))) for _ in range(8)]
    inst = ds.make_instance(E.PREC_FP64, n_cla_slots=40 * ds.n_tips, n_pmat_slots=2 * ds.n_tips - 3)
    srv = E.Server()
    srv.set_batch_window(100000)
    ids = [srv.submit(inst, p, b) for p, b in pop]
    got = [srv.wait(i)[0] for i in ids]
    st = srv.stats(reset=True)
    assert st["mean_batch"] <= len(pop) / 2, st  # spread over ticks, not all in the first
    ref_inst = ds.make_instance(E.PREC_FP64)
    for (p, b), g in zip(pop, got):
        ref = E.Tree(ref_inst, p, b).score()
        assert abs(g - ref) < 1e-9 * abs(ref)
    # the pool is whole again: another burst is served as well
    more = [sim.perturb(ds.parent, ds.blen, rng, ds.n_tips, 1) for _ in range(4)]
    got2 = [srv.wait(srv.submit(inst, p, b))[0] for p, b in more]
    for (p, b), g in zip(more, got2):
        ref = E.Tree(ref_inst, p, b).score()
        assert abs(g - ref) < 1e-9 * abs(ref)
    srv.close()


def test_request_too_large_for_the_instance_is_refused(gpu):
    ds = sim.gtr_gamma(n_tips=12, n_sites=300, seed=607)
    inst = ds.make_instance(E.PREC_FP64, n_cla_slots=40, n_pmat_slots=2 * ds.n_tips - 4)
    srv = E.Server()
    with pytest.raises(E.GleError) as ei:
        srv.submit(inst, ds.parent, ds.blen)
    assert ei.value.code == -3
    srv.close()
