"""Serving layer: continuous batching with a device CLA cache.

Each result must be the same alone and in the mix, and equal to the tree
layer's / oracle's answer for that tree.  A score is taken on the edge that
needs the fewest new CLAs given the cache, so alone and mixed results agree to
the form's rounding: 1e-12 relative in FP64, 1e-3 absolute in FP32."""
import threading

import numpy as np
import pytest

from garli_amd import engine as E
from garli_amd import sim
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _population(ds, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        out.append(sim.perturb(ds.parent, ds.blen, rng, ds.n_tips, int(rng.integers(1, 4))))
    return out


@pytest.mark.parametrize("prec", [E.PREC_FP64, E.PREC_FP32])
def test_alone_equals_in_the_mix(gpu, prec):
    ds = sim.gtr_gamma(n_tips=20, n_sites=1500, pinv=0.1, seed=101)
    inst = ds.make_instance(prec, n_cla_slots=600, n_pmat_slots=600)
    srv = E.Server()
    pop = _population(ds, 24, 5)
    # alone
    alone = []
    for p, b in pop[:6]:
        rid = srv.submit(inst, p, b)
        alone.append(srv.wait(rid)[0])
    # in the mix: everything submitted at once, from several threads
    ids = [None] * len(pop)

    def sub(k):
        ids[k] = srv.submit(inst, *pop[k])

    th = [threading.Thread(target=sub, args=(k,)) for k in range(len(pop))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    mix = [srv.wait(i)[0] for i in ids]
    for a, m in zip(alone, mix[:6]):
        assert abs(a - m) <= (1e-12 * abs(a) if prec == E.PREC_FP64 else 1e-3)
    st = srv.stats()
    # a request that arrives while the device is free starts a tick at once;
    # the ones that arrive meanwhile share the next (so a tick holds one or more)
    assert st["completed"] == len(pop) + 6 and st["mean_batch"] >= 1.0
    o = O.Oracle(ds)
    tol = 1e-9 if prec == E.PREC_FP64 else 1e-3
    for (p, b), got in zip(pop[::4], mix[::4]):
        ref = o.score(parent=p, blen=b)
        assert abs(got - ref) < (tol * abs(ref) if prec == E.PREC_FP64 else tol)
    srv.close()


def test_optimize_request_matches_tree_layer(gpu):
    ds = sim.gtr_gamma(n_tips=14, n_sites=800, seed=102)
    inst = ds.make_instance(E.PREC_FP64, n_cla_slots=400, n_pmat_slots=800)
    srv = E.Server()
    rid = srv.submit(inst, ds.parent, ds.blen, kind=E.REQ_OPTIMIZE, precision=0.001)
    lnl, _, bl = srv.wait(rid, n_nodes=2 * ds.n_tips - 2)
    srv.close()
    inst2 = ds.make_instance(E.PREC_FP64)
    t = E.Tree(inst2, ds.parent, ds.blen)
    _, tl = t.optimize_all_branches(0.001)
    np.testing.assert_allclose(bl[1:], t.branch_lengths()[1:], rtol=1e-10)
    assert abs(lnl - tl) < 1e-9 * abs(tl)


def test_admission_under_small_budget(gpu):
    """A cache smaller than the population's working set still serves every
    request correctly by admitting a few at a time and evicting LRU entries."""
    ds = sim.gtr_gamma(n_tips=16, n_sites=600, seed=103)
    inst = ds.make_instance(E.PREC_FP32, n_cla_slots=2 * 14 + 4 + 10, n_pmat_slots=400)
    srv = E.Server()
    pop = _population(ds, 12, 7)
    ids = [srv.submit(inst, p, b) for p, b in pop]
    got = [srv.wait(i)[0] for i in ids]
    o = O.Oracle(ds)
    for (p, b), g in zip(pop, got):
        assert abs(g - o.score(parent=p, blen=b)) < 1e-3
    srv.close()


def test_wait_all_collects_a_burst(gpu):
    """A burst of 40 submitted in one call and collected in one call: every
    lnL equals the oracle's, the burst starts with a small first tick, and the
    collected ids are released."""
    ds = sim.gtr_gamma(n_tips=14, n_sites=700, seed=109)
    inst = ds.make_instance(E.PREC_FP64, n_cla_slots=40 * 12 + 4, n_pmat_slots=2000)
    srv = E.Server()
    pop = _population(ds, 40, 12)
    ids = srv.submit_batch(inst, pop)
    got = srv.wait_all(ids)
    st = srv.stats()
    o = O.Oracle(ds)
    for (p, b), g in zip(pop, got):
        ref = o.score(parent=p, blen=b)
        assert abs(g - ref) < 1e-9 * abs(ref)
    assert st["completed"] == len(pop) and st["mean_batch"] < len(pop), st
    with pytest.raises(E.GleError):
        srv.wait_all(ids[:1])
    srv.close()


def test_submit_rejects_a_detached_cycle(gpu):
    """A parent array whose degrees are all right but whose nodes 10..13 form
    a cycle detached from the root (each with one tip) is refused at
    submission; the server then serves a valid tree."""
    ds = sim.gtr_gamma(n_tips=8, n_sites=200, seed=111)
    parent = np.array([-1, 0, 0, 9, 9, 10, 11, 12, 13, 0, 11, 12, 13, 10], dtype=np.int32)
    inst = ds.make_instance(E.PREC_FP64, n_cla_slots=40, n_pmat_slots=100)
    srv = E.Server()
    with pytest.raises(E.GleError):
        srv.submit(inst, parent, ds.blen)
    with pytest.raises(E.GleError):
        srv.submit_batch(inst, [(ds.parent, ds.blen), (parent, ds.blen)])
    got = srv.wait(srv.submit(inst, ds.parent, ds.blen))[0]
    ref = O.Oracle(ds).score(parent=ds.parent, blen=ds.blen)
    assert abs(got - ref) < 1e-9 * abs(ref)
    srv.close()


def test_admission_at_minimum_budget(gpu):
    """A score touches at most the N - 2 CLAs that point at its edge, so an
    instance with exactly N - 2 slots serves scores one at a time."""
    ds = sim.gtr_gamma(n_tips=16, n_sites=500, seed=106)
    inst = ds.make_instance(E.PREC_FP64, n_cla_slots=ds.n_tips - 2, n_pmat_slots=400)
    srv = E.Server()
    pop = _population(ds, 6, 8)
    ids = [srv.submit(inst, p, b) for p, b in pop]
    got = [srv.wait(i)[0] for i in ids]
    st = srv.stats()
    srv.close()
    o = O.Oracle(ds)
    for (p, b), g in zip(pop, got):
        ref = o.score(parent=p, blen=b)
        assert abs(g - ref) < 1e-9 * abs(ref)
    assert st["mean_batch"] == 1.0


def test_batch_window_gathers_a_burst(gpu):
    """With a batching window an idle server lets a burst of submissions
    gather before its first tick; the results are those served without it."""
    ds = sim.gtr_gamma(n_tips=14, n_sites=700, seed=107)
    pop = _population(ds, 12, 10)
    out = []
    for window in (0, 200000):
        inst = ds.make_instance(E.PREC_FP64, n_cla_slots=12 * 12 + 4, n_pmat_slots=600)
        srv = E.Server()
        srv.set_batch_window(window)
        ids = [srv.submit(inst, p, b) for p, b in pop]
        out.append([srv.wait(i)[0] for i in ids])
        st = srv.stats()
        srv.close()
        inst.close()
        if window:
            # the gathered burst of 12 goes out in at most two ticks
            assert st["mean_batch"] >= len(pop) / 2, st
    np.testing.assert_allclose(out[0], out[1], rtol=1e-12)


@pytest.mark.parametrize("kind", ["aa", "codon"])
def test_serving_state_spaces(gpu, kind):
    """20- and 61-state instances through the server: alone equals in the mix,
    and both equal the oracle's answer for the request's tree."""
    if kind == "aa":
        ds = sim.aa_gamma(n_tips=10, n_patterns=300, seed=104)
    else:
        ds = sim.codon_m3(n_tips=8, n_patterns=120, seed=105)
    srv = E.Server()
    pop = _population(ds, 10, 9)
    for prec in (E.PREC_FP64, E.PREC_FP32):
        inst = ds.make_instance(prec, n_cla_slots=300, n_pmat_slots=400)
        alone = [srv.wait(srv.submit(inst, p, b))[0] for p, b in pop[:3]]
        ids = [srv.submit(inst, p, b) for p, b in pop]
        mix = [srv.wait(i)[0] for i in ids]
        for a, m in zip(alone, mix[:3]):
            assert abs(a - m) <= (1e-12 * abs(a) if prec == E.PREC_FP64 else 1e-3)
        o = O.Oracle(ds)
        for (p, b), got in zip(pop[::3], mix[::3]):
            ref = o.score(parent=p, blen=b)
            assert abs(got - ref) < (1e-9 * abs(ref) if prec == E.PREC_FP64 else 1e-3)
    srv.close()


def test_codon_score_edges_read_kept_messages(gpu):
    """A score edge whose child side (or, with the sides swapped, partial side)
    a CLA op of the same tick contracts at the edge's length reads that kept
    message instead of contracting the side again.  A cold server's single
    tick reads no message in its CLA ops (kept messages reach the cache only
    after the tick), so every message read it counts is an edge's.  Results
    equal the FP64 tree layer to 1e-12 relative and the oracle to 1e-9 in
    FP64, and stay within 1e-3 of FP64 in the FP32 form."""
    ds = sim.codon_m3(n_tips=12, n_patterns=240, seed=137)
    pop = _population(ds, 12, 19)
    o = O.Oracle(ds)
    inst64 = ds.make_instance(E.PREC_FP64)
    ref64 = []
    for p, b in pop:
        t = E.Tree(inst64, p, b)
        ref64.append(t.score())
        t.close()
    inst64.close()
    for prec in (E.PREC_FP64, E.PREC_FP32):
        inst = ds.make_instance(prec, n_cla_slots=400, n_pmat_slots=600)
        srv = E.Server()
        got = srv.wait_all(srv.submit_batch(inst, pop))
        w, st = srv.work(), srv.stats()
        srv.close()
        inst.close()
        assert st["mean_batch"] == len(pop), st  # one tick
        assert w["msg_reads"] > 0, w
        for k in range(len(pop)):
            tol = 1e-12 * abs(ref64[k]) if prec == E.PREC_FP64 else 1e-3
            assert abs(got[k] - ref64[k]) <= tol, (prec, k, got[k], ref64[k])
        for (p, b), g in zip(pop[::4], got[::4]):
            ref = o.score(parent=p, blen=b)
            assert abs(g - ref) < (1e-9 * abs(ref) if prec == E.PREC_FP64 else 1e-3)


def test_codon_messages_are_kept_and_read(gpu):
    """Codon serving keeps a contracted child's message P(t) . x when its
    subtree is shared, and later CLAs read it instead of contracting the child
    again (gle_serve_work counts both).  Results with messages equal the FP64
    tree layer to 1e-12 relative (FP64 messages are the exact products) and
    the oracle to 1e-9; the FP32 form stays within 1e-3 of FP64.  A nucleotide
    instance keeps no messages."""
    ds = sim.codon_m3(n_tips=12, n_patterns=240, seed=131)
    pop = _population(ds, 20, 17)
    o = O.Oracle(ds)
    ref64 = []
    inst64 = ds.make_instance(E.PREC_FP64)
    for p, b in pop:
        t = E.Tree(inst64, p, b)
        ref64.append(t.score())
        t.close()
    inst64.close()
    for prec in (E.PREC_FP64, E.PREC_FP32):
        inst = ds.make_instance(prec, n_cla_slots=400, n_pmat_slots=600)
        srv = E.Server()
        srv.wait(srv.submit(inst, ds.parent, ds.blen))  # the parent's CLAs, cached
        first = [srv.wait(i)[0] for i in srv.submit_batch(inst, pop[:10])]
        second = [srv.wait(i)[0] for i in srv.submit_batch(inst, pop[10:])]
        got = first + second
        w = srv.work()
        assert w["msg_keeps"] > 0 and w["msg_reads"] > 0, w
        assert w["contractions"] + w["msg_reads"] > w["msg_reads"]
        for k, (p, b) in enumerate(pop):
            if prec == E.PREC_FP64:
                assert abs(got[k] - ref64[k]) <= 1e-12 * abs(ref64[k]), (k, got[k], ref64[k])
            else:
                assert abs(got[k] - ref64[k]) < 1e-3, (k, got[k], ref64[k])
        for (p, b), g in zip(pop[::5], got[::5]):
            ref = o.score(parent=p, blen=b)
            assert abs(g - ref) < (1e-9 * abs(ref) if prec == E.PREC_FP64 else 1e-3)
        srv.close()
        inst.close()
    nuc = sim.gtr_gamma(n_tips=12, n_sites=400, seed=132)
    inst = nuc.make_instance(E.PREC_FP32, n_cla_slots=300, n_pmat_slots=400)
    srv = E.Server()
    npop = _population(nuc, 8, 18)
    [srv.wait(i) for i in srv.submit_batch(inst, npop)]
    w = srv.work()
    assert w["cla_ops"] > 0 and w["msg_keeps"] == 0 and w["msg_reads"] == 0, w
    srv.close()
    inst.close()


@pytest.mark.parametrize("slots", [10, 24, 60])
def test_codon_messages_under_slot_pressure(gpu, slots):
    """Codon serving with few CLA slots: kept messages are evicted before CLAs
    and their slots reused while later requests read other messages, over
    several rounds of the same population; every FP64 result still equals
    the tree layer's to 1e-12 relative (a stale message would not)."""
    ds = sim.codon_m3(n_tips=12, n_patterns=200, seed=133)
    pop = _population(ds, 16, 19)
    inst64 = ds.make_instance(E.PREC_FP64)
    ref = []
    for p, b in pop:
        t = E.Tree(inst64, p, b)
        ref.append(t.score())
        t.close()
    inst64.close()
    inst = ds.make_instance(E.PREC_FP64, n_cla_slots=slots, n_pmat_slots=400)
    srv = E.Server()
    srv.wait(srv.submit(inst, ds.parent, ds.blen))
    for rnd in range(3):
        got = [srv.wait(i)[0] for i in srv.submit_batch(inst, pop[rnd::2] if rnd < 2 else pop)]
        want = ref[rnd::2] if rnd < 2 else ref
        for g, w in zip(got, want):
            assert abs(g - w) <= 1e-12 * abs(w), (slots, rnd, g, w)
    w = srv.work()
    if slots >= 60:  # with a little room, messages survive to be read
        assert w["msg_reads"] > 0, w
    srv.close()
    inst.close()


def test_batch_submit_equals_single_submits(gpu):
    """gle_serve_submit_batch queues a generation's offspring in one call; the
    results are those of one-by-one submissions, and a batch holding an
    invalid tree queues nothing."""
    ds = sim.gtr_gamma(n_tips=22, n_sites=900, pinv=0.1, seed=108)
    pop = _population(ds, 16, 12)
    inst = ds.make_instance(E.PREC_FP64, n_cla_slots=16 * 22, n_pmat_slots=600)
    srv = E.Server()
    one = [srv.wait(srv.submit(inst, p, b))[0] for p, b in pop]
    ids = srv.submit_batch(inst, pop)
    many = [srv.wait(i)[0] for i in ids]
    np.testing.assert_allclose(many, one, rtol=1e-12)
    bad = pop[0][0].copy()
    bad[3] = 2
    with pytest.raises(E.GleError):
        srv.submit_batch(inst, [pop[1], (bad, pop[0][1])])
    st = srv.stats()
    assert st["completed"] == 2 * len(pop)
    o = O.Oracle(ds)
    for (p, b), got in zip(pop[::5], many[::5]):
        assert abs(got - o.score(parent=p, blen=b)) < 1e-9 * abs(got)
    srv.close()


@pytest.mark.parametrize("kind", ["gtr", "aa"])
def test_fp32_heavy_part_in_ticks(gpu, kind):
    """FP32 instances with an FP64 heavy part (patterns of count > 1 here, so
    it is a large share): a serving tick queues the heavy part's levels and
    edges after the FP32 part's on one stream, reading the same staged lists
    and P slots, and sums both parts' results.  Scores and a branch-length
    pass must match the FP64 drop-in on the same trees."""
    if kind == "gtr":
        ds = sim.gtr_gamma(n_tips=16, n_sites=3000, pinv=0.1, seed=131)
    else:
        ds = sim.aa_gamma(n_tips=12, n_patterns=600, seed=132)
    inst = ds.make_instance(E.PREC_FP32, n_cla_slots=400, n_pmat_slots=800)
    inst.set_heavy_split(1)
    assert 0 < inst.heavy_patterns() < ds.n_patterns
    inst64 = ds.make_instance(E.PREC_FP64)
    srv = E.Server()
    pop = _population(ds, 12, 7)
    ids = [srv.submit(inst, p, b) for p, b in pop]
    opt = srv.submit(inst, ds.parent, ds.blen, kind=E.REQ_OPTIMIZE, precision=0.01)
    got = [srv.wait(i)[0] for i in ids]
    lnl_opt, _, bl = srv.wait(opt, n_nodes=2 * ds.n_tips - 2)
    srv.close()
    for (p, b), g in zip(pop, got):
        t = E.Tree(inst64, p, b)
        ref = t.score()
        t.close()
        assert abs(g - ref) < 1e-3
    # the served pass's score at its own lengths, and its end against the
    # FP64 drop-in's own pass from the same start
    t = E.Tree(inst64, ds.parent, bl)
    at_served = t.score()
    t.close()
    assert abs(lnl_opt - at_served) < 1e-3
    t = E.Tree(inst64, ds.parent, ds.blen)
    _, pass64 = t.optimize_all_branches(0.01)
    t.close()
    assert abs(lnl_opt - pass64) < 1e-3
    inst.close()
    inst64.close()
