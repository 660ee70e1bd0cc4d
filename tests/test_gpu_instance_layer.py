"""The instance layer (the BEAGLE-seam-shaped op-list ABI) driven directly.

This is what an adapter in place of GARLI's CalculationManager does
(subsetcalculationmanager.cpp:232-341 transition matrices, :595-682 CLA
batches by level, :765-901 scoring): tips from MakeAmbigStrings streams,
P/P'/P'' per branch, every directional CLA in one multi-level
gle_update_partials, then (lnL, d1, d2) on every edge.  Every result is
checked against the FP64 oracle: CLAs to 1e-12 relative with rescale counts
bit-exact, lnL and derivatives to 1e-9."""
import numpy as np
import pytest

from garli_amd import engine as E
from garli_amd import sim
from oracle import oracle as O
import oplist

pytestmark = pytest.mark.gpu


def _instance(ds, prec, streams=True, n_cla_slots=None):
    nn = 2 * ds.n_tips - 2
    inst = E.Instance(ds.n_tips, ds.n_patterns, ds.n_states, ds.n_rates, n_eigen=len(ds.eig),
                      n_cla_slots=n_cla_slots or 3 * (ds.n_tips - 2), n_pmat_slots=nn, precision=prec)
    if streams:
        for t, st in enumerate(sim.ambig_streams(ds)):
            inst.set_tip_ambig_stream(t, st)
    else:
        for t in range(ds.n_tips):
            inst.set_tip_states(t, ds.tips[t])
    inst.set_patterns(ds.counts, ds.last_const, ds.const_states)
    inst.set_model(ds.pi, ds.rate_probs, ds.pinv, ds.pinv_in_model)
    for e, (V, W, lam) in enumerate(ds.eig):
        inst.set_eigen(e, V, W, lam)
    inst.set_category_rates(ds.eig_of_rate, ds.rate_mult)
    return inst


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_ambig_streams_sweep_matches_oracle(gpu):
    """MakeAmbigStrings-format tips (5% IUPAC codes), one multi-level
    gle_update_partials of all 3(N-2) directional CLAs, every CLA and count
    against the oracle, then every edge's lnL, d1, d2 and site lnL."""
    ds = sim.gtr_gamma(n_tips=18, n_sites=1200, pinv=0.15, ambig_frac=0.05, mean_blen=0.4, seed=301)
    assert any(np.any(st < 0) for st in sim.ambig_streams(ds)), "no ambiguity codes in the streams"
    print("case gtr 18 taxa", ds.n_patterns, "patterns", flush=True)
    inst = _instance(ds, E.PREC_FP64)
    ops, ls, slot_of, pmat_of = oplist.full_sweep(ds.parent, ds.n_tips, rescale_every=4)
    assert len(ls) > 3 and sum(o[5] for o in ops) > 0
    nodes = sorted(pmat_of)
    inst.update_transition_matrices([pmat_of[n] for n in nodes], ds.blen[nodes], derivs=True)
    inst.rescale_flags(reset=True)
    assert inst.update_partials(ops, ls) == E.SUCCESS
    o = O.Oracle(ds, own_model=False)
    rescaled = 0
    for (a, b), s in slot_of.items():
        g, gu = inst.get_partials(s)
        r, ru = o.cla(a, b, rescale_every=4)
        np.testing.assert_array_equal(gu, ru, err_msg="counts of X(%d|%d)" % (a, b))
        np.testing.assert_allclose(g, r, rtol=1e-12, atol=0, err_msg="X(%d|%d)" % (a, b))
        rescaled += int((ru > 0).sum())
    assert rescaled > 0
    eops = oplist.edge_ops(ds.parent, ds.n_tips, slot_of, pmat_of)
    res, site = inst.edge_likelihoods(eops, want_sites=True)
    for (n, (l, d1, d2)) in zip(range(1, len(ds.parent)), res):
        rl, r1, r2 = o.derivs(n, rescale_every=4)
        assert _rel(l, rl) < 1e-9, (n, l, rl)
        assert abs(d1 - r1) <= 1e-9 * max(1.0, abs(r1)), (n, d1, r1)
        assert abs(d2 - r2) <= 1e-9 * max(1.0, abs(r2)), (n, d2, r2)
    ref, site_ref, _ = o.score(rescale_every=4, sites=True)
    np.testing.assert_allclose(site, site_ref, rtol=1e-10, atol=1e-10)
    assert _rel(float(np.dot(ds.counts, site)), ref) < 1e-9
    # one edge with its per-pattern values (OutputSiteDerivatives inputs)
    (l3, d13, d23), sv = inst.edge_site_values(eops[4][:3])
    assert _rel(l3, res[4][0]) < 1e-12 and abs(d13 - res[4][1]) <= 1e-9 * max(1.0, abs(d13))
    assert np.all(sv["under_partial"] >= 0)


def test_transition_matrices_match_oracle(gpu):
    ds = sim.gtr_gamma(n_tips=6, n_sites=100, pinv=0.2, seed=302)
    inst = _instance(ds, E.PREC_FP64)
    o = O.Oracle(ds, own_model=False)
    for k, t in enumerate([1e-8, 0.01, 0.3, 2.5, 40.0]):
        inst.update_transition_matrices([k], [t], derivs=True)
        got = inst.get_transition_matrices(k)
        P, D1, D2 = o.pmats(t)
        np.testing.assert_allclose(got[0], P, rtol=1e-11, atol=1e-15)
        np.testing.assert_allclose(got[1], D1, rtol=1e-9, atol=1e-13)
        np.testing.assert_allclose(got[2], D2, rtol=1e-9, atol=1e-12)


def test_rescale_retry_from_the_op_list_layer(gpu):
    """Discordant tips on near-zero branches with rescaling only every 80
    ranks: a rescaled site maximum falls below 1e-190 and gle_update_partials
    returns GLE_RESCALE_RETRY (the reference's throw(1)); lowering the
    frequency by 2 per retry, as Tree::Score does (tree.cpp:4077-4090),
    reaches the oracle's frequency and its lnL."""
    ds = sim.gtr_gamma(n_tips=64, n_sites=40, mean_blen=0.1, seed=81, ambig_frac=0.0)
    rng = np.random.default_rng(0)
    ds.tips[:] = (1 << rng.integers(0, 4, size=ds.tips.shape)).astype(np.uint8)
    blen = ds.blen.copy()
    blen[1:] = 1e-8
    inst = _instance(ds, E.PREC_FP64, streams=False)
    nodes = list(range(1, len(ds.parent)))
    mid = [n for n in range(1, len(ds.parent)) if ds.parent[n] == 0][1]  # the root's middle child
    every, codes = 80, []
    while True:
        # the CLAs Tree::Score needs on the edge (root, middle child)
        ops, ls, slot_of, pmat_of = oplist.full_sweep(ds.parent, ds.n_tips, rescale_every=every,
                                                      targets=[(0, mid), (mid, 0)])
        inst.update_transition_matrices([pmat_of[n] for n in nodes], blen[nodes], derivs=False)
        inst.rescale_flags(reset=True)
        rc = inst.update_partials(ops, ls)
        codes.append(rc)
        if rc == E.SUCCESS:
            break
        assert rc == E.RESCALE_RETRY
        every -= 2
        assert every >= 2
    assert codes[0] == E.RESCALE_RETRY and codes[-1] == E.SUCCESS
    ref, _, re = O.Oracle(ds).score(blen=blen, rescale_every=80, sites=True)
    assert re == every
    child = -mid if mid <= ds.n_tips else slot_of[(mid, 0)]
    res = inst.edge_likelihoods([(slot_of[(0, mid)], child, pmat_of[mid], 0)])
    assert _rel(res[0][0], ref) < 1e-9


def test_fp32_instance_layer_within_tolerance(gpu):
    """The FP32 form through the same op lists: |dlnL| < 1e-3 on every edge
    against the FP64 form (and its counts are powers of two, so not compared)."""
    ds = sim.gtr_gamma(n_tips=20, n_sites=3000, pinv=0.1, ambig_frac=0.05, seed=303)
    out = {}
    for prec in (E.PREC_FP64, E.PREC_FP32):
        inst = _instance(ds, prec)
        ops, ls, slot_of, pmat_of = oplist.full_sweep(ds.parent, ds.n_tips,
                                                      rescale_every=16 if prec == E.PREC_FP64 else 1)
        nodes = sorted(pmat_of)
        inst.update_transition_matrices([pmat_of[n] for n in nodes], ds.blen[nodes], derivs=True)
        inst.update_partials(ops, ls)
        out[prec] = inst.edge_likelihoods(oplist.edge_ops(ds.parent, ds.n_tips, slot_of, pmat_of))
        inst.close()
    assert np.max(np.abs(out[E.PREC_FP32][:, 0] - out[E.PREC_FP64][:, 0])) < 1e-3


def test_level_list_checks(gpu):
    """Level ranges that skip, reorder or overrun ops, and a level whose op
    reads a CLA written in the same level, are refused before any launch."""
    ds = sim.gtr_gamma(n_tips=8, n_sites=200, seed=304)
    inst = _instance(ds, E.PREC_FP64)
    ops, ls, slot_of, pmat_of = oplist.full_sweep(ds.parent, ds.n_tips)
    n = len(ops)
    for bad in ([0, n + 3, n], [0, 5, 3, n], [0, -2, n], [1, n]):
        with pytest.raises(E.GleError) as ei:
            inst.update_partials(ops, bad)
        assert ei.value.code == -1, bad
    with pytest.raises(E.GleError):
        inst.update_partials(ops, [0, n])  # one level: ops read CLAs written beside them
    nodes = sorted(pmat_of)
    inst.update_transition_matrices([pmat_of[n] for n in nodes], ds.blen[nodes], derivs=False)
    assert inst.update_partials(ops, ls) == E.SUCCESS
    # no levels given: the engine derives them from the topological order
    assert inst.update_partials(ops) == E.SUCCESS
