"""Topology moves in the tree layer keep the CLAs the move did not touch.

GARLI dirties only the CLAs on the paths through the edges an SPR or NNI
changed (SweepDirtynessOverTree, tree.cpp:4247) and keeps the rest shared
(CopyClaIndeces, tree.cpp:9209).  gle_tree_set_topology does the same by
matching each directional CLA's exact subtree.  The tests count the CLA
updates a rescore runs after a move, against the count worked out here from
the two trees' subtrees, and check the scores against the oracle."""
import numpy as np
import pytest

from garli_amd import engine as E
from garli_amd import sim
from oracle import oracle as O
import oplist

pytestmark = pytest.mark.gpu


def _names(parent, blen, n_tips):
    """Exact name of every directional CLA X(a|b): the set of (child name,
    branch length) pairs below it, tips named by their numbers."""
    nb = oplist.neighbours(parent)
    memo = {}

    def name(a, b):
        if 1 <= a <= n_tips:
            return ("tip", a)
        if (a, b) not in memo:
            kids = []
            for c in nb[a]:
                if c != b:
                    kids.append((name(c, a), float(blen[a] if parent[a] == c else blen[c])))
            memo[(a, b)] = ("clade", frozenset(kids))
        return memo[(a, b)]

    for a in nb:
        if not (1 <= a <= n_tips):
            for b in nb[a]:
                name(a, b)
    return memo, nb


def _closure(nb, n_tips, targets, clean):
    need, stack = set(), [t for t in targets if not (1 <= t[0] <= n_tips)]
    while stack:
        a, b = stack.pop()
        if (a, b) in clean or (a, b) in need:
            continue
        need.add((a, b))
        stack.extend((c, a) for c in nb[a] if c != b and not (1 <= c <= n_tips))
    return need


def _score_edge(nb, clean):
    """The tree layer's choice of scoring edge (ConditionalLikelihoodRateHet's
    ROOT case, tree.cpp:3686-3768): left, then right, then middle child."""
    left, middle, right = nb[0][0], nb[0][1], nb[0][2]
    if (0, left) in clean:
        return left
    if (0, right) in clean:
        return right
    return middle


def _expected_updates(p0, b0, p1, b1, n_tips):
    """CLA updates a rescore of (p1, b1) needs after a score of (p0, b0) on a fresh tree."""
    names0, nb0 = _names(p0, b0, n_tips)
    mid0 = _score_edge(nb0, set())
    clean0 = _closure(nb0, n_tips, [(0, mid0), (mid0, 0)], set())
    have = {names0[k] for k in clean0}
    names1, nb1 = _names(p1, b1, n_tips)
    clean1 = {k for k, v in names1.items() if v in have}
    e = _score_edge(nb1, clean1)
    return len(_closure(nb1, n_tips, [(0, e), (e, 0)], clean1)), len(clean1)


@pytest.mark.parametrize("move", ["nni", "spr"])
@pytest.mark.parametrize("prec", [E.PREC_FP64, E.PREC_FP32])
def test_move_recomputes_only_the_changed_paths(gpu, move, prec):
    ds = sim.gtr_gamma(n_tips=24, n_sites=1500, pinv=0.1, seed=401)
    o = O.Oracle(ds)
    rng = np.random.default_rng(5 if move == "nni" else 6)
    inst = ds.make_instance(prec)
    t = E.Tree(inst, ds.parent, ds.blen)
    t.score()
    fresh = t.cla_updates()
    assert fresh == ds.n_tips - 2
    p, b = ds.parent, ds.blen
    for k in range(4):
        p1, b1 = (sim.nni if move == "nni" else sim.spr)(p, b, rng, ds.n_tips)
        want, kept = _expected_updates(p, b, p1, b1, ds.n_tips) if k == 0 else (None, None)
        before = t.cla_updates()
        t.set_topology(p1, b1)
        lnl = t.score()
        ran = t.cla_updates() - before
        if k == 0:
            print("move", move, "kept", kept, "recomputed", ran, "of", ds.n_tips - 2, flush=True)
            assert ran == want, (ran, want)
            assert ran < ds.n_tips - 2 and kept > 0
        ref = o.score(parent=p1, blen=b1)
        if prec == E.PREC_FP64:
            assert abs(lnl - ref) < 1e-9 * abs(ref)
        else:
            assert abs(lnl - ref) < 1e-3
        p, b = p1, b1
    # derivatives on every edge of the moved tree (reused and new CLAs mixed)
    for n in range(1, 2 * ds.n_tips - 2, 5):
        l, d1, d2 = t.derivatives(n)
        rl, r1, r2 = o.derivs(n, parent=p, blen=b)
        tol = 1e-9 if prec == E.PREC_FP64 else 1e-3
        assert abs(l - rl) < (tol * abs(rl) if prec == E.PREC_FP64 else tol)
        assert abs(d1 - r1) <= tol * max(1.0, abs(r1)) * (1 if prec == E.PREC_FP64 else 10)


def test_move_then_copy_shares_and_stays_separate(gpu):
    """An offspring made by copying the parent and moving it (Individual::Mutate
    after CopyClaIndeces) leaves the parent's CLAs and score untouched."""
    ds = sim.gtr_gamma(n_tips=20, n_sites=900, seed=402)
    inst = ds.make_instance(E.PREC_FP64)
    a = E.Tree(inst, ds.parent, ds.blen)
    la = a.score()
    c = a.copy()
    rng = np.random.default_rng(3)
    p1, b1 = sim.perturb(ds.parent, ds.blen, rng, ds.n_tips, 2)
    c.set_topology(p1, b1)
    lc = c.score()
    assert c.cla_updates() < ds.n_tips - 2
    assert a.score() == la
    o = O.Oracle(ds)
    assert abs(lc - o.score(parent=p1, blen=b1)) < 1e-9 * abs(lc)


def test_same_topology_new_length_is_a_branch_change(gpu):
    """A parent array with one changed length recomputes exactly what
    gle_tree_set_branch_length would."""
    ds = sim.gtr_gamma(n_tips=16, n_sites=600, seed=403)
    inst = ds.make_instance(E.PREC_FP64)
    t1, t2 = E.Tree(inst, ds.parent, ds.blen), E.Tree(inst, ds.parent, ds.blen)
    t1.score()
    t2.score()
    b = ds.blen.copy()
    b[7] *= 1.7
    t1.set_branch_length(7, b[7])
    t2.set_topology(ds.parent, b)
    c1, c2 = t1.cla_updates(), t2.cla_updates()
    assert t1.score() == t2.score()
    assert t1.cla_updates() - c1 == t2.cla_updates() - c2


def test_bad_topology_is_refused_and_tree_unchanged(gpu):
    ds = sim.gtr_gamma(n_tips=10, n_sites=300, seed=404)
    inst = ds.make_instance(E.PREC_FP64)
    t = E.Tree(inst, ds.parent, ds.blen)
    l0 = t.score()
    bad = ds.parent.copy()
    bad[3] = 2  # a tip as a parent
    with pytest.raises(E.GleError):
        t.set_topology(bad, ds.blen)
    assert t.score() == l0
