"""Edge cases on the GPU against the oracle: ragged and tiny pattern counts
(padding inside the last tile), the smallest tree, 1..8 rate categories,
zero and very long branches, a taxon with no data, and single-class codon
models.  FP64: lnL and every edge's (lnL, d1, d2) to 1e-9 relative; FP32
form: |dlnL| < 1e-3."""
import numpy as np
import pytest

from garli_amd import engine as E
from garli_amd import sim
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PRECS = [E.PREC_FP64, E.PREC_FP32]


def _check(ds, prec, sweep=True, blen=None):
    blen = ds.blen if blen is None else blen
    t = E.Tree(ds.make_instance(prec), ds.parent, blen)
    o = O.Oracle(ds)
    ref = o.score(blen=blen)
    got = t.score()
    if prec == E.PREC_FP64:
        assert abs(got - ref) <= 1e-9 * abs(ref), (got, ref)
    else:
        assert abs(got - ref) < 1e-3, (got, ref)
    if sweep:
        sw = t.sweep()
        osw = o.sweep(blen=blen)
        if prec == E.PREC_FP64:
            np.testing.assert_allclose(sw[:, 0], osw[:, 0], rtol=1e-9)
            np.testing.assert_allclose(sw[:, 1:], osw[:, 1:], rtol=1e-7, atol=1e-7 * np.abs(osw[:, 1:]).max())
        else:
            assert np.max(np.abs(sw[:, 0] - osw[:, 0])) < 1e-3
    t.close()


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("npat", [1, 2, 255, 257])
def test_ragged_nucleotide_patterns(gpu, prec, npat):
    ds = sim.gtr_gamma(n_tips=7, n_sites=None, n_patterns=npat, pinv=0.1, seed=300 + npat)
    print("nuc patterns", ds.n_patterns, flush=True)
    _check(ds, prec)


@pytest.mark.parametrize("prec", PRECS)
def test_three_taxa(gpu, prec):
    """The smallest unrooted tree: the root and three tips, every CLA a
    tip-tip op."""
    ds = sim.gtr_gamma(n_tips=3, n_sites=400, pinv=0.2, seed=310)
    _check(ds, prec)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("ncat", [1, 2, 5, 8])
def test_nucleotide_category_counts(gpu, prec, ncat):
    ds = sim.gtr_gamma(n_tips=9, n_sites=700, ncat=ncat, seed=320 + ncat)
    _check(ds, prec)


@pytest.mark.parametrize("prec", PRECS)
def test_branch_length_extremes(gpu, prec):
    """Zero-length branches (P = I) next to very long ones (P rows -> pi)."""
    ds = sim.gtr_gamma(n_tips=10, n_sites=600, seed=330)
    blen = ds.blen.copy()
    blen[1] = 0.0
    blen[4] = 0.0
    blen[6] = 50.0
    blen[ds.n_tips + 2] = 30.0
    _check(ds, prec, blen=blen)


@pytest.mark.parametrize("prec", PRECS)
def test_taxon_without_data(gpu, prec):
    ds = sim.gtr_gamma(n_tips=8, n_sites=500, seed=340)
    ds.tips[3, :] = 15
    _check(ds, prec)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("npat,ncat", [(1, 1), (65, 3), (130, 8)])
def test_amino_acid_shapes(gpu, prec, npat, ncat):
    ds = sim.aa_gamma(n_tips=5, n_patterns=npat, ncat=ncat, seed=350 + npat)
    print("aa patterns", ds.n_patterns, "categories", ncat, flush=True)
    _check(ds, prec)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("omegas,npat", [((0.4,), 17), ((0.05, 0.5, 1.0, 2.0, 4.0), 40)])
def test_codon_shapes(gpu, prec, omegas, npat):
    ds = sim.codon_m3(n_tips=5, n_patterns=npat, omegas=omegas, seed=360 + npat)
    print("codon patterns", ds.n_patterns, "classes", len(omegas), flush=True)
    _check(ds, prec)


@pytest.mark.parametrize("prec", PRECS)
def test_model_change_refreshes_caches(gpu, prec):
    """A new Gamma shape or new frequencies after a score make the tree's and
    the server's cached CLAs and P(t) stale (GARLI's MakeAllNodesDirty); the
    next score uses the new model without an explicit invalidate."""
    import dataclasses
    ds = sim.gtr_gamma(n_tips=12, n_sites=900, pinv=0.1, seed=370)
    inst = ds.make_instance(prec, n_cla_slots=400, n_pmat_slots=400)
    t = E.Tree(inst, ds.parent, ds.blen)
    srv = E.Server()
    before = t.score()
    srv.wait(srv.submit(inst, ds.parent, ds.blen))
    rates = E.discrete_gamma(ds.n_rates, 1.7)
    ds2 = dataclasses.replace(ds, rate_mult=rates / (1.0 - ds.pinv), meta=dict(ds.meta, alpha=1.7))
    inst.set_category_rates(ds2.eig_of_rate, ds2.rate_mult)
    pi = ds.pi[::-1].copy()
    V, W, lam = E.model_gtr(pi, ds.meta["rel"])
    ds2 = dataclasses.replace(ds2, pi=pi, eig=[(V, W, lam)])
    inst.set_model(pi, ds.rate_probs, ds.pinv, ds.pinv_in_model)
    inst.set_eigen(0, V, W, lam)
    ref = O.Oracle(ds2).score()
    after = t.score()
    served = srv.wait(srv.submit(inst, ds.parent, ds.blen))[0]
    srv.close()
    tol = 1e-9 * abs(ref) if prec == E.PREC_FP64 else 1e-3
    assert abs(after - before) > 1.0
    assert abs(after - ref) < tol and abs(served - ref) < tol


def _caterpillar(n_tips, rng, mean_blen):
    """The deepest unrooted tree: every internal node has a tip child."""
    nn = 2 * n_tips - 2
    parent = np.full(nn, -1, np.int32)
    # root 0 holds tips 1 and 2 and the first internal node; each internal
    # node below holds one tip and the next internal node, the last two tips
    parent[1] = parent[2] = 0
    prev, tip = 0, 3
    for k in range(n_tips + 1, nn):
        parent[k] = prev
        parent[tip] = k
        prev, tip = k, tip + 1
    parent[tip] = prev
    blen = rng.exponential(mean_blen, size=nn)
    blen[0] = 0.0
    return parent, blen


@pytest.mark.parametrize("prec", PRECS)
def test_large_and_deep_trees(gpu, prec):
    """500 taxa (a Yule tree) and a 160-taxon caterpillar, whose dependency
    chain is as long as the tree: many levels, many rescales."""
    import dataclasses
    ds = sim.gtr_gamma(n_tips=500, n_sites=1500, mean_blen=0.05, seed=380)
    _check(ds, prec, sweep=False)
    rng = np.random.default_rng(381)
    cat = sim.gtr_gamma(n_tips=160, n_sites=800, mean_blen=0.3, seed=382)
    parent, blen = _caterpillar(160, rng, 0.3)
    cat = dataclasses.replace(cat, parent=parent, blen=blen)
    _check(cat, prec)
