"""Per-pattern values of an edge: the inputs of the reference's site
derivative dumps (Tree::OutputSiteDerivatives, tree.cpp:5567-5595).

Each pattern's lnL is the same on every edge and equals the oracle's site lnL;
the count-weighted sums of the per-pattern lnL, d1 and d2 are the edge's
totals; d1 is the central difference of the site lnL in the branch length; the
rescale counts are those of the edge's two CLAs."""
import numpy as np
import pytest

from garli_amd import engine as E
from garli_amd import sim
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _check_sums(t, node, counts, tol):
    (lnl, d1, d2), v = t.site_derivatives(node)
    l2, e1, e2 = t.derivatives(node)
    assert abs(lnl - l2) <= tol * abs(l2) and abs(d1 - e1) <= tol * max(1.0, abs(e1))
    assert abs(np.dot(counts, v["lnl"]) - lnl) <= tol * abs(lnl)
    assert abs(np.dot(counts, v["d1"]) - d1) <= tol * max(1.0, abs(d1))
    assert abs(np.dot(counts, v["d2"]) - d2) <= tol * max(1.0, abs(d2))
    return v


@pytest.mark.parametrize("kind", ["nucleotide", "aa", "codon"])
def test_site_values_fp64(gpu, kind):
    if kind == "nucleotide":
        ds = sim.gtr_gamma(n_tips=14, n_sites=900, pinv=0.2, seed=301)  # with ambiguity codes
    elif kind == "aa":
        ds = sim.aa_gamma(n_tips=10, n_patterns=400, seed=302)
    else:
        ds = sim.codon_m3(n_tips=8, n_patterns=150, seed=303)
    print(kind, ds.n_patterns, "patterns", flush=True)
    ref, site_ref, _ = O.Oracle(ds).score(sites=True)
    t = E.Tree(ds.make_instance(E.PREC_FP64), ds.parent, ds.blen)
    inner = next(n for n in range(ds.n_tips + 1, 2 * ds.n_tips - 2))
    for node in (1, inner):  # a tip child and an internal child
        v = _check_sums(t, node, ds.counts, 1e-10)
        np.testing.assert_allclose(v["lnl"], site_ref, rtol=1e-9)
        _, up = t.get_cla(ds.parent[node], node)
        np.testing.assert_array_equal(v["under_partial"], up)
        if node <= ds.n_tips:
            assert np.all(v["under_child"] == -1)
        else:
            _, uc = t.get_cla(node, ds.parent[node])
            np.testing.assert_array_equal(v["under_child"], uc)
    t.close()


def test_site_d1_is_the_derivative_of_site_lnl(gpu):
    ds = sim.gtr_gamma(n_tips=10, n_sites=600, pinv=0.1, seed=304)
    t = E.Tree(ds.make_instance(E.PREC_FP64), ds.parent, ds.blen)
    node = ds.n_tips + 2
    b = ds.blen[node]
    _, v = t.site_derivatives(node)
    h = 1e-5 * b
    t.set_branch_length(node, b + h)
    _, vp = t.site_derivatives(node)
    t.set_branch_length(node, b - h)
    _, vm = t.site_derivatives(node)
    fd = (vp["lnl"] - vm["lnl"]) / (2 * h)
    big = np.abs(v["d1"]) > 1e-2
    assert big.sum() > 50
    np.testing.assert_allclose(fd[big], v["d1"][big], rtol=1e-4)
    t.close()


def test_site_values_fp32_heavy_split(gpu):
    """FP32 form with heavy patterns in the FP64 part: every pattern's values
    come back in the caller's order, within the form's tolerance of FP64."""
    ds = sim.gtr_gamma(n_tips=12, n_sites=3000, pinv=0.3, seed=305)
    assert np.any(ds.counts > 4)
    t64 = E.Tree(ds.make_instance(E.PREC_FP64), ds.parent, ds.blen)
    inst = ds.make_instance(E.PREC_FP32)
    inst.set_heavy_split(4)
    assert inst.heavy_patterns() > 0
    t32 = E.Tree(inst, ds.parent, ds.blen)
    node = ds.n_tips + 3
    _, v64 = t64.site_derivatives(node)
    v32 = _check_sums(t32, node, ds.counts, 1e-9)
    np.testing.assert_allclose(v32["lnl"], v64["lnl"], atol=1e-5)
    np.testing.assert_allclose(v32["d1"], v64["d1"], atol=1e-4 * max(1.0, np.abs(v64["d1"]).max()))
    t32.close()
    t64.close()


def test_site_likelihoods_follow_the_last_score(gpu):
    """Site likelihoods are computed on request for the last score: their
    count-weighted sum is that score, and a tree changed since the score has
    none until it is scored again."""
    ds = sim.gtr_gamma(n_tips=12, n_sites=800, pinv=0.1, seed=306)
    t = E.Tree(ds.make_instance(E.PREC_FP64), ds.parent, ds.blen)
    with pytest.raises(E.GleError):
        t.site_likelihoods()
    lnl = t.score()
    site = t.site_likelihoods()
    assert abs(np.dot(ds.counts, site) - lnl) < 1e-10 * abs(lnl)
    np.testing.assert_array_equal(t.site_likelihoods(), site)  # asked again: the same values
    t.derivatives(ds.n_tips + 2)  # an evaluation elsewhere changes nothing
    np.testing.assert_array_equal(t.site_likelihoods(), site)
    t.set_branch_length(3, ds.blen[3] * 1.5)
    with pytest.raises(E.GleError):
        t.site_likelihoods()
    lnl2 = t.score()
    site2 = t.site_likelihoods()
    assert abs(np.dot(ds.counts, site2) - lnl2) < 1e-10 * abs(lnl2) and lnl2 != lnl
    ref = O.Oracle(ds).score(blen=t.branch_lengths(), sites=True)[1]
    np.testing.assert_allclose(site2, ref, rtol=1e-9)
    t.close()
