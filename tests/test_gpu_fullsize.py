"""The benchmark configurations at their full sizes, and the CPU-runnable one.

Nucleotide GTR+I+G4 on 128 taxa x 100,000 patterns, amino acids G4 on 64 x
20,000 and codon M3 on 32 x 5,000 (the workloads bench.py times): the FP64
score equals the oracle's (one full-size oracle score takes about a second),
and size-independent properties hold over the whole tree: lnL is the same on
every one of the 2N - 3 edges (root invariance), it does not depend on how
often CLAs are rescaled, and the FP32 form is within 1e-3 of the FP64 form
on every edge."""
import numpy as np
import pytest

from garli_amd import engine as E
from garli_amd import sim
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _nuc():
    return sim.gtr_gamma(n_tips=128, n_sites=None, n_patterns=100000, mean_blen=0.05, alpha=None, pinv=0.2,
                         ambig_frac=0.0, seed=2)


def _workload(kind):
    if kind == "nucleotide":
        return _nuc()
    if kind == "aa":
        return sim.aa_gamma(n_tips=64, n_patterns=20000, seed=3)
    return sim.codon_m3(n_tips=32, n_patterns=5000, seed=4)


@pytest.mark.parametrize("kind", ["nucleotide", "aa", "codon"])
def test_full_size_parity_and_invariants(gpu, kind):
    ds = _workload(kind)
    print(kind, ds.n_tips, "taxa", ds.n_patterns, "patterns", flush=True)
    ref = O.Oracle(ds).score()
    t64 = E.Tree(ds.make_instance(E.PREC_FP64), ds.parent, ds.blen)
    lnl = t64.score()
    assert abs(lnl - ref) < 1e-9 * abs(ref), (lnl, ref)
    sw64 = t64.sweep()
    # root invariance: every edge integrates to the same lnL
    assert np.max(np.abs(sw64[:, 0] - ref)) < 1e-9 * abs(ref)
    # rescale frequency does not change lnL (counts differ, the score does not)
    t64.set_rescale_every(4)
    t64.invalidate()
    assert abs(t64.score() - ref) < 1e-9 * abs(ref)
    t64.close()
    t32 = E.Tree(ds.make_instance(E.PREC_FP32), ds.parent, ds.blen)
    sw32 = t32.sweep()
    assert np.max(np.abs(sw32[:, 0] - sw64[:, 0])) < 1e-3
    # first derivatives: the FP32 form's error is small against the branch's scale
    d1err = np.abs(sw32[:, 1] - sw64[:, 1]) / np.maximum(1.0, np.abs(sw64[:, 1]))
    assert np.max(d1err) < 1e-3
    t32.close()


def test_cpu_configuration_score_and_newton_pass(gpu):
    """The CPU-runnable configuration: 24 taxa, 2,000 sites under GTR+G4 with
    5% ambiguity codes; full lnL, then one Newton pass over all 45 edges,
    against the oracle's own pass."""
    ds = sim.gtr_gamma()  # 24 taxa, 2000 sites, alpha 0.5, 5% IUPAC codes
    o = O.Oracle(ds)
    t = E.Tree(ds.make_instance(E.PREC_FP64), ds.parent, ds.blen)
    ref = o.score()
    assert abs(t.score() - ref) < 1e-9 * abs(ref)
    imp, lnl = t.optimize_all_branches(0.01)
    oimp, olnl, obl = o.optimize_all(0.01)
    np.testing.assert_allclose(t.branch_lengths()[1:], obl[1:], rtol=1e-7, atol=1e-10)
    assert abs(lnl - olnl) < 1e-9 * abs(olnl) and lnl > ref
    t.close()
