"""Model-parameter optimisation on a fixed tree, checked end to end against
the reference's own scoring tests.

trunk/tests/scoring/n.G4.conf and n.G5.conf run GARLI with
optimizeinputonly = 1 on data/z.11x2178.wtset.nex (GTR+I+G with 4 or 5 gamma
categories; relative rates, base frequencies, shape and pinv estimated;
branch lengths optimised) from the trees and parameters of data/n.G4.start /
n.G5.start.  runtests.sh then sums column 2 of the run's .sitelikes.log (the
tree's -lnL plus each listed character's -lnL) and halves it; the result must
be within 0.01 of data/expected.scr.  Here the same file is ingested
(its EXSET and WTSET included), optimised on the GPU, written by
gle_write_sitelikes and summed the same way."""
import dataclasses
import os

import numpy as np
import pytest

from garli_amd import engine as E
from garli_amd import sim
from oracle import oracle as O
import refdata

pytestmark = pytest.mark.gpu
DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "data")


def _start(ds):
    return E.GtrParams(ds.meta["rel"], ds.pi, ds.meta["alpha"], ds.pinv, ds.pinv_in_model)


def _dataset_at(ds, p, blen):
    """The same data under model p and branch lengths blen; the oracle builds
    its own Q, eigensystem and gamma rates from these parameters."""
    R = ds.n_rates
    pinv = p.pinv if p.pinv_in_model else 0.0
    return dataclasses.replace(ds, pi=np.array(p.pi), rate_probs=np.full(R, (1.0 - pinv) / R), pinv=pinv,
                               pinv_in_model=p.pinv_in_model, blen=np.asarray(blen),
                               meta=dict(ds.meta, rel=np.array(p.rel), alpha=p.alpha))


def _runtests_score(path):
    """runtests.sh: awk '{sum+=$2}' over the sitelikes file, divided by 2."""
    total = 0.0
    for ln in open(path):
        f = ln.split()
        try:
            total += float(f[1])
        except (IndexError, ValueError):
            pass
    return total / 2.0


@pytest.mark.parametrize("name", ["n.G4", "n.G5"])
def test_optimize_input_only_matches_expected_scr(gpu, name, tmp_path):
    plain = refdata.nucleotide_dataset(name)  # start tree and model of <name>.start
    expect = refdata.load()["expected_scr"][name + ".conf"]
    aln = E.Alignment(os.path.join(DATA, "z.11x2178.wtset.nex"))
    pats, counts, cs, lc, site_pat = aln.nucleotide_patterns()
    ds = dataclasses.replace(plain, tips=pats, counts=counts, const_states=cs, last_const=lc)
    t = E.Tree(ds.make_instance(E.PREC_FP64), ds.parent, ds.blen)
    start = _start(ds)
    t.set_gtr(start)
    assert abs(t.score() - ds.meta["score"]) < 2e-6  # the start file's own [!GarliScore]
    lnl, opt = t.optimize_gtr(start, precision=1e-4)
    assert lnl >= ds.meta["score"] - 1e-9
    per_pattern = t.site_likelihoods()
    per_char = np.where(site_pat >= 0, per_pattern[np.maximum(site_pat, 0)], 0.0)
    out = tmp_path / ("scr.%s.sitelikes.log" % name)
    E.write_sitelikes(str(out), lnl, per_char, site_pat >= 0)
    score = _runtests_score(out)
    print(name, "lnL %.6f, runtests score %.5f, expected.scr %.5f" % (lnl, score, expect), opt, flush=True)
    assert abs(score - expect) < 0.01
    # the optimum's score is the oracle's score of the same model and tree
    ref = O.Oracle(_dataset_at(ds, opt, t.branch_lengths())).score()
    assert abs(lnl - ref) < 1e-9 * abs(ref)
    t.close()


def test_optimizer_climbs_from_a_perturbed_start(gpu):
    """From rates, frequencies, shape, pinv and branch lengths moved away from
    the n.G4 optimum the optimiser returns to it (the start file's parameters
    are that optimum: its [!GarliScore] is already the maximum)."""
    ds = refdata.nucleotide_dataset("n.G4")
    rng = np.random.default_rng(3)
    far = E.GtrParams(ds.meta["rel"] * rng.uniform(0.5, 2.0, 5), [0.3, 0.2, 0.2, 0.3], 0.6, 0.1, True)
    t = E.Tree(ds.make_instance(E.PREC_FP64), ds.parent, ds.blen * rng.uniform(0.5, 2.0, len(ds.blen)))
    t.set_gtr(far)
    low = t.score()
    lnl, opt = t.optimize_gtr(far, precision=1e-4)
    print("perturbed start %.4f -> %.4f (reference optimum %.4f)" % (low, lnl, ds.meta["score"]), opt, flush=True)
    assert low < ds.meta["score"] - 50.0
    assert abs(lnl - ds.meta["score"]) < 0.01
    np.testing.assert_allclose(opt.rel, ds.meta["rel"], rtol=0.02)
    np.testing.assert_allclose(opt.pi, ds.pi, atol=2e-3)
    t.close()


def test_each_parameter_search_improves(gpu):
    """One parameter group at a time never lowers lnL, and a second full pass
    from the optimum gains less than its precision."""
    ds = refdata.nucleotide_dataset("n.G4")
    t = E.Tree(ds.make_instance(E.PREC_FP64), ds.parent, ds.blen)
    p = _start(ds)
    t.set_gtr(p)
    prev = t.score()
    for which in (E.OPT_BRANCHES, E.OPT_RATES, E.OPT_FREQS, E.OPT_ALPHA, E.OPT_PINV):
        lnl, p = t.optimize_gtr(p, which=which, precision=1e-3)
        assert lnl >= prev - 1e-9, (which, lnl, prev)
        prev = lnl
    lnl, p = t.optimize_gtr(p, precision=1e-3)
    again, _ = t.optimize_gtr(p, precision=1e-3)
    assert again - lnl < 0.01
    assert abs(p.pi.sum() - 1.0) < 1e-12 and np.all(p.rel > 0) and 0.0 <= p.pinv < 1.0
    t.close()


def test_fp32_form_optimum(gpu):
    """The FP32 bulk form reaches the same optimum within its lnL tolerance."""
    ds = refdata.nucleotide_dataset("n.G4")
    t = E.Tree(ds.make_instance(E.PREC_FP32), ds.parent, ds.blen)
    lnl, _ = t.optimize_gtr(_start(ds), precision=1e-3)
    assert abs(lnl - ds.meta["score"]) < 0.01
    t.close()


@pytest.mark.parametrize("name", ["c", "c.M3x2"])
def test_codon_optimizer_returns_to_the_reference_optimum(gpu, name):
    """trunk/tests/scoring/c.conf and c.M3x2.conf (optimizeinputonly = 1):
    their start files hold the optimum, which expected.scr repeats.  From
    exchangeabilities, omegas, omega proportions and branch lengths moved away
    from it, the optimiser climbs back to that score."""
    ds = refdata.codon_dataset(name)
    expect = refdata.load()["expected_scr"][name + ".conf"]
    rng = np.random.default_rng(11)
    rel = ds.meta["rel6"] * np.r_[rng.uniform(0.6, 1.6, 5), 1.0]
    omega = ds.meta["omegas"] * rng.uniform(0.5, 2.0, len(ds.meta["omegas"]))
    prob = np.full(len(omega), 1.0 / len(omega))
    far = E.CodonParams(rel, omega, prob, ds.pi)
    t = E.Tree(ds.make_instance(E.PREC_FP64), ds.parent, ds.blen * rng.uniform(0.7, 1.4, len(ds.blen)))
    t.set_codon(far)
    low = t.score()
    lnl, opt = t.optimize_codon(far, precision=1e-4)
    print(name, "perturbed start %.4f -> %.4f (reference %.4f)" % (low, lnl, -expect), opt, flush=True)
    assert low < -expect - 5.0
    assert abs(-lnl - expect) < 0.01  # runtests.sh's tolerance
    t.close()


@pytest.mark.parametrize("name", ["a", "a.G3", "a.G4"])
def test_amino_acid_scoring_matches_expected_scr(gpu, name, tmp_path):
    """trunk/tests/scoring/a.conf, a.G3.conf and a.G4.conf (optimizeinputonly
    = 1): the data ingested, translated and counted by the engine, the gamma
    shape and branch lengths optimised, the sitelikes file written and summed
    as runtests.sh does; within 0.01 of expected.scr.  (The start files'
    [!GarliScore] values came from other data: their printed frequencies
    differ from the data's in the fourth decimal.)"""
    ds = refdata.aa_dataset(name)
    expect = refdata.load()["expected_scr"][name + ".conf"]
    t = E.Tree(ds.make_instance(E.PREC_FP64), ds.parent, ds.blen)
    start = E.AaParams(ds.meta["rel"], ds.pi, ds.meta["alpha"])
    t.set_aa(start)
    low = t.score()
    ref0 = O.Oracle(ds).score()
    assert abs(low - ref0) < 1e-9 * abs(ref0)
    lnl, opt = t.optimize_aa(start, precision=1e-4)
    assert lnl >= low - 1e-9
    site = ds.meta["site"]
    per_pattern = t.site_likelihoods()
    per_char = np.where(site >= 0, per_pattern[np.maximum(site, 0)], 0.0)
    out = tmp_path / ("scr.%s.sitelikes.log" % name)
    E.write_sitelikes(str(out), lnl, per_char, site >= 0)
    score = _runtests_score(out)
    print(name, "start %.5f, lnL %.6f, runtests score %.5f, expected.scr %.5f" % (low, lnl, score, expect), opt,
          flush=True)
    assert abs(score - expect) < 0.01
    at_opt = dataclasses.replace(ds, blen=np.asarray(t.branch_lengths()), meta=dict(ds.meta, alpha=opt.alpha))
    ref = O.Oracle(at_opt).score()
    assert abs(lnl - ref) < 1e-9 * abs(ref)
    t.close()
    if name == "a.G4":  # the FP32 bulk form reaches the same score within its tolerance
        t32 = E.Tree(ds.make_instance(E.PREC_FP32), ds.parent, ds.blen)
        l32, _ = t32.optimize_aa(start, precision=1e-4)
        assert abs(l32 - lnl) < 2e-3
        t32.close()


def test_amino_acid_frequency_and_rate_estimation(gpu):
    """statefrequencies = estimate and ratematrix = estimate on the a.G3 data:
    frequencies estimated from a flat start reach at least the score of the
    empirical ones and land near them; a round over the 189 free
    exchangeabilities from WAG never lowers lnL and raises it."""
    ds = refdata.aa_dataset("a.G3")
    t = E.Tree(ds.make_instance(E.PREC_FP64), ds.parent, ds.blen)
    emp = E.AaParams(ds.meta["rel"], ds.pi, ds.meta["alpha"])
    l_emp, p_emp = t.optimize_aa(emp, which=E.OPT_ALPHA | E.OPT_BRANCHES, precision=1e-3)
    flat = E.AaParams(ds.meta["rel"], np.full(20, 0.05), p_emp.alpha)
    t.set_aa(flat)
    low = t.score()
    l_est, p_est = t.optimize_aa(flat, which=E.OPT_FREQS | E.OPT_ALPHA | E.OPT_BRANCHES, precision=1e-3)
    print("frequencies: flat %.4f -> estimated %.4f (empirical %.4f)" % (low, l_est, l_emp), flush=True)
    assert l_est >= l_emp - 1e-3 and l_est > low + 10.0
    assert abs(p_est.pi.sum() - 1.0) < 1e-12 and np.max(np.abs(p_est.pi - ds.pi)) < 0.02
    l_rates, p_rates = t.optimize_aa(p_emp, which=E.OPT_RATES, precision=1e9)  # exactly one round
    print("exchangeabilities: WAG %.4f -> %.4f" % (l_emp, l_rates), flush=True)
    assert l_rates > l_emp + 1.0 and np.all(p_rates.rel > 0) and p_rates.rel[189] == ds.meta["rel"][189]
    with pytest.raises(E.GleError):
        t.optimize_aa(p_emp, which=E.OPT_OMEGA)
    zero = E.AaParams(np.r_[0.0, ds.meta["rel"][1:]], ds.pi, 0.5)
    with pytest.raises(E.GleError):
        t.optimize_aa(zero, which=E.OPT_RATES)
    t.close()


def test_codon_frequency_estimation(gpu):
    """statefrequencies = estimateF (model.h:871-876; OptimizeEquilibriumFreqs,
    tree.cpp:7288-7303) on simulated M0 codon data, 8 taxa x 300 patterns,
    starting from the data's empirical frequencies: the 61 estimated
    frequencies raise lnL, sum to 1, and a second run from them gains almost
    nothing; the engine's score at the estimate equals the CPU oracle's, whose
    codon Q and eigensystem are built from those frequencies independently."""
    ds = sim.codon_m3(n_tips=8, n_patterns=300, omegas=(0.4,), seed=31)
    emp = E.empirical_freqs(ds.tips, ds.counts, 61)
    start = E.CodonParams(ds.meta["rel6"], np.array([0.4]), np.array([1.0]), emp)
    t = E.Tree(ds.make_instance(E.PREC_FP64), ds.parent, ds.blen)
    t.set_codon(start)
    s0 = t.score()
    lnl, est = t.optimize_codon(start, which=E.OPT_FREQS, precision=0.05)
    print("empirical %.4f -> estimated %.4f" % (s0, lnl), flush=True)
    assert lnl > s0 + 0.5
    assert abs(est.pi.sum() - 1.0) < 1e-12 and np.all(est.pi > 0)
    again, _ = t.optimize_codon(est, which=E.OPT_FREQS, precision=0.05)
    assert again - lnl < 0.5
    t.set_codon(est)
    mine = t.score()
    ref = O.Oracle(dataclasses.replace(ds, pi=est.pi)).score()
    assert abs(mine - ref) < 1e-9 * abs(ref), (mine, ref)
    t.close()
