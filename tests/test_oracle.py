"""Pins the CPU oracle: the reference's own golden scores, its own EISPACK code
(oracle/_ref), closed forms, library routines, brute force and invariants.
None of these checks retypes the oracle's formulas."""
import itertools

import numpy as np
import pytest
import scipy.linalg
import scipy.special

from oracle import oracle as O
from garli_amd import sim
import refdata


# ------------------------------------------------------------ golden scores
@pytest.mark.parametrize("name", ["n.G4", "n.G5"])
def test_reference_garliscore_nucleotide(name):
    """[!GarliScore] of trunk/tests/data/<name>.start (GTR+I+G on z.11x2178,
    the reference's own computed lnL, printed to 6 decimals)."""
    ds = refdata.nucleotide_dataset(name)
    lnl = O.Oracle(ds).score()
    assert abs(lnl - ds.meta["score"]) < 2e-6


@pytest.mark.parametrize("name", ["c.M3x2", "c"])
def test_reference_garliscore_codon(name):
    """Codon models (6-rate nucleotide exchangeabilities, omega mixture,
    empirical codon frequencies) against trunk/tests/data/<name>.start.  The
    printed omegas/rates carry 5 decimals, so the pin is ~1e-9 relative."""
    ds = refdata.codon_dataset(name)
    lnl = O.Oracle(ds).score()
    assert abs(lnl - ds.meta["score"]) < 2e-5


# ------------------------------------------------- eigen -> P(t) pins
def _models():
    rng = np.random.default_rng(5)
    pi4 = rng.dirichlet([10] * 4)
    rel = rng.lognormal(0, 0.5, 5)
    Q4, bm4 = O.qmat_gtr(pi4, rel)
    pi20 = rng.dirichlet([1] * 20) * 0.98 + 0.001
    pi20 /= pi20.sum()
    Q20, bm20 = O.qmat_aa(pi20, rng.lognormal(0, 1, 190))
    pi61 = rng.dirichlet([2] * 61) * 0.99 + 1e-4
    pi61 /= pi61.sum()
    Qc, bmc = O.qmat_codon(pi61, [1, 2.5, 1, 1, 2.5, 1], [0.1, 1.0, 3.0], [1 / 3] * 3)
    return [(Q4, bm4, pi4), (Q20, bm20, pi20), (Qc[2], bmc, pi61)]


@pytest.mark.parametrize("k", [0, 1, 2])
def test_pmat_matches_matrix_exponential(k):
    Q, bm, pi = _models()[k]
    V, W, lam = O.eigen(Q, pi)
    for t in (1e-6, 0.01, 0.3, 2.0):
        P, D1, D2 = O.pmat(V, W, lam, bm, t)
        E_ = scipy.linalg.expm(Q * bm * t)
        np.testing.assert_allclose(P, np.maximum(E_, 0), atol=2e-13)
        Qs = Q * bm
        np.testing.assert_allclose(D1, Qs @ E_, atol=1e-12)
        np.testing.assert_allclose(D2, Qs @ Qs @ E_, atol=1e-11)


@pytest.mark.parametrize("k", [0, 1, 2])
def test_pmat_matches_reference_eispack(k):
    """The reference's own linalg.cpp (EISPACK RG + LU inverse + c_ijk), built
    from its sources into oracle/_ref, gives the same P(t)."""
    ref = O.ref_linalg()
    if ref is None:
        pytest.skip("reference sources absent: oracle/_ref not built")
    import ctypes as C
    Q, bm, pi = _models()[k]
    n = len(pi)
    Qs = np.ascontiguousarray(Q * bm)
    lam, V, W = np.zeros(n), np.zeros(n * n), np.zeros(n * n)
    P = np.zeros(n * n)
    dp = C.POINTER(C.c_double)
    assert ref.ref_eigen(n, Qs.ctypes.data_as(dp), lam.ctypes.data_as(dp), V.ctypes.data_as(dp),
                         W.ctypes.data_as(dp)) == 0
    Vo, Wo, lo = O.eigen(Q, pi)
    for t in (0.001, 0.1, 1.5):
        ref.ref_pmat(n, V.ctypes.data_as(dp), W.ctypes.data_as(dp), lam.ctypes.data_as(dp), t,
                     P.ctypes.data_as(dp))
        Po, _, _ = O.pmat(Vo, Wo, lo, bm, t)
        np.testing.assert_allclose(Po, P.reshape(n, n), atol=1e-12)


def test_pmat_invariants_and_jc69_closed_form():
    pi = np.full(4, 0.25)
    Q, bm = O.qmat_gtr(pi, np.ones(5))
    V, W, lam = O.eigen(Q, pi)
    for t in (0.0, 0.05, 0.7):
        P, _, _ = O.pmat(V, W, lam, bm, t)
        e = np.exp(-4.0 * t / 3.0)
        np.testing.assert_allclose(np.diag(P), 0.25 + 0.75 * e, atol=1e-14)
        off = P[~np.eye(4, dtype=bool)]
        np.testing.assert_allclose(off, 0.25 - 0.25 * e, atol=1e-14)
    Q, bm, pi = _models()[1]
    V, W, lam = O.eigen(Q, pi)
    Pa, _, _ = O.pmat(V, W, lam, bm, 0.2)
    Pb, _, _ = O.pmat(V, W, lam, bm, 0.5)
    Pc, _, _ = O.pmat(V, W, lam, bm, 0.7)
    np.testing.assert_allclose(Pa @ Pb, Pc, atol=1e-13)                       # Chapman-Kolmogorov
    np.testing.assert_allclose(Pa.sum(1), 1.0, atol=1e-13)                    # stochastic rows
    np.testing.assert_allclose(pi[:, None] * Pa, (pi[:, None] * Pa).T, atol=1e-15)  # reversibility


def test_rate_normalisation():
    """Normalised Q has mean rate 1 (model.cpp:293, 669-686, 777)."""
    for Q, bm, pi in _models()[:2]:
        assert abs(-(pi * np.diag(Q)).sum() * bm - 1.0) < 1e-13
    rng = np.random.default_rng(9)
    pi = rng.dirichlet([2] * 61)
    om, pr = [0.1, 1.0, 3.0], [0.2, 0.5, 0.3]
    Qs, bm = O.qmat_codon(pi, [1.3, 2.0, 0.7, 1.1, 2.4, 1.0], om, pr)
    mean = sum(p * -(pi * np.diag(Qs[k])).sum() for k, p in enumerate(pr)) * bm
    assert abs(mean - 1.0) < 1e-13


def test_codon_lookup_genetic_code():
    lk, aa = O.codon_lookup()
    assert lk.shape == (61, 61) and (lk == lk.T).all()
    # 61 sense codons; ATG = Met (10), TGG = Trp (18): sense indices after removing
    # the stops TAA, TAG, TGA
    sense = [c for c in range(64) if c not in (48, 50, 56)]
    assert aa[sense.index(0b001110)] == 10 and aa[sense.index(0b111010)] == 18
    # each codon has <= 9 single-nucleotide neighbours; AAA->AAG is a synonymous transition
    assert ((lk & 1).sum(1) <= 9).all()
    a, g = sense.index(0), sense.index(2)
    assert lk[a, g] & 2 and not lk[a, g] & 4 and lk[a, g] & 16


def test_discrete_gamma_against_scipy():
    """Mean-of-category discrete gamma (Yang 1994) through scipy's incomplete
    gamma; the reference's AS 91 percentage points converge to 0.5e-6."""
    for alpha in (0.2, 0.5, 1.25046, 3.0):
        for K in (2, 4, 5):
            q = scipy.special.gammaincinv(alpha, np.arange(1, K) / K) / alpha
            cum = np.concatenate([[0.0], scipy.special.gammainc(alpha + 1, q * alpha), [1.0]])
            exact = K * np.diff(cum)
            got = O.discrete_gamma(K, alpha)
            np.testing.assert_allclose(got, exact, rtol=1e-5, atol=1e-7)
            assert abs(got.mean() - 1.0) < 1e-6


# ------------------------------------------------------- brute force lnL
def _brute(ds, parent, blen):
    """Sum over every assignment of internal states (tiny trees only)."""
    o = O.Oracle(ds, own_model=False)
    S, R, N = ds.n_states, ds.n_rates, ds.n_tips
    nn = 2 * N - 2
    internal = [0] + list(range(N + 1, nn))
    Ps = {n: o.pmats(blen[n])[0] for n in range(1, nn)}
    total = 0.0
    for p in range(ds.n_patterns):
        L = 0.0
        for r in range(R):
            lr = 0.0
            for assign in itertools.product(range(S), repeat=len(internal)):
                st = dict(zip(internal, assign))
                v = ds.pi[st[0]]
                for n in range(1, nn):
                    a = st[parent[n]]
                    if n in st:
                        v *= Ps[n][r][a, st[n]]
                    else:
                        code = ds.tips[n - 1, p]
                        if S == 4:
                            v *= 1.0 if code == 15 else sum(Ps[n][r][a, s] for s in range(4) if code >> s & 1)
                        else:
                            v *= 1.0 if code == S else Ps[n][r][a, code]
                lr += v
            L += ds.rate_probs[r] * lr
        if ds.pinv_in_model and p <= ds.last_const:
            cs = ds.const_states[p]
            L += ds.pinv * (sum(ds.pi[s] for s in range(4) if cs >> s & 1) if S == 4 else ds.pi[cs])
        total += ds.counts[p] * np.log(L)
    return total


@pytest.mark.parametrize("seed,ntax,pinv", [(3, 4, 0.0), (4, 5, 0.25), (5, 5, 0.0)])
def test_brute_force_nucleotide(seed, ntax, pinv):
    ds = sim.gtr_gamma(n_tips=ntax, n_sites=60, pinv=pinv, ambig_frac=0.15, seed=seed)
    got = O.Oracle(ds, own_model=False).score()
    assert abs(got - _brute(ds, ds.parent, ds.blen)) < 1e-10 * abs(got)


def test_brute_force_amino_acid():
    ds = sim.aa_gamma(n_tips=4, n_sites=25, n_patterns=None, ncat=2, seed=6)
    ds.tips[0, :3] = 20  # missing data
    got = O.Oracle(ds, own_model=False).score()
    assert abs(got - _brute(ds, ds.parent, ds.blen)) < 1e-10 * abs(got)


# --------------------------------------------------------- invariants
def test_rescale_and_root_invariance():
    """RunTests (population.cpp:896-954): rescaling frequency and the edge at
    which the tree is integrated do not change lnL."""
    ds = sim.gtr_gamma(n_tips=40, n_sites=400, mean_blen=0.6, pinv=0.1, seed=21)
    o = O.Oracle(ds)
    base = o.score()
    for every in (2, 4, 8, 30):
        assert abs(o.score(rescale_every=every) - base) < 1e-10 * abs(base)
    sw = o.sweep()
    assert np.ptp(sw[:, 0]) < 1e-10 * abs(base)
    assert abs(sw[0, 0] - base) < 1e-10 * abs(base)


def test_derivatives_match_finite_differences():
    """CalcEmpiricalDerivatives (optimization.cpp:1250-1320) is the reference's
    own check: central differences of the score."""
    ds = sim.gtr_gamma(n_tips=10, n_sites=500, pinv=0.2, seed=33)
    o = O.Oracle(ds)
    for n in (1, 7, 12):
        l0, d1, d2 = o.derivs(n)
        h = 1e-5 * max(ds.blen[n], 1e-3)
        bl = ds.blen.copy()
        bl[n] += h
        lp = o.score(blen=bl)
        bl[n] -= 2 * h
        lm = o.score(blen=bl)
        assert abs((lp - lm) / (2 * h) - d1) < 1e-5 * max(1.0, abs(d1))
        h2 = 1e-3 * max(ds.blen[n], 1e-2)  # larger step: second differences cancel badly
        bl = ds.blen.copy()
        bl[n] += h2
        lp = o.score(blen=bl)
        bl[n] -= 2 * h2
        lm = o.score(blen=bl)
        assert abs((lp - 2 * l0 + lm) / h2 ** 2 - d2) < 1e-3 * max(1.0, abs(d2))


def test_newton_pass_improves_and_converges():
    """OptimizeAllBranches never worsens lnL, and repeated passes converge."""
    ds = sim.gtr_gamma(n_tips=12, n_sites=800, seed=44)
    o = O.Oracle(ds)
    start = o.score()
    _, l1, bl1 = o.optimize_all(1e-4)
    _, l2, bl2 = o.optimize_all(1e-4, blen=bl1)
    _, l3, _ = o.optimize_all(1e-4, blen=bl2)
    assert start < l1 <= l2 + 1e-9 <= l3 + 2e-9
    assert (l3 - l2) < (l2 - l1) < (l1 - start)
