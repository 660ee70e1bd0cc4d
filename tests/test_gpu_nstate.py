"""Amino-acid (20-state) and codon (61-state) paths on the GPU against the
oracle, and the reference's own golden scores through the engine."""
import numpy as np
import pytest

from garli_amd import engine as E
from garli_amd import sim
from oracle import oracle as O
import refdata

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return abs(a - b) / abs(b)


@pytest.mark.parametrize("name", ["n.G4", "n.G5"])
def test_reference_garliscore_nucleotide_engine(gpu, name):
    ds = refdata.nucleotide_dataset(name)
    for prec, tol in ((E.PREC_FP64, 2e-6), (E.PREC_FP32, 1e-3)):
        t = E.Tree(ds.make_instance(prec), ds.parent, ds.blen)
        assert abs(t.score() - ds.meta["score"]) < tol


@pytest.mark.parametrize("name", ["c.M3x2", "c"])
def test_reference_garliscore_codon_engine(gpu, name):
    """Also the reference's scoring tests trunk/tests/scoring/c.conf and
    c.M3x2.conf: they score z.11x2178.nex (no weights, so runtests.sh's
    halved sitelikes sum is -lnL) from these start files, whose parameters
    are already optimal; expected.scr holds the same numbers, accepted within
    0.01."""
    ds = refdata.codon_dataset(name)
    expect = refdata.load()["expected_scr"][name + ".conf"]
    for prec, tol in ((E.PREC_FP64, 2e-5), (E.PREC_FP32, 1e-3)):
        t = E.Tree(ds.make_instance(prec), ds.parent, ds.blen)
        lnl = t.score()
        assert abs(lnl - ds.meta["score"]) < tol
        assert abs(-lnl - expect) < 0.01


@pytest.mark.parametrize("seed,ntax,npat,ncat", [(5, 6, 300, 4), (6, 17, 1500, 4), (7, 9, 500, 1)])
def test_aa_fp64_matches_oracle(gpu, seed, ntax, npat, ncat):
    ds = sim.aa_gamma(n_tips=ntax, n_patterns=npat, ncat=ncat, seed=seed)
    ds.tips[0, ::7] = 20  # missing data
    print("aa case", seed, ntax, ds.n_patterns, flush=True)
    t = E.Tree(ds.make_instance(E.PREC_FP64), ds.parent, ds.blen)
    o = O.Oracle(ds)
    assert _rel(t.score(), o.score()) < 1e-9
    sw = t.sweep()
    osw = o.sweep()
    np.testing.assert_allclose(sw[:, 0], osw[:, 0], rtol=1e-9)
    np.testing.assert_allclose(sw[:, 1:], osw[:, 1:], rtol=1e-8, atol=1e-8)


def test_aa_every_cla_and_counts(gpu):
    ds = sim.aa_gamma(n_tips=24, n_patterns=400, mean_blen=0.8, seed=9)
    t = E.Tree(ds.make_instance(E.PREC_FP64), ds.parent, ds.blen)
    o = O.Oracle(ds, own_model=False)  # same eigensystem: element-wise comparison of tiny entries
    rescaled = 0
    for a in (0, ds.n_tips + 1, ds.n_tips + 5, 2 * ds.n_tips - 3):
        nbrs = ([int(ds.parent[a])] if a else []) + [n for n in range(1, 2 * ds.n_tips - 2) if ds.parent[n] == a]
        for b in nbrs:
            g, gu = t.get_cla(a, b)
            r, ru = o.cla(a, b)
            np.testing.assert_array_equal(gu, ru)
            np.testing.assert_allclose(g, r, rtol=1e-12, atol=1e-300)
            rescaled += int((ru > 0).sum())
    assert rescaled > 0


def test_aa_fp32_within_tolerance(gpu):
    ds = sim.aa_gamma(n_tips=32, n_patterns=3000, seed=10)
    t = E.Tree(ds.make_instance(E.PREC_FP32), ds.parent, ds.blen)
    assert abs(t.score() - O.Oracle(ds).score()) < 1e-3


@pytest.mark.parametrize("seed,ntax,npat", [(12, 6, 150), (13, 12, 400)])
def test_codon_fp64_matches_oracle(gpu, seed, ntax, npat):
    ds = sim.codon_m3(n_tips=ntax, n_patterns=npat, seed=seed)
    ds.tips[1, ::5] = 61
    print("codon case", seed, ntax, ds.n_patterns, flush=True)
    t = E.Tree(ds.make_instance(E.PREC_FP64), ds.parent, ds.blen)
    o = O.Oracle(ds)
    assert _rel(t.score(), o.score()) < 1e-9
    for n in (1, ntax + 2):
        l, d1, d2 = t.derivatives(n)
        rl, r1, r2 = o.derivs(n)
        assert _rel(l, rl) < 1e-9
        assert abs(d1 - r1) < 1e-8 * max(1, abs(r1)) and abs(d2 - r2) < 1e-8 * max(1, abs(r2))


def test_codon_every_cla_and_counts(gpu):
    """FP64 MFMA codon CLAs against the oracle's sequential sums, element-wise
    to 1e-12 of each pattern's largest entry (the MFMA sums in another order,
    so entries far below the maximum carry that order's rounding); rescale
    counts equal."""
    ds = sim.codon_m3(n_tips=20, n_patterns=300, seed=15)
    ds.tips[2, ::3] = 61
    t = E.Tree(ds.make_instance(E.PREC_FP64), ds.parent, ds.blen)
    o = O.Oracle(ds, own_model=False)
    rescaled = 0
    for a in (0, ds.n_tips + 1, ds.n_tips + 7, 2 * ds.n_tips - 3):
        nbrs = ([int(ds.parent[a])] if a else []) + [n for n in range(1, 2 * ds.n_tips - 2) if ds.parent[n] == a]
        for b in nbrs:
            g, gu = t.get_cla(a, b)
            r, ru = o.cla(a, b)
            np.testing.assert_array_equal(gu, ru)
            scale = r.max(axis=(1, 2), keepdims=True)
            assert np.all(np.abs(g - r) <= 1e-12 * scale)
            rescaled += int((ru > 0).sum())
    assert rescaled > 0


def test_codon_fp32_within_tolerance(gpu):
    ds = sim.codon_m3(n_tips=16, n_patterns=800, seed=14)
    t = E.Tree(ds.make_instance(E.PREC_FP32), ds.parent, ds.blen)
    o = O.Oracle(ds)
    assert abs(t.score() - o.score()) < 1e-3
    sw = t.sweep()
    osw = o.sweep()
    assert np.max(np.abs(sw[:, 0] - osw[:, 0])) < 1e-3
