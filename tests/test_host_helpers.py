"""Engine host helpers (CPU side of libgle) against the oracle: tip encodings
and pattern compression bit-exact, model helpers to rounding."""
import numpy as np
import pytest

from garli_amd import engine as E
from oracle import oracle as O
import refdata

IUPAC = "ACGTUMRWSYKVHDBN-?acgtumrwsykvhdbn"


def test_nucleotide_datum_table():
    """NucleotideData::CharToDatum (sequencedata.cpp:672-713): IUPAC bit masks."""
    expect = {"A": 1, "C": 2, "G": 4, "T": 8, "U": 8, "M": 3, "R": 5, "W": 9, "S": 6, "Y": 10, "K": 12,
              "V": 7, "H": 11, "D": 13, "B": 14, "N": 15, "-": 15, "?": 15}
    for ch in IUPAC:
        assert E.nuc_char_to_datum(ch) == expect[ch.upper()] == O.lib().ora_nuc_char_to_datum(ch.encode())
    for ch in "XZJ*1":
        assert E.nuc_char_to_datum(ch) == -1 == O.lib().ora_nuc_char_to_datum(ch.encode())


def test_amino_acid_datum_table():
    order = "ACDEFGHIKLMNPQRSTVWY"
    for i, ch in enumerate(order):
        assert E.aa_char_to_datum(ch) == i == E.aa_char_to_datum(ch.lower())
        assert O.lib().ora_aa_char_to_datum(ch.encode()) == i
    for ch in "BZX-?":
        assert E.aa_char_to_datum(ch) == 20 == O.lib().ora_aa_char_to_datum(ch.encode())
    assert E.aa_char_to_datum("J") == -1


def test_ambiguity_stream_round_trip():
    """MakeAmbigStrings (sequencedata.cpp:278-362) <-> fixed 4-bit masks."""
    rng = np.random.default_rng(0)
    masks = rng.integers(1, 16, size=5000).astype(np.uint8)
    st = E.make_ambig_string(masks)
    assert np.array_equal(st, O.make_ambig_string(masks))
    back = E.ambig_string_to_masks(st, len(masks))
    assert np.array_equal(back, masks)
    # the stream's own grammar: single states 0..3, -4 total ambiguity, -n then n states
    assert E.make_ambig_string(np.array([1, 15, 5, 14], np.uint8)).tolist() == [0, -4, -2, 0, 2, -3, 1, 2, 3]


def test_ambiguity_stream_rejects_garbage():
    with pytest.raises(E.GleError):
        E.ambig_string_to_masks(np.array([-2, 0], np.int8), 1)
    with pytest.raises(E.GleError):
        E.ambig_string_to_masks(np.array([7], np.int8), 1)


@pytest.mark.parametrize("ms,seed", [(4, 1), (4, 2), (20, 3), (61, 4)])
def test_pattern_compression_matches_oracle(ms, seed):
    rng = np.random.default_rng(seed)
    ntax, nsite = 9, 700
    if ms == 4:
        m = (1 << rng.integers(0, 4, (ntax, nsite))).astype(np.uint8)
        m[rng.random(m.shape) < 0.08] = rng.integers(1, 16, int((rng.random(m.shape) < 0.08).sum()) or 1)[0]
        m[:, :40] = m[:1, :40]           # constant columns
        m[:, 40:45] = 15                 # all missing
        m[:3, 45:60] = 5                 # partial ambiguity in otherwise constant columns
        m[3:, 45:60] = 1
    else:
        m = rng.integers(0, 3, (ntax, nsite)).astype(np.uint8)
        m[rng.random(m.shape) < 0.05] = ms
        m[:, :30] = m[:1, :30]
        m[:, 30:33] = ms
    got = E.compress_patterns(m, ms)
    ref = O.compress_patterns(m, ms)
    P = got[0].shape[1]
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    assert got[3] == ref[3] and np.array_equal(got[2][:got[3] + 1], ref[2][:ref[3] + 1])
    assert np.array_equal(got[4], ref[4])
    # invariants: counts cover every non-missing site; columns map to identical patterns
    s2p = got[4]
    assert got[1].sum() == (s2p >= 0).sum()
    for s in np.flatnonzero(s2p >= 0)[:200]:
        assert np.array_equal(got[0][:, s2p[s]], m[:, s])
    assert len({got[0][:, p].tobytes() for p in range(P)}) == P


def test_compression_of_reference_alignment():
    d = refdata.load()
    codes = np.array([[E.nuc_char_to_datum(c) for c in seq] for seq in d["nuc_matrix"]], np.uint8)
    got = E.compress_patterns(codes, 4)
    ref = O.compress_patterns(codes, 4)
    assert got[0].shape == (11, 845) and np.array_equal(got[0], ref[0]) and got[3] == ref[3] == 10


def _check_same_pmats(Vp, Wp, lp, Vo, Wo, lo, mult=1.0):
    for t in (0.0, 0.01, 0.5, 3.0):
        Pp = (Vp * np.exp(lp * mult * t)) @ Wp
        Po, _, _ = O.pmat(Vo, Wo, lo, mult, t)
        np.testing.assert_allclose(np.maximum(Pp, 0), Po, atol=1e-13)


def test_model_helpers_match_oracle():
    rng = np.random.default_rng(8)
    pi = rng.dirichlet([5] * 4)
    rel = rng.lognormal(0, 0.5, 5)
    Vp, Wp, lp = E.model_gtr(pi, rel)
    Q, bm = O.qmat_gtr(pi, rel)
    Vo, Wo, lo = O.eigen(Q, pi)
    _check_same_pmats(Vp, Wp, lp, Vo, Wo, lo * bm)
    pi20 = rng.dirichlet([2] * 20)
    r190 = rng.lognormal(0, 1, 190)
    Vp, Wp, lp = E.model_aa(pi20, r190)
    Q, bm = O.qmat_aa(pi20, r190)
    Vo, Wo, lo = O.eigen(Q, pi20)
    _check_same_pmats(Vp, Wp, lp, Vo, Wo, lo * bm)
    pi61 = rng.dirichlet([2] * 61)
    rel6 = [1.2, 2.2, 0.8, 1.1, 2.9, 1.0]
    Vp, Wp, lp = E.model_codon(pi61, rel6, [0.1, 1.0, 3.0], [0.3, 0.3, 0.4])
    Qs, bm = O.qmat_codon(pi61, rel6, [0.1, 1.0, 3.0], [0.3, 0.3, 0.4])
    for k in range(3):
        Vo, Wo, lo = O.eigen(Qs[k], pi61)
        _check_same_pmats(Vp[k], Wp[k], lp[k], Vo, Wo, lo * bm)


def test_discrete_gamma_matches_oracle():
    for a in (0.05, 0.2, 0.5, 1.0, 1.25046, 7.5, 50.0):
        for k in (2, 3, 4, 5, 8):
            np.testing.assert_allclose(E.discrete_gamma(k, a), O.discrete_gamma(k, a), rtol=1e-13, atol=1e-15)
    with pytest.raises(E.GleError):
        E.discrete_gamma(4, -1.0)
