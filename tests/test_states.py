"""State matrices and state frequencies from alignments (host code, no GPU).

Codon states, the codon-to-amino-acid translation and the empirical
frequencies are checked against the reference's own data: z.11x2178.AA.nex is
the translation of z.11x2178.nex (its first 726 columns), c.start prints the
empirical codon frequencies of z.11x2178.nex to five decimals, and the
standard genetic code fixes the translation of ambiguous codons."""
import os

import numpy as np
import pytest

from garli_amd import engine as E
import refdata

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "data")


def _aln(name):
    return E.Alignment(os.path.join(DATA, name))


def _text_aln(tmp_path, rows, datatype="dna"):
    p = tmp_path / "t.nex"
    body = "\n".join("t%d %s" % (k, r) for k, r in enumerate(rows))
    p.write_text("#NEXUS\nbegin data; dimensions ntax=%d nchar=%d; format datatype=%s missing=? gap=-;\n"
                 "matrix\n%s\n;\nend;\n" % (len(rows), len(rows[0]), datatype, body))
    return E.Alignment(str(p))


def test_translation_is_the_reference_aa_file():
    aa = _aln("z.11x2178.nex").translate()
    ref = _aln("z.11x2178.AA.nex").states()
    assert aa.shape == (11, 726) and ref.shape == (11, 727)
    np.testing.assert_array_equal(aa, ref[:, :726])
    assert np.all(ref[:, 726] == 20)


def test_codon_states_and_printed_empirical_frequencies():
    a = _aln("z.11x2178.nex")
    cod, bf = a.codon_states()
    np.testing.assert_array_equal(cod, refdata.codon_matrix(refdata.load()["nuc_matrix"]))
    np.testing.assert_allclose(bf.sum(axis=1), 1.0, rtol=1e-15)
    pats, counts, _, _, site = a.codon_patterns()
    assert counts.sum() == 726 and np.all(site >= 0)
    pi = E.empirical_freqs(pats, counts, 61)
    for name in ("c", "c.M3x2"):
        printed = np.array(refdata.model_fields(refdata.load()["starts"][name]["model"])["e"])
        assert np.max(np.abs(pi - printed)) <= 5e-6 + 1e-12  # printed to 5 decimals


def test_empirical_aa_frequencies_of_file_and_translation_agree():
    p1 = E.empirical_freqs(*_aln("z.11x2178.AA.nex").aa_patterns()[:2], 20)
    p2 = E.empirical_freqs(*_aln("z.11x2178.nex").translated_patterns()[:2], 20)
    np.testing.assert_array_equal(p1, p2)
    assert abs(p1.sum() - 1.0) < 1e-14 and np.all(p1 > 0)


def test_stop_codons(tmp_path):
    """z.11x30.stop.nex: MorNa6 and ClownNa6 end in TAA."""
    a = _aln("z.11x30.stop.nex")
    for f in (a.codon_states, a.translate):
        with pytest.raises(E.GleError, match="stop codon TAA at codon 11"):
            f()
    cod, _ = a.codon_states(ignore_stops=True)
    aa = a.translate(ignore_stops=True)
    assert cod[0, 10] == 61 and cod[1, 10] == 61 and aa[0, 10] == 20 and aa[1, 10] == 20
    assert np.all(cod[2:, 10] == 61) and np.all(aa[2:, 10] == 20)  # '---'
    np.testing.assert_array_equal(aa[:, :10], _aln("z.11x30.AA.fas").states()[:, :10])


AA = "ACDEFGHIKLMNPQRSTVWY"


@pytest.mark.parametrize("codon,expect", [
    ("GCN", "A"), ("ATH", "I"), ("TAY", "Y"), ("YTR", "L"), ("MGR", "R"), ("GC-", "A"), ("AGY", "S"),
    ("TGR", None),   # TGA is a stop, TGG tryptophan: missing
    ("TA-", None),   # TAA / TAG stops among tyrosines: missing
    ("A-A", None),   # AAA lysine, ACA threonine ...: missing
    ("NNA", None),   # first two positions missing (TAA and TGA never examined)
    ("--T", None), ("???", None), ("ATG", "M"), ("TGG", "W"),
])
def test_translation_of_ambiguous_codons(tmp_path, codon, expect):
    """The standard code decides these: an ambiguous codon keeps its amino
    acid when every resolution codes for it."""
    a = _text_aln(tmp_path, [codon, "ATG"])
    aa = a.translate()
    assert aa[0, 0] == (20 if expect is None else AA.index(expect))
    cod, _ = a.codon_states()
    full = all(c in "ACGT" for c in codon)
    assert (cod[0, 0] != 61) == full


@pytest.mark.parametrize("codon", ["TAR", "TRA", "TAA"])
def test_all_resolutions_stop(tmp_path, codon):
    a = _text_aln(tmp_path, [codon, "ATG"])
    with pytest.raises(E.GleError, match="stop codon"):
        a.translate()
    assert a.translate(ignore_stops=True)[0, 0] == 20


def test_codon_base_frequencies_and_fx4(tmp_path):
    a = _text_aln(tmp_path, ["ATGGCC", "ATGGCA", "ATG-CA"])
    cod, bf = a.codon_states()
    # unambiguous codons: ATG x3, GCC, GCA (the third taxon's -CA is missing)
    np.testing.assert_allclose(bf[0], [3 / 5, 0, 2 / 5, 0])
    np.testing.assert_allclose(bf[1], [0, 2 / 5, 0, 3 / 5])
    np.testing.assert_allclose(bf[2], [1 / 5, 1 / 5, 3 / 5, 0])
    uniform = E.codon_freqs_fx4(np.full((3, 4), 0.25))
    np.testing.assert_allclose(uniform, 1.0 / 61)
    rng = np.random.default_rng(1)
    b = rng.dirichlet(np.ones(4), size=3)
    f3 = E.codon_freqs_fx4(b, f3x4=True)
    f1 = E.codon_freqs_fx4(b, f3x4=False)
    sense = [c for c in range(64) if c not in (48, 50, 56)]  # TAA, TAG, TGA
    pooled = b.mean(axis=0)
    want3 = np.array([b[0][c >> 4] * b[1][(c >> 2) & 3] * b[2][c & 3] for c in sense])
    want1 = np.array([pooled[c >> 4] * pooled[(c >> 2) & 3] * pooled[c & 3] for c in sense])
    np.testing.assert_allclose(f3, want3 / want3.sum(), rtol=1e-13)
    np.testing.assert_allclose(f1, want1 / want1.sum(), rtol=1e-13)
    with pytest.raises(E.GleError, match="F3x4"):
        E.codon_freqs_fx4(bf, f3x4=True)


def test_nucleotide_empirical_frequencies():
    """Without ambiguity: counts over total.  With R (A or G) cells the EM
    fixed point has a closed form: pA = a (a + g + r) / (N (a + g))."""
    rng = np.random.default_rng(5)
    m = rng.choice([1, 2, 4, 8], size=(6, 40)).astype(np.uint8)
    counts = rng.integers(1, 5, 40).astype(np.int32)
    pi = E.empirical_freqs(m, counts, 4)
    want = np.array([(counts * (m == b)).sum() for b in (1, 2, 4, 8)], float)
    np.testing.assert_allclose(pi, want / want.sum(), rtol=1e-14)
    a, g, r, c, t = 30, 10, 8, 20, 12
    cells = [1] * a + [4] * g + [5] * r + [2] * c + [8] * t + [15] * 7  # 7 fully missing cells
    m = np.array([cells], np.uint8)
    pi = E.empirical_freqs(m, np.ones(len(cells), np.int32), 4)
    N = a + g + r + c + t
    np.testing.assert_allclose(pi, [a * (a + g + r) / (N * (a + g)), c / N, g * (a + g + r) / (N * (a + g)), t / N],
                               atol=1e-8)
    # a base never seen: one pseudo-count each
    m = np.array([[1, 1, 2, 4]], np.uint8)
    np.testing.assert_allclose(E.empirical_freqs(m, np.ones(4, np.int32), 4), [3 / 8, 2 / 8, 2 / 8, 1 / 8])


def test_empirical_pseudocounts_and_nucleotide_guard():
    st = np.array([[0, 1, 2], [0, 1, 20]], np.uint8)
    pi = E.empirical_freqs(st, np.array([1, 2, 3], np.int32), 20)
    want = np.ones(20)
    want[0] += 2
    want[1] += 4
    want[2] += 3
    np.testing.assert_allclose(pi, want / want.sum())
    acgt = np.array([[0, 1, 5, 16, 0, 1]], np.uint8)  # A, C, G, T as amino acids
    with pytest.raises(E.GleError, match="nucleotide"):
        E.empirical_freqs(acgt, np.ones(6, np.int32), 20)
