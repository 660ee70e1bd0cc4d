"""Alignment ingestion and site-likelihood output (host code, no GPU).

The files under tests/golden/data are the reference's own test data
(garli/trunk/tests/data), copied by tests/golden/make_fixtures.py."""
import os

import numpy as np
import pytest

from garli_amd import engine as E
import refdata

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "data")


def _aln(name):
    return E.Alignment(os.path.join(DATA, name))


def test_nexus_matrix_matches_fixture():
    a = _aln("z.11x2178.nex")
    d = refdata.load()
    assert a.names == d["taxa"] and a.rows == d["nuc_matrix"] and a.datatype == E.DATA_DNA
    assert np.all(a.weights == 1)


def test_wtset_and_exset():
    """z.11x2178.wtset.nex: 'exset * zeroWeight = 1995 816 1012 1019 1022
    2180' and 'wtset * compensate = 1:1-749 751-1008 1010-1928 1930-.,
    2:1929 750, 4:1009' (the file's ASSUMPTIONS block)."""
    a = _aln("z.11x2178.wtset.nex")
    assert a.n_chars == 2180
    w = a.weights
    expect = np.ones(2180, np.int32)
    expect[[1929 - 1, 750 - 1]] = 2
    expect[1009 - 1] = 4
    expect[[1995 - 1, 816 - 1, 1012 - 1, 1019 - 1, 1022 - 1, 2180 - 1]] = 0
    np.testing.assert_array_equal(w, expect)


def _pattern_multiset(pats, counts):
    return sorted((bytes(pats[:, p]), int(c)) for p, c in enumerate(counts))


def test_weighted_patterns_equal_the_plain_alignment():
    """The file's own comment: the wtset exactly counteracts the exset, so its
    weighted patterns are those of z.11x2178.nex (whose last two characters it
    lacks: one all-missing, one excluded)."""
    pw, cw, _, lcw, site = _aln("z.11x2178.wtset.nex").nucleotide_patterns()
    pp, cp, _, lcp, _ = _aln("z.11x2178.nex").nucleotide_patterns()
    assert _pattern_multiset(pw, cw) == _pattern_multiset(pp, cp)
    assert lcw == lcp and cw.sum() == cp.sum()
    assert site[2180 - 1] == -1 and site[1995 - 1] == -1 and site[2179 - 1] == -1  # excluded / all missing
    assert site[1929 - 1] >= 0


def test_phylip_fasta_and_names():
    plain = _aln("z.11x2178.nex")
    phy = _aln("z.11x30.phy")
    assert phy.names == plain.names and phy.rows == [r[:30] for r in plain.rows]
    wacky = _aln("z.11x30.wackyNames.nex")
    assert wacky.names[:6] == ["Mor Na6", "Clown Na6", "Ara-Na6", "puff -Na6", "New_Zebra", "Ster_-Na6"]
    assert wacky.rows == phy.rows
    fas = _aln("z.11x30.AA.fas")
    aa = _aln("z.11x2178.AA.nex")
    assert fas.datatype == E.DATA_PROTEIN and aa.datatype == E.DATA_PROTEIN
    assert fas.names == aa.names and fas.rows == [r[:10] + "-" for r in aa.rows]
    stop = _aln("z.11x30.stop.nex")
    assert stop.n_chars == 33 and [r[:30] for r in stop.rows] == phy.rows


def test_bad_files_raise(tmp_path):
    p = tmp_path / "short.nex"
    p.write_text("#NEXUS\nbegin data; dimensions ntax=2 nchar=4; format datatype=dna;\n"
                 "matrix\na ACGT\nb ACG\n;\nend;\n")
    with pytest.raises(E.GleError):
        E.Alignment(str(p))
    q = tmp_path / "missing.nex"
    with pytest.raises(E.GleError):
        E.Alignment(str(q))


def test_nexus_interleave_matchchar_and_sets(tmp_path):
    p = tmp_path / "il.nex"
    p.write_text("#NEXUS\n[comment]\nBEGIN TAXA; DIMENSIONS NTAX=3; TAXLABELS one 'two b' three; END;\n"
                 "BEGIN CHARACTERS; DIMENSIONS NCHAR=8; FORMAT DATATYPE=DNA MISSING=N GAP=- MATCHCHAR=. "
                 "INTERLEAVE;\nMATRIX\none ACGT\n'two b' .T{AG}-\nthree N[x]CCA\n\none GGTT\n'two b' ....\n"
                 "three TTN(CT)\n;\nEND;\nBEGIN ASSUMPTIONS; EXSET * ex = 2 7-8; WTSET * w = 3: 1, 2: 3-4; END;\n")
    a = E.Alignment(str(p))
    assert a.names == ["one", "two b", "three"]
    assert a.rows == ["ACGTGGTT", "ATR-GGTT", "?CCATT?Y"]
    np.testing.assert_array_equal(a.weights, [3, 0, 2, 2, 1, 1, 0, 0])


def test_sitelikes_file_and_runtests_sum(tmp_path):
    """GARLI's sitelikes layout; runtests.sh sums column 2 (tree total plus
    every listed site) and halves it."""
    path = tmp_path / "x.sitelikes.log"
    site = np.array([-1.5, -2.25, -3.0, -4.125])
    inc = np.array([1, 0, 1, 1])
    E.write_sitelikes(str(path), -12.0, site, inc, tree_num=1)
    lines = path.read_text().splitlines()
    assert lines[0] == "Tree\t-lnL\tSite\t-lnL"
    assert lines[1] == "\t\t1\t1.50000000" and lines[2] == "\t\t2\t-" and lines[5] == "1\t12"
    col2 = 0.0
    for ln in lines:
        f = ln.split()
        try:
            col2 += float(f[1])
        except (IndexError, ValueError):
            pass
    assert col2 / 2 == pytest.approx((12.0 + 1.5 + 3.0 + 4.125) / 2)
    E.write_sitelikes(str(path), -12.0, site, inc, tree_num=2, append=True)
    assert path.read_text().count("Tree\t-lnL") == 1 and path.read_text().splitlines()[-1] == "2\t12"


def test_site_derivative_logs(tmp_path):
    """Tree::OutputSiteDerivatives' two files (tree.cpp:5567-5595): the
    ordered one per character, "gap" (with no line end, as the reference
    writes it) for a character without a pattern; the packed one per pattern;
    %.10g values, "-" for the child count of a tip child."""
    vals = {"lnl": np.array([-1.25, -2.5, -0.123456789012]), "d1": np.array([0.5, -1e-5, 3.0]),
            "d2": np.array([-0.25, -2.0, 1.0 / 3.0]), "under_partial": np.array([0, 7, 14], np.int32),
            "under_child": np.array([-1, -1, -1], np.int32)}
    o, pk = tmp_path / "orderedSiteDerivs.log", tmp_path / "packedSiteDerivs.log"
    E.write_site_derivatives(str(o), str(pk), [2, -1, 0, 1], vals)
    assert o.read_text() == ("Partition subset 1\nsite#\ttruelnL\td1\td2\tunder1\tunder2\n"
                             "1\t-0.123456789\t3\t0.3333333333\t14\t-\n"
                             "2\tgap\t-\t-\t-\t-"
                             "3\t-1.25\t0.5\t-0.25\t0\t-\n"
                             "4\t-2.5\t-1e-05\t-2\t7\t-\n")
    assert pk.read_text().splitlines()[1:] == ["packedIndex\ttruelnL\td1\td2\tunder1\tunder2",
                                               "0\t-1.25\t0.5\t-0.25\t0\t-", "1\t-2.5\t-1e-05\t-2\t7\t-",
                                               "2\t-0.123456789\t3\t0.3333333333\t14\t-"]
    vals["under_child"] = np.array([3, 0, 5], np.int32)
    E.write_site_derivatives(None, str(pk), [0, 1, 2], vals, subset=2)
    assert pk.read_text().splitlines()[0] == "Partition subset 2"
    assert pk.read_text().splitlines()[2] == "0\t-1.25\t0.5\t-0.25\t0\t3"
    with pytest.raises(E.GleError):
        E.write_site_derivatives(str(o), None, [0, 5], vals)
