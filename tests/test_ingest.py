"""Alignment ingestion and site-likelihood output (host code, no GPU).

The files under tests/golden/data are the reference's own test data
(garli/trunk/tests/data), copied by tests/golden/make_fixtures.py."""
import os
import re

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


# ------------------------------------------------------------ starting trees
def _splits(parent, n_taxa):
    """The bipartitions of a tree given as a parent array, each as the
    frozenset of taxa on the side without taxon 1."""
    nn = len(parent)
    below = {n: ({n} if 1 <= n <= n_taxa else set()) for n in range(nn)}
    order = sorted(range(1, nn), key=lambda n: -_depth(parent, n))
    for n in order:
        below[int(parent[n])] |= below[n]
    out = set()
    for n in range(n_taxa + 1, nn):
        s = below[n]
        side = frozenset(s if 1 not in s else set(range(1, n_taxa + 1)) - s)
        out.add(side)
    return out


def _depth(parent, n):
    d = 0
    while n != 0:
        n = int(parent[n])
        d += 1
    return d


def _newick_splits(newick, n_taxa):
    """Bipartitions written in a newick string of taxon numbers, read with a
    plain stack (no branch lengths needed), the two sides of the top merged."""
    s = re.sub(r"\[.*?\]", "", newick)
    s = re.sub(r":[0-9.eE+-]+", "", s).strip().rstrip(";")
    stack, groups = [[]], []
    tok = ""
    for ch in s + ",":
        if ch == "(":
            stack.append([])
        elif ch in ",)":
            if tok:
                stack[-1].append({int(tok)})
                tok = ""
            if ch == ")":
                g = set().union(*stack.pop())
                groups.append(g)
                stack[-1].append(g)
        elif not ch.isspace():
            tok += ch
    full = set(range(1, n_taxa + 1))
    out = set()
    for g in groups:
        side = frozenset(g if 1 not in g else full - g)
        if 1 < len(side) < n_taxa - 1:
            out.add(side)
    return out


def _valid_unrooted(parent, n_taxa):
    nn = 2 * n_taxa - 2
    assert len(parent) == nn and parent[0] == -1
    deg = np.zeros(nn, int)
    for n in range(1, nn):
        deg[n] += 1
        deg[parent[n]] += 1
    assert all(deg[n] == 1 for n in range(1, n_taxa + 1))
    assert deg[0] == 3 and all(deg[n] == 3 for n in range(n_taxa + 1, nn))
    for n in range(1, nn):
        _depth(parent, n)  # reaches the root


@pytest.mark.parametrize("name", ["a.G3", "a.G4", "a", "c.M3x2", "c", "n.G4", "n.G5"])
def test_start_file_trees_match_the_fixture(name):
    """Each 11-taxon start file through the library's tree-file reader gives
    the tree, GarliScore and GarliModel of the fixture (independent newick
    parse of the same file)."""
    d = refdata.load()
    st = d["starts"][name]
    t = E.read_start_trees(os.path.join(DATA, name + ".start"), d["taxa"])
    assert len(t) == 1
    p, b = refdata.parse_newick(st["newick"], len(d["taxa"]))
    np.testing.assert_array_equal(t[0]["parent"], p)
    np.testing.assert_array_equal(t[0]["blen"], b)
    assert t[0]["score"] == st["score"] and t[0]["model"] == st["model"]
    assert not t[0]["polytomies_resolved"]
    # without names: the TRANSLATE tokens are the taxon numbers
    t2 = E.read_start_trees(os.path.join(DATA, name + ".start"), n_taxa=len(d["taxa"]))
    np.testing.assert_array_equal(t2[0]["parent"], p)


def test_start_file_without_a_tree_is_refused():
    """n.start holds only a garli block (the model) and no TREES block."""
    with pytest.raises(E.GleError):
        E.read_start_trees(os.path.join(DATA, "n.start"), n_taxa=11)


def test_rooted_start_trees_become_trifurcations():
    """p.3diff.start and L.start (PAUP trees, bifurcating at the top, no
    lengths): MakeTrifurcatingRoot dissolves the internal child of the root
    (tree.cpp:761-797) and every length is the starting length 0.05, the
    kept tip's doubled."""
    for fname, ntax in (("p.3diff.start", 11), ("L.start", 30)):
        txt = open(os.path.join(DATA, fname)).read()
        newick = re.search(r"tree\s+\S+\s*=\s*(.*?;)", txt, re.S).group(1)
        t = E.read_start_trees(os.path.join(DATA, fname), n_taxa=ntax)[0]
        _valid_unrooted(t["parent"], ntax)
        assert _splits(t["parent"], ntax) == _newick_splits(newick, ntax)
        # "(1,(2,..." : tip 1 is kept, the right child dissolved into the root
        assert t["parent"][1] == 0 and t["parent"][2] == 0
        assert t["blen"][1] == pytest.approx(0.1)
        assert np.all(t["blen"][2:] == 0.05) and np.isnan(t["score"]) and t["model"] == ""
        # internal nodes numbered as the description opens them, the
        # dissolved one given up: "(3,((4,11),..." opens ntax+1 under the root
        assert t["parent"][ntax + 1] == 0 and t["parent"][3] == ntax + 1


def test_rooted_tree_with_internal_left_child(tmp_path):
    """A rooted tree whose left top child is internal: it is the one
    dissolved, its length added to the right child's; numbers are those of
    the description with the dissolved node's given up (EliminateNode)."""
    f = tmp_path / "r.tre"
    f.write_text("#NEXUS\nbegin trees;\ntree r = (((1:0.1,2:0.2):0.3,3:0.4):0.5,(4:0.6,5:0.7):0.8);\nend;\n")
    t = E.read_start_trees(str(f), n_taxa=5)[0]
    # (1,2) -> node 6, (4,5) -> node 7; the root holds 6, 3 and 7
    np.testing.assert_array_equal(t["parent"], [-1, 6, 6, 0, 7, 7, 0, 0])
    np.testing.assert_allclose(t["blen"], [0, 0.1, 0.2, 0.4, 0.6, 0.7, 0.3, 1.3])


def test_polytomies_are_resolved(tmp_path):
    """moore.start (86 taxa of moore.matK90-120.nex, PAUP tree with
    polytomies and a bifurcating top, names through TRANSLATE): fully
    resolved (the reader returns 1), every split of the description kept,
    new branches 0.05 long."""
    aln = _aln("moore.matK90-120.nex")
    assert len(aln.names) == 86
    txt = open(os.path.join(DATA, "moore.start")).read()
    newick = re.search(r"tree\s+\S+\s*=\s*(.*?;)", txt, re.S).group(1)
    t = E.read_start_trees(os.path.join(DATA, "moore.start"), aln.names)[0]
    assert t["polytomies_resolved"]
    _valid_unrooted(t["parent"], 86)
    want = _newick_splits(newick, 86)
    got = _splits(t["parent"], 86)
    assert want < got and len(got) == 86 - 3
    # rooted as well ("(1,(((..."): tip 1 carries both top lengths
    assert t["blen"][1] == pytest.approx(0.1) and np.all(t["blen"][2:] == 0.05)
    # a trifurcating top with four children: one join
    f = tmp_path / "p.tre"
    f.write_text("#NEXUS\nbegin trees;\ntree p = (1:0.1,2:0.2,3:0.3,(4:0.4,5:0.5):0.6);\nend;\n")
    t = E.read_start_trees(str(f), n_taxa=5)[0]
    assert t["polytomies_resolved"]
    np.testing.assert_array_equal(t["parent"], [-1, 7, 7, 0, 6, 6, 0, 0])
    np.testing.assert_allclose(t["blen"], [0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.05])


def test_tree_labels_lengths_and_errors(tmp_path):
    """Names matched through TRANSLATE or directly (quoted, '_' as a blank),
    lengths held to [1e-8, 100], several trees in a block, and the errors
    Tree::Tree raises (tree.cpp:309-576): a tip not in the data, a taxon
    twice, a taxon missing."""
    names = ["Homo sapiens", "Pan", "Gorilla", "Pongo"]
    f = tmp_path / "n.tre"
    f.write_text("#NEXUS\n[comment; with ; semicolons]\nbegin trees;\n translate a Pan, b 'Homo sapiens',"
                 " c Gorilla, d Pongo;\n tree one = [&U] (a:0,b:1000,(c:0.25,d:0.5):0.125);\n"
                 " tree two = [&U][!GarliScore -12.5] ('Homo sapiens':1,Pan:2,(Gorilla:3,Pongo:4):5);\nend;\n")
    t = E.read_start_trees(str(f), names)
    assert len(t) == 2
    np.testing.assert_array_equal(t[0]["parent"], [-1, 0, 0, 5, 5, 0])
    np.testing.assert_allclose(t[0]["blen"], [0, 100.0, 1e-8, 0.25, 0.5, 0.125])
    np.testing.assert_array_equal(t[1]["parent"], [-1, 0, 0, 5, 5, 0])
    np.testing.assert_allclose(t[1]["blen"], [0, 1, 2, 3, 4, 5])
    assert t[1]["score"] == -12.5 and np.isnan(t[0]["score"])
    for body in ("(1,2,(3,9));", "(1,2,(3,3));", "(1,2,3);", "(1,2,(3,4)"):
        g = tmp_path / "bad.tre"
        g.write_text("#NEXUS\nbegin trees;\ntree x = %s\nend;\n" % body)
        with pytest.raises(E.GleError):
            E.read_start_trees(str(g), n_taxa=4)
