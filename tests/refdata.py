"""Datasets built from the reference's own test fixtures (tests/golden/garli_testdata.json).

The alignments and start trees are garli/trunk/tests/data files; the scores
they carry ([!GarliScore]) were computed by the reference program.
"""
import json
import os
import re

import numpy as np

from garli_amd import engine as E
from garli_amd import sim

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "garli_testdata.json")


def load():
    with open(FIXTURE) as fh:
        return json.load(fh)


def parse_newick(newick, ntax):
    """Unrooted newick with numeric tip labels 1..N -> (parent, blen) in the
    engine's numbering: 0 root (the outermost node, 3 children), 1..N tips,
    N+1.. internal in preorder."""
    s = newick.strip().rstrip(";")
    pos = 0
    nodes = []  # (label or None, children, blen)

    def parse():
        nonlocal pos
        kids = []
        label = None
        if s[pos] == "(":
            pos += 1
            while True:
                kids.append(parse())
                if s[pos] == ",":
                    pos += 1
                    continue
                assert s[pos] == ")"
                pos += 1
                break
        m = re.match(r"[^:,()]*", s[pos:])
        txt = m.group(0)
        pos += len(txt)
        if txt:
            label = int(txt)
        bl = 0.0
        if pos < len(s) and s[pos] == ":":
            m = re.match(r":([0-9.eE+-]+)", s[pos:])
            bl = float(m.group(1))
            pos += len(m.group(0))
        nodes.append((label, kids, bl))
        return len(nodes) - 1

    root = parse()
    nn = 2 * ntax - 2
    parent = np.full(nn, -1, np.int32)
    blen = np.zeros(nn)
    ids = {}
    nxt = [ntax + 1]

    def number(k, par):
        label, kids, bl = nodes[k]
        if par is None:
            me = 0
        elif label is not None:
            me = label
        else:
            me = nxt[0]
            nxt[0] += 1
        ids[k] = me
        if par is not None:
            parent[me] = par
            blen[me] = bl
        for c in kids:
            number(c, me)

    assert len(nodes[root][1]) == 3, "expected an unrooted tree with a trifurcating root"
    number(root, None)
    return parent, blen


def model_fields(model):
    toks = model.split()
    out, key = {}, None
    for t in toks:
        if re.match(r"^[a-z]$", t):
            key = t
            out[key] = []
        else:
            out[key].append(float(t))
    return out


def start_tree(start_name, taxa):
    """The first tree of the start file <start_name>.start (a copy under
    tests/golden/data) as the library's tree-file reader gives it: parent,
    blen, score and model, tips numbered by the order of taxa."""
    return E.read_start_trees(os.path.join(HERE, "golden", "data", start_name + ".start"), taxa)[0]


def nucleotide_dataset(start_name, ncat=None):
    d = load()
    st = start_tree(start_name, d["taxa"])
    mf = model_fields(st["model"])
    ntax = len(d["taxa"])
    codes = np.array([[E.nuc_char_to_datum(c) for c in seq] for seq in d["nuc_matrix"]], dtype=np.uint8)
    pats, counts, cs, lc, _ = E.compress_patterns(codes, 4)
    parent, blen = st["parent"], st["blen"]
    pi = np.array(mf["e"])
    r = np.array(mf["r"])
    rel = r[:5] / r[5]
    alpha = mf["a"][0] if "a" in mf else None
    pinv = mf["p"][0] if "p" in mf else 0.0
    if ncat is None:
        ncat = {"n.G4": 4, "n.G5": 5}.get(start_name, 1 if alpha is None else 4)
    V, W, lam = E.model_gtr(pi, rel)
    rates = E.discrete_gamma(ncat, alpha) if ncat > 1 else np.ones(1)
    pin = pinv > 0
    w = np.full(ncat, (1.0 - pinv) / ncat)
    mult = rates / (1.0 - pinv) if pin else rates
    return sim.Dataset(start_name, ntax, 4, parent, blen, pats, counts, cs, lc, pi, w, pinv, pin,
                       [(V, W, lam)], np.zeros(ncat, np.int32), mult,
                       dict(alpha=alpha if alpha else 1.0, rel=rel, rates=rates, score=st["score"],
                            raw_matrix=codes))


_TO_INDEX = {1: 0, 2: 1, 4: 2, 8: 3}


def codon_matrix(nuc_rows):
    """CodonData::FillCodonMatrixFromDNA (sequencedata.cpp:364-470): codon
    16*n1+4*n2+n3 from unambiguous positions, any gap/ambiguity -> missing (61);
    sense codons renumbered 0..60 in codon order (stops removed)."""
    _, aa_of = _genetic_code()
    sense = [c for c in range(64) if aa_of[c] != 20]
    idx61 = {c: k for k, c in enumerate(sense)}
    ntax = len(nuc_rows)
    ncod = len(nuc_rows[0]) // 3
    out = np.full((ntax, ncod), 61, np.uint8)
    for t, row in enumerate(nuc_rows):
        for k in range(ncod):
            m = [E.nuc_char_to_datum(ch) for ch in row[3 * k:3 * k + 3]]
            if all(x in _TO_INDEX for x in m):
                c = 16 * _TO_INDEX[m[0]] + 4 * _TO_INDEX[m[1]] + _TO_INDEX[m[2]]
                assert aa_of[c] != 20, "stop codon in test data"
                out[t, k] = idx61[c]
    return out


def _genetic_code():
    from oracle import oracle as O
    lk, aa61 = O.codon_lookup()
    tcag = "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG"
    m = [2, 1, 3, 0]
    order = "ACDEFGHIKLMNPQRSTVWY"
    aa_of = [20 if tcag[16 * m[c >> 4] + 4 * m[(c >> 2) & 3] + m[c & 3]] == "*"
             else order.index(tcag[16 * m[c >> 4] + 4 * m[(c >> 2) & 3] + m[c & 3]]) for c in range(64)]
    return lk, aa_of


def empirical_codon_freqs(codons):
    """CodonData::CalcEmpiricalFreqs CODON_TABLE branch (sequencedata.cpp:228-275)."""
    f = np.zeros(61)
    for v in codons.reshape(-1):
        if v != 61:
            f[v] += 1.0
    total = f.sum()
    if (f == 0).any():
        f += 1.0
        total += 61.0
    return f / total


def codon_dataset(start_name):
    d = load()
    st = start_tree(start_name, d["taxa"])
    mf = model_fields(st["model"])
    ntax = len(d["taxa"])
    codons = codon_matrix(d["nuc_matrix"])
    pats, counts, cs, lc, _ = E.compress_patterns(codons, 61)
    parent, blen = st["parent"], st["blen"]
    pi = empirical_codon_freqs(codons)
    rel6 = np.array(mf["r"])
    o = mf["o"]
    omegas, probs = np.array(o[0::2]), np.array(o[1::2])
    V, W, lam = E.model_codon(pi, rel6, omegas, probs)
    nw = len(omegas)
    eig = [(V[k], W[k], lam[k]) for k in range(nw)]
    return sim.Dataset(start_name, ntax, 61, parent, blen, pats, counts, cs, lc, pi, probs, 0.0, False, eig,
                       np.arange(nw, dtype=np.int32), np.ones(nw),
                       dict(rel6=rel6, omegas=omegas, score=st["score"], printed_freqs=np.array(mf["e"])))


def aa_dataset(start_name):
    """trunk/tests/scoring/a.G3.conf, a.G4.conf and a.conf: amino acids read
    from z.11x2178.AA.nex (a.G3) or translated from the codons of
    z.11x2178.nex (datatype = codon-aminoacid: a.G4, a), by the engine's own
    ingestion; statefrequencies = empirical from those patterns; WAG (a.G3,
    a) or the fixed matrix of a.G4.start's garli block; 3, 4 or 1 gamma
    categories at the start file's shape."""
    d = load()
    st = start_tree(start_name, d["taxa"])
    mf = model_fields(st["model"])
    data = os.path.join(HERE, "golden", "data")
    if start_name == "a.G3":
        aln = E.Alignment(os.path.join(data, "z.11x2178.AA.nex"))
        pats, counts, cs, lc, site = aln.aa_patterns()
    else:
        aln = E.Alignment(os.path.join(data, "z.11x2178.nex"))
        pats, counts, cs, lc, site = aln.translated_patterns()
    assert aln.names == d["taxa"]
    rel = np.array(d["starts"]["a.G4"]["garli_block"]["r"] if start_name == "a.G4" else d["wag_rel190"])
    pi = E.empirical_freqs(pats, counts, 20)
    ncat = {"a.G3": 3, "a.G4": 4, "a": 1}[start_name]
    alpha = mf["a"][0] if "a" in mf else 1.0
    parent, blen = st["parent"], st["blen"]
    V, W, lam = E.model_aa(pi, rel)
    rates = E.discrete_gamma(ncat, alpha) if ncat > 1 else np.ones(1)
    return sim.Dataset(start_name, len(d["taxa"]), 20, parent, blen, pats, counts, cs, lc, pi,
                       np.full(ncat, 1.0 / ncat), 0.0, False, [(V, W, lam)], np.zeros(ncat, np.int32), rates,
                       dict(alpha=alpha, rel=rel, rates=rates, score=st["score"], site=site))
