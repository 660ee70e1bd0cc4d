"""Extract the reference's own test data into a committed fixture.

Reads (as data, never executed) garli/trunk/tests/data/ from the reference
checkout: the 11-taxon alignments z.11x2178.nex / z.11x2178.AA.nex, the
starting trees *.start with their [!GarliScore] and [!GarliModel] comments,
and tests/data/expected.scr; and, from garli/trunk/src/model.cpp, the numbers
of the published WAG amino-acid exchangeability matrix (Whelan & Goldman
2001) that ratematrix = wag multiplies into Q (Model::MultiplyByWAGAAMatrix),
as 190 values in upper-triangle row order over ACDEFGHIKLMNPQRSTVWY.  Writes
tests/golden/garli_testdata.json and copies the alignment and starting-tree
files the ingestion tests read into tests/golden/data/.

    python tests/golden/make_fixtures.py [/path/to/garli/trunk/tests/data]
"""
import json
import os
import re
import sys

DEFAULT = "/root/reference/garli/trunk/tests/data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "garli_testdata.json")
DATA_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DATA_FILES = ["z.11x2178.nex", "z.11x2178.wtset.nex", "z.11x2178.AA.nex", "z.11x30.phy", "z.11x30.AA.fas",
              "z.11x30.wackyNames.nex", "z.11x30.stop.nex", "moore.matK90-120.nex",
              # starting-tree files the tree-file reader reads
              "a.G3.start", "a.G4.start", "a.start", "c.M3x2.start", "c.start", "n.G4.start", "n.G5.start",
              "n.start", "p.3diff.start", "L.start", "moore.start"]


def read_matrix(path):
    txt = open(path).read()
    m = re.search(r"\bmatrix\b(.*?);", txt, re.S | re.I)
    names, seqs = [], {}
    for line in m.group(1).splitlines():
        line = line.strip()
        if not line or line.startswith("["):
            continue
        parts = line.split()
        name, seq = parts[0], "".join(parts[1:])
        if name not in seqs:
            names.append(name)
            seqs[name] = ""
        seqs[name] += seq
    return names, [seqs[n] for n in names]


def read_start(path):
    txt = open(path).read()
    tr = re.search(r"translate(.*?);", txt, re.S)
    translate = {}
    for item in tr.group(1).split(","):
        k, v = item.split()
        translate[int(k)] = v
    t = re.search(r"tree\s+\S+\s*=\s*(.*?;)", txt, re.S).group(1)
    score = re.search(r"GarliScore\s+(-?[0-9.]+)", t)
    model = re.search(r"GarliModel\s+([^\]]*)\]", t)
    newick = re.sub(r"\[.*?\]", "", t).strip()
    # a "begin garli;" block after the trees holds parameters the GarliModel
    # comment does not print (e.g. the 190 fixed amino-acid exchangeabilities)
    garli = {}
    gb = re.search(r"begin\s+garli\s*;(.*?)end\s*;", txt, re.S | re.I)
    if gb:
        for stmt in re.split(r"[;\n]", gb.group(1)):
            toks = stmt.split()
            if len(toks) > 1 and re.match(r"^[a-z]$", toks[0]):
                garli[toks[0]] = [float(x) for x in toks[1:]]
    return {"translate": translate, "newick": newick,
            "score": float(score.group(1)) if score else None,
            "model": model.group(1).strip() if model else None,
            "garli_block": garli}


def read_wag(model_cpp):
    """The WAG exchangeabilities as data: every `qmatOffset[i][j] *= v;` of
    Model::MultiplyByWAGAAMatrix, checked to fill a symmetric 20 x 20 matrix."""
    txt = open(model_cpp).read()
    body = txt[txt.index("void Model::MultiplyByWAGAAMatrix"):]
    body = body[:body.index("\n}")]
    r = [[0.0] * 20 for _ in range(20)]
    for i, j, v in re.findall(r"qmatOffset\[(\d+)\]\[(\d+)\]\s*\*=\s*([0-9.eE+-]+)\s*;", body):
        r[int(i)][int(j)] = float(v)
    upper = [r[i][j] for i in range(20) for j in range(i + 1, 20)]
    assert all(r[i][j] == r[j][i] for i in range(20) for j in range(20)) and all(x > 0 for x in upper)
    return upper


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else DEFAULT
    names, nuc = read_matrix(os.path.join(src, "z.11x2178.nex"))
    anames, aa = read_matrix(os.path.join(src, "z.11x2178.AA.nex"))
    starts = {}
    for f in sorted(os.listdir(src)):
        if f.endswith(".start"):
            try:
                starts[f[:-6]] = read_start(os.path.join(src, f))
            except (AttributeError, ValueError):
                pass
    expected = {}
    for line in open(os.path.join(src, "expected.scr")):
        p = line.split()
        if len(p) == 2:
            expected[p[0]] = float(p[1])
    wag = read_wag(os.path.join(src, "..", "..", "src", "model.cpp"))
    out = {"source": "garli/trunk/tests/data (reference test fixtures)",
           "taxa": names, "nuc_matrix": nuc, "aa_taxa": anames, "aa_matrix": aa,
           "starts": starts, "expected_scr": expected, "wag_rel190": wag}
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=1)
    # alignment files the ingestion tests read, copied as data
    os.makedirs(DATA_OUT, exist_ok=True)
    for f in DATA_FILES:
        with open(os.path.join(src, f), "rb") as fi, open(os.path.join(DATA_OUT, f), "wb") as fo:
            fo.write(fi.read())
    print("wrote", OUT, len(nuc), "taxa x", len(nuc[0]), "nt;", len(starts), "start trees")


if __name__ == "__main__":
    main()
