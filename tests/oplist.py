"""Op lists in the shape GARLI's BEAGLE seam hands to its back end.

CalculationManager::DetermineRequiredOperations / AccumulateOpsOnPath
(trunk/src/calculationmanager.cpp:522-670) turn dirty CLA holders into
ClaOperations grouped by dependency level (BlockingOperationsSet,
calculationmanager.h:118-123), one TransMatOperation per branch and a
ScoringOperation per edge; PerformClaOperationBatch
(subsetcalculationmanager.cpp:595-682) sends each level to the library.
This module builds the same lists for a whole tree (every directional CLA),
addressing CLAs by slot and tips as -(tip + 1), the way the C ABI takes them.
"""
import numpy as np


def neighbours(parent):
    nn = len(parent)
    nb = {n: [] for n in range(nn)}
    for n in range(1, nn):
        nb[n].append(int(parent[n]))
    for n in range(1, nn):
        nb[int(parent[n])].append(n)
    return nb


def full_sweep(parent, n_tips, rescale_every=16, targets=None):
    """Every directional CLA X(a|b) of the tree, or with targets = [(a, b), ..]
    only the CLAs those need (AccumulateOpsOnPath's closure).

    Returns (ops, level_start, slot_of, pmat_of):
      ops          list of (dest, child1, pmat1, child2, pmat2, rescale) by level
      level_start  level ranges (len = levels + 1)
      slot_of      {(a, b): CLA slot}
      pmat_of      {node n: transition-matrix slot of the edge (n, parent[n])}
    rescale follows the reference's rescaleRank rule: II 2 + rL + rR,
    IT r + 2, TT 2 (tree.cpp:7840, 8534, 8079), reset when it reaches
    rescale_every."""
    nn = len(parent)
    nb = neighbours(parent)

    def tip(n):
        return 1 <= n <= n_tips

    def edge(a, b):
        return a if parent[a] == b else b

    pmat_of = {n: n - 1 for n in range(1, nn)}
    slot_of, level, rank = {}, {}, {}
    if targets is None:
        for a in range(nn):
            if tip(a):
                continue
            for b in nb[a]:
                slot_of[(a, b)] = len(slot_of)
    else:
        stack = [t for t in targets if not tip(t[0])]
        while stack:
            a, b = stack.pop()
            if (a, b) in slot_of:
                continue
            slot_of[(a, b)] = len(slot_of)
            stack.extend((c, a) for c in nb[a] if c != b and not tip(c))

    def lev(a, b):
        if tip(a):
            return -1
        if (a, b) not in level:
            level[(a, b)] = 1 + max(lev(c, a) for c in nb[a] if c != b)
        return level[(a, b)]

    def rk(a, b):
        if tip(a):
            return 0
        if (a, b) not in rank:
            r = 2 + sum(rk(c, a) for c in nb[a] if c != b)
            rank[(a, b)] = 0 if r >= rescale_every else r
        return rank[(a, b)]

    entries = []
    for (a, b), s in slot_of.items():
        x, y = [c for c in nb[a] if c != b]
        ref = [(-c if tip(c) else slot_of[(c, a)]) for c in (x, y)]
        r = 2 + rk(x, a) + rk(y, a)
        entries.append((lev(a, b), (s, ref[0], pmat_of[edge(a, x)], ref[1], pmat_of[edge(a, y)],
                                    1 if r >= rescale_every else 0)))
    entries.sort(key=lambda e: e[0])
    ops = [e[1] for e in entries]
    levels = [e[0] for e in entries]
    level_start = [0] + [i for i in range(1, len(levels)) if levels[i] != levels[i - 1]] + [len(ops)]
    return ops, np.array(level_start, np.int32), slot_of, pmat_of


def edge_ops(parent, n_tips, slot_of, pmat_of, derivs=1):
    """(partial, child, pmat, derivs) for the edge above every node n >= 1,
    partial side = parent[n]: the op of CalcDerivativesRateHet(anc, n)."""
    out = []
    for n in range(1, len(parent)):
        a = int(parent[n])
        child = -n if 1 <= n <= n_tips else slot_of[(n, a)]
        out.append((slot_of[(a, n)], child, pmat_of[n], derivs))
    return out
