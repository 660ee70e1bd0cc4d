"""Seeded synthetic workloads for the GARLI likelihood engine.

Trees are Yule trees in GARLI's node numbering (0 = root with three children,
1..N tips, N+1..2N-3 internal nodes; parent[0] = -1).  Alignments are simulated
down the tree under the same model that is later scored, compressed to site
patterns with counts (PatternManager semantics, via libgle's host helper), and
carry tip data in the reference encodings (nucleotide 4-bit masks, state
indices with S = missing otherwise).

Workloads (sizes are parameters; the defaults are the benchmark sizes):
  * gtr_gamma:  nucleotide GTR+G4 (+I), IUPAC ambiguity codes on a fraction of cells
  * aa_gamma:   20-state amino acid with random symmetric exchangeabilities, G4
  * codon_m3:   GY94 codon model, omega classes {0.1, 1, 3}, kappa 2.5
"""
from dataclasses import dataclass, field

import numpy as np

from . import engine as E

IUPAC_AMBIG = "MRWSYKVHDBN"


@dataclass
class Dataset:
    name: str
    n_tips: int
    n_states: int
    parent: np.ndarray
    blen: np.ndarray
    tips: np.ndarray            # [N][P] uint8 codes
    counts: np.ndarray
    const_states: np.ndarray
    last_const: int
    pi: np.ndarray
    rate_probs: np.ndarray
    pinv: float
    pinv_in_model: bool
    eig: list                   # [(V, Vinv, lam)] per eigensystem (lam scaled)
    eig_of_rate: np.ndarray
    rate_mult: np.ndarray
    meta: dict = field(default_factory=dict)

    @property
    def n_patterns(self):
        return self.tips.shape[1]

    @property
    def n_rates(self):
        return len(self.rate_probs)

    def make_instance(self, precision=E.PREC_FP64, n_cla_slots=None, n_pmat_slots=None):
        inst = E.Instance(self.n_tips, self.n_patterns, self.n_states, self.n_rates, n_eigen=len(self.eig),
                          n_cla_slots=n_cla_slots, n_pmat_slots=n_pmat_slots, precision=precision)
        for t in range(self.n_tips):
            inst.set_tip_states(t, self.tips[t])
        inst.set_patterns(self.counts, self.last_const, self.const_states)
        inst.set_model(self.pi, self.rate_probs, self.pinv, self.pinv_in_model)
        for e, (V, W, lam) in enumerate(self.eig):
            inst.set_eigen(e, V, W, lam)
        inst.set_category_rates(self.eig_of_rate, self.rate_mult)
        return inst


def yule_tree(n_tips, rng, mean_blen):
    """Random Yule topology, unrooted, with i.i.d. Exp(mean) branch lengths."""
    assert n_tips >= 3
    children = {0: [1, 2]}
    leaves = [1, 2]
    nxt = 3
    while len(leaves) < n_tips:
        i = int(rng.integers(len(leaves)))
        leaf = leaves[i]
        children[leaf] = [nxt, nxt + 1]
        leaves[i] = nxt
        leaves.append(nxt + 1)
        nxt += 2
    # unroot: drop node 0 and join its two children
    a, b = children.pop(0)
    if a not in children:
        a, b = b, a
    children[a] = children[a] + [b]
    new_root = a
    # renumber: root 0, tips 1..N (random order), internals N+1.. in BFS order
    order = list(rng.permutation(leaves))
    ids = {new_root: 0}
    for k, leaf in enumerate(order):
        ids[leaf] = k + 1
    nxt_int = n_tips + 1
    queue = [new_root]
    while queue:
        x = queue.pop(0)
        for c in children.get(x, []):
            if c not in ids:
                ids[c] = nxt_int
                nxt_int += 1
            queue.append(c)
    nn = 2 * n_tips - 2
    parent = np.full(nn, -1, dtype=np.int32)
    for x, cs in children.items():
        for c in cs:
            parent[ids[c]] = ids[x]
    blen = rng.exponential(mean_blen, size=nn)
    blen = np.maximum(blen, 1e-8)
    blen[0] = 0.0
    return parent, blen


def _pmats(eig, eig_of_rate, rate_mult, t):
    out = []
    for r in range(len(rate_mult)):
        V, W, lam = eig[eig_of_rate[r]]
        P = (V * np.exp(lam * rate_mult[r] * t)) @ W
        P = np.maximum(P, 0.0)
        out.append(P / P.sum(axis=1, keepdims=True))
    return out


def _simulate_states(parent, blen, pi, eig, eig_of_rate, rate_mult, rate_probs, pinv, n_sites, rng):
    """States [nodes][sites] evolved from the root; invariant sites share one state."""
    nn = len(parent)
    S = len(pi)
    kids = [[] for _ in range(nn)]
    for n in range(1, nn):
        kids[parent[n]].append(n)
    w = np.asarray(rate_probs, dtype=float)
    probs = np.concatenate([[pinv], w * (1.0 - pinv) / w.sum()]) if pinv > 0 else w / w.sum()
    cls = rng.choice(len(probs), size=n_sites, p=probs)
    inv = cls == 0 if pinv > 0 else np.zeros(n_sites, bool)
    rate = (cls - 1 if pinv > 0 else cls).clip(0)
    states = np.zeros((nn, n_sites), dtype=np.int32)
    states[0] = rng.choice(S, size=n_sites, p=pi)
    stack = [0]
    while stack:
        a = stack.pop()
        for c in kids[a]:
            Ps = _pmats(eig, eig_of_rate, rate_mult, blen[c])
            cum = np.stack([np.cumsum(P, axis=1) for P in Ps])  # [R][S][S]
            rows = cum[rate, states[a]]                          # [sites][S]
            u = rng.random(n_sites)[:, None]
            new = (u > rows).sum(axis=1).clip(0, S - 1)
            states[c] = np.where(inv, states[a], new)
            stack.append(c)
    return states


def _first_k_unique(cols, k):
    """Index of the first site at which the k-th distinct column appears."""
    v = np.ascontiguousarray(cols).view(np.dtype((np.void, cols.shape[1])))[:, 0]
    _, first = np.unique(v, return_index=True)
    if len(first) < k:
        return None
    return int(np.sort(first)[k - 1]) + 1


def _finish(name, parent, blen, tip_codes, S, pi, rate_probs, pinv, pinv_in_model, eig, eig_of_rate, rate_mult,
            meta):
    ms = 4 if S == 4 else S
    pats, counts, cs, lc, _ = E.compress_patterns(tip_codes, ms)
    return Dataset(name, tip_codes.shape[0], S, parent, blen, pats, counts, cs, lc, np.asarray(pi, float),
                   np.asarray(rate_probs, float), float(pinv), bool(pinv_in_model), eig,
                   np.asarray(eig_of_rate, np.int32), np.asarray(rate_mult, float), meta)


def _simulate_until(parent, blen, pi, eig, eor, mult, w, pinv, n_tips, rng, n_sites, n_patterns, code_fn):
    if n_patterns is None:
        st = _simulate_states(parent, blen, pi, eig, eor, mult, w, pinv, n_sites, rng)
        return code_fn(st[1:n_tips + 1])
    chunks, have = [], None
    batch = int(n_patterns * 1.25) + 64
    for _ in range(64):
        st = _simulate_states(parent, blen, pi, eig, eor, mult, w, pinv, batch, rng)
        chunks.append(code_fn(st[1:n_tips + 1]))
        have = np.concatenate(chunks, axis=1)
        cut = _first_k_unique(have.T, n_patterns)
        if cut is not None:
            return have[:, :cut]
    raise RuntimeError("could not reach %d unique patterns" % n_patterns)


def gtr_gamma(n_tips=24, n_sites=2000, n_patterns=None, mean_blen=0.1, alpha=0.5, pinv=0.0, ncat=4,
              ambig_frac=0.05, seed=1, name="gtr_gamma"):
    rng = np.random.default_rng(seed)
    parent, blen = yule_tree(n_tips, rng, mean_blen)
    pi = rng.dirichlet([10.0] * 4)
    r = rng.lognormal(0.0, 0.5, size=6)
    rel = r[:5] / r[5]
    if alpha is None:
        alpha = float(rng.uniform(0.2, 1.5))
    V, W, lam = E.model_gtr(pi, rel)
    rates = E.discrete_gamma(ncat, alpha) if ncat > 1 else np.ones(1)
    pin = pinv > 0
    w = np.full(ncat, (1.0 - pinv) / ncat)
    mult = rates / (1.0 - pinv) if pin else rates
    eig = [(V, W, lam)]
    eor = np.zeros(ncat, np.int32)

    def code(states):
        masks = (1 << states).astype(np.uint8)
        if ambig_frac > 0:
            hit = rng.random(masks.shape) < ambig_frac
            amb = np.array([E.nuc_char_to_datum(c) for c in IUPAC_AMBIG], dtype=np.uint8)
            masks = np.where(hit, amb[rng.integers(len(amb), size=masks.shape)], masks)
        return masks

    tips = _simulate_until(parent, blen, pi, eig, eor, mult, w, pinv, n_tips, rng, n_sites, n_patterns, code)
    meta = dict(alpha=alpha, rel=rel, rates=rates, seed=seed, n_sites=tips.shape[1])
    return _finish(name, parent, blen, tips, 4, pi, w, pinv, pin, eig, eor, mult, meta)


def aa_gamma(n_tips=64, n_sites=None, n_patterns=20000, mean_blen=0.1, alpha=0.5, ncat=4, seed=3,
             name="aa_gamma"):
    rng = np.random.default_rng(seed)
    parent, blen = yule_tree(n_tips, rng, mean_blen)
    pi = rng.dirichlet([1.0] * 20)
    pi = np.maximum(pi, 1e-4)
    pi /= pi.sum()
    rel = rng.lognormal(0.0, 1.0, size=190)
    V, W, lam = E.model_aa(pi, rel)
    rates = E.discrete_gamma(ncat, alpha) if ncat > 1 else np.ones(1)
    w = np.full(ncat, 1.0 / ncat)
    eig = [(V, W, lam)]
    eor = np.zeros(ncat, np.int32)
    tips = _simulate_until(parent, blen, pi, eig, eor, rates, w, 0.0, n_tips, rng, n_sites, n_patterns,
                           lambda s: s.astype(np.uint8))
    meta = dict(alpha=alpha, rel=rel, rates=rates, seed=seed)
    return _finish(name, parent, blen, tips, 20, pi, w, 0.0, False, eig, eor, rates, meta)


def codon_m3(n_tips=32, n_sites=None, n_patterns=5000, mean_blen=0.1, kappa=2.5, omegas=(0.1, 1.0, 3.0),
             seed=4, name="codon_m3"):
    rng = np.random.default_rng(seed)
    parent, blen = yule_tree(n_tips, rng, mean_blen)
    pi = rng.dirichlet([2.0] * 61)
    pi = np.maximum(pi, 1e-5)
    pi /= pi.sum()
    op = np.full(len(omegas), 1.0 / len(omegas))
    rel6 = np.array([1.0, kappa, 1.0, 1.0, kappa, 1.0])
    V, W, lam = E.model_codon(pi, rel6, np.asarray(omegas, float), op)
    eig = [(V[k], W[k], lam[k]) for k in range(len(omegas))]
    eor = np.arange(len(omegas), dtype=np.int32)
    mult = np.ones(len(omegas))
    tips = _simulate_until(parent, blen, pi, eig, eor, mult, op, 0.0, n_tips, rng, n_sites, n_patterns,
                           lambda s: s.astype(np.uint8))
    meta = dict(kappa=kappa, rel6=rel6, omegas=omegas, seed=seed)
    return _finish(name, parent, blen, tips, 61, pi, op, 0.0, False, eig, eor, mult, meta)


def ambig_streams(ds):
    """Reference-format nucleotide ambiguity streams, one per taxon."""
    return [E.make_ambig_string(ds.tips[t]) for t in range(ds.n_tips)]


# ------------------------------------------------------------ topology moves
def _adjacency(parent, blen):
    nn = len(parent)
    adj = {n: {} for n in range(nn)}
    for n in range(1, nn):
        a = int(parent[n])
        adj[n][a] = blen[n]
        adj[a][n] = blen[n]
    return adj


def _reroot(adj, nn):
    """Parent array rooted at node 0 (which must have degree 3)."""
    parent = np.full(nn, -1, np.int32)
    blen = np.zeros(nn)
    seen = {0}
    stack = [0]
    while stack:
        a = stack.pop()
        for c, l in adj[a].items():
            if c not in seen:
                seen.add(c)
                parent[c] = a
                blen[c] = l
                stack.append(c)
    assert len(seen) == nn
    return parent, blen


def nni(parent, blen, rng, n_tips):
    """Nearest-neighbour interchange across a random internal edge
    (the move of Tree::NNIMutate, tree.cpp:6970)."""
    adj = _adjacency(parent, blen)
    internal = [n for n in adj if not (1 <= n <= n_tips)]
    edges = [(u, v) for u in internal for v in adj[u] if v > u and not (1 <= v <= n_tips)]
    u, v = edges[int(rng.integers(len(edges)))]
    b = [x for x in adj[u] if x != v][int(rng.integers(2))]
    c = [x for x in adj[v] if x != u][int(rng.integers(2))]
    lb, lc = adj[u].pop(b), adj[v].pop(c)
    del adj[b][u], adj[c][v]
    adj[u][c] = adj[c][u] = lc
    adj[v][b] = adj[b][v] = lb
    return _reroot(adj, len(parent))


def spr(parent, blen, rng, n_tips):
    """Subtree prune and regraft (Tree::SPRMutate, tree.cpp:2571-2708): detach the
    subtree below a random edge at its internal attachment node p, join p's other
    two neighbours, and reinsert p at the midpoint of a random edge outside the
    subtree."""
    adj = _adjacency(parent, blen)
    nn = len(parent)
    for _ in range(100):
        p = int(rng.choice([n for n in adj if not (1 <= n <= n_tips)]))
        s = int(rng.choice(list(adj[p])))
        x, y = [q for q in adj[p] if q != s]
        # nodes of the pruned subtree (s's side of edge (p, s))
        sub, stack = {s}, [s]
        while stack:
            a = stack.pop()
            for q in adj[a]:
                if q != p and q not in sub:
                    sub.add(q)
                    stack.append(q)
        cand = [(a, q) for a in adj if a not in sub and a != p for q in adj[a]
                if q not in sub and q != p and a < q and {a, q} != {x, y}]
        if not cand:
            continue
        lx, ly = adj[p].pop(x), adj[p].pop(y)
        del adj[x][p], adj[y][p]
        adj[x][y] = adj[y][x] = lx + ly
        cand = [(a, q) for (a, q) in cand if q in adj[a]]
        if not cand:
            return _reroot(_adjacency(parent, blen), nn)
        a, q = cand[int(rng.integers(len(cand)))]
        l = adj[a].pop(q)
        del adj[q][a]
        adj[a][p] = adj[p][a] = l / 2
        adj[q][p] = adj[p][q] = l / 2
        return _reroot(adj, nn)
    return parent.copy(), blen.copy()


def perturb(parent, blen, rng, n_tips, n_moves):
    """1..n random NNI/SPR moves plus one branch-length change: a GA offspring."""
    for _ in range(n_moves):
        parent, blen = (nni if rng.random() < 0.5 else spr)(parent, blen, rng, n_tips)
    blen = blen.copy()
    k = int(rng.integers(1, len(blen)))
    blen[k] = max(1e-8, blen[k] * float(rng.uniform(0.5, 1.5)))
    return parent, blen
