// gle_tree.cpp -- GARLI's tree-scoring surface on top of the device engine.
//
// A Tree owns the directional CLAs of an unrooted binary tree: for an internal
// node a and a neighbour b, X(a|b) is the conditional likelihood of the part of
// the tree on a's side of edge (a,b).  GARLI names these DOWN (b = anc),
// UPLEFT (b = left) and UPRIGHT (b = right) (treenode.h:45-47); the root's DOWN
// excludes its middle child (tree.cpp:3643-3647).  CLAs live in device slots
// shared between trees by reference count, the ClaManager scheme
// (clamanager.h:96-441): copying a tree shares every clean CLA, and a CLA is
// recomputed into a fresh slot when it turns dirty (copy-on-write).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <unordered_map>
#include <memory>
#include <mutex>
#include <vector>

#include "gle_internal.h"
#include "gle_newton.h"

namespace gle {

// The slot pool lives in the instance (pool_for, guarded by pool_mutex).
static std::recursive_mutex g_tree_rmu;

struct Tree {
  int inst_h = -1;  // the instance handle; trees of a destroyed instance are orphaned
  Instance *I = nullptr;
  ClaPool *pool = nullptr;
  int N = 0, nn = 0;
  std::vector<int> parent;
  std::vector<double> blen;
  std::vector<std::array<int, 3>> nb;  // anc first (if any), then children ascending
  std::vector<int> nnb;
  std::vector<std::array<int, 3>> slot;  // CLA slot of X(a|nb[a][k]), -1 = dirty
  std::vector<int> pm;                   // pmat slot per node's parent edge
  std::vector<double> pm_len;            // length the slot holds (NaN = stale)
  std::vector<char> pm_has_derivs;
  int rescale_every = 16;
  int schedule = 0;
  unsigned long long gen = 0;  // instance generation the cached CLAs and P(t) belong to
  double lnl = 0.0;
  // The last score is current (nothing changed since): its site lnL is
  // computed when asked for, on the same edge, instead of copied back with
  // every score.
  bool site_ready = false;
  int score_pa = 0, score_pb = 0;  // the edge the last score was taken on
  std::vector<double> site;
  long long cla_updates = 0;  // CLA updates this tree has computed (all calls)
  // scratch
  std::vector<int> level;
  std::vector<DevClaOp> ops;
  std::vector<int> op_level;
};

static std::vector<std::unique_ptr<Tree>> g_trees;

static Tree *get_tree(int h) {
  if (h < 0 || h >= int(g_trees.size())) return nullptr;
  return g_trees[h].get();
}

static bool is_tip(const Tree &t, int n) { return n >= 1 && n <= t.N; }
static int nb_index(const Tree &t, int a, int b) {
  for (int k = 0; k < t.nnb[a]; k++)
    if (t.nb[a][k] == b) return k;
  return -1;
}
static int edge_node(const Tree &t, int a, int b) { return t.parent[a] == b ? a : b; }

static void release(Tree &t, int a, int k) {
  int &s = t.slot[a][k];
  if (s < 0) return;
  if (--t.pool->refs[s] == 0) t.pool->free_slot(s);
  s = -1;
}
static void dirty_all(Tree &t) {
  t.site_ready = false;
  for (int a = 0; a < t.nn; a++)
    if (!is_tip(t, a))
      for (int k = 0; k < t.nnb[a]; k++) release(t, a, k);
}
// After a device error reported at a wait: nothing queued before it counts as
// computed, so every CLA and P(t) slot of the tree is recomputed next time.
static int forget_device_state(Tree &t, int rc) {
  dirty_all(t);
  std::fill(t.pm_len.begin(), t.pm_len.end(), std::nan(""));
  return rc;
}
// SweepDirtynessOverTree: every X(x|y) whose side holds edge (n, parent(n))
static void dirty_from(Tree &t, int from, int came) {
  for (int j = 0; j < t.nnb[from]; j++) {
    const int nbr = t.nb[from][j];
    if (nbr == came) continue;
    if (!is_tip(t, from)) release(t, from, j);
    dirty_from(t, nbr, from);
  }
}
static void dirty_edge(Tree &t, int n) {
  t.site_ready = false;
  const int a = t.parent[n];
  dirty_from(t, n, a);
  dirty_from(t, a, n);
}

// Topology of a parent array: neighbours (anc first, then children
// ascending) and degrees, checked to be an unrooted binary tree on the
// instance's tips with node 0 as the root.
static int topology(int N, const int *parent, const double *blen, std::vector<std::array<int, 3>> &nb,
                    std::vector<int> &nnb) {
  const int nn = 2 * N - 2;
  nb.assign(nn, {-1, -1, -1});
  nnb.assign(nn, 0);
  if (parent[0] != -1) return set_error(GLE_ERR_ARG, "node 0 must be the root (parent -1)");
  for (int n = 1; n < nn; n++) {
    const int a = parent[n];
    if (a < 0 || a >= nn || a == n || (a >= 1 && a <= N))
      return set_error(GLE_ERR_ARG, "node %d: bad parent %d", n, a);
    if (!(blen[n] >= 0.0)) return set_error(GLE_ERR_ARG, "node %d: bad branch length", n);
    nb[n][nnb[n]++] = a;
  }
  for (int n = 1; n < nn; n++) {
    const int a = parent[n];
    if (nnb[a] >= 3) return set_error(GLE_ERR_ARG, "node %d has too many children", a);
    nb[a][nnb[a]++] = n;
  }
  for (int n = 0; n < nn; n++) {
    const int want = (n >= 1 && n <= N) ? 1 : 3;
    if (nnb[n] != want) return set_error(GLE_ERR_ARG, "node %d has degree %d, expected %d", n, nnb[n], want);
  }
  // every node must reach the root (no cycles)
  for (int n = 1; n < nn; n++) {
    int x = n, steps = 0;
    while (x != 0 && steps <= nn) {
      x = parent[x];
      steps++;
    }
    if (x != 0) return set_error(GLE_ERR_ARG, "node %d does not reach the root", n);
  }
  return GLE_SUCCESS;
}

static int build_tree(Tree &t, int inst_h, Instance *I, const int *parent, const double *blen) {
  const int N = instance_tips(I), nn = 2 * N - 2;
  t.inst_h = inst_h;
  t.I = I;
  t.pool = pool_for(I);
  t.N = N;
  t.nn = nn;
  t.parent.assign(parent, parent + nn);
  t.blen.assign(blen, blen + nn);
  int rc = topology(N, parent, blen, t.nb, t.nnb);
  if (rc) return rc;
  t.slot.assign(nn, {-1, -1, -1});
  if (int(t.pool->pm_free.size()) < nn) return set_error(GLE_ERR_NOMEM, "out of transition-matrix slots");
  t.pm.assign(nn, -1);
  for (int n = 1; n < nn; n++) {
    t.pm[n] = t.pool->pm_free.back();
    t.pool->pm_free.pop_back();
  }
  t.pm_len.assign(nn, std::nan(""));
  t.pm_has_derivs.assign(nn, 0);
  t.rescale_every = instance_precision(I) == GLE_PREC_FP64 ? 16 : 1;
  t.gen = instance_generation(I);
  t.level.assign(size_t(nn) * 3, -1);
  return GLE_SUCCESS;
}

// ------------------------------------------------------ op-list generation

// Collect the ops that make X(a|b) clean (post-order), recording levels.
// Returns the level of X(a|b) (-1 for tips / already clean CLAs).
static int collect(Tree &t, int a, int b, std::vector<std::array<int, 3>> &want) {
  if (is_tip(t, a)) return -1;
  const int k = nb_index(t, a, b);
  if (t.slot[a][k] >= 0) return -1;
  int &lv = t.level[size_t(a) * 3 + k];
  if (lv >= 0) return lv;
  int lmax = -1;
  for (int j = 0; j < t.nnb[a]; j++) {
    const int c = t.nb[a][j];
    if (c == b) continue;
    lmax = std::max(lmax, collect(t, c, a, want));
  }
  lv = lmax + 1;
  want.push_back({a, k, lv});
  return lv;
}

// Pmat bookkeeping: make sure the given edges hold P (and P', P'' if derivs).
// wait = false: queued only (see engine_update_pmats); the slots count as
// current at once, and forget_device_state() takes that back if the wait
// that follows reports an error.
static int refresh_pmats(Tree &t, const std::vector<int> &edges, bool derivs, bool wait = true) {
  std::vector<int> idx, nodes;
  std::vector<double> len;
  for (int n : edges) {
    const bool ok = t.pm_len[n] == t.blen[n] && (!derivs || t.pm_has_derivs[n]);
    if (ok || std::find(nodes.begin(), nodes.end(), n) != nodes.end()) continue;
    nodes.push_back(n);
    idx.push_back(t.pm[n]);
    len.push_back(t.blen[n]);
  }
  if (idx.empty()) return GLE_SUCCESS;
  int rc = engine_update_pmats(t.I, idx.data(), len.data(), int(idx.size()), derivs ? 1 : 0, wait);
  if (rc) return rc;
  for (int n : nodes) {
    t.pm_len[n] = t.blen[n];
    t.pm_has_derivs[n] = derivs ? 1 : 0;
  }
  return GLE_SUCCESS;
}

// Compute the listed directional CLAs (already in dependency order with levels).
// On any failure the slots handed out here go back to the pool and the CLAs
// stay dirty, so a later call recomputes them; the levels collect() recorded
// are cleared on every path.  wait = false only queues the updates (the edge
// evaluation that follows on the same streams waits for them, once).
static int run_ops(Tree &t, std::vector<std::array<int, 3>> &want, bool wait = true) {
  if (want.empty()) return GLE_SUCCESS;
  std::vector<std::array<int, 2>> assigned;  // (a, k) whose slot this call set
  auto fail = [&](int rc) {
    for (auto &ak : assigned) release(t, ak[0], ak[1]);
    for (auto &w : want) t.level[size_t(w[0]) * 3 + w[1]] = -1;
    return rc;
  };
  const bool fused = t.schedule == 1 && engine_supports_fused(t.I);
  // Level schedule: sort by dependency level.  Fused schedule: keep collect()'s
  // depth-first post-order, so a CLA is consumed soon after it is written.
  if (!fused)
    std::stable_sort(want.begin(), want.end(), [](const auto &x, const auto &y) { return x[2] < y[2]; });
  std::vector<int> edges;
  for (auto &w : want) {
    const int a = w[0], k = w[1], b = t.nb[a][k];
    for (int j = 0; j < t.nnb[a]; j++)
      if (t.nb[a][j] != b) edges.push_back(edge_node(t, a, t.nb[a][j]));
  }
  int rc = refresh_pmats(t, edges, false);
  if (rc) return fail(rc);
  std::vector<DevClaOp> ops;
  std::vector<int> level_start;
  ops.reserve(want.size());
  int cur = -1;
  for (auto &w : want) {
    const int a = w[0], k = w[1], b = t.nb[a][k];
    int x = -1, y = -1;
    for (int j = 0; j < t.nnb[a]; j++) {
      const int c = t.nb[a][j];
      if (c == b) continue;
      if (x < 0) x = c;
      else y = c;
    }
    auto child_ref = [&](int c, int &rank) {
      if (is_tip(t, c)) {
        rank = 0;
        return -c;  // tip index c-1 encoded as -(c-1)-1
      }
      const int s = t.slot[c][nb_index(t, c, a)];
      rank = t.pool->rank[s];
      return s;
    };
    int rx = 0, ry = 0;
    DevClaOp op;
    op.c1 = child_ref(x, rx);
    op.c2 = child_ref(y, ry);
    op.pm1 = t.pm[edge_node(t, a, x)];
    op.pm2 = t.pm[edge_node(t, a, y)];
    // rescaleRank: II 2+rL+rR, IT r+2, TT 2 (tree.cpp:7840, 8534, 8079)
    int rank = 2 + rx + ry;
    op.flags = 0;
    if (rank >= t.rescale_every) {
      op.flags = kOpRescale;
      rank = 0;
    }
    if (!t.pool->has_free()) return fail(set_error(GLE_ERR_NOMEM, "out of CLA slots"));
    const int s = t.pool->take_slot();
    t.pool->refs[s] = 1;
    t.pool->rank[s] = rank;
    t.slot[a][k] = s;
    assigned.push_back({a, k});
    op.dest = s;
    if (w[2] != cur) {
      level_start.push_back(int(ops.size()));
      cur = w[2];
    }
    ops.push_back(op);
  }
  level_start.push_back(int(ops.size()));
  rc = engine_update_partials(t.I, ops.data(), int(ops.size()), level_start.data(), int(level_start.size()) - 1,
                              fused, wait);
  if (rc) return fail(rc);
  for (auto &w : want) t.level[size_t(w[0]) * 3 + w[1]] = -1;
  t.cla_updates += static_cast<long long>(ops.size());
  return GLE_SUCCESS;
}

static int child_code(Tree &t, int c, int a) {
  if (is_tip(t, c)) return -c;
  return t.slot[c][nb_index(t, c, a)];
}

// One attempt of an edge evaluation.  Returns GLE_RESCALE_RETRY when the
// reference would have thrown from RescaleRateHet.
// The CLA updates, the edge's P(t) and the edge are queued back to back and
// waited for once, with the rescale flags read in the same wait (a wait per
// step had left the device idle between them).
static int edge_attempt(Tree &t, int pa, int pb, bool derivs, double *res, double *site, double *site5) {
  int flags = 0, rc = engine_flags(t.I, nullptr, 1);  // reset, queued ahead of the updates
  if (rc) return rc;
  std::vector<std::array<int, 3>> want;  // collect() levels are cleared by run_ops on every path
  collect(t, pa, pb, want);
  collect(t, pb, pa, want);
  rc = run_ops(t, want, false);
  if (rc) return forget_device_state(t, rc);
  const int en = edge_node(t, pa, pb);
  rc = refresh_pmats(t, {en}, derivs, false);
  if (rc) return forget_device_state(t, rc);
  DevEdgeOp op;
  op.partial = t.slot[pa][nb_index(t, pa, pb)];
  op.child = child_code(t, pb, pa);
  op.pm = t.pm[en];
  op.derivs = derivs ? 1 : 0;
  rc = site5 ? engine_edge_sites(t.I, op, res, site5, &flags) : engine_edges(t.I, &op, 1, res, site, &flags);
  if (rc) return forget_device_state(t, rc);
  if (instance_precision(t.I) == GLE_PREC_FP64) {
    const bool thrown = ((flags & GLE_FLAG_REDUCE_RESCALE) && t.rescale_every > 2) ||
                        ((flags & GLE_FLAG_BAIL_OUT) && t.rescale_every <= 2);
    if (thrown) return GLE_RESCALE_RETRY;
  }
  return GLE_SUCCESS;
}

// Retry loop of Tree::Score / BranchLike (tree.cpp:4077-4090)
static int edge_eval(Tree &t, int pa, int pb, bool derivs, double *res, double *site, double *site5 = nullptr) {
  for (;;) {
    const int rc = edge_attempt(t, pa, pb, derivs, res, site, site5);
    if (rc != GLE_RESCALE_RETRY) return rc;
    dirty_all(t);
    t.rescale_every -= 2;
    if (t.rescale_every < 2)
      return set_error(GLE_ERR_RESCALE_FATAL, "Problem with rescaling during tree scoring.");
  }
}

// ConditionalLikelihoodRateHet(ROOT, root) partial/child choice (tree.cpp:3686-3768)
static void root_edge(Tree &t, int &pa, int &pb) {
  const int left = t.nb[0][0], middle = t.nb[0][1], right = t.nb[0][2];
  pa = 0;
  if (t.slot[0][0] >= 0) pb = left;
  else if (t.slot[0][2] >= 0) pb = right;
  else pb = middle;
}

static int tree_score(Tree &t, double *lnl) {
  int pa, pb;
  root_edge(t, pa, pb);
  double res[3];
  t.site.clear();
  t.site_ready = false;
  const int rc = edge_eval(t, pa, pb, false, res, nullptr);
  if (rc) return rc;
  t.site_ready = true;
  t.score_pa = pa;  // site likelihoods are taken on this same edge later
  t.score_pb = pb;
  t.lnl = res[0];
  *lnl = res[0];
  return GLE_SUCCESS;
}

static int tree_derivs(Tree &t, int nd, double *lnl, double *d1, double *d2) {
  if (nd <= 0 || nd >= t.nn) return set_error(GLE_ERR_ARG, "node %d has no parent edge", nd);
  double res[3];
  const int rc = edge_eval(t, t.parent[nd], nd, true, res, nullptr);
  if (rc) return rc;
  t.lnl = res[0];
  *lnl = res[0];
  *d1 = res[1];
  *d2 = res[2];
  return GLE_SUCCESS;
}

static void set_blen(Tree &t, int nd, double v) {
  if (t.blen[nd] == v) return;
  t.blen[nd] = v;
  dirty_edge(t, nd);
}

// NewtonRaphsonOptimizeBranchLength (optimization.cpp:1359-1972) on one branch.
static int newton(Tree &t, int nd, double prec, double *improve) {
  NewtonBranch nb;
  nb.start(t.blen[nd], prec);
  for (;;) {
    double lnL, d1, d2, next = 0.0;
    const int rc = tree_derivs(t, nd, &lnL, &d1, &d2);
    if (rc) return rc;
    const NewtonBranch::Action act = nb.step(lnL, d1, d2, t.blen[nd], &next);
    if (act != NewtonBranch::kFinish) set_blen(t, nd, next);
    if (act != NewtonBranch::kContinue) break;
  }
  *improve = nb.total;
  return GLE_SUCCESS;
}

static void preorder(const Tree &t, int n, std::vector<int> &out) {
  out.push_back(n);
  for (int j = 0; j < t.nnb[n]; j++)
    if (t.nb[n][j] != t.parent[n]) preorder(t, t.nb[n][j], out);
}

static int optimize_all(Tree &t, double prec, double *improve, double *lnl) {
  std::vector<int> order;
  for (int j = 0; j < t.nnb[0]; j++) preorder(t, t.nb[0][j], order);
  double total = 0.0;
  for (int n : order) {
    double imp = 0.0;
    const int rc = newton(t, n, prec, &imp);
    if (rc) return rc;
    total += imp;
  }
  *improve = total;
  dirty_all(t);
  return tree_score(t, lnl);
}

// Every directional CLA, then (lnL, d1, d2) on every edge in one batched launch.
static int sweep(Tree &t, double *out) {
  for (int attempt = 0;; attempt++) {
    dirty_all(t);
    int flags = 0, rc = engine_flags(t.I, nullptr, 1);  // reset, queued ahead of the updates
    if (rc) return rc;
    std::vector<std::array<int, 3>> want;
    for (int a = 0; a < t.nn; a++)
      if (!is_tip(t, a))
        for (int k = 0; k < t.nnb[a]; k++) collect(t, a, t.nb[a][k], want);
    rc = run_ops(t, want, false);  // queued; the edges below wait for everything once
    if (rc) return forget_device_state(t, rc);
    std::vector<int> edges;
    for (int n = 1; n < t.nn; n++) edges.push_back(n);
    rc = refresh_pmats(t, edges, true, false);
    if (rc) return forget_device_state(t, rc);
    std::vector<DevEdgeOp> ops(t.nn - 1);
    for (int n = 1; n < t.nn; n++) {
      const int a = t.parent[n];
      ops[n - 1] = {t.slot[a][nb_index(t, a, n)], child_code(t, n, a), t.pm[n], 1};
    }
    rc = engine_edges(t.I, ops.data(), int(ops.size()), out, nullptr, &flags);
    if (rc) return forget_device_state(t, rc);
    const bool thrown = instance_precision(t.I) == GLE_PREC_FP64 &&
                        (((flags & GLE_FLAG_REDUCE_RESCALE) && t.rescale_every > 2) ||
                         ((flags & GLE_FLAG_BAIL_OUT) && t.rescale_every <= 2));
    if (!thrown) return GLE_SUCCESS;
    t.rescale_every -= 2;
    if (t.rescale_every < 2) return set_error(GLE_ERR_RESCALE_FATAL, "Problem with rescaling during sweep.");
  }
}

static void destroy_tree(Tree &t) {
  dirty_all(t);
  for (int n = 1; n < t.nn; n++)
    if (t.pm[n] >= 0) t.pool->pm_free.push_back(t.pm[n]);
}

// Exact names of directional CLAs: X(a|b) is named by its two (child name,
// branch-length bits) pairs in a fixed order, tips by their node numbers, and
// names are interned into integers.  Two CLAs get the same name exactly when
// the subtrees on their side are identical (tips, topology, lengths), which
// is exactly when one can stand for the other.
struct CladeNames {
  struct Key {
    uint64_t x, lx, y, ly;
    bool operator==(const Key &o) const { return x == o.x && lx == o.lx && y == o.y && ly == o.ly; }
  };
  struct KeyHash {
    size_t operator()(const Key &k) const {
      uint64_t h = k.x * 0x9E3779B97F4A7C15ull ^ (k.lx + 0x632BE59BD9B4E019ull);
      h = (h ^ (h >> 29)) * 0xBF58476D1CE4E5B9ull ^ (k.y * 0x94D049BB133111EBull) ^ k.ly;
      return size_t(h ^ (h >> 31));
    }
  };
  std::unordered_map<Key, uint64_t, KeyHash> ids;
  uint64_t next = uint64_t(1) << 40;  // above every tip number
  uint64_t name(uint64_t ix, double lx, uint64_t iy, double ly) {
    uint64_t bx, by;
    std::memcpy(&bx, &lx, 8);
    std::memcpy(&by, &ly, 8);
    Key k{ix, bx, iy, by};
    if (k.x > k.y || (k.x == k.y && k.lx > k.ly)) {
      std::swap(k.x, k.y);
      std::swap(k.lx, k.ly);
    }
    auto it = ids.find(k);
    if (it != ids.end()) return it->second;
    ids.emplace(k, next);
    return next++;
  }
};

static uint64_t clade_name(const Tree &t, int a, int b, CladeNames &cn, std::vector<std::array<uint64_t, 3>> &memo) {
  if (is_tip(t, a)) return uint64_t(a);
  uint64_t &m = memo[a][nb_index(t, a, b)];
  if (m) return m;
  int x = -1, y = -1;
  for (int j = 0; j < t.nnb[a]; j++) {
    const int c = t.nb[a][j];
    if (c == b) continue;
    if (x < 0) x = c;
    else y = c;
  }
  const uint64_t ix = clade_name(t, x, a, cn, memo), iy = clade_name(t, y, a, cn, memo);
  m = cn.name(ix, t.blen[edge_node(t, a, x)], iy, t.blen[edge_node(t, a, y)]);
  return m;
}

// A topology move (SPRMutate / NNIMutate, tree.cpp:2571-2708, 6970) followed
// by SweepDirtynessOverTree (tree.cpp:4247): the tree takes the new parent
// array and lengths and keeps every clean CLA whose side of the tree the move
// left unchanged -- however the nodes were renumbered -- so only the CLAs on
// the paths through the changed edges are recomputed.  P(t) slots stay with
// their nodes and are refreshed where a node's length changed.
static int set_topology(Tree &t, const int *parent, const double *blen) {
  std::vector<std::array<int, 3>> nb;
  std::vector<int> nnb;
  int rc = topology(t.N, parent, blen, nb, nnb);
  if (rc) return rc;
  CladeNames cn;
  std::unordered_map<uint64_t, int> clean;  // name -> slot of the old tree's clean CLAs
  {
    std::vector<std::array<uint64_t, 3>> memo(t.nn, {0, 0, 0});
    for (int a = 0; a < t.nn; a++)
      if (!is_tip(t, a))
        for (int k = 0; k < t.nnb[a]; k++)
          if (t.slot[a][k] >= 0) clean.emplace(clade_name(t, a, t.nb[a][k], cn, memo), t.slot[a][k]);
  }
  const std::vector<std::array<int, 3>> old_slot = t.slot;
  t.parent.assign(parent, parent + t.nn);
  t.blen.assign(blen, blen + t.nn);
  t.nb.swap(nb);
  t.nnb.swap(nnb);
  std::vector<std::array<uint64_t, 3>> memo(t.nn, {0, 0, 0});
  for (int a = 0; a < t.nn; a++) {
    t.slot[a] = {-1, -1, -1};
    if (is_tip(t, a)) continue;
    for (int k = 0; k < t.nnb[a]; k++) {
      auto it = clean.find(clade_name(t, a, t.nb[a][k], cn, memo));
      if (it != clean.end()) {
        t.slot[a][k] = it->second;
        t.pool->refs[it->second]++;
      }
    }
  }
  for (const auto &row : old_slot)  // the old tree's references
    for (int s : row)
      if (s >= 0 && --t.pool->refs[s] == 0) t.pool->free_slot(s);
  t.site_ready = false;
  return GLE_SUCCESS;
}

}  // namespace gle

using namespace gle;

// A change of the instance's data or model since the tree last computed makes
// every cached CLA and P(t) of the tree stale.
static void sync_generation(Tree &t) {
  const unsigned long long g = instance_generation(t.I);
  if (g == t.gen) return;
  dirty_all(t);
  std::fill(t.pm_len.begin(), t.pm_len.end(), std::nan(""));
  std::fill(t.pm_has_derivs.begin(), t.pm_has_derivs.end(), 0);
  t.gen = g;
}

#define GET_TREE(h)                                                        \
  std::lock_guard<std::recursive_mutex> lk_(g_tree_rmu);                   \
  std::lock_guard<std::mutex> pk_(pool_mutex());                           \
  Tree *T = get_tree(h);                                                   \
  if (!T) return set_error(GLE_ERR_ARG, "bad tree handle %d", (h));        \
  if (get_instance(T->inst_h) != T->I)                                     \
    return set_error(GLE_ERR_STATE, "the tree's instance was destroyed");  \
  sync_generation(*T);

extern "C" int gle_tree_create(int inst, const int *parent, const double *blen, int *tree) {
  if (!parent || !blen || !tree) return set_error(GLE_ERR_ARG, "null argument");
  Instance *I = get_instance(inst);
  if (!I) return set_error(GLE_ERR_ARG, "bad instance handle %d", inst);
  std::lock_guard<std::recursive_mutex> lk(g_tree_rmu);
  std::lock_guard<std::mutex> pk(pool_mutex());
  auto t = std::make_unique<Tree>();
  const int rc = build_tree(*t, inst, I, parent, blen);
  if (rc) {
    if (!t->pm.empty()) destroy_tree(*t);
    return rc;
  }
  g_trees.push_back(std::move(t));
  *tree = int(g_trees.size()) - 1;
  return GLE_SUCCESS;
}

extern "C" int gle_tree_copy(int h, int *out) {
  if (!out) return set_error(GLE_ERR_ARG, "null argument");
  GET_TREE(h);
  auto t = std::make_unique<Tree>();
  const int rc = build_tree(*t, T->inst_h, T->I, T->parent.data(), T->blen.data());
  if (rc) {
    if (!t->pm.empty()) destroy_tree(*t);
    return rc;
  }
  for (int a = 0; a < T->nn; a++)
    for (int k = 0; k < 3; k++) {
      const int s = T->slot[a][k];
      if (s >= 0) {
        t->slot[a][k] = s;
        T->pool->refs[s]++;
      }
    }
  t->rescale_every = T->rescale_every;
  t->schedule = T->schedule;
  g_trees.push_back(std::move(t));
  *out = int(g_trees.size()) - 1;
  return GLE_SUCCESS;
}

extern "C" int gle_tree_destroy(int h) {
  std::lock_guard<std::recursive_mutex> lk(g_tree_rmu);
  std::lock_guard<std::mutex> pk(pool_mutex());
  Tree *T = get_tree(h);
  if (!T) return set_error(GLE_ERR_ARG, "bad tree handle %d", h);
  if (get_instance(T->inst_h) == T->I) destroy_tree(*T);  // else the pool went with the instance
  g_trees[h].reset();
  return GLE_SUCCESS;
}

extern "C" int gle_tree_set_branch_length(int h, int node, double v) {
  GET_TREE(h);
  if (node <= 0 || node >= T->nn || !(v >= 1e-8 && v <= 100.0))
    return set_error(GLE_ERR_ARG, "bad node %d or length %g (allowed [1e-8, 100])", node, v);
  set_blen(*T, node, v);
  return GLE_SUCCESS;
}

extern "C" int gle_tree_set_topology(int h, const int *parent, const double *blen) {
  if (!parent || !blen) return set_error(GLE_ERR_ARG, "null argument");
  GET_TREE(h);
  return set_topology(*T, parent, blen);
}

extern "C" int gle_tree_cla_updates(int h, long long *count) {
  if (!count) return set_error(GLE_ERR_ARG, "null output");
  GET_TREE(h);
  *count = T->cla_updates;
  return GLE_SUCCESS;
}

extern "C" int gle_tree_get_branch_lengths(int h, double *out) {
  if (!out) return set_error(GLE_ERR_ARG, "null output");
  GET_TREE(h);
  std::copy(T->blen.begin(), T->blen.end(), out);
  return GLE_SUCCESS;
}

extern "C" int gle_tree_invalidate(int h) {
  GET_TREE(h);
  dirty_all(*T);
  return GLE_SUCCESS;
}

extern "C" int gle_tree_score(int h, double *lnl) {
  if (!lnl) return set_error(GLE_ERR_ARG, "null output");
  GET_TREE(h);
  return tree_score(*T, lnl);
}

extern "C" int gle_tree_site_likelihoods(int h, double *out) {
  if (!out) return set_error(GLE_ERR_ARG, "null output");
  GET_TREE(h);
  Tree &t = *T;
  if (!t.site_ready) return set_error(GLE_ERR_STATE, "score the tree first");
  if (t.site.empty()) {
    // the score's own edge: root_edge() now could pick another one, since
    // later derivative calls may have made other root CLAs clean
    std::vector<double> site(size_t(instance_patterns(t.I)));
    double res[3];
    const int rc = edge_eval(t, t.score_pa, t.score_pb, false, res, site.data());
    if (rc) return rc;
    t.site.swap(site);
    t.site_ready = true;  // a rescale retry inside edge_eval changes no value
  }
  std::copy(t.site.begin(), t.site.end(), out);
  return GLE_SUCCESS;
}

extern "C" int gle_tree_derivatives(int h, int node, double *lnl, double *d1, double *d2) {
  if (!lnl || !d1 || !d2) return set_error(GLE_ERR_ARG, "null output");
  GET_TREE(h);
  return tree_derivs(*T, node, lnl, d1, d2);
}

extern "C" int gle_tree_site_derivatives(int h, int node, double *out3, double *site_lnl, double *site_d1,
                                         double *site_d2, int *under_partial, int *under_child) {
  if (!out3) return set_error(GLE_ERR_ARG, "null output");
  GET_TREE(h);
  Tree &t = *T;
  if (node <= 0 || node >= t.nn) return set_error(GLE_ERR_ARG, "node %d has no parent edge", node);
  const size_t P = size_t(instance_patterns(t.I));
  std::vector<double> s5(5 * std::max<size_t>(P, 1));
  const int rc = edge_eval(t, t.parent[node], node, true, out3, nullptr, s5.data());
  if (rc) return rc;
  t.lnl = out3[0];
  for (size_t q = 0; q < P; q++) {
    if (site_lnl) site_lnl[q] = s5[q];
    if (site_d1) site_d1[q] = s5[P + q];
    if (site_d2) site_d2[q] = s5[2 * P + q];
    if (under_partial) under_partial[q] = int(s5[3 * P + q]);
    if (under_child) under_child[q] = int(s5[4 * P + q]);
  }
  return GLE_SUCCESS;
}

extern "C" int gle_tree_optimize_all_branches(int h, double prec, double *improve, double *lnl) {
  if (!improve || !lnl || !(prec > 0.0)) return set_error(GLE_ERR_ARG, "bad argument");
  GET_TREE(h);
  return optimize_all(*T, prec, improve, lnl);
}

extern "C" int gle_tree_sweep(int h, double *out) {
  if (!out) return set_error(GLE_ERR_ARG, "null output");
  GET_TREE(h);
  return sweep(*T, out);
}

extern "C" int gle_tree_get_cla(int h, int a, int b, double *out, int *u) {
  if (!out || !u) return set_error(GLE_ERR_ARG, "null output");
  GET_TREE(h);
  if (a < 0 || a >= T->nn || is_tip(*T, a)) return set_error(GLE_ERR_ARG, "node %d is not internal", a);
  const int k = nb_index(*T, a, b);
  if (k < 0) return set_error(GLE_ERR_ARG, "%d is not a neighbour of %d", b, a);
  std::vector<std::array<int, 3>> want;
  collect(*T, a, b, want);
  int rc = run_ops(*T, want);
  if (rc) return rc;
  return engine_get_partials(T->I, T->slot[a][k], out, u);
}

extern "C" int gle_tree_rescale_every(int h) {
  GET_TREE(h);
  return T->rescale_every;
}

extern "C" int gle_tree_set_rescale_every(int h, int every) {
  GET_TREE(h);
  if (every < 2) return set_error(GLE_ERR_ARG, "rescaleEvery must be >= 2");
  T->rescale_every = every;
  T->site_ready = false;
  return GLE_SUCCESS;
}

extern "C" int gle_tree_set_schedule(int h, int mode) {
  GET_TREE(h);
  if (mode != 0 && mode != 1) return set_error(GLE_ERR_ARG, "schedule must be 0 or 1");
  T->schedule = mode;
  return GLE_SUCCESS;
}

// ------------------------------------------------------------------ GTR+I+G
// Model parameters on a fixed tree: the nst=6 nucleotide model of the
// reference (Model::UpdateQMat + CalcEigenStuff, model.cpp:233-294, 827-856;
// DiscreteGamma, :2170-2201; pinv in the model, model.h:1733-1740) and its
// optimisation on a fixed topology (Population::OptimizeInputOnly ->
// BetterFinalOptimization, population.cpp:2566; OptimizeRelativeNucRates,
// tree.cpp:7307; OptimizeBoundedParameter, optimization.cpp:332): rounds of
// one-dimensional maximisations, one parameter at a time, with a Newton pass
// over all branches, until a round gains less than the precision.  Each
// trial value is a full model push and rescore on the device.

// Discrete-gamma categories with pinv (model.cpp:1478-1486, 2170-2201;
// model.h:1733-1740, 2051-2064) and one eigensystem for every category.
static int set_rate_het(Tree &t, const double *V, const double *W, const double *lam, const double *pi, double alpha,
                        double pinv_value, int pinv_in_model) {
  const int R = instance_rates(t.I);
  std::vector<double> rates(R, 1.0), mult(R), w(R);
  int rc;
  if (R > 1 && (rc = discrete_gamma(R, alpha, rates.data()))) return rc;
  const double pinv = pinv_in_model ? pinv_value : 0.0;
  for (int r = 0; r < R; r++) {
    mult[r] = pinv_in_model ? rates[r] / (1.0 - pinv) : rates[r];
    w[r] = (1.0 - pinv) / R;
  }
  std::vector<int> eor(R, 0);
  if ((rc = gle_set_eigen(t.inst_h, 0, V, W, lam)) || (rc = gle_set_category_rates(t.inst_h, eor.data(), mult.data())) ||
      (rc = gle_set_model(t.inst_h, pi, w.data(), pinv, pinv_in_model)))
    return rc;
  sync_generation(t);
  return GLE_SUCCESS;
}

static int apply_gtr(Tree &t, const gle_gtr_params &m) {
  if (instance_states(t.I) != 4) return set_error(GLE_ERR_UNSUPPORTED, "GTR parameters need a nucleotide instance");
  double V[16], W[16], lam[4];
  int rc = gle_model_gtr(m.pi, m.rel, V, W, lam);
  if (rc) return rc;
  return set_rate_het(t, V, W, lam, m.pi, m.alpha, m.pinv, m.pinv_in_model);
}

// UpdateQMatAminoAcid (model.cpp:735-780): 190 exchangeabilities (a named
// matrix or user values) times the target frequencies; rate heterogeneity as
// for nucleotides.
static int apply_aa(Tree &t, const gle_aa_params &m) {
  if (instance_states(t.I) != 20) return set_error(GLE_ERR_UNSUPPORTED, "amino-acid parameters need a 20-state instance");
  double V[400], W[400], lam[20];
  int rc = gle_model_aa(m.pi, m.rel, V, W, lam);
  if (rc) return rc;
  return set_rate_het(t, V, W, lam, m.pi, m.alpha, m.pinv, m.pinv_in_model);
}

// Brent's one-dimensional minimiser (golden section with parabolic steps,
// Brent 1973, ch. 5) applied to -f on [a, b] from x; returns the best point
// and its value in *xbest, *fbest (f maximised).
template <typename F>
static int brent_max(F &&f, double a, double b, double x, double xtol, int max_eval, double *xbest, double *fbest) {
  const double cgold = 0.3819660112501051;
  double fx;
  int rc = f(x, &fx);
  if (rc) return rc;
  fx = -fx;
  double v = x, w = x, fv = fx, fw = fx, d = 0.0, e = 0.0;
  for (int it = 0; it < max_eval; it++) {
    const double xm = 0.5 * (a + b), tol1 = xtol * std::fabs(x) + 1e-10, tol2 = 2.0 * tol1;
    if (std::fabs(x - xm) <= tol2 - 0.5 * (b - a)) break;
    bool golden = true;
    if (std::fabs(e) > tol1) {
      double r = (x - w) * (fx - fv), q = (x - v) * (fx - fw), pp = (x - v) * q - (x - w) * r;
      q = 2.0 * (q - r);
      if (q > 0.0) pp = -pp;
      q = std::fabs(q);
      const double etemp = e;
      e = d;
      if (!(std::fabs(pp) >= std::fabs(0.5 * q * etemp) || pp <= q * (a - x) || pp >= q * (b - x))) {
        d = pp / q;
        const double u = x + d;
        if (u - a < tol2 || b - u < tol2) d = xm >= x ? tol1 : -tol1;
        golden = false;
      }
    }
    if (golden) {
      e = (x >= xm ? a : b) - x;
      d = cgold * e;
    }
    const double u = std::fabs(d) >= tol1 ? x + d : x + (d > 0 ? tol1 : -tol1);
    double fu;
    if ((rc = f(u, &fu))) return rc;
    fu = -fu;
    if (fu <= fx) {
      (u >= x ? a : b) = x;
      v = w;
      fv = fw;
      w = x;
      fw = fx;
      x = u;
      fx = fu;
    } else {
      (u < x ? a : b) = u;
      if (fu <= fw || w == x) {
        v = w;
        fv = fw;
        w = u;
        fw = fu;
      } else if (fu <= fv || v == x || v == w) {
        v = u;
        fv = fu;
      }
    }
  }
  *xbest = x;
  *fbest = -fx;
  return GLE_SUCCESS;
}

extern "C" int gle_tree_set_gtr(int h, const gle_gtr_params *m) {
  if (!m) return set_error(GLE_ERR_ARG, "null model");
  GET_TREE(h);
  return apply_gtr(*T, *m);
}

// One coordinate of a model: value through `set`, searched by brent_max on a
// transformed axis in [lo, hi] from x0; the best point is kept when it beats
// `best`, and the instance is left at the accepted model.
template <typename Params, typename Apply, typename Set>
static int line_search(Tree &t, Params &cur, double &best, Apply &&apply, double x0, double lo, double hi,
                       Set &&set) {
  auto f = [&](double x, double *val) -> int {
    Params trial = cur;
    set(trial, x);
    int rc = apply(t, trial);
    if (rc) return rc;
    return tree_score(t, val);
  };
  double xb, fb;
  int rc = brent_max(f, lo, hi, x0, 1e-5, 60, &xb, &fb);
  if (rc) return rc;
  if (fb > best) {
    set(cur, xb);
    best = fb;
  }
  return apply(t, cur);
}

// at most a factor 5 from the current value per round, as the reference
// limits one pass (maxChangeProportion, tree.cpp:7330)
static const double kStep = std::log(5.0);
// codon frequencies: the reference's factor per pass (maxChangeProportion, tree.cpp:7294)
static const double kFreqStep = std::log(1.2);

// The parameter loop shared by the nucleotide and amino-acid models: rounds
// of a Newton pass over the branches, then one Brent search per free
// exchangeability (log axis, within [rel_lo, rel_hi]; the last n_rel_fixed
// entries stay put), per frequency (log-ratio to the last state), for the
// gamma shape (log axis, [0.05, 999]) and for pinv ([0, 0.999 x the constant
// fraction]), until a round gains less than prec.
template <typename Params, typename Apply>
static int optimize_rate_het_model(Tree &t, Params &cur, int n_rel, int n_states, double rel_lo, double rel_hi,
                                   int which, double prec, Apply &&apply) {
  const int R = instance_rates(t.I);
  int rc = apply(t, cur);
  if (rc) return rc;
  double best;
  if ((rc = tree_score(t, &best))) return rc;
  for (int round = 0; round < 200; round++) {
    const double start = best;
    if (which & GLE_OPT_BRANCHES) {
      double imp, l;
      if ((rc = optimize_all(t, prec, &imp, &l))) return rc;
      best = l;
    }
    if (which & GLE_OPT_RATES)
      for (int i = 0; i < n_rel; i++) {
        const double x0 = std::log(cur.rel[i]);
        rc = line_search(t, cur, best, apply, x0, std::max(x0 - kStep, std::log(rel_lo)),
                         std::min(x0 + kStep, std::log(rel_hi)), [i](Params &q, double x) { q.rel[i] = std::exp(x); });
        if (rc) return rc;
      }
    if (which & GLE_OPT_FREQS)
      for (int i = 0; i < n_states - 1; i++) {
        const double x0 = std::log(cur.pi[i] / cur.pi[n_states - 1]);
        rc = line_search(t, cur, best, apply, x0, x0 - kStep, x0 + kStep, [i, n_states](Params &q, double x) {
          double u[64], sum = 0.0;
          for (int k = 0; k < n_states - 1; k++) u[k] = q.pi[k] / q.pi[n_states - 1];
          u[i] = std::exp(x);
          u[n_states - 1] = 1.0;
          for (int k = 0; k < n_states; k++) sum += u[k];
          for (int k = 0; k < n_states; k++) q.pi[k] = u[k] / sum;
        });
        if (rc) return rc;
      }
    if ((which & GLE_OPT_ALPHA) && R > 1) {
      const double x0 = std::log(cur.alpha);
      rc = line_search(t, cur, best, apply, x0, std::max(x0 - kStep, std::log(0.05)),
                       std::min(x0 + kStep, std::log(999.0)), [](Params &q, double x) { q.alpha = std::exp(x); });
      if (rc) return rc;
    }
    if ((which & GLE_OPT_PINV) && cur.pinv_in_model) {
      const double hi = 0.999 * instance_constant_fraction(t.I);
      if (hi > 0.0) {
        rc = line_search(t, cur, best, apply, std::min(cur.pinv, hi), 0.0, hi,
                         [](Params &q, double x) { q.pinv = x; });
        if (rc) return rc;
      }
    }
    if (!(best - start >= prec)) break;
  }
  return GLE_SUCCESS;
}

extern "C" int gle_tree_optimize_gtr(int h, gle_gtr_params *m, int which, double prec, double *lnl) {
  if (!m || !lnl || !(prec > 0.0)) return set_error(GLE_ERR_ARG, "bad argument");
  GET_TREE(h);
  gle_gtr_params cur = *m;
  // five free rates with GT = 1, bounds of the reference's rate search
  int rc = optimize_rate_het_model(*T, cur, 5, 4, 1e-5, 999.0, which, prec, apply_gtr);
  if (rc) return rc;
  *m = cur;
  return tree_score(*T, lnl);
}

extern "C" int gle_tree_set_aa(int h, const gle_aa_params *m) {
  if (!m) return set_error(GLE_ERR_ARG, "null model");
  GET_TREE(h);
  return apply_aa(*T, *m);
}

extern "C" int gle_tree_optimize_aa(int h, gle_aa_params *m, int which, double prec, double *lnl) {
  if (!m || !lnl || !(prec > 0.0)) return set_error(GLE_ERR_ARG, "bad argument");
  if (which & (GLE_OPT_OMEGA | GLE_OPT_OMEGA_PROBS)) return set_error(GLE_ERR_ARG, "no omegas in an amino-acid model");
  GET_TREE(h);
  gle_aa_params cur = *m;
  if (which & GLE_OPT_RATES)
    for (double r : cur.rel)
      if (!(r > 0.0)) return set_error(GLE_ERR_ARG, "estimating exchangeabilities needs them all positive");
  // 189 free exchangeabilities relative to the last (Y-V), within 1e-5 and 1e5 of it
  const double ref = cur.rel[189];
  int rc = optimize_rate_het_model(*T, cur, 189, 20, 1e-5 * ref, 1e5 * ref, which, prec, apply_aa);
  if (rc) return rc;
  *m = cur;
  return tree_score(*T, lnl);
}

// ------------------------------------------------------------------- codons
// GARLI's codon models on a fixed tree: nst=6 exchangeabilities between
// codons one nucleotide apart, an omega class per rate category with its
// proportion (M0 .. M3, model.cpp:442-687, 1514-1530) and codon frequencies;
// optimisation of the exchangeabilities (OptimizeRelativeNucRates,
// tree.cpp:7307), of the omegas and their proportions
// (OptimizeOmegaParameters, tree.cpp:7164) and of the frequencies
// (OptimizeEquilibriumFreqs, tree.cpp:7288) with Newton branch passes.

static int apply_codon(Tree &t, const gle_codon_params &m) {
  const int R = instance_rates(t.I);
  if (instance_states(t.I) != 61) return set_error(GLE_ERR_UNSUPPORTED, "codon parameters need a 61-state instance");
  if (m.n_omega != R) return set_error(GLE_ERR_ARG, "%d omega classes for %d rate categories", m.n_omega, R);
  std::vector<double> V(size_t(R) * 3721), W(size_t(R) * 3721), lam(size_t(R) * 61);
  int rc = gle_model_codon(m.pi, m.rel, m.omega, m.omega_prob, R, V.data(), W.data(), lam.data());
  if (rc) return rc;
  std::vector<int> eor(R);
  std::vector<double> mult(R, 1.0);
  for (int k = 0; k < R; k++) {
    eor[k] = k;
    if ((rc = gle_set_eigen(t.inst_h, k, V.data() + size_t(k) * 3721, W.data() + size_t(k) * 3721,
                            lam.data() + size_t(k) * 61)))
      return rc;
  }
  if ((rc = gle_set_category_rates(t.inst_h, eor.data(), mult.data())) ||
      (rc = gle_set_model(t.inst_h, m.pi, m.omega_prob, 0.0, 0)))
    return rc;
  sync_generation(t);
  return GLE_SUCCESS;
}

extern "C" int gle_tree_set_codon(int h, const gle_codon_params *m) {
  if (!m) return set_error(GLE_ERR_ARG, "null model");
  GET_TREE(h);
  return apply_codon(*T, *m);
}

extern "C" int gle_tree_optimize_codon(int h, gle_codon_params *m, int which, double prec, double *lnl) {
  if (!m || !lnl || !(prec > 0.0)) return set_error(GLE_ERR_ARG, "bad argument");
  GET_TREE(h);
  Tree &t = *T;
  const int K = m->n_omega;
  gle_codon_params cur = *m;
  int rc = apply_codon(t, cur);
  if (rc) return rc;
  double best;
  if ((rc = tree_score(t, &best))) return rc;
  for (int round = 0; round < 200; round++) {
    const double start = best;
    if (which & GLE_OPT_BRANCHES) {
      double imp, l;
      if ((rc = optimize_all(t, prec, &imp, &l))) return rc;
      best = l;
    }
    if (which & GLE_OPT_RATES)  // AC .. CT relative to GT
      for (int i = 0; i < 5; i++) {
        const double x0 = std::log(cur.rel[i] / cur.rel[5]);
        rc = line_search(t, cur, best, apply_codon, x0, std::max(x0 - kStep, std::log(1e-3)),
                         std::min(x0 + kStep, std::log(999.0)),
                         [i](gle_codon_params &q, double x) { q.rel[i] = std::exp(x) * q.rel[5]; });
        if (rc) return rc;
      }
    if (which & GLE_OPT_OMEGA)
      for (int k = 0; k < K; k++) {
        const double x0 = std::log(cur.omega[k]);
        rc = line_search(t, cur, best, apply_codon, x0, std::max(x0 - kStep, std::log(1e-4)),
                         std::min(x0 + kStep, std::log(999.0)),
                         [k](gle_codon_params &q, double x) { q.omega[k] = std::exp(x); });
        if (rc) return rc;
      }
    if (which & GLE_OPT_FREQS)
      // statefrequencies = estimateF (model.h:871-876; "estimate" refuses codon
      // models, model.h:866-867): OptimizeEquilibriumFreqs (tree.cpp:7288-7303),
      // each frequency in turn within a factor 1.2 per round, here as the
      // log-ratio of codon i's frequency to the last codon's
      for (int i = 0; i < 60; i++) {
        const double x0 = std::log(cur.pi[i] / cur.pi[60]);
        rc = line_search(t, cur, best, apply_codon, x0, x0 - kFreqStep, x0 + kFreqStep,
                         [i](gle_codon_params &q, double x) {
                           double u[61], sum = 0.0;
                           for (int k = 0; k < 61; k++) u[k] = q.pi[k] / q.pi[60];
                           u[i] = std::exp(x);
                           u[60] = 1.0;
                           for (int k = 0; k < 61; k++) sum += u[k];
                           for (int k = 0; k < 61; k++) q.pi[k] = u[k] / sum;
                         });
        if (rc) return rc;
      }
    if ((which & GLE_OPT_OMEGA_PROBS) && K > 1)  // log-ratio of each proportion to the last class's
      for (int k = 0; k < K - 1; k++) {
        const double x0 = std::log(cur.omega_prob[k] / cur.omega_prob[K - 1]);
        rc = line_search(t, cur, best, apply_codon, x0, std::max(x0 - kStep, std::log(1e-4)),
                         std::min(x0 + kStep, std::log(1e4)), [k, K](gle_codon_params &q, double x) {
                           double u[8];
                           for (int j = 0; j < K; j++) u[j] = q.omega_prob[j] / q.omega_prob[K - 1];
                           u[k] = std::exp(x);
                           double sum = 0.0;
                           for (int j = 0; j < K; j++) sum += u[j];
                           for (int j = 0; j < K; j++) q.omega_prob[j] = u[j] / sum;
                         });
        if (rc) return rc;
      }
    if (!(best - start >= prec)) break;
  }
  *m = cur;
  return tree_score(t, lnl);
}
