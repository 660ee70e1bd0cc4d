// gle_serve.cpp -- continuous-batching scoring service for a GA population.
//
// GARLI scores its population one individual at a time (population.cpp:4555);
// each individual is a parent tree plus a topology move and a branch-length
// pass (Individual::Mutate, individual.cpp:127-222).  The service takes such
// requests from any number of client threads and runs them together:
//
//  * CLA cache.  Every directional CLA X(a|b) is named by an exact subtree
//    id: ids are interned from (id of one child, bits of its branch length,
//    id of the other child, bits of its branch length), the pair ordered, and
//    tips are their node numbers.  Two CLAs get the same id exactly when
//    their subtrees are identical (tips, topology, branch-length bits), so no
//    collision can hand a request another subtree's CLA.  Any request that
//    needs a subtree some earlier request computed -- the unmodified part of a
//    perturbed parent, or a subtree shared by other individuals -- reuses the
//    device slot.  This generalises ClaManager's holder sharing
//    (clamanager.h:96-441): a changed branch changes the ids of exactly the
//    CLAs SweepDirtynessOverTree would dirty.
//  * P(t) cache keyed by branch-length bits (P, P', P'' per slot).
//  * Admission by HBM budget: a request is admitted when its working set fits
//    in the instance's CLA slots alongside the requests already running;
//    least-recently-used cache entries are evicted on demand.  A tick whose
//    new CLAs or P(t) would not fit runs the requests that do and keeps the
//    rest for the next tick.
//  * Continuous batching: each tick gathers the next step of every admitted
//    request (a score, or one Newton iteration of a branch-length pass),
//    dedupes shared CLAs, and issues one P(t) launch, one CLA launch per
//    dependency level and one edge launch for all of them.  Requests that
//    arrive while a tick runs join the next tick.
//  * A score is taken on the edge that needs the fewest new CLAs given the
//    cache (lnL is the same on every edge), next to the perturbation of an
//    offspring.  Every CLA and edge reduction is deterministic, but the edge
//    depends on what is cached, so a request's result alone and in a mix
//    agree to the precision form's rounding rather than bit for bit.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gle_internal.h"
#include "gle_newton.h"

namespace gle {

static inline uint64_t mix64(uint64_t x) {  // SplitMix64 finaliser
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
static inline uint64_t dbits(double v) {
  uint64_t u;
  std::memcpy(&u, &v, sizeof u);
  return u;
}

// The exact name of an internal subtree: its two (child id, branch-length
// bits) pairs in a fixed order.
struct SubKey {
  uint64_t x, lx, y, ly;
  bool operator==(const SubKey &o) const { return x == o.x && lx == o.lx && y == o.y && ly == o.ly; }
};
struct SubKeyHash {
  size_t operator()(const SubKey &k) const {
    return size_t(mix64(k.x ^ mix64(k.lx ^ mix64(k.y ^ mix64(k.ly)))));
  }
};
static SubKey sub_key(uint64_t ix, double lx, uint64_t iy, double ly) {
  SubKey k{ix, dbits(lx), iy, dbits(ly)};
  if (k.x > k.y || (k.x == k.y && k.lx > k.ly)) {
    std::swap(k.x, k.y);
    std::swap(k.lx, k.ly);
  }
  return k;
}
// Internal subtree ids start above every tip number.
constexpr uint64_t kFirstSubtreeId = uint64_t(1) << 40;
// Interned names are kept while their CLAs come and go; past this many the
// table and the CLA cache are dropped together (a rare full refill).
constexpr size_t kInternCap = size_t(1) << 22;

struct CacheEntry {
  int slot = -1, rank = 0;
  uint64_t last = 0;
};

struct ServeInstance {
  int h = -1;
  Instance *I = nullptr;
  int rescale_every = 16;
  std::unordered_map<SubKey, uint64_t, SubKeyHash> intern;  // subtree name -> id
  uint64_t next_id = kFirstSubtreeId;
  std::unordered_map<uint64_t, CacheEntry> cla;  // subtree id -> CLA slot
  std::unordered_map<uint64_t, CacheEntry> pm;   // blen bits -> P, P', P'' slot
  int reserved = 0;                              // working-set slots of admitted requests
  unsigned long long gen = 0;                    // instance generation the caches belong to

  uint64_t intern_id(const SubKey &k) {
    auto it = intern.find(k);
    if (it != intern.end()) return it->second;
    const uint64_t id = next_id++;
    intern.emplace(k, id);
    return id;
  }
  uint64_t find_id(const SubKey &k) const {
    auto it = intern.find(k);
    return it == intern.end() ? 0 : it->second;
  }
  // every cached CLA back to the pool (caller holds pool_mutex)
  void drop_clas(ClaPool *pool) {
    for (auto &c : cla)
      if (--pool->refs[c.second.slot] == 0) pool->free_slot(c.second.slot);
    cla.clear();
  }
};

enum { kReqScore = 0, kReqOptimize = 1 };

// per-request memo of subtree ids: kUnset = not looked up yet, 0 = no id yet
constexpr uint64_t kUnset = ~uint64_t(0);

struct Req {
  int id = 0, inst = 0, kind = 0, N = 0, nn = 0;
  double prec = 0.0;
  std::vector<int> parent;
  std::vector<double> blen;
  std::vector<std::array<int, 3>> nb;
  std::vector<int> nnb;
  std::vector<int> order;
  size_t bi = 0;
  NewtonBranch newton;
  bool scoring = true;  // final score phase
  double lnl = 0.0, improve = 0.0;
  int rc = GLE_SUCCESS, need = 0;
  int waits = 0;          // ticks in a row this request was deferred (goes first next time)
  bool done = false;      // set by the worker while it owns the request
  bool finished = false;  // published under Server::mu once the worker let go
  std::chrono::steady_clock::time_point t_submit, t_done;
  // per-tick scratch
  std::vector<std::array<uint64_t, 3>> sid;
  int edge_a = 0, edge_b = 0;
  bool derivs = false;
};

struct Server {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req *> pending;
  std::vector<Req *> active;
  std::unordered_map<int, std::unique_ptr<Req>> reqs;
  std::map<int, ServeInstance> insts;
  std::thread worker;
  bool stop = false;
  int next_id = 1;
  uint64_t tick = 0;
  std::vector<double> latency_ms;
  long long batched_requests = 0, ticks = 0, deferred = 0;
  int batch_window_us = 0;  // an idle server waits this long for a burst to gather
  // where tick time goes: building op lists on the host, and from the first
  // launch until the results are back
  double host_ms = 0.0, device_ms = 0.0;
};

static std::mutex g_servers_mu;
static std::vector<std::unique_ptr<Server>> g_servers;

static int tree_topology(Req &r) {
  r.nn = 2 * r.N - 2;
  r.nb.assign(r.nn, {-1, -1, -1});
  r.nnb.assign(r.nn, 0);
  if (int(r.parent.size()) != r.nn || r.parent[0] != -1) return set_error(GLE_ERR_ARG, "bad parent array");
  for (int n = 1; n < r.nn; n++) {
    const int a = r.parent[n];
    if (a < 0 || a >= r.nn || a == n || (a >= 1 && a <= r.N)) return set_error(GLE_ERR_ARG, "node %d: bad parent", n);
    if (!(r.blen[n] >= 0.0)) return set_error(GLE_ERR_ARG, "node %d: bad length", n);
    r.nb[n][r.nnb[n]++] = a;
  }
  for (int n = 1; n < r.nn; n++) {
    const int a = r.parent[n];
    if (r.nnb[a] >= 3) return set_error(GLE_ERR_ARG, "node %d has too many neighbours", a);
    r.nb[a][r.nnb[a]++] = n;
  }
  for (int n = 0; n < r.nn; n++)
    if (r.nnb[n] != ((n >= 1 && n <= r.N) ? 1 : 3)) return set_error(GLE_ERR_ARG, "node %d: wrong degree", n);
  for (int n = 1; n < r.nn; n++) {
    int x = n, steps = 0;
    while (x != 0 && steps++ <= r.nn) x = r.parent[x];
    if (x != 0) return set_error(GLE_ERR_ARG, "node %d does not reach the root", n);
  }
  return GLE_SUCCESS;
}

static bool req_tip(const Req &r, int n) { return n >= 1 && n <= r.N; }
static int req_nbk(const Req &r, int a, int b) {
  for (int k = 0; k < r.nnb[a]; k++)
    if (r.nb[a][k] == b) return k;
  return -1;
}
static double req_len(const Req &r, int a, int b) { return r.parent[a] == b ? r.blen[a] : r.blen[b]; }
static void req_children(const Req &r, int a, int b, int &x, int &y) {
  x = y = -1;
  for (int j = 0; j < r.nnb[a]; j++) {
    const int c = r.nb[a][j];
    if (c == b) continue;
    if (x < 0) x = c;
    else y = c;
  }
}

// The id X(a|b) already has (0: its subtree was never named, so no CLA of it
// can be cached), memoised per tick in r.sid.
static uint64_t known_id(Req &r, const ServeInstance &si, int a, int b) {
  if (req_tip(r, a)) return uint64_t(a);
  uint64_t &m = r.sid[a][req_nbk(r, a, b)];
  if (m != kUnset) return m;
  int x, y;
  req_children(r, a, b, x, y);
  const uint64_t ix = known_id(r, si, x, a), iy = ix ? known_id(r, si, y, a) : 0;
  m = (ix && iy) ? si.find_id(sub_key(ix, req_len(r, a, x), iy, req_len(r, a, y))) : 0;
  return m;
}

// ------------------------------------------------------------ one tick
struct TickNew {
  int slot, rank, level;
};

struct TickBuilder {
  ServeInstance *si;
  ClaPool *pool;
  uint64_t tick;
  std::unordered_map<uint64_t, TickNew> fresh;
  std::vector<uint64_t> fresh_order;          // ids in the order they joined `fresh`
  std::vector<std::pair<int, DevClaOp>> ops;  // (level, op)
  std::vector<int> pm_idx;
  std::vector<double> pm_len;
  std::unordered_map<uint64_t, int> pm_fresh;
  int err = GLE_SUCCESS;

  struct Mark {
    size_t fresh, ops, pm;
  };
  Mark mark() const { return {fresh_order.size(), ops.size(), pm_idx.size()}; }
  // undo everything one request added since m: its new CLA slots and P(t)
  // slots go back to the pool
  void rollback(const Mark &m) {
    for (size_t i = m.fresh; i < fresh_order.size(); i++) {
      auto it = fresh.find(fresh_order[i]);
      const int s = it->second.slot;
      if (--pool->refs[s] == 0) pool->free_slot(s);
      fresh.erase(it);
    }
    fresh_order.resize(m.fresh);
    ops.resize(m.ops);
    for (size_t i = m.pm; i < pm_idx.size(); i++) {
      pm_fresh.erase(dbits(pm_len[i]));
      pool->pm_free.push_back(pm_idx[i]);
    }
    pm_idx.resize(m.pm);
    pm_len.resize(m.pm);
    err = GLE_SUCCESS;
  }
  // every new CLA and P(t) slot of the tick back to the pool (a failed tick)
  void release_all() { rollback(Mark{0, 0, 0}); }

  int take_slot() {
    if (!pool->has_free()) {  // evict the least recently used entry not used this tick
      auto victim = si->cla.end();
      for (auto it = si->cla.begin(); it != si->cla.end(); ++it)
        if (it->second.last < tick && (victim == si->cla.end() || it->second.last < victim->second.last))
          victim = it;
      if (victim == si->cla.end()) return -1;
      const int s = victim->second.slot;
      if (--pool->refs[s] == 0) pool->free_slot(s);
      si->cla.erase(victim);
      if (!pool->has_free()) return -1;
    }
    const int s = pool->take_slot();
    pool->refs[s] = 1;
    return s;
  }

  int pm_slot(double len) {
    const uint64_t key = dbits(len);
    auto it = si->pm.find(key);
    if (it != si->pm.end()) {
      it->second.last = tick;
      return it->second.slot;
    }
    auto f = pm_fresh.find(key);
    if (f != pm_fresh.end()) return f->second;
    if (pool->pm_free.empty()) {
      auto victim = si->pm.end();
      for (auto jt = si->pm.begin(); jt != si->pm.end(); ++jt)
        if (jt->second.last < tick && (victim == si->pm.end() || jt->second.last < victim->second.last))
          victim = jt;
      if (victim == si->pm.end()) return -1;
      pool->pm_free.push_back(victim->second.slot);
      si->pm.erase(victim);
    }
    const int s = pool->pm_free.back();
    pool->pm_free.pop_back();
    pm_fresh[key] = s;
    pm_idx.push_back(s);
    pm_len.push_back(len);
    return s;
  }

  // device reference of X(a|b): CLA slot, or tip code -(tip)-1; fills rank,
  // level and the subtree id (memoised in r.sid)
  int resolve(Req &r, int a, int b, int &rank, int &level) {
    if (req_tip(r, a)) {
      rank = 0;
      level = -1;
      return -a;
    }
    const uint64_t known = known_id(r, *si, a, b);
    if (known) {
      auto it = si->cla.find(known);
      if (it != si->cla.end()) {
        it->second.last = tick;
        rank = it->second.rank;
        level = -1;
        return it->second.slot;
      }
      auto f = fresh.find(known);
      if (f != fresh.end()) {
        rank = f->second.rank;
        level = f->second.level;
        return f->second.slot;
      }
    }
    int x, y;
    req_children(r, a, b, x, y);
    int rx, ry, lx, ly;
    DevClaOp op;
    op.c1 = resolve(r, x, a, rx, lx);
    if (err) return 0;
    op.c2 = resolve(r, y, a, ry, ly);
    if (err) return 0;
    const uint64_t ix = req_tip(r, x) ? uint64_t(x) : r.sid[x][req_nbk(r, x, a)];
    const uint64_t iy = req_tip(r, y) ? uint64_t(y) : r.sid[y][req_nbk(r, y, a)];
    const uint64_t id = si->intern_id(sub_key(ix, req_len(r, a, x), iy, req_len(r, a, y)));
    r.sid[a][req_nbk(r, a, b)] = id;
    op.pm1 = pm_slot(req_len(r, a, x));
    op.pm2 = op.pm1 < 0 ? -1 : pm_slot(req_len(r, a, y));
    const int slot = op.pm2 < 0 ? -1 : take_slot();
    if (slot < 0) {
      err = set_error(GLE_ERR_NOMEM, "serving cache exhausted: raise n_cla_slots / n_pmat_slots");
      return 0;
    }
    rank = 2 + rx + ry;
    op.flags = 0;
    if (rank >= si->rescale_every) {
      op.flags = kOpRescale;
      rank = 0;
    }
    op.dest = slot;
    level = std::max(lx, ly) + 1;
    fresh[id] = {slot, rank, level};
    fresh_order.push_back(id);
    ops.push_back({level, op});
    return slot;
  }
};

// Scoring edge of a request.  lnL is the same on every edge (the pruning
// algorithm's root invariance; the reference's RunTests checks it,
// population.cpp:924-954), so a score is taken where the fewest CLAs are
// missing from the cache and this tick: next to the perturbation, for an
// offspring whose parent's CLAs are cached.  Ties keep the reference's edge
// (root, middle child).  The choice depends on what is cached, so a result
// alone and in a mix agree to the form's rounding, not bit for bit.
static void choose_score_edge(Req &r, const TickBuilder &tb) {
  std::vector<std::array<int, 3>> memo(r.nn, {-1, -1, -1});
  std::function<int(int, int)> cost = [&](int a, int b) -> int {  // CLAs missing for X(a|b)
    if (req_tip(r, a)) return 0;
    int &m = memo[a][req_nbk(r, a, b)];
    if (m >= 0) return m;
    const uint64_t id = known_id(r, *tb.si, a, b);
    if (id && (tb.si->cla.count(id) || tb.fresh.count(id))) return m = 0;
    int c = 1;
    for (int j = 0; j < r.nnb[a]; j++)
      if (r.nb[a][j] != b) c += cost(r.nb[a][j], a);
    return m = c;
  };
  r.edge_a = 0;
  r.edge_b = r.nb[0][1];
  int best = cost(r.edge_a, r.edge_b) + cost(r.edge_b, r.edge_a);
  for (int n = 1; n < r.nn && best > 0; n++) {
    const int a = r.parent[n];
    const int c = cost(n, a) + cost(a, n);
    if (c < best) {
      best = c;
      r.edge_a = a;
      r.edge_b = n;
    }
  }
}

// Next evaluation of a request: (edge, derivs?).
static void next_edge(Req &r, const TickBuilder &tb) {
  if (r.scoring) {
    choose_score_edge(r, tb);
    r.derivs = false;
  } else {
    const int nd = r.order[r.bi];
    r.edge_a = r.parent[nd];
    r.edge_b = nd;
    r.derivs = true;
  }
}

static void advance(Req &r, const double *res) {
  if (r.scoring) {
    r.lnl = res[0];
    r.done = true;
    return;
  }
  const int nd = r.order[r.bi];
  double next = 0.0;
  const NewtonBranch::Action act = r.newton.step(res[0], res[1], res[2], r.blen[nd], &next);
  if (act != NewtonBranch::kFinish) r.blen[nd] = next;
  if (act != NewtonBranch::kContinue) {
    r.improve += r.newton.total;
    if (++r.bi < r.order.size()) r.newton.start(r.blen[r.order[r.bi]], r.prec);
    else r.scoring = true;
  }
}

// Runs one batched step of the requests in `batch` (all on instance si).  A
// request whose new CLAs or P(t) do not fit beside those the tick already
// holds waits for the next tick (it stays active and unchanged); one that
// does not fit on its own fails.  On a device error every request of the tick
// fails and every slot the tick took goes back to the pool.  Returns how many
// requests were deferred.
static int run_tick(Server &sv, ServeInstance &si, std::vector<Req *> &batch, double &host_ms, double &device_ms) {
  using clock = std::chrono::steady_clock;
  auto ms_since = [](clock::time_point a) { return std::chrono::duration<double, std::milli>(clock::now() - a).count(); };
  auto fail_all = [](std::vector<Req *> &v, int rc) {
    for (Req *r : v) {
      r->rc = rc;
      r->done = true;
    }
  };
  for (;;) {
    const auto t_build = clock::now();
    auto t_launch = t_build;
    std::vector<DevClaOp> ops;
    std::vector<int> level_start;
    std::vector<DevEdgeOp> eops;
    std::vector<Req *> in_tick;
    int deferred = 0;
    std::unique_lock<std::mutex> pk(pool_mutex());
    ClaPool *pool = pool_for(si.I);
    if (si.gen != instance_generation(si.I) || si.intern.size() > kInternCap) {
      // data or model changed (or the name table is full): drop both caches
      si.drop_clas(pool);
      for (auto &c : si.pm) pool->pm_free.push_back(c.second.slot);
      si.pm.clear();
      si.intern.clear();
      si.gen = instance_generation(si.I);
    }
    TickBuilder tb{&si, pool, sv.tick, {}, {}, {}, {}, {}, {}, GLE_SUCCESS};
    for (Req *r : batch) {
      const TickBuilder::Mark m = tb.mark();
      r->sid.assign(r->nn, {kUnset, kUnset, kUnset});
      next_edge(*r, tb);
      int ra, la, rb, lb;
      DevEdgeOp e;
      e.partial = tb.resolve(*r, r->edge_a, r->edge_b, ra, la);
      e.child = tb.err ? 0 : tb.resolve(*r, r->edge_b, r->edge_a, rb, lb);
      e.pm = tb.err ? -1 : tb.pm_slot(req_len(*r, r->edge_a, r->edge_b));
      e.derivs = r->derivs ? 1 : 0;
      if (!tb.err && e.pm < 0) tb.err = set_error(GLE_ERR_NOMEM, "serving cache exhausted: out of pmat slots");
      if (tb.err) {
        const int err = tb.err;
        tb.rollback(m);
        if (in_tick.empty()) {  // does not fit even alone
          r->rc = err;
          r->done = true;
        } else {
          deferred++;  // next tick, with the slots this tick holds free again
          r->waits++;
        }
        continue;
      }
      eops.push_back(e);
      in_tick.push_back(r);
      r->waits = 0;
    }
    if (in_tick.empty()) return deferred;
    std::stable_sort(tb.ops.begin(), tb.ops.end(),
                     [](const std::pair<int, DevClaOp> &x, const std::pair<int, DevClaOp> &y) {
                       return x.first < y.first;
                     });
    int cur = -1;
    for (auto &lo : tb.ops) {
      if (lo.first != cur) {
        level_start.push_back(int(ops.size()));
        cur = lo.first;
      }
      ops.push_back(lo.second);
    }
    level_start.push_back(int(ops.size()));
    host_ms += ms_since(t_build);
    t_launch = clock::now();
    // P, P', P'' for new branch lengths (one launch)
    int rc = engine_update_pmats(si.I, tb.pm_idx.data(), tb.pm_len.data(), int(tb.pm_idx.size()), 1);
    if (rc) {
      tb.release_all();
      fail_all(in_tick, rc);
      return deferred;
    }
    for (size_t i = 0; i < tb.pm_idx.size(); i++) si.pm[dbits(tb.pm_len[i])] = {tb.pm_idx[i], 0, sv.tick};
    tb.pm_idx.clear();  // now owned by the P(t) cache
    tb.pm_len.clear();
    tb.pm_fresh.clear();
    pk.unlock();
    int flags = 0;
    rc = engine_flags(si.I, nullptr, 1);
    if (!rc)
      rc = engine_update_partials(si.I, ops.data(), int(ops.size()), level_start.data(),
                                  int(level_start.size()) - 1, false);
    std::vector<double> res(3 * eops.size());
    if (!rc) rc = engine_edges(si.I, eops.data(), int(eops.size()), res.data(), nullptr);
    if (!rc) rc = engine_flags(si.I, &flags, 1);
    device_ms += ms_since(t_launch);
    pk.lock();
    if (rc) {
      tb.release_all();
      fail_all(in_tick, rc);
      return deferred;
    }
    const bool thrown = instance_precision(si.I) == GLE_PREC_FP64 &&
                        (((flags & GLE_FLAG_REDUCE_RESCALE) && si.rescale_every > 2) ||
                         ((flags & GLE_FLAG_BAIL_OUT) && si.rescale_every <= 2));
    if (thrown) {
      // Tree::Score's retry: drop every cached CLA (they were built with the old
      // rescaleEvery) and run the step again with rescaleEvery - 2.
      tb.release_all();
      si.drop_clas(pool);
      si.rescale_every -= 2;
      if (si.rescale_every < 2) {
        fail_all(in_tick, set_error(GLE_ERR_RESCALE_FATAL, "Problem with rescaling in serving."));
        return deferred;
      }
      continue;
    }
    for (uint64_t id : tb.fresh_order) {
      const TickNew &f = tb.fresh.at(id);
      si.cla[id] = {f.slot, f.rank, sv.tick};
    }
    for (size_t i = 0; i < in_tick.size(); i++) advance(*in_tick[i], &res[3 * i]);
    return deferred;
  }
}

// CLA slots a request can hold in one tick: a step evaluates one edge and
// touches at most the N - 2 directional CLAs that point at it (cached or
// new).  A branch-length pass keeps room for both directions so that its
// CLAs survive between Newton steps.
static int working_set(const Req &r) { return r.kind == kReqScore ? r.N - 2 : 2 * (r.N - 2) + 4; }

static void worker_loop(Server *sv) {
  std::vector<std::pair<ServeInstance *, std::vector<Req *>>> by_inst;
  for (;;) {
    std::vector<Req *> work;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->cv_work.wait(lk, [&] { return sv->stop || !sv->pending.empty() || !sv->active.empty(); });
      if (sv->stop) return;
      if (sv->active.empty() && sv->batch_window_us > 0) {
        // idle until now: let a burst of submissions gather into the first
        // tick, until no request has arrived for a fifth of the window (at
        // least 20 us) or the window has passed
        using clock = std::chrono::steady_clock;
        const auto deadline = clock::now() + std::chrono::microseconds(sv->batch_window_us);
        const auto gap = std::chrono::microseconds(std::max(20, sv->batch_window_us / 5));
        size_t seen = sv->pending.size();
        for (;;) {
          const auto until = std::min(deadline, clock::now() + gap);
          sv->cv_work.wait_until(lk, until, [&] { return sv->stop || sv->pending.size() != seen; });
          if (sv->stop) return;
          if (sv->pending.size() == seen || clock::now() >= deadline) break;
          seen = sv->pending.size();
        }
      }
      // admission by HBM budget: FIFO while the working sets fit the instance's CLA slots
      while (!sv->pending.empty()) {
        Req *r = sv->pending.front();
        ServeInstance &si = sv->insts[r->inst];
        const int cap = instance_cla_slots(si.I);
        if (si.reserved + r->need > cap && !sv->active.empty()) break;
        si.reserved += r->need;
        sv->pending.pop_front();
        sv->active.push_back(r);
      }
      work = sv->active;
      sv->tick++;
      // the instances are looked up here, under the lock that submissions
      // insert under (map references stay valid across later inserts)
      for (Req *r : work) {
        ServeInstance *si = &sv->insts[r->inst];
        auto it = std::find_if(by_inst.begin(), by_inst.end(), [&](const auto &e) { return e.first == si; });
        if (it == by_inst.end()) by_inst.push_back({si, {r}});
        else it->second.push_back(r);
      }
    }
    double host_ms = 0.0, device_ms = 0.0;
    int deferred = 0;
    for (auto &kv : by_inst) {
      // requests deferred last time go first, so none waits behind the same others forever
      std::stable_sort(kv.second.begin(), kv.second.end(), [](const Req *a, const Req *b) { return a->waits > b->waits; });
      deferred += run_tick(*sv, *kv.first, kv.second, host_ms, device_ms);
    }
    by_inst.clear();
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->ticks++;
      sv->host_ms += host_ms;
      sv->device_ms += device_ms;
      sv->batched_requests += long(work.size()) - deferred;
      sv->deferred += deferred;
      std::vector<Req *> still;
      const auto now = std::chrono::steady_clock::now();
      for (Req *r : sv->active) {
        if (r->done) {
          r->t_done = now;
          r->finished = true;
          sv->insts[r->inst].reserved -= r->need;
          sv->latency_ms.push_back(std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count());
        } else
          still.push_back(r);
      }
      sv->active.swap(still);
    }
    sv->cv_done.notify_all();
  }
}

static Server *get_server(int h) {
  std::lock_guard<std::mutex> lk(g_servers_mu);
  if (h < 0 || h >= int(g_servers.size())) return nullptr;
  return g_servers[h].get();
}

}  // namespace gle

using namespace gle;

extern "C" int gle_serve_create(int *server) {
  if (!server) return set_error(GLE_ERR_ARG, "null argument");
  auto sv = std::make_unique<Server>();
  Server *raw = sv.get();
  raw->worker = std::thread(worker_loop, raw);
  std::lock_guard<std::mutex> lk(g_servers_mu);
  g_servers.push_back(std::move(sv));
  *server = int(g_servers.size()) - 1;
  return GLE_SUCCESS;
}

extern "C" int gle_serve_timing(int h, double *host_ms, double *device_ms, int reset) {
  Server *sv = get_server(h);
  if (!sv || !host_ms || !device_ms) return set_error(GLE_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  *host_ms = sv->host_ms;
  *device_ms = sv->device_ms;
  if (reset) sv->host_ms = sv->device_ms = 0.0;
  return GLE_SUCCESS;
}

extern "C" int gle_serve_set_batch_window(int h, int microseconds) {
  Server *sv = get_server(h);
  if (!sv || microseconds < 0) return set_error(GLE_ERR_ARG, "bad server handle or window");
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->batch_window_us = microseconds;
  return GLE_SUCCESS;
}

extern "C" int gle_serve_destroy(int h) {
  Server *sv = get_server(h);
  if (!sv) return set_error(GLE_ERR_ARG, "bad server handle %d", h);
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  sv->cv_work.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  {
    std::lock_guard<std::mutex> pk(pool_mutex());
    for (auto &kv : sv->insts) {
      if (get_instance(kv.second.h) != kv.second.I) continue;  // destroyed: its pool went with it
      ClaPool *pool = pool_for(kv.second.I);
      for (auto &c : kv.second.cla)
        if (--pool->refs[c.second.slot] == 0) pool->free_slot(c.second.slot);
      for (auto &p : kv.second.pm) pool->pm_free.push_back(p.second.slot);
    }
  }
  std::lock_guard<std::mutex> lk(g_servers_mu);
  g_servers[h].reset();
  return GLE_SUCCESS;
}

extern "C" int gle_serve_submit(int h, int inst, const int *parent, const double *blen, int kind, double precision,
                                int *req) {
  Server *sv = get_server(h);
  if (!sv) return set_error(GLE_ERR_ARG, "bad server handle %d", h);
  Instance *I = get_instance(inst);
  if (!I || !parent || !blen || !req || (kind != kReqScore && kind != kReqOptimize) ||
      (kind == kReqOptimize && !(precision > 0.0)))
    return set_error(GLE_ERR_ARG, "bad request");
  auto r = std::make_unique<Req>();
  r->inst = inst;
  r->kind = kind;
  r->N = instance_tips(I);
  r->prec = precision;
  r->parent.assign(parent, parent + 2 * r->N - 2);
  r->blen.assign(blen, blen + 2 * r->N - 2);
  int rc = tree_topology(*r);
  if (rc) return rc;
  r->need = working_set(*r);
  if (kind == kReqOptimize) {
    // OptimizeAllBranches order: preorder from each root child (optimization.cpp:97-105)
    std::vector<int> stack;
    for (int j = r->nnb[0] - 1; j >= 0; j--) stack.push_back(r->nb[0][j]);
    while (!stack.empty()) {
      const int n = stack.back();
      stack.pop_back();
      r->order.push_back(n);
      for (int j = r->nnb[n] - 1; j >= 0; j--)
        if (r->nb[n][j] != r->parent[n]) stack.push_back(r->nb[n][j]);
    }
    r->scoring = false;
    r->newton.start(r->blen[r->order[0]], precision);
  }
  r->t_submit = std::chrono::steady_clock::now();
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    ServeInstance &si = sv->insts[inst];
    if (!si.I) {
      si.h = inst;
      si.I = I;
      si.rescale_every = instance_precision(I) == GLE_PREC_FP64 ? 16 : 1;
      si.gen = instance_generation(I);
    }
    if (r->need > instance_cla_slots(I))
      return set_error(GLE_ERR_NOMEM, "request needs %d CLA slots, instance has %d", r->need,
                       instance_cla_slots(I));
    // a step may need a P(t) for every one of its 2N - 3 branch lengths
    if (instance_pmat_slots(I) < 2 * r->N - 3)
      return set_error(GLE_ERR_NOMEM, "request needs %d transition-matrix slots, instance has %d", 2 * r->N - 3,
                       instance_pmat_slots(I));
    r->id = sv->next_id++;
    *req = r->id;
    sv->pending.push_back(r.get());
    sv->reqs[r->id] = std::move(r);
  }
  sv->cv_work.notify_one();
  return GLE_SUCCESS;
}

extern "C" int gle_serve_poll(int h, int req, int *done) {
  Server *sv = get_server(h);
  if (!sv || !done) return set_error(GLE_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(req);
  if (it == sv->reqs.end()) return set_error(GLE_ERR_ARG, "unknown request %d", req);
  *done = it->second->finished ? 1 : 0;
  return GLE_SUCCESS;
}

extern "C" int gle_serve_wait(int h, int req, double *lnl, double *blen_out, double *latency_ms) {
  Server *sv = get_server(h);
  if (!sv || !lnl) return set_error(GLE_ERR_ARG, "bad argument");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(req);
  if (it == sv->reqs.end()) return set_error(GLE_ERR_ARG, "unknown request %d", req);
  Req *r = it->second.get();
  sv->cv_done.wait(lk, [&] { return r->finished; });
  const int rc = r->rc;
  *lnl = r->lnl;
  if (blen_out) std::copy(r->blen.begin(), r->blen.end(), blen_out);
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  sv->reqs.erase(it);
  return rc;
}

extern "C" int gle_serve_stats(int h, long long *completed, double *p50_ms, double *p99_ms, double *mean_batch,
                               int reset) {
  Server *sv = get_server(h);
  if (!sv || !completed || !p50_ms || !p99_ms || !mean_batch) return set_error(GLE_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  std::vector<double> v = sv->latency_ms;
  *completed = (long long)v.size();
  std::sort(v.begin(), v.end());
  auto pct = [&](double q) {
    if (v.empty()) return 0.0;
    const size_t i = std::min(v.size() - 1, size_t(std::ceil(q * double(v.size()))) - (q > 0 ? 1 : 0));
    return v[i];
  };
  *p50_ms = pct(0.50);
  *p99_ms = pct(0.99);
  *mean_batch = sv->ticks ? double(sv->batched_requests) / double(sv->ticks) : 0.0;
  if (reset) {
    sv->latency_ms.clear();
    sv->ticks = 0;
    sv->batched_requests = 0;
  }
  return GLE_SUCCESS;
}
