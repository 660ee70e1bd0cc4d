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
//  * Message cache (codon instances).  A CLA is the product of its two
//    children's messages P(t) . x; a message is named by (child subtree id,
//    branch-length bits) and so does not depend on where the child hangs.
//    The CLA kernel keeps the messages it forms in free slots, and a later CLA
//    whose child has a kept message reads it instead of contracting the child
//    again: a CLA on the path of a move has one child off the path, and a
//    subtree an NNI or SPR moved keeps its message, so most new codon CLAs of
//    an offspring need one 61 x 61 contraction per category or none instead
//    of two.  Messages take only free slots and are evicted before CLAs.
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
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <set>
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
static SubKey sub_key(uint64_t ix, double lx, uint64_t iy, double ly) {
  SubKey k{ix, dbits(lx), iy, dbits(ly)};
  if (k.x > k.y || (k.x == k.y && k.lx > k.ly)) {
    std::swap(k.x, k.y);
    std::swap(k.lx, k.ly);
  }
  return k;
}
// The exact name of a message P(t) . x: (0, 0, child id, length bits), which
// no subtree name equals (a subtree has two nonzero child ids).
static SubKey msg_key(uint64_t ic, double len) { return SubKey{0, 0, ic, dbits(len)}; }
// Internal subtree ids start above every tip number.
constexpr uint64_t kFirstSubtreeId = uint64_t(1) << 40;
// Interned names are kept while their CLAs come and go; past this many the
// table and the CLA cache are dropped together (a rare full refill).
constexpr size_t kInternCap = size_t(1) << 20;

// Interned subtree names: an open-addressing table (linear probing,
// power-of-two capacity, at most half full); id 0 marks an empty cell.
class NameTable {
  struct Cell {
    SubKey k;
    uint64_t id;
  };
  std::vector<Cell> t_;
  size_t n_ = 0, mask_ = 0;
  static size_t hash(const SubKey &k) {
    return size_t(mix64(k.x * 0x9E3779B97F4A7C15ull ^ k.lx) ^ mix64(k.y * 0xC2B2AE3D27D4EB4Full + k.ly));
  }
  void grow() {
    std::vector<Cell> old;
    old.swap(t_);
    t_.assign(std::max<size_t>(1024, 2 * old.size()), Cell{{0, 0, 0, 0}, 0});
    mask_ = t_.size() - 1;
    for (const Cell &c : old)
      if (c.id)
        for (size_t i = hash(c.k) & mask_;; i = (i + 1) & mask_)
          if (!t_[i].id) {
            t_[i] = c;
            break;
          }
  }

 public:
  size_t size() const { return n_; }
  void clear() {
    std::vector<Cell>().swap(t_);
    n_ = mask_ = 0;
  }
  uint64_t find(const SubKey &k) const {
    if (t_.empty()) return 0;
    for (size_t i = hash(k) & mask_;; i = (i + 1) & mask_) {
      if (!t_[i].id) return 0;
      if (t_[i].k == k) return t_[i].id;
    }
  }
  uint64_t intern(const SubKey &k, uint64_t &next) {
    if (2 * (n_ + 1) > t_.size()) grow();
    for (size_t i = hash(k) & mask_;; i = (i + 1) & mask_) {
      if (!t_[i].id) {
        t_[i] = Cell{k, next};
        n_++;
        return next++;
      }
      if (t_[i].k == k) return t_[i].id;
    }
  }
};

struct CacheEntry {
  int slot = -1, rank = 0;
  uint64_t last = 0;
  std::list<uint64_t>::iterator pos;  // place in the cache's use order
};

// A device-slot cache in least-recently-used order: a hit moves its entry to
// the back, so eviction takes the front (O(1) instead of a scan).
struct LruCache {
  std::unordered_map<uint64_t, CacheEntry> map;
  std::list<uint64_t> order;
  CacheEntry *find(uint64_t key) {
    auto it = map.find(key);
    return it == map.end() ? nullptr : &it->second;
  }
  bool has(uint64_t key) const { return map.count(key) != 0; }
  void touch(CacheEntry *e, uint64_t tick) {
    e->last = tick;
    order.splice(order.end(), order, e->pos);
  }
  void put(uint64_t key, int slot, int rank, uint64_t tick) {
    auto it = map.find(key);
    if (it != map.end()) order.erase(it->second.pos);
    order.push_back(key);
    CacheEntry &e = map[key];
    e.slot = slot;
    e.rank = rank;
    e.last = tick;
    e.pos = std::prev(order.end());
  }
  // slot of the least recently used entry not used in tick `tick`, removed; -1 if none
  int evict(uint64_t tick) {
    if (order.empty()) return -1;
    auto it = map.find(order.front());
    if (it->second.last >= tick) return -1;
    const int s = it->second.slot;
    order.pop_front();
    map.erase(it);
    return s;
  }
  template <typename F>
  void clear(F &&release) {
    for (auto &kv : map) release(kv.second.slot);
    map.clear();
    order.clear();
  }
};

struct ServeInstance {
  int h = -1;
  Instance *I = nullptr;
  int rescale_every = 16;
  NameTable intern;  // subtree name -> id
  uint64_t next_id = kFirstSubtreeId;
  LruCache cla;  // subtree id -> CLA slot
  LruCache msg;  // message name -> slot holding P(t) . x of that child (rank = the child's)
  LruCache pm;   // blen bits -> P, P', P'' slot
  bool msgs = false;  // codon instance: keep and read messages
  int reserved = 0;                              // working-set slots of admitted requests
  int next_stage = 0;                            // staging buffers of the next part queued
  int parts_in_flight = 0;                       // parts queued and not yet collected (worker only)
  unsigned long long gen = 0;                    // instance generation the caches belong to
  unsigned long long epoch = 0;                  // bumped whenever the CLA cache is dropped
  unsigned long long names = 0;                  // bumped whenever the name table is cleared

  uint64_t intern_id(const SubKey &k) { return intern.intern(k, next_id); }
  uint64_t find_id(const SubKey &k) const { return intern.find(k); }
  // every cached CLA and message back to the pool (caller holds pool_mutex)
  void drop_clas(ClaPool *pool) {
    auto release = [pool](int s) {
      if (--pool->refs[s] == 0) pool->free_slot(s);
    };
    cla.clear(release);
    msg.clear(release);
  }
  void drop_pmats(ClaPool *pool) {
    pm.clear([pool](int s) { pool->pm_free.push_back(s); });
  }
};

enum { kReqScore = 0, kReqOptimize = 1 };

// per-request memo of subtree ids: kUnset = not looked up yet, 0 = no id yet.
// An id is a name, valid until the name table is cleared, so the memo is
// kept across a request's ticks; a branch-length change forgets the ids
// whose subtree holds that branch, and a 0 is looked up again each tick.
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
  bool in_flight = false; // its step is queued on the device (worker only)
  bool done = false;      // set by the worker while it owns the request
  bool finished = false;  // published under Server::mu once the worker let go
  std::chrono::steady_clock::time_point t_submit, t_done;
  std::vector<std::array<uint64_t, 3>> sid;  // X(a|b) id memo, sid[a][index of b among a's neighbours]
  std::vector<std::array<uint64_t, 3>> shape;  // X(a|b) content hash (the tick's sharing count), same indexing
  unsigned long long sid_names = ~0ull;      // the instance's name-table generation sid belongs to
  // per-tick scratch
  int edge_a = 0, edge_b = 0;
  bool derivs = false;
};

struct Server {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req *> pending;
  std::vector<Req *> active;
  std::unordered_map<int, std::unique_ptr<Req>> reqs;
  // requests their clients collected, freed by the worker while a tick it
  // queued runs (freeing a generation's requests inline cost its client
  // ~50 us after the last result)
  std::vector<std::unique_ptr<Req>> collected;
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
  // CLA work: ops, children contracted on the device (internal children not
  // read from a message), messages read and kept
  long long n_ops = 0, n_contract = 0, n_msg_read = 0, n_msg_keep = 0;
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
  // every node reaches the root: walk up until a node known to reach it,
  // then mark the walked path (linear overall; a cycle exceeds nn steps)
  std::vector<char> reaches(size_t(r.nn), 0);
  reaches[0] = 1;
  for (int n = 1; n < r.nn; n++) {
    int x = n, steps = 0;
    while (!reaches[size_t(x)] && steps++ <= r.nn) x = r.parent[x];
    if (!reaches[size_t(x)]) return set_error(GLE_ERR_ARG, "node %d does not reach the root", n);
    for (int y = n; !reaches[size_t(y)]; y = r.parent[y]) reaches[size_t(y)] = 1;
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
// can be cached), memoised in r.sid.
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

// Scores in a tick from which their edge choice weighs subtrees by sharing.
constexpr int kShareMinScores = 8;

// Counts of 64-bit keys (nonzero), open addressing with linear probing.
class KeyCount {
  std::vector<std::pair<uint64_t, int>> t_;
  size_t mask_ = 0;

 public:
  void reset(size_t n) {
    size_t cap = 64;
    while (cap < 2 * n) cap <<= 1;
    t_.assign(cap, {0, 0});
    mask_ = cap - 1;
  }
  void add(uint64_t k) {
    if (t_.empty()) return;
    for (size_t i = size_t(mix64(k)) & mask_;; i = (i + 1) & mask_)
      if (t_[i].first == k || t_[i].first == 0) {
        t_[i].first = k;
        t_[i].second++;
        return;
      }
  }
  int get(uint64_t k) const {
    if (t_.empty()) return 0;
    for (size_t i = size_t(mix64(k)) & mask_;; i = (i + 1) & mask_) {
      if (t_[i].first == k) return t_[i].second;
      if (t_[i].first == 0) return 0;
    }
  }
};

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
  std::vector<char> pm_new;  // per pm_idx entry: 1 a new slot, 0 a cached one refreshed with P', P''
  bool derivs = false;       // the tick computes P' and P'' as well
  // per op (parallel to ops): each child contracted on the device, named by
  // (id, length) with its rescale rank; id 0 = a tip or a message read
  struct Kid {
    uint64_t id = 0;
    double len = 0.0;
    int rank = 0;
    bool cached = false;  // its CLA was in the cache before this tick
  };
  std::vector<std::array<Kid, 2>> kids;
  std::unordered_map<uint64_t, int> uses;  // per subtree id: children contracted from it this tick
  KeyCount sharing;  // per subtree content hash: the tick's scores whose trees contain it
  std::vector<std::pair<uint64_t, TickNew>> fresh_msgs;  // messages kept this tick

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
    kids.resize(m.ops);
    for (size_t i = m.pm; i < pm_idx.size(); i++) {
      pm_fresh.erase(dbits(pm_len[i]));
      if (pm_new[i]) pool->pm_free.push_back(pm_idx[i]);
    }
    pm_idx.resize(m.pm);
    pm_len.resize(m.pm);
    pm_new.resize(m.pm);
    err = GLE_SUCCESS;
  }
  // every new CLA, message and P(t) slot of the tick back to the pool (a failed tick)
  void release_all() {
    for (auto &fm : fresh_msgs)
      if (--pool->refs[fm.second.slot] == 0) pool->free_slot(fm.second.slot);
    fresh_msgs.clear();
    rollback(Mark{0, 0, 0});
  }

  int take_slot() {
    while (!pool->has_free()) {  // evict least recently used entries not used this tick, messages first
      int s = si->msg.evict(tick);
      if (s < 0) s = si->cla.evict(tick);
      if (s < 0) return -1;
      if (--pool->refs[s] == 0) pool->free_slot(s);
    }
    const int s = pool->take_slot();
    pool->refs[s] = 1;
    return s;
  }

  // After every request of the tick is laid out: a child the tick contracts
  // keeps its message in a free slot (or one an unused message frees) when
  // its subtree is shared -- cached before this tick (the unmodified part of
  // a parent), or contracted more than once in it -- or when a score edge of
  // the tick will read it (`want`: (subtree id, length bits)), and the
  // message is not kept already.  An offspring's own new subtrees are seldom
  // met again, and their messages would only cost writes.  Messages never
  // take a slot from a CLA, so admission's working sets are unaffected.
  void keep_messages(const std::vector<std::pair<uint64_t, uint64_t>> &want) {  // want: sorted
    std::unordered_map<uint64_t, int> kept;
    for (size_t i = 0; i < ops.size(); i++)
      for (int k = 0; k < 2; k++) {
        const Kid &kd = kids[i][k];
        if (!kd.id ||
            !(kd.cached || uses[kd.id] > 1 ||
              std::binary_search(want.begin(), want.end(), std::make_pair(kd.id, dbits(kd.len)))))
          continue;
        const uint64_t name = si->intern_id(msg_key(kd.id, kd.len));
        if (si->msg.has(name) || kept.count(name)) continue;
        if (!pool->has_free()) {
          const int v = si->msg.evict(tick);
          if (v < 0) return;
          if (--pool->refs[v] == 0) pool->free_slot(v);
          if (!pool->has_free()) return;
        }
        const int s = pool->take_slot();
        pool->refs[s] = 1;
        (k ? ops[i].second.m2 : ops[i].second.m1) = s;
        kept[name] = s;
        fresh_msgs.push_back({name, TickNew{s, kd.rank, 0}});
      }
  }

  // A child of a new CLA: a kept message of it (codon instances), else its
  // CLA slot or tip code as resolve() gives them
  int child(Req &r, int c, int a, int &rank, int &level, bool &msg) {
    msg = false;
    if (si->msgs && !req_tip(r, c)) {
      const uint64_t ic = known_id(r, *si, c, a);
      const uint64_t name = ic ? si->find_id(msg_key(ic, req_len(r, a, c))) : 0;
      if (name)
        if (CacheEntry *e = si->msg.find(name)) {
          si->msg.touch(e, tick);
          rank = e->rank;
          level = -1;
          msg = true;
          return e->slot;
        }
    }
    const int s = resolve(r, c, a, rank, level);
    if (!err && !req_tip(r, c)) uses[r.sid[c][req_nbk(r, c, a)]]++;
    return s;
  }

  // P(t) slot of a branch length; need_derivs when an edge's Newton step
  // reads P' and P'' from it.  The tick computes the derivative matrices
  // only when one of its steps needs them (P(t) cache entries remember
  // whether they hold them, in CacheEntry::rank).
  int pm_slot(double len, bool need_derivs = false) {
    const uint64_t key = dbits(len);
    if (need_derivs) derivs = true;
    if (CacheEntry *e = si->pm.find(key)) {
      si->pm.touch(e, tick);
      if (!need_derivs || e->rank) return e->slot;
      if (!pm_fresh.count(key)) {  // cached without P', P'': recompute it in place
        pm_fresh[key] = e->slot;
        pm_idx.push_back(e->slot);
        pm_len.push_back(len);
        pm_new.push_back(0);
      }
      return e->slot;
    }
    auto f = pm_fresh.find(key);
    if (f != pm_fresh.end()) return f->second;
    if (pool->pm_free.empty()) {
      const int v = si->pm.evict(tick);
      if (v < 0) return -1;
      pool->pm_free.push_back(v);
    }
    const int s = pool->pm_free.back();
    pool->pm_free.pop_back();
    pm_fresh[key] = s;
    pm_idx.push_back(s);
    pm_len.push_back(len);
    pm_new.push_back(1);
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
      if (CacheEntry *e = si->cla.find(known)) {
        si->cla.touch(e, tick);
        rank = e->rank;
        level = -1;
        return e->slot;
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
    bool g1, g2;
    DevClaOp op;
    op.c1 = child(r, x, a, rx, lx, g1);
    if (err) return 0;
    op.c2 = child(r, y, a, ry, ly, g2);
    if (err) return 0;
    const uint64_t ix = req_tip(r, x) ? uint64_t(x) : r.sid[x][req_nbk(r, x, a)];
    const uint64_t iy = req_tip(r, y) ? uint64_t(y) : r.sid[y][req_nbk(r, y, a)];
    const double lenx = req_len(r, a, x), leny = req_len(r, a, y);
    const uint64_t id = si->intern_id(sub_key(ix, lenx, iy, leny));
    r.sid[a][req_nbk(r, a, b)] = id;
    op.pm1 = g1 ? 0 : pm_slot(lenx);  // a message child needs no P(t)
    op.pm2 = op.pm1 < 0 ? -1 : g2 ? 0 : pm_slot(leny);
    const int slot = op.pm2 < 0 ? -1 : take_slot();
    if (slot < 0) {
      err = set_error(GLE_ERR_NOMEM, "serving cache exhausted: raise n_cla_slots / n_pmat_slots");
      return 0;
    }
    rank = 2 + rx + ry;
    op.flags = (g1 ? kOpMsg1 : 0) | (g2 ? kOpMsg2 : 0);
    if (rank >= si->rescale_every) {
      op.flags |= kOpRescale;
      rank = 0;
    }
    op.dest = slot;
    level = std::max(lx, ly) + 1;
    fresh[id] = {slot, rank, level};
    fresh_order.push_back(id);
    ops.push_back({level, op});
    std::array<Kid, 2> kd;
    if (!g1 && !req_tip(r, x)) kd[0] = Kid{ix, lenx, rx, lx < 0 && !fresh.count(ix)};
    if (!g2 && !req_tip(r, y)) kd[1] = Kid{iy, leny, ry, ly < 0 && !fresh.count(iy)};
    kids.push_back(kd);
    return slot;
  }
};

// Content hash of X(a|b) in request r (tips, topology, branch-length bits),
// memoised in r.shape; a key for counting how many of a tick's requests
// contain the same subtree, not a name (names are interned exactly).
static uint64_t shape_hash(Req &r, int a, int b) {
  if (req_tip(r, a)) return mix64(uint64_t(a) * 0x9E3779B97F4A7C15ull);
  uint64_t &m = r.shape[a][req_nbk(r, a, b)];
  if (m) return m;
  int x, y;
  req_children(r, a, b, x, y);
  uint64_t px = mix64(shape_hash(r, x, a) ^ dbits(req_len(r, a, x)) * 0xC2B2AE3D27D4EB4Full);
  uint64_t py = mix64(shape_hash(r, y, a) ^ dbits(req_len(r, a, y)) * 0xC2B2AE3D27D4EB4Full);
  if (px > py) std::swap(px, py);
  m = mix64(px * 0x165667B19E3779F9ull + py) | 1;
  return m;
}

// Scoring edge of a request.  lnL is the same on every edge (the pruning
// algorithm's root invariance; the reference's RunTests checks it,
// population.cpp:924-954), so a score is taken where the fewest new CLAs are
// needed: next to the perturbation, for an offspring whose parent's CLAs are
// cached.  A missing CLA counts 1/k when k of the tick's scores contain the
// same subtree (it is computed once for all of them), so an offspring does
// not avoid the parent's uncached subtrees that the rest of the generation
// needs too.  Ties keep the reference's edge (root, middle child).  The
// choice depends on what is cached, so a result alone and in a mix agree to
// the form's rounding, not bit for bit.
// Weighted CLAs missing for X(a|b), memoised in memo (< 0: not yet)
static double missing_clas(Req &r, const TickBuilder &tb, std::vector<std::array<double, 3>> &memo, int a, int b) {
  if (req_tip(r, a)) return 0.0;
  const int k = req_nbk(r, a, b);
  double &m = memo[a][k];
  if (m >= 0.0) return m;
  const uint64_t id = known_id(r, *tb.si, a, b);
  if (id && (tb.si->cla.has(id) || tb.fresh.count(id))) return m = 0.0;
  // (a lone score in its tick has no sharing counts: each missing CLA is 1)
  const int shared = r.shape.size() == size_t(r.nn) ? tb.sharing.get(shape_hash(r, a, b)) : 1;
  double c = 1.0 / std::max(1, shared);
  for (int j = 0; j < r.nnb[a]; j++)
    if (r.nb[a][j] != b) c += missing_clas(r, tb, memo, r.nb[a][j], a);
  return m = c;
}

static void choose_score_edge(Req &r, const TickBuilder &tb) {
  static thread_local std::vector<std::array<double, 3>> memo;
  memo.assign(r.nn, {-1.0, -1.0, -1.0});
  auto cost = [&](int a, int b) { return missing_clas(r, tb, memo, a, b); };
  r.edge_a = 0;
  r.edge_b = r.nb[0][1];
  double best = cost(r.edge_a, r.edge_b) + cost(r.edge_b, r.edge_a);
  for (int n = 1; n < r.nn && best > 0.0; n++) {
    const int a = r.parent[n];
    const double c = cost(n, a) + cost(a, n);
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

// Forget the ids that name branch n's length: X(x|y) for every x on the far
// side of y from the branch (x's subtree away from y holds it), found by
// walking away from each end of the branch.
static void forget_ids_through(Req &r, int n) {
  if (r.sid.size() != size_t(r.nn)) return;
  std::vector<std::pair<int, int>> stack = {{n, r.parent[n]}, {r.parent[n], n}};
  while (!stack.empty()) {
    const auto [x, from] = stack.back();
    stack.pop_back();
    if (req_tip(r, x)) continue;
    for (int j = 0; j < r.nnb[x]; j++) {
      const int y = r.nb[x][j];
      if (y == from) continue;
      r.sid[x][j] = kUnset;
      stack.push_back({y, x});
    }
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
  if (act != NewtonBranch::kFinish && next != r.blen[nd]) {
    r.blen[nd] = next;
    forget_ids_through(r, nd);
  }
  if (act != NewtonBranch::kContinue) {
    r.improve += r.newton.total;
    if (++r.bi < r.order.size()) r.newton.start(r.blen[r.order[r.bi]], r.prec);
    else r.scoring = true;
  }
}

// One instance's share of a tick: the requests it carries, the CLAs it
// computes (named, so a failed tick can take them back out of the cache) and
// the cache epoch it was built on.
// Up to two parts of an instance are in flight, so the host builds the next
// while the device runs the current one.  Every part's work is on its
// instance's own stream, so a later part that reads CLAs an earlier one
// writes, or reuses a slot an earlier one read, runs after it; parts of
// different instances run and are collected independently.
struct TickPart {
  ServeInstance *si = nullptr;
  std::vector<Req *> reqs;
  std::vector<uint64_t> fresh_ids;
  unsigned long long epoch = 0;
  int n_edges = 0;
  int stage = 0;  // the instance's staging buffers it was queued with
};

// Builds one instance's part of a tick from `batch` and queues it on the
// device (no wait).  A request whose new CLAs or P(t) do not fit beside those
// the part already holds waits for a later tick, unchanged and holding
// nothing; one that does not fit on its own fails.  The part's new CLAs and
// P(t) enter the caches at once, so the next tick can use them.
// CLA work of the ticks built, published under Server::mu (see Server)
struct TickWork {
  long long ops = 0, contract = 0, msg_read = 0, msg_keep = 0;
};

static int build_part(Server &sv, ServeInstance &si, std::vector<Req *> &batch, uint64_t tick, int stage,
                      TickPart &part, double &host_ms, TickWork &work) {
  using clock = std::chrono::steady_clock;
  const auto t_build = clock::now();
  std::vector<DevClaOp> ops;
  std::vector<int> level_start;
  std::vector<DevEdgeOp> eops;
  int deferred = 0;
  std::lock_guard<std::mutex> pk(pool_mutex());
  ClaPool *pool = pool_for(si.I);
  if (si.gen != instance_generation(si.I) || si.intern.size() > kInternCap) {
    // data or model changed (or the name table is full): drop both caches
    si.drop_clas(pool);
    si.drop_pmats(pool);
    si.intern.clear();
    si.names++;
    si.gen = instance_generation(si.I);
    si.epoch++;
  }
  TickBuilder tb{&si, pool, tick, {}, {}, {}, {}, {}, {}, GLE_SUCCESS};
  {
    // room for every request's directional CLAs, so a cold tick's maps do
    // not rehash as they grow
    const size_t cap = batch.size() * size_t(instance_tips(si.I));
    tb.fresh.reserve(cap);
    tb.uses.reserve(cap);
    tb.pm_fresh.reserve(cap);
    tb.fresh_order.reserve(cap);
    tb.ops.reserve(cap);
    tb.kids.reserve(cap);
  }
  {
    // how many of the tick's scores contain each directional subtree: for a
    // tick of many scores (a generation's offspring), where the counts steer
    // edges towards subtrees computed once for all; a small tick (a request
    // stream, Newton steps) skips the pass and its cost
    int scores = 0;
    for (const Req *r : batch) scores += r->scoring ? 1 : 0;
    if (scores >= kShareMinScores) {
      tb.sharing.reset(size_t(scores) * 2 * size_t(instance_tips(si.I)) * 2);
      for (Req *r : batch) {
        if (!r->scoring) continue;
        r->shape.assign(r->nn, {0, 0, 0});
        for (int a = 0; a < r->nn; a++)
          if (!req_tip(*r, a))
            for (int k = 0; k < r->nnb[a]; k++) tb.sharing.add(shape_hash(*r, a, r->nb[a][k]));
      }
    }
  }
  // each edge's two sides: subtree ids (0 for a tip) and the edge's length
  struct EdgeSides {
    uint64_t p, c;
    double len;
  };
  std::vector<EdgeSides> sides;
  part.si = &si;
  part.epoch = si.epoch;
  for (Req *r : batch) {
    const TickBuilder::Mark m = tb.mark();
    if (r->sid_names != si.names || r->sid.size() != size_t(r->nn)) {
      r->sid.assign(r->nn, {kUnset, kUnset, kUnset});
      r->sid_names = si.names;
    } else {
      for (auto &e : r->sid)  // another request may have named it since
        for (uint64_t &v : e)
          if (v == 0) v = kUnset;
    }
    next_edge(*r, tb);
    int ra, la, rb, lb;
    DevEdgeOp e;
    e.partial = tb.resolve(*r, r->edge_a, r->edge_b, ra, la);
    e.child = tb.err ? 0 : tb.resolve(*r, r->edge_b, r->edge_a, rb, lb);
    e.pm = tb.err ? -1 : tb.pm_slot(req_len(*r, r->edge_a, r->edge_b), r->derivs);
    e.derivs = r->derivs ? 1 : 0;
    if (!tb.err && e.pm < 0) tb.err = set_error(GLE_ERR_NOMEM, "serving cache exhausted: out of pmat slots");
    if (tb.err) {
      const int err = tb.err;
      tb.rollback(m);
      if (part.reqs.empty()) {  // does not fit even alone
        r->rc = err;
        r->done = true;
      } else {
        deferred++;
        r->waits++;
      }
      continue;
    }
    eops.push_back(e);
    sides.push_back({req_tip(*r, r->edge_a) ? 0 : r->sid[r->edge_a][req_nbk(*r, r->edge_a, r->edge_b)],
                     req_tip(*r, r->edge_b) ? 0 : r->sid[r->edge_b][req_nbk(*r, r->edge_b, r->edge_a)],
                     req_len(*r, r->edge_a, r->edge_b)});
    part.reqs.push_back(r);
    r->waits = 0;
  }
  if (part.reqs.empty()) {
    host_ms += std::chrono::duration<double, std::milli>(clock::now() - t_build).count();
    return deferred;
  }
  // Codon scores (a tick without derivatives): an edge whose child side, or
  // else whose partial side, a CLA op of the tick contracts at the edge's
  // length reads that message instead of contracting the side again (the
  // edge sum is symmetric in its two sides for a reversible model, so the
  // partial side's message serves with the sides swapped).
  const bool edge_msgs = si.msgs && !tb.derivs;
  std::vector<std::pair<uint64_t, uint64_t>> want;  // sorted
  if (edge_msgs) {
    std::vector<std::pair<uint64_t, uint64_t>> formed;  // (child id, length bits) the tick contracts
    formed.reserve(2 * tb.kids.size());
    for (const auto &kd : tb.kids)
      for (const auto &k : kd)
        if (k.id) formed.push_back({k.id, dbits(k.len)});
    std::sort(formed.begin(), formed.end());
    auto has = [&](const std::pair<uint64_t, uint64_t> &x) {
      return std::binary_search(formed.begin(), formed.end(), x);
    };
    for (const EdgeSides &e : sides) {
      const std::pair<uint64_t, uint64_t> c{e.c, dbits(e.len)}, p{e.p, dbits(e.len)};
      if (e.c && has(c)) want.push_back(c);
      else if (e.c && e.p && has(p)) want.push_back(p);
    }
    std::sort(want.begin(), want.end());
  }
  if (si.msgs) tb.keep_messages(want);
  if (edge_msgs) {
    std::unordered_map<uint64_t, int> fresh;
    for (const auto &fm : tb.fresh_msgs) fresh[fm.first] = fm.second.slot;
    auto msg_slot = [&](uint64_t id, double len) {
      const uint64_t name = id ? si.find_id(msg_key(id, len)) : 0;
      if (!name) return -1;
      const auto it = fresh.find(name);
      if (it != fresh.end()) return it->second;
      if (CacheEntry *c = si.msg.find(name)) {
        si.msg.touch(c, tick);
        return c->slot;
      }
      return -1;
    };
    for (size_t i = 0; i < eops.size(); i++) {
      const EdgeSides &e = sides[i];
      if (!e.c) continue;  // a tip child is a gather, not a contraction
      const int sc = msg_slot(e.c, e.len);
      const int sp = sc >= 0 || !e.p ? -1 : msg_slot(e.p, e.len);
      if (sc >= 0) {
        eops[i].child = sc;
      } else if (sp >= 0) {
        eops[i].partial = eops[i].child;
        eops[i].child = sp;
      } else {
        continue;
      }
      eops[i].msg = 1;
      work.msg_read++;
    }
  }
  for (size_t i = 0; i < tb.ops.size(); i++) {
    const DevClaOp &o = tb.ops[i].second;
    work.ops++;
    work.contract += (tb.kids[i][0].id != 0) + (tb.kids[i][1].id != 0);
    work.msg_read += ((o.flags & kOpMsg1) != 0) + ((o.flags & kOpMsg2) != 0);
    work.msg_keep += (o.m1 >= 0) + (o.m2 >= 0);
  }
  std::stable_sort(tb.ops.begin(), tb.ops.end(),
                   [](const std::pair<int, DevClaOp> &x, const std::pair<int, DevClaOp> &y) { return x.first < y.first; });
  int cur = -1;
  for (auto &lo : tb.ops) {
    if (lo.first != cur) {
      level_start.push_back(int(ops.size()));
      cur = lo.first;
    }
    ops.push_back(lo.second);
  }
  level_start.push_back(int(ops.size()));
  const int rc = engine_tick_enqueue(si.I, stage, tb.pm_idx.data(), tb.pm_len.data(), int(tb.pm_idx.size()),
                                     tb.derivs ? 1 : 0, ops.data(), int(ops.size()), level_start.data(),
                                     int(level_start.size()) - 1, eops.data(), int(eops.size()));
  if (rc) {
    tb.release_all();
    for (Req *r : part.reqs) {
      r->rc = rc;
      r->done = true;
    }
    part.reqs.clear();
  } else {
    for (size_t i = 0; i < tb.pm_idx.size(); i++) si.pm.put(dbits(tb.pm_len[i]), tb.pm_idx[i], tb.derivs, tick);
    for (uint64_t id : tb.fresh_order) {
      const TickNew &f = tb.fresh.at(id);
      si.cla.put(id, f.slot, f.rank, tick);
      part.fresh_ids.push_back(id);
    }
    for (auto &fm : tb.fresh_msgs) si.msg.put(fm.first, fm.second.slot, fm.second.rank, tick);
    part.n_edges = int(eops.size());
    for (Req *r : part.reqs) r->in_flight = true;
  }
  host_ms += std::chrono::duration<double, std::milli>(clock::now() - t_build).count();
  return deferred;
}

// Waits for one part and applies its results.  A part built on a cache the
// instance has since dropped is discarded (its requests run their step
// again).  A device error fails the part's requests and drops the caches;
// the reference's rescale throw (tree.cpp:3436-3449, FP64) drops the caches
// and lowers rescaleEvery by 2 (Tree::Score's retry, tree.cpp:4077-4090) and
// the requests run their step again.
static void collect_part(TickPart &part) {
  ServeInstance &si = *part.si;
  si.parts_in_flight--;
  std::vector<double> res(3 * size_t(std::max(part.n_edges, 1)));
  int flags = 0;
  const int rc = engine_tick_collect(si.I, part.stage, res.data(), &flags);
  for (Req *r : part.reqs) r->in_flight = false;
  const bool stale = part.epoch != si.epoch;
  const bool thrown = !rc && instance_precision(si.I) == GLE_PREC_FP64 &&
                      (((flags & GLE_FLAG_REDUCE_RESCALE) && si.rescale_every > 2) ||
                       ((flags & GLE_FLAG_BAIL_OUT) && si.rescale_every <= 2));
  if (rc || thrown) {
    std::lock_guard<std::mutex> pk(pool_mutex());
    if (!stale) {
      si.drop_clas(pool_for(si.I));
      si.epoch++;
    }
    if (thrown && !stale) si.rescale_every -= 2;
    const bool fatal = thrown && si.rescale_every < 2;
    if (rc || fatal) {
      const int e = rc ? rc : set_error(GLE_ERR_RESCALE_FATAL, "Problem with rescaling in serving.");
      for (Req *r : part.reqs) {
        r->rc = e;
        r->done = true;
      }
    }
    return;
  }
  if (stale) return;
  for (size_t i = 0; i < part.reqs.size(); i++) advance(*part.reqs[i], &res[3 * i]);
}

// CLA slots a request can hold in one tick: a step evaluates one edge and
// touches at most the N - 2 directional CLAs that point at it (cached or
// new).  A branch-length pass keeps room for both directions so that its
// CLAs survive between Newton steps.
static int working_set(const Req &r) { return r.kind == kReqScore ? r.N - 2 : 2 * (r.N - 2) + 4; }

// Admission by HBM budget: per instance in arrival order, while the working
// sets fit the instance's CLA slots (the first request of an idle instance
// always goes in).  A request that waits holds back later ones of its own
// instance only.  Caller holds sv.mu.
static void admit(Server &sv) {
  std::vector<int> blocked;
  for (auto it = sv.pending.begin(); it != sv.pending.end();) {
    Req *r = *it;
    if (std::find(blocked.begin(), blocked.end(), r->inst) != blocked.end()) {
      ++it;
      continue;
    }
    ServeInstance &si = sv.insts[r->inst];
    if (si.reserved + r->need > instance_cla_slots(si.I) && si.reserved > 0) {
      blocked.push_back(r->inst);
      ++it;
      continue;
    }
    si.reserved += r->need;
    sv.active.push_back(r);
    it = sv.pending.erase(it);
  }
}

static void worker_loop(Server *sv) {
  std::deque<TickPart> inflight;  // parts queued on the device, in the order they were queued
  for (;;) {
    std::vector<std::pair<ServeInstance *, std::vector<Req *>>> by_inst;
    bool stop = false;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      auto has_ready = [&] {
        for (Req *r : sv->active)
          if (!r->in_flight && !r->done) return true;
        return false;
      };
      sv->cv_work.wait(lk, [&] { return sv->stop || !sv->pending.empty() || !inflight.empty() || has_ready(); });
      stop = sv->stop;
      if (!stop && sv->active.empty() && inflight.empty() && sv->batch_window_us > 0) {
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
          if (sv->stop) break;
          if (sv->pending.size() == seen || clock::now() >= deadline) break;
          seen = sv->pending.size();
        }
        stop = sv->stop;
      }
      if (!stop) {
        admit(*sv);
        // An instance runs one part while its next is built and queued.  A
        // large burst (a generation's offspring) reaching an idle device
        // starts with a quarter of it, at most 8 requests, so the device
        // starts early and the rest is built while that part runs.
        int ready = 0;
        for (Req *r : sv->active) ready += !(r->in_flight || r->done);
        const int chunk = (inflight.empty() && ready >= 16) ? std::min((ready + 3) / 4, 8) : ready;
        int taken = 0;
        // the instances are looked up here, under the lock that submissions
        // insert under (map references stay valid across later inserts)
        for (Req *r : sv->active) {
          if (r->in_flight || r->done || taken >= chunk) continue;
          ServeInstance *si = &sv->insts[r->inst];
          if (si->parts_in_flight >= 2) continue;
          taken++;
          auto it = std::find_if(by_inst.begin(), by_inst.end(), [&](const auto &e) { return e.first == si; });
          if (it == by_inst.end()) by_inst.push_back({si, {r}});
          else it->second.push_back(r);
        }
        if (!by_inst.empty()) sv->tick++;
      }
    }
    if (stop) {  // let the queued work finish before the caches are released
      for (TickPart &p : inflight) collect_part(p);
      return;
    }
    double host_ms = 0.0, device_ms = 0.0;
    TickWork work;
    long long in_tick = 0;
    int deferred = 0;
    if (!by_inst.empty()) {
      const uint64_t tick = sv->tick;  // written by this thread only
      for (auto &kv : by_inst) {
        // requests deferred last time go first, so none waits behind the same others forever
        std::stable_sort(kv.second.begin(), kv.second.end(),
                         [](const Req *a, const Req *b) { return a->waits > b->waits; });
        TickPart part;
        part.stage = kv.first->next_stage;
        deferred += build_part(*sv, *kv.first, kv.second, tick, part.stage, part, host_ms, work);
        in_tick += long(part.reqs.size());
        if (!part.reqs.empty()) {
          kv.first->next_stage = (part.stage + 1) % kTickStages;
          kv.first->parts_in_flight++;
          inflight.push_back(std::move(part));
        }
      }
      if (in_tick) {
        // parts are queued: free the requests clients have collected while
        // the device runs them
        std::vector<std::unique_ptr<Req>> done_reqs;
        {
          std::lock_guard<std::mutex> lk(sv->mu);
          done_reqs.swap(sv->collected);
        }
      }
    } else if (!inflight.empty()) {
      // collect every part the device has finished, each instance's parts in
      // the order they were queued, so an instance whose part is done goes
      // on without waiting for other instances' larger parts; with none
      // finished, look again (new requests are picked up meanwhile)
      const auto t0 = std::chrono::steady_clock::now();
      std::vector<const ServeInstance *> waiting;  // instances with an unfinished earlier part
      bool any = false;
      for (auto it = inflight.begin(); it != inflight.end();) {
        const ServeInstance *si = it->si;
        if (std::find(waiting.begin(), waiting.end(), si) != waiting.end() || !engine_tick_ready(si->I, it->stage)) {
          waiting.push_back(si);
          ++it;
          continue;
        }
        collect_part(*it);
        it = inflight.erase(it);
        any = true;
      }
      if (!any) std::this_thread::yield();
      device_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      if (in_tick) sv->ticks++;
      sv->host_ms += host_ms;
      sv->n_ops += work.ops;
      sv->n_contract += work.contract;
      sv->n_msg_read += work.msg_read;
      sv->n_msg_keep += work.msg_keep;
      sv->device_ms += device_ms;
      sv->batched_requests += in_tick;
      sv->deferred += deferred;
      std::vector<Req *> still;
      const auto now = std::chrono::steady_clock::now();
      bool any_done = false;
      for (Req *r : sv->active) {
        if (r->done && !r->in_flight) {
          r->t_done = now;
          r->finished = true;
          any_done = true;
          sv->insts[r->inst].reserved -= r->need;
          sv->latency_ms.push_back(std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count());
        } else
          still.push_back(r);
      }
      sv->active.swap(still);
      if (any_done) sv->cv_done.notify_all();
    }
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

extern "C" int gle_serve_work(int h, long long *cla_ops, long long *contractions, long long *msg_reads,
                              long long *msg_keeps, int reset) {
  Server *sv = get_server(h);
  if (!sv || !cla_ops || !contractions || !msg_reads || !msg_keeps) return set_error(GLE_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  *cla_ops = sv->n_ops;
  *contractions = sv->n_contract;
  *msg_reads = sv->n_msg_read;
  *msg_keeps = sv->n_msg_keep;
  if (reset) sv->n_ops = sv->n_contract = sv->n_msg_read = sv->n_msg_keep = 0;
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
      kv.second.drop_clas(pool);
      kv.second.drop_pmats(pool);
    }
  }
  std::lock_guard<std::mutex> lk(g_servers_mu);
  g_servers[h].reset();
  return GLE_SUCCESS;
}

// A request for instance I (validated), not yet queued.
static int make_request(Instance *I, int inst, const int *parent, const double *blen, int kind, double precision,
                        std::unique_ptr<Req> &out) {
  if (!parent || !blen || (kind != kReqScore && kind != kReqOptimize) || (kind == kReqOptimize && !(precision > 0.0)))
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
  if (r->need > instance_cla_slots(I))
    return set_error(GLE_ERR_NOMEM, "request needs %d CLA slots, instance has %d", r->need, instance_cla_slots(I));
  // a step may need a P(t) for every one of its 2N - 3 branch lengths
  if (instance_pmat_slots(I) < 2 * r->N - 3)
    return set_error(GLE_ERR_NOMEM, "request needs %d transition-matrix slots, instance has %d", 2 * r->N - 3,
                     instance_pmat_slots(I));
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
  out = std::move(r);
  return GLE_SUCCESS;
}

// Queues validated requests (caller holds sv.mu).
static void enqueue_requests(Server &sv, int inst, Instance *I, std::vector<std::unique_ptr<Req>> &rs, int *ids) {
  ServeInstance &si = sv.insts[inst];
  if (!si.I) {
    si.h = inst;
    si.I = I;
    si.rescale_every = instance_precision(I) == GLE_PREC_FP64 ? 16 : 1;
    si.gen = instance_generation(I);
    si.msgs = instance_states(I) == 61;
  }
  const auto now = std::chrono::steady_clock::now();
  for (size_t k = 0; k < rs.size(); k++) {
    Req *r = rs[k].get();
    r->t_submit = now;
    r->id = sv.next_id++;
    ids[k] = r->id;
    sv.pending.push_back(r);
    sv.reqs[r->id] = std::move(rs[k]);
  }
}

extern "C" int gle_serve_submit(int h, int inst, const int *parent, const double *blen, int kind, double precision,
                                int *req) {
  Server *sv = get_server(h);
  if (!sv) return set_error(GLE_ERR_ARG, "bad server handle %d", h);
  Instance *I = get_instance(inst);
  if (!I || !req) return set_error(GLE_ERR_ARG, "bad request");
  std::vector<std::unique_ptr<Req>> rs(1);
  const int rc = make_request(I, inst, parent, blen, kind, precision, rs[0]);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    enqueue_requests(*sv, inst, I, rs, req);
  }
  sv->cv_work.notify_one();
  return GLE_SUCCESS;
}

extern "C" int gle_serve_submit_batch(int h, int inst, int n, const int *parents, const double *blens, int kind,
                                      double precision, int *reqs) {
  Server *sv = get_server(h);
  if (!sv) return set_error(GLE_ERR_ARG, "bad server handle %d", h);
  Instance *I = get_instance(inst);
  if (!I || n < 0 || (n > 0 && (!parents || !blens || !reqs))) return set_error(GLE_ERR_ARG, "bad request batch");
  const size_t nn = size_t(2 * instance_tips(I) - 2);
  std::vector<std::unique_ptr<Req>> rs(n);
  for (int k = 0; k < n; k++) {
    const int rc = make_request(I, inst, parents + k * nn, blens + k * nn, kind, precision, rs[k]);
    if (rc) return rc;  // nothing of the batch was queued
  }
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    enqueue_requests(*sv, inst, I, rs, reqs);
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
  sv->collected.push_back(std::move(it->second));
  sv->reqs.erase(it);
  return rc;
}

extern "C" int gle_serve_wait_all(int h, int n, const int *reqs, double *lnl) {
  Server *sv = get_server(h);
  if (!sv || n < 0 || (n > 0 && (!reqs || !lnl))) return set_error(GLE_ERR_ARG, "bad argument");
  std::unique_lock<std::mutex> lk(sv->mu);
  std::vector<Req *> rs(size_t(n), nullptr);
  for (int i = 0; i < n; i++) {
    auto it = sv->reqs.find(reqs[i]);
    if (it == sv->reqs.end()) return set_error(GLE_ERR_ARG, "unknown request %d", reqs[i]);
    rs[size_t(i)] = it->second.get();
  }
  sv->cv_done.wait(lk, [&] {
    for (const Req *r : rs)
      if (!r->finished) return false;
    return true;
  });
  int rc = GLE_SUCCESS;
  for (int i = 0; i < n; i++) {
    const Req *r = rs[size_t(i)];
    lnl[i] = r->lnl;
    if (!rc && r->rc) rc = r->rc;
  }
  for (int i = 0; i < n; i++) {
    auto it = sv->reqs.find(reqs[i]);
    if (it == sv->reqs.end()) continue;  // listed twice
    sv->collected.push_back(std::move(it->second));
    sv->reqs.erase(it);
  }
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
