// gle_internal.h -- shared declarations of libgle (not installed).
#pragma once
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "gle.h"

namespace gle {

int set_error(int code, const char *fmt, ...);

// host model/data helpers (gle_host_model.cpp)
int reversible_eigen(int n, const double *Q, const double *pi, double *V, double *Vinv, double *lam);
int discrete_gamma(int K, double alpha, double *rates);
int nuc_datum(char c);
int aa_datum(char c);
// amino acid of codon 16*n1 + 4*n2 + n3 (A, C, G, T = 0..3) under the
// standard code: ACDEFGHIKLMNPQRSTVWY = 0..19, stop = 20
int standard_code_aa(int codon64);
int make_ambig_string(const unsigned char *masks, int n, signed char *out);
int ambig_string_to_masks(const signed char *st, int len, int n, unsigned char *masks);
int compress_patterns(int ntax, int nsite, int ms, const unsigned char *matrix, unsigned char *out, int *counts,
                      int *const_states, int *last_const, int *site_to_pattern);

// Rescale constants of the FP64 reference (trunk/src/tree.cpp:159-172).
constexpr int kRescaleLen = 90;
struct RescaleTables {
  double thresh[kRescaleLen];
  double mult[kRescaleLen];
  int incr[kRescaleLen];
  double rescale_below, reduce_below, bail_below;
};
const RescaleTables &rescale_tables();

// One CLA update as the device sees it (child < 0 means tip -(child+1)).
// Codon instances only: with kOpMsg1 (kOpMsg2) set, c1 (c2) is the slot of a
// stored message P(t) . x of that child instead of the child's CLA, and pm1
// (pm2) is unused; m1 (m2) >= 0 asks the kernel to keep the message it forms
// for child 1 (2) in that slot, with the child's rescale counts.
struct DevClaOp {
  int dest, c1, pm1, c2, pm2, flags;
  int m1 = -1, m2 = -1;
};
// One edge integration.
struct DevEdgeOp {
  int partial, child, pm, derivs;
  int msg = 0;  // codon, no derivatives: child is a kept message P(t) . x (pm unused)
};
enum : int { kOpRescale = 1, kOpMsg1 = 2, kOpMsg2 = 4 };

// Engine entry points used by the tree and serving layers (gle_engine.hip).
struct Instance;
Instance *get_instance(int h);
int instance_states(const Instance *);
int instance_rates(const Instance *);
int instance_patterns(const Instance *);
double instance_constant_fraction(const Instance *);  // count-weighted share of constant patterns
int instance_tips(const Instance *);
int instance_precision(const Instance *);
unsigned long long instance_generation(const Instance *);
int instance_cla_slots(const Instance *);
int instance_pmat_slots(const Instance *);
// wait = false: queue the work and return (the next evaluation on the
// instance's streams runs after it; errors of the queued kernels surface at
// that evaluation's wait)
int engine_update_pmats(Instance *, const int *idx, const double *t, int n, int derivs, bool wait = true);
// ops in level order; level_start[n_levels+1]; fused = one launch for all ops
int engine_update_partials(Instance *, const DevClaOp *ops, int n_ops, const int *level_start, int n_levels,
                           bool fused, bool wait = true);
// flags (optional): the rescale flags, read back with the results in the same wait
int engine_edges(Instance *, const DevEdgeOp *ops, int n, double *out, double *site_lnl, int *flags = nullptr);
// one edge with its per-pattern values: site5 = 5 x patterns (lnL, d1, d2,
// partial-side and child-side rescale counts, -1 for a tip child)
int engine_edge_sites(Instance *, const DevEdgeOp &op, double *out3, double *site5, int *flags = nullptr);
int engine_flags(Instance *, int *flags, int reset);
// Staging areas per instance for serving ticks in flight at once.
constexpr int kTickStages = 3;
// One serving tick on an instance (and its heavy part), queued with no wait
// into staging area sidx (0 .. kTickStages - 1, one per tick in flight): P (and
// P', P'' when pm_derivs) for the listed lengths, the rescale-flag reset, the CLA ops by level, the
// edges.  engine_tick_collect waits for it and returns n_e x (lnL, d1, d2)
// and the rescale flags.
int engine_tick_enqueue(Instance *, int sidx, const int *pm_idx, const double *pm_len, int n_pm, int pm_derivs,
                        const DevClaOp *ops, int n_ops, const int *level_start, int n_levels, const DevEdgeOp *eops,
                        int n_e);
int engine_tick_collect(Instance *, int sidx, double *res, int *flags);
// True when the tick in stage sidx has finished on the device (both parts),
// or none is in flight there: engine_tick_collect will then not wait.
bool engine_tick_ready(Instance *, int sidx);
int engine_get_partials(Instance *, int slot, double *out, int *u);
int engine_sync(Instance *);
bool engine_supports_fused(const Instance *);

// Slot allocator shared by the tree and serving layers (one per instance):
// CLA slots with reference counts and the rescale rank of the CLA each holds,
// and transition-matrix slots.  Guarded by pool_mutex().
struct ClaPool {
  std::vector<int> refs;
  std::vector<int> rank;
  std::vector<int> pm_free;
  // free CLA slots, lowest first: the same tree state always gets the same
  // slots, so a repeated op list is recognised (and replayed from a graph)
  std::set<int> free_set;
  void free_slot(int s) { free_set.insert(s); }
  bool has_free() const { return !free_set.empty(); }
  int take_slot() {
    const int s = *free_set.begin();
    free_set.erase(free_set.begin());
    return s;
  }
};
ClaPool *pool_for(Instance *I);
std::mutex &pool_mutex();

}  // namespace gle
