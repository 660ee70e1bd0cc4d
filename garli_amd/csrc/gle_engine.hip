// gle_engine.hip -- device side of libgle: instances, kernels, instance-layer C ABI.
//
// Device data layout (per instance, all in HBM):
//   tips   u8  [N][Ppad]           nucleotide 4-bit masks (1..15) or state (S = missing)
//   cla    T   [slots][Ppad][R][Sp] the reference's pattern-major layout (condlike.h:20-43);
//                                   Sp = S except codon (61 -> 64).  Nucleotides use a
//                                   rate-blocked order instead: [slots][Ppad/64][R][64][4],
//                                   so each (64 patterns, category) is one contiguous 1 KB
//                                   (FP32) run and a wave's 16-byte loads coalesce fully.
//                                   Amino acids: [slots][Ppad/64][R][20][64], state-major
//                                   inside a 64-pattern block, so every per-state load of a
//                                   wave is one contiguous 256-byte (FP32) run.
//   u      i32 [slots][Ppad]       per-pattern rescale counts (FP64: natural-log units
//                                   exactly as the reference; FP32: powers of two)
//   pm64   f64 [pmats][3][R][S][S] P, P', P'' (row = parent state, model.cpp:1271-1358)
//   pmpad  f64 [pmats][3][R][Sp][Sp] a second copy of P, P', P'' in the layout a kernel family
//                                   wants: codons zero-padded to 64 x 64 in MFMA fragment
//                                   order (codon_frag); amino
//                                   acids transposed (column j of P contiguous)
//   counts i32 [Ppad], cst i32 [Ppad] (constant-pattern state/mask, -1 otherwise)
// Ppad rounds P up to a multiple of 256 so every 256-thread tile is whole.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "gle_internal.h"

namespace gle {

// ------------------------------------------------------------------ errors
static thread_local std::string g_last_error;
int set_error(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}
#define GLE_HIP(call)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return set_error(GLE_ERR_DEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                                    \
  } while (0)

const RescaleTables &rescale_tables() {
  static RescaleTables t = [] {
    RescaleTables r;
    r.rescale_below = std::exp(-24.0);
    r.reduce_below = 1.0e-190;
    r.bail_below = 1.0e-250;
    const double maxmult = 1.0 / r.bail_below;
    for (int i = 0; i < kRescaleLen; i++) {
      r.incr[i] = i * 7 - int(std::log(r.rescale_below));
      r.thresh[i] = std::exp(double(-r.incr[i]));
      r.mult[i] = std::min(std::exp(double(r.incr[i])), maxmult);
    }
    return r;
  }();
  return t;
}

__constant__ double c_thresh[kRescaleLen];
__constant__ double c_mult[kRescaleLen];
__constant__ int c_incr[kRescaleLen];
__constant__ double c_rescale_below, c_reduce_below, c_bail_below;

// Scratch of the allocation-free edge entry point (gle_edge_likelihoods_noalloc):
// the uploaded op list, the per-block partial sums and the per-op results of an
// instance and of its FP64 heavy part, in place of the buffers
// gle_edge_likelihoods grows with hipMalloc.  Every GPU the library runs on
// holds its own copy until the process ends, and every such call on one GPU
// shares it.  In host code `scratch` only names the symbol: kernels and copies
// use the device address hipGetSymbolAddress gives for the current device,
// kept in g_scratch_dev.
constexpr size_t kNoAllocScratchBytes = size_t(8) << 20;
constexpr int kNoAllocMaxOps = 1024;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static unsigned char *g_scratch_dev[kMaxDevices];

// Once per device, with that device current (called from device set-up).
static int scratch_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return set_error(GLE_ERR_DEVICE, "device %d beyond the scratch table", dev);
  if (g_scratch_dev[dev]) return GLE_SUCCESS;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return set_error(GLE_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                     hipGetErrorString(e));
  g_scratch_dev[dev] = static_cast<unsigned char *>(p);
  return GLE_SUCCESS;
}

// --------------------------------------------------------------- instance
struct Instance {
  int N = 0, P = 0, Ppad = 0, S = 0, Sp = 0, R = 0, n_eig = 0, n_slots = 0, n_pm = 0, prec = 0;
  hipStream_t stream = nullptr;
  unsigned char *d_tips = nullptr;
  int *d_counts = nullptr, *d_cst = nullptr, *d_u = nullptr, *d_flags = nullptr, *d_eig_of_rate = nullptr;
  void *d_cla = nullptr;
  double *d_pm64 = nullptr;
  double *d_pmpad = nullptr;  // codons (padded) and amino acids (transposed) only
  double *d_pdiag = nullptr;  // FP32 amino acids / codons: exp(s t), s exp(s t), s^2 exp(s t) per [pmat][R][3][S],
                              // s = lam * rate
  double *d_eigfrag = nullptr;  // FP32 codons: V^-1 and V^T of each eigensystem in codon fragment order
  double *d_V = nullptr, *d_Vinv = nullptr, *d_lam = nullptr, *d_model = nullptr;  // model: pi[S] w[R] mult[R]
  // staging
  DevClaOp *d_ops = nullptr;
  int ops_cap = 0;
  DevEdgeOp *d_eops = nullptr;
  int eops_cap = 0;
  double *d_part = nullptr;
  size_t part_cap = 0;
  // a heavy part's edge partials in serving ticks, which run on the main
  // part's stream (d_part stays with the work on the heavy part's own stream)
  double *d_tpart = nullptr;
  size_t tpart_cap = 0;
  double *d_eres = nullptr;
  int eres_cap = 0;
  double *d_site = nullptr;
  int *d_pmidx = nullptr;
  double *d_pmt = nullptr;
  int pm_cap = 0;
  double *h_stage = nullptr;  // pinned
  size_t h_stage_cap = 0;
  int *h_flag = nullptr;  // pinned: the rescale flags read back with an edge evaluation's results
  std::vector<void *> retired;       // device buffers outgrown by serving ticks, freed with the instance
  std::vector<void *> retired_host;  // pinned host buffers likewise
  double pinv = 0.0;
  int pinv_in_model = 0, last_const = -1;
  // kernel timing (HIP events on this instance's stream): 0 CLA, 1 edge, 2 pmat
  bool timing = false;
  struct KStat {
    long long launches = 0;
    double ms = 0.0, bytes = 0.0;
  } kst[3];
  std::vector<hipEvent_t> ev_free;
  struct Pending {
    hipEvent_t a, b;
    int kind;
    double bytes;
  };
  std::vector<Pending> pending;
  std::mutex mu;  // one caller at a time per instance (staging buffers, stream)
  std::unique_ptr<ClaPool> pool;  // slot allocator of the tree and serving layers
  // FP32 instances: patterns whose count exceeds heavy_threshold run in an FP64
  // sub-instance that receives the same slots, P(t) and op lists, so FP32
  // rounding is never multiplied by a large pattern weight.
  // A pattern of count c carries c times its FP32 per-site error (at most
  // about 1e-6 in lnL here): above 100 that could exceed 1e-4, so such
  // patterns run in the FP64 part.  0 keeps every pattern in FP32.
  int heavy_threshold = 100;
  int fused_min_patterns = 1 << 19;  // smallest padded pattern count that runs fused walks
  // bumped by every change of data or model: cached CLAs and P(t) of trees and
  // servers built before it are stale (MakeAllNodesDirty, tree.cpp)
  unsigned long long generation = 0;
  hipEvent_t tev_a = nullptr, tev_b = nullptr;  // events for the next timed launch
  // what d_ops / d_eops hold, so that a repeated list is not uploaded again
  std::vector<DevClaOp> ops_held;
  std::vector<DevEdgeOp> eops_held;
  const void *ops_held_buf = nullptr, *eops_held_buf = nullptr;
  // hipGraph replay of repeated level-launch sequences (enqueue_partials): a
  // few recent op lists, each captured on its second use
  struct GraphEntry {
    std::vector<DevClaOp> ops;
    std::vector<int> levels;
    const DevClaOp *dev = nullptr;
    unsigned long long gen = 0, last = 0;
    hipGraphExec_t exec = nullptr;
  };
  static constexpr int kGraphEntries = 4;
  bool graphs = true;
  std::vector<GraphEntry> g_cache;
  unsigned long long g_clock = 0;
  int g_replays = 0;
  void drop_graphs() {
    for (auto &g : g_cache)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g_cache.clear();
  }
  // asynchronous serving ticks (engine_tick_enqueue / engine_tick_collect):
  // pinned staging areas used in turn, each with the event its tick ends on
  struct TickStage {
    unsigned char *h = nullptr;  // pinned: P(t) list, op lists, flags and results
    unsigned char *d = nullptr;  // the same layout on the device
    size_t cap = 0;
    hipEvent_t done = nullptr;
    int n_edges = 0;
    double *res = nullptr;  // n_edges x 3, then the heavy part's n_edges x 3 (with_heavy)
    int *flags = nullptr;   // this part's rescale flags, then the heavy part's
    bool with_heavy = false;
    bool busy = false;
  } stage[kTickStages];
  std::unique_ptr<Instance> heavy;
  int P_orig = 0;
  std::vector<int> light_idx, heavy_idx;  // original pattern indices of each part
  // host copies (the split is re-derived whenever the patterns change)
  std::vector<unsigned char> h_tips;  // [N][P_orig]
  std::vector<int> h_counts, h_cst;
  int h_last_const = -1;
  bool h_patterns = false;
  std::vector<double> h_pi, h_w, h_V, h_Vinv, h_lam, h_mult;
  std::vector<int> h_eor;
  std::vector<char> h_eig_set;
  bool h_model = false, h_rates = false;
  size_t slot_elems() const { return size_t(Ppad) * R * Sp; }
  size_t elem_bytes() const { return prec == GLE_PREC_FP64 ? 8 : 4; }
};

static std::mutex g_mu;
static std::vector<std::unique_ptr<Instance>> g_instances;
static int g_device = -1;

Instance *get_instance(int h) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (h < 0 || h >= int(g_instances.size())) return nullptr;
  return g_instances[h].get();
}
int instance_states(const Instance *I) { return I->S; }
int instance_rates(const Instance *I) { return I->R; }
int instance_patterns(const Instance *I) { return I->P_orig; }
double instance_constant_fraction(const Instance *I) {
  double all = 0.0, cst = 0.0;
  for (int q = 0; q < I->P_orig; q++) {
    all += I->h_counts[q];
    if (I->h_cst[q] >= 0) cst += I->h_counts[q];
  }
  return all > 0.0 ? cst / all : 0.0;
}
int instance_tips(const Instance *I) { return I->N; }
int instance_precision(const Instance *I) { return I->prec; }
unsigned long long instance_generation(const Instance *I) { return I->generation; }
int instance_cla_slots(const Instance *I) { return I->n_slots; }
int instance_pmat_slots(const Instance *I) { return I->n_pm; }

// Guarded by pool_mutex() (callers hold it).
ClaPool *pool_for(Instance *I) {
  if (!I->pool) {
    I->pool = std::make_unique<ClaPool>();
    I->pool->refs.assign(I->n_slots, 0);
    I->pool->rank.assign(I->n_slots, 0);
    for (int s = I->n_slots - 1; s >= 0; s--) I->pool->free_slot(s);
    for (int s = I->n_pm - 1; s >= 0; s--) I->pool->pm_free.push_back(s);
  }
  return I->pool.get();
}

std::mutex &pool_mutex() {
  static std::mutex m;
  return m;
}

static int scratch_setup(int dev);

static int ensure_device(int dev = 0) {
  if (g_device >= 0) return GLE_SUCCESS;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
    return set_error(GLE_ERR_DEVICE, "no HIP device visible: the GARLI engine runs only on a GPU");
  GLE_HIP(hipSetDevice(dev));
  const int rs = scratch_setup(dev);
  if (rs) return rs;
  g_device = dev;
  const RescaleTables &t = rescale_tables();
  GLE_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_thresh), t.thresh, sizeof t.thresh));
  GLE_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_mult), t.mult, sizeof t.mult));
  GLE_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_incr), t.incr, sizeof t.incr));
  GLE_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_rescale_below), &t.rescale_below, sizeof(double)));
  GLE_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_reduce_below), &t.reduce_below, sizeof(double)));
  GLE_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_bail_below), &t.bail_below, sizeof(double)));
  return GLE_SUCCESS;
}

template <typename X>
static int grow(X *&ptr, int &cap, int need) {
  if (need <= cap) return GLE_SUCCESS;
  if (ptr) GLE_HIP(hipFree(ptr));
  ptr = nullptr;
  int c = std::max(need, 2 * cap);
  GLE_HIP(hipMalloc(&ptr, sizeof(X) * size_t(c)));
  cap = c;
  return GLE_SUCCESS;
}

// HIP-event timing of individual launches; bytes are algorithmic bytes.
static hipEvent_t take_event(Instance *I) {
  if (!I->ev_free.empty()) {
    hipEvent_t e = I->ev_free.back();
    I->ev_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}
// timed_begin arms a pair of events that the next GLE_LAUNCH hands to
// hipExtLaunchKernelGGL, which stamps them at that kernel's own start and end
// (no queue latency in the interval); timed_end is a no-op kept for symmetry.
static void timed_begin(Instance *I, int kind, double bytes) {
  if (!I->timing) return;
  Instance::Pending p{take_event(I), take_event(I), kind, bytes};
  I->pending.push_back(p);
  I->tev_a = p.a;
  I->tev_b = p.b;
}
static void timed_end(Instance *I) { I->tev_a = I->tev_b = nullptr; }
#define GLE_LAUNCH(I, K, grid, block, shmem, ...)                                                           \
  do {                                                                                                      \
    if ((I)->tev_a) {                                                                                       \
      hipExtLaunchKernelGGL(K, grid, block, shmem, (I)->stream, (I)->tev_a, (I)->tev_b, 0, __VA_ARGS__);    \
      (I)->tev_a = (I)->tev_b = nullptr;                                                                    \
    } else {                                                                                                \
      hipLaunchKernelGGL(K, grid, block, shmem, (I)->stream, __VA_ARGS__);                                  \
    }                                                                                                       \
  } while (0)
static void timed_collect(Instance *I) {  // call after the stream drained
  for (auto &p : I->pending) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      I->kst[p.kind].launches++;
      I->kst[p.kind].ms += ms;
      I->kst[p.kind].bytes += p.bytes;
    }
    I->ev_free.push_back(p.a);
    I->ev_free.push_back(p.b);
  }
  I->pending.clear();
}

// Algorithmic bytes of one CLA op / edge op per pattern (true S, not padded).
static double cla_op_bytes(const Instance *I, const DevClaOp &o) {
  const double rsb = double(I->R) * I->S * I->elem_bytes();
  const int internal = (o.c1 >= 0) + (o.c2 >= 0);
  const double per = internal == 2 ? 3 * rsb + 12 : internal == 1 ? 2 * rsb + 1 + 8 : rsb + 2 + 4;
  return per * I->P;
}
static double edge_op_bytes(const Instance *I, const DevEdgeOp &o) {
  const double rsb = double(I->R) * I->S * I->elem_bytes();
  const double per = o.child >= 0 ? 2 * rsb + 8 + 4 : rsb + 1 + 4 + 4;
  return per * I->P;
}

// =====================================================================
// Device helpers
// =====================================================================

// Reference-order dot products for one P row.  FP64 paths run with
// contraction off so the arithmetic is the reference's mul-then-add.
__device__ __forceinline__ double dot4_paired(const double *p, const double *x) {
#pragma clang fp contract(off)
  return ((p[0] * x[0]) + (p[1] * x[1])) + ((p[2] * x[2]) + (p[3] * x[3]));
}
__device__ __forceinline__ double dot4_seq(const double *p, const double *x) {
#pragma clang fp contract(off)
  return p[0] * x[0] + p[1] * x[1] + p[2] * x[2] + p[3] * x[3];
}

__device__ __forceinline__ double mul_nc(double a, double b) {
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ double add_nc(double a, double b) {
#pragma clang fp contract(off)
  return a + b;
}

template <typename T>
__device__ __forceinline__ void load4(const T *src, T *v) {
  if constexpr (sizeof(T) == 4) {
    float4 a = *reinterpret_cast<const float4 *>(src);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  } else {
    double2 a = reinterpret_cast<const double2 *>(src)[0], b = reinterpret_cast<const double2 *>(src)[1];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
  }
}
template <typename T>
__device__ __forceinline__ void store4(T *dst, const T *v) {
  if constexpr (sizeof(T) == 4) {
    *reinterpret_cast<float4 *>(dst) = make_float4(v[0], v[1], v[2], v[3]);
  } else {
    reinterpret_cast<double2 *>(dst)[0] = make_double2(v[0], v[1]);
    reinterpret_cast<double2 *>(dst)[1] = make_double2(v[2], v[3]);
  }
}

// The FP32 store with the nontemporal hint: a CLA written by one op is read
// back by a later launch, not by this one (about 1% more bandwidth in the
// nucleotide CLA kernel; the FP64 form measured 1.5% slower with it, so it
// keeps the plain store).
typedef float gle_nt_f4 __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __forceinline__ void store4_stream(T *dst, const T *v) {
  if constexpr (sizeof(T) == 4) {
    const gle_nt_f4 a = {v[0], v[1], v[2], v[3]};
    __builtin_nontemporal_store(a, reinterpret_cast<gle_nt_f4 *>(dst));
  } else {
    store4(dst, v);
  }
}

// Layouts of the padded FP64 copy of P, P', P'' (pmpad).
enum { kPadRowMajor = 0, kPadTransposed = 1, kPadCodonFrag = 2 };
// Codon fragment order: entry P[i][j] of a 64x64 block sits where lane
// l = 16 (j / 16) + i % 16 finds it at step kk = j % 16 of output tile
// t = i / 16 of the codon MFMA kernels, so that every B-fragment load of a
// wave reads 64 consecutive doubles.
__host__ __device__ __forceinline__ int codon_frag(int i, int j) {
  return ((i >> 4) * 16 + (j & 15)) * 64 + (j >> 4) * 16 + (i & 15);
}

// =====================================================================
// P(t), P', P'' (model.cpp:1271-1358, 1469-1597) -- FP64
// grid (n, R, z); V/Vinv of the category's eigensystem staged in LDS.  Each
// thread forms a 2x2 tile of the S x S result: per input state k it reads two
// entries of V's column and two of V^-1's row and updates 4 (P only) or 12
// independent sums, so the LDS latency of one k overlaps the arithmetic of
// the others.  Every entry is summed in the reference's order, the product
// V_ik W_kj first (contraction off), so the values do not depend on the
// tiling.  The z blocks of a matrix split its (ceil(S/2))^2 tiles 256 apiece.
// =====================================================================
__host__ __device__ __forceinline__ int pmat_tiles(int S) { return ((S + 1) / 2) * ((S + 1) / 2); }

__global__ __launch_bounds__(256) void k_pmat(int S, int Sp, int R, const int *__restrict__ idx,
                                              const double *__restrict__ tt, const int *__restrict__ eig_of_rate,
                                              const double *__restrict__ rate_mult, const double *__restrict__ V,
                                              const double *__restrict__ Vinv, const double *__restrict__ lam,
                                              double *__restrict__ pm64, double *__restrict__ pmpad,
                                              double *__restrict__ pdiag, int derivs, int layout) {
#pragma clang fp contract(off)
  extern __shared__ double sm[];
  const int SS = S * S, op = blockIdx.x, r = blockIdx.y;
  const int e = eig_of_rate[r];
  double *sV = sm, *sW = sm + SS, *ex = sm + 2 * SS, *e1 = ex + S, *e2 = e1 + S;
  for (int q = threadIdx.x; q < SS; q += blockDim.x) {
    sV[q] = V[size_t(e) * SS + q];
    sW[q] = Vinv[size_t(e) * SS + q];
  }
  const double t = tt[op], mult = rate_mult[r];
  const int slot = idx[op];
  for (int k = threadIdx.x; k < S; k += blockDim.x) {
    const double s = lam[size_t(e) * S + k] * mult;
    ex[k] = exp(s * t);
    e1[k] = s * ex[k];
    e2[k] = s * e1[k];
    if (pdiag && blockIdx.z == 0) {
      double *d = pdiag + (size_t(slot) * R + r) * 3 * S;
      d[k] = ex[k];
      d[S + k] = e1[k];
      d[2 * S + k] = e2[k];
    }
  }
  __syncthreads();
  const int half = (S + 1) / 2, tile = blockIdx.z * blockDim.x + threadIdx.x;
  if (tile >= half * half) return;
  const int i0 = 2 * (tile / half), j0 = 2 * (tile % half);
  // the second row / column of an edge tile repeats the first (in bounds), and is not stored
  const int i1 = i0 + 1 < S ? i0 + 1 : i0, j1 = j0 + 1 < S ? j0 + 1 : j0;
  double sp[2][2] = {{0.0, 0.0}, {0.0, 0.0}}, s1[2][2] = {{0.0, 0.0}, {0.0, 0.0}}, s2[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  if (derivs) {
    for (int k = 0; k < S; k++) {
      const double va[2] = {sV[i0 * S + k], sV[i1 * S + k]}, wb[2] = {sW[k * S + j0], sW[k * S + j1]};
      const double f0 = ex[k], f1 = e1[k], f2 = e2[k];
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          const double x = va[a] * wb[b];
          sp[a][b] += x * f0;
          s1[a][b] += x * f1;
          s2[a][b] += x * f2;
        }
    }
  } else {
    for (int k = 0; k < S; k++) {
      const double va[2] = {sV[i0 * S + k], sV[i1 * S + k]}, wb[2] = {sW[k * S + j0], sW[k * S + j1]};
      const double f0 = ex[k];
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) sp[a][b] += (va[a] * wb[b]) * f0;
    }
  }
  double *P = pm64 + (size_t(slot) * 3 * R + r) * SS;
  double *D1 = pm64 + (size_t(slot) * 3 * R + R + r) * SS;
  double *D2 = pm64 + (size_t(slot) * 3 * R + 2 * R + r) * SS;
  double *Pp = pmpad ? pmpad + (size_t(slot) * 3 * R + r) * Sp * Sp : nullptr;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int i = i0 + a, j = j0 + b;
      if (i >= S || j >= S) continue;
      const int q = i * S + j;
      const double pv = sp[a][b] > 0.0 ? sp[a][b] : 0.0;
      P[q] = pv;
      const int pq = layout == kPadTransposed ? j * Sp + i : layout == kPadCodonFrag ? codon_frag(i, j) : i * Sp + j;
      if (Pp) Pp[pq] = pv;
      if (derivs) {
        D1[q] = s1[a][b];
        D2[q] = s2[a][b];
        if (Pp) {
          Pp[size_t(R) * Sp * Sp + pq] = s1[a][b];
          Pp[size_t(2 * R) * Sp * Sp + pq] = s2[a][b];
        }
      }
    }
}

// 16 wave-uniform doubles (one category's 4x4 P, 128-byte aligned) as two
// 64-byte blocks, so the compiler issues wide scalar loads.
struct alignas(64) Dbl8 {
  double v[8];
};
__device__ __forceinline__ void load_uniform16(const double *__restrict__ src, double *v) {
  const Dbl8 a = reinterpret_cast<const Dbl8 *>(src)[0], b = reinterpret_cast<const Dbl8 *>(src)[1];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    v[q] = a.v[q];
    v[8 + q] = b.v[q];
  }
}
// Tip factor of mask m for one P row: 1 for total ambiguity, else the
// ascending-order sum of the row's entries over the mask's states (one entry
// for an unambiguous tip).  tree.cpp:7941-8047, 8387-8446: the ambiguity
// stream lists a tip's states in ascending order.
__device__ __forceinline__ double tip_sum(const double *row, const double *bit, int m) {
  const double s = __builtin_fma(bit[3], row[3], __builtin_fma(bit[2], row[2],
                   __builtin_fma(bit[1], row[1], bit[0] * row[0])));
  return m == 15 ? 1.0 : s;
}

// Offset of (pattern p, category r, state 0) within an amino-acid CLA slot;
// state s is 64 elements further on.
__host__ __device__ __forceinline__ size_t aa_off(int p, int r, int R) {
  return (size_t(p >> 6) * R + r) * 1280 + size_t(p & 63);
}

// Offset of (pattern p, category r) within a nucleotide CLA slot.
__host__ __device__ __forceinline__ size_t nuc_off(int p, int r, int R) {
  return (size_t(p >> 6) * R + r) * 256 + size_t(p & 63) * 4;
}

// =====================================================================
// Nucleotide CLA update (S = 4): one thread per pattern holding every rate
// category, so the per-pattern rescale needs no cross-lane traffic.  The op's
// P matrices are the same for the whole wave: with every pointer argument
// __restrict__ they are fetched by scalar loads and enter the FMAs as SGPR
// operands, leaving the vector memory path to CLA, tip and count traffic.
// tree.cpp:7843-7910 (II), 7941-8047 (IT), 8387-8446 (TT).
// =====================================================================

// FP32 form: CLAs are stored in FP32 but every update computes in FP64 from the
// FP64 P (FP64 FMA issues at the FP32 rate here), so the only FP32 rounding is
// the store.  That store adds a dither uniform in [-ulp/2, ulp/2) of the FP32
// result before rounding to nearest: the rounding is unbiased, and patterns
// that carry equal values (every pattern below a cherry with the same two tip
// states, for one) get independent errors.  A plain round-to-nearest gives
// them all the same error, which then adds up linearly over the alignment.
// The dither is a hash of the pattern index, the element and the value, so
// results are deterministic and independent of slot assignment and batching.
__device__ __forceinline__ unsigned mix32(unsigned h) {
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}
// x * 2^scale rounded to FP32 with the dither drawn from x itself, so a value
// scaled before its store rounds as the unscaled value would (FP32 normals).
__device__ __forceinline__ float round_dither(double x, unsigned key, int scale = 0) {
  if (x == 0.0) return 0.0f;
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
  const unsigned h = mix32(key * 0x9e3779b1u ^ unsigned(b >> 29));
  const double y = scale ? __builtin_amdgcn_ldexp(x, scale) : x;
  const int e = __builtin_amdgcn_frexp_exp(y);  // y = m * 2^e, m in [0.5, 1)
  const double ulp = __builtin_amdgcn_ldexp(1.0, e - 24 > -149 ? e - 24 : -149);
  return float(y + double(int(h)) * 0x1p-32 * ulp);
}
// The same dithered rounding done on the bit pattern, for kernels that store
// many values per pattern: adding an integer k to the bits of a positive normal
// double moves it by k units of 2^-29 FP32 ulps (within its binade), so with
// k = (r >> 3) - 2^28 uniform in [-2^28, 2^28) the round to nearest that
// follows is the dithered one.  Zero stays zero.  r comes from dither_stream.
__device__ __forceinline__ float round_dither_bits(double x, unsigned r) {
  const long long b = __double_as_longlong(x) + (long long)(int(r >> 3) - (1 << 28));
  return x == 0.0 ? 0.0f : float(__longlong_as_double(b));
}
// Element k's random word from a per-pattern seed h (one mix32 per pattern).
__device__ __forceinline__ unsigned dither_stream(unsigned h, unsigned k) {
  return (h ^ (k * 0x9e3779b9u)) * 0x85ebca6bu;
}
// FP32-form rescale on FP64 values: exact power of two once the max drops below 2^-32.
template <int W>
__device__ __forceinline__ int rescale_pow2(double *v, double m) {
  if (!(m < 0x1p-32 && m > 0.0)) return 0;
  const int e = 1 - __builtin_amdgcn_frexp_exp(m);  // m * 2^e in [1, 2)
#pragma unroll
  for (int q = 0; q < W; q++) v[q] = __builtin_amdgcn_ldexp(v[q], e);
  return e;
}

// grid (Ppad/256, fused ? 1 : n_ops); fused_n > 0 walks ops[0..fused_n) in order.
// pm: FP64 P slots [slot][3][R][4][4] (P, P', P'').
template <typename T, int R>
__global__ __launch_bounds__(256) void k_cla_nuc(T *__restrict__ cla, int *__restrict__ ua,
                                                 const unsigned char *__restrict__ tips,
                                                 const double *__restrict__ pm, const DevClaOp *__restrict__ ops,
                                                 int *__restrict__ flags, int Ppad, int P, int fused_n) {
  constexpr int kPmStride = 3 * R * 16;
  if (int(blockIdx.x * 256 + threadIdx.x) >= P) return;  // no cross-lane work below
  const size_t slot = size_t(Ppad) * R * 4;
  const int first = fused_n ? 0 : blockIdx.y, last = fused_n ? fused_n : blockIdx.y + 1;
  for (int oi = first; oi < last; ++oi) {
    // the thread's pattern and offset, formed again in each op from an index
    // the compiler cannot see through (an empty asm), so they are not hoisted
    // out of the op loop and held across it
    int t = int(threadIdx.x);
    asm volatile("" : "+v"(t));
    const int p = blockIdx.x * 256 + t;
    const size_t off = nuc_off(p, 0, R);
    const DevClaOp op = ops[oi];
    T x1[R][4], x2[R][4];
    int m1 = 15, m2 = 15, u = 0;
    if (op.c1 >= 0) {
#pragma unroll
      for (int r = 0; r < R; r++) load4(cla + op.c1 * slot + off + 256 * r, x1[r]);
      u += ua[op.c1 * size_t(Ppad) + p];
    } else
      m1 = tips[size_t(-op.c1 - 1) * Ppad + p];
    if (op.c2 >= 0) {
#pragma unroll
      for (int r = 0; r < R; r++) load4(cla + op.c2 * slot + off + 256 * r, x2[r]);
      u += ua[op.c2 * size_t(Ppad) + p];
    } else
      m2 = tips[size_t(-op.c2 - 1) * Ppad + p];
    const double *pm1 = pm + size_t(op.pm1) * kPmStride, *pm2 = pm + size_t(op.pm2) * kPmStride;
    // tip masks as 0/1 factors: fma(bit, p, acc) adds p exactly when the state
    // is in the mask and leaves acc unchanged otherwise, so the ascending-order
    // sum of the reference comes out bit for bit without per-lane branches
    double b1[4], b2[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      b1[q] = double((m1 >> q) & 1);
      b2[q] = double((m2 >> q) & 1);
    }
    T *dst = cla + op.dest * slot + off;
    double mx = 0.0;
    unsigned seed = 0;  // FP32 form: dither seed of this pattern and op
    // each category is stored as soon as it is formed (few live registers);
    // the rare rescale re-reads and scales what this thread stored
#pragma unroll
    for (int r = 0; r < R; r++) {
      double P1[16], P2[16], out[4];
      load_uniform16(pm1 + 16 * r, P1);
      load_uniform16(pm2 + 16 * r, P2);
      if constexpr (sizeof(T) == 8) {
        // reference summation order: paired for four categories when the other
        // child is a CLA or an unambiguous tip
        const bool paired1 = (R == 4) && (op.c2 >= 0 || __popc(m2) == 1);
        const bool paired2 = (R == 4) && (op.c1 >= 0 || __popc(m1) == 1);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const double v1 = op.c1 >= 0 ? (paired1 ? dot4_paired(P1 + 4 * i, x1[r]) : dot4_seq(P1 + 4 * i, x1[r]))
                                       : tip_sum(P1 + 4 * i, b1, m1);
          const double v2 = op.c2 >= 0 ? (paired2 ? dot4_paired(P2 + 4 * i, x2[r]) : dot4_seq(P2 + 4 * i, x2[r]))
                                       : tip_sum(P2 + 4 * i, b2, m2);
          out[i] = mul_nc(v1, v2);
        }
      } else {
#pragma unroll
        for (int i = 0; i < 4; i++) {
          double v1, v2;
          if (op.c1 >= 0) {
            const double *q = P1 + 4 * i;
            v1 = __builtin_fma(q[0], double(x1[r][0]), __builtin_fma(q[1], double(x1[r][1]),
                 __builtin_fma(q[2], double(x1[r][2]), q[3] * double(x1[r][3]))));
          } else
            v1 = tip_sum(P1 + 4 * i, b1, m1);
          if (op.c2 >= 0) {
            const double *q = P2 + 4 * i;
            v2 = __builtin_fma(q[0], double(x2[r][0]), __builtin_fma(q[1], double(x2[r][1]),
                 __builtin_fma(q[2], double(x2[r][2]), q[3] * double(x2[r][3]))));
          } else
            v2 = tip_sum(P2 + 4 * i, b2, m2);
          out[i] = v1 * v2;
        }
      }
      T o[4];
      if constexpr (sizeof(T) == 4) {
        // one seed per pattern and op: the pattern index and the first value
        if (r == 0)
          seed = mix32(unsigned(p) * 0x9e3779b1u ^
                       unsigned(static_cast<unsigned long long>(__double_as_longlong(out[0])) >> 29));
      }
#pragma unroll
      for (int i = 0; i < 4; i++) {
        mx = fmax(mx, out[i]);
        if constexpr (sizeof(T) == 8)
          o[i] = out[i];
        else
          o[i] = round_dither_bits(out[i], dither_stream(seed, unsigned(4 * r + i)));
      }
      store4_stream(dst + 256 * r, o);
    }
    if (op.flags & kOpRescale) {
      if constexpr (sizeof(T) == 8) {
        // RescaleRateHet (tree.cpp:3423-3471) on the stored slice
        if (mx < c_rescale_below) {
          if (mx < c_reduce_below) atomicOr(flags, mx < c_bail_below ? 3 : 1);
          int idx = 0;
          while (idx + 1 < kRescaleLen && c_thresh[idx + 1] > mx) idx++;
          const double mult = c_mult[idx];
          for (int r = 0; r < R; r++) {
            T v[4];
            load4(dst + 256 * r, v);
#pragma unroll
            for (int i = 0; i < 4; i++) v[i] = mul_nc(v[i], mult);
            store4(dst + 256 * r, v);
          }
          u += c_incr[idx];
        }
      } else if (mx < 0x1p-32 && mx > 0.0) {
        const int e = 1 - __builtin_amdgcn_frexp_exp(mx);  // max * 2^e in [1, 2)
        for (int r = 0; r < R; r++) {
          T v[4];
          load4(dst + 256 * r, v);
#pragma unroll
          for (int i = 0; i < 4; i++) v[i] = ldexpf(v[i], e);
          store4(dst + 256 * r, v);
        }
        u += e;
      }
    }
    ua[op.dest * size_t(Ppad) + p] = u;
  }
}

// =====================================================================
// Amino-acid CLA update (S = 20).  A block is R waves over the same 64
// patterns and wave r takes category r, so the op's P is wave-uniform.  The
// products accumulate input state by input state: for q = 0..19 every lane
// adds P[f][q] * x[q] to its 20 sums, with column q of P (contiguous in the
// transposed copy pmT) arriving by scalar loads in the FP32 form and from an
// LDS copy of the wave's two matrices in the FP64 form (whose longer
// mul-then-add chains hide the LDS latency better than scalar-cache misses).
// One child at a time, the live registers are its 20 sums plus the other
// child's; the sums run in the
// reference's q order, so the FP64 form keeps its mul-then-add sequence
// (tree.cpp:7843-7910, 8082-8160, 8537-8599).  The 20 products are held until
// the op's rescale decision (per pattern, over the R waves through a small
// LDS table) and then stored once, scaled where needed.
// =====================================================================
template <typename T, int R>
__global__ __launch_bounds__(64 * R, sizeof(T) == 4 ? 1 : 3) void k_cla_aa(T *__restrict__ cla, int *__restrict__ ua,
                                                    const unsigned char *__restrict__ tips,
                                                    const double *__restrict__ pmT, const DevClaOp *__restrict__ ops,
                                                    int *__restrict__ flags, int Ppad, int P, int fused_n) {
  constexpr int S = 20;
  constexpr bool kF32 = sizeof(T) == 4;
  constexpr bool kLdsP = !kF32;
  __shared__ double smax[R][64];
  __shared__ double sPT[kLdsP ? R : 1][2][kLdsP ? S * S : 2];  // FP64: the wave's P^T of both children
  const int r = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const size_t slot = size_t(Ppad) * R * S;
  const int first = fused_n ? 0 : blockIdx.y, last = fused_n ? fused_n : blockIdx.y + 1;
  // the lane's pattern and offset, formed again in each op from a lane index
  // the compiler cannot see through (an empty asm), so they are not hoisted
  // out of the op loop and held or spilled across it
  int lane = 0, p = 0;
  bool live = false;
  size_t off = 0;
  // acc[f] = (P x)[f] for one child: a CLA (input state by input state) or a tip
  auto child = [&](int c, int pmslot, int ci, double(&acc)[S]) {
    // PT[q * S + f] = P[f][q]
    const double *PT = kLdsP ? sPT[kLdsP ? r : 0][ci] : pmT + (size_t(pmslot) * 3 * R + r) * S * S;
    if (c >= 0) {
      T x[S];
#pragma unroll
      for (int q = 0; q < S; q++) x[q] = cla[c * slot + off + 64 * q];  // p < Ppad: inside the slot
      // FP32 form: lanes past the last pattern compute on whatever their
      // padded rows hold (nothing of theirs is stored or compared with
      // another pattern), and each sum starts from its first product
      if constexpr (!kF32) {
#pragma unroll
        for (int q = 0; q < S; q++) x[q] = live ? x[q] : T(0);
#pragma unroll
        for (int f = 0; f < S; f++) acc[f] = 0.0;
      }
#pragma unroll
      for (int q = 0; q < S; q++) {
        const double xq = double(x[q]);
        double col[S];
#pragma unroll
        for (int f = 0; f < S; f++) col[f] = PT[q * S + f];
#pragma unroll
        for (int f = 0; f < S; f++) {
          if constexpr (kF32)
            acc[f] = q == 0 ? col[f] * xq : __builtin_fma(col[f], xq, acc[f]);
          else
            acc[f] = add_nc(acc[f], mul_nc(col[f], xq));
        }
      }
    } else {
      const int tp = tips[size_t(-c - 1) * Ppad + p];
      const int t = live ? tp : S;
      if (t < S) {
        const double *row = PT + t * S;  // P[f][t] for f = 0..19
#pragma unroll
        for (int f = 0; f < S; f++) acc[f] = row[f];
      } else {
#pragma unroll
        for (int f = 0; f < S; f++) acc[f] = 1.0;
      }
    }
  };
  for (int oi = first; oi < last; ++oi) {
    if (oi != first) __syncthreads();  // smax of the previous op has been read
    lane = int(threadIdx.x & 63);
    asm volatile("" : "+v"(lane));
    p = blockIdx.x * 64 + lane;
    live = p < P;
    off = aa_off(p, r, R);
    const DevClaOp op = ops[oi];
    if constexpr (kLdsP) {
      const double2 *g1 = reinterpret_cast<const double2 *>(pmT + (size_t(op.pm1) * 3 * R + r) * S * S);
      const double2 *g2 = reinterpret_cast<const double2 *>(pmT + (size_t(op.pm2) * 3 * R + r) * S * S);
      double2 *l1 = reinterpret_cast<double2 *>(sPT[r][0]), *l2 = reinterpret_cast<double2 *>(sPT[r][1]);
      for (int k = lane; k < S * S / 2; k += 64) {
        l1[k] = g1[k];
        l2[k] = g2[k];
      }
      __syncthreads();
    }
    int u = 0;
    if (live) {
      if (op.c1 >= 0) u += ua[op.c1 * size_t(Ppad) + p];
      if (op.c2 >= 0) u += ua[op.c2 * size_t(Ppad) + p];
    }
    double a1[S], a2[S];
    child(op.c1, op.pm1, 0, a1);
    child(op.c2, op.pm2, 1, a2);
    T *dst = cla + op.dest * slot + off;
    double o[S], m = 0.0;
#pragma unroll
    for (int f = 0; f < S; f++) {
      o[f] = mul_nc(a1[f], a2[f]);
      m = fmax(m, o[f]);
    }
    // The rescale is decided before anything is stored (per pattern, over
    // the R waves through a small LDS table), so each value is written once.
    int e = 0;           // FP32 form: exact power-of-two scale
    double mult = 1.0;   // FP64 form: the reference's table multiplier
    bool scale = false;
    if (op.flags & kOpRescale) {
      smax[r][lane] = m;
      __syncthreads();
#pragma unroll
      for (int k = 0; k < R; k++) m = fmax(m, smax[k][lane]);
      if (live) {
        if constexpr (kF32) {
          if (m < 0x1p-32 && m > 0.0) {
            e = 1 - __builtin_amdgcn_frexp_exp(m);  // max * 2^e in [1, 2)
            u += e;
            scale = true;
          }
        } else if (m < c_rescale_below) {
          // RescaleRateHet / RescaleRateHetNState (tree.cpp:3423-3471, 3492-3598)
          if (m < c_reduce_below && r == 0) atomicOr(flags, m < c_bail_below ? 3 : 1);
          int idx = 0;
          while (idx + 1 < kRescaleLen && c_thresh[idx + 1] > m) idx++;
          mult = c_mult[idx];
          u += c_incr[idx];
          scale = true;
        }
      }
    }
    if (live) {
      if constexpr (kF32) {
        // dither seed of this (pattern, category) and op, from the unscaled first value
        const unsigned seed = mix32((unsigned(p) * R + unsigned(r)) * 0x9e3779b1u ^
                                    unsigned(static_cast<unsigned long long>(__double_as_longlong(o[0])) >> 29));
#pragma unroll
        for (int f = 0; f < S; f++)  // e = 0 where the pattern is not scaled
          dst[64 * f] = round_dither_bits(__builtin_amdgcn_ldexp(o[f], e), dither_stream(seed, unsigned(f)));
      } else {
#pragma unroll
        for (int f = 0; f < S; f++) dst[64 * f] = scale ? mul_nc(o[f], mult) : o[f];
      }
    }
    if (live && r == 0) ua[op.dest * size_t(Ppad) + p] = u;
  }
}

// =====================================================================
// Codon CLA update (61 states padded to 64) on FP64 MFMA, both forms.
// One wave owns 16 patterns of an op.  For each (category, child) the
// contraction out[p][i] = sum_j P[i][j] x[p][j] runs as 4 output tiles x 16
// steps of v_mfma_f64_16x16x4_f64.  The K index is permuted so that at step kk
// the four K rows of the instruction are j = 16g + kk (g = lane/16): each
// lane's A and B fragments are then 16 contiguous values of one CLA row / one
// P row.  v_mfma_f64_16x16x4_f64 D layout: lane l, element v holds row
// l/16 + 4v (pattern), column l%16 (state).  P comes from the zero-padded FP64 copy
// pmpad [slot][3][R][64*64] (P, P', P'') in codon fragment order, so the B
// fragment of step kk is one coalesced 512-byte load.  The MFMA sums in its own order,
// so FP64 codon CLAs match the reference's sequential sums to rounding
// (about 1e-16 relative), not bit for bit.
// =====================================================================
typedef double gle_doublex4 __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ void load16(const T *src, double *x) {
#pragma unroll
  for (int q = 0; q < 16; q += 4) {
    T v[4];
    load4(src + q, v);
    x[q] = double(v[0]);
    x[q + 1] = double(v[1]);
    x[q + 2] = double(v[2]);
    x[q + 3] = double(v[3]);
  }
}

// x = this lane's 16 CLA values (pattern p0 + lane%16, states 16 (lane/16)
// .. +15) of internal child c, zero beyond the last pattern.
template <typename T>
__device__ __forceinline__ void codon_load_x(const T *__restrict__ cla, int c, size_t slot, int R, int r, int lo,
                                             int g, int p0, int P, double x[16]) {
  const int p = p0 + lo;
  if (p < P)
    load16(cla + c * slot + (size_t(p) * R + r) * 64 + 16 * g, x);
  else
#pragma unroll
    for (int q = 0; q < 16; q++) x[q] = 0.0;
}
// Tip states of the patterns p0 + g + 4v this lane's D rows hold (61 =
// missing, also beyond the last pattern).
__device__ __forceinline__ void codon_tip_states(const unsigned char *__restrict__ tips, int c, int Ppad, int P,
                                                 int g, int p0, int st[4]) {
  const unsigned char *tip = tips + size_t(-c - 1) * Ppad;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int p = p0 + g + 4 * v;
    const int s = tip[p];  // p < Ppad: inside the tip row
    st[v] = p < P ? s : 61;
  }
}
// Output tile t (states 16t .. 16t+15) of Pr . x for the wave's 16 patterns.
__device__ __forceinline__ gle_doublex4 codon_mfma_tile(const double *__restrict__ Pr, const double x[16], int t,
                                                        int l) {
  const double *B = Pr + t * 1024 + l;
  double b[16];
#pragma unroll
  for (int kk = 0; kk < 16; kk++) b[kk] = B[kk * 64];
  // all 16 loads in flight before the first MFMA waits on one
  __builtin_amdgcn_sched_barrier(0);
  gle_doublex4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int kk = 0; kk < 16; kk++) d = __builtin_amdgcn_mfma_f64_16x16x4f64(x[kk], b[kk], d, 0, 0, 0);
  return d;
}
// The same tile for a tip child: column st of Pr, `missing` for a missing state.
__device__ __forceinline__ gle_doublex4 codon_tip_tile(const double *__restrict__ Pr, const int st[4], int t, int lo,
                                                       double missing) {
  gle_doublex4 d;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const double e = Pr[codon_frag(16 * t + lo, st[v] < 61 ? st[v] : 0)];
    d[v] = st[v] < 61 ? e : missing;
  }
  return d;
}

// Child c's output tiles 0..3 into ov: Pr . x for an internal child, the tip
// column for a tip child, or (msg) the product an earlier op kept in slot c.
// ov[t] = tile (MUL false) or ov[t] *= tile (MUL true), so the second child's
// tiles are never all held.  mkeep >= 0 keeps an internal child's tiles in
// that slot (a message, read back by later ops with kOpMsg).  Each tile's 16
// B fragments are loaded and its 16-step MFMA chain runs on them (the
// compiler interleaves a tile's loads with the chain's first steps; loading
// the next tile's fragments ahead of the chain spilled).  The FP32 form holds
// x in FP32 and widens it at each MFMA.  The message reads are issued
// unconditionally and back to back (rows p0 .. p0+15 lie inside the slot's
// Ppad rows): guarded, they had compiled to one branch and one wait per
// load.  The FP32 form keeps
// a message rounded to nearest without the dither: a message is a contraction
// of stored (dithered) CLA values, so its values differ from pattern to
// pattern and their roundings are already independent; the dither is for
// stores whose values repeat across patterns (a tip-tip product).
template <typename T, bool MUL>
__device__ __forceinline__ void codon_child_into(T *__restrict__ cla, const double *__restrict__ Pr, int c,
                                                 bool msg, int mkeep, const int ts[4], size_t slot, int R, int r,
                                                 int l, int p0, int P, gle_doublex4 ov[4]) {
  const int lo = l & 15, g = l >> 4;
  auto put = [&](int t, const gle_doublex4 &d) {
#pragma unroll
    for (int v = 0; v < 4; v++) ov[t][v] = MUL ? mul_nc(ov[t][v], d[v]) : d[v];
  };
  if (msg) {
    const T *src = cla + c * slot;
    T m[4][4];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int v = 0; v < 4; v++) m[t][v] = src[(size_t(p0 + g + 4 * v) * R + r) * 64 + 16 * t + lo];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      gle_doublex4 d;
#pragma unroll
      for (int v = 0; v < 4; v++) d[v] = p0 + g + 4 * v < P ? double(m[t][v]) : 0.0;
      put(t, d);
    }
    return;
  }
  if (c < 0) {
#pragma unroll
    for (int t = 0; t < 4; t++) put(t, codon_tip_tile(Pr, ts, t, lo, 1.0));
    return;
  }
  T x[16];
  {
    const int p = p0 + lo;
    if (p < P) {
      const T *src = cla + c * slot + (size_t(p) * R + r) * 64 + 16 * g;
#pragma unroll
      for (int q = 0; q < 16; q += 4) load4(src + q, x + q);
    } else {
#pragma unroll
      for (int q = 0; q < 16; q++) x[q] = T(0);
    }
  }
  const double *B = Pr + l;
#pragma unroll
  for (int t = 0; t < 4; t++) {
    double b[16];
#pragma unroll
    for (int kk = 0; kk < 16; kk++) b[kk] = B[t * 1024 + kk * 64];
    gle_doublex4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < 16; kk++) d = __builtin_amdgcn_mfma_f64_16x16x4f64(double(x[kk]), b[kk], d, 0, 0, 0);
    if (mkeep >= 0) {
      T *dst = cla + mkeep * slot;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int p = p0 + g + 4 * v;
        if (p < P) dst[(size_t(p) * R + r) * 64 + 16 * t + lo] = T(d[v]);
      }
    }
    put(t, d);
  }
}

// grid (Ppad/16, fused ? 1 : n_ops); a block is R waves over the same 16
// patterns and wave r takes category r, so that a wave's serial MFMA chain is
// one category long (small dependency levels fill the GPU R times faster).
// The products are held until the op's rescale decision (each pattern's
// maximum over its 16 lanes and over the R waves, through LDS) and stored
// once, scaled where needed.
// Ops may read children as kept messages and keep the messages they form
// (DevClaOp::m1, m2: serving ticks); the tree layer's ops set neither, and
// its sweeps run this same kernel (a form without the message paths took
// more registers, three waves per SIMD against four, and was slower).
template <typename T, int R>
__global__ __launch_bounds__(64 * R, 3) void k_cla_codon(T *__restrict__ cla, int *__restrict__ ua,
                                                         const unsigned char *__restrict__ tips,
                                                         const double *__restrict__ pmpad,
                                                         const DevClaOp *__restrict__ ops, int *__restrict__ flags,
                                                         int Ppad, int P, int fused_n) {
  constexpr bool kF32 = sizeof(T) == 4;
  __shared__ double smax[R][16];
  const int r = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int p0 = blockIdx.x * 16;
  const size_t slot = size_t(Ppad) * R * 64;
  const int first = fused_n ? 0 : blockIdx.y, last = fused_n ? fused_n : blockIdx.y + 1;
  for (int oi = first; oi < last; ++oi) {
    if (oi != first) __syncthreads();  // the previous op's stores are read by other lanes
    // The lane index is made opaque to the compiler in each op (an empty asm,
    // no instruction), so the per-lane offsets and dither keys derived from
    // it are formed where they are used instead of being hoisted out of the
    // op loop and spilled at the register cap.
    int l = int(threadIdx.x & 63);
    asm volatile("" : "+v"(l));
    const int lo = l & 15, g = l >> 4;
    const DevClaOp op = ops[oi];
    T *dst = cla + op.dest * slot;
    int ts1[4], ts2[4];
    if (op.c1 < 0) codon_tip_states(tips, op.c1, Ppad, P, g, p0, ts1);
    if (op.c2 < 0) codon_tip_states(tips, op.c2, Ppad, P, g, p0, ts2);
    double mx[4] = {0.0, 0.0, 0.0, 0.0};
    double ov[4][4];  // [pattern row v][state tile t], held until the rescale decision
    {
      gle_doublex4 a1[4];  // child 1's tiles, then the product
      const bool g1 = (op.flags & kOpMsg1) != 0, g2 = (op.flags & kOpMsg2) != 0;
      codon_child_into<T, false>(cla, g1 ? nullptr : pmpad + (size_t(op.pm1) * 3 * R + r) * 4096, op.c1, g1,
                                 op.m1, ts1, slot, R, r, l, p0, P, a1);
      codon_child_into<T, true>(cla, g2 ? nullptr : pmpad + (size_t(op.pm2) * 3 * R + r) * 4096, op.c2, g2,
                                op.m2, ts2, slot, R, r, l, p0, P, a1);
#pragma unroll
      for (int v = 0; v < 4; v++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const double o = 16 * t + lo < 61 ? a1[t][v] : 0.0;
          ov[v][t] = o;
          mx[v] = fmax(mx[v], o);
        }
    }
    // the children's rescale counts for wave 0's stores, loaded back to back
    // before the rescale exchange (rows p0 .. p0+15 lie inside the Ppad rows
    // of a slot's counts; rows past P are not stored)
    int u1[4] = {0, 0, 0, 0}, u2[4] = {0, 0, 0, 0};
    if (r == 0) {
      if (op.c1 >= 0)
#pragma unroll
        for (int v = 0; v < 4; v++) u1[v] = ua[size_t(op.c1) * Ppad + p0 + g + 4 * v];
      if (op.c2 >= 0)
#pragma unroll
        for (int v = 0; v < 4; v++) u2[v] = ua[size_t(op.c2) * Ppad + p0 + g + 4 * v];
    }
    if (op.flags & kOpRescale) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        double m = mx[v];
        for (int o = 1; o < 16; o <<= 1) m = fmax(m, __shfl_xor(m, o));
        mx[v] = m;
        if (lo == 0) smax[r][g + 4 * v] = m;
      }
      __syncthreads();
#pragma unroll
      for (int v = 0; v < 4; v++) {
        double m = smax[0][g + 4 * v];
        for (int q = 1; q < R; q++) m = fmax(m, smax[q][g + 4 * v]);
        mx[v] = m;
      }
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int p = p0 + g + 4 * v;
      int u = u1[v] + u2[v];
      // the pattern's rescale (over its R x 64 values) is applied as the values are stored
      int e = 0;
      double mult = 1.0;
      bool scale = false;
      if ((op.flags & kOpRescale) && p < P) {
        const double m = mx[v];
        if constexpr (kF32) {
          if (m < 0x1p-32 && m > 0.0) {
            e = 1 - __builtin_amdgcn_frexp_exp(m);
            u += e;
            scale = true;
          }
        } else if (m < c_rescale_below) {
          if (m < c_reduce_below && lo == 0 && r == 0) atomicOr(flags, m < c_bail_below ? 3 : 1);
          int idx = 0;
          while (idx + 1 < kRescaleLen && c_thresh[idx + 1] > m) idx++;
          mult = c_mult[idx];
          u += c_incr[idx];
          scale = true;
        }
      }
      if (p < P) {
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const int st = 16 * t + lo;
          T out;
          if constexpr (kF32)
            out = round_dither(ov[v][t], unsigned(p) * 512u + unsigned(r * 64 + st), scale ? e : 0);
          else
            out = scale ? mul_nc(ov[v][t], mult) : ov[v][t];
          dst[(size_t(p) * R + r) * 64 + st] = out;
        }
      }
      if (p < P && lo == 0 && r == 0) {
        ua[size_t(op.dest) * Ppad + p] = u;
        // a kept message carries its child's rescale counts
        if (op.m1 >= 0 && op.c1 >= 0) ua[size_t(op.m1) * Ppad + p] = u1[v];
        if (op.m2 >= 0 && op.c2 >= 0) ua[size_t(op.m2) * Ppad + p] = u2[v];
      }
    }
  }
}

// =====================================================================
// Edge integration + derivatives, FP64 arithmetic after the CLA loads.
// One thread per (pattern, rate); the rate-0 lane combines the rates in
// order, applies the invariant-site term and the log, and the block reduces
// (count*lnL, count*d1, count*d2) into per-block partials.
// =====================================================================
template <typename T>
struct EdgeArgs {
  const T *cla;
  const int *u;
  const unsigned char *tips;
  const double *pm64;  // [slot][3][R][S][S]
  const double *model; // pi[S], w[R]
  const int *counts, *cst;
  double pinv;
  int pinv_in_model;
  size_t slot_stride;
  int Ppad, P, R, S, Sp;
  double *partials;  // [op][nblk][3]
  double *site;      // per-pattern lnL for op 0, or null
};

__device__ __forceinline__ void block_reduce3(double a, double b, double c, double *out) {
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o);
    b += __shfl_down(b, o);
    c += __shfl_down(c, o);
  }
  __shared__ double red[3][8];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
    red[0][w] = a;
    red[1][w] = b;
    red[2][w] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0, s1 = 0, s2 = 0;
    for (int k = 0; k < int(blockDim.x >> 6); k++) {
      s0 += red[0][k];
      s1 += red[1][k];
      s2 += red[2][k];
    }
    out[0] = s0;
    out[1] = s1;
    out[2] = s2;
  }
}

// Combine one site: L/D1/D2 are the rate-summed, frequency-weighted values.
// FP64 form: natural-log counts as tree.cpp:6118-6139 / optimization.cpp:3166-3187.
// FP32 form: power-of-two counts; the invariant term is added at true scale.
template <bool POW2>
__device__ __forceinline__ void site_terms(double L, double D1, double D2, int u, double pinv_term, bool nstate,
                                           int count, double &lnl, double &g1, double &g2, double &ul, double &r1o,
                                           double &r2o) {
#pragma clang fp contract(off)
  double sL, r1, r2;
  if (!POW2) {
    sL = L;
    if (pinv_term > 0.0) sL = (u == 0) ? L + pinv_term : L + pinv_term * exp(double(u));
    ul = log(sL) - double(u);
    r1 = D1 / sL;
    r2 = (D2 / sL) - r1 * r1;
  } else {
    // invariant-site patterns are summed at true scale (2^-u); others keep
    // the scale and subtract u ln 2.  Branch-free, one log, one division.
    const bool inv_site = pinv_term > 0.0;
    const double sc = inv_site ? ldexp(1.0, -u) : 1.0;
    sL = inv_site ? L * sc + pinv_term : L;
    const double rs = 1.0 / sL;
    ul = log(sL) - (inv_site ? 0.0 : double(u) * 0.69314718055994530942);
    r1 = D1 * sc * rs;
    r2 = D2 * sc * rs - r1 * r1;
  }
  r1o = r1;
  r2o = r2;
  if (nstate && !(ul < 0.0)) {  // NState scoring skips non-negative site lnL (tree.cpp:6473)
    lnl = g1 = g2 = 0.0;
    return;
  }
  lnl = double(count) * ul;
  g1 = double(count) * r1;
  g2 = double(count) * r2;
}

// Per-pattern values of an edge evaluated with site output, in five sections
// of Ppad doubles: site lnL, d1 = D1/L, d2 = D2/L - d1^2, and the rescale
// counts of the partial side and of the child side (-1 for a tip child).
__device__ __forceinline__ void site_out(double *site, int Ppad, int p, double ul, double r1, double r2, int u1,
                                         int u2) {
  site[p] = ul;
  site[Ppad + p] = r1;
  site[2 * size_t(Ppad) + p] = r2;
  site[3 * size_t(Ppad) + p] = double(u1);
  site[4 * size_t(Ppad) + p] = double(u2);
}

// One thread per pattern over all categories.  P, P', P'' of the op and the
// model are wave-uniform and reach the FMAs through scalar loads (every pointer
// argument is __restrict__).  Per state the categories accumulate in order, as
// GetScorePartial*RateHet / GetDerivsPartial* do (tree.cpp:6208-6230,
// optimization.cpp:3123-3158).  grid (Ppad/256, n_ops).
//
// FP32 form with derivatives (D): the spectral route of k_edge_aa.  With
// P = V diag(e) V^-1 (e = exp(s t), s = lambda times the category rate),
// sum_s pi_s x_s (M y)_s = sum_k a_k z_k f_k for M = P, P', P'' and
// f = e, s e, s^2 e, where a = V^T (pi x) and z = V^-1 y: two 4 x 4 products
// and three 4-term sums per category instead of three 4 x 4 products (about a
// third fewer FP64 instructions in a kernel that is partly bound by their
// issue).  The FP64 form keeps the reference's order of operations.
template <typename T, int R, bool D, bool SPEC = false>
__global__ __launch_bounds__(256) void k_edge_nuc(const T *__restrict__ cla, const int *__restrict__ ua,
                                                  const unsigned char *__restrict__ tips,
                                                  const double *__restrict__ pm64, const double *__restrict__ model,
                                                  const int *__restrict__ counts, const int *__restrict__ cst,
                                                  const DevEdgeOp *__restrict__ ops, double *__restrict__ partials,
                                                  double *__restrict__ site, double pinv, int pinv_in_model,
                                                  int Ppad, int P, const double *__restrict__ Vg,
                                                  const double *__restrict__ Wg, const double *__restrict__ pdiag) {
#pragma clang fp contract(off)
  const DevEdgeOp op = ops[blockIdx.y];
  const double *Pb = pm64 + size_t(op.pm) * 3 * R * 16;
  const size_t slot = size_t(Ppad) * R * 4;
  const int p = blockIdx.x * 256 + threadIdx.x;
  const double *pi = model, *w = model + 4;
  double lnl = 0, g1 = 0, g2 = 0;
  if (p < P) {
    double L[4] = {0, 0, 0, 0}, G1[4] = {0, 0, 0, 0}, G2[4] = {0, 0, 0, 0};
    int u = ua[op.partial * size_t(Ppad) + p];
    const T *xp = cla + op.partial * slot + nuc_off(p, 0, R);
    const bool internal = op.child >= 0;
    int m = 15;
    T xt[R][4], yt[R][4];
#pragma unroll
    for (int r = 0; r < R; r++) load4(xp + 256 * r, xt[r]);
    if (internal) {
      const T *yp = cla + op.child * slot + nuc_off(p, 0, R);
#pragma unroll
      for (int r = 0; r < R; r++) load4(yp + 256 * r, yt[r]);
      u += ua[op.child * size_t(Ppad) + p];
    } else {
      m = tips[size_t(-op.child - 1) * Ppad + p];
    }
    double bit[4];
#pragma unroll
    for (int q = 0; q < 4; q++) bit[q] = double((m >> q) & 1);
    // FP64 form: the reference's sequential sums; FP32 form: fused multiply-adds
    auto dot = [](const double *q, const double *v) {
      if constexpr (sizeof(T) == 8)
        return dot4_seq(q, v);
      else
        return __builtin_fma(q[3], v[3], __builtin_fma(q[2], v[2], __builtin_fma(q[1], v[1], q[0] * v[0])));
    };
    constexpr bool spectral = SPEC && D && sizeof(T) == 4;
    double Ls = 0.0, D1s = 0.0, D2s = 0.0;  // spectral route: the category-weighted sums
    if constexpr (spectral) {
      // (the host takes this route only when one eigensystem serves every category)
      double Vm[16], Wm[16];
      load_uniform16(Vg, Vm);  // V[i][k]
      load_uniform16(Wg, Wm);  // V^-1[k][j]
#pragma unroll
      for (int r = 0; r < R; r++) {
        const double *f = pdiag + (size_t(op.pm) * R + r) * 12;
        double xp[4], a[4], z[4];
#pragma unroll
        for (int st = 0; st < 4; st++) xp[st] = double(xt[r][st]) * pi[st];
#pragma unroll
        for (int k = 0; k < 4; k++)
          a[k] = __builtin_fma(Vm[12 + k], xp[3], __builtin_fma(Vm[8 + k], xp[2],
                 __builtin_fma(Vm[4 + k], xp[1], Vm[k] * xp[0])));
        if (internal) {
          const double y[4] = {double(yt[r][0]), double(yt[r][1]), double(yt[r][2]), double(yt[r][3])};
#pragma unroll
          for (int k = 0; k < 4; k++)
            z[k] = __builtin_fma(Wm[4 * k + 3], y[3], __builtin_fma(Wm[4 * k + 2], y[2],
                   __builtin_fma(Wm[4 * k + 1], y[1], Wm[4 * k] * y[0])));
        } else {
#pragma unroll
          for (int k = 0; k < 4; k++)
            z[k] = __builtin_fma(bit[3], Wm[4 * k + 3], __builtin_fma(bit[2], Wm[4 * k + 2],
                   __builtin_fma(bit[1], Wm[4 * k + 1], bit[0] * Wm[4 * k])));
        }
        double l0 = 0.0, l1 = 0.0, l2 = 0.0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const double c = a[k] * z[k];
          l0 = __builtin_fma(c, f[k], l0);
          l1 = __builtin_fma(c, f[4 + k], l1);
          l2 = __builtin_fma(c, f[8 + k], l2);
        }
        if (!internal && m == 15) {  // P 1 = 1: the partial side alone, no derivative
          l0 = xp[0] + xp[1] + xp[2] + xp[3];
          l1 = l2 = 0.0;
        }
        Ls = __builtin_fma(l0, w[r], Ls);
        D1s = __builtin_fma(l1, w[r], D1s);
        D2s = __builtin_fma(l2, w[r], D2s);
      }
    }
    // P, then P' and P'' when the launch computes derivatives (D)
    constexpr int nm = D ? 3 : 1;
#pragma unroll
    for (int r = 0; r < (spectral ? 0 : R); r++) {
      const double x[4] = {double(xt[r][0]), double(xt[r][1]), double(xt[r][2]), double(xt[r][3])};
      const double y[4] = {double(yt[r][0]), double(yt[r][1]), double(yt[r][2]), double(yt[r][3])};
#pragma unroll
      for (int mi = 0; mi < nm; mi++) {
        double M[16], t[4];
        load_uniform16(Pb + (mi * R + r) * 16, M);
        if (internal) {
#pragma unroll
          for (int s = 0; s < 4; s++) t[s] = dot(M + 4 * s, y) * x[s] * w[r];
        } else {
          // tip: ascending-order sums over the mask's states; a missing tip
          // contributes x * w to L and nothing to the derivatives
#pragma unroll
          for (int s = 0; s < 4; s++) t[s] = (mi > 0 && m == 15) ? 0.0 : tip_sum(M + 4 * s, bit, m) * x[s] * w[r];
        }
        // constant indices keep L, G1, G2 in registers
#pragma unroll
        for (int s = 0; s < 4; s++) {
          if (mi == 0)
            L[s] += t[s];
          else if (mi == 1)
            G1[s] += t[s];
          else
            G2[s] += t[s];
        }
      }
    }
    const double Lt = spectral ? Ls : L[0] * pi[0] + L[1] * pi[1] + L[2] * pi[2] + L[3] * pi[3];
    const double D1 = spectral ? D1s : G1[0] * pi[0] + G1[1] * pi[1] + G1[2] * pi[2] + G1[3] * pi[3];
    const double D2 = spectral ? D2s : G2[0] * pi[0] + G2[1] * pi[1] + G2[2] * pi[2] + G2[3] * pi[3];
    double pt = 0.0;
    const int cs = cst[p];
    if (pinv_in_model && cs >= 0) {
      double btot = 0.0;
      for (int s = 0; s < 4; s++)
        if (cs & (1 << s)) btot += pi[s];
      pt = pinv * btot;
    }
    double ul, r1, r2;
    site_terms<sizeof(T) == 4>(Lt, D1, D2, u, pt, false, counts[p], lnl, g1, g2, ul, r1, r2);
    // the two sides' counts are read again here rather than kept live (the
    // kernel sits just under the 128-VGPR line)
    if (site && blockIdx.y == 0)
      site_out(site, Ppad, p, ul, r1, r2, ua[op.partial * size_t(Ppad) + p],
               internal ? ua[op.child * size_t(Ppad) + p] : -1);
  }
  block_reduce3(lnl, g1, g2, partials + (size_t(blockIdx.y) * gridDim.x + blockIdx.x) * 3);
}

// Codon edge: per category, Y = P.y (and P'.y, P''.y) on FP64 MFMA in the
// layout of k_cla_codon, then each lane weights its (pattern, state) cells by
// x * pi, the 16 lanes of a pattern row reduce, and the categories combine in
// order with their weights (GetScorePartialNState / GetDerivsPartialNState,
// tree.cpp:6431-6480, optimization.cpp:3230-3290).  grid (Ppad/16, n_ops); a
// block is R waves over the same 16 patterns and wave r takes category r; the
// per-category row sums meet in LDS, where wave 0 weights them in category
// order and finishes the sites.  D = false: no op of the launch wants
// derivatives, and the four tiles are unrolled so that a tile's P fragments
// load during the previous tile's MFMAs; that form fits four waves per SIMD
// (its products in VGPRs rather than accumulation registers).
template <typename T, int R, bool D>
__global__ __launch_bounds__(64 * R, D ? 3 : 4) void k_edge_codon(const T *__restrict__ cla, const int *__restrict__ ua,
                                                   const unsigned char *__restrict__ tips,
                                                   const double *__restrict__ pmpad, const double *__restrict__ pdiag,
                                                   const double *__restrict__ eigfrag,
                                                   const int *__restrict__ eig_of_rate,
                                                   const double *__restrict__ model,
                                                   const int *__restrict__ counts, const int *__restrict__ cst,
                                                   const DevEdgeOp *__restrict__ ops, double *__restrict__ partials,
                                                   double *__restrict__ site, double pinv, int pinv_in_model,
                                                   int Ppad, int P) {
  __shared__ double srow[3][R][16];
  const int l = threadIdx.x & 63, lo = l & 15, g = l >> 4;
  const int r = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const int p0 = blockIdx.x * 16;
  const size_t slot = size_t(Ppad) * R * 64;
  const DevEdgeOp op = ops[blockIdx.y];
  const double *pi = model, *w = model + 61;
  const int nm = D && op.derivs ? 3 : 1;
  double sL[4] = {0, 0, 0, 0}, s1[4] = {0, 0, 0, 0}, s2[4] = {0, 0, 0, 0};
  int ts[4];
  if (op.child < 0) codon_tip_states(tips, op.child, Ppad, P, g, p0, ts);
  {
    double acc[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};  // [matrix][v]: sum over this lane's states
    double y[16];
    if (op.child >= 0 && !op.msg) codon_load_x(cla, op.child, slot, R, r, lo, g, p0, P, y);
    // FP32 form, derivatives on an internal child: spectral route.  With
    // P = V diag(ex) V^-1, z = V^-1 y and a = V^T (x * pi) are two MFMA
    // contractions per tile (instead of P y, P' y and P'' y), and each cell
    // adds a z times ex, s ex and s^2 ex.  The FP64 form keeps the P matrices.
    const bool spectral = D && sizeof(T) == 4 && op.derivs && op.child >= 0;
    if (spectral) {
      double xa[16];
      codon_load_x(cla, op.partial, slot, R, r, lo, g, p0, P, xa);
#pragma unroll
      for (int kk = 0; kk < 16; kk++) {
        const int st = 16 * g + kk;
        xa[kk] *= st < 61 ? pi[st < 61 ? st : 0] : 0.0;
      }
      const double *Wf = eigfrag + size_t(eig_of_rate[r]) * 8192, *VTf = Wf + 4096;
      const double *dg = pdiag + (size_t(op.pm) * R + r) * 3 * 61;
#pragma unroll 1
      for (int t = 0; t < 4; t++) {
        const int st = 16 * t + lo, sc = st < 61 ? st : 0;
        const double on = st < 61 ? 1.0 : 0.0;
        const gle_doublex4 Z = codon_mfma_tile(Wf, y, t, l);
        const gle_doublex4 A = codon_mfma_tile(VTf, xa, t, l);
        const double f0 = on * dg[sc], f1 = on * dg[61 + sc], f2 = on * dg[122 + sc];
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const double c = A[v] * Z[v];
          acc[0][v] += c * f0;
          acc[1][v] += c * f1;
          acc[2][v] += c * f2;
        }
      }
    }
    auto tile = [&](int t) {
      const int st = 16 * t + lo;
      double xw[4];  // x * pi of the partial side for this lane's cells
      const double ps = st < 61 ? pi[st < 61 ? st : 0] : 0.0;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        // unconditional load (rows up to p0+15 lie inside the slot's Ppad rows)
        const int p = p0 + g + 4 * v;
        const double xv = double(cla[op.partial * slot + (size_t(p) * R + r) * 64 + st]);
        xw[v] = p < P ? xv * ps : 0.0;
      }
#pragma unroll
      for (int mi = 0; mi < 3; mi++) {
        if (mi >= nm) break;
        const double *Pm = pmpad + ((size_t(op.pm) * 3 + mi) * R + r) * 4096;
        // a missing tip contributes no derivative
        const gle_doublex4 Y = op.child >= 0 ? codon_mfma_tile(Pm, y, t, l)
                                             : codon_tip_tile(Pm, ts, t, lo, mi == 0 ? 1.0 : 0.0);
#pragma unroll
        for (int v = 0; v < 4; v++) acc[mi][v] += Y[v] * xw[v];
      }
    };
    if constexpr (D) {
#pragma unroll 1
      for (int t = 0; t < (spectral ? 0 : 4); t++) tile(t);
    } else {
      // Y = P y for the four tiles first, with no scheduling barrier between
      // them, then the x * pi weighting
      const double *Pm = pmpad + (size_t(op.pm) * 3 * R + r) * 4096;
      gle_doublex4 Y[4];
      if (op.child >= 0 && op.msg) {
        // the child is a kept message P(t) . x (a serving tick's score whose
        // child side a CLA op of the tick contracted at this length): read,
        // not contracted again; rows p0 .. p0+15 lie inside the slot's Ppad
        // rows, and rows past P weigh zero
        const T *src = cla + op.child * slot;
        T m[4][4];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
          for (int v = 0; v < 4; v++) m[t][v] = src[(size_t(p0 + g + 4 * v) * R + r) * 64 + 16 * t + lo];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
          for (int v = 0; v < 4; v++) Y[t][v] = p0 + g + 4 * v < P ? double(m[t][v]) : 0.0;
      } else if (op.child >= 0) {
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const double *B = Pm + t * 1024 + l;
          double b[16];
#pragma unroll
          for (int kk = 0; kk < 16; kk++) b[kk] = B[kk * 64];
          gle_doublex4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int kk = 0; kk < 16; kk++) d = __builtin_amdgcn_mfma_f64_16x16x4f64(y[kk], b[kk], d, 0, 0, 0);
          Y[t] = d;
        }
      } else {
#pragma unroll
        for (int t = 0; t < 4; t++) Y[t] = codon_tip_tile(Pm, ts, t, lo, 1.0);
      }
      // the partial side's 16 values are loaded unconditionally and back to
      // back (rows up to p0+15 lie inside the slot's Ppad rows); rows past P
      // and the padding states weigh zero
      T xp[4][4];
#pragma unroll
      for (int t = 0; t < 4; t++)
#pragma unroll
        for (int v = 0; v < 4; v++) xp[t][v] = cla[op.partial * slot + (size_t(p0 + g + 4 * v) * R + r) * 64 + 16 * t + lo];
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const int st = 16 * t + lo;
        const double ps = st < 61 ? pi[st < 61 ? st : 0] : 0.0;
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const double xw = p0 + g + 4 * v < P ? double(xp[t][v]) * ps : 0.0;
          acc[0][v] += Y[t][v] * xw;
        }
      }
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
      double a0 = acc[0][v], a1 = acc[1][v], a2 = acc[2][v];
      for (int o = 1; o < 16; o <<= 1) {
        a0 += __shfl_xor(a0, o);
        a1 += __shfl_xor(a1, o);
        a2 += __shfl_xor(a2, o);
      }
      if (lo == 0) {
        srow[0][r][g + 4 * v] = a0;
        srow[1][r][g + 4 * v] = a1;
        srow[2][r][g + 4 * v] = a2;
      }
    }
  }
  __syncthreads();
  if (r == 0 && lo == 0)
#pragma unroll
    for (int v = 0; v < 4; v++)
      for (int q = 0; q < R; q++) {
        sL[v] += srow[0][q][g + 4 * v] * w[q];
        s1[v] += srow[1][q][g + 4 * v] * w[q];
        s2[v] += srow[2][q][g + 4 * v] * w[q];
      }
  double lnl = 0, g1 = 0, g2 = 0;
  if (r == 0 && lo == 0) {
    // the four patterns' counts, loaded back to back (every array here has
    // Ppad rows)
    int u1v[4], u2v[4], csv[4], cnv[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int p = p0 + g + 4 * v;
      u1v[v] = ua[size_t(op.partial) * Ppad + p];
      u2v[v] = op.child >= 0 ? ua[size_t(op.child) * Ppad + p] : -1;
      csv[v] = cst[p];
      cnv[v] = counts[p];
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int p = p0 + g + 4 * v;
      if (p >= P) continue;
      const int u1 = u1v[v], u2 = u2v[v];
      const int cs = csv[v];
      const bool inv = pinv_in_model && cs >= 0 && cs < 61;
      const double pcs = pi[inv ? cs : 0];
      const double pt = inv ? pinv * pcs : 0.0;
      double l1, d1, d2, ul, r1, r2;
      site_terms<sizeof(T) == 4>(sL[v], s1[v], s2[v], u1 + (u2 > 0 ? u2 : 0), pt, true, cnv[v], l1, d1, d2, ul,
                                 r1, r2);
      lnl += l1;
      g1 += d1;
      g2 += d2;
      if (site && blockIdx.y == 0) site_out(site, Ppad, p, ul, r1, r2, u1, u2);
    }
  }
  block_reduce3(lnl, g1, g2, partials + (size_t(blockIdx.y) * gridDim.x + blockIdx.x) * 3);
}

// Amino-acid edge: the block is R waves over 64 patterns, wave r takes
// category r (P, P', P'' wave-uniform, staged in LDS), and the per-category
// sums meet in LDS, where wave 0 weights them in category order and finishes
// the site (GetScorePartialNState / GetDerivsPartialNState, tree.cpp:6431-6480,
// optimization.cpp:3230-3290).  grid (Ppad/64, n_ops).
template <typename T, int R>
__global__ __launch_bounds__(64 * R) void k_edge_aa(const T *__restrict__ cla, const int *__restrict__ ua,
                                                    const unsigned char *__restrict__ tips,
                                                    const double *__restrict__ pmT, const double *__restrict__ pdiag,
                                                    const double *__restrict__ Vg, const double *__restrict__ Wg,
                                                    const int *__restrict__ eig_of_rate,
                                                    const double *__restrict__ model,
                                                    const int *__restrict__ counts, const int *__restrict__ cst,
                                                    const DevEdgeOp *__restrict__ ops, double *__restrict__ partials,
                                                    double *__restrict__ site, double pinv, int pinv_in_model,
                                                    int Ppad, int P) {
#pragma clang fp contract(off)
  constexpr int S = 20;
  __shared__ double sred[3][R][64];
  // the wave's P, P', P'' (transposed) staged in LDS and read as broadcast
  // columns: scalar loads of them missed the scalar cache (0.76 -> 0.66 ms
  // per FP32 sweep, 1.29 -> 0.82 ms FP64)
  __shared__ double sM[R][3][S * S];
  const int lane = threadIdx.x & 63;
  const int r = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const int p = blockIdx.x * 64 + lane;
  const DevEdgeOp op = ops[blockIdx.y];
  const size_t slot = size_t(Ppad) * R * S, off = aa_off(p, r, R);
  // transposed P, P', P'' of category r: PT[(mi * R) * S * S + q * S + f] = M_mi[f][q]
  const double *PT = pmT + (size_t(op.pm) * 3 * R + r) * S * S;
  const double *pi = model, *w = model + S;
  // FP32 form, derivatives on an internal child: spectral route.  With
  // P = V diag(ex) V^-1, z = V^-1 y and a = V^T (x * pi), the three sums are
  // sum_k a_k z_k {ex_k, s ex_k, s^2 ex_k}: two 20 x 20 products per pattern
  // and category instead of three (P' y and P'' y need no matrices of their
  // own).  The FP64 form keeps the reference's order of operations.
  const bool spectral = sizeof(T) == 4 && op.derivs && op.child >= 0;
  if (spectral) {
    // sM[r][0]: V (row f = V[f][0..S)); sM[r][1]: V^-1 transposed (row q = V^-1[0..S)[q])
    const int e = eig_of_rate[r];
    const double *Ve = Vg + size_t(e) * S * S, *We = Wg + size_t(e) * S * S;
    for (int k = lane; k < S * S; k += 64) {
      sM[r][0][k] = Ve[k];
      sM[r][1][(k % S) * S + k / S] = We[k];
    }
    __syncthreads();
  } else {
    const int nmat = op.derivs ? 3 : 1;
    for (int mi = 0; mi < nmat; mi++) {
      const double2 *g = reinterpret_cast<const double2 *>(PT + size_t(mi) * R * S * S);
      double2 *d = reinterpret_cast<double2 *>(sM[r][mi]);
      for (int k = lane; k < S * S / 2; k += 64) d[k] = g[k];
    }
    __syncthreads();
  }
  auto mat = [&](int mi) -> const double * { return sM[r][mi]; };  // matrix mi of this category
  double rr[3] = {0, 0, 0};  // this category's L, D1, D2 for the pattern
  if (p < P) {
    T x[S];
#pragma unroll
    for (int f = 0; f < S; f++) x[f] = cla[op.partial * slot + off + 64 * f];
    const int nm = op.derivs ? 3 : 1;
    if (spectral) {
      const double *Vs = sM[r][0], *WT = sM[r][1];
      const double *dg = pdiag + (size_t(op.pm) * R + r) * 3 * S;  // ex, s ex, s^2 ex
      double z[S];
      {
        T y[S];
#pragma unroll
        for (int q = 0; q < S; q++) y[q] = cla[op.child * slot + off + 64 * q];
#pragma unroll
        for (int k = 0; k < S; k++) z[k] = 0.0;
#pragma unroll
        for (int q = 0; q < S; q++) {
          const double yq = double(y[q]);
#pragma unroll
          for (int k = 0; k < S; k++) z[k] = __builtin_fma(WT[q * S + k], yq, z[k]);
        }
      }
      double av[S];
#pragma unroll
      for (int k = 0; k < S; k++) av[k] = 0.0;
#pragma unroll
      for (int f = 0; f < S; f++) {
        const double xf = double(x[f]) * pi[f];
#pragma unroll
        for (int k = 0; k < S; k++) av[k] = __builtin_fma(Vs[f * S + k], xf, av[k]);
      }
      double l0 = 0.0, l1 = 0.0, l2 = 0.0;
#pragma unroll
      for (int k = 0; k < S; k++) {
        const double c = av[k] * z[k];
        l0 = __builtin_fma(c, dg[k], l0);
        l1 = __builtin_fma(c, dg[S + k], l1);
        l2 = __builtin_fma(c, dg[2 * S + k], l2);
      }
      rr[0] = l0;
      rr[1] = l1;
      rr[2] = l2;
    } else if (op.child >= 0) {
      T y[S];
#pragma unroll
      for (int q = 0; q < S; q++) y[q] = cla[op.child * slot + off + 64 * q];
      // (M y)[f] input state by input state (the reference's q order), one
      // matrix at a time, then the x * pi weighting in f order
#pragma unroll 1
      for (int mi = 0; mi < nm; mi++) {
        const double *MT = mat(mi);
        double acc[S];
#pragma unroll
        for (int f = 0; f < S; f++) acc[f] = 0.0;
#pragma unroll
        for (int q = 0; q < S; q++) {
          const double yq = double(y[q]);
          double col[S];
#pragma unroll
          for (int f = 0; f < S; f++) col[f] = MT[q * S + f];
#pragma unroll
          for (int f = 0; f < S; f++) {
            if constexpr (sizeof(T) == 8)
              acc[f] += col[f] * yq;
            else
              acc[f] = __builtin_fma(col[f], yq, acc[f]);
          }
        }
        double v = 0.0;
#pragma unroll
        for (int f = 0; f < S; f++) v += acc[f] * double(x[f]) * pi[f];
        if (mi == 0)
          rr[0] = v;
        else if (mi == 1)
          rr[1] = v;
        else
          rr[2] = v;
      }
    } else {
      const int t = tips[size_t(-op.child - 1) * Ppad + p];
      if (t < S) {
#pragma unroll 1
        for (int mi = 0; mi < nm; mi++) {
          const double *row = mat(mi) + t * S;  // M[f][t], f = 0..19
          double v = 0.0;
#pragma unroll
          for (int f = 0; f < S; f++) v += row[f] * double(x[f]) * pi[f];
          if (mi == 0)
            rr[0] = v;
          else if (mi == 1)
            rr[1] = v;
          else
            rr[2] = v;
        }
      } else {
#pragma unroll
        for (int f = 0; f < S; f++) rr[0] += double(x[f]) * pi[f];
      }
    }
  }
  const double rL = rr[0], r1 = rr[1], r2 = rr[2];
  sred[0][r][lane] = rL;
  sred[1][r][lane] = r1;
  sred[2][r][lane] = r2;
  __syncthreads();
  double lnl = 0, g1 = 0, g2 = 0;
  if (r == 0 && p < P) {
    double sL = 0, s1 = 0, s2 = 0;
#pragma unroll
    for (int k = 0; k < R; k++) {
      sL += sred[0][k][lane] * w[k];
      s1 += sred[1][k][lane] * w[k];
      s2 += sred[2][k][lane] * w[k];
    }
    const int u1 = ua[op.partial * size_t(Ppad) + p];
    const int u2 = op.child >= 0 ? ua[op.child * size_t(Ppad) + p] : -1;
    double pt = 0.0;
    const int cs = cst[p];
    if (pinv_in_model && cs >= 0 && cs < S) pt = pinv * pi[cs];
    double ul, r1, r2;
    site_terms<sizeof(T) == 4>(sL, s1, s2, u1 + (u2 > 0 ? u2 : 0), pt, true, counts[p], lnl, g1, g2, ul, r1, r2);
    if (site && blockIdx.y == 0) site_out(site, Ppad, p, ul, r1, r2, u1, u2);
  }
  block_reduce3(lnl, g1, g2, partials + (size_t(blockIdx.y) * gridDim.x + blockIdx.x) * 3);
}

__global__ __launch_bounds__(64) void k_edge_reduce(const double *__restrict__ part, int nblk,
                                                    double *__restrict__ out) {
  const int op = blockIdx.x, l = threadIdx.x;
  double a = 0, b = 0, c = 0;
  for (int k = l; k < nblk; k += 64) {
    const double *q = part + (size_t(op) * nblk + k) * 3;
    a += q[0];
    b += q[1];
    c += q[2];
  }
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o);
    b += __shfl_down(b, o);
    c += __shfl_down(c, o);
  }
  if (l == 0) {
    out[3 * op] = a;
    out[3 * op + 1] = b;
    out[3 * op + 2] = c;
  }
}

// The edges of both parts of a serving tick reduced in one launch: block
// op < n reduces op of part a (na partials each), block n + k op k of part b
// (nb each), in k_edge_reduce's order; results to out[3 * block].
__global__ __launch_bounds__(64) void k_edge_reduce2(const double *__restrict__ pa, int na,
                                                     const double *__restrict__ pb, int nb, int n,
                                                     double *__restrict__ out) {
  const int blk = blockIdx.x, l = threadIdx.x;
  const bool second = blk >= n;
  const double *part = second ? pb : pa;
  const int op = second ? blk - n : blk, nblk = second ? nb : na;
  double a = 0, b = 0, c = 0;
  for (int k = l; k < nblk; k += 64) {
    const double *q = part + (size_t(op) * nblk + k) * 3;
    a += q[0];
    b += q[1];
    c += q[2];
  }
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o);
    b += __shfl_down(b, o);
    c += __shfl_down(c, o);
  }
  if (l == 0) {
    out[3 * blk] = a;
    out[3 * blk + 1] = b;
    out[3 * blk + 2] = c;
  }
}

// Plain device copy of known size: calibrates the HBM byte counters
// (FETCH_SIZE / WRITE_SIZE) against the CLA kernels' measured traffic.
__global__ __launch_bounds__(256) void k_copy_calibrate(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                        size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

// FP64 MFMA with register operands only (four independent accumulator
// chains per wave): the achievable v_mfma_f64_16x16x4_f64 rate, against which
// the codon kernels' MFMA utilisation is read beside the published peak.
__global__ __launch_bounds__(256) void k_mfma_calibrate(double *__restrict__ out, int iters) {
  gle_doublex4 d0 = {0.0, 0.0, 0.0, 0.0}, d1 = d0, d2 = d0, d3 = d0;
  const double a = threadIdx.x * 1e-3, b = blockIdx.x * 1e-3 + 1.0;
  for (int i = 0; i < iters; i++) {
    d0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, d0, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, d1, 0, 0, 0);
    d2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, d2, 0, 0, 0);
    d3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, d3, 0, 0, 0);
  }
  out[size_t(blockIdx.x) * blockDim.x + threadIdx.x] = (d0[0] + d1[1]) + (d2[2] + d3[3]);
}
// The same for v_mfma_f64_4x4x4_4b_f64 (four 4x4x4 blocks, 512 flops), eight chains.
__global__ __launch_bounds__(256) void k_mfma_calibrate44(double *__restrict__ out, int iters) {
  double d[8];
#pragma unroll
  for (int c = 0; c < 8; c++) d[c] = 0.0;
  const double a = threadIdx.x * 1e-3, b = blockIdx.x * 1e-3 + 1.0;
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int c = 0; c < 8; c++) d[c] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, d[c], 0, 0, 0);
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < 8; c++) s += d[c];
  out[size_t(blockIdx.x) * blockDim.x + threadIdx.x] = s;
}

// =====================================================================
// Host-side launchers
// =====================================================================

// Every CLA kernel can walk a whole topologically ordered op list in one launch:
// a (pattern, rate) lane reads only what it wrote itself, and the kernels that
// share data across lanes (LDS-staged P, the codon MFMA tiles) barrier per op.
// The fused walk gives each thread a whole op list, so the grid holds only one
// thread per pattern: worthwhile only once that fills the chip several times
// over (256 CUs x 4 SIMDs x 8 waves x 64 lanes = 512k patterns).  Below that,
// launching per dependency level exposes far more parallelism.
bool engine_supports_fused(const Instance *I) { return I->Ppad >= I->fused_min_patterns; }

// Dependency levels of a topologically ordered op list: an op's level is one
// more than the level of the ops that wrote its children.  Returns false when a
// slot is written twice or written after being read (then order must be kept).
static bool level_ops(const DevClaOp *ops, int n, std::vector<DevClaOp> &sorted, std::vector<int> &start) {
  std::unordered_map<int, int> lvl_of_slot;
  std::unordered_set<int> read;
  std::vector<int> lvl(n);
  int maxl = 0;
  for (int i = 0; i < n; i++) {
    const DevClaOp &o = ops[i];
    if (lvl_of_slot.count(o.dest) || read.count(o.dest)) return false;
    int l = 0;
    for (int c : {o.c1, o.c2}) {
      if (c < 0) continue;
      read.insert(c);
      auto it = lvl_of_slot.find(c);
      if (it != lvl_of_slot.end()) l = std::max(l, it->second + 1);
    }
    lvl[i] = l;
    lvl_of_slot[o.dest] = l;
    maxl = std::max(maxl, l);
  }
  start.assign(maxl + 2, 0);
  for (int i = 0; i < n; i++) start[lvl[i] + 1]++;
  for (int l = 0; l <= maxl; l++) start[l + 1] += start[l];
  sorted.resize(n);
  std::vector<int> pos(start.begin(), start.end() - 1);
  for (int i = 0; i < n; i++) sorted[pos[lvl[i]]++] = ops[i];
  return true;
}

// Work on an instance and its heavy-pattern part goes out on both streams
// before either is waited for, so the two parts overlap on the device.
static int finish(Instance *I) {
  GLE_HIP(hipGetLastError());
  GLE_HIP(hipStreamSynchronize(I->stream));
  timed_collect(I);
  return GLE_SUCCESS;
}
static int finish_both(Instance *I, int rc) {
  if (I->heavy) {
    const int rh = finish(I->heavy.get());
    if (!rc) rc = rh;
  }
  const int rm = finish(I);
  return rc ? rc : rm;
}

static int enqueue_pmats(Instance *I, const int *idx, const double *t, int n, int derivs) {
  if (n <= 0) return GLE_SUCCESS;
  for (int i = 0; i < n; i++)
    if (idx[i] < 0 || idx[i] >= I->n_pm || !(t[i] >= 0.0))
      return set_error(GLE_ERR_ARG, "pmat op %d: slot %d / length %g out of range", i, idx[i], t[i]);
  if (n > I->pm_cap) {
    if (I->d_pmidx) GLE_HIP(hipFree(I->d_pmidx));
    if (I->d_pmt) GLE_HIP(hipFree(I->d_pmt));
    I->d_pmidx = nullptr;
    I->d_pmt = nullptr;
    const int cap = std::max(n, 2 * I->pm_cap);
    GLE_HIP(hipMalloc(&I->d_pmidx, sizeof(int) * cap));
    GLE_HIP(hipMalloc(&I->d_pmt, sizeof(double) * cap));
    I->pm_cap = cap;
  }
  GLE_HIP(hipMemcpyAsync(I->d_pmidx, idx, sizeof(int) * n, hipMemcpyHostToDevice, I->stream));
  GLE_HIP(hipMemcpyAsync(I->d_pmt, t, sizeof(double) * n, hipMemcpyHostToDevice, I->stream));
  const int S = I->S;
  const size_t shmem = sizeof(double) * (2 * S * S + 3 * S);
  timed_begin(I, 2, double(n) * I->R * S * S * 8.0 * (derivs ? 3 : 1));
  GLE_LAUNCH(I, k_pmat, dim3(n, I->R, (pmat_tiles(S) + 255) / 256), dim3(256), shmem, S, I->Sp, I->R, I->d_pmidx, I->d_pmt,
                     I->d_eig_of_rate, I->d_model + S + I->R, I->d_V, I->d_Vinv, I->d_lam, I->d_pm64, I->d_pmpad,
                     I->d_pdiag, derivs, S == 20 ? kPadTransposed : S == 61 ? kPadCodonFrag : kPadRowMajor);
  timed_end(I);
  GLE_HIP(hipGetLastError());
  return GLE_SUCCESS;
}

// wait = false queues the work and returns: the caller's next evaluation on
// the same streams waits for it (the tree layer then synchronises once, with
// the edge results).  The device copies of idx and t are stream-ordered, so a
// later call may reuse them at once.
int engine_update_pmats(Instance *I, const int *idx, const double *t, int n, int derivs, bool wait) {
  std::lock_guard<std::mutex> lk_(I->mu);
  int rc = enqueue_pmats(I, idx, t, n, derivs);
  if (!rc && I->heavy) rc = enqueue_pmats(I->heavy.get(), idx, t, n, derivs);
  return wait ? finish_both(I, rc) : rc;
}

template <typename T, int R>
static void launch_nuc(Instance *I, const DevClaOp *d_ops, int n, bool fused) {
  GLE_LAUNCH(I, (k_cla_nuc<T, R>), dim3(I->Ppad / 256, fused ? 1 : n), dim3(256), 0,
                     reinterpret_cast<T *>(I->d_cla), I->d_u, I->d_tips, I->d_pm64, d_ops, I->d_flags, I->Ppad,
                     I->P, fused ? n : 0);
}

template <typename T>
static int launch_nuc_r(Instance *I, const DevClaOp *d_ops, int n, bool fused) {
  switch (I->R) {
    case 1: launch_nuc<T, 1>(I, d_ops, n, fused); break;
    case 2: launch_nuc<T, 2>(I, d_ops, n, fused); break;
    case 3: launch_nuc<T, 3>(I, d_ops, n, fused); break;
    case 4: launch_nuc<T, 4>(I, d_ops, n, fused); break;
    case 5: launch_nuc<T, 5>(I, d_ops, n, fused); break;
    case 6: launch_nuc<T, 6>(I, d_ops, n, fused); break;
    case 7: launch_nuc<T, 7>(I, d_ops, n, fused); break;
    case 8: launch_nuc<T, 8>(I, d_ops, n, fused); break;
    default: return set_error(GLE_ERR_UNSUPPORTED, "rate categories must be 1..8");
  }
  return GLE_SUCCESS;
}

template <typename T, int R>
static void launch_codon(Instance *I, const DevClaOp *d_ops, int n, bool fused) {
  GLE_LAUNCH(I, (k_cla_codon<T, R>), dim3(I->Ppad / 16, fused ? 1 : n), dim3(64 * R), 0,
             reinterpret_cast<T *>(I->d_cla), I->d_u, I->d_tips, I->d_pmpad, d_ops, I->d_flags, I->Ppad, I->P,
             fused ? n : 0);
}

template <typename T>
static int launch_codon_r(Instance *I, const DevClaOp *d_ops, int n, bool fused) {
  switch (I->R) {
    case 1: launch_codon<T, 1>(I, d_ops, n, fused); break;
    case 2: launch_codon<T, 2>(I, d_ops, n, fused); break;
    case 3: launch_codon<T, 3>(I, d_ops, n, fused); break;
    case 4: launch_codon<T, 4>(I, d_ops, n, fused); break;
    case 5: launch_codon<T, 5>(I, d_ops, n, fused); break;
    case 6: launch_codon<T, 6>(I, d_ops, n, fused); break;
    case 7: launch_codon<T, 7>(I, d_ops, n, fused); break;
    case 8: launch_codon<T, 8>(I, d_ops, n, fused); break;
    default: return set_error(GLE_ERR_UNSUPPORTED, "rate categories must be 1..8");
  }
  return GLE_SUCCESS;
}

template <typename T, int R>
static void launch_aa(Instance *I, const DevClaOp *d_ops, int n, bool fused) {
  GLE_LAUNCH(I, (k_cla_aa<T, R>), dim3(I->Ppad / 64, fused ? 1 : n), dim3(64 * R), 0,
                     reinterpret_cast<T *>(I->d_cla), I->d_u, I->d_tips, I->d_pmpad, d_ops, I->d_flags, I->Ppad,
                     I->P, fused ? n : 0);
}

template <typename T>
static int launch_aa_r(Instance *I, const DevClaOp *d_ops, int n, bool fused) {
  switch (I->R) {
    case 1: launch_aa<T, 1>(I, d_ops, n, fused); break;
    case 2: launch_aa<T, 2>(I, d_ops, n, fused); break;
    case 3: launch_aa<T, 3>(I, d_ops, n, fused); break;
    case 4: launch_aa<T, 4>(I, d_ops, n, fused); break;
    case 5: launch_aa<T, 5>(I, d_ops, n, fused); break;
    case 6: launch_aa<T, 6>(I, d_ops, n, fused); break;
    case 7: launch_aa<T, 7>(I, d_ops, n, fused); break;
    case 8: launch_aa<T, 8>(I, d_ops, n, fused); break;
    default: return set_error(GLE_ERR_UNSUPPORTED, "rate categories must be 1..8");
  }
  return GLE_SUCCESS;
}

template <typename T>
static int launch_cla(Instance *I, const DevClaOp *d_ops, int n, bool fused) {
  switch (I->S) {
    case 4: return launch_nuc_r<T>(I, d_ops, n, fused);
    case 20: return launch_aa_r<T>(I, d_ops, n, fused);
    case 61: return launch_codon_r<T>(I, d_ops, n, fused);
    default: return set_error(GLE_ERR_UNSUPPORTED, "state count %d has no kernel", I->S);
  }
}

static int enqueue_partials(Instance *I, const DevClaOp *ops, int n_ops, const int *level_start, int n_levels,
                            bool fused, bool allow_graph = true) {
  if (n_ops <= 0) return GLE_SUCCESS;
  fused = fused && engine_supports_fused(I);
  std::vector<DevClaOp> lv_ops;
  std::vector<int> lv_start;
  if (!fused && n_levels <= 0 && n_ops > 1 && level_ops(ops, n_ops, lv_ops, lv_start)) {
    ops = lv_ops.data();
    level_start = lv_start.data();
    n_levels = int(lv_start.size()) - 1;
  }
  for (int i = 0; i < n_ops; i++) {
    const DevClaOp &o = ops[i];
    auto bad_child = [&](int c) { return c >= I->n_slots || (c < 0 && -c - 1 >= I->N); };
    if (o.dest < 0 || o.dest >= I->n_slots || bad_child(o.c1) || bad_child(o.c2) || o.pm1 < 0 ||
        o.pm1 >= I->n_pm || o.pm2 < 0 || o.pm2 >= I->n_pm || o.dest == o.c1 || o.dest == o.c2)
      return set_error(GLE_ERR_ARG, "cla op %d out of range (dest %d c1 %d c2 %d pm %d %d)", i, o.dest, o.c1,
                       o.c2, o.pm1, o.pm2);
  }
  const int ops_cap0 = I->ops_cap;
  int rc = grow(I->d_ops, I->ops_cap, n_ops);
  if (rc) return rc;
  if (I->ops_cap != ops_cap0) I->ops_held_buf = nullptr;  // a new buffer holds nothing yet
  if (!(I->ops_held_buf == I->d_ops && I->ops_held.size() == size_t(n_ops) &&
        std::memcmp(I->ops_held.data(), ops, sizeof(DevClaOp) * n_ops) == 0)) {
    GLE_HIP(hipMemcpyAsync(I->d_ops, ops, sizeof(DevClaOp) * n_ops, hipMemcpyHostToDevice, I->stream));
    I->ops_held.assign(ops, ops + n_ops);
    I->ops_held_buf = I->d_ops;
  }
  auto bytes_of = [&](int b, int e) {
    double t = 0.0;
    for (int i = b; i < e; i++) t += cla_op_bytes(I, ops[i]);
    return t;
  };
  // An op list met twice in a row (same ops and levels, same device buffers,
  // same instance generation) is captured into a hipGraph on its second use
  // and replayed after that: one launch instead of one per level.  The graph
  // holds kernel nodes only; the op list is copied in above, outside it.
  std::vector<int> levels;
  if (!fused && n_levels > 0) levels.assign(level_start, level_start + n_levels + 1);
  const bool graphable = allow_graph && I->graphs && !I->timing && !fused && n_levels > 0;
  Instance::GraphEntry *entry = nullptr;
  if (graphable) {
    for (auto &g : I->g_cache)
      if (g.dev == I->d_ops && g.gen == I->generation && g.levels == levels && g.ops.size() == size_t(n_ops) &&
          std::memcmp(g.ops.data(), ops, sizeof(DevClaOp) * n_ops) == 0) {
        entry = &g;
        break;
      }
    if (entry) {
      entry->last = ++I->g_clock;
      if (entry->exec) {
        GLE_HIP(hipGraphLaunch(entry->exec, I->stream));
        I->g_replays++;
        return GLE_SUCCESS;
      }
      // second use: capture below
    } else {
      if (int(I->g_cache.size()) >= Instance::kGraphEntries) {
        auto lru = std::min_element(I->g_cache.begin(), I->g_cache.end(),
                                    [](const auto &a, const auto &b) { return a.last < b.last; });
        if (lru->exec) (void)hipGraphExecDestroy(lru->exec);
        I->g_cache.erase(lru);
      }
      Instance::GraphEntry g;
      g.ops.assign(ops, ops + n_ops);
      g.levels = levels;
      g.dev = I->d_ops;
      g.gen = I->generation;
      g.last = ++I->g_clock;
      I->g_cache.push_back(std::move(g));
    }
  }
  auto launch_all = [&]() -> int {
    if (fused) {
      timed_begin(I, 0, bytes_of(0, n_ops));
      const int r = I->prec == GLE_PREC_FP64 ? launch_cla<double>(I, I->d_ops, n_ops, true)
                                             : launch_cla<float>(I, I->d_ops, n_ops, true);
      timed_end(I);
      return r;
    }
    const int nl = n_levels > 0 ? n_levels : n_ops;
    for (int l = 0; l < nl; l++) {
      const int b = n_levels > 0 ? level_start[l] : l, e = n_levels > 0 ? level_start[l + 1] : l + 1;
      if (e <= b) continue;
      timed_begin(I, 0, bytes_of(b, e));
      const int r = I->prec == GLE_PREC_FP64 ? launch_cla<double>(I, I->d_ops + b, e - b, false)
                                             : launch_cla<float>(I, I->d_ops + b, e - b, false);
      timed_end(I);
      if (r) return r;
    }
    return GLE_SUCCESS;
  };
  bool captured = false;
  if (entry && hipStreamBeginCapture(I->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
    rc = launch_all();
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(I->stream, &graph);
    hipError_t ei = hipErrorUnknown;
    if (ec == hipSuccess && !rc) ei = hipGraphInstantiate(&entry->exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (rc) return rc;  // a launch argument was refused before anything ran
    captured = ec == hipSuccess && ei == hipSuccess;
    if (captured) GLE_HIP(hipGraphLaunch(entry->exec, I->stream));
  }
  if (entry && !captured) {
    // capture is unavailable here: nothing ran, so launch eagerly, and stay
    // eager from now on
    (void)hipGetLastError();
    entry->exec = nullptr;
    I->drop_graphs();
    I->graphs = false;
  }
  if (!captured) rc = launch_all();
  if (rc) return rc;
  GLE_HIP(hipGetLastError());
  return GLE_SUCCESS;
}

// wait = false: queued only, as engine_update_pmats
int engine_update_partials(Instance *I, const DevClaOp *ops, int n_ops, const int *level_start, int n_levels,
                           bool fused, bool wait) {
  std::lock_guard<std::mutex> lk_(I->mu);
  // the bulk first: its stream starts working while the small part is issued
  int rc = enqueue_partials(I, ops, n_ops, level_start, n_levels, fused);
  if (!rc && I->heavy) rc = enqueue_partials(I->heavy.get(), ops, n_ops, level_start, n_levels, fused);
  return wait ? finish_both(I, rc) : rc;
}

// derivs: some op of the launch wants (d1, d2); the nucleotide kernel then
// computes them for every op (the others ignore them), the codon kernel takes
// its derivative-free form only when none does.
template <typename T>
static int launch_edges(Instance *I, const DevEdgeOp *d_ops, int n, int nblk, bool derivs, double *d_part) {
  EdgeArgs<T> a;
  a.cla = reinterpret_cast<const T *>(I->d_cla);
  a.u = I->d_u;
  a.tips = I->d_tips;
  a.pm64 = I->d_pm64;
  a.model = I->d_model;
  a.counts = I->d_counts;
  a.cst = I->d_cst;
  a.pinv = I->pinv;
  a.pinv_in_model = I->pinv_in_model;
  a.slot_stride = I->slot_elems();
  a.Ppad = I->Ppad;
  a.P = I->P;
  a.R = I->R;
  a.S = I->S;
  a.Sp = I->Sp;
  a.partials = d_part;
  a.site = I->d_site;
  dim3 grid(nblk, n);
  if (I->S == 4) {
    // the FP32 form's spectral derivative route needs one eigensystem for all categories
    const bool nuc_spec = I->prec == GLE_PREC_FP32 && I->n_eig == 1 && I->d_pdiag;
#define GLE_EDGE_NUC(RR)                                                                                          \
  if (derivs && nuc_spec)                                                                                         \
    GLE_LAUNCH(I, (k_edge_nuc<T, RR, true, true>), grid, dim3(256), 0, a.cla, a.u, a.tips, a.pm64,               \
                       a.model, a.counts, a.cst, d_ops, a.partials, a.site, a.pinv, a.pinv_in_model, a.Ppad, a.P,  \
                       I->d_V, I->d_Vinv, I->d_pdiag);                                                            \
  else if (derivs)                                                                                                \
    GLE_LAUNCH(I, (k_edge_nuc<T, RR, true>), grid, dim3(256), 0, a.cla, a.u, a.tips, a.pm64,                     \
                       a.model, a.counts, a.cst, d_ops, a.partials, a.site, a.pinv, a.pinv_in_model, a.Ppad, a.P,  \
                       nullptr, nullptr, nullptr);                                                                \
  else                                                                                                            \
    GLE_LAUNCH(I, (k_edge_nuc<T, RR, false>), grid, dim3(256), 0, a.cla, a.u, a.tips, a.pm64,                    \
                       a.model, a.counts, a.cst, d_ops, a.partials, a.site, a.pinv, a.pinv_in_model, a.Ppad, a.P,  \
                       nullptr, nullptr, nullptr)
    switch (I->R) {
      case 1: GLE_EDGE_NUC(1); break;
      case 2: GLE_EDGE_NUC(2); break;
      case 3: GLE_EDGE_NUC(3); break;
      case 4: GLE_EDGE_NUC(4); break;
      case 5: GLE_EDGE_NUC(5); break;
      case 6: GLE_EDGE_NUC(6); break;
      case 7: GLE_EDGE_NUC(7); break;
      case 8: GLE_EDGE_NUC(8); break;
      default: return set_error(GLE_ERR_UNSUPPORTED, "rate categories must be 1..8");
    }
#undef GLE_EDGE_NUC
  } else if (I->S == 20) {
#define GLE_EDGE_AA(RR)                                                                                           \
  GLE_LAUNCH(I, (k_edge_aa<T, RR>), grid, dim3(64 * RR), 0, a.cla, a.u, a.tips, I->d_pmpad, I->d_pdiag,         \
                     I->d_V, I->d_Vinv, I->d_eig_of_rate, a.model, a.counts, a.cst, d_ops, a.partials, a.site, a.pinv, \
                     a.pinv_in_model, a.Ppad, a.P)
    switch (I->R) {
      case 1: GLE_EDGE_AA(1); break;
      case 2: GLE_EDGE_AA(2); break;
      case 3: GLE_EDGE_AA(3); break;
      case 4: GLE_EDGE_AA(4); break;
      case 5: GLE_EDGE_AA(5); break;
      case 6: GLE_EDGE_AA(6); break;
      case 7: GLE_EDGE_AA(7); break;
      case 8: GLE_EDGE_AA(8); break;
      default: return set_error(GLE_ERR_UNSUPPORTED, "rate categories must be 1..8");
    }
#undef GLE_EDGE_AA
  } else if (I->S == 61) {
#define GLE_EDGE_CODON(RR)                                                                                        \
  if (derivs)                                                                                                     \
    GLE_LAUNCH(I, (k_edge_codon<T, RR, true>), grid, dim3(64 * RR), 0, a.cla, a.u, a.tips, I->d_pmpad,            \
               I->d_pdiag, I->d_eigfrag, I->d_eig_of_rate, a.model, a.counts, a.cst, d_ops, a.partials, a.site,    \
               a.pinv, a.pinv_in_model, a.Ppad, a.P);                                                             \
  else                                                                                                            \
    GLE_LAUNCH(I, (k_edge_codon<T, RR, false>), grid, dim3(64 * RR), 0, a.cla, a.u, a.tips, I->d_pmpad,           \
               I->d_pdiag, I->d_eigfrag, I->d_eig_of_rate, a.model, a.counts, a.cst, d_ops, a.partials, a.site,    \
               a.pinv, a.pinv_in_model, a.Ppad, a.P)
    switch (I->R) {
      case 1: GLE_EDGE_CODON(1); break;
      case 2: GLE_EDGE_CODON(2); break;
      case 3: GLE_EDGE_CODON(3); break;
      case 4: GLE_EDGE_CODON(4); break;
      case 5: GLE_EDGE_CODON(5); break;
      case 6: GLE_EDGE_CODON(6); break;
      case 7: GLE_EDGE_CODON(7); break;
      case 8: GLE_EDGE_CODON(8); break;
      default: return set_error(GLE_ERR_UNSUPPORTED, "rate categories must be 1..8");
    }
#undef GLE_EDGE_CODON
  } else
    return set_error(GLE_ERR_UNSUPPORTED, "state count %d has no kernel", I->S);
  return GLE_SUCCESS;
}

static int edge_blocks(const Instance *I) {
  return I->S == 4 ? I->Ppad / 256 : I->S == 61 ? I->Ppad / 16 : I->Ppad / 64;
}

// Device buffers of one edge launch: the uploaded op list, the per-block
// partial sums [op][block][3] and the per-op results [op][3].
struct EdgeBufs {
  DevEdgeOp *ops;
  double *part, *res;
};

// Edge kernels and the per-op reduction into the results; d_site when `site`.
// With ext the launch uses those buffers as they are (nothing is allocated,
// the op list is always uploaded and the launch is not timed); otherwise the
// instance's own buffers, grown as needed.
static int enqueue_edges(Instance *I, const DevEdgeOp *ops, int n, bool site, const EdgeBufs *ext = nullptr) {
  for (int i = 0; i < n; i++) {
    const DevEdgeOp &o = ops[i];
    if (o.partial < 0 || o.partial >= I->n_slots || o.child >= I->n_slots || (o.child < 0 && -o.child - 1 >= I->N) ||
        o.pm < 0 || o.pm >= I->n_pm || o.child == o.partial)
      return set_error(GLE_ERR_ARG, "edge op %d out of range", i);
  }
  const int nblk = edge_blocks(I);
  bool derivs = false;
  for (int i = 0; i < n; i++) derivs = derivs || ops[i].derivs;
  if (ext) {
    GLE_HIP(hipMemcpyAsync(ext->ops, ops, sizeof(DevEdgeOp) * n, hipMemcpyHostToDevice, I->stream));
    double *site_save = I->d_site;
    if (!site) I->d_site = nullptr;
    const int rc = I->prec == GLE_PREC_FP64 ? launch_edges<double>(I, ext->ops, n, nblk, derivs, ext->part)
                                            : launch_edges<float>(I, ext->ops, n, nblk, derivs, ext->part);
    I->d_site = site_save;
    if (rc) return rc;
    hipLaunchKernelGGL(k_edge_reduce, dim3(n), dim3(64), 0, I->stream, ext->part, nblk, ext->res);
    GLE_HIP(hipGetLastError());
    return GLE_SUCCESS;
  }
  const int eops_cap0 = I->eops_cap;
  int rc = grow(I->d_eops, I->eops_cap, n);
  if (rc) return rc;
  if (I->eops_cap != eops_cap0) I->eops_held_buf = nullptr;  // a new buffer holds nothing yet
  rc = grow(I->d_eres, I->eres_cap, 3 * n);
  if (rc) return rc;
  const size_t need = size_t(n) * nblk * 3;
  if (need > I->part_cap) {
    if (I->d_part) GLE_HIP(hipFree(I->d_part));
    GLE_HIP(hipMalloc(&I->d_part, sizeof(double) * need));
    I->part_cap = need;
  }
  if (!(I->eops_held_buf == I->d_eops && I->eops_held.size() == size_t(n) &&
        std::memcmp(I->eops_held.data(), ops, sizeof(DevEdgeOp) * n) == 0)) {
    GLE_HIP(hipMemcpyAsync(I->d_eops, ops, sizeof(DevEdgeOp) * n, hipMemcpyHostToDevice, I->stream));
    I->eops_held.assign(ops, ops + n);
    I->eops_held_buf = I->d_eops;
  }
  double *site_save = I->d_site;
  if (!site) I->d_site = nullptr;
  double eb = 0.0;
  for (int i = 0; i < n; i++) eb += edge_op_bytes(I, ops[i]);
  timed_begin(I, 1, eb);
  rc = I->prec == GLE_PREC_FP64 ? launch_edges<double>(I, I->d_eops, n, nblk, derivs, I->d_part)
                                : launch_edges<float>(I, I->d_eops, n, nblk, derivs, I->d_part);
  timed_end(I);
  I->d_site = site_save;
  if (rc) return rc;
  hipLaunchKernelGGL(k_edge_reduce, dim3(n), dim3(64), 0, I->stream, I->d_part, nblk, I->d_eres);
  GLE_HIP(hipGetLastError());
  return GLE_SUCCESS;
}

// site: host buffer of nsec x P values, the first nsec of (lnL, d1, d2,
// partial and child rescale counts)
static int copy_edges_out(Instance *I, int n, double *out, double *site, int nsec, const double *d_res = nullptr) {
  GLE_HIP(hipMemcpyAsync(out, d_res ? d_res : I->d_eres, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, I->stream));
  if (site)
    for (int k = 0; k < nsec; k++)
      GLE_HIP(hipMemcpyAsync(site + size_t(k) * I->P, I->d_site + size_t(k) * I->Ppad, sizeof(double) * I->P,
                             hipMemcpyDeviceToHost, I->stream));
  return GLE_SUCCESS;
}

// Edge results of an instance and its heavy-pattern part; with site (nsec x
// patterns, the caller's pattern order) also the per-pattern values of the
// ops' first edge: its lnL (nsec = 1) or all five sections (nsec = 5).
// Pinned room for n results (3 doubles each) of an instance, grown on demand
// (no copy of the instance is in flight: every evaluation waits before returning).
static int stage_results(Instance *X, int n) {
  const size_t need = 3 * size_t(n);
  if (need <= X->h_stage_cap) return GLE_SUCCESS;
  if (X->h_stage) GLE_HIP(hipHostFree(X->h_stage));
  X->h_stage = nullptr;
  X->h_stage_cap = 0;
  const size_t cap = std::max(need, size_t(1024));
  GLE_HIP(hipHostMalloc(reinterpret_cast<void **>(&X->h_stage), sizeof(double) * cap));
  X->h_stage_cap = cap;
  return GLE_SUCCESS;
}

static int edges_impl(Instance *I, const DevEdgeOp *ops, int n, double *out, double *site, int nsec,
                      int *flags) {
  std::lock_guard<std::mutex> lk_(I->mu);
  if (n <= 0) return GLE_SUCCESS;
  Instance *H = I->heavy.get();
  const bool scatter = site && H;  // the two parts' patterns go back to the caller's order
  std::vector<double> site_l(scatter ? nsec * size_t(std::max(I->P, 1)) : 0),
      site_h(scatter ? nsec * size_t(std::max(H->P, 1)) : 0);
  // kernels of both parts first; the results come back through pinned memory,
  // so the copies queue behind them without a wait each (a copy into pageable
  // memory had cost a host round trip per part)
  int rc = enqueue_edges(I, ops, n, site != nullptr);
  if (!rc && H) rc = enqueue_edges(H, ops, n, site != nullptr);
  if (!rc) rc = stage_results(I, n);
  if (!rc && H) rc = stage_results(H, n);
  if (!rc) rc = copy_edges_out(I, n, I->h_stage, site ? (scatter ? site_l.data() : site) : nullptr, nsec);
  if (!rc && H) rc = copy_edges_out(H, n, H->h_stage, site ? site_h.data() : nullptr, nsec);
  // the rescale flags of both parts come back with the results (one wait)
  if (!rc && flags) rc = hipMemcpyAsync(I->h_flag, I->d_flags, sizeof(int), hipMemcpyDeviceToHost, I->stream) ==
                                 hipSuccess ? GLE_SUCCESS : set_error(GLE_ERR_DEVICE, "flag copy failed");
  if (!rc && flags && H)
    rc = hipMemcpyAsync(H->h_flag, H->d_flags, sizeof(int), hipMemcpyDeviceToHost, H->stream) == hipSuccess
             ? GLE_SUCCESS
             : set_error(GLE_ERR_DEVICE, "flag copy failed");
  rc = finish_both(I, rc);
  if (rc) return rc;
  if (flags) *flags = *I->h_flag | (H ? *H->h_flag : 0);
  for (int i = 0; i < 3 * n; i++) out[i] = I->h_stage[i] + (H ? H->h_stage[i] : 0.0);
  if (scatter) {
    const size_t Pt = I->light_idx.size() + I->heavy_idx.size();
    for (int k = 0; k < nsec; k++) {
      for (size_t j = 0; j < I->light_idx.size(); j++) site[k * Pt + I->light_idx[j]] = site_l[k * size_t(I->P) + j];
      for (size_t j = 0; j < I->heavy_idx.size(); j++) site[k * Pt + I->heavy_idx[j]] = site_h[k * size_t(H->P) + j];
    }
  }
  return GLE_SUCCESS;
}

// ---------------------------------------------------------------- serving ticks
// The edge-partials capacity a tick grows from: what the instance holds, or 64 edges' worth.
static size_t part_cap_floor(const Instance *X) {
  return std::max(X->part_cap, size_t(64) * edge_blocks(X) * 3);
}

// The CLA levels and the edges of one part of a tick, on X's stream, reading
// the staged op lists; the rescale flags go to d_flags.  The edge partials go
// to the part's own buffer, or (heavy) to the one kept for its serving ticks,
// returned with their count per edge for the reduction.
static int tick_part_kernels(Instance *X, bool heavy, const DevClaOp *ops, const DevClaOp *d_ops,
                             const int *level_start, int n_levels, const DevEdgeOp *eops, const DevEdgeOp *d_eops,
                             int n_e, int *d_flags, const double *&parts_out, int &nblk_out) {
  int *flags_save = X->d_flags;
  X->d_flags = d_flags;
  int rc = GLE_SUCCESS;
  for (int l = 0; l < n_levels && !rc; l++) {
    const int b = level_start[l], e = level_start[l + 1];
    if (e <= b) continue;
    double bytes = 0.0;
    for (int i = b; i < e; i++) bytes += cla_op_bytes(X, ops[i]);
    timed_begin(X, 0, bytes);
    rc = X->prec == GLE_PREC_FP64 ? launch_cla<double>(X, d_ops + b, e - b, false)
                                  : launch_cla<float>(X, d_ops + b, e - b, false);
    timed_end(X);
  }
  X->d_flags = flags_save;
  if (rc || n_e == 0) return rc;
  const int nblk = edge_blocks(X);
  const size_t part = size_t(n_e) * nblk * 3;
  double *&d_part = heavy ? X->d_tpart : X->d_part;
  size_t &part_cap = heavy ? X->tpart_cap : X->part_cap;
  if (part > part_cap) {
    // a tick grows the buffer without hipFree, which would wait for the
    // whole device (the tick in flight included) before this tick's edges
    // could be queued; the old buffer is freed with the instance
    if (d_part) X->retired.push_back(d_part);
    d_part = nullptr;
    part_cap = 0;
    const size_t cap = std::max(part, 2 * part_cap_floor(X));
    GLE_HIP(hipMalloc(&d_part, sizeof(double) * cap));
    part_cap = cap;
  }
  bool derivs = false;
  double eb = 0.0;
  for (int i = 0; i < n_e; i++) {
    derivs = derivs || eops[i].derivs;
    eb += edge_op_bytes(X, eops[i]);
  }
  double *site_save = X->d_site;
  X->d_site = nullptr;
  timed_begin(X, 1, eb);
  rc = X->prec == GLE_PREC_FP64 ? launch_edges<double>(X, d_eops, n_e, nblk, derivs, d_part)
                                : launch_edges<float>(X, d_eops, n_e, nblk, derivs, d_part);
  timed_end(X);
  X->d_site = site_save;
  parts_out = d_part;
  nblk_out = nblk;
  return rc;
}

// One tick of the serving layer on an instance and its heavy part (H, or
// null), enqueued without any wait on the instance's stream: P, P', P'' for
// the new lengths, the CLA levels, the edges and their reduction of the
// instance, then those of the heavy part, which read the same staged lists
// and the instance's FP64 P slots (the same values: one P(t) kernel writes
// both parts' layouts), and one copy of both parts' results and rescale flags
// into pinned memory, and an event.  The lists are staged in pinned memory
// first, so every copy is asynchronous.  Queuing the heavy part on its own
// stream took a second upload, P(t) launch, download and event per tick
// (about 2.4 us of host time each) for its few patterns.
static int tick_enqueue_part(Instance *X, Instance *H, int sidx, const int *pm_idx, const double *pm_len, int n_pm,
                             int pm_derivs, const DevClaOp *ops, int n_ops, const int *level_start, int n_levels,
                             const DevEdgeOp *eops, int n_e) {
  Instance::TickStage &st = X->stage[sidx];
  if (st.busy) return set_error(GLE_ERR_STATE, "tick stage %d still in flight", sidx);
  for (int i = 0; i < n_pm; i++)
    if (pm_idx[i] < 0 || pm_idx[i] >= X->n_pm || !(pm_len[i] >= 0.0))
      return set_error(GLE_ERR_ARG, "tick pmat %d out of range", i);
  for (int i = 0; i < n_ops; i++) {
    const DevClaOp &o = ops[i];
    auto bad = [&](int c) { return c >= X->n_slots || (c < 0 && -c - 1 >= X->N); };
    // a message child names a slot and needs no P(t); kept messages go to other slots
    const bool g1 = o.flags & kOpMsg1, g2 = o.flags & kOpMsg2;
    auto bad_m = [&](int m, int c) { return m != -1 && (m < 0 || m >= X->n_slots || m == o.dest || c < 0); };
    if (o.dest < 0 || o.dest >= X->n_slots || bad(o.c1) || bad(o.c2) ||
        (!g1 && (o.pm1 < 0 || o.pm1 >= X->n_pm)) || (!g2 && (o.pm2 < 0 || o.pm2 >= X->n_pm)) ||
        (g1 && o.c1 < 0) || (g2 && o.c2 < 0) || o.dest == o.c1 || o.dest == o.c2 || bad_m(o.m1, o.c1) ||
        bad_m(o.m2, o.c2) || (o.m1 >= 0 && (g1 || o.m1 == o.c2 || o.m1 == o.m2)) || (o.m2 >= 0 && (g2 || o.m2 == o.c1)) ||
        ((g1 || g2 || o.m1 >= 0 || o.m2 >= 0) && X->S != 61))
      return set_error(GLE_ERR_ARG, "tick cla op %d out of range", i);
  }
  bool any_derivs = false, any_msg = false;
  for (int i = 0; i < n_e; i++) {
    const DevEdgeOp &o = eops[i];
    if (o.partial < 0 || o.partial >= X->n_slots || o.child >= X->n_slots || (o.child < 0 && -o.child - 1 >= X->N) ||
        o.pm < 0 || o.pm >= X->n_pm || o.child == o.partial || (o.msg && (o.child < 0 || X->S != 61)))
      return set_error(GLE_ERR_ARG, "tick edge op %d out of range", i);
    any_derivs = any_derivs || o.derivs;
    any_msg = any_msg || o.msg;
  }
  // a message child is read only by the derivative-free edge kernel
  if (any_msg && any_derivs) return set_error(GLE_ERR_ARG, "tick edges read messages in a launch with derivatives");
  // one staging layout in pinned host memory and on the device:
  // [P(t) slots | lengths | CLA ops | edge ops | flags (2 ints) | results |
  // heavy part's results]; the upload carries everything up to the (zeroed)
  // flags, the download the flags and the results
  auto a8 = [](size_t b) { return (b + 7) & ~size_t(7); };
  const int parts = H ? 2 : 1;
  const size_t o_len = a8(sizeof(int) * n_pm), o_ops = o_len + sizeof(double) * n_pm,
               o_eops = o_ops + a8(sizeof(DevClaOp) * n_ops), o_flags = o_eops + a8(sizeof(DevEdgeOp) * n_e),
               o_res = o_flags + 8, need = o_res + sizeof(double) * 3 * n_e * parts;
  if (need > st.cap) {
    const size_t cap = std::max(need + need / 2, size_t(64) << 10);
    // retired rather than freed: either free would wait for the other
    // stage's tick still in flight
    if (st.h) X->retired_host.push_back(st.h);
    if (st.d) X->retired.push_back(st.d);
    st.h = st.d = nullptr;
    st.cap = 0;
    GLE_HIP(hipHostMalloc(reinterpret_cast<void **>(&st.h), cap));
    GLE_HIP(hipMalloc(reinterpret_cast<void **>(&st.d), cap));
    st.cap = cap;
  }
  if (!st.done) GLE_HIP(hipEventCreateWithFlags(&st.done, hipEventDisableTiming));
  std::copy(pm_idx, pm_idx + n_pm, reinterpret_cast<int *>(st.h));
  std::copy(pm_len, pm_len + n_pm, reinterpret_cast<double *>(st.h + o_len));
  std::copy(ops, ops + n_ops, reinterpret_cast<DevClaOp *>(st.h + o_ops));
  std::copy(eops, eops + n_e, reinterpret_cast<DevEdgeOp *>(st.h + o_eops));
  std::memset(st.h + o_flags, 0, 8);
  st.res = reinterpret_cast<double *>(st.h + o_res);
  st.flags = reinterpret_cast<int *>(st.h + o_flags);
  GLE_HIP(hipMemcpyAsync(st.d, st.h, o_res, hipMemcpyHostToDevice, X->stream));
  const DevClaOp *d_ops = reinterpret_cast<const DevClaOp *>(st.d + o_ops);
  const DevEdgeOp *d_eops = reinterpret_cast<const DevEdgeOp *>(st.d + o_eops);
  int *d_flags = reinterpret_cast<int *>(st.d + o_flags);
  if (n_pm > 0) {
    const int S = X->S;
    const size_t shmem = sizeof(double) * (2 * S * S + 3 * S);
    timed_begin(X, 2, double(n_pm) * X->R * S * S * 8.0 * (pm_derivs ? 3 : 1));
    GLE_LAUNCH(X, k_pmat, dim3(n_pm, X->R, (pmat_tiles(S) + 255) / 256), dim3(256), shmem, S, X->Sp, X->R,
               reinterpret_cast<const int *>(st.d), reinterpret_cast<const double *>(st.d + o_len),
               X->d_eig_of_rate, X->d_model + S + X->R, X->d_V, X->d_Vinv, X->d_lam, X->d_pm64, X->d_pmpad,
               X->d_pdiag, pm_derivs, S == 20 ? kPadTransposed : S == 61 ? kPadCodonFrag : kPadRowMajor);
    timed_end(X);
  }
  // the CLA kernels raise their rescale flags in this tick's own words
  double *d_res = reinterpret_cast<double *>(st.d + o_res);
  const double *pa = nullptr, *pb = nullptr;
  int na = 0, nb = 0;
  int rc = tick_part_kernels(X, false, ops, d_ops, level_start, n_levels, eops, d_eops, n_e, d_flags, pa, na);
  if (rc) return rc;
  if (H) {
    // the heavy part's kernels on this stream, reading this part's P slots
    hipStream_t hs = H->stream;
    double *hpm64 = H->d_pm64, *hpmpad = H->d_pmpad;
    H->stream = X->stream;
    H->d_pm64 = X->d_pm64;
    H->d_pmpad = X->d_pmpad;
    rc = tick_part_kernels(H, true, ops, d_ops, level_start, n_levels, eops, d_eops, n_e, d_flags + 1, pb, nb);
    H->stream = hs;
    H->d_pm64 = hpm64;
    H->d_pmpad = hpmpad;
    if (rc) return rc;
  }
  // both parts' edges reduced by one launch, the heavy part's results after
  if (n_e > 0) {
    if (H)
      hipLaunchKernelGGL(k_edge_reduce2, dim3(2 * n_e), dim3(64), 0, X->stream, pa, na, pb, nb, n_e, d_res);
    else
      hipLaunchKernelGGL(k_edge_reduce, dim3(n_e), dim3(64), 0, X->stream, pa, na, d_res);
  }
  GLE_HIP(hipGetLastError());
  GLE_HIP(hipMemcpyAsync(st.h + o_flags, st.d + o_flags, 8 + sizeof(double) * 3 * n_e * parts, hipMemcpyDeviceToHost,
                         X->stream));
  GLE_HIP(hipEventRecord(st.done, X->stream));
  st.n_edges = n_e;
  st.with_heavy = H != nullptr;
  st.busy = true;
  return GLE_SUCCESS;
}

int engine_tick_enqueue(Instance *I, int sidx, const int *pm_idx, const double *pm_len, int n_pm, int pm_derivs,
                        const DevClaOp *ops, int n_ops, const int *level_start, int n_levels,
                        const DevEdgeOp *eops, int n_e) {
  std::lock_guard<std::mutex> lk_(I->mu);
  if (sidx < 0 || sidx >= kTickStages) return set_error(GLE_ERR_ARG, "tick stage %d", sidx);
  const int rc = tick_enqueue_part(I, I->heavy.get(), sidx, pm_idx, pm_len, n_pm, pm_derivs, ops, n_ops, level_start,
                                   n_levels, eops, n_e);
  if (rc) {  // drain what was queued so the stage can be reused
    (void)hipStreamSynchronize(I->stream);
    I->stage[sidx].busy = false;
  }
  return rc;
}

// Waits for the tick in stage sidx and returns its n x (lnL, d1, d2) (both
// parts summed) and the OR of the parts' rescale flags.
int engine_tick_collect(Instance *I, int sidx, double *res, int *flags) {
  std::lock_guard<std::mutex> lk_(I->mu);
  int rc = GLE_SUCCESS, f = 0;
  Instance::TickStage &st = I->stage[sidx];
  if (st.busy) {
    const hipError_t e = hipEventSynchronize(st.done);
    st.busy = false;
    if (e != hipSuccess) {
      rc = set_error(GLE_ERR_DEVICE, "serving tick failed: %s", hipGetErrorString(e));
    } else {
      timed_collect(I);
      const int n3 = 3 * st.n_edges;
      for (int i = 0; i < n3; i++) res[i] = st.res[i];
      f = st.flags[0];
      if (st.with_heavy) {
        if (I->heavy) timed_collect(I->heavy.get());
        for (int i = 0; i < n3; i++) res[i] += st.res[n3 + i];
        f |= st.flags[1];
      }
    }
  }
  if (flags) *flags = f;
  if (!rc) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = set_error(GLE_ERR_DEVICE, "serving tick: %s", hipGetErrorString(e));
  }
  return rc;
}

bool engine_tick_ready(Instance *I, int sidx) {
  std::lock_guard<std::mutex> lk_(I->mu);
  if (sidx < 0 || sidx >= kTickStages) return true;
  const Instance::TickStage &st = I->stage[sidx];
  // (both parts' work is on the instance's stream; an error is reported by
  // the collect, which then does not wait either)
  return !(st.busy && hipEventQuery(st.done) == hipErrorNotReady);
}

int engine_edges(Instance *I, const DevEdgeOp *ops, int n, double *out, double *site_lnl, int *flags) {
  return edges_impl(I, ops, n, out, site_lnl, 1, flags);
}

int engine_edge_sites(Instance *I, const DevEdgeOp &op, double *out3, double *site5, int *flags) {
  return edges_impl(I, &op, 1, out3, site5, 5, flags);
}

// Rescale flags of an instance and its heavy part.  flags may be null (reset
// only, no wait).  The reset is queued on the streams, ahead of later work.
int engine_flags(Instance *I, int *flags, int reset) {
  std::lock_guard<std::mutex> lk_(I->mu);
  Instance *H = I->heavy.get();
  if (flags) {
    int f = 0, hf = 0;
    GLE_HIP(hipMemcpyAsync(&f, I->d_flags, sizeof(int), hipMemcpyDeviceToHost, I->stream));
    if (H) GLE_HIP(hipMemcpyAsync(&hf, H->d_flags, sizeof(int), hipMemcpyDeviceToHost, H->stream));
    GLE_HIP(hipStreamSynchronize(I->stream));
    if (H) GLE_HIP(hipStreamSynchronize(H->stream));
    *flags = f | hf;
  }
  if (reset) {
    GLE_HIP(hipMemsetAsync(I->d_flags, 0, sizeof(int), I->stream));
    if (H) GLE_HIP(hipMemsetAsync(H->d_flags, 0, sizeof(int), H->stream));
  }
  return GLE_SUCCESS;
}

static int engine_get_partials_one(Instance *I, int slot, double *out, int *u);

// One CLA slot in the caller's pattern order.  In a split FP32 instance the
// heavy patterns' counts are natural-log units (FP64 scheme), the others powers of two.
int engine_get_partials(Instance *I, int slot, double *out, int *u) {
  if (I->light_idx.empty() && !I->heavy) return engine_get_partials_one(I, slot, out, u);
  const size_t W = size_t(I->R) * I->S;
  std::vector<double> ol(std::max<size_t>(1, I->light_idx.size() * W)), oh;
  std::vector<int> ul(std::max<size_t>(1, I->light_idx.size())), uh;
  int rc = engine_get_partials_one(I, slot, ol.data(), ul.data());
  if (rc) return rc;
  if (I->heavy) {
    oh.resize(std::max<size_t>(1, I->heavy_idx.size() * W));
    uh.resize(std::max<size_t>(1, I->heavy_idx.size()));
    rc = engine_get_partials_one(I->heavy.get(), slot, oh.data(), uh.data());
    if (rc) return rc;
  }
  for (size_t k = 0; k < I->light_idx.size(); k++) {
    std::copy(ol.begin() + k * W, ol.begin() + (k + 1) * W, out + size_t(I->light_idx[k]) * W);
    u[I->light_idx[k]] = ul[k];
  }
  for (size_t k = 0; k < I->heavy_idx.size(); k++) {
    std::copy(oh.begin() + k * W, oh.begin() + (k + 1) * W, out + size_t(I->heavy_idx[k]) * W);
    u[I->heavy_idx[k]] = uh[k];
  }
  return GLE_SUCCESS;
}

static int engine_get_partials_one(Instance *I, int slot, double *out, int *u) {
  std::lock_guard<std::mutex> lk_(I->mu);
  if (slot < 0 || slot >= I->n_slots) return set_error(GLE_ERR_ARG, "slot %d out of range", slot);
  const size_t W = I->S == 61 ? size_t(I->P) * I->R * I->Sp : I->slot_elems();
  std::vector<unsigned char> tmp(W * I->elem_bytes());
  const char *base = reinterpret_cast<const char *>(I->d_cla) + slot * I->slot_elems() * I->elem_bytes();
  GLE_HIP(hipMemcpy(tmp.data(), base, tmp.size(), hipMemcpyDeviceToHost));
  GLE_HIP(hipMemcpy(u, I->d_u + size_t(slot) * I->Ppad, sizeof(int) * I->P, hipMemcpyDeviceToHost));
  for (int p = 0; p < I->P; p++)
    for (int r = 0; r < I->R; r++)
      for (int s = 0; s < I->S; s++) {
        const size_t src = I->S == 4    ? nuc_off(p, r, I->R) + s
                           : I->S == 20 ? aa_off(p, r, I->R) + 64 * size_t(s)
                                        : (size_t(p) * I->R + r) * I->Sp + s;
        const size_t dst = (size_t(p) * I->R + r) * I->S + s;
        out[dst] = I->prec == GLE_PREC_FP64 ? reinterpret_cast<const double *>(tmp.data())[src]
                                            : double(reinterpret_cast<const float *>(tmp.data())[src]);
      }
  return GLE_SUCCESS;
}

int engine_sync(Instance *I) {
  if (I->heavy) {
    const int rc = engine_sync(I->heavy.get());
    if (rc) return rc;
  }
  std::lock_guard<std::mutex> lk_(I->mu);
  GLE_HIP(hipStreamSynchronize(I->stream));
  return GLE_SUCCESS;
}

static int destroy(Instance *I) {
  I->drop_graphs();
  if (I->heavy) {
    destroy(I->heavy.get());
    I->heavy.reset();
  }
  void *ptrs[] = {I->d_tips, I->d_counts, I->d_cst, I->d_u, I->d_flags, I->d_eig_of_rate, I->d_cla,
                  I->d_pm64, I->d_pmpad,  I->d_pdiag, I->d_eigfrag, I->d_V, I->d_Vinv, I->d_lam, I->d_model,
                  I->d_ops,
                  I->d_eops, I->d_part,  I->d_tpart, I->d_eres, I->d_site, I->d_pmidx, I->d_pmt};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (I->h_stage) (void)hipHostFree(I->h_stage);
  if (I->h_flag) (void)hipHostFree(I->h_flag);
  for (void *p : I->retired) (void)hipFree(p);
  for (void *p : I->retired_host) (void)hipHostFree(p);
  I->retired.clear();
  I->retired_host.clear();
  for (auto &st : I->stage) {
    if (st.h) (void)hipHostFree(st.h);
    if (st.d) (void)hipFree(st.d);
    if (st.done) (void)hipEventDestroy(st.done);
  }
  for (auto &p : I->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (hipEvent_t e : I->ev_free) (void)hipEventDestroy(e);
  if (I->stream) (void)hipStreamDestroy(I->stream);
  return GLE_SUCCESS;
}

}  // namespace gle

// =====================================================================
// C ABI: instance layer
// =====================================================================
using namespace gle;

extern "C" const char *gle_last_error(void) { return g_last_error.c_str(); }

extern "C" int gle_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

extern "C" int gle_init(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
    return set_error(GLE_ERR_DEVICE, "no HIP device visible: the GARLI engine runs only on a GPU");
  if (device < 0 || device >= n) return set_error(GLE_ERR_ARG, "device %d out of range (%d visible)", device, n);
  g_device = -1;
  return ensure_device(device);
}

static int create_internal(const gle_instance_desc *d, std::unique_ptr<Instance> &out) {
  auto I = std::make_unique<Instance>();
  I->N = d->n_tips;
  I->P = d->n_patterns;
  I->P_orig = d->n_patterns;
  I->Ppad = (d->n_patterns + 255) / 256 * 256;
  I->S = d->n_states;
  I->Sp = d->n_states == 61 ? 64 : d->n_states;
  I->R = d->n_rate_cats;
  I->n_eig = d->n_eigen;
  I->n_slots = d->n_cla_slots;
  I->n_pm = d->n_pmat_slots;
  I->prec = d->precision;
  const int S = I->S, R = I->R;
  auto fail = [&](int rc) {
    destroy(I.get());
    return rc;
  };
#define GLE_ALLOC(ptr, bytes)                                                                         \
  do {                                                                                                \
    if (hipMalloc(&(ptr), (bytes)) != hipSuccess)                                                     \
      return fail(set_error(GLE_ERR_NOMEM, "device allocation of %zu bytes failed", size_t(bytes))); \
  } while (0)
  if (hipStreamCreateWithFlags(&I->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(set_error(GLE_ERR_DEVICE, "stream creation failed"));
  if (hipHostMalloc(reinterpret_cast<void **>(&I->h_flag), 64) != hipSuccess)
    return fail(set_error(GLE_ERR_NOMEM, "pinned allocation failed"));
  GLE_ALLOC(I->d_tips, size_t(I->N) * I->Ppad);
  GLE_ALLOC(I->d_counts, sizeof(int) * I->Ppad);
  GLE_ALLOC(I->d_cst, sizeof(int) * I->Ppad);
  GLE_ALLOC(I->d_u, sizeof(int) * size_t(I->n_slots) * I->Ppad);
  GLE_ALLOC(I->d_flags, sizeof(int));
  GLE_ALLOC(I->d_eig_of_rate, sizeof(int) * R);
  GLE_ALLOC(I->d_cla, I->slot_elems() * I->elem_bytes() * I->n_slots);
  GLE_ALLOC(I->d_pm64, sizeof(double) * size_t(I->n_pm) * 3 * R * S * S);
  if (S == 61 || S == 20) GLE_ALLOC(I->d_pmpad, sizeof(double) * size_t(I->n_pm) * 3 * R * I->Sp * I->Sp);
  if (I->prec == GLE_PREC_FP32) GLE_ALLOC(I->d_pdiag, sizeof(double) * size_t(I->n_pm) * R * 3 * S);
  if (S == 61 && I->prec == GLE_PREC_FP32) GLE_ALLOC(I->d_eigfrag, sizeof(double) * size_t(I->n_eig) * 2 * 4096);
  GLE_ALLOC(I->d_V, sizeof(double) * size_t(I->n_eig) * S * S);
  GLE_ALLOC(I->d_Vinv, sizeof(double) * size_t(I->n_eig) * S * S);
  GLE_ALLOC(I->d_lam, sizeof(double) * size_t(I->n_eig) * S);
  GLE_ALLOC(I->d_model, sizeof(double) * (S + 2 * R));
  GLE_ALLOC(I->d_site, sizeof(double) * 5 * size_t(I->Ppad));
  // Serving ticks' staging (both stages), their events and edge partials
  // for 64 edges, made here so that an instance's first served requests do
  // not wait for pinned allocations; a larger tick grows them.
  for (auto &st : I->stage) {
    const size_t cap = size_t(256) << 10;
    if (hipHostMalloc(reinterpret_cast<void **>(&st.h), cap) != hipSuccess)
      return fail(set_error(GLE_ERR_NOMEM, "pinned allocation failed"));
    GLE_ALLOC(st.d, cap);
    st.cap = cap;
    if (hipEventCreateWithFlags(&st.done, hipEventDisableTiming) != hipSuccess)
      return fail(set_error(GLE_ERR_DEVICE, "event creation failed"));
  }
  I->part_cap = size_t(64) * edge_blocks(I.get()) * 3;
  GLE_ALLOC(I->d_part, sizeof(double) * I->part_cap);
#undef GLE_ALLOC
  // zero everything the kernels may read before the caller sets it
  const unsigned char missing = (unsigned char)(S == 4 ? 15 : S);
  if (hipMemset(I->d_tips, missing, size_t(I->N) * I->Ppad) != hipSuccess ||
      hipMemset(I->d_counts, 0, sizeof(int) * I->Ppad) != hipSuccess ||
      hipMemset(I->d_cst, 0xff, sizeof(int) * I->Ppad) != hipSuccess ||
      hipMemset(I->d_u, 0, sizeof(int) * size_t(I->n_slots) * I->Ppad) != hipSuccess ||
      hipMemset(I->d_flags, 0, sizeof(int)) != hipSuccess ||
      hipMemset(I->d_cla, 0, I->slot_elems() * I->elem_bytes() * I->n_slots) != hipSuccess ||
      hipMemset(I->d_pm64, 0, sizeof(double) * size_t(I->n_pm) * 3 * R * S * S) != hipSuccess ||
      (I->d_pmpad && hipMemset(I->d_pmpad, 0, sizeof(double) * size_t(I->n_pm) * 3 * R * I->Sp * I->Sp) !=
                         hipSuccess) ||
      (I->d_pdiag && hipMemset(I->d_pdiag, 0, sizeof(double) * size_t(I->n_pm) * R * 3 * S) != hipSuccess) ||
      (I->d_eigfrag && hipMemset(I->d_eigfrag, 0, sizeof(double) * size_t(I->n_eig) * 2 * 4096) != hipSuccess))
    return fail(set_error(GLE_ERR_DEVICE, "device memset failed"));
  I->h_tips.assign(size_t(I->N) * I->P_orig, missing);
  I->h_pi.assign(S, 1.0 / S);
  I->h_w.assign(R, 1.0 / R);
  I->h_mult.assign(R, 1.0);
  I->h_eor.assign(R, 0);
  I->h_V.assign(size_t(I->n_eig) * S * S, 0.0);
  I->h_Vinv.assign(size_t(I->n_eig) * S * S, 0.0);
  I->h_lam.assign(size_t(I->n_eig) * S, 0.0);
  I->h_eig_set.assign(I->n_eig, 0);
  std::vector<double> model(S + 2 * R, 0.0);
  for (int q = 0; q < S; q++) model[q] = 1.0 / S;
  for (int r = 0; r < R; r++) {
    model[S + r] = 1.0 / R;
    model[S + R + r] = 1.0;
  }
  if (hipMemcpy(I->d_eig_of_rate, I->h_eor.data(), sizeof(int) * R, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(I->d_model, model.data(), sizeof(double) * model.size(), hipMemcpyHostToDevice) != hipSuccess)
    return fail(set_error(GLE_ERR_DEVICE, "device upload failed"));
  out = std::move(I);
  return GLE_SUCCESS;
}

// ---- pushing host state to a device instance (the main one or its heavy part)
static int push_tip_row(Instance *X, const Instance *M, int tip, const std::vector<int> *idx) {
  const unsigned char missing = (unsigned char)(X->S == 4 ? 15 : X->S);
  std::vector<unsigned char> m(X->Ppad, missing);
  const unsigned char *row = M->h_tips.data() + size_t(tip) * M->P_orig;
  for (int k = 0; k < X->P; k++) m[k] = row[idx ? (*idx)[k] : k];
  GLE_HIP(hipMemcpy(X->d_tips + size_t(tip) * X->Ppad, m.data(), X->Ppad, hipMemcpyHostToDevice));
  return GLE_SUCCESS;
}

static int push_patterns(Instance *X, const Instance *M, const std::vector<int> *idx) {
  std::vector<int> c(X->Ppad, 0), k(X->Ppad, -1);
  for (int q = 0; q < X->P; q++) {
    const int o = idx ? (*idx)[q] : q;
    c[q] = M->h_counts[o];
    k[q] = M->h_cst[o];
  }
  GLE_HIP(hipMemcpy(X->d_counts, c.data(), sizeof(int) * X->Ppad, hipMemcpyHostToDevice));
  GLE_HIP(hipMemcpy(X->d_cst, k.data(), sizeof(int) * X->Ppad, hipMemcpyHostToDevice));
  return GLE_SUCCESS;
}

static int push_model(Instance *X, const Instance *M) {
  const int S = X->S, R = X->R;
  std::vector<double> m(S + 2 * R);
  for (int q = 0; q < S; q++) m[q] = M->h_pi[q];
  for (int r = 0; r < R; r++) {
    m[S + r] = M->h_w[r];
    m[S + R + r] = M->h_mult[r];
  }
  GLE_HIP(hipMemcpy(X->d_model, m.data(), sizeof(double) * m.size(), hipMemcpyHostToDevice));
  GLE_HIP(hipMemcpy(X->d_eig_of_rate, M->h_eor.data(), sizeof(int) * R, hipMemcpyHostToDevice));
  GLE_HIP(hipMemcpy(X->d_V, M->h_V.data(), sizeof(double) * M->h_V.size(), hipMemcpyHostToDevice));
  GLE_HIP(hipMemcpy(X->d_Vinv, M->h_Vinv.data(), sizeof(double) * M->h_Vinv.size(), hipMemcpyHostToDevice));
  GLE_HIP(hipMemcpy(X->d_lam, M->h_lam.data(), sizeof(double) * M->h_lam.size(), hipMemcpyHostToDevice));
  if (X->d_eigfrag) {
    // V^-1 (output k, input q) and V^T (output k, input f) as zero-padded 64 x 64
    // blocks in fragment order, the layout codon_mfma_tile reads
    std::vector<double> f(size_t(X->n_eig) * 2 * 4096, 0.0);
    for (int e = 0; e < X->n_eig; e++)
      for (int i = 0; i < 61; i++)
        for (int j = 0; j < 61; j++) {
          f[size_t(e) * 8192 + codon_frag(i, j)] = M->h_Vinv[size_t(e) * 3721 + i * 61 + j];
          f[size_t(e) * 8192 + 4096 + codon_frag(i, j)] = M->h_V[size_t(e) * 3721 + j * 61 + i];
        }
    GLE_HIP(hipMemcpy(X->d_eigfrag, f.data(), sizeof(double) * f.size(), hipMemcpyHostToDevice));
  }
  X->pinv = M->pinv;
  X->pinv_in_model = M->pinv_in_model;
  return GLE_SUCCESS;
}

// Split the patterns by weight (FP32 instances) and push tips/counts to both parts.
static int apply_patterns(Instance *I) {
  I->light_idx.clear();
  I->heavy_idx.clear();
  if (I->prec == GLE_PREC_FP32 && I->heavy_threshold > 0)
    for (int q = 0; q < I->P_orig; q++)
      (I->h_counts[q] > I->heavy_threshold ? I->heavy_idx : I->light_idx).push_back(q);
  int rc;
  if (I->heavy_idx.empty()) {
    I->light_idx.clear();
    if (I->heavy) {
      destroy(I->heavy.get());
      I->heavy.reset();
    }
    I->P = I->P_orig;
    if ((rc = push_patterns(I, I, nullptr))) return rc;
    for (int t = 0; t < I->N; t++)
      if ((rc = push_tip_row(I, I, t, nullptr))) return rc;
    return GLE_SUCCESS;
  }
  I->P = int(I->light_idx.size());
  if (!I->heavy || I->heavy->P != int(I->heavy_idx.size())) {
    if (I->heavy) {
      destroy(I->heavy.get());
      I->heavy.reset();
    }
    gle_instance_desc d;
    d.n_tips = I->N;
    d.n_patterns = int(I->heavy_idx.size());
    d.n_states = I->S;
    d.n_rate_cats = I->R;
    d.n_eigen = I->n_eig;
    d.n_cla_slots = I->n_slots;
    d.n_pmat_slots = I->n_pm;
    d.precision = GLE_PREC_FP64;
    if ((rc = create_internal(&d, I->heavy))) return rc;
    I->heavy->heavy_threshold = 0;
    I->heavy->timing = I->timing;
    I->heavy->fused_min_patterns = I->fused_min_patterns;
    // P(t) slots already filled for the main part stay valid for the new one
    GLE_HIP(hipStreamSynchronize(I->stream));
    GLE_HIP(hipMemcpy(I->heavy->d_pm64, I->d_pm64, sizeof(double) * size_t(I->n_pm) * 3 * I->R * I->S * I->S,
                      hipMemcpyDeviceToDevice));
    if (I->d_pmpad)
      GLE_HIP(hipMemcpy(I->heavy->d_pmpad, I->d_pmpad, sizeof(double) * size_t(I->n_pm) * 3 * I->R * I->Sp * I->Sp,
                        hipMemcpyDeviceToDevice));
  }
  if ((rc = push_model(I->heavy.get(), I))) return rc;
  if ((rc = push_patterns(I, I, &I->light_idx))) return rc;
  if ((rc = push_patterns(I->heavy.get(), I, &I->heavy_idx))) return rc;
  for (int t = 0; t < I->N; t++) {
    if ((rc = push_tip_row(I, I, t, &I->light_idx))) return rc;
    if ((rc = push_tip_row(I->heavy.get(), I, t, &I->heavy_idx))) return rc;
  }
  return GLE_SUCCESS;
}

static int push_tip(Instance *I, int tip) {
  int rc = push_tip_row(I, I, tip, I->light_idx.empty() ? nullptr : &I->light_idx);
  if (!rc && I->heavy) rc = push_tip_row(I->heavy.get(), I, tip, &I->heavy_idx);
  return rc;
}

static int push_model_all(Instance *I) {
  int rc = push_model(I, I);
  if (!rc && I->heavy) rc = push_model(I->heavy.get(), I);
  return rc;
}

extern "C" int gle_create_instance(const gle_instance_desc *d, int *handle) {
  if (!d || !handle) return set_error(GLE_ERR_ARG, "null argument");
  if (d->n_tips < 2 || d->n_patterns < 1 || d->n_states < 2 || d->n_states > 64 || d->n_rate_cats < 1 ||
      d->n_rate_cats > 8 || d->n_eigen < 1 || d->n_eigen > d->n_rate_cats || d->n_cla_slots < 1 ||
      d->n_pmat_slots < 1 || (d->precision != GLE_PREC_FP64 && d->precision != GLE_PREC_FP32))
    return set_error(GLE_ERR_ARG, "invalid instance description");
  if (d->n_states != 4 && d->n_states != 20 && d->n_states != 61)
    return set_error(GLE_ERR_UNSUPPORTED, "state count %d: kernels exist for 4, 20 and 61", d->n_states);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    int rc = ensure_device();
    if (rc) return rc;
  }
  std::unique_ptr<Instance> I;
  const int rc = create_internal(d, I);
  if (rc) return rc;
  {
    // the CLA slot pool (its free list holds every slot) is made with the
    // instance rather than by its first tree or served tick
    std::lock_guard<std::mutex> pk(pool_mutex());
    pool_for(I.get());
  }
  std::lock_guard<std::mutex> lk(g_mu);
  g_instances.push_back(std::move(I));
  *handle = int(g_instances.size()) - 1;
  return GLE_SUCCESS;
}

extern "C" int gle_destroy_instance(int h) {
  std::lock_guard<std::mutex> pk(pool_mutex());
  std::lock_guard<std::mutex> lk(g_mu);
  if (h < 0 || h >= int(g_instances.size()) || !g_instances[h]) return set_error(GLE_ERR_ARG, "bad instance");
  destroy(g_instances[h].get());
  g_instances[h].reset();
  return GLE_SUCCESS;
}

#define GET_INSTANCE(h)                                                  \
  Instance *I = get_instance(h);                                         \
  if (!I) return set_error(GLE_ERR_ARG, "bad instance handle %d", (h));

extern "C" int gle_set_tip_ambig_stream(int h, int tip, const signed char *stream, int len) {
  GET_INSTANCE(h);
  I->generation++;
  if (I->S != 4) return set_error(GLE_ERR_ARG, "ambiguity streams are nucleotide data");
  if (tip < 0 || tip >= I->N || !stream) return set_error(GLE_ERR_ARG, "bad tip");
  std::vector<unsigned char> m(I->P_orig, 15);
  int used = ambig_string_to_masks(stream, len, I->P_orig, m.data());
  if (used < 0) return set_error(GLE_ERR_ARG, "malformed ambiguity stream for tip %d", tip);
  std::copy(m.begin(), m.end(), I->h_tips.begin() + size_t(tip) * I->P_orig);
  return push_tip(I, tip);
}

extern "C" int gle_set_tip_states(int h, int tip, const unsigned char *states) {
  GET_INSTANCE(h);
  I->generation++;
  if (tip < 0 || tip >= I->N || !states) return set_error(GLE_ERR_ARG, "bad tip");
  for (int q = 0; q < I->P_orig; q++) {
    const int v = states[q];
    if ((I->S == 4 && (v < 1 || v > 15)) || (I->S != 4 && v > I->S))
      return set_error(GLE_ERR_ARG, "tip %d pattern %d: code %d out of range", tip, q, v);
  }
  std::copy(states, states + I->P_orig, I->h_tips.begin() + size_t(tip) * I->P_orig);
  return push_tip(I, tip);
}

extern "C" int gle_set_patterns(int h, const int *counts, int last_const, const int *cst) {
  GET_INSTANCE(h);
  I->generation++;
  if (!counts || last_const < -1 || last_const >= I->P_orig || (last_const >= 0 && !cst))
    return set_error(GLE_ERR_ARG, "bad pattern info");
  for (int q = 0; q < I->P_orig; q++)
    if (counts[q] < 0) return set_error(GLE_ERR_ARG, "negative count");
  I->h_counts.assign(counts, counts + I->P_orig);
  I->h_cst.assign(I->P_orig, -1);
  for (int q = 0; q <= last_const; q++) I->h_cst[q] = cst[q];
  I->h_last_const = last_const;
  I->h_patterns = true;
  I->last_const = last_const;
  return apply_patterns(I);
}

extern "C" int gle_set_heavy_split(int h, int threshold) {
  GET_INSTANCE(h);
  I->generation++;
  if (threshold < 0) return set_error(GLE_ERR_ARG, "threshold must be >= 0");
  I->heavy_threshold = threshold;
  return I->h_patterns ? apply_patterns(I) : GLE_SUCCESS;
}

extern "C" int gle_set_fused_threshold(int h, int min_patterns) {
  GET_INSTANCE(h);
  if (min_patterns < 0) return set_error(GLE_ERR_ARG, "threshold must be >= 0");
  I->fused_min_patterns = min_patterns;
  if (I->heavy) I->heavy->fused_min_patterns = min_patterns;
  return GLE_SUCCESS;
}

extern "C" int gle_set_graph_replay(int h, int enable, int *replays) {
  GET_INSTANCE(h);
  std::lock_guard<std::mutex> lk_(I->mu);
  for (Instance *X : {I, I->heavy.get()}) {
    if (!X) continue;
    X->graphs = enable != 0;
    X->drop_graphs();
  }
  if (replays) *replays = I->g_replays;
  return GLE_SUCCESS;
}

extern "C" int gle_get_heavy_split(int h, int *n_heavy) {
  GET_INSTANCE(h);
  if (!n_heavy) return set_error(GLE_ERR_ARG, "null output");
  *n_heavy = int(I->heavy_idx.size());
  return GLE_SUCCESS;
}

extern "C" int gle_set_model(int h, const double *pi, const double *w, double pinv, int pinv_in_model) {
  GET_INSTANCE(h);
  I->generation++;
  if (!pi || !w || pinv < 0.0 || pinv >= 1.0) return set_error(GLE_ERR_ARG, "bad model");
  I->h_pi.assign(pi, pi + I->S);
  I->h_w.assign(w, w + I->R);
  I->pinv = pinv;
  I->pinv_in_model = pinv_in_model ? 1 : 0;
  I->h_model = true;
  return push_model_all(I);
}

extern "C" int gle_set_eigen(int h, int e, const double *V, const double *Vinv, const double *lam) {
  GET_INSTANCE(h);
  I->generation++;
  if (e < 0 || e >= I->n_eig || !V || !Vinv || !lam) return set_error(GLE_ERR_ARG, "bad eigensystem");
  const size_t SS = size_t(I->S) * I->S;
  std::copy(V, V + SS, I->h_V.begin() + e * SS);
  std::copy(Vinv, Vinv + SS, I->h_Vinv.begin() + e * SS);
  std::copy(lam, lam + I->S, I->h_lam.begin() + size_t(e) * I->S);
  I->h_eig_set[e] = 1;
  return push_model_all(I);
}

extern "C" int gle_set_category_rates(int h, const int *eig_of_rate, const double *mult) {
  GET_INSTANCE(h);
  I->generation++;
  if (!eig_of_rate || !mult) return set_error(GLE_ERR_ARG, "null argument");
  for (int r = 0; r < I->R; r++)
    if (eig_of_rate[r] < 0 || eig_of_rate[r] >= I->n_eig || !(mult[r] >= 0.0))
      return set_error(GLE_ERR_ARG, "bad category %d", r);
  I->h_eor.assign(eig_of_rate, eig_of_rate + I->R);
  I->h_mult.assign(mult, mult + I->R);
  I->h_rates = true;
  return push_model_all(I);
}

extern "C" int gle_update_transition_matrices(int h, const int *idx, const double *t, int n, int derivs) {
  GET_INSTANCE(h);
  if (n < 0 || (n > 0 && (!idx || !t))) return set_error(GLE_ERR_ARG, "bad pmat op list");
  return engine_update_pmats(I, idx, t, n, derivs);
}

extern "C" int gle_get_transition_matrices(int h, int pm, double *out) {
  GET_INSTANCE(h);
  if (pm < 0 || pm >= I->n_pm || !out) return set_error(GLE_ERR_ARG, "bad pmat slot");
  const size_t n = size_t(3) * I->R * I->S * I->S;
  GLE_HIP(hipMemcpy(out, I->d_pm64 + pm * n, sizeof(double) * n, hipMemcpyDeviceToHost));
  return GLE_SUCCESS;
}

// The caller's level ranges: non-decreasing, covering [0, n); the ops of one
// level run concurrently, so none of them may read or overwrite a CLA that
// another op of the same level writes (BlockingOperationsSet semantics,
// calculationmanager.h:118-123).
static int check_levels(const gle_cla_op *ops, int n, const int *level_start, int n_levels) {
  if (level_start[0] != 0 || level_start[n_levels] != n)
    return set_error(GLE_ERR_ARG, "level ranges must cover the op list");
  for (int l = 0; l < n_levels; l++)
    if (level_start[l + 1] < level_start[l] || level_start[l + 1] > n)
      return set_error(GLE_ERR_ARG, "level %d range [%d, %d) is not inside [0, %d) in order", l, level_start[l],
                       level_start[l + 1], n);
  std::unordered_set<int> dest;
  for (int l = 0; l < n_levels; l++) {
    dest.clear();
    for (int i = level_start[l]; i < level_start[l + 1]; i++)
      if (!dest.insert(ops[i].dest).second)
        return set_error(GLE_ERR_ARG, "level %d writes slot %d twice", l, ops[i].dest);
    for (int i = level_start[l]; i < level_start[l + 1]; i++)
      for (int c : {ops[i].child1, ops[i].child2})
        if (c >= 0 && dest.count(c))
          return set_error(GLE_ERR_ARG, "op %d reads slot %d, which its own level %d writes", i, c, l);
  }
  return GLE_SUCCESS;
}

extern "C" int gle_update_partials(int h, const gle_cla_op *ops, int n, const int *level_start, int n_levels) {
  if (n < 0 || (n > 0 && !ops) || (n_levels > 0 && !level_start) || n_levels < 0)
    return set_error(GLE_ERR_ARG, "bad op list");
  if (n_levels > 0) {
    const int rc = check_levels(ops, n, level_start, n_levels);
    if (rc) return rc;
  }
  GET_INSTANCE(h);
  std::vector<DevClaOp> d(n);
  for (int i = 0; i < n; i++)
    d[i] = {ops[i].dest, ops[i].child1, ops[i].pmat1, ops[i].child2, ops[i].pmat2, ops[i].rescale ? kOpRescale : 0};
  int rc = engine_update_partials(I, d.data(), n, level_start, n_levels, false);
  if (rc || I->prec != GLE_PREC_FP64) return rc;
  // RescaleRateHet's throw(1) (tree.cpp:3436-3449): a rescaled site whose
  // maximum fell below 1e-190 since the flags were last reset
  int flags = 0;
  rc = engine_flags(I, &flags, 0);
  if (rc) return rc;
  return (flags & GLE_FLAG_REDUCE_RESCALE) ? GLE_RESCALE_RETRY : GLE_SUCCESS;
}

extern "C" int gle_get_partials(int h, int slot, double *out, int *u) {
  GET_INSTANCE(h);
  if (!out || !u) return set_error(GLE_ERR_ARG, "null output");
  return engine_get_partials(I, slot, out, u);
}

extern "C" int gle_edge_likelihoods(int h, const gle_edge_op *ops, int n, double *out, double *site) {
  GET_INSTANCE(h);
  if (n < 0 || (n > 0 && (!ops || !out))) return set_error(GLE_ERR_ARG, "bad edge op list");
  std::vector<DevEdgeOp> d(n);
  for (int i = 0; i < n; i++) d[i] = {ops[i].partial, ops[i].child, ops[i].pmat, ops[i].derivs ? 1 : 0};
  return engine_edges(I, d.data(), n, out, site);
}

static size_t round256(size_t b) { return (b + 255) & ~size_t(255); }
// Scratch bytes one part (an instance or its heavy part) needs for n edge ops,
// and its buffers carved from `base` in the same order.
static size_t edge_scratch_bytes(const Instance *X, int n) {
  return round256(sizeof(DevEdgeOp) * size_t(n)) + round256(sizeof(double) * 3 * size_t(n)) +
         round256(sizeof(double) * 3 * size_t(n) * size_t(edge_blocks(X)));
}
static EdgeBufs carve_edge_scratch(unsigned char *base, int n) {
  EdgeBufs b;
  b.ops = reinterpret_cast<DevEdgeOp *>(base);
  base += round256(sizeof(DevEdgeOp) * size_t(n));
  b.res = reinterpret_cast<double *>(base);
  base += round256(sizeof(double) * 3 * size_t(n));
  b.part = reinterpret_cast<double *>(base);
  return b;
}

// gle_edge_likelihoods with every device buffer taken from the library's
// scratch array: no allocation on any path.  Each block writes its own
// partial sums and each op its own result, so nothing has to start at zero.
extern "C" int gle_edge_likelihoods_noalloc(int h, const gle_edge_op *ops, int n, double *out, double *site) {
  GET_INSTANCE(h);
  if (n < 0 || (n > 0 && (!ops || !out))) return set_error(GLE_ERR_ARG, "bad edge op list");
  if (n > kNoAllocMaxOps)
    return set_error(GLE_ERR_ARG, "%d edge ops: the allocation-free entry point takes at most %d", n,
                     kNoAllocMaxOps);
  if (n == 0) return GLE_SUCCESS;
  DevEdgeOp d[kNoAllocMaxOps];
  for (int i = 0; i < n; i++) d[i] = {ops[i].partial, ops[i].child, ops[i].pmat, ops[i].derivs ? 1 : 0};
  std::lock_guard<std::mutex> lk_(I->mu);
  Instance *H = I->heavy.get();
  if (site && H)
    return set_error(GLE_ERR_UNSUPPORTED,
                     "site values of a split FP32 instance need gle_edge_likelihoods (host-side reordering)");
  const size_t main_bytes = edge_scratch_bytes(I, n);
  const size_t need = main_bytes + (H ? edge_scratch_bytes(H, n) : 0);
  if (need > kNoAllocScratchBytes)
    return set_error(GLE_ERR_ARG, "%d edge ops over %d patterns need %zu bytes of scratch; the maximum is %zu", n,
                     I->P_orig, need, kNoAllocScratchBytes);
  int dev = 0;
  GLE_HIP(hipGetDevice(&dev));
  int rc = scratch_setup(dev);  // set up already, unless this device was never initialised by the library
  if (rc) return rc;
  unsigned char *base = g_scratch_dev[dev];
  const EdgeBufs bm = carve_edge_scratch(base, n);
  rc = enqueue_edges(I, d, n, site != nullptr, &bm);
  EdgeBufs bh{};
  if (!rc && H) {
    bh = carve_edge_scratch(base + main_bytes, n);
    rc = enqueue_edges(H, d, n, false, &bh);
  }
  double out_h[3 * kNoAllocMaxOps];
  if (!rc) rc = copy_edges_out(I, n, out, site, 1, bm.res);
  if (!rc && H) rc = copy_edges_out(H, n, out_h, nullptr, 1, bh.res);
  rc = finish_both(I, rc);
  if (rc) return rc;
  if (H)
    for (int i = 0; i < 3 * n; i++) out[i] += out_h[i];
  return GLE_SUCCESS;
}

extern "C" int gle_edge_site_values(int h, const gle_edge_op *op, double *out3, double *site_lnl, double *site_d1,
                                    double *site_d2, int *under_partial, int *under_child) {
  GET_INSTANCE(h);
  if (!op || !out3) return set_error(GLE_ERR_ARG, "null argument");
  const DevEdgeOp d = {op->partial, op->child, op->pmat, 1};
  const size_t P = size_t(instance_patterns(I));
  std::vector<double> s5(5 * std::max<size_t>(P, 1));
  const int rc = engine_edge_sites(I, d, out3, s5.data());
  if (rc) return rc;
  for (size_t q = 0; q < P; q++) {
    if (site_lnl) site_lnl[q] = s5[q];
    if (site_d1) site_d1[q] = s5[P + q];
    if (site_d2) site_d2[q] = s5[2 * P + q];
    if (under_partial) under_partial[q] = int(s5[3 * P + q]);
    if (under_child) under_child[q] = int(s5[4 * P + q]);
  }
  return GLE_SUCCESS;
}

extern "C" int gle_get_rescale_flags(int h, int *flags, int reset) {
  GET_INSTANCE(h);
  if (!flags) return set_error(GLE_ERR_ARG, "null flags");
  return engine_flags(I, flags, reset);
}

extern "C" int gle_set_timing(int h, int enable) {
  GET_INSTANCE(h);
  I->timing = enable != 0;
  if (I->heavy) I->heavy->timing = I->timing;
  return GLE_SUCCESS;
}

extern "C" int gle_get_kernel_stats(int h, int kind, long long *launches, double *ms, double *bytes, int reset) {
  GET_INSTANCE(h);
  if (kind < 0 || kind > 5 || !launches || !ms || !bytes) return set_error(GLE_ERR_ARG, "bad stats request");
  Instance *X = kind < 3 ? I : I->heavy.get();
  const int k = kind % 3;
  if (!X) {
    *launches = 0;
    *ms = *bytes = 0.0;
    return GLE_SUCCESS;
  }
  *launches = X->kst[k].launches;
  *ms = X->kst[k].ms;
  *bytes = X->kst[k].bytes;
  if (reset) X->kst[k] = Instance::KStat();
  return GLE_SUCCESS;
}

extern "C" int gle_copy_calibrate(size_t bytes, int reps, double *ms_per_copy) {
  if (bytes < 16 || reps < 1 || !ms_per_copy) return set_error(GLE_ERR_ARG, "bad calibration request");
  {
    std::lock_guard<std::mutex> lk(g_mu);
    int rc = ensure_device();
    if (rc) return rc;
  }
  const size_t n = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  GLE_HIP(hipMalloc(&a, n * 16));
  if (hipMalloc(&b, n * 16) != hipSuccess) {
    (void)hipFree(a);
    return set_error(GLE_ERR_NOMEM, "calibration buffer");
  }
  (void)hipMemset(a, 0, n * 16);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipLaunchKernelGGL(k_copy_calibrate, dim3(4096), dim3(256), 0, nullptr, a, b, n);
  (void)hipEventRecord(e0, nullptr);
  for (int i = 0; i < reps; i++) hipLaunchKernelGGL(k_copy_calibrate, dim3(4096), dim3(256), 0, nullptr, a, b, n);
  (void)hipEventRecord(e1, nullptr);
  const hipError_t err = hipEventSynchronize(e1);
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  *ms_per_copy = double(ms) / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (err != hipSuccess) return set_error(GLE_ERR_DEVICE, "calibration copy failed: %s", hipGetErrorString(err));
  return GLE_SUCCESS;
}

extern "C" int gle_mfma_calibrate_shape(int shape, double *tflops) {
  if (!tflops || (shape != 0 && shape != 1)) return set_error(GLE_ERR_ARG, "bad calibration request");
  {
    std::lock_guard<std::mutex> lk(g_mu);
    int rc = ensure_device();
    if (rc) return rc;
  }
  constexpr int kBlocks = 2048, kThreads = 256, kIters = 500;  // 8 waves per SIMD on 256 CUs
  double *out = nullptr;
  GLE_HIP(hipMalloc(&out, sizeof(double) * kBlocks * kThreads));
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  auto k = shape == 0 ? k_mfma_calibrate : k_mfma_calibrate44;
  hipLaunchKernelGGL(k, dim3(kBlocks), dim3(kThreads), 0, nullptr, out, kIters);
  (void)hipEventRecord(e0, nullptr);
  hipLaunchKernelGGL(k, dim3(kBlocks), dim3(kThreads), 0, nullptr, out, kIters);
  (void)hipEventRecord(e1, nullptr);
  const hipError_t err = hipEventSynchronize(e1);
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  // flops per wave and iteration: 4 chains x 2048 (16x16x4) or 8 chains x 512 (4x4x4, 4 blocks)
  const double flops = double(kBlocks) * (kThreads / 64) * kIters * (shape == 0 ? 4 * 2048.0 : 8 * 512.0);
  *tflops = flops / (double(ms) * 1e-3) / 1e12;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(out);
  if (err != hipSuccess) return set_error(GLE_ERR_DEVICE, "MFMA calibration failed: %s", hipGetErrorString(err));
  return GLE_SUCCESS;
}

extern "C" int gle_mfma_calibrate(double *tflops) { return gle_mfma_calibrate_shape(0, tflops); }

extern "C" int gle_synchronize(int h) {
  GET_INSTANCE(h);
  return engine_sync(I);
}
