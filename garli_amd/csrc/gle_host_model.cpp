// gle_host_model.cpp -- host-side model and data helpers of libgle.
//
// Q-matrix builds, the reversible eigensystem, discrete-gamma rates, tip
// encodings and pattern compression.  These are the parts of GARLI's Model and
// DataPartition that feed the device path; all FP64, all O(S^3) or one-time.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "gle.h"
#include "gle_internal.h"

namespace gle {

// ---------------------------------------------------------------- Q builds

// model.cpp:256-293: GTR with GT fixed to 1; normaliser 0.5 / sum_{i<j} pi_i q_ij.
static void qmat_gtr(const double *pi, const double *rel, double *q, double *bm) {
  double r6[6] = {rel[0], rel[1], rel[2], rel[3], rel[4], 1.0};
  static const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {1, 2, 3, 2, 3, 3};
  std::fill(q, q + 16, 0.0);
  for (int k = 0; k < 6; k++) {
    q[ii[k] * 4 + jj[k]] = r6[k] * pi[jj[k]];
    q[jj[k] * 4 + ii[k]] = r6[k] * pi[ii[k]];
  }
  for (int x = 0; x < 4; x++) {
    double s = 0.0;
    for (int y = 0; y < 4; y++)
      if (y != x) s += q[x * 4 + y];
    q[x * 4 + x] = -s;
  }
  *bm = 0.5 / ((q[1] * pi[0]) + (q[2] * pi[0]) + (q[3] * pi[0]) + (q[6] * pi[1]) + (q[7] * pi[1]) +
               (q[11] * pi[2]));
}

// model.cpp:735-780
static void qmat_aa(const double *pi, const double *rel, double *q, double *bm) {
  const int n = 20;
  int k = 0;
  std::vector<double> R(n * n, 0.0);
  for (int a = 0; a < n; a++)
    for (int b = a + 1; b < n; b++) R[a * n + b] = R[b * n + a] = rel[k++];
  double w = 0.0;
  for (int a = 0; a < n; a++) {
    double s = 0.0;
    for (int b = 0; b < n; b++) {
      q[a * n + b] = (a == b) ? 0.0 : pi[b] * R[a * n + b];
      if (a != b) s += q[a * n + b];
    }
    q[a * n + a] = -s;
    w += s * pi[a];
  }
  *bm = 1.0 / w;
}

// Standard genetic code, codons numbered 16*n1+4*n2+n3 with A,C,G,T = 0..3;
// amino acids numbered alphabetically by one-letter code, stop = 20.
static int codon_aa(int c) {
  // T,C,A,G-ordered public table, re-indexed to A,C,G,T
  static const char tcag[] = "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG";
  static const int m[4] = {2, 1, 3, 0};
  static const char order[] = "ACDEFGHIKLMNPQRSTVWY";
  char a = tcag[16 * m[c >> 4] + 4 * m[(c >> 2) & 3] + m[c & 3]];
  const char *p = std::strchr(order, a);
  return (a == '*' || !p) ? 20 : int(p - order);
}

int standard_code_aa(int codon64) { return codon_aa(codon64); }

// model.cpp:296-687: one-position changes only; the changed nucleotide pair's
// exchangeability (rel6 = AC AG AT CG CT GT), omega_w on nonsynonymous changes;
// normaliser over the omega mixture.
static void qmat_codon(const double *pi, const double *rel6, const double *omega, const double *oprob, int nw,
                       double *Q, double *bm) {
  static const int pair_index[4][4] = {{-1, 0, 1, 2}, {0, -1, 3, 4}, {1, 3, -1, 5}, {2, 4, 5, -1}};
  int sense[61], n = 0;
  for (int c = 0; c < 64; c++)
    if (codon_aa(c) != 20) sense[n++] = c;
  double tot = 0.0;
  for (int w = 0; w < nw; w++) {
    double *q = Q + size_t(w) * 61 * 61, wsum = 0.0;
    for (int i = 0; i < 61; i++) {
      double rs = 0.0;
      for (int j = 0; j < 61; j++) {
        double v = 0.0;
        if (i != j) {
          int a = sense[i], b = sense[j], nd = 0, x = 0, y = 0;
          for (int s = 0; s < 3; s++) {
            int xa = (a >> (4 - 2 * s)) & 3, xb = (b >> (4 - 2 * s)) & 3;
            if (xa != xb) { nd++; x = xa; y = xb; }
          }
          if (nd == 1) {
            v = pi[j] * rel6[pair_index[x][y]];
            if (codon_aa(a) != codon_aa(b)) v *= omega[w];
          }
        }
        q[i * 61 + j] = v;
        rs += v;
      }
      q[i * 61 + i] = -rs;
      wsum += rs * pi[i];
    }
    tot += wsum * oprob[w];
  }
  *bm = 1.0 / tot;
}

// ------------------------------------------------------------ eigensystem

// Reversible Q: B = D^{1/2} Q D^{-1/2} is symmetric.  Two-sided cyclic Jacobi on B
// gives B = W diag(mu) W^T; then Q = (D^{-1/2} W) diag(mu) (W^T D^{1/2}).
int reversible_eigen(int n, const double *Q, const double *pi, double *V, double *Vinv, double *lam) {
  std::vector<double> B(n * n), W(n * n, 0.0), sp(n);
  for (int i = 0; i < n; i++) {
    if (!(pi[i] > 0.0)) return GLE_ERR_ARG;
    sp[i] = std::sqrt(pi[i]);
    W[i * n + i] = 1.0;
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      B[i * n + j] = 0.5 * (sp[i] * Q[i * n + j] / sp[j] + sp[j] * Q[j * n + i] / sp[i]);
  for (int sweep = 0; sweep < 64; sweep++) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < n; i++) {
      diag += B[i * n + i] * B[i * n + i];
      for (int j = i + 1; j < n; j++) off += 2.0 * B[i * n + j] * B[i * n + j];
    }
    if (off <= 1e-32 * (diag + off)) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        const double bpq = B[p * n + q];
        if (std::fabs(bpq) < 1e-300) continue;
        const double tau = (B[q * n + q] - B[p * n + p]) / (2.0 * bpq);
        const double t = std::copysign(1.0, tau) / (std::fabs(tau) + std::hypot(1.0, tau));
        const double c = 1.0 / std::hypot(1.0, t), s = t * c;
        for (int k = 0; k < n; k++) {
          const double x = B[k * n + p], y = B[k * n + q];
          B[k * n + p] = c * x - s * y;
          B[k * n + q] = s * x + c * y;
        }
        for (int k = 0; k < n; k++) {
          const double x = B[p * n + k], y = B[q * n + k];
          B[p * n + k] = c * x - s * y;
          B[q * n + k] = s * x + c * y;
        }
        for (int k = 0; k < n; k++) {
          const double x = W[k * n + p], y = W[k * n + q];
          W[k * n + p] = c * x - s * y;
          W[k * n + q] = s * x + c * y;
        }
      }
  }
  for (int k = 0; k < n; k++) lam[k] = B[k * n + k];
  for (int i = 0; i < n; i++)
    for (int k = 0; k < n; k++) {
      V[i * n + k] = W[i * n + k] / sp[i];
      Vinv[k * n + i] = W[i * n + k] * sp[i];
    }
  return GLE_SUCCESS;
}

// ------------------------------------------------------ discrete gamma
// Published algorithms the reference uses (model.cpp:1883-2201):
// ACM 291 log-gamma, AS 32 incomplete gamma ratio, AS 111 normal and AS 91
// chi-square percentage points.

static double ln_gamma(double a) {
  double x = a, f = 0.0;
  if (x < 7.0) {
    f = 1.0;
    double z = x - 1.0;
    while (++z < 7.0) f *= z;
    x = z;
    f = -std::log(f);
  }
  const double z = 1.0 / (x * x);
  return f + (x - 0.5) * std::log(x) - x + 0.918938533204673 +
         (((-0.000595238095238 * z + 0.000793650793651) * z - 0.002777777777778) * z + 0.083333333333333) / x;
}

static double inc_gamma(double x, double p, double lg) {
  const double acc = 1e-8, overflow = 1e30;
  if (x == 0.0) return 0.0;
  if (x < 0.0 || p <= 0.0) return -1.0;
  const double factor = std::exp(p * std::log(x) - x - lg);
  if (!(x > 1.0 && x >= p)) {
    double g = 1.0, term = 1.0, rn = p;
    do {
      rn += 1.0;
      term *= x / rn;
      g += term;
    } while (term > acc);
    return g * factor / p;
  }
  double a = 1.0 - p, b = a + x + 1.0, term = 0.0;
  double pn[6] = {1.0, x, x + 1.0, x * b, 0.0, 0.0};
  double g = pn[2] / pn[3];
  while (true) {
    a += 1.0;
    b += 2.0;
    term += 1.0;
    const double an = a * term;
    pn[4] = b * pn[2] - an * pn[0];
    pn[5] = b * pn[3] - an * pn[1];
    if (pn[5] != 0.0) {
      const double rn = pn[4] / pn[5], dif = std::fabs(g - rn);
      if (dif <= acc && dif <= acc * rn) return 1.0 - factor * g;
      g = rn;
    }
    for (int i = 0; i < 4; i++) pn[i] = pn[i + 2];
    if (std::fabs(pn[4]) >= overflow)
      for (int i = 0; i < 4; i++) pn[i] /= overflow;
  }
}

static double normal_point(double prob) {
  const double a[5] = {-0.322232431088, -1.0, -0.342242088547, -0.0204231210245, -0.453642210148e-4};
  const double b[5] = {0.0993484626060, 0.588581570495, 0.531103462366, 0.103537752850, 0.0038560700634};
  const double p1 = prob < 0.5 ? prob : 1.0 - prob;
  if (p1 < 1e-20) return -9999.0;
  const double y = std::sqrt(std::log(1.0 / (p1 * p1)));
  const double z = y + ((((y * a[4] + a[3]) * y + a[2]) * y + a[1]) * y + a[0]) /
                           ((((y * b[4] + b[3]) * y + b[2]) * y + b[1]) * y + b[0]);
  return prob < 0.5 ? -z : z;
}

static double chi2_point(double p, double v) {
  const double e = 0.5e-6, aa = 0.6931471805;
  if (p < 0.000002 || p > 0.999998 || v <= 0.0) return -1.0;
  const double g = ln_gamma(0.5 * v), xx = 0.5 * v, c = xx - 1.0;
  double ch, q;
  if (v < -1.24 * std::log(p)) {
    ch = std::pow(p * xx * std::exp(g + xx * aa), 1.0 / xx);
    if (ch - e < 0.0) return ch;
  } else if (v <= 0.32) {
    ch = 0.4;
    const double a = std::log(1.0 - p);
    do {
      q = ch;
      const double p1 = 1.0 + ch * (4.67 + ch), p2 = ch * (6.73 + ch * (6.66 + ch));
      const double t = -0.5 + (4.67 + 2.0 * ch) / p1 - (6.73 + ch * (13.32 + 3.0 * ch)) / p2;
      ch -= (1.0 - std::exp(a + g + 0.5 * ch + c * aa) * p2 / p1) / t;
    } while (std::fabs(q / ch - 1.0) - 0.01 > 0.0);
  } else {
    const double x = normal_point(p), p1 = 0.222222 / v;
    ch = v * std::pow(x * std::sqrt(p1) + 1.0 - p1, 3.0);
    if (ch > 2.2 * v + 6.0) ch = -2.0 * (std::log(1.0 - p) - c * std::log(0.5 * ch) + g);
  }
  do {
    q = ch;
    const double p1 = 0.5 * ch;
    double t = inc_gamma(p1, xx, g);
    if (t < 0.0) return -1.0;
    const double p2 = p - t;
    t = p2 * std::exp(xx * aa + g + p1 - c * std::log(ch));
    const double b = t / ch, a = 0.5 * t - b * c;
    const double s1 = (210.0 + a * (140.0 + a * (105.0 + a * (84.0 + a * (70.0 + 60.0 * a))))) / 420.0;
    const double s2 = (420.0 + a * (735.0 + a * (966.0 + a * (1141.0 + 1278.0 * a)))) / 2520.0;
    const double s3 = (210.0 + a * (462.0 + a * (707.0 + 932.0 * a))) / 2520.0;
    const double s4 = (252.0 + a * (672.0 + 1182.0 * a) + c * (294.0 + a * (889.0 + 1740.0 * a))) / 5040.0;
    const double s5 = (84.0 + 264.0 * a + c * (175.0 + 606.0 * a)) / 2520.0;
    const double s6 = (120.0 + c * (346.0 + 127.0 * c)) / 5040.0;
    ch += t * (1 + 0.5 * t * s1 - b * c * (s1 - b * (s2 - b * (s3 - b * (s4 - b * (s5 - b * s6))))));
  } while (std::fabs(q / ch - 1.0) > e);
  return ch;
}

int discrete_gamma(int K, double alpha, double *rates) {
  if (K < 1 || !(alpha > 0.0)) return GLE_ERR_ARG;
  if (K == 1) {
    rates[0] = 1.0;
    return GLE_SUCCESS;
  }
  std::vector<double> cut(K - 1);
  const double lg1 = ln_gamma(alpha + 1.0), scale = double(K);
  for (int i = 0; i < K - 1; i++) {
    const double x = chi2_point((i + 1.0) / K, 2.0 * alpha) / (2.0 * alpha);
    cut[i] = inc_gamma(x * alpha, alpha + 1.0, lg1);
  }
  rates[0] = cut[0] * scale;
  for (int i = 1; i < K - 1; i++) rates[i] = (cut[i] - cut[i - 1]) * scale;
  rates[K - 1] = (1.0 - cut[K - 2]) * scale;
  return GLE_SUCCESS;
}

// ------------------------------------------------------------ encodings

int nuc_datum(char c) {
  switch (c >= 'a' && c <= 'z' ? char(c - 32) : c) {
    case 'A': return 1;
    case 'C': return 2;
    case 'G': return 4;
    case 'T': case 'U': return 8;
    case 'M': return 3;
    case 'R': return 5;
    case 'S': return 6;
    case 'V': return 7;
    case 'W': return 9;
    case 'Y': return 10;
    case 'H': return 11;
    case 'K': return 12;
    case 'D': return 13;
    case 'B': return 14;
    case 'N': case '-': case '?': return 15;
    default: return -1;
  }
}

int aa_datum(char c) {
  static const char order[] = "ACDEFGHIKLMNPQRSTVWY";
  const char u = (c >= 'a' && c <= 'z') ? char(c - 32) : c;
  if (u == 'B' || u == 'Z' || u == 'X' || u == '-' || u == '?') return 20;
  if (u == 0) return -1;
  const char *p = std::strchr(order, u);
  return p ? int(p - order) : -1;
}

int make_ambig_string(const unsigned char *masks, int n, signed char *out) {
  int k = 0;
  for (int j = 0; j < n; j++) {
    const int m = masks[j] & 15, bits = __builtin_popcount(unsigned(m));
    if (bits == 1) out[k++] = (signed char)__builtin_ctz(unsigned(m));
    else if (bits == 0 || bits == 4) out[k++] = -4;
    else {
      out[k++] = (signed char)(-bits);
      for (int s = 0; s < 4; s++)
        if (m & (1 << s)) out[k++] = (signed char)s;
    }
  }
  return k;
}

int ambig_string_to_masks(const signed char *st, int len, int n, unsigned char *masks) {
  int k = 0;
  for (int j = 0; j < n; j++) {
    if (k >= len) return GLE_ERR_ARG;
    const int c = st[k++];
    if (c >= 0 && c < 4) masks[j] = (unsigned char)(1 << c);
    else if (c == -4) masks[j] = 15;
    else if (c == -2 || c == -3) {
      int m = 0;
      for (int s = 0; s < -c; s++) {
        if (k >= len || st[k] < 0 || st[k] > 3) return GLE_ERR_ARG;
        m |= 1 << st[k++];
      }
      masks[j] = (unsigned char)m;
    } else
      return GLE_ERR_ARG;
  }
  return k;
}

// datamatr.cpp:82-209: number of states of a column, the primary sort key.
static int column_states(const unsigned char *col, int ntax, int ms) {
  if (ms == 4) {
    unsigned char con = 15;
    bool partial = false;
    unsigned seen = 0;
    for (int t = 0; t < ntax; t++) {
      const unsigned char c = col[t];
      con &= c;
      if (c == 15) continue;
      if (c & (c - 1)) partial = true;
      else seen |= c;
    }
    if (partial) return con ? (con == 15 ? 0 : 1) : 2;
    const int k = __builtin_popcount(seen);
    return k;
  }
  std::vector<char> seen(ms, 0);
  int k = 0;
  for (int t = 0; t < ntax; t++)
    if (col[t] != ms && !seen[col[t]]) {
      seen[col[t]] = 1;
      k++;
    }
  return k;
}

int compress_patterns(int ntax, int nsite, int ms, const unsigned char *matrix, unsigned char *out, int *counts,
                      int *const_states, int *last_const, int *site_to_pattern) {
  if (ntax < 1 || nsite < 0 || ms < 2 || ms > 255) return GLE_ERR_ARG;
  std::vector<unsigned char> cols(size_t(ntax) * nsite);
  std::vector<int> ns(nsite), order(nsite);
  for (int s = 0; s < nsite; s++) {
    for (int t = 0; t < ntax; t++) {
      const unsigned char v = matrix[size_t(t) * nsite + s];
      if ((ms == 4 && (v < 1 || v > 15)) || (ms != 4 && v > ms)) return GLE_ERR_ARG;
      cols[size_t(s) * ntax + t] = v;
    }
    ns[s] = column_states(&cols[size_t(s) * ntax], ntax, ms);
  }
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    if (ns[a] != ns[b]) return ns[a] < ns[b];
    return std::memcmp(&cols[size_t(a) * ntax], &cols[size_t(b) * ntax], size_t(ntax)) < 0;
  });
  int P = 0;
  *last_const = -1;
  std::vector<unsigned char> packed;
  packed.reserve(size_t(nsite) * ntax);
  for (int i = 0; i < nsite;) {
    const unsigned char *ci = &cols[size_t(order[i]) * ntax];
    int j = i;
    while (j < nsite && std::memcmp(&cols[size_t(order[j]) * ntax], ci, size_t(ntax)) == 0) j++;
    if (ns[order[i]] == 0) {
      for (int k = i; k < j; k++) site_to_pattern[order[k]] = -1;
    } else {
      packed.insert(packed.end(), ci, ci + ntax);
      counts[P] = j - i;
      for (int k = i; k < j; k++) site_to_pattern[order[k]] = P;
      int cs = 0;
      if (ns[order[i]] == 1) {
        *last_const = P;
        if (ms == 4) {
          cs = 15;
          for (int t = 0; t < ntax; t++) cs &= ci[t];
        } else {
          cs = ms;
          for (int t = 0; t < ntax && cs == ms; t++) cs = ci[t];
        }
      }
      const_states[P] = cs;
      P++;
    }
    i = j;
  }
  for (int t = 0; t < ntax; t++)
    for (int p = 0; p < P; p++) out[size_t(t) * P + p] = packed[size_t(p) * ntax + t];
  return P;
}

}  // namespace gle

// ------------------------------------------------------------------ C ABI
using namespace gle;

extern "C" int gle_model_gtr(const double *pi, const double *rel, double *V, double *Vinv, double *lam) {
  if (!pi || !rel || !V || !Vinv || !lam) return set_error(GLE_ERR_ARG, "null argument");
  double q[16], bm;
  qmat_gtr(pi, rel, q, &bm);
  int rc = reversible_eigen(4, q, pi, V, Vinv, lam);
  if (rc) return set_error(rc, "state frequencies must be positive");
  for (int k = 0; k < 4; k++) lam[k] *= bm;
  return GLE_SUCCESS;
}

extern "C" int gle_model_aa(const double *pi, const double *rel, double *V, double *Vinv, double *lam) {
  if (!pi || !rel || !V || !Vinv || !lam) return set_error(GLE_ERR_ARG, "null argument");
  std::vector<double> q(400);
  double bm;
  qmat_aa(pi, rel, q.data(), &bm);
  int rc = reversible_eigen(20, q.data(), pi, V, Vinv, lam);
  if (rc) return set_error(rc, "state frequencies must be positive");
  for (int k = 0; k < 20; k++) lam[k] *= bm;
  return GLE_SUCCESS;
}

extern "C" int gle_model_codon(const double *pi, const double *rel6, const double *omega, const double *oprob,
                               int nw, double *V, double *Vinv, double *lam) {
  if (!pi || !rel6 || !omega || !oprob || nw < 1 || !V || !Vinv || !lam) return set_error(GLE_ERR_ARG, "bad argument");
  std::vector<double> Q(size_t(nw) * 61 * 61);
  double bm;
  qmat_codon(pi, rel6, omega, oprob, nw, Q.data(), &bm);
  for (int w = 0; w < nw; w++) {
    int rc = reversible_eigen(61, Q.data() + size_t(w) * 3721, pi, V + size_t(w) * 3721, Vinv + size_t(w) * 3721,
                              lam + w * 61);
    if (rc) return set_error(rc, "codon frequencies must be positive");
    for (int k = 0; k < 61; k++) lam[w * 61 + k] *= bm;
  }
  return GLE_SUCCESS;
}

extern "C" int gle_discrete_gamma(int n, double alpha, double *rates) {
  if (!rates) return set_error(GLE_ERR_ARG, "null rates");
  int rc = discrete_gamma(n, alpha, rates);
  return rc ? set_error(rc, "bad discrete gamma arguments") : GLE_SUCCESS;
}
extern "C" int gle_nuc_char_to_datum(char c) { return nuc_datum(c); }
extern "C" int gle_aa_char_to_datum(char c) { return aa_datum(c); }
extern "C" int gle_make_ambig_string(const unsigned char *m, int n, signed char *out) {
  if (!m || !out || n < 0) return set_error(GLE_ERR_ARG, "bad argument");
  return make_ambig_string(m, n, out);
}
extern "C" int gle_ambig_string_to_masks(const signed char *st, int len, int n, unsigned char *masks) {
  if (!st || !masks || n < 0) return set_error(GLE_ERR_ARG, "bad argument");
  int rc = ambig_string_to_masks(st, len, n, masks);
  return rc < 0 ? set_error(rc, "malformed ambiguity stream") : rc;
}
extern "C" int gle_compress_patterns(int ntax, int nsite, int ms, const unsigned char *matrix, unsigned char *out,
                                     int *counts, int *cs, int *lc, int *s2p) {
  if (!matrix || !out || !counts || !cs || !lc || !s2p) return set_error(GLE_ERR_ARG, "null argument");
  int rc = compress_patterns(ntax, nsite, ms, matrix, out, counts, cs, lc, s2p);
  return rc < 0 ? set_error(rc, "bad matrix for pattern compression") : rc;
}

// ------------------------------------------------------ state frequencies

// AminoacidData / CodonData::CalcEmpiricalFreqs (sequencedata.cpp:192-275):
// every non-missing state of every taxon counted with its pattern's count; one
// pseudo-count per state when a state is never observed.  Amino-acid data
// whose A, C, G and T make up more than 90% is refused as nucleotides
// (sequencedata.cpp:210).
extern "C" int gle_empirical_freqs(int ntax, int npat, int ns, const unsigned char *st, const int *counts,
                                   double *pi) {
  if (ntax < 1 || npat < 0 || ns < 2 || ns > 255 || !st || !counts || !pi) return set_error(GLE_ERR_ARG, "bad argument");
  std::vector<double> f(size_t(ns), 0.0);
  double total = 0.0;
  for (int t = 0; t < ntax; t++)
    for (int p = 0; p < npat; p++) {
      const int v = st[size_t(t) * npat + p];
      if (v == ns) continue;
      if (v > ns) return set_error(GLE_ERR_ARG, "state %d outside 0..%d", v, ns);
      f[v] += counts[p];
      total += counts[p];
    }
  if (!(total > 0.0)) return set_error(GLE_ERR_ARG, "no observed states");
  if (ns == 20 && (f[0] + f[1] + f[5] + f[16]) / total > 0.90)
    return set_error(GLE_ERR_ARG, "amino-acid model on what looks like nucleotide data");
  if (std::find(f.begin(), f.end(), 0.0) != f.end()) {
    for (double &x : f) x += 1.0;
    total += ns;
  }
  for (int j = 0; j < ns; j++) pi[j] = f[j] / total;
  return GLE_SUCCESS;
}

// NucleotideData::CalcEmpiricalFreqs (sequencedata.cpp:61-190).  Cells with
// one base count for it; partly ambiguous cells (2 or 3 bases) are shared
// among their bases, first equally and then in proportion to the current
// frequencies, repeated until no frequency moves by more than 1e-8 (the
// fixed point of that EM iteration); fully ambiguous cells are skipped.  One
// pseudo-count per base when a base is never seen.  The reference loops per
// cell; here the ambiguous cells are first summed per bit set, which changes
// only the rounding.
extern "C" int gle_empirical_nuc_freqs(int ntax, int npat, const unsigned char *masks, const int *counts, double *pi) {
  if (ntax < 1 || npat < 0 || !masks || !counts || !pi) return set_error(GLE_ERR_ARG, "bad argument");
  double single[4] = {0, 0, 0, 0}, amb[4] = {0, 0, 0, 0}, by_mask[16] = {0}, n_single = 0.0, n_amb = 0.0;
  for (int t = 0; t < ntax; t++)
    for (int p = 0; p < npat; p++) {
      const int m = masks[size_t(t) * npat + p] & 15, bits = __builtin_popcount(unsigned(m));
      const double c = counts[p];
      if (bits == 1) {
        single[__builtin_ctz(unsigned(m))] += c;
        n_single += c;
      } else if (bits == 2 || bits == 3) {
        for (int s = 0; s < 4; s++)
          if (m >> s & 1) amb[s] += c / bits;
        n_amb += c;
        by_mask[m] += c;
      }
    }
  for (int j = 0; j < 4; j++)
    if (single[j] + amb[j] == 0.0) {
      for (double &x : single) x += 1.0;
      n_single += 4.0;
      break;
    }
  const double total = n_single + n_amb;
  if (!(total > 0.0)) return set_error(GLE_ERR_ARG, "no observed bases");
  for (int j = 0; j < 4; j++) pi[j] = (single[j] + amb[j]) / total;
  if (n_amb > 0.0)
    for (int it = 0; it < 100000; it++) {
      double a[4] = {0, 0, 0, 0};
      for (int m = 1; m < 16; m++) {
        if (by_mask[m] == 0.0) continue;
        double frac = 0.0;
        for (int s = 0; s < 4; s++)
          if (m >> s & 1) frac += pi[s];
        for (int s = 0; s < 4; s++)
          if (m >> s & 1) a[s] += by_mask[m] * (pi[s] / frac);
      }
      bool moved = false;
      for (int j = 0; j < 4; j++) {
        const double v = (single[j] + a[j]) / total;
        if (std::fabs(v - pi[j]) > 1e-8) moved = true;
        pi[j] = v;
      }
      if (!moved) break;
    }
  return GLE_SUCCESS;
}

// CodonData::CalcF1x4Freqs / CalcF3x4Freqs (sequencedata.cpp:605-660):
// codon frequencies as products of base frequencies, pooled over the three
// positions (F1x4) or per position (F3x4), renormalised over the 61 sense
// codons.  base_freqs holds positions 1, 2, 3 (A, C, G, T each).
extern "C" int gle_codon_freqs_fx4(const double *bf, int f3x4, double *pi61) {
  if (!bf || !pi61) return set_error(GLE_ERR_ARG, "null argument");
  double pos[3][4];
  for (int k = 0; k < 3; k++)
    for (int b = 0; b < 4; b++) pos[k][b] = f3x4 ? bf[4 * k + b] : (bf[b] + bf[4 + b] + bf[8 + b]) / 3.0;
  if (f3x4)
    for (int k = 0; k < 3; k++)
      for (int b = 0; b < 4; b++)
        if (pos[k][b] == 0.0) return set_error(GLE_ERR_ARG, "F3x4: base %d never seen at codon position %d", b, k + 1);
  double total = 0.0;
  int n = 0;
  for (int c = 0; c < 64; c++) {
    if (codon_aa(c) == 20) continue;
    pi61[n] = pos[0][c >> 4] * pos[1][(c >> 2) & 3] * pos[2][c & 3];
    total += pi61[n++];
  }
  if (!(total > 0.0)) return set_error(GLE_ERR_ARG, "F1x4/F3x4: no base frequencies");
  for (int j = 0; j < 61; j++) pi61[j] /= total;
  return GLE_SUCCESS;
}
