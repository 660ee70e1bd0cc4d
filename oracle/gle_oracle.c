/*
 * gle_oracle.c -- FP64 CPU oracle for the GARLI tree-likelihood path.
 * TEST INFRASTRUCTURE ONLY (see gle_oracle.h): never linked into the engine.
 *
 * Each function cites the reference lines it restates (paths relative to
 * garli/trunk/src).  Loop orders and summation orders follow the reference so
 * that an FP64 engine with contraction disabled can be compared to ~1 ulp.
 */
#include "gle_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORA_RESCALE_LEN 90 /* tree.h:48 */

/* ======================================================================
 * Substitution models
 * ==================================================================== */

/* model.cpp:256-293 (GTR, GT exchangeability fixed at 1) */
void ora_qmat_gtr(const double pi[4], const double rel[5], double *Q, double *blen_mult) {
  double q[4][4];
  memset(q, 0, sizeof q);
  q[0][1] = rel[0] * pi[1];
  q[0][2] = rel[1] * pi[2];
  q[0][3] = rel[2] * pi[3];
  q[1][2] = rel[3] * pi[2];
  q[1][3] = rel[4] * pi[3];
  q[2][3] = pi[3];
  q[1][0] = rel[0] * pi[0];
  q[2][0] = rel[1] * pi[0];
  q[2][1] = rel[3] * pi[1];
  q[3][0] = rel[2] * pi[0];
  q[3][1] = rel[4] * pi[1];
  q[3][2] = pi[2];
  for (int x = 0; x < 4; x++) {
    double sum = 0.0;
    for (int y = 0; y < 4; y++)
      if (x != y) sum += q[x][y];
    q[x][x] = -sum;
  }
  *blen_mult = 0.5 / ((q[0][1] * pi[0]) + (q[0][2] * pi[0]) + (q[0][3] * pi[0]) + (q[1][2] * pi[1]) +
                      (q[1][3] * pi[1]) + (q[2][3] * pi[2]));
  memcpy(Q, q, sizeof q);
}

/* model.cpp:735-780 (estimated / user AA matrix: rel190 in upper-triangle row order) */
void ora_qmat_aa(const double pi[20], const double *rel, double *Q, double *blen_mult) {
  const int n = 20;
  for (int f = 0; f < n; f++)
    for (int t = 0; t < n; t++) Q[f * n + t] = pi[t];
  int r = 0;
  for (int f = 0; f < n - 1; f++)
    for (int t = f + 1; t < n; t++) Q[f * n + t] *= rel[r++];
  r = 0;
  for (int t = 0; t < n - 1; t++)
    for (int f = t + 1; f < n; f++) Q[f * n + t] *= rel[r++];
  double wsum = 0.0;
  for (int f = 0; f < n; f++) {
    double sum = 0.0;
    for (int t = 0; t < n; t++)
      if (f != t) sum += Q[f * n + t];
    Q[f * n + f] = -sum;
    wsum += sum * pi[f];
  }
  *blen_mult = 1.0 / wsum;
}

/* Standard genetic code in T,C,A,G order (public table); amino acids are
 * renumbered alphabetically by one-letter code (A C D E F G H I K L M N P Q R S T
 * V W Y = 0..19, stop = 20), which is the reference's numbering
 * (sequencedata.h:217-283, sequencedata.cpp:715-780). */
static const char *ORA_TCAG_CODE = "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG";
static int ora_aa_index(char a) {
  static const char *order = "ACDEFGHIKLMNPQRSTVWY";
  const char *p = strchr(order, a);
  return (a == '*' || !p) ? 20 : (int)(p - order);
}
static int ora_codon64_aa(int c64) { /* codon index 16*n1+4*n2+n3, A=0 C=1 G=2 T=3 */
  static const int to_tcag[4] = {2, 1, 3, 0};
  int n1 = c64 >> 4, n2 = (c64 >> 2) & 3, n3 = c64 & 3;
  return ora_aa_index(ORA_TCAG_CODE[16 * to_tcag[n1] + 4 * to_tcag[n2] + to_tcag[n3]]);
}

/* model.cpp:296-440: bit1 one-nucleotide change, bit2 transition, bit4
 * nonsynonymous, bits 8..256 which nucleotide pair changed; stops removed. */
void ora_codon_lookup(int *lookup, int *aa_of_codon61) {
  int idx[64], n = 0;
  for (int c = 0; c < 64; c++) idx[c] = (ora_codon64_aa(c) == 20) ? -1 : n++;
  for (int c = 0; c < 64; c++)
    if (idx[c] >= 0 && aa_of_codon61) aa_of_codon61[idx[c]] = ora_codon64_aa(c);
  for (int a = 0; a < 64; a++) {
    if (idx[a] < 0) continue;
    for (int b = 0; b < 64; b++) {
      if (idx[b] < 0) continue;
      int code = 0, ndiff = 0, pa = 0, pb = 0;
      for (int pos = 0; pos < 3; pos++) {
        int x = (a >> (2 * (2 - pos))) & 3, y = (b >> (2 * (2 - pos))) & 3;
        if (x != y) { ndiff++; pa = x; pb = y; }
      }
      if (ndiff == 1) {
        int lo = pa < pb ? pa : pb, hi = pa < pb ? pb : pa;
        code = 1;
        if (((lo + hi) & 1) == 0) code |= 2;
        static const int pairbit[4][4] = {{0, 8, 16, 32}, {0, 0, 64, 128}, {0, 0, 0, 256}, {0, 0, 0, 0}};
        code |= pairbit[lo][hi];
        if (ora_codon64_aa(a) != ora_codon64_aa(b)) code |= 4;
      }
      lookup[idx[a] * 61 + idx[b]] = code;
    }
  }
}

/* model.cpp:442-687: Q_ij = pi_j * r(pair changed) * (omega_w if nonsynonymous) for
 * single-nucleotide changes; rel6 = {AC, AG, AT, CG, CT, GT} (nst = 6; nst = 2 is
 * rel6 = {1, kappa, 1, 1, kappa, 1}); one Q per omega class. */
void ora_qmat_codon(const double *pi, const double *rel6, const double *omega, const double *omega_prob,
                    int nw, double *Q, double *blen_mult) {
  const int n = 61;
  static int lookup[61 * 61];
  ora_codon_lookup(lookup, NULL);
  for (int w = 0; w < nw; w++) {
    double *q = Q + (size_t)w * n * n;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        int l = lookup[i * n + j];
        if (l == 0) { q[i * n + j] = 0.0; continue; }
        q[i * n + j] = pi[j];
        if (l & 8) q[i * n + j] *= rel6[0];
        else if (l & 16) q[i * n + j] *= rel6[1];
        else if (l & 32) q[i * n + j] *= rel6[2];
        else if (l & 64) q[i * n + j] *= rel6[3];
        else if (l & 128) q[i * n + j] *= rel6[4];
        else if (l & 256) q[i * n + j] *= rel6[5];
        if (l & 4) q[i * n + j] *= omega[w];
      }
  }
  double bm = 0.0;
  for (int w = 0; w < nw; w++) {
    double *q = Q + (size_t)w * n * n;
    double wsum = 0.0;
    for (int x = 0; x < n; x++) {
      double sum = 0.0;
      for (int y = 0; y < n; y++)
        if (x != y) sum += q[x * n + y];
      q[x * n + x] = -sum;
      wsum += sum * pi[x];
    }
    bm += wsum * omega_prob[w];
  }
  *blen_mult = 1.0 / bm;
}

/* Eigensystem of a reversible Q: A = D^1/2 Q D^-1/2 is symmetric; cyclic Jacobi
 * gives A = U L U^T, so V = D^-1/2 U and Vinv = U^T D^1/2 (model.cpp:827-856 uses
 * EISPACK RG on Q; any basis gives the same P(t)). */
int ora_eigen(int n, const double *Q, const double *pi, double *V, double *Vinv, double *lam) {
  double *A = (double *)malloc(sizeof(double) * n * n);
  double *U = (double *)malloc(sizeof(double) * n * n);
  for (int i = 0; i < n; i++) {
    if (!(pi[i] > 0.0)) { free(A); free(U); return 1; }
    for (int j = 0; j < n; j++) U[i * n + j] = (i == j);
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double aij = sqrt(pi[i]) * Q[i * n + j] / sqrt(pi[j]);
      double aji = sqrt(pi[j]) * Q[j * n + i] / sqrt(pi[i]);
      A[i * n + j] = 0.5 * (aij + aji);
    }
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        tot += A[i * n + j] * A[i * n + j];
        if (i != j) off += A[i * n + j] * A[i * n + j];
      }
    if (off <= 1e-34 * tot) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = A[p * n + q];
        if (apq == 0.0) continue;
        double app = A[p * n + p], aqq = A[q * n + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) { /* columns p,q */
          double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) { /* rows p,q */
          double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          double ukp = U[k * n + p], ukq = U[k * n + q];
          U[k * n + p] = c * ukp - s * ukq;
          U[k * n + q] = s * ukp + c * ukq;
        }
      }
  }
  for (int k = 0; k < n; k++) lam[k] = A[k * n + k];
  for (int i = 0; i < n; i++)
    for (int k = 0; k < n; k++) {
      V[i * n + k] = U[i * n + k] / sqrt(pi[i]);
      Vinv[k * n + i] = U[i * n + k] * sqrt(pi[i]);
    }
  free(A);
  free(U);
  return 0;
}

/* model.cpp:1271-1358 (P, P', P'') and 1469-1597 (P only): x = V_ik*Vinv_kj,
 * summed over k in order; P clamped at 0, derivatives not. */
void ora_pmat(int n, const double *V, const double *Vinv, const double *lam, double mult, double t,
              double *P, double *D1, double *D2) {
  double ex[64], e1[64], e2[64];
  for (int k = 0; k < n; k++) {
    double s = lam[k] * mult;
    ex[k] = exp(s * t);
    e1[k] = s * ex[k];
    e2[k] = s * e1[k];
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double sp = 0.0, s1 = 0.0, s2 = 0.0;
      for (int k = 0; k < n; k++) {
        double x = V[i * n + k] * Vinv[k * n + j];
        sp += x * ex[k];
        s1 += x * e1[k];
        s2 += x * e2[k];
      }
      P[i * n + j] = sp > 0.0 ? sp : 0.0;
      if (D1) D1[i * n + j] = s1;
      if (D2) D2[i * n + j] = s2;
    }
}

void ora_pmats(const ora_ctx *c, double t, double *P, double *D1, double *D2) {
  const int S = c->nstates, SS = S * S;
  for (int r = 0; r < c->nrates; r++) {
    int e = c->eig_of_rate ? c->eig_of_rate[r] : 0;
    ora_pmat(S, c->V + (size_t)e * SS, c->Vinv + (size_t)e * SS, c->lam + (size_t)e * S, c->rate_mult[r], t,
             P + r * SS, D1 ? D1 + r * SS : NULL, D2 ? D2 + r * SS : NULL);
  }
}

/* Pike & Hill (1966) ACM 291 log-gamma (model.cpp LnGamma) */
double ora_lngamma(double alpha) {
  double x = alpha, f = 0.0, z;
  if (x < 7) {
    f = 1.0;
    z = x - 1.0;
    while (++z < 7.0) f *= z;
    x = z;
    f = -log(f);
  }
  z = 1.0 / (x * x);
  return f + (x - 0.5) * log(x) - x + 0.918938533204673 +
         (((-0.000595238095238 * z + 0.000793650793651) * z - 0.002777777777778) * z + 0.083333333333333) / x;
}

/* Bhattacharjee (1970) AS 32 incomplete gamma ratio, accuracy 1e-8 */
double ora_incomplete_gamma(double x, double alpha, double g) {
  const double acc = 1e-8, big = 1e30;
  double p = alpha, gin, rn, a, b, an, dif, term, pn[6];
  if (x == 0.0) return 0.0;
  if (x < 0 || p <= 0) return -1.0;
  double factor = exp(p * log(x) - x - g);
  if (!(x > 1 && x >= p)) { /* series */
    gin = 1.0;
    term = 1.0;
    rn = p;
    do {
      rn += 1.0;
      term *= x / rn;
      gin += term;
    } while (term > acc);
    return gin * factor / p;
  }
  /* continued fraction */
  a = 1.0 - p;
  b = a + x + 1.0;
  term = 0.0;
  pn[0] = 1.0;
  pn[1] = x;
  pn[2] = x + 1;
  pn[3] = x * b;
  gin = pn[2] / pn[3];
  for (;;) {
    a += 1.0;
    b += 2.0;
    term += 1.0;
    an = a * term;
    for (int i = 0; i < 2; i++) pn[i + 4] = b * pn[i + 2] - an * pn[i];
    if (pn[5] != 0) {
      rn = pn[4] / pn[5];
      dif = fabs(gin - rn);
      if (!(dif > acc) && dif <= acc * rn) return 1.0 - factor * gin;
      gin = rn;
    }
    for (int i = 0; i < 4; i++) pn[i] = pn[i + 2];
    if (!(fabs(pn[4]) < big))
      for (int i = 0; i < 4; i++) pn[i] /= big;
  }
}

/* Odeh & Evans (1974) AS 70/111 normal percentage point */
double ora_point_normal(double prob) {
  const double a0 = -0.322232431088, a1 = -1.0, a2 = -0.342242088547, a3 = -0.0204231210245,
               a4 = -0.453642210148e-4, b0 = 0.0993484626060, b1 = 0.588581570495, b2 = 0.531103462366,
               b3 = 0.103537752850, b4 = 0.0038560700634;
  double p = prob, p1 = (p < 0.5 ? p : 1 - p);
  if (p1 < 1e-20) return -9999;
  double y = sqrt(log(1 / (p1 * p1)));
  double z = y + ((((y * a4 + a3) * y + a2) * y + a1) * y + a0) / ((((y * b4 + b3) * y + b2) * y + b1) * y + b0);
  return p < 0.5 ? -z : z;
}

/* Best & Roberts (1975) AS 91 chi-square percentage point, e = 0.5e-6 */
double ora_point_chi2(double prob, double v) {
  const double e = 0.5e-6, aa = 0.6931471805;
  double p = prob, g, xx, c, ch, a = 0, q, p1, p2, t, x, b, s1, s2, s3, s4, s5, s6;
  if (p < 0.000002 || p > 0.999998 || v <= 0.0) return -1.0;
  g = ora_lngamma(v * 0.5);
  xx = v / 2.0;
  c = xx - 1.0;
  if (v < -1.24 * log(p)) {
    ch = pow((p * xx * exp(g + xx * aa)), 1.0 / xx);
    if (ch - e < 0) return ch;
  } else if (v <= 0.32) {
    ch = 0.4;
    a = log(1.0 - p);
    do {
      q = ch;
      p1 = 1.0 + ch * (4.67 + ch);
      p2 = ch * (6.73 + ch * (6.66 + ch));
      t = -0.5 + (4.67 + 2.0 * ch) / p1 - (6.73 + ch * (13.32 + 3.0 * ch)) / p2;
      ch -= (1.0 - exp(a + g + 0.5 * ch + c * aa) * p2 / p1) / t;
    } while (fabs(q / ch - 1.0) - 0.01 > 0.0);
  } else {
    x = ora_point_normal(p);
    p1 = 0.222222 / v;
    ch = v * pow((x * sqrt(p1) + 1.0 - p1), 3.0);
    if (ch > 2.2 * v + 6.0) ch = -2.0 * (log(1.0 - p) - c * log(0.5 * ch) + g);
  }
  do {
    q = ch;
    p1 = 0.5 * ch;
    t = ora_incomplete_gamma(p1, xx, g);
    if (t < 0.0) return -1.0;
    p2 = p - t;
    t = p2 * exp(xx * aa + g + p1 - c * log(ch));
    b = t / ch;
    a = 0.5 * t - b * c;
    s1 = (210.0 + a * (140.0 + a * (105.0 + a * (84.0 + a * (70.0 + 60.0 * a))))) / 420.0;
    s2 = (420.0 + a * (735.0 + a * (966.0 + a * (1141.0 + 1278.0 * a)))) / 2520.0;
    s3 = (210.0 + a * (462.0 + a * (707.0 + 932.0 * a))) / 2520.0;
    s4 = (252.0 + a * (672.0 + 1182.0 * a) + c * (294.0 + a * (889.0 + 1740.0 * a))) / 5040.0;
    s5 = (84.0 + 264.0 * a + c * (175.0 + 606.0 * a)) / 2520.0;
    s6 = (120.0 + c * (346.0 + 127.0 * c)) / 5040.0;
    ch += t * (1 + 0.5 * t * s1 - b * c * (s1 - b * (s2 - b * (s3 - b * (s4 - b * (s5 - b * s6))))));
  } while (fabs(q / ch - 1.0) > e);
  return ch;
}

/* model.cpp:2170-2201: mean-of-category discrete gamma (rates only) */
int ora_discrete_gamma(int K, double shape, double *rates) {
  if (K == 1) { rates[0] = 1.0; return 0; }
  double props[64];
  double factor = shape / shape * K, lnga1 = ora_lngamma(shape + 1);
  for (int i = 0; i < K - 1; i++) props[i] = ora_point_chi2((i + 1.0) / K, 2.0 * shape) / (2.0 * shape);
  for (int i = 0; i < K - 1; i++) props[i] = ora_incomplete_gamma(props[i] * shape, shape + 1, lnga1);
  rates[0] = props[0] * factor;
  rates[K - 1] = (1 - props[K - 2]) * factor;
  for (int i = 1; i < K - 1; i++) rates[i] = (props[i] - props[i - 1]) * factor;
  return 0;
}

/* ======================================================================
 * Data encoding
 * ==================================================================== */

/* sequencedata.cpp:672-713 */
int ora_nuc_char_to_datum(char ch) {
  switch (ch) {
    case 'A': case 'a': return 1;
    case 'C': case 'c': return 2;
    case 'G': case 'g': return 4;
    case 'T': case 't': case 'U': case 'u': return 8;
    case 'M': case 'm': return 3;
    case 'R': case 'r': return 5;
    case 'S': case 's': return 6;
    case 'V': case 'v': return 7;
    case 'W': case 'w': return 9;
    case 'Y': case 'y': return 10;
    case 'H': case 'h': return 11;
    case 'K': case 'k': return 12;
    case 'D': case 'd': return 13;
    case 'B': case 'b': return 14;
    case 'N': case 'n': case '-': case '?': return 15;
    default: return -1;
  }
}

/* sequencedata.cpp:715-780 (B, Z, X, gap and ? are all missing = 20) */
int ora_aa_char_to_datum(char ch) {
  static const char *order = "ACDEFGHIKLMNPQRSTVWY";
  char u = (ch >= 'a' && ch <= 'z') ? (char)(ch - 32) : ch;
  if (u == 'B' || u == 'Z' || u == 'X' || u == '-' || u == '?') return 20;
  const char *p = strchr(order, u);
  return (p && u) ? (int)(p - order) : -1;
}

/* sequencedata.cpp:278-362 */
int ora_make_ambig_string(const unsigned char *m, int n, signed char *out) {
  int k = 0;
  for (int j = 0; j < n; j++) {
    int b = m[j], ns = 0, code = 0;
    for (int s = 0; s < 4; s++)
      if (b & (1 << s)) { ns++; code = s; }
    if (ns == 1)
      out[k++] = (signed char)code;
    else if (ns == 4 || ns == 0)
      out[k++] = -4;
    else {
      out[k++] = (signed char)-ns;
      for (int s = 0; s < 4; s++)
        if (b & (1 << s)) out[k++] = (signed char)s;
    }
  }
  return k;
}

/* datamatr.cpp:82-209 -- the number of states used as the primary sort key */
static int ora_pattern_nstates(const unsigned char *col, int ntax, int max_states) {
  int counts[64] = {0};
  unsigned char full = (unsigned char)(max_states == 4 ? 15 : max_states);
  if (max_states == 4) {
    unsigned char con = 15;
    int ambig = 0;
    for (int t = 0; t < ntax; t++) {
      unsigned char c = col[t];
      con &= c;
      if (c != full && (c & (c - 1))) ambig = 1;
      else if (c != full) counts[(c > 1) + (c > 2) + (c > 4)]++;
    }
    if (ambig) {
      if (con) return con == 15 ? 0 : 1;
      return 2;
    }
  } else {
    for (int t = 0; t < ntax; t++)
      if (col[t] != full) counts[col[t]]++;
  }
  int ns = 0;
  for (int s = 0; s < max_states; s++) ns += counts[s] > 0;
  return ns < 2 ? ns : (ns > 2 ? ns : 2);
}

typedef struct { int nstates, site; const unsigned char *col; int ntax; } ora_pat;
static int ora_pat_cmp(const void *a, const void *b) {
  const ora_pat *x = (const ora_pat *)a, *y = (const ora_pat *)b;
  if (x->nstates != y->nstates) return x->nstates < y->nstates ? -1 : 1;
  int c = memcmp(x->col, y->col, (size_t)x->ntax);
  if (c) return c;
  return x->site < y->site ? -1 : (x->site > y->site); /* stable */
}

/* datamatr.cpp:243-375: sort by (#states, state vector), collapse, drop missing,
 * constant patterns first; nuc constant state = AND of the taxon masks. */
int ora_compress_patterns(int ntax, int nsite, int max_states, const unsigned char *matrix,
                          unsigned char *out, int *counts, int *const_states, int *last_const,
                          int *site_to_pattern) {
  unsigned char *cols = (unsigned char *)malloc((size_t)ntax * nsite);
  ora_pat *pats = (ora_pat *)malloc(sizeof(ora_pat) * (size_t)nsite);
  for (int s = 0; s < nsite; s++) {
    for (int t = 0; t < ntax; t++) cols[(size_t)s * ntax + t] = matrix[(size_t)t * nsite + s];
    pats[s].col = cols + (size_t)s * ntax;
    pats[s].ntax = ntax;
    pats[s].site = s;
    pats[s].nstates = ora_pattern_nstates(pats[s].col, ntax, max_states);
  }
  qsort(pats, (size_t)nsite, sizeof(ora_pat), ora_pat_cmp);
  int P = 0;
  *last_const = -1;
  for (int i = 0; i < nsite;) {
    int j = i;
    while (j < nsite && memcmp(pats[j].col, pats[i].col, (size_t)ntax) == 0) j++;
    if (pats[i].nstates == 0) {
      for (int k = i; k < j; k++) site_to_pattern[pats[k].site] = -1;
    } else {
      for (int t = 0; t < ntax; t++) out[(size_t)t * nsite + P] = pats[i].col[t];
      counts[P] = j - i;
      for (int k = i; k < j; k++) site_to_pattern[pats[k].site] = P;
      if (pats[i].nstates == 1) {
        *last_const = P;
        if (max_states == 4) {
          int c = 15;
          for (int t = 0; t < ntax; t++) c &= pats[i].col[t];
          const_states[P] = c;
        } else {
          int c = max_states, t = 0;
          do { c = pats[i].col[t]; t++; } while (c == max_states && t < ntax);
          const_states[P] = c;
        }
      } else
        const_states[P] = 0;
      P++;
    }
    i = j;
  }
  /* out was written with row stride nsite; repack to stride P */
  for (int t = 1; t < ntax; t++) memmove(out + (size_t)t * P, out + (size_t)t * nsite, (size_t)P);
  free(cols);
  free(pats);
  return P;
}

/* ======================================================================
 * CLA kernels (reference layout)
 * ==================================================================== */

/* advance a nucleotide ambiguity stream by one pattern (tree.cpp:125-136) */
static const signed char *ora_adv(const signed char *d) {
  if (*d > -1 || *d == -4) return d + 1;
  int n = -*d;
  return d + 1 + n;
}

/* tree.cpp:7668-7841 (nuc) and 7843-7910 (NState) */
void ora_cla_ii(const ora_ctx *c, double *dest, int *du, const double *LCL, const int *Lu,
                const double *RCL, const int *Ru, const double *Lpr, const double *Rpr) {
  const int S = c->nstates, R = c->nrates, P = c->npat;
  if (S == 4) {
    for (int i = 0; i < P; i++) {
      const double *l = LCL + (size_t)i * 4 * R, *rr = RCL + (size_t)i * 4 * R;
      double *d = dest + (size_t)i * 4 * R;
      for (int r = 0; r < R; r++) {
        const double *lp = Lpr + 16 * r, *rp = Rpr + 16 * r, *lc = l + 4 * r, *rc = rr + 4 * r;
        for (int s = 0; s < 4; s++) {
          double L1, R1;
          if (R == 4) {
            L1 = ((lp[4 * s] * lc[0]) + (lp[4 * s + 1] * lc[1])) + ((lp[4 * s + 2] * lc[2]) + (lp[4 * s + 3] * lc[3]));
            R1 = ((rp[4 * s] * rc[0]) + (rp[4 * s + 1] * rc[1])) + ((rp[4 * s + 2] * rc[2]) + (rp[4 * s + 3] * rc[3]));
          } else {
            L1 = lp[4 * s] * lc[0] + lp[4 * s + 1] * lc[1] + lp[4 * s + 2] * lc[2] + lp[4 * s + 3] * lc[3];
            R1 = rp[4 * s] * rc[0] + rp[4 * s + 1] * rc[1] + rp[4 * s + 2] * rc[2] + rp[4 * s + 3] * rc[3];
          }
          d[4 * r + s] = L1 * R1;
        }
      }
    }
  } else {
    for (int i = 0; i < P; i++)
      for (int r = 0; r < R; r++) {
        const double *lc = LCL + ((size_t)i * R + r) * S, *rc = RCL + ((size_t)i * R + r) * S;
        double *d = dest + ((size_t)i * R + r) * S;
        for (int f = 0; f < S; f++) {
          double L1 = 0.0, R1 = 0.0;
          for (int t = 0; t < S; t++) {
            L1 += Lpr[r * S * S + f * S + t] * lc[t];
            R1 += Rpr[r * S * S + f * S + t] * rc[t];
          }
          d[f] = L1 * R1;
        }
      }
  }
  for (int i = 0; i < P; i++) du[i] = Lu[i] + Ru[i];
}

/* tree.cpp:8291-8535 (nuc) and 8537-8599 (NState); pr1 = internal child's P */
void ora_cla_it(const ora_ctx *c, double *dest, int *du, const double *CL, const int *CLu,
                const double *pr1, const double *pr2, int tip) {
  const int S = c->nstates, R = c->nrates, P = c->npat;
  if (S == 4) {
    const signed char *d2 = c->nuc_tip[tip];
    for (int i = 0; i < P; i++) {
      const double *cl = CL + (size_t)i * 4 * R;
      double *d = dest + (size_t)i * 4 * R;
      if (*d2 > -1) {
        for (int r = 0; r < R; r++)
          for (int s = 0; s < 4; s++) {
            const double *p = pr1 + 16 * r + 4 * s, *x = cl + 4 * r;
            double L1;
            if (R == 4) L1 = ((p[0] * x[0] + p[1] * x[1]) + (p[2] * x[2] + p[3] * x[3]));
            else L1 = (p[0] * x[0] + p[1] * x[1] + p[2] * x[2] + p[3] * x[3]);
            d[4 * r + s] = L1 * pr2[(*d2) + 4 * s + 16 * r];
          }
      } else if (*d2 == -4) {
        for (int r = 0; r < R; r++)
          for (int s = 0; s < 4; s++) {
            const double *p = pr1 + 16 * r + 4 * s, *x = cl + 4 * r;
            d[4 * r + s] = (p[0] * x[0] + p[1] * x[1] + p[2] * x[2] + p[3] * x[3]);
          }
      } else {
        int ns = -*d2;
        const signed char *st = d2 + 1;
        for (int q = 0; q < 4 * R; q++) d[q] = 0.0;
        for (int k = 0; k < ns; k++)
          for (int r = 0; r < R; r++)
            for (int s = 0; s < 4; s++) d[4 * r + s] += pr2[st[k] + 4 * s + 16 * r];
        for (int r = 0; r < R; r++)
          for (int s = 0; s < 4; s++) {
            const double *p = pr1 + 16 * r + 4 * s, *x = cl + 4 * r;
            d[4 * r + s] *= (p[0] * x[0] + p[1] * x[1] + p[2] * x[2] + p[3] * x[3]);
          }
      }
      d2 = ora_adv(d2);
    }
  } else {
    const unsigned char *d2 = c->ns_tip[tip];
    for (int i = 0; i < P; i++)
      for (int r = 0; r < R; r++) {
        const double *cl = CL + ((size_t)i * R + r) * S;
        double *d = dest + ((size_t)i * R + r) * S;
        for (int f = 0; f < S; f++) {
          double x = 0.0;
          for (int t = 0; t < S; t++) x += pr1[r * S * S + f * S + t] * cl[t];
          d[f] = (d2[i] < S ? x * pr2[r * S * S + d2[i] + f * S] : x);
        }
      }
  }
  for (int i = 0; i < P; i++) du[i] = CLu[i];
}

/* tree.cpp:7912-8080 (nuc) and 8082-8160 (NState) */
void ora_cla_tt(const ora_ctx *c, double *dest, int *du, const double *Lpr, const double *Rpr,
                int tipL, int tipR) {
  const int S = c->nstates, R = c->nrates, P = c->npat;
  if (S == 4) {
    const signed char *L = c->nuc_tip[tipL], *Rd = c->nuc_tip[tipR];
    for (int i = 0; i < P; i++) {
      double *d = dest + (size_t)i * 4 * R;
      if (*L > -1 && *Rd > -1) {
        for (int r = 0; r < R; r++)
          for (int s = 0; s < 4; s++) d[4 * r + s] = Lpr[*L + 4 * s + 16 * r] * Rpr[*Rd + 4 * s + 16 * r];
      } else if ((*L == -4 && *Rd == -4) || (*L == -4 && *Rd > -1) || (*Rd == -4 && *L > -1)) {
        for (int r = 0; r < R; r++)
          for (int s = 0; s < 4; s++) {
            if (*L == -4 && *Rd == -4) d[4 * r + s] = 1.0;
            else if (*L == -4) d[4 * r + s] = Rpr[*Rd + 4 * s + 16 * r];
            else d[4 * r + s] = Lpr[*L + 4 * s + 16 * r];
          }
      } else {
        /* partial ambiguity on at least one side */
        if (*L > -1) {
          for (int r = 0; r < R; r++)
            for (int s = 0; s < 4; s++) d[4 * r + s] = Lpr[*L + 4 * s + 16 * r];
        } else if (*L == -4) {
          for (int q = 0; q < 4 * R; q++) d[q] = 1.0;
        } else {
          int ns = -*L;
          for (int q = 0; q < 4 * R; q++) d[q] = 0.0;
          for (int k = 0; k < ns; k++)
            for (int r = 0; r < R; r++)
              for (int s = 0; s < 4; s++) d[4 * r + s] += Lpr[L[1 + k] + 4 * s + 16 * r];
        }
        if (*Rd > -1) {
          for (int r = 0; r < R; r++)
            for (int s = 0; s < 4; s++) d[4 * r + s] *= Rpr[*Rd + 4 * s + 16 * r];
        } else if (*Rd != -4) {
          int ns = -*Rd;
          double tmp[64];
          for (int q = 0; q < 4 * R; q++) tmp[q] = 0.0;
          for (int k = 0; k < ns; k++)
            for (int r = 0; r < R; r++)
              for (int s = 0; s < 4; s++) tmp[4 * r + s] += Rpr[Rd[1 + k] + 4 * s + 16 * r];
          for (int q = 0; q < 4 * R; q++) d[q] *= tmp[q];
        }
      }
      L = ora_adv(L);
      Rd = ora_adv(Rd);
    }
  } else {
    const unsigned char *L = c->ns_tip[tipL], *Rd = c->ns_tip[tipR];
    for (int i = 0; i < P; i++) {
      double *d = dest + (size_t)i * R * S;
      for (int r = 0; r < R; r++)
        for (int f = 0; f < S; f++) {
          double v;
          if (L[i] < S && Rd[i] < S) v = Lpr[L[i] + f * S + r * S * S] * Rpr[Rd[i] + f * S + r * S * S];
          else if (L[i] == S && Rd[i] == S) v = 1.0;
          else if (L[i] == S) v = Rpr[Rd[i] + f * S + r * S * S];
          else v = Lpr[L[i] + f * S + r * S * S];
          d[r * S + f] = v;
        }
    }
  }
  for (int i = 0; i < P; i++) du[i] = 0;
}

/* rescale tables, tree.cpp:159-172 (FP64 branch) */
static double ora_thresh[ORA_RESCALE_LEN], ora_mult[ORA_RESCALE_LEN];
static int ora_incr[ORA_RESCALE_LEN], ora_tables_ready = 0;
static void ora_tables(void) {
  if (ora_tables_ready) return;
  double below = exp(-24.0), bail = 1.0e-250, maxmult = 1.0 / bail;
  for (int i = 0; i < ORA_RESCALE_LEN; i++) {
    ora_incr[i] = i * 7 - (int)log(below);
    ora_thresh[i] = exp((double)(-ora_incr[i]));
    double m = exp((double)ora_incr[i]);
    ora_mult[i] = m < maxmult ? m : maxmult;
  }
  ora_tables_ready = 1;
}

/* tree.cpp:3370-3598 */
int ora_rescale(const ora_ctx *c, double *cla, int *u) {
  ora_tables();
  const int S = c->nstates, R = c->nrates, P = c->npat, W = S * R;
  const double rescale_below = exp(-24.0), reduce_below = 1.0e-190, bail = 1.0e-250;
  for (int i = 0; i < P; i++) {
    double *d = cla + (size_t)i * W, large1;
    if (S == 4) {
      double large2;
      large1 = (d[0] > d[2] ? d[0] : d[2]);
      large2 = (d[1] > d[3] ? d[1] : d[3]);
      large1 = (large1 > large2 ? large1 : large2);
      for (int r = 1; r < R; r++) {
        large2 = (d[0 + r * 4] > d[2 + r * 4] ? d[0 + r * 4] : d[2 + r * 4]);
        large1 = (large1 > large2 ? large1 : large2);
        large2 = (d[1 + r * 4] > d[3 + r * 4] ? d[1 + r * 4] : d[3 + r * 4]);
        large1 = (large1 > large2 ? large1 : large2);
      }
    } else {
      large1 = (d[0] > d[1]) ? d[0] : d[1];
      for (int s = 2; s < W; s++) large1 = (d[s] > large1) ? d[s] : large1;
    }
    if (large1 < rescale_below) {
      if (large1 < reduce_below) {
        if (c->rescale_every > 2) return 1;
        if (large1 < bail) return 2;
      }
      int index = 0;
      while (((index + 1) < ORA_RESCALE_LEN) && (ora_thresh[index + 1] > large1)) index++;
      u[i] += ora_incr[index];
      double mult = ora_mult[index];
      for (int q = 0; q < W; q++) d[q] *= mult;
    }
  }
  return 0;
}

static double ora_btot(const ora_ctx *c, int p) {
  double b = 0.0;
  int m = c->const_states[p];
  for (int s = 0; s < 4; s++)
    if (m & (1 << s)) b += c->pi[s];
  return b;
}

/* tree.cpp:6040-6271 (nuc) and 5597-5838, 6273-6500 (NState; Mk conditioning out of scope) */
double ora_score_edge(const ora_ctx *c, const double *partial, const int *pu, const double *CL1,
                      const int *cu, int child_tip, const double *pr, double *site_lnl) {
  const int S = c->nstates, R = c->nrates, P = c->npat;
  const double *rp = c->rate_probs, *fr = c->pi, prI = c->pinv;
  double tot = 0.0;
  if (S == 4) {
    const signed char *ld = CL1 ? NULL : c->nuc_tip[child_tip];
    for (int i = 0; i < P; i++) {
      const double *pa = partial + (size_t)i * 4 * R;
      double La = 0, Lc = 0, Lg = 0, Lt = 0;
      int u = pu[i] + (CL1 ? cu[i] : 0);
      if (CL1) {
        const double *cl = CL1 + (size_t)i * 4 * R;
        for (int r = 0; r < R; r++) {
          const double *p = pr + 16 * r, *x = cl + 4 * r, *q = pa + 4 * r;
          La += (p[0] * x[0] + p[1] * x[1] + p[2] * x[2] + p[3] * x[3]) * q[0] * rp[r];
          Lc += (p[4] * x[0] + p[5] * x[1] + p[6] * x[2] + p[7] * x[3]) * q[1] * rp[r];
          Lg += (p[8] * x[0] + p[9] * x[1] + p[10] * x[2] + p[11] * x[3]) * q[2] * rp[r];
          Lt += (p[12] * x[0] + p[13] * x[1] + p[14] * x[2] + p[15] * x[3]) * q[3] * rp[r];
        }
      } else if (*ld > -1) {
        for (int r = 0; r < R; r++) {
          const double *q = pa + 4 * r;
          La += pr[(*ld) + 16 * r] * q[0] * rp[r];
          Lc += pr[(*ld + 4) + 16 * r] * q[1] * rp[r];
          Lg += pr[(*ld + 8) + 16 * r] * q[2] * rp[r];
          Lt += pr[(*ld + 12) + 16 * r] * q[3] * rp[r];
        }
      } else if (*ld == -4) {
        for (int r = 0; r < R; r++) {
          const double *q = pa + 4 * r;
          La += q[0] * rp[r];
          Lc += q[1] * rp[r];
          Lg += q[2] * rp[r];
          Lt += q[3] * rp[r];
        }
      } else {
        int ns = -*ld;
        for (int k = 0; k < ns; k++) {
          int st = ld[1 + k];
          for (int r = 0; r < R; r++) {
            La += pr[st + 16 * r] * pa[4 * r] * rp[r];
            Lc += pr[(st + 4) + 16 * r] * pa[1 + 4 * r] * rp[r];
            Lg += pr[(st + 8) + 16 * r] * pa[2 + 4 * r] * rp[r];
            Lt += pr[(st + 12) + 16 * r] * pa[3 + 4 * r] * rp[r];
          }
        }
      }
      double siteL;
      if (c->pinv_in_model && i <= c->last_const) {
        double btot = ora_btot(c, i);
        if (u == 0) siteL = ((La * fr[0] + Lc * fr[1] + Lg * fr[2] + Lt * fr[3]) + prI * btot);
        else siteL = ((La * fr[0] + Lc * fr[1] + Lg * fr[2] + Lt * fr[3]) + (prI * btot * exp((double)u)));
      } else
        siteL = ((La * fr[0] + Lc * fr[1] + Lg * fr[2] + Lt * fr[3]));
      double ul = CL1 ? (log(siteL) - pu[i] - cu[i]) : (log(siteL) - pu[i]);
      tot += (c->counts[i] * ul);
      if (site_lnl) site_lnl[i] = ul;
      if (ld) ld = ora_adv(ld);
    }
  } else {
    const unsigned char *ld = CL1 ? NULL : c->ns_tip[child_tip];
    for (int i = 0; i < P; i++) {
      double siteL = 0.0;
      int u = pu[i] + (CL1 ? cu[i] : 0);
      for (int r = 0; r < R; r++) {
        const double *pa = partial + ((size_t)i * R + r) * S;
        const double *p = pr + r * S * S;
        double rateL = 0.0;
        if (CL1) {
          const double *cl = CL1 + ((size_t)i * R + r) * S;
          for (int f = 0; f < S; f++) {
            double tmp = 0.0;
            for (int t = 0; t < S; t++) tmp += p[f * S + t] * cl[t];
            rateL += tmp * pa[f] * fr[f];
          }
        } else if (ld[i] < S) {
          for (int f = 0; f < S; f++) rateL += p[f * S + ld[i]] * pa[f] * fr[f];
        } else {
          for (int f = 0; f < S; f++) rateL += pa[f] * fr[f];
        }
        siteL += rateL * c->rate_probs[r];
      }
      if (c->pinv_in_model && i <= c->last_const) {
        if (u == 0) siteL += prI * fr[c->const_states[i]];
        else siteL += prI * fr[c->const_states[i]] * exp((double)u);
      }
      double ul = log(siteL) - u;
      if (ul < 0.0) tot += c->counts[i] * ul;
      if (site_lnl) site_lnl[i] = ul;
    }
  }
  return tot;
}

/* optimization.cpp:2412-2590, 3063-3228 (nuc) and 2592-3062, 3230-3613 (NState) */
void ora_derivs_edge(const ora_ctx *c, const double *partial, const int *pu, const double *CL1,
                     const int *cu, int child_tip, const double *pr, const double *d1m,
                     const double *d2m, double *lnl_out, double *d1_out, double *d2_out) {
  const int S = c->nstates, R = c->nrates, P = c->npat;
  const double *rp = c->rate_probs, *fr = c->pi, prI = c->pinv;
  double tot1 = 0, tot2 = 0, totL = 0;
  if (S == 4) {
    const signed char *ld = CL1 ? NULL : c->nuc_tip[child_tip];
    for (int i = 0; i < P; i++) {
      const double *pa = partial + (size_t)i * 4 * R;
      double L[4] = {0, 0, 0, 0}, D1[4] = {0, 0, 0, 0}, D2[4] = {0, 0, 0, 0};
      if (CL1) {
        const double *cl = CL1 + (size_t)i * 4 * R;
        for (int r = 0; r < R; r++) {
          const double *x = cl + 4 * r, *q = pa + 4 * r;
          for (int s = 0; s < 4; s++) {
            const double *p = pr + 16 * r + 4 * s, *a = d1m + 16 * r + 4 * s, *b = d2m + 16 * r + 4 * s;
            L[s] += (p[0] * x[0] + p[1] * x[1] + p[2] * x[2] + p[3] * x[3]) * q[s] * rp[r];
            D1[s] += (a[0] * x[0] + a[1] * x[1] + a[2] * x[2] + a[3] * x[3]) * q[s] * rp[r];
            D2[s] += (b[0] * x[0] + b[1] * x[1] + b[2] * x[2] + b[3] * x[3]) * q[s] * rp[r];
          }
        }
      } else if (*ld > -1) {
        for (int r = 0; r < R; r++)
          for (int s = 0; s < 4; s++) {
            L[s] += pr[(*ld) + 4 * s + 16 * r] * pa[4 * r + s] * rp[r];
            D1[s] += d1m[(*ld) + 4 * s + 16 * r] * pa[4 * r + s] * rp[r];
            D2[s] += d2m[(*ld) + 4 * s + 16 * r] * pa[4 * r + s] * rp[r];
          }
      } else if (*ld == -4) {
        for (int r = 0; r < R; r++)
          for (int s = 0; s < 4; s++) L[s] += pa[4 * r + s] * rp[r];
      } else {
        int ns = -*ld;
        for (int k = 0; k < ns; k++) {
          int st = ld[1 + k];
          for (int r = 0; r < R; r++)
            for (int s = 0; s < 4; s++) {
              L[s] += pr[st + 4 * s + 16 * r] * pa[s + 4 * r] * rp[r];
              D1[s] += d1m[st + 4 * s + 16 * r] * pa[s + 4 * r] * rp[r];
              D2[s] += d2m[st + 4 * s + 16 * r] * pa[s + 4 * r] * rp[r];
            }
        }
      }
      double siteL, ucomb = (double)pu[i] + (CL1 ? (double)cu[i] : 0.0);
      if (c->pinv_in_model && i <= c->last_const) {
        double btot = ora_btot(c, i);
        if (CL1) siteL = ((L[0] * fr[0] + L[1] * fr[1] + L[2] * fr[2] + L[3] * fr[3]) + (prI * btot) * exp((double)cu[i] + pu[i]));
        else siteL = ((L[0] * fr[0] + L[1] * fr[1] + L[2] * fr[2] + L[3] * fr[3]) + (prI * btot) * exp((double)pu[i]));
      } else
        siteL = ((L[0] * fr[0] + L[1] * fr[1] + L[2] * fr[2] + L[3] * fr[3]));
      double sD1 = ((D1[0] * fr[0] + D1[1] * fr[1] + D1[2] * fr[2] + D1[3] * fr[3])) / siteL;
      double ul = log(siteL) - ucomb;
      if (CL1) ul = log(siteL) - cu[i] - pu[i];
      totL += ul * c->counts[i];
      tot1 += c->counts[i] * sD1;
      double sD2 = ((D2[0] * fr[0] + D2[1] * fr[1] + D2[2] * fr[2] + D2[3] * fr[3]) / siteL) - (sD1 * sD1);
      tot2 += c->counts[i] * sD2;
      if (ld) ld = ora_adv(ld);
    }
  } else {
    const unsigned char *ld = CL1 ? NULL : c->ns_tip[child_tip];
    for (int i = 0; i < P; i++) {
      double siteL = 0, sD1 = 0, sD2 = 0;
      for (int r = 0; r < R; r++) {
        const double *pa = partial + ((size_t)i * R + r) * S;
        const int ro = r * S * S;
        double rL = 0, r1 = 0, r2 = 0;
        if (CL1) {
          const double *cl = CL1 + ((size_t)i * R + r) * S;
          for (int f = 0; f < S; f++) {
            double tL = 0, t1 = 0, t2 = 0;
            for (int t = 0; t < S; t++) {
              tL += pr[ro + f * S + t] * cl[t];
              t1 += d1m[ro + f * S + t] * cl[t];
              t2 += d2m[ro + f * S + t] * cl[t];
            }
            rL += tL * pa[f] * fr[f];
            r1 += t1 * pa[f] * fr[f];
            r2 += t2 * pa[f] * fr[f];
          }
        } else if (ld[i] < S) {
          for (int f = 0; f < S; f++) {
            rL += pr[ro + f * S + ld[i]] * pa[f] * fr[f];
            r1 += d1m[ro + f * S + ld[i]] * pa[f] * fr[f];
            r2 += d2m[ro + f * S + ld[i]] * pa[f] * fr[f];
          }
        } else {
          for (int f = 0; f < S; f++) rL += pa[f] * fr[f];
        }
        siteL += rL * rp[r];
        sD1 += r1 * rp[r];
        sD2 += r2 * rp[r];
      }
      if (c->pinv_in_model && i <= c->last_const) {
        if (CL1) siteL += (prI * fr[c->const_states[i]] * exp((double)pu[i]) * exp((double)cu[i]));
        else siteL += (prI * fr[c->const_states[i]] * (exp((double)pu[i])));
      }
      double ul = log(siteL) - pu[i] - (CL1 ? cu[i] : 0);
      if (ul < 0.0) {
        totL += ul * c->counts[i];
        sD1 /= siteL;
        tot1 += c->counts[i] * sD1;
        sD2 = ((sD2 / siteL) - sD1 * sD1);
        tot2 += c->counts[i] * sD2;
      }
    }
  }
  *d1_out = tot1;
  *d2_out = tot2;
  *lnl_out = totL;
}

/* ======================================================================
 * Tree level: directional CLAs with memoisation and dirtiness
 * ==================================================================== */

typedef struct {
  const ora_ctx *c;
  const ora_tree *t;
  int nnodes, N;
  int nb[1024][3], nnb[1024]; /* neighbours: anc first (if any), then children ascending */
  double *cla[1024][3];
  int *u[1024][3];
  int rank[1024][3];
  int valid[1024][3];
  int fail;
} ora_memo;

static int ora_is_tip(const ora_memo *m, int n) { return n >= 1 && n <= m->N; }

static void ora_memo_init(ora_memo *m, const ora_ctx *c, const ora_tree *t) {
  memset(m, 0, sizeof *m);
  m->c = c;
  m->t = t;
  m->N = t->ntax;
  m->nnodes = 2 * t->ntax - 2;
  for (int n = 0; n < m->nnodes; n++) {
    if (t->parent[n] >= 0) m->nb[n][m->nnb[n]++] = t->parent[n];
  }
  for (int n = 0; n < m->nnodes; n++)
    if (t->parent[n] >= 0) {
      int a = t->parent[n];
      m->nb[a][m->nnb[a]++] = n;
    }
}
static void ora_memo_free(ora_memo *m) {
  for (int n = 0; n < m->nnodes; n++)
    for (int k = 0; k < 3; k++) {
      free(m->cla[n][k]);
      free(m->u[n][k]);
    }
}
static int ora_nb_index(const ora_memo *m, int a, int b) {
  for (int k = 0; k < m->nnb[a]; k++)
    if (m->nb[a][k] == b) return k;
  return -1;
}
static double ora_edge_len(const ora_memo *m, int a, int b) {
  return m->t->parent[a] == b ? m->t->blen[a] : m->t->blen[b];
}

/* X(a | b): tree.cpp:3600-3684 (direction logic) + 3905-4001 (UpdateCLAs) */
static int ora_get(ora_memo *m, int a, int b) {
  int k = ora_nb_index(m, a, b);
  if (m->valid[a][k]) return k;
  const ora_ctx *c = m->c;
  int x = -1, y = -1;
  for (int j = 0; j < m->nnb[a]; j++) {
    int nbr = m->nb[a][j];
    if (nbr == b) continue;
    if (x < 0) x = nbr; else y = nbr;
  }
  size_t W = (size_t)c->npat * c->nrates * c->nstates;
  if (!m->cla[a][k]) {
    m->cla[a][k] = (double *)malloc(sizeof(double) * W);
    m->u[a][k] = (int *)malloc(sizeof(int) * c->npat);
  }
  double *P1 = (double *)malloc(sizeof(double) * c->nrates * c->nstates * c->nstates);
  double *P2 = (double *)malloc(sizeof(double) * c->nrates * c->nstates * c->nstates);
  ora_pmats(c, ora_edge_len(m, a, x), P1, NULL, NULL);
  ora_pmats(c, ora_edge_len(m, a, y), P2, NULL, NULL);
  int xt = ora_is_tip(m, x), yt = ora_is_tip(m, y), rank;
  if (!xt && !yt) {
    int kx = ora_get(m, x, a), ky = ora_get(m, y, a);
    ora_cla_ii(c, m->cla[a][k], m->u[a][k], m->cla[x][kx], m->u[x][kx], m->cla[y][ky], m->u[y][ky], P1, P2);
    rank = 2 + m->rank[x][kx] + m->rank[y][ky];
  } else if (xt && yt) {
    ora_cla_tt(c, m->cla[a][k], m->u[a][k], P1, P2, x - 1, y - 1);
    rank = 2;
  } else if (xt) {
    int ky = ora_get(m, y, a);
    ora_cla_it(c, m->cla[a][k], m->u[a][k], m->cla[y][ky], m->u[y][ky], P2, P1, x - 1);
    rank = m->rank[y][ky] + 2;
  } else {
    int kx = ora_get(m, x, a);
    ora_cla_it(c, m->cla[a][k], m->u[a][k], m->cla[x][kx], m->u[x][kx], P1, P2, y - 1);
    rank = m->rank[x][kx] + 2;
  }
  free(P1);
  free(P2);
  if (rank >= c->rescale_every) {
    int rc = ora_rescale(c, m->cla[a][k], m->u[a][k]);
    if (rc) m->fail = rc;
    rank = 0;
  }
  m->rank[a][k] = rank;
  m->valid[a][k] = 1;
  return k;
}

/* SweepDirtynessOverTree: invalidate every X(a|b) whose subtree holds edge (n, parent(n)) */
static void ora_dirty_from(ora_memo *m, int from, int came) {
  for (int j = 0; j < m->nnb[from]; j++) {
    int nbr = m->nb[from][j];
    if (nbr == came) continue;
    if (!ora_is_tip(m, from)) {
      int k = ora_nb_index(m, from, nbr);
      m->valid[from][k] = 0;
    }
    ora_dirty_from(m, nbr, from);
  }
}
static void ora_dirty_edge(ora_memo *m, int n) {
  int a = m->t->parent[n];
  ora_dirty_from(m, n, a);
  ora_dirty_from(m, a, n);
}
static void ora_dirty_all(ora_memo *m) { memset(m->valid, 0, sizeof m->valid); }

static int ora_memo_score(ora_memo *m, double *lnl, double *site) {
  const ora_ctx *c = m->c;
  int root = 0, middle = m->nb[0][1];
  int kp = ora_get(m, root, middle);
  const double *ccl = NULL;
  const int *cu = NULL;
  if (!ora_is_tip(m, middle)) {
    int kc = ora_get(m, middle, root);
    ccl = m->cla[middle][kc];
    cu = m->u[middle][kc];
  }
  if (m->fail) return m->fail;
  double *P = (double *)malloc(sizeof(double) * c->nrates * c->nstates * c->nstates);
  ora_pmats(c, m->t->blen[middle], P, NULL, NULL);
  *lnl = ora_score_edge(c, m->cla[root][kp], m->u[root][kp], ccl, cu, middle - 1, P, site);
  free(P);
  return 0;
}

static int ora_memo_derivs(ora_memo *m, int nd, double *lnl, double *d1, double *d2) {
  const ora_ctx *c = m->c;
  int anc = m->t->parent[nd];
  int kp = ora_get(m, anc, nd);
  const double *ccl = NULL;
  const int *cu = NULL;
  if (!ora_is_tip(m, nd)) {
    int kc = ora_get(m, nd, anc);
    ccl = m->cla[nd][kc];
    cu = m->u[nd][kc];
  }
  if (m->fail) return m->fail;
  size_t SS = (size_t)c->nrates * c->nstates * c->nstates;
  double *P = (double *)malloc(sizeof(double) * SS * 3);
  ora_pmats(c, m->t->blen[nd], P, P + SS, P + 2 * SS);
  ora_derivs_edge(c, m->cla[anc][kp], m->u[anc][kp], ccl, cu, nd - 1, P, P + SS, P + 2 * SS, lnl, d1, d2);
  free(P);
  return 0;
}

/* Score with the reference's retry (tree.cpp:4077-4090) */
int ora_tree_score(const ora_ctx *c0, const ora_tree *t, int *re_io, double *lnl, double *site) {
  ora_ctx c = *c0;
  c.rescale_every = *re_io;
  ora_memo *m = (ora_memo *)malloc(sizeof(ora_memo));
  int rc;
  for (;;) {
    ora_memo_init(m, &c, t);
    rc = ora_memo_score(m, lnl, site);
    ora_memo_free(m);
    if (rc == 0) break;
    c.rescale_every -= 2;
    if (c.rescale_every < 2) break;
  }
  *re_io = c.rescale_every;
  free(m);
  return rc;
}

int ora_tree_derivs(const ora_ctx *c0, const ora_tree *t, int node, int *re_io, double *lnl,
                    double *d1, double *d2) {
  ora_ctx c = *c0;
  c.rescale_every = *re_io;
  ora_memo *m = (ora_memo *)malloc(sizeof(ora_memo));
  int rc;
  for (;;) {
    ora_memo_init(m, &c, t);
    rc = ora_memo_derivs(m, node, lnl, d1, d2);
    ora_memo_free(m);
    if (rc == 0) break;
    c.rescale_every -= 2;
    if (c.rescale_every < 2) break;
  }
  *re_io = c.rescale_every;
  free(m);
  return rc;
}

/* every directional CLA once (memoised), then (lnL, d1, d2) on every edge */
int ora_tree_sweep(const ora_ctx *c0, const ora_tree *t, int *re_io, double *out) {
  ora_ctx c = *c0;
  c.rescale_every = *re_io;
  ora_memo *m = (ora_memo *)malloc(sizeof(ora_memo));
  int rc = 0;
  for (;;) {
    ora_memo_init(m, &c, t);
    for (int n = 1; n < m->nnodes && !rc; n++)
      rc = ora_memo_derivs(m, n, out + 3 * (n - 1), out + 3 * (n - 1) + 1, out + 3 * (n - 1) + 2);
    ora_memo_free(m);
    if (rc == 0) break;
    rc = 0;
    c.rescale_every -= 2;
    if (c.rescale_every < 2) { rc = 1; break; }
  }
  *re_io = c.rescale_every;
  free(m);
  return rc;
}

int ora_tree_cla(const ora_ctx *c, const ora_tree *t, int a, int b, double *out, int *u) {
  ora_memo *m = (ora_memo *)malloc(sizeof(ora_memo));
  ora_memo_init(m, c, t);
  int k = ora_nb_index(m, a, b);
  if (k < 0 || ora_is_tip(m, a)) { free(m); return -1; }
  k = ora_get(m, a, b);
  size_t W = (size_t)c->npat * c->nrates * c->nstates;
  memcpy(out, m->cla[a][k], sizeof(double) * W);
  memcpy(u, m->u[a][k], sizeof(int) * c->npat);
  int rc = m->fail;
  ora_memo_free(m);
  free(m);
  return rc;
}

/* optimization.cpp:1359-1972, double-precision branch without debug output.
 * Returns totalEstImprove; the tree's lnL after the call is left in *lnl_io. */
static double ora_newton(ora_memo *m, int nd, double prec, double *lnl_io, int *fail) {
  const ora_ctx *c = m->c;
  const double min_b = c->min_brlen, max_b = c->max_brlen;
  double totalEstImprove = 0.0, v_onEntry = m->t->blen[nd], v = v_onEntry, v_prev = v;
  double lnL = *lnl_io, initialL = 0, d1, d2, abs_d1_prev = 3.4e38;
  double knownMin = min_b, knownMax = max_b;
  int iter = 0, negProposalNum = 0;
  (void)abs_d1_prev;
  for (;;) {
    if (ora_memo_derivs(m, nd, &lnL, &d1, &d2)) { *fail = 1; return totalEstImprove; }
    if (iter == 0) initialL = lnL;
    double dlen = m->t->blen[nd];
    double estDeltaNR = -d1 / d2;
    double estScoreDelta = d1 * estDeltaNR + (d2 * estDeltaNR * estDeltaNR * 0.5);
    if (d1 <= 0.0 && dlen < knownMax) knownMax = dlen;
    else if (d1 > 0.0 && dlen > knownMin) knownMin = dlen;
    double abs_d1 = fabs(d1);
    if (d2 >= 0.0) {
      if (d1 <= 0.0) {
        double proposed;
        if (fabs(dlen - min_b) < 1.0e-8) { *lnl_io = lnL; return totalEstImprove; }
        if (knownMin == min_b) {
          if (dlen <= 1.0e-4) proposed = min_b;
          else if (dlen <= 0.005) proposed = 1.0e-4;
          else if (dlen <= 0.05) proposed = dlen * 0.1;
          else proposed = dlen * 0.25;
        } else
          proposed = (knownMin + dlen) * 0.5;
        if (iter > 0 || proposed == min_b) {
          double estImp = d1 * (proposed - dlen) + (d2 * (proposed - dlen) * (proposed - dlen) * 0.5);
          if (estImp < prec && lnL >= initialL - 1.0e-4) { *lnl_io = lnL; return totalEstImprove; }
        }
        v = proposed;
        totalEstImprove += prec;
      } else {
        double proposed;
        if (knownMax == max_b) {
          if (dlen < 0.1) proposed = dlen * 10.0;
          else proposed = fmin(dlen * 2.0, max_b);
        } else
          proposed = (knownMax + dlen) * 0.5;
        if (iter > 0) {
          double estImp = d1 * (proposed - dlen) + (d2 * (proposed - dlen) * (proposed - dlen) * 0.5);
          if (estImp < prec && lnL >= initialL - 1.0e-4) { *lnl_io = lnL; return totalEstImprove; }
        }
        v = proposed;
        totalEstImprove += prec;
      }
    } else {
      if (d1 < 0.0 && fabs(dlen - min_b) < 1.0e-8) { *lnl_io = lnL; return totalEstImprove; }
      if (d1 > 0.0 && fabs(dlen - max_b) < 1.0e-8) { *lnl_io = lnL; return totalEstImprove; }
      if (estScoreDelta < prec && (iter == 0 || lnL + ((iter < 10 ? 1 : iter) * 1.0e-7) >= initialL)) {
        *lnl_io = lnL;
        return totalEstImprove;
      }
      int noNR = 0;
      if (iter > 10) {
        if (d1 > 0.0) {
          if (fabs(knownMax - max_b) < 1e-8) {
            if ((iter > 20) && (dlen > (max_b / 2.0))) v = max_b;
            else if (v < 0.2) v = fmin((v + max_b) * 0.5, v * 5.0);
            else v = fmin((v + max_b) * 0.5, v * 2.0);
            noNR = 1;
          } else if (((iter - 20) > 0) && ((iter - 20) % 10 == 0))
            v = fmin((v + knownMax) * 0.5, v * 100.0);
        } else {
          if (fabs(knownMin - min_b) < 1e-8) {
            if (iter > 20) v = min_b;
            else v = (v + min_b) * 0.5;
            noNR = 1;
          }
        }
      }
      if (!noNR) v += estDeltaNR;
      if (v <= knownMin) {
        negProposalNum++;
        if (knownMin == min_b) {
          double dm = min_b - dlen, sdm = (dm * d1 + (dm * dm * d2 * 0.5));
          if (sdm < prec) { *lnl_io = lnL; return totalEstImprove; }
          else if (negProposalNum == 1 && dlen > 1e-4 && v_prev != 1e-4) {
            if (dlen < .005) v = 1e-4;
            else if (dlen < 0.05) v = dlen * 0.1;
            else v = dlen * 0.25;
            double delta = v - dlen;
            totalEstImprove += (delta * d1 + (delta * delta * d2 * 0.5));
          } else {
            v = min_b;
            totalEstImprove += sdm;
          }
        } else {
          double proposed = (knownMin + dlen) * 0.5, dm = proposed - dlen;
          double sdm = (dm * d1 + (dm * dm * d2 * 0.5));
          if (sdm < prec) { *lnl_io = lnL; return totalEstImprove; }
          v = proposed;
          totalEstImprove += sdm;
        }
      } else if (v >= knownMax) {
        if (knownMax == max_b) {
          double dm = max_b - dlen, sdm = (dm * d1 + (dm * dm * d2 * 0.5));
          if (sdm < prec) { *lnl_io = lnL; return totalEstImprove; }
          v = max_b;
          totalEstImprove += sdm;
        } else {
          double proposed = (knownMax + dlen) * 0.5, dm = proposed - dlen;
          double sdm = (dm * d1 + (dm * dm * d2 * 0.5));
          if (sdm < prec) { *lnl_io = lnL; return totalEstImprove; }
          v = proposed;
          totalEstImprove += sdm;
        }
      } else
        totalEstImprove += estScoreDelta;
      abs_d1_prev = abs_d1;
    }
    /* SetBranchLength (tree.h:425) + SweepDirtynessOverTree */
    m->t->blen[nd] = v;
    ora_dirty_edge(m, nd);
    v_prev = v;
    iter++;
    if (iter > 100) { *lnl_io = lnL; return totalEstImprove; }
  }
}

static void ora_preorder(ora_memo *m, int n, int *order, int *k) {
  order[(*k)++] = n;
  for (int j = 0; j < m->nnb[n]; j++) {
    int ch = m->nb[n][j];
    if (ch == m->t->parent[n]) continue;
    ora_preorder(m, ch, order, k);
  }
}

/* optimization.cpp:97-105 with RecursivelyOptimizeBranches (:1042-1060): preorder */
double ora_tree_optimize_all(const ora_ctx *c0, ora_tree *t, double prec, int *re_io, double *lnl) {
  ora_ctx c = *c0;
  c.rescale_every = *re_io;
  ora_memo *m = (ora_memo *)malloc(sizeof(ora_memo));
  ora_memo_init(m, &c, t);
  int order[1024], k = 0, fail = 0;
  for (int j = 0; j < m->nnb[0]; j++) ora_preorder(m, m->nb[0][j], order, &k);
  double cur = 0.0, improve = 0.0;
  ora_memo_score(m, &cur, NULL);
  for (int i = 0; i < k && !fail; i++) improve += ora_newton(m, order[i], prec, &cur, &fail);
  ora_dirty_all(m);
  ora_memo_score(m, lnl, NULL);
  ora_memo_free(m);
  free(m);
  return improve;
}
