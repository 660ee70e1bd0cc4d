/*
 * gle_oracle.h -- FP64 CPU oracle for the GARLI tree-likelihood path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped engine (garli_amd/) may
 * link, load or call this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, always as the checker.
 *
 * It restates, in plain C, the reference's FP64 likelihood arithmetic:
 *   Q builds          trunk/src/model.cpp:233-294 (GTR), 296-440 (codon lookup),
 *                     442-687 (GY94/M3), 735-780 (AA)
 *   eigen -> P(t)     trunk/src/model.cpp:827-856, 1271-1358, 1469-1597
 *   discrete gamma    trunk/src/model.cpp:1883-2201 (AS 91 / AS 111 / AS 32 /
 *                     ACM 291, the published algorithms the reference uses)
 *   CLA kernels       trunk/src/tree.cpp:7668-8160 (II, TT), 8291-8599 (IT)
 *   rescaling         trunk/src/tree.cpp:139-173 (tables), 3370-3598
 *   edge scoring      trunk/src/tree.cpp:6040-6500
 *   derivatives       trunk/src/optimization.cpp:1099-1194, 2412-3613
 *   Newton branch opt trunk/src/optimization.cpp:97-105, 1042-1060, 1359-1972
 *   tip encoding      trunk/src/sequencedata.cpp:278-362, 672-780
 *   pattern packing   trunk/src/datamatr.cpp:51-375
 *
 * Layouts follow the reference: CLA arr[(p*R + r)*S + s] with int u[p];
 * P matrices pr[r*S*S + from*S + to] (row = parent state).
 */
#ifndef GLE_ORACLE_H
#define GLE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- substitution model ---------------- */
void ora_qmat_gtr(const double pi[4], const double rel[5], double *Q, double *blen_mult);
void ora_qmat_aa(const double pi[20], const double *rel190, double *Q, double *blen_mult);
void ora_codon_lookup(int *lookup61x61, int *aa_of_codon61);
void ora_qmat_codon(const double *pi61, const double *rel6, const double *omega, const double *omega_prob,
                    int n_omega, double *Q /* n_omega*61*61 */, double *blen_mult);
/* Reversible eigensystem of Q (symmetrised by pi, cyclic Jacobi).  Returns 0 on success. */
int ora_eigen(int n, const double *Q, const double *pi, double *V, double *Vinv, double *lam);
/* P(t) (clamped >= 0), and optionally dP/dt, d2P/dt2 (unclamped), for one rate:
 * P_ij = sum_k V_ik Vinv_kj exp(lam_k * mult * t), lam already scaled by blen_mult. */
void ora_pmat(int n, const double *V, const double *Vinv, const double *lam, double mult, double t,
              double *P, double *D1, double *D2);
double ora_lngamma(double x);
double ora_incomplete_gamma(double x, double alpha, double lngamma_alpha);
double ora_point_normal(double p);
double ora_point_chi2(double p, double v);
int ora_discrete_gamma(int ncat, double alpha, double *rates);

/* ---------------- data encoding ---------------- */
int ora_nuc_char_to_datum(char c);      /* 1..15, or -1 for an unknown character */
int ora_aa_char_to_datum(char c);       /* 0..20 (20 = missing), -1 unknown */
/* Variable-length signed-char ambiguity stream for one taxon (MakeAmbigStrings).
 * Returns stream length; out must hold 5*n bytes. */
int ora_make_ambig_string(const unsigned char *masks, int n, signed char *out);
/* Pattern compression (ProcessPatterns): matrix[tax*nsite + site] of datums
 * (nuc: 4-bit masks, 15 missing; other: state or max_states missing).
 * Writes out[tax*P + p], counts[p], const_states[p] (only p<=last_const meaningful),
 * site_to_pattern[site] (-1 for all-missing sites).  Returns P (or <0 on error). */
int ora_compress_patterns(int ntax, int nsite, int max_states, const unsigned char *matrix,
                          unsigned char *out, int *counts, int *const_states, int *last_const,
                          int *site_to_pattern);

/* ---------------- likelihood context ---------------- */
typedef struct ora_ctx {
  int ntax, npat, nstates, nrates;
  const double *pi;            /* [S] */
  const double *rate_probs;    /* [R] category weights (1-pinv)/R, or omega probs */
  double pinv;
  int pinv_in_model;
  const int *counts;           /* [P] */
  int last_const;
  const int *const_states;     /* [last_const+1]: nuc bitmask, else state index */
  int n_eigen;
  const double *V, *Vinv, *lam;/* [n_eigen][S*S], [n_eigen][S*S], [n_eigen][S] */
  const int *eig_of_rate;      /* [R] */
  const double *rate_mult;     /* [R] multiplier on lam (rateMult/(1-pinv) or 1) */
  const signed char *const *nuc_tip;   /* [ntax] ambiguity streams (S == 4) */
  const unsigned char *const *ns_tip;  /* [ntax] per-pattern states (S != 4) */
  int rescale_every;           /* 16 in the FP64 reference */
  double min_brlen, max_brlen; /* 1e-8, 100 */
} ora_ctx;

/* P, P', P'' for every rate: out[r*S*S + ...]. D1/D2 may be NULL. */
void ora_pmats(const ora_ctx *c, double t, double *P, double *D1, double *D2);

/* one-CLA kernels; tip arguments follow the reference's argument order */
void ora_cla_ii(const ora_ctx *c, double *dest, int *du, const double *L, const int *Lu,
                const double *R, const int *Ru, const double *PL, const double *PR);
void ora_cla_it(const ora_ctx *c, double *dest, int *du, const double *CL, const int *CLu,
                const double *Pint, const double *Ptip, int tip);
void ora_cla_tt(const ora_ctx *c, double *dest, int *du, const double *PL, const double *PR,
                int tipL, int tipR);
/* Rescale in place; returns 0, or 1 when some site fell below reduceRescaleBelow
 * (the reference's throw(1) when rescaleEvery > 2), 2 below bailOutBelow. */
int ora_rescale(const ora_ctx *c, double *cla, int *u);
/* edge integration: child is a CLA (child_cla != NULL) or tip index child_tip.
 * Returns lnL; site_lnl (optional) receives per-pattern unscaled lnL. */
double ora_score_edge(const ora_ctx *c, const double *partial, const int *pu,
                      const double *child_cla, const int *cu, int child_tip,
                      const double *P, double *site_lnl);
void ora_derivs_edge(const ora_ctx *c, const double *partial, const int *pu,
                     const double *child_cla, const int *cu, int child_tip,
                     const double *P, const double *D1, const double *D2,
                     double *lnl, double *d1, double *d2);

/* ---------------- tree level ----------------
 * Unrooted tree with nodes 0..2N-3: 0 = root (3 children), 1..N = tips,
 * N+1..2N-3 = internal.  parent[0] = -1; blen[n] = length of n's edge to parent.
 * Children are taken in ascending node order (left, [middle], right). */
typedef struct ora_tree {
  int ntax;
  const int *parent;
  double *blen;
} ora_tree;

/* Tree::Score(0): returns 0 and lnL, or nonzero on unrecoverable rescale failure.
 * rescale_every_io is lowered by 2 per retry, as in the reference. */
int ora_tree_score(const ora_ctx *c, const ora_tree *t, int *rescale_every_io, double *lnl,
                   double *site_lnl);
/* Tree::CalcDerivativesRateHet(nd->anc, nd) */
int ora_tree_derivs(const ora_ctx *c, const ora_tree *t, int node, int *rescale_every_io,
                    double *lnl, double *d1, double *d2);
/* all directional CLAs, then out[3*(n-1)] = (lnL, d1, d2) for the edge of node n */
int ora_tree_sweep(const ora_ctx *c, const ora_tree *t, int *rescale_every_io, double *out);
/* directional CLA X(a | b) for internal a and neighbour b, reference layout */
int ora_tree_cla(const ora_ctx *c, const ora_tree *t, int a, int b, double *out, int *u);
/* Tree::OptimizeAllBranches(prec): one Newton pass over all edges (preorder);
 * updates t->blen; returns total estimated improvement; *lnl = final Score */
double ora_tree_optimize_all(const ora_ctx *c, ora_tree *t, double prec, int *rescale_every_io,
                             double *lnl);

#ifdef __cplusplus
}
#endif
#endif
