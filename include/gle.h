/*
 * gle.h -- C ABI of the MI355X GARLI likelihood engine (libgle.so).
 *
 * The engine replaces GARLI's likelihood back ends (the native CPU kernels in
 * trunk/src/tree.cpp and optimization.cpp, and the BEAGLE seam in
 * trunk/src/calculationmanager.cpp / subsetcalculationmanager.cpp).  Every
 * entry point takes plain pointers and sizes; results are FP64.
 *
 * Two layers:
 *   1. instance layer: op lists in the shape of GARLI's CalculationManager
 *      (ClaOperation / TransMatOperation / ScoringOperation,
 *      trunk/src/calculationmanager.h:39-123);
 *   2. tree layer: GARLI's tree-scoring surface (Tree::Score,
 *      Tree::CalcDerivativesRateHet, Tree::SetBranchLength,
 *      Tree::OptimizeAllBranches), with ClaManager-style CLA caching.
 * A serving layer batches scoring requests from many trees.
 *
 * Return codes: GLE_SUCCESS (0), GLE_RESCALE_RETRY (1, the reference's
 * throw(1) from RescaleRateHet, trunk/src/tree.cpp:3436-3449), negative values
 * are errors; gle_last_error() describes the last failure on this thread.
 */
#ifndef GLE_H
#define GLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define GLE_SUCCESS 0
#define GLE_RESCALE_RETRY 1
#define GLE_ERR_ARG (-1)
#define GLE_ERR_DEVICE (-2)
#define GLE_ERR_NOMEM (-3)
#define GLE_ERR_RESCALE_FATAL (-4)
#define GLE_ERR_UNSUPPORTED (-5)
#define GLE_ERR_STATE (-6)

#define GLE_PREC_FP64 0 /* drop-in: FP64 CLAs, reference rescale tables */
#define GLE_PREC_FP32 1 /* bulk: FP32 CLAs, exact power-of-two rescale counts */

/* bit flags reported by gle_get_rescale_flags */
#define GLE_FLAG_REDUCE_RESCALE 1 /* some site max < 1e-190 (FP64) */
#define GLE_FLAG_BAIL_OUT 2       /* some site max < 1e-250 (FP64) */

const char *gle_last_error(void);
/* Select the HIP device (default 0).  Fails with GLE_ERR_DEVICE if none. */
int gle_init(int device);
/* Number of visible HIP devices (0 on a host without GPU). */
int gle_device_count(void);

/* ---------------------------------------------------------------------------
 * Model and data helpers (host-side, FP64).  They replace Model::UpdateQMat
 * (model.cpp:233-294, 442-687, 735-780), Model::CalcEigenStuff (:827-856),
 * Model::DiscreteGamma (:2170-2201), NucleotideData/AminoacidData::CharToDatum
 * (sequencedata.cpp:672-780), NucleotideData::MakeAmbigStrings (:278-362) and
 * PatternManager::ProcessPatterns (datamatr.cpp:288-375).
 * ------------------------------------------------------------------------- */
/* GTR: rel = {AC, AG, AT, CG, CT} (GT = 1).  V, Vinv: 4x4 row-major; lam scaled
 * by the branch-length normaliser (model.cpp:845-846). */
int gle_model_gtr(const double *pi, const double *rel, double *V, double *Vinv, double *lam);
/* AA with 190 exchangeabilities in upper-triangle row order (model.cpp:747-760). */
int gle_model_aa(const double *pi, const double *rel190, double *V, double *Vinv, double *lam);
/* Codon model (61 sense codons, standard code): single-nucleotide changes with
 * exchangeabilities rel6 = {AC, AG, AT, CG, CT, GT} (GY94 with kappa is
 * {1, kappa, 1, 1, kappa, 1}), omega on nonsynonymous changes; one eigensystem
 * per omega class: V, Vinv [n_omega][61*61], lam [n_omega][61]. */
int gle_model_codon(const double *pi61, const double *rel6, const double *omega, const double *omega_prob,
                    int n_omega, double *V, double *Vinv, double *lam);
/* Mean-of-category discrete gamma rates (Yang 1994) as the reference computes them. */
int gle_discrete_gamma(int n_cat, double alpha, double *rates);
int gle_nuc_char_to_datum(char c); /* 1..15 bitmask, -1 unknown */
int gle_aa_char_to_datum(char c);  /* 0..19, 20 = missing, -1 unknown */
/* reference nucleotide ambiguity stream from 4-bit masks; out holds >= 5*n bytes */
int gle_make_ambig_string(const unsigned char *masks, int n, signed char *out);
/* inverse: stream -> masks (1..15); returns bytes consumed or <0 */
int gle_ambig_string_to_masks(const signed char *stream, int stream_len, int n, unsigned char *masks);
/* Pattern compression; same contract as PatternManager::ProcessPatterns:
 * matrix[tax*nsite+site]; out[tax*P+p]; returns P. */
int gle_compress_patterns(int ntax, int nsite, int max_states, const unsigned char *matrix,
                          unsigned char *out, int *counts, int *const_states, int *last_const,
                          int *site_to_pattern);

/* ---------------------------------------------------------------------------
 * Ingestion and output (host-side).  They replace the character-data part of
 * GarliReader / SequenceData::CreateMatrixFromNCL (garlireader.cpp,
 * sequencedata.cpp) and Tree::OutputSiteLikelihoods (tree.cpp:5501-5565).
 * ------------------------------------------------------------------------- */
enum { GLE_DATA_DNA = 0, GLE_DATA_PROTEIN = 1, GLE_DATA_OTHER = 2 };
/* NEXUS (DATA/CHARACTERS, TAXA labels, the starred EXSET and WTSET of an
 * ASSUMPTIONS block), relaxed PHYLIP or FASTA, told apart by content. */
int gle_alignment_read(const char *path, int *handle);
int gle_alignment_dims(int h, int *ntax, int *nchar, int *datatype);
/* name of taxon i into name[cap] (NUL-terminated); returns its length */
int gle_alignment_taxon(int h, int i, char *name, int cap);
/* the nchar symbols of taxon i (upper case; '?' missing, '-' gap; matchchar
 * resolved; a NEXUS state set as its IUPAC code) */
int gle_alignment_row(int h, int i, char *row);
/* per character: the WTSET weight (1 without one), 0 when the EXSET excludes it */
int gle_alignment_weights(int h, int *weights);
int gle_alignment_free(int h);
/* Per-character states of a DNA or protein alignment, out[ntax][nchar]: the
 * nucleotide bit set (A 1, C 2, G 4, T 8, missing 15; CharToDatum,
 * sequencedata.cpp:672-713) or the amino-acid index (ACDEFGHIKLMNPQRSTVWY =
 * 0..19, missing 20). */
int gle_alignment_states(int h, unsigned char *out);
/* Codon states of a DNA alignment (CodonData::FillCodonMatrixFromDNA,
 * sequencedata.cpp:364-470; standard code): out[ntax][nchar / 3] = sense
 * codon 0..60 (AAA, AAC, .., TTT without the stops), 61 for a codon with a
 * gap or ambiguity.  A stop codon fails with GLE_ERR_ARG, or is missing with
 * ignore_stops.  base_freqs (NULL allowed) receives the base frequencies at
 * codon positions 1, 2 and 3 (A, C, G, T each) for gle_codon_freqs_fx4. */
int gle_alignment_codons(int h, int ignore_stops, unsigned char *out, double *base_freqs);
/* Amino acids translated from a DNA alignment's codons
 * (AminoacidData::FillAminoacidMatrixFromDNA, sequencedata.cpp:473-590;
 * datatype = codon-aminoacid): out[ntax][nchar / 3], 20 = missing, an
 * ambiguous codon keeps its amino acid when every resolution agrees. */
int gle_alignment_translate(int h, int ignore_stops, unsigned char *out);
/* Empirical state frequencies of amino-acid or codon patterns
 * (CalcEmpiricalFreqs, sequencedata.cpp:192-275): states[ntax][n_patterns]
 * with n_states = missing, weighted by counts. */
int gle_empirical_freqs(int ntax, int n_patterns, int n_states, const unsigned char *states, const int *counts,
                        double *pi);
/* Empirical base frequencies of nucleotide bit-set patterns, ambiguity
 * shared out iteratively (NucleotideData::CalcEmpiricalFreqs,
 * sequencedata.cpp:61-190). */
int gle_empirical_nuc_freqs(int ntax, int n_patterns, const unsigned char *masks, const int *counts, double *pi);
/* F1x4 (f3x4 = 0) or F3x4 codon frequencies from gle_alignment_codons'
 * base_freqs (CalcF1x4Freqs / CalcF3x4Freqs, sequencedata.cpp:605-660). */
int gle_codon_freqs_fx4(const double *base_freqs, int f3x4, double *pi61);
/* orderedSiteDerivs / packedSiteDerivs logs in Tree::OutputSiteDerivatives's
 * layout (tree.cpp:5567-5595): "Partition subset <subset>" and a header, then
 * per character (site_pattern[c] = its pattern, or -1: "gap") and per pattern
 * the site lnL, d1, d2 (10 significant digits) and the rescale counts
 * (under_child NULL: "-").  As in the reference, a gap line ends without a
 * newline.  Either path may be NULL. */
int gle_write_site_derivatives(const char *ordered_path, const char *packed_path, int subset, int nchar,
                               const int *site_pattern, int n_patterns, const double *site_lnl, const double *site_d1,
                               const double *site_d2, const int *under_partial, const int *under_child);
/* <ofprefix>.sitelikes.log as GARLI writes it: the header unless append, one
 * line per character (-lnL, or "-" where included[s] == 0), then
 * "tree_num <tab> -lnl" (Tree::OutputSiteLikelihoods, tree.cpp:5501-5565;
 * population.cpp:3459-3478). */
int gle_write_sitelikes(const char *path, int append, int tree_num, double lnl, int nchar, const double *site_lnl,
                        const int *included);
/* Starting trees from a NEXUS TREES block, as GarliReader hands them to
 * Population::SeedPopulationWithStartingTree (garlireader.cpp:793, 1245;
 * population.cpp:1186): the TRANSLATE table, then each "tree name = newick;".
 * gle_treefile_tree gives tree k as the tree layer's arrays for n_taxa taxa
 * the way Tree::Tree(const char *, ..) builds it (tree.cpp:309-576): node 0
 * the root; taxon i (names[i - 1], or the label read as a number when names
 * is NULL or holds no match) node i; internal nodes n_taxa + 1.. in the order
 * the description opens them; parent[2 n_taxa - 2] (-1 at the root) and
 * blen (missing: 0.05, DEF_STARTING_BRLEN; given: held to [1e-8, 100],
 * defs.h:106-108).  A bifurcating top becomes a trifurcation by
 * MakeTrifurcatingRoot's rule (tree.cpp:761-797).  A polytomy is resolved as
 * ArbitrarilyBifurcate does (tree.cpp:855-923) with the node's first two
 * children joined where the reference draws the pair at random; the call
 * then returns 1.  score: the [!GarliScore s] comment, NaN without one.
 * gle_treefile_model copies the [!GarliModel ..] comment into buf[cap] and
 * returns its length (0 without one). */
int gle_treefile_read(const char *path, int *handle);
int gle_treefile_count(int h);
int gle_treefile_tree(int h, int k, int n_taxa, const char *const *names, int *parent, double *blen, double *score);
int gle_treefile_model(int h, int k, char *buf, int cap);
int gle_treefile_free(int h);

/* ---------------------------------------------------------------------------
 * Instance layer.
 * ------------------------------------------------------------------------- */
typedef struct gle_instance_desc {
  int n_tips;       /* taxa */
  int n_patterns;   /* compressed patterns */
  int n_states;     /* 4, 20, 61, or any S <= 64 */
  int n_rate_cats;  /* gamma categories or omega classes (GARLI NRateCats) */
  int n_eigen;      /* eigensystems (1, or n_omega for codon M3) */
  int n_cla_slots;  /* directional CLA buffers */
  int n_pmat_slots; /* transition-matrix slots (each holds P, P', P'') */
  int precision;    /* GLE_PREC_FP64 or GLE_PREC_FP32 */
} gle_instance_desc;

/* replaces CalculationManager/ClaManager allocation (calculationmanager.cpp:268-289,
 * condlike.cpp:122-146) */
int gle_create_instance(const gle_instance_desc *desc, int *handle);
int gle_destroy_instance(int handle);
/* Tip data in the reference's own encodings: a nucleotide ambiguity stream
 * (sequencedata.cpp:278-362), or for other data one state per pattern with
 * n_states meaning missing (sequencedata.h:614,677). */
int gle_set_tip_ambig_stream(int h, int tip, const signed char *stream, int stream_len);
int gle_set_tip_states(int h, int tip, const unsigned char *states);
/* counts per pattern; constant-pattern states (nuc bitmask / state index) for
 * patterns 0..last_const (DataPartition::GetCounts/LastConstant/GetConstStates) */
int gle_set_patterns(int h, const int *counts, int last_const, const int *const_states);
/* FP32 instances keep patterns whose count exceeds `threshold` (default 100:
 * count times the FP32 per-site error, about 1e-6 at most, stays under 1e-4) in an
 * FP64 part that runs the same op lists, so FP32 rounding is never multiplied
 * by a large pattern weight; 0 keeps every pattern in FP32.  No reference
 * counterpart (the reference has one precision per build, defs.h:86-112). */
int gle_set_heavy_split(int h, int threshold);
int gle_get_heavy_split(int h, int *n_heavy);
/* Trees with schedule 1 walk a whole op list in one launch only when the
 * instance has at least `min_patterns` padded patterns (default 2^19, enough
 * threads to fill the GPU); otherwise they launch per dependency level.  No
 * reference counterpart (a launch-policy knob). */
int gle_set_fused_threshold(int h, int min_patterns);
/* A level-launch sequence submitted twice in a row is captured into a hipGraph
 * and replayed from then on (default on).  enable = 0 keeps every launch
 * eager; either call drops the cached graph.  *replays (may be null) returns
 * how many updates were replayed from a graph so far.  No reference
 * counterpart (a launch-policy knob). */
int gle_set_graph_replay(int h, int enable, int *replays);
/* state freqs [S], category weights [R] (Model::GetRateProbs, model.h:2051-2064), pinv */
int gle_set_model(int h, const double *pi, const double *rate_probs, double pinv, int pinv_in_model);
/* eigensystem e: V, Vinv [S*S], lam [S] (already scaled).  Transition
 * matrices computed before the call describe the old model: recompute them
 * (gle_update_transition_matrices) before the next edge or CLA update, as
 * the reference's Model::UpdateQMat / CalcPmats order requires. */
int gle_set_eigen(int h, int e, const double *V, const double *Vinv, const double *lam);
/* per category: which eigensystem and the multiplier on lam
 * (rateMult/(1-pinv) with pinv in the model, rateMult otherwise, 1 for omega classes;
 * model.cpp:1474-1491) */
int gle_set_category_rates(int h, const int *eig_of_rate, const double *rate_mult);

/* Model::CalcPmats / CalcDerivatives (model.cpp:866-916, 1271-1358) on the GPU
 * in FP64: for each i, slot pmat_idx[i] <- P(t[i]) [and P', P'' if with_derivs]. */
int gle_update_transition_matrices(int h, const int *pmat_idx, const double *t, int n, int with_derivs);
/* Copy P (and P', P'' if derivs) of one slot back: out [3][R][S][S] */
int gle_get_transition_matrices(int h, int pmat, double *out);

/* One CLA update (ClaOperation; UpdateCLAs tree.cpp:3905-4001).
 * child >= 0: CLA slot; child < 0: tip index -(child+1).  rescale != 0 applies
 * RescaleRateHet[NState] (tree.cpp:3370-3598) after the product. */
typedef struct gle_cla_op {
  int dest, child1, pmat1, child2, pmat2, rescale;
} gle_cla_op;
/* ops sorted by dependency level; level_start[n_levels+1] gives the ranges
 * (BlockingOperationsSet, calculationmanager.h:118-123): non-decreasing from 0
 * to n_ops, and no op may read or overwrite a slot another op of its own
 * level writes (GLE_ERR_ARG otherwise).  n_levels = 0: the engine derives the
 * levels of a topologically ordered list itself.  FP64 instances return
 * GLE_RESCALE_RETRY, with every CLA written, when the rescale flags gathered
 * since their last reset (gle_get_rescale_flags) hold GLE_FLAG_REDUCE_RESCALE:
 * the reference's throw(1) from RescaleRateHet (tree.cpp:3436-3449), after
 * which the caller lowers its rescale frequency and recomputes. */
int gle_update_partials(int h, const gle_cla_op *ops, int n_ops, const int *level_start, int n_levels);
/* debugging/parity: one CLA slot back in the reference layout [P][R][S] + u[P] */
int gle_get_partials(int h, int slot, double *out, int *u);

/* Edge integration (ScoringOperation; GetScorePartial* tree.cpp:5597-6500 and
 * GetDerivsPartial* optimization.cpp:2412-3613).  child >= 0 CLA slot, < 0 tip.
 * out[3*i] = {lnL, d1, d2} (d1, d2 zero unless derivs).  site_lnl, if non-NULL,
 * receives the per-pattern unscaled lnL of op 0. */
typedef struct gle_edge_op {
  int partial, child, pmat, derivs;
} gle_edge_op;
int gle_edge_likelihoods(int h, const gle_edge_op *ops, int n, double *out, double *site_lnl);
/* gle_edge_likelihoods for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc); it has no counterpart in the reference.  The op list, the
 * per-block partial sums and the results live in the library's 8 MiB
 * per-device scratch array instead of buffers grown with hipMalloc: at most
 * 1024 ops, and ops x ceil(patterns / block) x 24 bytes must fit beside them
 * (block = 256 patterns for 4 states, 64 for 20, 16 for 61), else GLE_ERR_ARG.
 * site_lnl is refused (GLE_ERR_UNSUPPORTED) for an FP32 instance with an FP64
 * heavy part.  Calls that overlap on one GPU race on the scratch array. */
int gle_edge_likelihoods_noalloc(int h, const gle_edge_op *ops, int n, double *out, double *site_lnl);
/* One edge with derivatives and its per-pattern values, the inputs of
 * Tree::OutputSiteDerivatives (tree.cpp:5567-5595; filled under
 * OUTPUT_SITEDERIVS in GetDerivsPartialTerminal / Internal,
 * optimization.cpp:2412-3228): out3 = {lnL, d1, d2}; per pattern (any output
 * may be NULL) the unscaled site lnL, d1 = D1/L and d2 = D2/L - d1^2 (not
 * multiplied by the pattern count) and the rescale counts of the partial and
 * child sides (-1 for a tip child).  Counts are in the instance's units: the
 * reference's for the FP64 form, powers of two for the FP32 form.  The
 * op's transition-matrix slot must hold P' and P'' (computed with_derivs). */
int gle_edge_site_values(int h, const gle_edge_op *op, double *out3, double *site_lnl, double *site_d1,
                         double *site_d2, int *under_partial, int *under_child);
/* rescale flags accumulated since the last reset (see GLE_FLAG_*) */
int gle_get_rescale_flags(int h, int *flags, int reset);
int gle_synchronize(int h);
/* Per-launch HIP-event timing on the instance's stream.  kind: 0 CLA update
 * kernels, 1 edge kernels, 2 P(t) kernel; 3..5 the same for the FP64 part of
 * a split FP32 instance.  bytes are algorithmic HBM bytes
 * (true S, not padded).  No reference counterpart: the reference's Profiler
 * buckets (tree.cpp:41-47, utility.h:240-323) time whole CPU kernels instead. */
int gle_set_timing(int h, int enable);
/* Plain device-to-device copy of `bytes` (read + write), `reps` times; returns
 * the mean ms per copy.  No reference counterpart: it calibrates hardware
 * byte counters and the achievable HBM bandwidth for the roofline. */
int gle_copy_calibrate(size_t bytes, int reps, double *ms_per_copy);
/* Achievable FP64 MFMA rate (TFLOP/s) of v_mfma_f64_16x16x4_f64 with
 * register operands.  No reference counterpart: it calibrates the codon
 * kernels' MFMA utilisation. */
int gle_mfma_calibrate(double *tflops);
/* As gle_mfma_calibrate for shape 0 (v_mfma_f64_16x16x4_f64, four chains
 * per wave) or shape 1 (v_mfma_f64_4x4x4_4b_f64, eight chains per wave). */
int gle_mfma_calibrate_shape(int shape, double *tflops);
int gle_get_kernel_stats(int h, int kind, long long *launches, double *ms, double *bytes, int reset);

/* ---------------------------------------------------------------------------
 * Tree layer: GARLI's scoring surface over one instance.
 * Nodes 0..2N-3: 0 = root (3 children), 1..N tips, N+1..2N-3 internal;
 * parent[0] = -1, blen[n] = length of n's edge to its parent.  Children are
 * ordered by ascending node number (left, [middle], right).
 * ------------------------------------------------------------------------- */
int gle_tree_create(int inst, const int *parent, const double *blen, int *tree);
/* new tree sharing every clean CLA with the source (Tree::CopyClaIndeces,
 * tree.cpp:9209; copy-on-write through holder refcounts, clamanager.h:324-347) */
int gle_tree_copy(int tree, int *out);
int gle_tree_destroy(int tree);
/* Tree::SetBranchLength (tree.h:425) + SweepDirtynessOverTree (tree.cpp:4247) */
int gle_tree_set_branch_length(int tree, int node, double blen);
int gle_tree_get_branch_lengths(int tree, double *blen);
/* A topology move: Tree::SPRMutate / NNIMutate (tree.cpp:2571-2708, 6970)
 * followed by SweepDirtynessOverTree (tree.cpp:4247).  The tree takes the new
 * parent array and lengths (as in gle_tree_create) and keeps every clean CLA
 * whose side of the tree is unchanged -- matched by the exact subtree (tips,
 * topology and branch-length bits), whatever the node numbers -- so only the
 * CLAs on the paths through the changed edges are recomputed. */
int gle_tree_set_topology(int tree, const int *parent, const double *blen);
/* CLA updates the tree has computed so far (every call that ran ops). */
int gle_tree_cla_updates(int tree, long long *count);
/* Tree::MakeAllNodesDirty (tree.cpp): every CLA of this tree recomputed on next use */
int gle_tree_invalidate(int tree);
/* Tree::Score (tree.cpp:4003-4094), with the rescale retry; *lnl = Tree::lnL */
int gle_tree_score(int tree, double *lnl);
/* per-pattern site lnL of the last score (OutputSiteLikelihoods source),
 * computed on this call on the score's edge; GLE_ERR_STATE if the tree has
 * changed since that score */
int gle_tree_site_likelihoods(int tree, double *site_lnl);
/* Tree::CalcDerivativesRateHet(nd->anc, nd) (optimization.cpp:1099-1194) */
int gle_tree_derivatives(int tree, int node, double *lnl, double *d1, double *d2);
/* As gle_tree_derivatives, with the per-pattern values of gle_edge_site_values
 * for the edge above `node` (partial side: its parent). */
int gle_tree_site_derivatives(int tree, int node, double *out3, double *site_lnl, double *site_d1, double *site_d2,
                              int *under_partial, int *under_child);
/* Tree::OptimizeAllBranches (optimization.cpp:97-105, 1359-1972) */
int gle_tree_optimize_all_branches(int tree, double precision, double *improve, double *lnl);
/* Whole-tree sweep: every directional CLA, then (lnL, d1, d2) on every edge
 * (edge of node n at out[3*(n-1)]).  The batched form of the reference's
 * per-edge derivative calls. */
int gle_tree_sweep(int tree, double *edge_out);
/* directional CLA X(a|b) in the reference layout (parity/debug) */
int gle_tree_get_cla(int tree, int a, int b, double *out, int *u);
/* current rescaleEvery (16 at creation; lowered by 2 per retry) */
int gle_tree_rescale_every(int tree);
/* override rescaleEvery (tests force the rescale-retry path from the host) */
int gle_tree_set_rescale_every(int tree, int every);
/* Launch schedule for sweeps: 0 = one launch per dependency level,
 * 1 = one fused launch for the whole op list (pattern-local kernels only). */
int gle_tree_set_schedule(int tree, int mode);

/* GARLI's nst=6 nucleotide model with gamma rates and invariant sites.
 * rel = {AC, AG, AT, CG, CT} relative to GT = 1; alpha is used when the
 * instance has more than one rate category, pinv when pinv_in_model. */
typedef struct gle_gtr_params {
  double rel[5];
  double pi[4];
  double alpha;
  double pinv;
  int pinv_in_model;
} gle_gtr_params;
enum {
  GLE_OPT_RATES = 1,
  GLE_OPT_FREQS = 2,
  GLE_OPT_ALPHA = 4,
  GLE_OPT_PINV = 8,
  GLE_OPT_BRANCHES = 16,
  GLE_OPT_OMEGA = 32,
  GLE_OPT_OMEGA_PROBS = 64
};
/* Sets the tree's instance to the model *m: Q and its eigensystem, discrete
 * gamma category rates, rate weights (1 - pinv) / R and pinv; replaces
 * Model::UpdateQMat + CalcEigenStuff + DiscreteGamma for nst=6
 * (model.cpp:233-294, 827-856, 2170-2201). */
int gle_tree_set_gtr(int tree, const gle_gtr_params *m);
/* Maximises lnL on the tree's fixed topology over the parameters selected by
 * `which` (GLE_OPT_*; GLE_OPT_BRANCHES adds a Newton pass over all branches
 * each round): rounds of one-dimensional Brent searches, one parameter at a
 * time, until a round gains less than `precision`.  *m is updated to the
 * optimum and *lnl is its score.  Replaces the parameter loop of
 * Population::OptimizeInputOnly / BetterFinalOptimization (population.cpp:2566)
 * with OptimizeRelativeNucRates (tree.cpp:7307) and OptimizeBoundedParameter
 * (optimization.cpp:332). */
int gle_tree_optimize_gtr(int tree, gle_gtr_params *m, int which, double precision, double *lnl);

/* GARLI's amino-acid models (UpdateQMatAminoAcid, model.cpp:735-780):
 * exchangeabilities rel in upper-triangle row order over
 * ACDEFGHIKLMNPQRSTVWY (a named matrix such as WAG, or user / estimated
 * values; only ratios matter), frequencies pi, and rate heterogeneity as in
 * gle_gtr_params. */
typedef struct gle_aa_params {
  double rel[190];
  double pi[20];
  double alpha;
  double pinv;
  int pinv_in_model;
} gle_aa_params;
/* As gle_tree_set_gtr for a 20-state instance. */
int gle_tree_set_aa(int tree, const gle_aa_params *m);
/* As gle_tree_optimize_gtr for a 20-state instance: GLE_OPT_RATES (the first
 * 189 exchangeabilities relative to the last), GLE_OPT_FREQS, GLE_OPT_ALPHA,
 * GLE_OPT_PINV, GLE_OPT_BRANCHES.  Replaces the OptimizeInputOnly parameter
 * loop (population.cpp:2566) for ratematrix = fixed / wag / estimate. */
int gle_tree_optimize_aa(int tree, gle_aa_params *m, int which, double precision, double *lnl);

/* GARLI's codon models: exchangeabilities rel = {AC, AG, AT, CG, CT, GT}
 * between codons one nucleotide apart (GY94 with kappa: {1, k, 1, 1, k, 1}),
 * one omega class per rate category of the instance with its proportion
 * (n_omega = the instance's categories, at most 8), and 61 codon frequencies
 * (held fixed). */
typedef struct gle_codon_params {
  double rel[6];
  int n_omega;
  double omega[8];
  double omega_prob[8];
  double pi[61];
} gle_codon_params;
/* Sets the tree's instance to the codon model *m (one eigensystem per omega
 * class; replaces Model::UpdateQMatCodon + CalcEigenStuff, model.cpp:442-687,
 * 827-856). */
int gle_tree_set_codon(int tree, const gle_codon_params *m);
/* As gle_tree_optimize_gtr for codon models: GLE_OPT_RATES (AC..CT relative
 * to GT), GLE_OPT_OMEGA, GLE_OPT_OMEGA_PROBS, GLE_OPT_FREQS (the 61 codon
 * frequencies, the reference's statefrequencies = estimateF; its "estimate"
 * refuses codon models, model.h:866-876), GLE_OPT_BRANCHES.  Replaces
 * OptimizeRelativeNucRates (tree.cpp:7307), OptimizeOmegaParameters
 * (tree.cpp:7164) and OptimizeEquilibriumFreqs (tree.cpp:7288) in the
 * OptimizeInputOnly loop (population.cpp:2566). */
int gle_tree_optimize_codon(int tree, gle_codon_params *m, int which, double precision, double *lnl);

/* ---------------------------------------------------------------------------
 * Serving layer: continuous batching of scoring / branch-optimisation requests
 * from a GA population (the reference scores individuals one at a time,
 * population.cpp:4532-4589; Individual::CalcFitness, individual.cpp:224).
 * CLAs are cached on the device by subtree signature and reused across
 * requests; P(t) by branch length.  Requests run on the instance they name;
 * an instance served here must not be used by other calls at the same time.
 * ------------------------------------------------------------------------- */
#define GLE_REQ_SCORE 0    /* Tree::Score of the given tree */
#define GLE_REQ_OPTIMIZE 1 /* Tree::OptimizeAllBranches(precision), then Score */
int gle_serve_create(int *server);
int gle_serve_destroy(int server);
/* parent/blen as in gle_tree_create; returns a request id */
int gle_serve_submit(int server, int inst, const int *parent, const double *blen, int kind, double precision,
                     int *req);
/* n requests on one instance in one call (a GA generation's offspring,
 * Population::NextGeneration, population.cpp:4532-4589): parents and blens
 * hold n consecutive arrays of 2N - 2 entries; reqs[n] receives the ids.  If
 * any request is invalid, none is queued. */
int gle_serve_submit_batch(int server, int inst, int n, const int *parents, const double *blens, int kind,
                           double precision, int *reqs);
int gle_serve_poll(int server, int req, int *done);
/* blocks until done; blen_out (optional) receives the final branch lengths */
int gle_serve_wait(int server, int req, double *lnl, double *blen_out, double *latency_ms);
/* Blocks until all n requests are done and writes lnl[i] for reqs[i] (a GA
 * generation's offspring collected in one call); returns the first failed
 * request's error after every request was collected.  No reference
 * counterpart (GARLI scores one individual at a time). */
int gle_serve_wait_all(int server, int n, const int *reqs, double *lnl);
/* latency percentiles over completed requests; mean requests per batched step */
int gle_serve_stats(int server, long long *completed, double *p50_ms, double *p99_ms, double *mean_batch, int reset);
/* An idle server (nothing in flight) lets a burst of requests (a GA
 * generation's offspring) gather into one tick: after the first submission it
 * waits until no request has arrived for a fifth of `microseconds` (at least
 * 20 us) or `microseconds` have passed; 0 (the default) starts at once.  No
 * reference counterpart (GARLI scores one individual at a time). */
int gle_serve_set_batch_window(int server, int microseconds);
/* Milliseconds the worker spent building tick op lists on the host, and
 * collecting finished parts or polling for them while it had nothing to
 * build (time the host waited on the device), since the last reset. */
int gle_serve_timing(int server, double *host_ms, double *device_ms, int reset);
/* CLA work of the ticks built since the last reset: CLA ops; children
 * contracted on the device by CLA ops; children of CLA ops and sides of score
 * edges read from a kept message instead (a codon instance keeps a child's
 * term P(t) . x, named by the child's subtree and branch length, in a free
 * slot and reads it in place of contracting that child again); and messages
 * kept.  Serving statistics, no reference counterpart. */
int gle_serve_work(int server, long long *cla_ops, long long *contractions, long long *msg_reads,
                   long long *msg_keeps, int reset);

#ifdef __cplusplus
}
#endif
#endif
