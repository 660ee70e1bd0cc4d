"""ctypes bindings of libgle (include/gle.h).

The classes mirror the two C-ABI layers: :class:`Instance` is the op-list layer
(GARLI's CalculationManager seam, trunk/src/calculationmanager.h:39-123) and
:class:`Tree` is GARLI's tree-scoring surface (Tree::Score, tree.cpp:4003;
Tree::CalcDerivativesRateHet, optimization.cpp:1099; Tree::SetBranchLength,
tree.h:425; Tree::OptimizeAllBranches, optimization.cpp:97).

There is no CPU fallback: if libgle.so is missing, or no GPU is visible, the
calls raise.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgle.so")

PREC_FP64 = 0
PREC_FP32 = 1
SUCCESS = 0
RESCALE_RETRY = 1

_lib = None


class GleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("libgle error %d: %s" % (code, msg))
        self.code = code


class ClaOp(C.Structure):
    _fields_ = [("dest", C.c_int), ("child1", C.c_int), ("pmat1", C.c_int),
                ("child2", C.c_int), ("pmat2", C.c_int), ("rescale", C.c_int)]


class EdgeOp(C.Structure):
    _fields_ = [("partial", C.c_int), ("child", C.c_int), ("pmat", C.c_int), ("derivs", C.c_int)]


class InstanceDesc(C.Structure):
    _fields_ = [("n_tips", C.c_int), ("n_patterns", C.c_int), ("n_states", C.c_int),
                ("n_rate_cats", C.c_int), ("n_eigen", C.c_int), ("n_cla_slots", C.c_int),
                ("n_pmat_slots", C.c_int), ("precision", C.c_int)]


_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_U8 = C.POINTER(C.c_uint8)
_S8 = C.POINTER(C.c_int8)

# every symbol include/gle.h declares, with its argument types
SIGNATURES = {
    "gle_last_error": ([], C.c_char_p),
    "gle_init": ([C.c_int], C.c_int),
    "gle_device_count": ([], C.c_int),
    "gle_model_gtr": ([_D, _D, _D, _D, _D], C.c_int),
    "gle_model_aa": ([_D, _D, _D, _D, _D], C.c_int),
    "gle_model_codon": ([_D, _D, _D, _D, C.c_int, _D, _D, _D], C.c_int),
    "gle_discrete_gamma": ([C.c_int, C.c_double, _D], C.c_int),
    "gle_nuc_char_to_datum": ([C.c_char], C.c_int),
    "gle_aa_char_to_datum": ([C.c_char], C.c_int),
    "gle_make_ambig_string": ([_U8, C.c_int, _S8], C.c_int),
    "gle_ambig_string_to_masks": ([_S8, C.c_int, C.c_int, _U8], C.c_int),
    "gle_compress_patterns": ([C.c_int, C.c_int, C.c_int, _U8, _U8, _I, _I, _I, _I], C.c_int),
    "gle_alignment_read": ([C.c_char_p, C.POINTER(C.c_int)], C.c_int),
    "gle_alignment_dims": ([C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "gle_alignment_taxon": ([C.c_int, C.c_int, C.c_char_p, C.c_int], C.c_int),
    "gle_alignment_row": ([C.c_int, C.c_int, C.c_char_p], C.c_int),
    "gle_alignment_weights": ([C.c_int, C.POINTER(C.c_int)], C.c_int),
    "gle_alignment_free": ([C.c_int], C.c_int),
    "gle_alignment_states": ([C.c_int, C.POINTER(C.c_uint8)], C.c_int),
    "gle_alignment_codons": ([C.c_int, C.c_int, C.POINTER(C.c_uint8), _D], C.c_int),
    "gle_alignment_translate": ([C.c_int, C.c_int, C.POINTER(C.c_uint8)], C.c_int),
    "gle_empirical_freqs": ([C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint8), C.POINTER(C.c_int), _D], C.c_int),
    "gle_empirical_nuc_freqs": ([C.c_int, C.c_int, C.POINTER(C.c_uint8), C.POINTER(C.c_int), _D], C.c_int),
    "gle_codon_freqs_fx4": ([_D, C.c_int, _D], C.c_int),
    "gle_write_sitelikes": ([C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_int, _D, C.POINTER(C.c_int)], C.c_int),
    "gle_write_site_derivatives": ([C.c_char_p, C.c_char_p, C.c_int, C.c_int, _I, C.c_int, _D, _D, _D, _I, _I],
                                   C.c_int),
    "gle_treefile_read": ([C.c_char_p, C.POINTER(C.c_int)], C.c_int),
    "gle_treefile_count": ([C.c_int], C.c_int),
    "gle_treefile_tree": ([C.c_int, C.c_int, C.c_int, C.POINTER(C.c_char_p), _I, _D, _D], C.c_int),
    "gle_treefile_model": ([C.c_int, C.c_int, C.c_char_p, C.c_int], C.c_int),
    "gle_treefile_free": ([C.c_int], C.c_int),
    "gle_create_instance": ([C.POINTER(InstanceDesc), _I], C.c_int),
    "gle_destroy_instance": ([C.c_int], C.c_int),
    "gle_set_tip_ambig_stream": ([C.c_int, C.c_int, _S8, C.c_int], C.c_int),
    "gle_set_tip_states": ([C.c_int, C.c_int, _U8], C.c_int),
    "gle_set_patterns": ([C.c_int, _I, C.c_int, _I], C.c_int),
    "gle_set_model": ([C.c_int, _D, _D, C.c_double, C.c_int], C.c_int),
    "gle_set_heavy_split": ([C.c_int, C.c_int], C.c_int),
    "gle_get_heavy_split": ([C.c_int, _I], C.c_int),
    "gle_set_fused_threshold": ([C.c_int, C.c_int], C.c_int),
    "gle_set_graph_replay": ([C.c_int, C.c_int, _I], C.c_int),
    "gle_set_eigen": ([C.c_int, C.c_int, _D, _D, _D], C.c_int),
    "gle_set_category_rates": ([C.c_int, _I, _D], C.c_int),
    "gle_update_transition_matrices": ([C.c_int, _I, _D, C.c_int, C.c_int], C.c_int),
    "gle_get_transition_matrices": ([C.c_int, C.c_int, _D], C.c_int),
    "gle_update_partials": ([C.c_int, C.POINTER(ClaOp), C.c_int, _I, C.c_int], C.c_int),
    "gle_get_partials": ([C.c_int, C.c_int, _D, _I], C.c_int),
    "gle_edge_likelihoods": ([C.c_int, C.POINTER(EdgeOp), C.c_int, _D, _D], C.c_int),
    "gle_edge_likelihoods_noalloc": ([C.c_int, C.POINTER(EdgeOp), C.c_int, _D, _D], C.c_int),
    "gle_edge_site_values": ([C.c_int, C.POINTER(EdgeOp), _D, _D, _D, _D, _I, _I], C.c_int),
    "gle_get_rescale_flags": ([C.c_int, _I, C.c_int], C.c_int),
    "gle_synchronize": ([C.c_int], C.c_int),
    "gle_set_timing": ([C.c_int, C.c_int], C.c_int),
    "gle_copy_calibrate": ([C.c_size_t, C.c_int, _D], C.c_int),
    "gle_mfma_calibrate": ([_D], C.c_int),
    "gle_mfma_calibrate_shape": ([C.c_int, _D], C.c_int),
    "gle_get_kernel_stats": ([C.c_int, C.c_int, C.POINTER(C.c_longlong), _D, _D, C.c_int], C.c_int),
    "gle_tree_create": ([C.c_int, _I, _D, _I], C.c_int),
    "gle_tree_set_topology": ([C.c_int, _I, _D], C.c_int),
    "gle_tree_cla_updates": ([C.c_int, C.POINTER(C.c_longlong)], C.c_int),
    "gle_tree_copy": ([C.c_int, _I], C.c_int),
    "gle_tree_destroy": ([C.c_int], C.c_int),
    "gle_tree_set_branch_length": ([C.c_int, C.c_int, C.c_double], C.c_int),
    "gle_tree_get_branch_lengths": ([C.c_int, _D], C.c_int),
    "gle_tree_invalidate": ([C.c_int], C.c_int),
    "gle_tree_score": ([C.c_int, _D], C.c_int),
    "gle_tree_site_likelihoods": ([C.c_int, _D], C.c_int),
    "gle_tree_derivatives": ([C.c_int, C.c_int, _D, _D, _D], C.c_int),
    "gle_tree_site_derivatives": ([C.c_int, C.c_int, _D, _D, _D, _D, _I, _I], C.c_int),
    "gle_tree_optimize_all_branches": ([C.c_int, C.c_double, _D, _D], C.c_int),
    "gle_tree_sweep": ([C.c_int, _D], C.c_int),
    "gle_tree_get_cla": ([C.c_int, C.c_int, C.c_int, _D, _I], C.c_int),
    "gle_tree_rescale_every": ([C.c_int], C.c_int),
    "gle_tree_set_rescale_every": ([C.c_int, C.c_int], C.c_int),
    "gle_tree_set_schedule": ([C.c_int, C.c_int], C.c_int),
    "gle_tree_set_gtr": ([C.c_int, C.c_void_p], C.c_int),
    "gle_tree_optimize_gtr": ([C.c_int, C.c_void_p, C.c_int, C.c_double, _D], C.c_int),
    "gle_tree_set_codon": ([C.c_int, C.c_void_p], C.c_int),
    "gle_tree_set_aa": ([C.c_int, C.c_void_p], C.c_int),
    "gle_tree_optimize_aa": ([C.c_int, C.c_void_p, C.c_int, C.c_double, _D], C.c_int),
    "gle_tree_optimize_codon": ([C.c_int, C.c_void_p, C.c_int, C.c_double, _D], C.c_int),
    "gle_serve_create": ([_I], C.c_int),
    "gle_serve_destroy": ([C.c_int], C.c_int),
    "gle_serve_submit": ([C.c_int, C.c_int, _I, _D, C.c_int, C.c_double, _I], C.c_int),
    "gle_serve_submit_batch": ([C.c_int, C.c_int, C.c_int, _I, _D, C.c_int, C.c_double, _I], C.c_int),
    "gle_serve_poll": ([C.c_int, C.c_int, _I], C.c_int),
    "gle_serve_wait": ([C.c_int, C.c_int, _D, _D, _D], C.c_int),
    "gle_serve_wait_all": ([C.c_int, C.c_int, _I, _D], C.c_int),
    "gle_serve_stats": ([C.c_int, C.POINTER(C.c_longlong), _D, _D, _D, C.c_int], C.c_int),
    "gle_serve_set_batch_window": ([C.c_int, C.c_int], C.c_int),
    "gle_serve_timing": ([C.c_int, _D, _D, C.c_int], C.c_int),
    "gle_serve_work": ([C.c_int] + [C.POINTER(C.c_longlong)] * 4 + [C.c_int], C.c_int),
}


def lib():
    """Load libgle.so (built by garli_amd.build); raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgle.so not built: run `python -m garli_amd.build` "
                               "(the engine has no CPU fallback)")
        l = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(l, name)
            f.argtypes = args
            f.restype = res
        _lib = l
    return _lib


def _check(rc, allow=()):
    if rc < 0 or (rc > 0 and rc not in allow):
        raise GleError(rc, lib().gle_last_error().decode())
    return rc


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def init(device=0):
    _check(lib().gle_init(device))


def device_count():
    return lib().gle_device_count()


def mfma_calibrate(shape=0):
    """Achievable FP64 MFMA TFLOP/s (register operands, 8 waves per SIMD) of
    v_mfma_f64_16x16x4_f64 (shape 0) or v_mfma_f64_4x4x4_4b_f64 (shape 1)."""
    tf = C.c_double(0)
    _check(lib().gle_mfma_calibrate_shape(int(shape), C.byref(tf)))
    return tf.value


def copy_calibrate(nbytes, reps=10):
    """Mean ms of a plain device copy of nbytes (reads nbytes, writes nbytes)."""
    ms = C.c_double(0)
    _check(lib().gle_copy_calibrate(int(nbytes), int(reps), C.byref(ms)))
    return ms.value


# ------------------------------------------------------------ model helpers
def model_gtr(pi, rel):
    V = np.zeros(16)
    W = np.zeros(16)
    lam = np.zeros(4)
    pi, rel = _d(pi), _d(rel)
    _check(lib().gle_model_gtr(_p(pi, C.c_double), _p(rel, C.c_double), _p(V, C.c_double),
                               _p(W, C.c_double), _p(lam, C.c_double)))
    return V.reshape(4, 4), W.reshape(4, 4), lam


def model_aa(pi, rel190):
    V = np.zeros(400)
    W = np.zeros(400)
    lam = np.zeros(20)
    pi, rel = _d(pi), _d(rel190)
    _check(lib().gle_model_aa(_p(pi, C.c_double), _p(rel, C.c_double), _p(V, C.c_double),
                              _p(W, C.c_double), _p(lam, C.c_double)))
    return V.reshape(20, 20), W.reshape(20, 20), lam


def model_codon(pi61, rel6, omega, omega_prob):
    """rel6 = (AC, AG, AT, CG, CT, GT); GY94 with kappa is (1, kappa, 1, 1, kappa, 1)."""
    nw = len(omega)
    V = np.zeros(nw * 3721)
    W = np.zeros(nw * 3721)
    lam = np.zeros(nw * 61)
    pi, r6, om, op = _d(pi61), _d(rel6), _d(omega), _d(omega_prob)
    _check(lib().gle_model_codon(_p(pi, C.c_double), _p(r6, C.c_double), _p(om, C.c_double), _p(op, C.c_double),
                                 nw, _p(V, C.c_double), _p(W, C.c_double), _p(lam, C.c_double)))
    return V.reshape(nw, 61, 61), W.reshape(nw, 61, 61), lam.reshape(nw, 61)


def discrete_gamma(ncat, alpha):
    r = np.zeros(ncat)
    _check(lib().gle_discrete_gamma(ncat, float(alpha), _p(r, C.c_double)))
    return r


def nuc_char_to_datum(ch):
    return lib().gle_nuc_char_to_datum(ch.encode() if isinstance(ch, str) else ch)


def aa_char_to_datum(ch):
    return lib().gle_aa_char_to_datum(ch.encode() if isinstance(ch, str) else ch)


def make_ambig_string(masks):
    m = np.ascontiguousarray(masks, dtype=np.uint8)
    out = np.zeros(5 * len(m) + 1, dtype=np.int8)
    n = _check(lib().gle_make_ambig_string(_p(m, C.c_uint8), len(m), _p(out, C.c_int8)), allow=range(1 << 30))
    return out[:n].copy()


def ambig_string_to_masks(stream, n):
    s = np.ascontiguousarray(stream, dtype=np.int8)
    out = np.zeros(n, dtype=np.uint8)
    _check(lib().gle_ambig_string_to_masks(_p(s, C.c_int8), len(s), n, _p(out, C.c_uint8)), allow=range(1 << 30))
    return out


DATA_DNA, DATA_PROTEIN, DATA_OTHER = 0, 1, 2


def read_start_trees(path, taxon_names=None, n_taxa=None):
    """The trees of a NEXUS TREES block (a GARLI starting-tree file) read by
    gle_treefile_read: a list of dicts with the tree's name-free arrays
    parent / blen (node 0 the root, taxon i node i, in taxon_names' order),
    score (the GarliScore comment, NaN without one), model (the GarliModel
    comment, "" without one) and polytomies_resolved."""
    if n_taxa is None:
        n_taxa = len(taxon_names)
    h = C.c_int(-1)
    _check(lib().gle_treefile_read(os.fsencode(path), C.byref(h)))
    try:
        n = lib().gle_treefile_count(h)
        _check(n, allow=range(1 << 30))
        names = None
        if taxon_names is not None:
            names = (C.c_char_p * n_taxa)(*[t.encode() for t in taxon_names])
        out = []
        buf = C.create_string_buffer(1 << 16)
        for k in range(n):
            parent = np.zeros(2 * n_taxa - 2, dtype=np.int32)
            blen = np.zeros(2 * n_taxa - 2)
            score = C.c_double()
            rc = _check(lib().gle_treefile_tree(h, k, n_taxa, names, _p(parent, C.c_int), _p(blen, C.c_double),
                                                C.byref(score)), allow=(1,))
            m = lib().gle_treefile_model(h, k, buf, len(buf))
            _check(m, allow=range(1 << 30))
            out.append(dict(parent=parent, blen=blen, score=score.value, model=buf.value.decode(),
                            polytomies_resolved=rc == 1))
        return out
    finally:
        lib().gle_treefile_free(h)


class Alignment:
    """Character data read by gle_alignment_read (NEXUS, PHYLIP or FASTA):
    taxon names, rows of symbols and per-character weights (the WTSET weight,
    0 where the EXSET excludes the character)."""

    def __init__(self, path):
        h = C.c_int(-1)
        _check(lib().gle_alignment_read(os.fsencode(path), C.byref(h)))
        try:
            nt, nc, dt = C.c_int(), C.c_int(), C.c_int()
            _check(lib().gle_alignment_dims(h, C.byref(nt), C.byref(nc), C.byref(dt)))
            self.datatype = dt.value
            self.names, self.rows = [], []
            buf = C.create_string_buffer(4096)
            row = C.create_string_buffer(nc.value + 1)
            for i in range(nt.value):
                _check(lib().gle_alignment_taxon(h, i, buf, len(buf)), allow=range(1 << 30))
                self.names.append(buf.value.decode())
                _check(lib().gle_alignment_row(h, i, row))
                self.rows.append(row.raw[:nc.value].decode())
            self.weights = np.zeros(nc.value, dtype=np.int32)
            _check(lib().gle_alignment_weights(h, _p(self.weights, C.c_int)))
            self._h = h.value
            h = None
        finally:
            if h is not None:
                lib().gle_alignment_free(h)

    def close(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            lib().gle_alignment_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_chars(self):
        return len(self.weights)

    def states(self):
        """Per-character states [ntax][nchar]: nucleotide bit sets (A 1, C 2,
        G 4, T 8, missing 15) or amino-acid indices (missing 20)."""
        out = np.zeros((len(self.rows), self.n_chars), dtype=np.uint8)
        _check(lib().gle_alignment_states(self._h, _p(out, C.c_uint8)))
        return out

    def _codon_dims(self):
        if self.datatype != DATA_DNA or self.n_chars % 3:
            raise GleError(-1, "codons need DNA data with a whole number of codons")
        return len(self.rows), self.n_chars // 3

    def codon_states(self, ignore_stops=False):
        """Sense-codon states [ntax][nchar/3] (0..60, missing 61) and the base
        frequencies at the three codon positions [3][4] (datatype = codon)."""
        out = np.zeros(self._codon_dims(), dtype=np.uint8)
        bf = np.zeros(12)
        _check(lib().gle_alignment_codons(self._h, int(bool(ignore_stops)), _p(out, C.c_uint8), _p(bf, C.c_double)))
        return out, bf.reshape(3, 4)

    def translate(self, ignore_stops=False):
        """Amino acids of the codons [ntax][nchar/3] (missing 20; datatype =
        codon-aminoacid)."""
        out = np.zeros(self._codon_dims(), dtype=np.uint8)
        _check(lib().gle_alignment_translate(self._h, int(bool(ignore_stops)), _p(out, C.c_uint8)))
        return out

    def nucleotide_patterns(self):
        """Compressed nucleotide patterns with the weights applied:
        (patterns [ntax][P], counts, const_states, last_const, pattern of each
        character or -1 where it is excluded or holds no data)."""
        if self.datatype != DATA_DNA:
            raise GleError(-1, "not nucleotide data")
        return _weighted_patterns(self.states(), self.weights, 4)

    def aa_patterns(self):
        """As nucleotide_patterns for a protein alignment (20 states)."""
        if self.datatype != DATA_PROTEIN:
            raise GleError(-1, "not protein data")
        return _weighted_patterns(self.states(), self.weights, 20)

    def codon_patterns(self, ignore_stops=False):
        """Codon patterns (61 states); a pattern of each codon.  Weight sets
        do not apply to codons (nor to the translation, as in the
        reference's CodonData / AminoacidData, sequencedata.cpp:369-371)."""
        cod, _ = self.codon_states(ignore_stops)
        return _weighted_patterns(cod, np.ones(cod.shape[1], np.int32), 61)

    def translated_patterns(self, ignore_stops=False):
        """Amino-acid patterns of the translated codons (20 states)."""
        aa = self.translate(ignore_stops)
        return _weighted_patterns(aa, np.ones(aa.shape[1], np.int32), 20)


def _weighted_patterns(states, weights, n_states):
    keep = np.nonzero(weights > 0)[0]
    pats, counts, cs, lc, s2p = compress_patterns(np.ascontiguousarray(states[:, keep]), n_states)
    w = np.zeros(pats.shape[1], dtype=np.int64)
    np.add.at(w, s2p[s2p >= 0], weights[keep][s2p >= 0])
    site = np.full(len(weights), -1, dtype=np.int32)
    site[keep] = s2p
    return pats, w.astype(np.int32), cs, lc, site


def empirical_freqs(patterns, counts, n_states):
    """Empirical frequencies of amino-acid (20) or codon (61) patterns; for
    nucleotide bit-set patterns (n_states = 4) ambiguity is shared out as the
    reference does."""
    st = np.ascontiguousarray(patterns, dtype=np.uint8)
    c = np.ascontiguousarray(counts, dtype=np.int32)
    ntax, npat = st.shape
    pi = np.zeros(n_states)
    if n_states == 4:
        _check(lib().gle_empirical_nuc_freqs(ntax, npat, _p(st, C.c_uint8), _p(c, C.c_int), _p(pi, C.c_double)))
    else:
        _check(lib().gle_empirical_freqs(ntax, npat, n_states, _p(st, C.c_uint8), _p(c, C.c_int),
                                         _p(pi, C.c_double)))
    return pi


def codon_freqs_fx4(base_freqs, f3x4=False):
    """F1x4 or F3x4 codon frequencies from codon-position base frequencies."""
    bf = np.ascontiguousarray(base_freqs, dtype=np.float64).reshape(12)
    pi = np.zeros(61)
    _check(lib().gle_codon_freqs_fx4(_p(bf, C.c_double), int(bool(f3x4)), _p(pi, C.c_double)))
    return pi


def _site_values(call, P):
    out = np.zeros(3)
    v = {k: np.zeros(P) for k in ("lnl", "d1", "d2")}
    u = {k: np.zeros(P, dtype=np.int32) for k in ("under_partial", "under_child")}
    _check(call(_p(out, C.c_double), _p(v["lnl"], C.c_double), _p(v["d1"], C.c_double), _p(v["d2"], C.c_double),
                _p(u["under_partial"], C.c_int), _p(u["under_child"], C.c_int)))
    v.update(u)
    return tuple(out), v


def write_site_derivatives(ordered_path, packed_path, site_pattern, values, subset=1):
    """orderedSiteDerivs / packedSiteDerivs logs from Tree.site_derivatives'
    values; site_pattern maps each character to its pattern (-1: gap)."""
    sp = np.ascontiguousarray(site_pattern, dtype=np.int32)
    lnl, d1, d2 = (np.ascontiguousarray(values[k], dtype=np.float64) for k in ("lnl", "d1", "d2"))
    u1 = np.ascontiguousarray(values["under_partial"], dtype=np.int32)
    u2 = np.ascontiguousarray(values["under_child"], dtype=np.int32)
    tip_child = bool(np.all(u2 < 0))
    _check(lib().gle_write_site_derivatives(
        None if ordered_path is None else os.fsencode(ordered_path),
        None if packed_path is None else os.fsencode(packed_path), int(subset), len(sp), _p(sp, C.c_int), len(lnl),
        _p(lnl, C.c_double), _p(d1, C.c_double), _p(d2, C.c_double), _p(u1, C.c_int),
        None if tip_child else _p(u2, C.c_int)))


def write_sitelikes(path, lnl, site_lnl, included, tree_num=1, append=False):
    """<ofprefix>.sitelikes.log in GARLI's format (per character -lnL or "-",
    then the tree's -lnL)."""
    s = np.ascontiguousarray(site_lnl, dtype=np.float64)
    inc = np.ascontiguousarray(included, dtype=np.int32)
    _check(lib().gle_write_sitelikes(os.fsencode(path), int(append), int(tree_num), float(lnl), len(s),
                                     _p(s, C.c_double), _p(inc, C.c_int)))


def compress_patterns(matrix, max_states):
    """PatternManager::ProcessPatterns: returns (patterns [ntax][P], counts,
    const_states, last_const, site_to_pattern)."""
    m = np.ascontiguousarray(matrix, dtype=np.uint8)
    ntax, nsite = m.shape
    out = np.zeros((ntax, nsite), dtype=np.uint8)
    counts = np.zeros(nsite, dtype=np.int32)
    cs = np.zeros(nsite, dtype=np.int32)
    lc = np.zeros(1, dtype=np.int32)
    s2p = np.zeros(nsite, dtype=np.int32)
    P = _check(lib().gle_compress_patterns(ntax, nsite, max_states, _p(m, C.c_uint8), _p(out, C.c_uint8),
                                           _p(counts, C.c_int), _p(cs, C.c_int), _p(lc, C.c_int),
                                           _p(s2p, C.c_int)), allow=range(1 << 30))
    pats = out.reshape(-1)[: ntax * P].reshape(ntax, P).copy()
    return pats, counts[:P].copy(), cs[:P].copy(), int(lc[0]), s2p


# ----------------------------------------------------------------- instance
class Instance:
    """One data partition on the device (a BEAGLE-style instance)."""

    def __init__(self, n_tips, n_patterns, n_states, n_rates, n_eigen=1, n_cla_slots=None,
                 n_pmat_slots=None, precision=PREC_FP64):
        self.n_tips, self.n_patterns, self.n_states, self.n_rates = n_tips, n_patterns, n_states, n_rates
        self.precision = precision
        if n_cla_slots is None:
            n_cla_slots = 6 * (n_tips - 2) + 16   # room for a tree and a copy-on-write copy
        if n_pmat_slots is None:
            n_pmat_slots = 8 * n_tips + 16
        d = InstanceDesc(n_tips, n_patterns, n_states, n_rates, n_eigen, n_cla_slots, n_pmat_slots, precision)
        h = C.c_int(-1)
        _check(lib().gle_create_instance(C.byref(d), C.byref(h)))
        self.h = h.value

    def close(self):
        if self.h is not None and self.h >= 0:
            lib().gle_destroy_instance(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tip_states(self, tip, states):
        s = np.ascontiguousarray(states, dtype=np.uint8)
        assert len(s) == self.n_patterns
        _check(lib().gle_set_tip_states(self.h, tip, _p(s, C.c_uint8)))

    def set_tip_ambig_stream(self, tip, stream):
        s = np.ascontiguousarray(stream, dtype=np.int8)
        _check(lib().gle_set_tip_ambig_stream(self.h, tip, _p(s, C.c_int8), len(s)))

    def set_patterns(self, counts, last_const, const_states):
        c = np.ascontiguousarray(counts, dtype=np.int32)
        k = np.ascontiguousarray(const_states, dtype=np.int32)
        _check(lib().gle_set_patterns(self.h, _p(c, C.c_int), int(last_const), _p(k, C.c_int)))

    def set_heavy_split(self, threshold):
        """FP32 instances: patterns with count > threshold run in FP64 (0 = none)."""
        _check(lib().gle_set_heavy_split(self.h, int(threshold)))

    def set_fused_threshold(self, min_patterns):
        """Fused single-launch walks only for instances with >= min_patterns (padded)."""
        _check(lib().gle_set_fused_threshold(self.h, int(min_patterns)))

    def set_graph_replay(self, enable):
        """Replay repeated level-launch sequences from a captured hipGraph; returns
        the number of replays so far (the cached graph is dropped either way)."""
        n = C.c_int(0)
        _check(lib().gle_set_graph_replay(self.h, int(bool(enable)), C.byref(n)))
        return n.value

    def heavy_patterns(self):
        n = C.c_int(0)
        _check(lib().gle_get_heavy_split(self.h, C.byref(n)))
        return n.value

    def set_model(self, pi, rate_probs, pinv=0.0, pinv_in_model=False):
        pi, w = _d(pi), _d(rate_probs)
        _check(lib().gle_set_model(self.h, _p(pi, C.c_double), _p(w, C.c_double), float(pinv), int(pinv_in_model)))

    def set_eigen(self, e, V, Vinv, lam):
        V, W, l = _d(V), _d(Vinv), _d(lam)
        _check(lib().gle_set_eigen(self.h, e, _p(V, C.c_double), _p(W, C.c_double), _p(l, C.c_double)))

    def set_category_rates(self, eig_of_rate, mult):
        e = np.ascontiguousarray(eig_of_rate, dtype=np.int32)
        m = _d(mult)
        _check(lib().gle_set_category_rates(self.h, _p(e, C.c_int), _p(m, C.c_double)))

    def update_transition_matrices(self, idx, t, derivs=False):
        i = np.ascontiguousarray(idx, dtype=np.int32)
        tt = _d(t)
        _check(lib().gle_update_transition_matrices(self.h, _p(i, C.c_int), _p(tt, C.c_double), len(i), int(derivs)))

    def get_transition_matrices(self, slot):
        S, R = self.n_states, self.n_rates
        out = np.zeros(3 * R * S * S)
        _check(lib().gle_get_transition_matrices(self.h, slot, _p(out, C.c_double)))
        return out.reshape(3, R, S, S)

    def update_partials(self, ops, level_start=None):
        """One batch of CLA ops (dest, child1, pmat1, child2, pmat2, rescale);
        returns RESCALE_RETRY when the reference would have thrown from
        RescaleRateHet (FP64), SUCCESS otherwise."""
        arr = (ClaOp * len(ops))(*[ClaOp(*o) for o in ops])
        if level_start is None:
            return _check(lib().gle_update_partials(self.h, arr, len(ops), None, 0), allow=(RESCALE_RETRY,))
        ls = np.ascontiguousarray(level_start, dtype=np.int32)
        return _check(lib().gle_update_partials(self.h, arr, len(ops), _p(ls, C.c_int), len(ls) - 1),
                      allow=(RESCALE_RETRY,))

    def get_partials(self, slot):
        out = np.zeros(self.n_patterns * self.n_rates * self.n_states)
        u = np.zeros(self.n_patterns, dtype=np.int32)
        _check(lib().gle_get_partials(self.h, slot, _p(out, C.c_double), _p(u, C.c_int)))
        return out.reshape(self.n_patterns, self.n_rates, self.n_states), u

    def edge_likelihoods(self, ops, want_sites=False):
        arr = (EdgeOp * len(ops))(*[EdgeOp(*o) for o in ops])
        out = np.zeros(3 * len(ops))
        site = np.zeros(self.n_patterns) if want_sites else None
        _check(lib().gle_edge_likelihoods(self.h, arr, len(ops), _p(out, C.c_double),
                                          _p(site, C.c_double) if want_sites else None))
        return (out.reshape(-1, 3), site) if want_sites else out.reshape(-1, 3)

    def edge_likelihoods_noalloc(self, ops, want_sites=False):
        """edge_likelihoods through the allocation-free entry point (device
        buffers from the library's per-device scratch array)."""
        arr = (EdgeOp * len(ops))(*[EdgeOp(*o) for o in ops])
        out = np.zeros(3 * len(ops))
        site = np.zeros(self.n_patterns) if want_sites else None
        _check(lib().gle_edge_likelihoods_noalloc(self.h, arr, len(ops), _p(out, C.c_double),
                                                  _p(site, C.c_double) if want_sites else None))
        return (out.reshape(-1, 3), site) if want_sites else out.reshape(-1, 3)

    def edge_site_values(self, op):
        """One edge (partial, child, pmat) with derivatives: ((lnL, d1, d2),
        per-pattern dict of lnl, d1, d2, under_partial, under_child)."""
        e = EdgeOp(op[0], op[1], op[2], 1)
        return _site_values(lambda *a: lib().gle_edge_site_values(self.h, C.byref(e), *a), self.n_patterns)

    def set_timing(self, enable=True):
        _check(lib().gle_set_timing(self.h, int(enable)))

    def kernel_stats(self, kind, reset=False):
        """(launches, total ms, algorithmic bytes) for kind 0 CLA, 1 edge, 2 P(t)."""
        n, ms, b = C.c_longlong(0), C.c_double(0), C.c_double(0)
        _check(lib().gle_get_kernel_stats(self.h, kind, C.byref(n), C.byref(ms), C.byref(b), int(reset)))
        return n.value, ms.value, b.value

    def rescale_flags(self, reset=True):
        f = C.c_int(0)
        _check(lib().gle_get_rescale_flags(self.h, C.byref(f), int(reset)))
        return f.value


class Tree:
    """GARLI's Tree scoring surface over an :class:`Instance`."""

    def __init__(self, inst, parent=None, blen=None, _handle=None):
        self.inst = inst
        if _handle is not None:
            self.h = _handle
            return
        par = np.ascontiguousarray(parent, dtype=np.int32)
        bl = _d(blen)
        assert len(par) == 2 * inst.n_tips - 2 and len(bl) == len(par)
        h = C.c_int(-1)
        _check(lib().gle_tree_create(inst.h, _p(par, C.c_int), _p(bl, C.c_double), C.byref(h)))
        self.h = h.value

    def copy(self):
        h = C.c_int(-1)
        _check(lib().gle_tree_copy(self.h, C.byref(h)))
        return Tree(self.inst, _handle=h.value)

    def close(self):
        if self.h is not None:
            lib().gle_tree_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def invalidate(self):
        _check(lib().gle_tree_invalidate(self.h))

    def score(self):
        v = C.c_double(0)
        _check(lib().gle_tree_score(self.h, C.byref(v)))
        return v.value

    def site_likelihoods(self):
        out = np.zeros(self.inst.n_patterns)
        _check(lib().gle_tree_site_likelihoods(self.h, _p(out, C.c_double)))
        return out

    def derivatives(self, node):
        a, b, c = C.c_double(0), C.c_double(0), C.c_double(0)
        _check(lib().gle_tree_derivatives(self.h, node, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def site_derivatives(self, node):
        """Derivatives on the edge above `node` with their per-pattern values:
        ((lnL, d1, d2), dict of lnl, d1, d2, under_partial, under_child)."""
        return _site_values(lambda *a: lib().gle_tree_site_derivatives(self.h, int(node), *a), self.inst.n_patterns)

    def set_branch_length(self, node, v):
        _check(lib().gle_tree_set_branch_length(self.h, node, float(v)))

    def set_topology(self, parent, blen):
        """A topology move: the tree takes the new parent array and lengths and
        keeps every CLA whose side of the tree is unchanged."""
        par = np.ascontiguousarray(parent, dtype=np.int32)
        bl = _d(blen)
        assert len(par) == 2 * self.inst.n_tips - 2 and len(bl) == len(par)
        _check(lib().gle_tree_set_topology(self.h, _p(par, C.c_int), _p(bl, C.c_double)))

    def cla_updates(self):
        """CLA updates this tree has computed so far."""
        n = C.c_longlong(0)
        _check(lib().gle_tree_cla_updates(self.h, C.byref(n)))
        return n.value

    def branch_lengths(self):
        out = np.zeros(2 * self.inst.n_tips - 2)
        _check(lib().gle_tree_get_branch_lengths(self.h, _p(out, C.c_double)))
        return out

    def optimize_all_branches(self, precision):
        imp, lnl = C.c_double(0), C.c_double(0)
        _check(lib().gle_tree_optimize_all_branches(self.h, float(precision), C.byref(imp), C.byref(lnl)))
        return imp.value, lnl.value

    def sweep(self):
        out = np.zeros(3 * (2 * self.inst.n_tips - 3))
        _check(lib().gle_tree_sweep(self.h, _p(out, C.c_double)))
        return out.reshape(-1, 3)

    def get_cla(self, a, b):
        out = np.zeros(self.inst.n_patterns * self.inst.n_rates * self.inst.n_states)
        u = np.zeros(self.inst.n_patterns, dtype=np.int32)
        _check(lib().gle_tree_get_cla(self.h, a, b, _p(out, C.c_double), _p(u, C.c_int)))
        return out.reshape(self.inst.n_patterns, self.inst.n_rates, self.inst.n_states), u

    def rescale_every(self):
        return lib().gle_tree_rescale_every(self.h)

    def set_rescale_every(self, every):
        _check(lib().gle_tree_set_rescale_every(self.h, int(every)))

    def set_schedule(self, mode):
        _check(lib().gle_tree_set_schedule(self.h, int(mode)))

    def set_gtr(self, params):
        """Set the instance to GTR+I+G parameters (a GtrParams)."""
        c = params.to_c()
        _check(lib().gle_tree_set_gtr(self.h, C.byref(c)))

    def set_codon(self, params):
        """Set the instance to codon-model parameters (a CodonParams)."""
        c = params.to_c()
        _check(lib().gle_tree_set_codon(self.h, C.byref(c)))

    def optimize_codon(self, params, which=None, precision=0.01):
        """Maximise lnL over exchangeabilities, omegas and omega proportions
        (and branch lengths) by default; OPT_FREQS adds the 61 codon
        frequencies (statefrequencies = estimateF).  Returns (lnL, optimised
        CodonParams)."""
        if which is None:
            which = OPT_RATES | OPT_OMEGA | OPT_OMEGA_PROBS | OPT_BRANCHES
        c = params.to_c()
        out = C.c_double(0)
        _check(lib().gle_tree_optimize_codon(self.h, C.byref(c), int(which), float(precision), C.byref(out)))
        return out.value, CodonParams.from_c(c)

    def set_aa(self, params):
        """Set the instance to amino-acid model parameters (an AaParams)."""
        c = params.to_c()
        _check(lib().gle_tree_set_aa(self.h, C.byref(c)))

    def optimize_aa(self, params, which=None, precision=0.01):
        """Maximise lnL over the selected amino-acid model parameters (by
        default the gamma shape, pinv and branch lengths: a named or fixed
        matrix with fixed frequencies); returns (lnL, optimised AaParams)."""
        if which is None:
            which = OPT_ALPHA | OPT_PINV | OPT_BRANCHES
        c = params.to_c()
        out = C.c_double(0)
        _check(lib().gle_tree_optimize_aa(self.h, C.byref(c), int(which), float(precision), C.byref(out)))
        return out.value, AaParams.from_c(c)

    def optimize_gtr(self, params, which=None, precision=0.01):
        """Maximise lnL over the selected model parameters (and all branch
        lengths with OPT_BRANCHES) on the fixed topology; returns
        (lnL, optimised GtrParams)."""
        if which is None:
            which = OPT_RATES | OPT_FREQS | OPT_ALPHA | OPT_PINV | OPT_BRANCHES
        c = params.to_c()
        out = C.c_double(0)
        _check(lib().gle_tree_optimize_gtr(self.h, C.byref(c), int(which), float(precision), C.byref(out)))
        return out.value, GtrParams.from_c(c)


OPT_RATES, OPT_FREQS, OPT_ALPHA, OPT_PINV, OPT_BRANCHES, OPT_OMEGA, OPT_OMEGA_PROBS = 1, 2, 4, 8, 16, 32, 64


class _GtrC(C.Structure):
    _fields_ = [("rel", C.c_double * 5), ("pi", C.c_double * 4), ("alpha", C.c_double), ("pinv", C.c_double),
                ("pinv_in_model", C.c_int)]


class GtrParams:
    """GARLI's nst=6 nucleotide model: rel = (AC, AG, AT, CG, CT) with GT = 1,
    base frequencies, gamma shape and proportion of invariant sites."""

    def __init__(self, rel, pi, alpha=1.0, pinv=0.0, pinv_in_model=False):
        self.rel = np.asarray(rel, dtype=np.float64).copy()
        self.pi = np.asarray(pi, dtype=np.float64).copy()
        self.alpha = float(alpha)
        self.pinv = float(pinv)
        self.pinv_in_model = bool(pinv_in_model)

    def to_c(self):
        c = _GtrC()
        c.rel[:] = [float(x) for x in self.rel]
        c.pi[:] = [float(x) for x in self.pi]
        c.alpha, c.pinv, c.pinv_in_model = self.alpha, self.pinv, int(self.pinv_in_model)
        return c

    @classmethod
    def from_c(cls, c):
        return cls(list(c.rel), list(c.pi), c.alpha, c.pinv, bool(c.pinv_in_model))

    def __repr__(self):
        return "GtrParams(rel=%s, pi=%s, alpha=%.5f, pinv=%.5f)" % (np.round(self.rel, 5), np.round(self.pi, 5),
                                                                     self.alpha, self.pinv)


class _AaC(C.Structure):
    _fields_ = [("rel", C.c_double * 190), ("pi", C.c_double * 20), ("alpha", C.c_double), ("pinv", C.c_double),
                ("pinv_in_model", C.c_int)]


class AaParams:
    """GARLI's amino-acid models: 190 exchangeabilities in upper-triangle row
    order over ACDEFGHIKLMNPQRSTVWY, frequencies, gamma shape and pinv."""

    def __init__(self, rel, pi, alpha=1.0, pinv=0.0, pinv_in_model=False):
        self.rel = np.asarray(rel, dtype=np.float64).copy()
        self.pi = np.asarray(pi, dtype=np.float64).copy()
        if self.rel.shape != (190,) or self.pi.shape != (20,):
            raise ValueError("AaParams needs 190 exchangeabilities and 20 frequencies")
        self.alpha = float(alpha)
        self.pinv = float(pinv)
        self.pinv_in_model = bool(pinv_in_model)

    def to_c(self):
        c = _AaC()
        c.rel[:] = [float(x) for x in self.rel]
        c.pi[:] = [float(x) for x in self.pi]
        c.alpha, c.pinv, c.pinv_in_model = self.alpha, self.pinv, int(self.pinv_in_model)
        return c

    @classmethod
    def from_c(cls, c):
        return cls(list(c.rel), list(c.pi), c.alpha, c.pinv, bool(c.pinv_in_model))

    def __repr__(self):
        return "AaParams(alpha=%.5f, pinv=%.5f)" % (self.alpha, self.pinv)


class _CodonC(C.Structure):
    _fields_ = [("rel", C.c_double * 6), ("n_omega", C.c_int), ("omega", C.c_double * 8),
                ("omega_prob", C.c_double * 8), ("pi", C.c_double * 61)]


class CodonParams:
    """GARLI's codon models: exchangeabilities rel = (AC, AG, AT, CG, CT, GT),
    omega classes with their proportions, fixed codon frequencies pi[61]."""

    def __init__(self, rel, omega, omega_prob, pi):
        self.rel = np.asarray(rel, dtype=np.float64).copy()
        self.omega = np.asarray(omega, dtype=np.float64).copy()
        self.omega_prob = np.asarray(omega_prob, dtype=np.float64).copy()
        self.pi = np.asarray(pi, dtype=np.float64).copy()

    def to_c(self):
        c = _CodonC()
        c.rel[:] = [float(x) for x in self.rel]
        c.n_omega = len(self.omega)
        for k, (w, p) in enumerate(zip(self.omega, self.omega_prob)):
            c.omega[k], c.omega_prob[k] = float(w), float(p)
        c.pi[:] = [float(x) for x in self.pi]
        return c

    @classmethod
    def from_c(cls, c):
        k = c.n_omega
        return cls(list(c.rel), list(c.omega)[:k], list(c.omega_prob)[:k], list(c.pi))

    def __repr__(self):
        return "CodonParams(rel=%s, omega=%s, omega_prob=%s)" % (np.round(self.rel, 5), np.round(self.omega, 5),
                                                                 np.round(self.omega_prob, 5))


REQ_SCORE = 0
REQ_OPTIMIZE = 1


class Server:
    """Continuous-batching request service (gle_serve_*)."""

    def __init__(self):
        h = C.c_int(-1)
        _check(lib().gle_serve_create(C.byref(h)))
        self.h = h.value

    def close(self):
        if self.h is not None:
            lib().gle_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, inst, parent, blen, kind=REQ_SCORE, precision=0.0):
        par = np.ascontiguousarray(parent, dtype=np.int32)
        bl = _d(blen)
        rid = C.c_int(-1)
        _check(lib().gle_serve_submit(self.h, inst.h, _p(par, C.c_int), _p(bl, C.c_double), int(kind),
                                      float(precision), C.byref(rid)))
        return rid.value

    def submit_batch(self, inst, trees, kind=REQ_SCORE, precision=0.0):
        """Submit [(parent, blen), ...] on one instance in one call; returns the ids."""
        nn = 2 * inst.n_tips - 2
        ps, bs = zip(*trees) if len(trees) else ((), ())
        par = np.ascontiguousarray(np.concatenate(ps) if ps else np.zeros(0), dtype=np.int32)
        bl = _d(np.concatenate(bs) if bs else np.zeros(0))
        assert len(par) == nn * len(trees)
        ids = np.zeros(len(trees), dtype=np.int32)
        _check(lib().gle_serve_submit_batch(self.h, inst.h, len(trees), _p(par, C.c_int), _p(bl, C.c_double),
                                            int(kind), float(precision), _p(ids, C.c_int)))
        return [int(i) for i in ids]

    def poll(self, rid):
        d = C.c_int(0)
        _check(lib().gle_serve_poll(self.h, rid, C.byref(d)))
        return bool(d.value)

    def wait(self, rid, n_nodes=None):
        lnl, lat = C.c_double(0), C.c_double(0)
        bl = np.zeros(n_nodes) if n_nodes else None
        _check(lib().gle_serve_wait(self.h, rid, C.byref(lnl), _p(bl, C.c_double) if n_nodes else None,
                                    C.byref(lat)))
        return (lnl.value, lat.value, bl) if n_nodes else (lnl.value, lat.value)

    def wait_all(self, rids):
        """Wait for every request in rids; returns their lnL values in order."""
        ids = np.ascontiguousarray(rids, dtype=np.int32)
        out = np.zeros(len(ids))
        _check(lib().gle_serve_wait_all(self.h, len(ids), _p(ids, C.c_int), _p(out, C.c_double)))
        return out

    def timing(self, reset=False):
        """(host ms building ticks, ms from first launch to results) so far."""
        a, b = C.c_double(0), C.c_double(0)
        _check(lib().gle_serve_timing(self.h, C.byref(a), C.byref(b), int(reset)))
        return {"host_ms": a.value, "device_ms": b.value}

    def work(self, reset=False):
        """CLA work of the ticks so far: ops, children contracted on the device,
        children read from kept messages (codon instances) and messages kept."""
        v = [C.c_longlong(0) for _ in range(4)]
        _check(lib().gle_serve_work(self.h, *[C.byref(x) for x in v], int(reset)))
        return dict(zip(("cla_ops", "contractions", "msg_reads", "msg_keeps"), [x.value for x in v]))

    def set_batch_window(self, microseconds):
        """An idle server waits this long for a burst of submissions to gather."""
        _check(lib().gle_serve_set_batch_window(self.h, int(microseconds)))

    def stats(self, reset=False):
        n = C.c_longlong(0)
        p50, p99, mb = C.c_double(0), C.c_double(0), C.c_double(0)
        _check(lib().gle_serve_stats(self.h, C.byref(n), C.byref(p50), C.byref(p99), C.byref(mb), int(reset)))
        return {"completed": n.value, "p50_ms": p50.value, "p99_ms": p99.value, "mean_batch": mb.value}
