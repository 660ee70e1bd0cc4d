"""gle_edge_likelihoods_noalloc against gle_edge_likelihoods.

The allocation-free entry point takes its device buffers from the library's
per-device scratch array.  Its results must equal the allocating entry
point's bit for bit (same kernels, same reduction order), on two calls in a
row, the second with a smaller input, so that anything the first call left in
the scratch array cannot show through."""
import numpy as np
import pytest

from garli_amd import engine as E
from garli_amd import sim
import oplist

pytestmark = pytest.mark.gpu


def _swept(ds, prec):
    inst = ds.make_instance(prec, n_cla_slots=3 * (ds.n_tips - 2), n_pmat_slots=2 * ds.n_tips)
    ops, ls, slot_of, pmat_of = oplist.full_sweep(ds.parent, ds.n_tips,
                                                  rescale_every=16 if prec == E.PREC_FP64 else 1)
    nodes = sorted(pmat_of)
    inst.update_transition_matrices([pmat_of[n] for n in nodes], ds.blen[nodes], derivs=True)
    inst.update_partials(ops, ls)
    return inst, oplist.edge_ops(ds.parent, ds.n_tips, slot_of, pmat_of)


@pytest.mark.parametrize("prec", [E.PREC_FP64, E.PREC_FP32])
def test_noalloc_equals_allocating_entry_point(gpu, prec):
    big = sim.gtr_gamma(n_tips=40, n_sites=20000, pinv=0.2, seed=501)
    small = sim.gtr_gamma(n_tips=9, n_sites=700, pinv=0.1, seed=502)
    inst_b, eops_b = _swept(big, prec)
    inst_s, eops_s = _swept(small, prec)
    if prec == E.PREC_FP32:
        assert inst_b.heavy_patterns() > 0, "the FP64 heavy part is not exercised"
    print("case prec", prec, "big", len(eops_b), "ops x", big.n_patterns, "small", len(eops_s), "x",
          small.n_patterns, flush=True)
    # first call: every edge of the large instance; second: a few edges of the small one
    a1 = inst_b.edge_likelihoods_noalloc(eops_b)
    r1 = inst_b.edge_likelihoods(eops_b)
    np.testing.assert_array_equal(a1, r1)
    sub = eops_s[:5]
    a2 = inst_s.edge_likelihoods_noalloc(sub)
    r2 = inst_s.edge_likelihoods(sub)
    np.testing.assert_array_equal(a2, r2)
    # site values (no heavy part in the small instance)
    a3, s3 = inst_s.edge_likelihoods_noalloc(sub[:1], want_sites=True)
    r3, t3 = inst_s.edge_likelihoods(sub[:1], want_sites=True)
    np.testing.assert_array_equal(a3, r3)
    np.testing.assert_array_equal(s3, t3)
    # and the large one again after the small one, scoring only edges
    lnl_only = [(p, c, m, 0) for (p, c, m, _) in eops_b[::7]]
    np.testing.assert_array_equal(inst_b.edge_likelihoods_noalloc(lnl_only), inst_b.edge_likelihoods(lnl_only))
    inst_b.close()
    inst_s.close()


def test_noalloc_refuses_oversized_lists(gpu):
    ds = sim.gtr_gamma(n_tips=6, n_sites=200, seed=503)
    inst, eops = _swept(ds, E.PREC_FP64)
    with pytest.raises(E.GleError) as ei:
        inst.edge_likelihoods_noalloc(eops * 200)  # 1800 ops > 1024
    assert ei.value.code == -1 and b"at most" in E.lib().gle_last_error()
    np.testing.assert_array_equal(inst.edge_likelihoods_noalloc(eops), inst.edge_likelihoods(eops))
