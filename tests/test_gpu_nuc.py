"""Nucleotide path on the GPU against the FP64 oracle.

FP64 drop-in: lnL and derivatives within 1e-9 relative of the oracle, CLAs
within 1e-12 relative, rescale counts bit-exact.  FP32 bulk: |dlnL| < 1e-3.
"""
import numpy as np
import pytest

from garli_amd import engine as E
from garli_amd import sim
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("seed,ntax,nsites,pinv,ncat", [
    (11, 5, 200, 0.0, 4), (12, 12, 800, 0.2, 4), (13, 24, 2000, 0.0, 4), (14, 9, 300, 0.3, 1),
    (15, 16, 500, 0.1, 3), (16, 10, 400, 0.0, 6)])
def test_fp64_score_matches_oracle(gpu, seed, ntax, nsites, pinv, ncat):
    ds = sim.gtr_gamma(n_tips=ntax, n_sites=nsites, pinv=pinv, ncat=ncat, seed=seed)
    print("case", seed, ntax, ds.n_patterns, flush=True)
    inst = ds.make_instance(E.PREC_FP64)
    t = E.Tree(inst, ds.parent, ds.blen)
    lnl = t.score()
    ref, site_ref, _ = O.Oracle(ds).score(sites=True)
    assert _rel(lnl, ref) < 1e-9, (lnl, ref)
    site = t.site_likelihoods()
    np.testing.assert_allclose(site, site_ref, rtol=1e-10, atol=1e-10)


def test_fp64_every_cla_and_count(gpu):
    """Deep tree with long branches forces rescaling; every directional CLA and
    its integer rescale counts match the oracle (counts bit-exact)."""
    ds = sim.gtr_gamma(n_tips=40, n_sites=600, mean_blen=0.6, seed=21)
    inst = ds.make_instance(E.PREC_FP64)
    t = E.Tree(inst, ds.parent, ds.blen)
    o = O.Oracle(ds, own_model=False)  # same eigensystem: element-wise comparison
    nn = 2 * ds.n_tips - 2
    nb = {n: [] for n in range(nn)}
    for n in range(1, nn):
        nb[n].append(int(ds.parent[n]))
        nb[int(ds.parent[n])].append(n)
    total_rescaled = 0
    for a in range(nn):
        if 1 <= a <= ds.n_tips:
            continue
        for b in nb[a]:
            g, gu = t.get_cla(a, b)
            r, ru = o.cla(a, b)
            np.testing.assert_array_equal(gu, ru)
            np.testing.assert_allclose(g, r, rtol=1e-12, atol=0)
            total_rescaled += int((ru > 0).sum())
    assert total_rescaled > 0, "workload did not exercise rescaling"
    assert _rel(t.score(), o.score()) < 1e-9


def test_fp64_derivatives_every_edge(gpu):
    ds = sim.gtr_gamma(n_tips=14, n_sites=700, pinv=0.15, seed=31)
    inst = ds.make_instance(E.PREC_FP64)
    t = E.Tree(inst, ds.parent, ds.blen)
    o = O.Oracle(ds)
    for n in range(1, 2 * ds.n_tips - 2):
        l, d1, d2 = t.derivatives(n)
        rl, r1, r2 = o.derivs(n)
        assert _rel(l, rl) < 1e-9
        assert abs(d1 - r1) <= 1e-9 * max(1.0, abs(r1))
        assert abs(d2 - r2) <= 1e-9 * max(1.0, abs(r2))


def test_sweep_matches_per_edge(gpu):
    ds = sim.gtr_gamma(n_tips=20, n_sites=900, pinv=0.1, seed=41)
    inst = ds.make_instance(E.PREC_FP64)
    t = E.Tree(inst, ds.parent, ds.blen)
    sw = t.sweep()
    o = O.Oracle(ds)
    for n in range(1, 2 * ds.n_tips - 2):
        rl, r1, r2 = o.derivs(n)
        assert _rel(sw[n - 1, 0], rl) < 1e-9
        assert abs(sw[n - 1, 1] - r1) <= 1e-9 * max(1.0, abs(r1))
        assert abs(sw[n - 1, 2] - r2) <= 1e-9 * max(1.0, abs(r2))
    # root invariance: every edge gives the same total lnL (RunTests, population.cpp:924-954)
    assert np.ptp(sw[:, 0]) < 1e-9 * abs(sw[0, 0])


@pytest.mark.parametrize("mode", [0, 1])
def test_fp32_bulk_within_tolerance(gpu, mode):
    ds = sim.gtr_gamma(n_tips=32, n_sites=4000, pinv=0.2, mean_blen=0.08, seed=51)
    inst = ds.make_instance(E.PREC_FP32)
    t = E.Tree(inst, ds.parent, ds.blen)
    t.set_schedule(mode)
    ref = O.Oracle(ds).score()
    assert abs(t.score() - ref) < 1e-3
    sw = t.sweep()
    o = O.Oracle(ds)
    for n in (1, 5, 17, 40):
        rl, r1, r2 = o.derivs(n)
        assert abs(sw[n - 1, 0] - rl) < 1e-3
        assert abs(sw[n - 1, 1] - r1) < 1e-3 * max(1.0, abs(r1))
        assert abs(sw[n - 1, 2] - r2) < 1e-3 * max(1.0, abs(r2))


def test_fp32_heavy_pattern_split(gpu):
    """Heavily weighted patterns run in FP64 inside an FP32 instance; per-site
    results come back in the caller's pattern order whatever the split."""
    ds = sim.gtr_gamma(n_tips=24, n_sites=40000, pinv=0.3, mean_blen=0.05, seed=52)
    ref = O.Oracle(ds).score()
    f64 = E.Tree(ds.make_instance(E.PREC_FP64), ds.parent, ds.blen)
    f64.score()
    site64 = f64.site_likelihoods()
    inst = ds.make_instance(E.PREC_FP32)
    inst.set_heavy_split(4)
    n_heavy = inst.heavy_patterns()
    assert 0 < n_heavy < ds.n_patterns
    t = E.Tree(inst, ds.parent, ds.blen)
    got = t.score()
    assert abs(got - ref) < 1e-3
    site = t.site_likelihoods()
    heavy = ds.counts > 4
    np.testing.assert_allclose(site[heavy], site64[heavy], rtol=1e-12)
    np.testing.assert_allclose(site, site64, rtol=1e-5)
    # switching the split off and on again re-partitions in place
    inst.set_heavy_split(0)
    assert inst.heavy_patterns() == 0
    t.invalidate()
    all32 = t.score()
    assert abs(all32 - ref) < 0.1
    inst.set_heavy_split(4)
    assert inst.heavy_patterns() == n_heavy
    t.invalidate()
    assert t.score() == got


@pytest.mark.parametrize("prec", [E.PREC_FP64, E.PREC_FP32])
def test_fused_schedule_equals_levels(gpu, prec):
    ds = sim.gtr_gamma(n_tips=30, n_sites=1500, seed=61)
    inst = ds.make_instance(prec)
    inst.set_fused_threshold(0)  # force the single-launch walk at this size
    a = E.Tree(inst, ds.parent, ds.blen)
    b = E.Tree(inst, ds.parent, ds.blen)
    b.set_schedule(1)
    np.testing.assert_array_equal(a.sweep(), b.sweep())


def test_newton_pass_matches_oracle(gpu):
    ds = sim.gtr_gamma(n_tips=12, n_sites=600, seed=71)
    inst = ds.make_instance(E.PREC_FP64)
    t = E.Tree(inst, ds.parent, ds.blen)
    imp, lnl = t.optimize_all_branches(0.001)
    oimp, olnl, obl = O.Oracle(ds).optimize_all(0.001)
    np.testing.assert_allclose(t.branch_lengths()[1:], obl[1:], rtol=1e-7, atol=1e-10)
    assert _rel(lnl, olnl) < 1e-9
    assert lnl >= O.Oracle(ds).score() - 1e-6  # optimisation never worsens lnL


def test_rescale_retry_matches_oracle(gpu):
    """Discordant tips on near-zero branches drive site maxima below 1e-190
    between rescales; with rescaleEvery raised from the host the engine reports
    the reference's throw(1) and Tree::Score's retry lowers rescaleEvery by 2
    until the tree scores (tree.cpp:4077-4090)."""
    ds = sim.gtr_gamma(n_tips=64, n_sites=40, mean_blen=0.1, seed=81, ambig_frac=0.0)
    rng = np.random.default_rng(0)
    ds.tips[:] = (1 << rng.integers(0, 4, size=ds.tips.shape)).astype(np.uint8)
    blen = ds.blen.copy()
    blen[1:] = 1e-8
    inst = ds.make_instance(E.PREC_FP64)
    t = E.Tree(inst, ds.parent, blen)
    t.set_rescale_every(80)
    lnl = t.score()
    ref, _, re = O.Oracle(ds).score(blen=blen, rescale_every=80, sites=True)
    assert re < 80
    assert t.rescale_every() == re
    assert _rel(lnl, ref) < 1e-9


def test_copy_on_write(gpu):
    ds = sim.gtr_gamma(n_tips=16, n_sites=500, seed=91)
    inst = ds.make_instance(E.PREC_FP64, n_cla_slots=200)
    a = E.Tree(inst, ds.parent, ds.blen)
    la = a.score()
    b = a.copy()
    b.set_branch_length(5, 0.3)
    lb = b.score()
    assert a.score() == la
    bl = ds.blen.copy()
    bl[5] = 0.3
    assert _rel(lb, O.Oracle(ds).score(blen=bl)) < 1e-9
    assert _rel(la, O.Oracle(ds).score()) < 1e-9


def test_out_of_cla_slots_then_recovers(gpu):
    """Two trees on an instance with CLA slots for about one score: the second
    tree's score runs out of slots (GLE_ERR_NOMEM) part way through its op
    list.  The slots it had taken go back and nothing it named stays marked
    computed, so once the first tree is closed the second scores equal to the
    oracle, and so do a third and the first tree's topology again."""
    ds = sim.gtr_gamma(n_tips=14, n_sites=500, seed=93)
    inst = ds.make_instance(E.PREC_FP64, n_cla_slots=ds.n_tips + 2, n_pmat_slots=4 * ds.n_tips)
    o = O.Oracle(ds)
    t1 = E.Tree(inst, ds.parent, ds.blen)
    l1 = t1.score()
    assert _rel(l1, o.score()) < 1e-9
    rng = np.random.default_rng(4)
    p2, b2 = sim.perturb(ds.parent, ds.blen, rng, ds.n_tips, 3)
    t2 = E.Tree(inst, p2, b2)
    with pytest.raises(E.GleError) as ei:
        t2.score()
    assert ei.value.code == -3
    t1.close()
    assert _rel(t2.score(), o.score(parent=p2, blen=b2)) < 1e-9
    t2.close()
    p3, b3 = sim.perturb(ds.parent, ds.blen, rng, ds.n_tips, 2)
    t3 = E.Tree(inst, p3, b3)
    assert _rel(t3.score(), o.score(parent=p3, blen=b3)) < 1e-9
    t3.close()
    t4 = E.Tree(inst, ds.parent, ds.blen)
    assert t4.score() == l1


@pytest.mark.parametrize("prec", [E.PREC_FP64, E.PREC_FP32])
def test_graph_replay_equals_eager(gpu, prec):
    """Repeated sweeps replay the captured level launches; every replay gives
    the eager result bit for bit, with eager work and syncs in between."""
    ds = sim.gtr_gamma(n_tips=24, n_sites=3000, pinv=0.2, seed=63)
    inst = ds.make_instance(prec)
    t = E.Tree(inst, ds.parent, ds.blen)
    inst.set_graph_replay(False)
    eager = t.sweep()
    t.invalidate()
    eager_score = t.score()
    inst.set_graph_replay(True)
    for k in range(5):
        np.testing.assert_array_equal(t.sweep(), eager)
        t.invalidate()
        assert t.score() == eager_score
    replays = inst.set_graph_replay(True)
    assert replays > 0
