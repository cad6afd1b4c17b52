"""GPU parity of the synonymous-rate classes (hy_eval_srv) and of the
batched many-genes entry (hy_eval_batch) against the FP64 CPU oracle."""
import numpy as np
import pytest

from helpers import branch_lengths, grad_err, make_engine, rel, sim_problem
from oracle import hyoracle as ho

pytestmark = pytest.mark.gpu

FP64, FP32 = 0, 1
OMEGAS, WEIGHTS = [0.1, 0.8, 3.0], [0.6, 0.35, 0.05]


@pytest.mark.parametrize("n,L,seed,conv,M", [(8, 150, 5, 1, 3), (16, 400, 6, 0, 2), (24, 333, 7, 1, 4)])
def test_srv_matches_oracle(n, L, seed, conv, M):
    tree, aln, sim = sim_problem(n, L, seed, omegas=OMEGAS, weights=WEIGHTS, gaps=0.1)
    N = len(tree.nodes)
    bl = branch_lengths(tree) * (1.0 if conv else sim["conv"])
    th = sim["theta"]
    rng = np.random.default_rng(seed)
    rates = rng.uniform(0.2, 3.0, size=M)
    vw = rng.dirichlet(np.ones(M))
    mix = ho.Mixture(OMEGAS, np.zeros(N, dtype=np.int32), [[(k, w) for k, w in enumerate(WEIGHTS)]], conv)
    ref = ho.srv_eval(ho.Problem.from_refs(tree, aln), ho.Code(0), ho.params(1.0, 1.0, th), aln.pos_freqs, aln.cf3x4,
                      mix, bl, rates, vw)
    eng = make_engine(tree, aln)
    eng.set_mixture(3, np.zeros(N, dtype=np.int32), [[0, 1, 2]], convention=conv)
    for prec, tl, tg in ((FP64, 1e-9, 1e-7), (FP32, 1e-5, 1e-4)):
        r = eng.eval_srv(bl, OMEGAS, WEIGHTS, th, rates, vw, precision=prec, site=True)
        assert rel(r["log_likelihood"], ref["log_likelihood"]) < tl, prec
        assert np.max(np.abs(r["site_ll"] - ref["site_ll"])) < tl * 10 * np.max(np.abs(ref["site_ll"])), prec
        for k in ("grad_bl", "grad_omega", "grad_w", "grad_theta", "grad_syn_rates", "grad_syn_weights"):
            assert grad_err(r[k], ref[k]) < tg, (prec, k)


def test_srv_single_class_equals_eval():
    tree, aln, sim = sim_problem(12, 200, 8, omegas=OMEGAS, weights=WEIGHTS)
    N = len(tree.nodes)
    bl = branch_lengths(tree)
    eng = make_engine(tree, aln)
    eng.set_mixture(3, np.zeros(N, dtype=np.int32), [[0, 1, 2]], convention=1)
    a = eng.eval(bl, OMEGAS, WEIGHTS, sim["theta"], precision=FP64)
    b = eng.eval_srv(bl, OMEGAS, WEIGHTS, sim["theta"], [1.0], [1.0], precision=FP64)
    # same kernels; the spectra of each pass are refined from the previous
    # call's (warm start), which moves results at the 1e-10 level
    assert rel(b["log_likelihood"], a["log_likelihood"]) < 1e-12
    for k in ("grad_bl", "grad_omega", "grad_w", "grad_theta"):
        assert grad_err(b[k], a[k]) < 1e-8, k


def batch_genes(group, sizes, seed0, conv=1):
    genes = []
    for g, (n, L, K) in enumerate(sizes):
        om = list(np.linspace(0.1, 2.5, K))
        w = list(np.full(K, 1.0 / K))
        tree, aln, sim = sim_problem(n, L, seed0 + g, omegas=om, weights=w, gaps=0.1 * (g % 2))
        N = len(tree.nodes)
        eng = make_engine(tree, aln, eng=group[g])
        eng.set_mixture(K, np.zeros(N, dtype=np.int32), [list(range(K))], convention=conv)
        eng.set_warm_start(False)
        mix = ho.Mixture(om, np.zeros(N, dtype=np.int32), [[(k, x) for k, x in enumerate(w)]], conv)
        bl = branch_lengths(tree)
        ref = ho.mixture_eval(ho.Problem.from_refs(tree, aln), ho.Code(0), ho.params(1.0, 1.0, sim["theta"]),
                              aln.pos_freqs, aln.cf3x4, mix, bl)
        genes.append((eng, (bl, om, w, sim["theta"]), ref))
    return genes


SIZES = [(5, 40, 1), (9, 130, 2), (17, 75, 3), (33, 260, 4), (7, 1, 2), (12, 500, 1)]


def test_batch_fused_matches_oracle_and_individual():
    """A group's batch runs as one fused evaluation over all genes (one launch
    per stage); every gene matches the oracle and its own hy_eval."""
    from hyphy3.engine import create_group, eval_batch
    group = create_group(len(SIZES), n_stream_sets=2)
    genes = batch_genes(group, SIZES, 300)
    engines = [g[0] for g in genes]
    for prec, tl, tg, tl1, tg1 in ((FP64, 1e-9, 1e-7, 1e-12, 1e-9), (FP32, 1e-5, 1e-4, 1e-6, 1e-5)):
        for rep in range(2):
            res = eval_batch(engines, [g[1] for g in genes], precision=prec, site=True)
            for (eng, prm, ref), r in zip(genes, res):
                assert rel(r["log_likelihood"], ref["log_likelihood"]) < tl, (prec, rep)
                np.testing.assert_allclose(r["site_ll"], ref["site_ll"], rtol=tl * 10)
                for k in ("grad_bl", "grad_omega", "grad_w", "grad_theta"):
                    assert grad_err(r[k], ref[k]) < tg, (prec, k)
        for (eng, prm, ref), r in zip(genes, res):
            one = eng.eval(*prm, precision=prec)
            assert rel(r["log_likelihood"], one["log_likelihood"]) < tl1
            for k in ("grad_bl", "grad_omega", "grad_w", "grad_theta"):
                assert grad_err(r[k], one[k]) < tg1, (prec, k)
    # the reusable flat-array evaluator runs the same fused path
    from hyphy3.engine import Batch
    b = Batch(engines)
    rb = b.eval([g[1] for g in genes], precision=FP64, site=True)
    rf = b.eval_flat(b.bl.copy(), b.om.copy(), b.cw.copy(), b.th.copy(), precision=FP64)
    for (eng, prm, ref), r1, r2 in zip(genes, rb, rf):
        assert rel(r1["log_likelihood"], ref["log_likelihood"]) < 1e-9
        assert rel(r2["log_likelihood"], r1["log_likelihood"]) < 1e-12
        assert grad_err(r1["grad_bl"], ref["grad_bl"]) < 1e-7


def test_batch_engine_by_engine_is_bit_identical():
    """With fusion switched off (HY_NO_FUSE=1 at group creation) a batch is
    each gene's own hy_eval (same kernels; cold spectra, so repeated
    evaluations are bit-identical)."""
    import os
    from hyphy3.engine import create_group, eval_batch
    os.environ["HY_NO_FUSE"] = "1"
    try:
        group = create_group(len(SIZES), n_stream_sets=2)
    finally:
        del os.environ["HY_NO_FUSE"]
    genes = batch_genes(group, SIZES, 300)
    engines = [g[0] for g in genes]
    for prec in (FP64, FP32):
        res = eval_batch(engines, [g[1] for g in genes], precision=prec, site=True)
        for (eng, prm, ref), r in zip(genes, res):
            one = eng.eval(*prm, precision=prec, site=True)
            assert r["log_likelihood"] == one["log_likelihood"]
            assert np.array_equal(r["site_ll"], one["site_ll"])
            for k in ("grad_bl", "grad_omega", "grad_w", "grad_theta"):
                assert np.array_equal(r[k], one[k]), k


def test_batch_mixed_conventions_fall_back():
    """Genes that cannot share one fused problem (different time
    conventions) are evaluated engine by engine, still correct."""
    from hyphy3.engine import create_group, eval_batch
    group = create_group(3, n_stream_sets=2)
    genes = batch_genes(group[:2], SIZES[:2], 700, conv=1) + batch_genes(group[2:], SIZES[2:3], 710, conv=0)
    res = eval_batch([g[0] for g in genes], [g[1] for g in genes], precision=FP64)
    for (eng, prm, ref), r in zip(genes, res):
        assert rel(r["log_likelihood"], ref["log_likelihood"]) < 1e-9
        assert grad_err(r["grad_bl"], ref["grad_bl"]) < 1e-7
