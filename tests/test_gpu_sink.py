"""BUSTED's error sink on the MI355X engine (GPU) against the CPU oracle.

Test branches mix omega components with the sink, (1 - eps) sum_k w_k P_k +
eps P_sink (busted.hpp:571-577), background branches use one matrix; logL and
the full gradient (including the sink weight's) in both precisions and every
tile width, plus the argument checks of the sink's API."""
import numpy as np
import pytest

from helpers import branch_lengths, grad_err, make_engine, rel
from oracle import data_ref as dr
from oracle import hyoracle as ho
from test_gpu_sweep import random_alignment, random_newick
from hyphy3.engine import HY_SINK, HY_TIME_CONV, HyError

pytestmark = pytest.mark.gpu

S = len(dr.Code().sense_codons)


@pytest.mark.parametrize("seed,n,L,K", [(0, 12, 40, 2), (1, 31, 97, 3), (2, 48, 130, 1)])
def test_sink_mixture_parity(seed, n, L, K):
    rng = np.random.default_rng(500 + seed)
    tree = dr.TreeRef.from_newick(random_newick(rng, n))
    aln = dr.AlignmentRef(["T%d" % i for i in range(n)], random_alignment(rng, n, L, gaps=0.05 * seed))
    N = len(tree.nodes)
    bc = (rng.uniform(size=N) < 0.4).astype(np.int32)
    omegas = [0.7] + list(rng.uniform(0.05, 4.0, size=K))
    w = rng.dirichlet(np.ones(K))
    eps = 0.03 + 0.05 * seed
    f = rng.dirichlet(np.ones(S))
    theta = rng.uniform(0.5, 2.0, size=5)
    comp_w = np.concatenate([[1.0], (1.0 - eps) * w, [eps]])
    classes = [[(0, 1.0)], [(1 + k, comp_w[1 + k]) for k in range(K)] + [(-1, eps)]]
    mix = ho.Mixture(omegas, bc, classes, 0, sink_freqs=f)
    bl = branch_lengths(tree)
    ref = ho.mixture_eval(ho.Problem.from_refs(tree, aln), ho.Code(0), ho.params(1.0, 1.0, theta), aln.pos_freqs,
                          aln.cf3x4, mix, bl)
    eng = make_engine(tree, aln)
    try:
        eng.set_mixture(K + 1, bc, [[0], list(range(1, K + 1)) + [HY_SINK]])
        eng.set_sink(f)
        for T in (16, 32, 64):
            eng.set_tile(T)
            for prec, tl, tg in ((0, 1e-9, 1e-7), (1, 1e-5, 1e-4)):
                r = eng.eval(bl, omegas, comp_w, theta, precision=prec)
                msg = (seed, n, L, K, T, prec)
                assert rel(r["log_likelihood"], ref["log_likelihood"]) < tl, msg
                for k in ("grad_bl", "grad_omega", "grad_w", "grad_theta"):
                    assert grad_err(r[k], ref[k]) < tg, msg + (k,)
    finally:
        eng.close()


def test_sink_api_checks():
    rng = np.random.default_rng(7)
    n, L = 6, 10
    tree = dr.TreeRef.from_newick(random_newick(rng, n))
    aln = dr.AlignmentRef(["T%d" % i for i in range(n)], random_alignment(rng, n, L, gaps=0.0))
    N = len(tree.nodes)
    eng = make_engine(tree, aln)
    try:
        with pytest.raises(HyError):
            eng.set_mixture(1, np.zeros(N, dtype=np.int32), [[0, HY_SINK]], convention=HY_TIME_CONV)
        eng.set_mixture(1, np.zeros(N, dtype=np.int32), [[0, HY_SINK]])
        with pytest.raises(HyError):  # no sink matrix yet
            eng.eval(branch_lengths(tree), [0.5], [0.9, 0.1], np.ones(5))
        with pytest.raises(HyError):
            eng.set_sink(np.full(S, 2.0))
        with pytest.raises(ValueError):  # 61 values are read: a shorter array is refused on the host
            eng.set_sink(np.full(S - 3, 1.0 / S))
        with pytest.raises(ValueError):  # the 12 position frequencies are not codon frequencies
            eng.set_sink(np.full(12, 1.0 / 12))
        with pytest.raises(ValueError):
            eng.set_model(0, aln.pos_freqs, aln.cf3x4[:-1])
        eng.set_sink(np.full(S, 1.0 / S))
        r = eng.eval(branch_lengths(tree), [0.5], [0.9, 0.1], np.ones(5))
        assert np.isfinite(r["log_likelihood"])
    finally:
        eng.close()


def test_sink_state_reset_by_set_problem():
    """A handle reused for another problem keeps nothing of the previous
    problem's error sink: a sink mixture needs hy_set_sink again, and the
    default one-component mixture has no sink (the E-step accepts it)."""
    from helpers import tree_arrays
    rng = np.random.default_rng(11)
    n, L = 8, 24
    tree = dr.TreeRef.from_newick(random_newick(rng, n))
    aln = dr.AlignmentRef(["T%d" % i for i in range(n)], random_alignment(rng, n, L, gaps=0.0))
    N = len(tree.nodes)
    bl = branch_lengths(tree)
    eng = make_engine(tree, aln)
    try:
        eng.set_mixture(1, np.zeros(N, dtype=np.int32), [[0, HY_SINK]])
        eng.set_sink(np.full(S, 1.0 / S))
        assert np.isfinite(eng.eval(bl, [0.5], [0.9, 0.1], np.ones(5))["log_likelihood"])
        tree2 = dr.TreeRef.from_newick(random_newick(rng, n + 3))
        aln2 = dr.AlignmentRef(["T%d" % i for i in range(n + 3)], random_alignment(rng, n + 3, L + 5, gaps=0.0))
        parent, children, l2t = tree_arrays(tree2, aln2)
        eng.set_problem(parent, children, l2t, aln2.states, aln2.pmasks, aln2.weights, S=S)
        eng.set_model(0, aln2.pos_freqs, aln2.cf3x4)
        bl2 = branch_lengths(tree2)
        # (b) default mixture: no sink, so the expectation step is available
        w, v = eng.busted_expected_weights(bl2, [0.5], [1.0], np.ones(5), test_class=0)
        assert np.allclose(w, [1.0]) and np.allclose(v, [1.0])
        ref = ho.mixture_eval(ho.Problem.from_refs(tree2, aln2), ho.Code(0), ho.params(1.0, 0.5, np.ones(5)),
                              aln2.pos_freqs, aln2.cf3x4,
                              ho.Mixture([0.5], np.zeros(len(tree2.nodes), dtype=np.int32), [[(0, 1.0)]], 0), bl2)
        r = eng.eval(bl2, [0.5], [1.0], np.ones(5), precision=0)
        assert rel(r["log_likelihood"], ref["log_likelihood"]) < 1e-9
        # (a) a sink mixture on the new problem without hy_set_sink is refused
        eng.set_mixture(1, np.zeros(len(tree2.nodes), dtype=np.int32), [[0, HY_SINK]])
        with pytest.raises(HyError):
            eng.eval(bl2, [0.5], [0.9, 0.1], np.ones(5))
    finally:
        eng.close()
