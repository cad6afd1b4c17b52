"""Measured tile/parts autotuning (hy_autotune, hy_set_tuning): the chosen
configuration is the one later evaluations use, results stay within the
oracle tolerances under it, and hy_set_tile still takes precedence."""
import numpy as np
import pytest

from helpers import branch_lengths, grad_err, make_engine, rel, sim_problem
from oracle import hyoracle as ho

pytestmark = pytest.mark.gpu


def test_autotune_choice_is_used_and_exact():
    from hyphy3.engine import HY_PREC_FP32, HY_PREC_FP64
    omegas, weights = [0.2, 1.4], [0.75, 0.25]
    tree, aln, sim = sim_problem(24, 900, 61, omegas=omegas, weights=weights)
    N = len(tree.nodes)
    bl = branch_lengths(tree)
    mix = ho.Mixture(omegas, np.zeros(N, dtype=np.int32), [[(0, weights[0]), (1, weights[1])]], 1)
    ref = ho.mixture_eval(ho.Problem.from_refs(tree, aln), ho.Code(0), ho.params(1.0, 1.0, sim["theta"]),
                          aln.pos_freqs, aln.cf3x4, mix, bl)
    eng = make_engine(tree, aln)
    eng.set_mixture(2, np.zeros(N, dtype=np.int32), [[0, 1]], convention=1)
    for prec, tl, tg in ((HY_PREC_FP32, 1e-5, 1e-4), (HY_PREC_FP64, 1e-9, 1e-7)):
        tile, parts, ms = eng.autotune(bl, omegas, weights, sim["theta"], precision=prec, reps=2)
        print("precision", prec, "tile", tile, "parts", parts, "ms %.3f" % ms)
        assert tile in ((16, 32, 64) if prec == HY_PREC_FP32 else (16, 32)) and parts in (0, 1) and ms > 0.0
        r = eng.eval(bl, omegas, weights, sim["theta"], precision=prec)
        assert eng.last_tile() == tile
        assert (eng.last_parts() > 0) == (parts == 1)
        assert rel(r["log_likelihood"], ref["log_likelihood"]) < tl
        for k in ("grad_bl", "grad_omega", "grad_w", "grad_theta"):
            assert grad_err(r[k], ref[k]) < tg, (prec, k)
    # an explicit tile overrides the tuning; 0 returns to it
    eng.set_tile(64)
    eng.eval(bl, omegas, weights, sim["theta"], precision=HY_PREC_FP32)
    assert eng.last_tile() == 64
    # FP64 has no 64-pattern kernels (their outside pass would spill): 32
    r = eng.eval(bl, omegas, weights, sim["theta"], precision=HY_PREC_FP64)
    assert eng.last_tile() == 32
    assert rel(r["log_likelihood"], ref["log_likelihood"]) < 1e-9
    eng.set_tile(0)
    eng.set_tuning(HY_PREC_FP64, 64, 1)
    eng.eval(bl, omegas, weights, sim["theta"], precision=HY_PREC_FP64)
    assert eng.last_tile() == 32
    eng.set_tuning(HY_PREC_FP32, 32, 0)
    r = eng.eval(bl, omegas, weights, sim["theta"], precision=HY_PREC_FP32)
    assert eng.last_tile() == 32 and eng.last_parts() == 0
    assert rel(r["log_likelihood"], ref["log_likelihood"]) < 1e-5
    eng.close()


def test_batch_autotune_choice_is_used():
    """hy_autotune_batch: a fused batch of genes keeps its measured tile and
    its results stay equal to per-gene evaluations."""
    from hyphy3.engine import HY_PREC_FP32, HY_PREC_FP64, Batch, create_group
    probs = [sim_problem(n, L, 90 + i) for i, (n, L) in enumerate(((8, 120), (14, 300), (11, 90), (20, 200)))]
    engines = create_group(len(probs))
    params = []
    for (tree, aln, sim), eng in zip(probs, engines):
        make_engine(tree, aln, eng=eng)
        N = len(tree.nodes)
        eng.set_mixture(1, np.zeros(N, dtype=np.int32), [[0]], convention=1)
        params.append((branch_lengths(tree), [0.4], [1.0], sim["theta"]))
    batch = Batch(engines)
    for prec, tol in ((HY_PREC_FP32, 1e-5), (HY_PREC_FP64, 1e-12)):
        batch.eval(params, precision=prec)
        tile, ms = batch.autotune(precision=prec, reps=1)
        assert tile in ((16, 32, 64) if prec == HY_PREC_FP32 else (16, 32)) and ms > 0.0
        out = batch.eval(params, precision=prec)
        for (tree, aln, sim), prm, r in zip(probs, params, out):
            ref = make_engine(tree, aln)
            ref.set_mixture(1, np.zeros(len(tree.nodes), dtype=np.int32), [[0]], convention=1)
            rr = ref.eval(*prm, precision=HY_PREC_FP64)
            assert rel(r["log_likelihood"], rr["log_likelihood"]) < max(tol, 1e-9)
            assert grad_err(r["grad_bl"], rr["grad_bl"]) < (1e-4 if prec == HY_PREC_FP32 else 1e-7)
            ref.close()
    for e in engines:
        e.close()
