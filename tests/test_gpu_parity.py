"""MI355X engine vs the FP64 CPU oracle on the same seeded inputs (GPU)."""
import numpy as np
import pytest

from helpers import (branch_lengths, grad_err, load_cd2_reduced, make_engine, rel, sim_problem)
from oracle import hyoracle as ho

pytestmark = pytest.mark.gpu

FP64, FP32 = 0, 1
GOLDEN_MG94 = -92.77744887734814


def test_cd2_reduced_point_logl_fp64_fp32():
    tree, aln = load_cd2_reduced()
    eng = make_engine(tree, aln)
    bl = branch_lengths(tree)
    theta = [1, 1, 1, 2, 1]
    prob = ho.Problem.from_refs(tree, aln)
    ref = ho.mg94_loglik(prob, ho.Code(0), ho.params(1.0, 0.5, theta), aln.pos_freqs, aln.cf3x4, bl)
    r64 = eng.eval(bl, [0.5], [1.0], theta, precision=FP64, grad=False, site=True)
    r32 = eng.eval(bl, [0.5], [1.0], theta, precision=FP32, grad=False)
    print("cd2_reduced lnL fp64", r64["log_likelihood"], "fp32", r32["log_likelihood"], "oracle", ref)
    assert rel(r64["log_likelihood"], ref) < 1e-9
    assert rel(r64["log_likelihood"], GOLDEN_MG94) < 1e-8
    assert rel(r32["log_likelihood"], ref) < 1e-5


def test_cd2_reduced_joint_gradients():
    """compute_joint_frechet_gradients semantics (3/mu time, K=1) at the
    reference test's point (tests/test_core.cpp:149-223)."""
    tree, aln = load_cd2_reduced()
    eng = make_engine(tree, aln)
    eng.set_mixture(1, None, [[0]], convention=1)
    th = [0.8, 1.2, 0.6, 1.8, 0.9]
    bl = np.array([0.0 if n.id == tree.root else max(1e-4, n.bl) for n in tree.nodes])
    prob = ho.Problem.from_refs(tree, aln)
    ref = ho.joint_frechet(prob, ho.Code(0), ho.params(1.0, 0.45, th), aln.pos_freqs, aln.cf3x4, bl)
    for prec, tl, tg in ((FP64, 1e-9, 1e-7), (FP32, 1e-5, 1e-4)):
        r = eng.eval(bl, [0.45], [1.0], th, precision=prec)
        print(prec, r["log_likelihood"], ref["log_likelihood"])
        assert rel(r["log_likelihood"], ref["log_likelihood"]) < tl
        assert grad_err(r["grad_bl"], ref["grad_branches"]) < tg
        assert rel(r["grad_omega"][0], ref["grad_omega"]) < tg
        assert grad_err(r["grad_theta"], ref["grad_theta"]) < tg


@pytest.mark.parametrize("K", [1, 3])
def test_sim16_mixture_fp64_fp32(K):
    omegas = [0.5] if K == 1 else [0.1, 0.8, 3.0]
    weights = [1.0] if K == 1 else [0.6, 0.35, 0.05]
    tree, aln, sim = sim_problem(16, 500, 101, omegas=omegas, weights=weights)
    eng = make_engine(tree, aln)
    N = len(tree.nodes)
    eng.set_mixture(K, np.zeros(N, dtype=np.int32), [list(range(K))], convention=0)
    bl = branch_lengths(tree) * sim["conv"]
    prob = ho.Problem.from_refs(tree, aln)
    mix = ho.Mixture(omegas, np.zeros(N, dtype=np.int32), [[(k, w) for k, w in enumerate(weights)]], 0)
    ref = ho.mixture_eval(prob, ho.Code(0), ho.params(1.0, 1.0, sim["theta"]), aln.pos_freqs, aln.cf3x4, mix, bl)
    for prec, tl, tg in ((FP64, 1e-9, 1e-7), (FP32, 1e-5, 1e-4)):
        r = eng.eval(bl, omegas, weights, sim["theta"], precision=prec, site=True)
        print("K", K, "prec", prec, r["log_likelihood"], ref["log_likelihood"])
        assert rel(r["log_likelihood"], ref["log_likelihood"]) < tl
        assert grad_err(r["grad_bl"], ref["grad_bl"]) < tg
        assert grad_err(r["grad_omega"], ref["grad_omega"]) < tg
        assert grad_err(r["grad_w"], ref["grad_w"]) < tg
        assert grad_err(r["grad_theta"], ref["grad_theta"]) < tg
        np.testing.assert_allclose(r["site_ll"], ref["site_ll"], rtol=1e-6 if prec else 1e-11, atol=1e-9)


@pytest.mark.parametrize("T", [16, 32, 64])
def test_tile_widths_agree(T):
    tree, aln, sim = sim_problem(12, 300, 7)
    eng = make_engine(tree, aln)
    eng.set_tile(T)
    bl = branch_lengths(tree)
    prob = ho.Problem.from_refs(tree, aln)
    ll, g = ho.branch_length_gradients(prob, ho.Code(0), ho.params(1.0, 0.5, sim["theta"]), aln.pos_freqs,
                                       aln.cf3x4, bl)
    r = eng.eval(bl, [0.5], [1.0], sim["theta"], precision=FP64)
    assert rel(r["log_likelihood"], ll) < 1e-10
    assert grad_err(r["grad_bl"], g) < 1e-8


def test_warm_start_matches_cold():
    """Warm-started spectra (refined from the previous call) agree with cold
    Jacobi solves; a large parameter jump falls back to the cold solve."""
    omegas, weights = [0.2, 1.5], [0.7, 0.3]
    tree, aln, sim = sim_problem(16, 400, 11, omegas=omegas, weights=weights)
    N = len(tree.nodes)
    engines = []
    for warm in (True, False):
        e = make_engine(tree, aln)
        e.set_mixture(2, np.zeros(N, dtype=np.int32), [[0, 1]], convention=1)
        e.set_warm_start(warm)
        engines.append(e)
    bl = branch_lengths(tree)
    for step in range(5):
        f = 1.0 + 1e-3 * step
        om = [o * f for o in omegas] if step < 4 else [o * 3.0 for o in omegas]
        th = np.asarray(sim["theta"]) * (2.0 - f)
        rw = engines[0].eval(bl, om, weights, th, precision=FP64)
        rc = engines[1].eval(bl, om, weights, th, precision=FP64)
        st = engines[0].eigen_status()
        print("step", step, "status", st, rw["log_likelihood"], rc["log_likelihood"])
        if 0 < step < 4:
            assert (st >= 100).all()
        if step == 4:
            assert (st < 100).all()
        assert rel(rw["log_likelihood"], rc["log_likelihood"]) < 1e-12
        for k in ("grad_bl", "grad_omega", "grad_w", "grad_theta"):
            assert grad_err(rw[k], rc[k]) < 1e-9, k


def test_warm_start_fp32_tolerance():
    """FP32 evaluations accept warm-started spectra at a looser predicted
    correction (1e-10 instead of 1e-14); their results stay within the FP32
    tolerances of cold solves, and FP64 evaluations that follow on the same
    engine are refined to the FP64 tolerance again."""
    omegas, weights = [0.3, 2.0], [0.6, 0.4]
    tree, aln, sim = sim_problem(16, 400, 21, omegas=omegas, weights=weights)
    N = len(tree.nodes)
    engines = []
    for warm in (True, False):
        e = make_engine(tree, aln)
        e.set_mixture(2, np.zeros(N, dtype=np.int32), [[0, 1]], convention=1)
        e.set_warm_start(warm)
        engines.append(e)
    bl = branch_lengths(tree)
    for step in range(6):
        f = 1.0 + 1e-3 * step
        om = [o * f for o in omegas]
        th = np.asarray(sim["theta"]) * (2.0 - f)
        prec = FP32 if step < 4 else FP64
        rw = engines[0].eval(bl, om, weights, th, precision=prec)
        rc = engines[1].eval(bl, om, weights, th, precision=prec)
        st = engines[0].eigen_status()
        print("step", step, "prec", prec, "status", st, rw["log_likelihood"], rc["log_likelihood"])
        if step > 0:
            assert (st >= 100).all()
        tl, tg = (1e-6, 1e-4) if prec == FP32 else (1e-12, 1e-9)
        assert rel(rw["log_likelihood"], rc["log_likelihood"]) < tl
        for k in ("grad_bl", "grad_omega", "grad_w", "grad_theta"):
            assert grad_err(rw[k], rc[k]) < tg, k


def test_multihit_rates_forward():
    """Double and triple nucleotide changes (delta, psi; rate_matrix.hpp:205-219)
    in the forward pass, FP64 and FP32, against the oracle."""
    tree, aln, sim = sim_problem(10, 180, 12)
    N = len(tree.nodes)
    prob = ho.Problem.from_refs(tree, aln)
    bl = branch_lengths(tree) * sim["conv"]
    eng = make_engine(tree, aln)
    eng.set_mixture(1, np.zeros(N, dtype=np.int32), [[0]], convention=0)
    for delta, psi in ((0.3, 0.0), (0.0, 0.15), (0.25, 0.08)):
        p = ho.params(1.0, 0.6, sim["theta"], delta, psi)
        ref = ho.mg94_loglik(prob, ho.Code(0), p, aln.pos_freqs, aln.cf3x4, bl)
        for prec, tl in ((FP64, 1e-10), (FP32, 1e-5)):
            r = eng.eval(bl, [0.6], [1.0], sim["theta"], delta=delta, psi=psi, precision=prec, grad=False)
            assert rel(r["log_likelihood"], ref) < tl, (delta, psi, prec)


def test_lean_inside_on_many_tiles():
    """More FP32 tiles than two per CU switch the inside pass to its lean
    variant (six LDS slots, tips from HBM, four workgroups per CU); results
    match the oracle and the wide-tile path."""
    omegas, weights = [0.2, 1.5], [0.7, 0.3]
    tree, aln, sim = sim_problem(10, 9000, 77, omegas=omegas, weights=weights, gaps=0.05)
    N = len(tree.nodes)
    P = len(aln.weights)
    assert P > 32 * 256  # > 2 tiles of 16 per CU on 256 CUs
    bl = branch_lengths(tree)
    mix = ho.Mixture(omegas, np.zeros(N, dtype=np.int32), [[(k, w) for k, w in enumerate(weights)]], 1)
    ref = ho.mixture_eval(ho.Problem.from_refs(tree, aln), ho.Code(0), ho.params(1.0, 1.0, sim["theta"]),
                          aln.pos_freqs, aln.cf3x4, mix, bl, threads=8)
    eng = make_engine(tree, aln)
    eng.set_mixture(2, np.zeros(N, dtype=np.int32), [[0, 1]], convention=1)
    r = eng.eval(bl, omegas, weights, sim["theta"], precision=FP32)
    assert eng.last_tile() == 16
    assert rel(r["log_likelihood"], ref["log_likelihood"]) < 1e-5
    for k in ("grad_bl", "grad_omega", "grad_w", "grad_theta"):
        assert grad_err(r[k], ref[k]) < 1e-4, k
    eng.set_tile(32)
    r32 = eng.eval(bl, omegas, weights, sim["theta"], precision=FP32)
    assert rel(r["log_likelihood"], r32["log_likelihood"]) < 1e-6


def test_graph_cap_bounds_live_graphs():
    """At most HY_MAX_GRAPHS instantiated evaluation graphs exist in the
    process; engines past the cap launch eagerly with the same results, and
    destroying engines returns their graphs' slots."""
    import gc
    import os
    from hyphy3.engine import graph_count
    gc.collect()
    live0, _ = graph_count()
    os.environ["HY_MAX_GRAPHS"] = str(live0 + 3)
    engines = []
    try:
        tree, aln, sim = sim_problem(12, 150, 77)
        bl = branch_lengths(tree)
        ll_ref, g_ref = ho.branch_length_gradients(ho.Problem.from_refs(tree, aln), ho.Code(0),
                                                   ho.params(1.0, 0.5, sim["theta"]), aln.pos_freqs, aln.cf3x4, bl)
        for i in range(5):
            eng = make_engine(tree, aln)
            engines.append(eng)
            for rep in range(3):
                r = eng.eval(bl, [0.5], [1.0], sim["theta"], precision=FP64)
                assert rel(r["log_likelihood"], ll_ref) < 1e-9 and grad_err(r["grad_bl"], g_ref) < 1e-7, (i, rep)
                live, cap = graph_count()
                assert cap == live0 + 3 and live <= cap, (i, rep, live, cap)
        assert graph_count()[0] == live0 + 3  # the cap was reached: later engines ran eagerly
    finally:
        for e in engines:
            e.close()
        del os.environ["HY_MAX_GRAPHS"]
    assert graph_count()[0] == live0
