"""Checkpointed evaluations (GPU): with a memory budget below what every
node's partials need, the engine keeps only the top part's and the part
roots' inside partials and recomputes each part's interior before that part's
outside pass and adjoint.  Results against the FP64 oracle (FP64 1e-9 / 1e-7,
FP32 1e-5 / 1e-4) and against the same engine's resident layout; the partial
buffers shrink; the workspace read-back is refused while checkpointing."""
import numpy as np
import pytest

from helpers import branch_lengths, grad_err, make_engine, rel, sim_problem
from oracle import data_ref as dr
from oracle import hyoracle as ho
from test_gpu_sweep import random_alignment, random_newick
from hyphy3.engine import HY_TIME_CONV, HyError

pytestmark = pytest.mark.gpu

FP64, FP32 = 0, 1
KEYS = ("grad_bl", "grad_omega", "grad_w", "grad_theta")
OMEGAS, WEIGHTS = [0.1, 0.8, 3.0], [0.6, 0.35, 0.05]


def problems():
    tree, aln, sim = sim_problem(128, 600, 414, omegas=OMEGAS, weights=WEIGHTS, gaps=0.1)
    yield "128 taxa x 600 codons", tree, aln, sim["theta"]
    rng = np.random.default_rng(415)
    tree = dr.TreeRef.from_newick(random_newick(rng, 90))  # with polytomies (identity edges)
    aln = dr.AlignmentRef(["T%d" % i for i in range(90)], random_alignment(rng, 90, 140, gaps=0.15))
    yield "90 taxa with polytomies", tree, aln, rng.uniform(0.5, 2.0, size=5)


@pytest.mark.parametrize("case", [0, 1])
def test_checkpointed_matches_oracle_and_resident(case):
    label, tree, aln, theta = list(problems())[case]
    N = len(tree.nodes)
    bl = branch_lengths(tree)
    mix = ho.Mixture(OMEGAS, np.zeros(N, dtype=np.int32), [[(k, w) for k, w in enumerate(WEIGHTS)]], 1)
    ref = ho.mixture_eval(ho.Problem.from_refs(tree, aln), ho.Code(0), ho.params(1.0, 1.0, theta), aln.pos_freqs,
                          aln.cf3x4, mix, bl, threads=8)
    res, ck = make_engine(tree, aln), make_engine(tree, aln)
    try:
        for e in (res, ck):
            e.set_mixture(3, np.zeros(N, dtype=np.int32), [[0, 1, 2]], convention=HY_TIME_CONV)
        ck.set_memory_budget(1)  # below any layout: checkpoint whatever the size
        for prec, tl, tg in ((FP64, 1e-9, 1e-7), (FP32, 1e-5, 1e-4)):
            print(f"[checkpoint] {label} precision {prec}", flush=True)
            r0 = res.eval(bl, OMEGAS, WEIGHTS, theta, precision=prec)
            r1 = ck.eval(bl, OMEGAS, WEIGHTS, theta, precision=prec)
            n_parts, ck_bytes = ck.last_checkpoint()
            n0, res_bytes = res.last_checkpoint()
            print(f"[checkpoint] parts {n_parts}, partial bytes {ck_bytes} (resident {res_bytes})", flush=True)
            assert n0 == 0 and n_parts >= 2 and ck_bytes < res_bytes
            assert rel(r1["log_likelihood"], ref["log_likelihood"]) < tl, (label, prec)
            for k in KEYS:
                assert grad_err(r1[k], ref[k]) < tg, (label, prec, k)
            # same arithmetic per node as the resident layout, other launch grouping
            assert rel(r1["log_likelihood"], r0["log_likelihood"]) < tl
            for k in KEYS:
                assert grad_err(r1[k], r0[k]) < tg, (label, prec, k)
        # repeated (graph replay) and forward-only evaluations
        a = ck.eval(bl, OMEGAS, WEIGHTS, theta, precision=FP64)
        b = ck.eval(bl, OMEGAS, WEIGHTS, theta, precision=FP64, grad=False)
        assert rel(a["log_likelihood"], ref["log_likelihood"]) < 1e-9
        assert rel(b["log_likelihood"], ref["log_likelihood"]) < 1e-9
        # synonymous-rate classes run through the same passes
        sr, sv = [0.5, 1.7], [0.7, 0.3]
        s0 = res.eval_srv(bl, OMEGAS, WEIGHTS, theta, sr, sv, precision=FP64)
        s1 = ck.eval_srv(bl, OMEGAS, WEIGHTS, theta, sr, sv, precision=FP64)
        assert rel(s1["log_likelihood"], s0["log_likelihood"]) < 1e-9
        assert grad_err(s1["grad_bl"], s0["grad_bl"]) < 1e-7
        # the workspace read-back needs every node's partials
        P = np.broadcast_to(np.eye(61), (N, 61, 61)).copy()
        with pytest.raises(HyError):
            ck.inside_outside_from_P(P, aln.cf3x4)
        # back to the resident layout when the budget allows it
        ck.set_memory_budget(0)
        r2 = ck.eval(bl, OMEGAS, WEIGHTS, theta, precision=FP64)
        assert ck.last_checkpoint()[0] == 0
        assert rel(r2["log_likelihood"], ref["log_likelihood"]) < 1e-9
    finally:
        res.close()
        ck.close()
