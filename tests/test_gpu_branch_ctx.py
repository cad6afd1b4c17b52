"""aBSREL's local projection on the GPU (hy_branch_eval) against the
oracle's restatement of compute_branch_context_matrices and
eval_branch_ll_from_matrices (absrel.hpp:438-455, 723-780): internal and tip
branches, gaps, ambiguity codes and an unmatched leaf, several candidates per
call (the branch's own matrix, another MG94 matrix, an arbitrary positive
matrix, the zero matrix), per-pattern products, both precisions, and the
refusals when no gradient evaluation left the partials."""
import numpy as np
import pytest

from helpers import branch_lengths, make_engine, rel, sim_problem
from oracle import data_ref as dr
from oracle import hyoracle as ho

pytestmark = pytest.mark.gpu

FP64, FP32 = 0, 1


def raw_stack(tree, aln, omega, theta, bl):
    """P(bl) of every branch under the engine's raw time convention."""
    p = ho.params(1.0, omega, theta)
    Q, _ = ho.mg94_Q(ho.Code(0), p, aln.pos_freqs, aln.cf3x4)
    ev, U, V = ho.decompose(Q, aln.cf3x4)
    return np.stack([np.eye(61) if n.id == tree.root else ho.transition(ev, U, V, bl[n.id]) for n in tree.nodes])


def candidates(aln, theta, rng, own):
    p = ho.params(1.0, 2.2, theta)
    Q, _ = ho.mg94_Q(ho.Code(0), p, aln.pos_freqs, aln.cf3x4)
    ev, U, V = ho.decompose(Q, aln.cf3x4)
    return np.stack([own, ho.transition(ev, U, V, 0.23), rng.uniform(0.05, 1.0, (61, 61)), np.zeros((61, 61))])


def check_problem(tree, aln, omega, theta, seed, nodes_per_kind=2):
    bl = branch_lengths(tree)
    prob = ho.Problem.from_refs(tree, aln)
    P = raw_stack(tree, aln, omega, theta, bl)
    pi = aln.cf3x4
    eng = make_engine(tree, aln)
    rng = np.random.default_rng(seed)
    internal = [n.id for n in tree.nodes if not n.is_leaf and n.id != tree.root][:nodes_per_kind]
    tips = [n.id for n in tree.nodes if n.is_leaf][:nodes_per_kind]
    for prec, tol in ((FP64, 1e-9), (FP32, 1e-5)):
        r = eng.eval(bl, [omega], [1.0], theta, precision=prec)
        for node in internal + tips:
            Vm, Dm, off = ho.branch_context(prob, P, pi, node)
            C = candidates(aln, theta, rng, P[node])
            lnl, L, e = eng.branch_eval(node, C, precision=prec, site=True)
            assert rel(lnl[0], r["log_likelihood"]) < tol, (prec, node, lnl[0], r["log_likelihood"])
            for c in range(len(C)):
                ref = ho.branch_ll(prob, C[c], Vm, Dm, off)
                assert rel(lnl[c], ref) < tol, (prec, node, c, lnl[c], ref)
            # per-pattern products: ln L_c,p against the oracle's ln(V_p . P_c D_p) plus
            # the pattern's log-scale offset (the zero candidate has no logarithm)
            logsc = np.zeros(prob.num_patterns)
            for p0 in range(0, prob.num_patterns, 256):
                ws = ho.workspace(prob, P, pi, p0, min(p0 + 256, prob.num_patterns), outside=True)
                logsc[p0:p0 + len(ws["L"])] = ws["logL"] - np.log(ws["L"])
            for c in range(3):
                ref_p = np.log(np.einsum("sp,sp->p", Vm, C[c] @ Dm)) + logsc
                got_p = np.log(L[c]) + e * np.log(2.0)
                np.testing.assert_allclose(got_p, ref_p, rtol=0, atol=(1e-9 if prec == FP64 else 1e-5) * 50)
            assert np.all(L[3] == 0.0)
    eng.close()


def test_branch_eval_simulated_with_gaps():
    tree, aln, sim = sim_problem(14, 120, 61, gaps=0.15)
    check_problem(tree, aln, 0.5, sim["theta"], 7)


def test_branch_eval_ambiguity_and_unmatched_leaf():
    seqs = {
        "A": "ATGAAACCCGGGTTTCATNNNAGA",
        "B": "ATGAARCCYGGG---CATACGAGA",
        "C": "ATRAAACC?GGGTTTCA-ACGTAA",
        "D": "ATGAAACCCGGNTTTCATACGAGA",
        "E": "AYGAAACCCGGGTTKCATACGWGA",
    }
    aln = dr.AlignmentRef(list(seqs), list(seqs.values()))
    tree = dr.TreeRef.from_newick("((A:0.1,B:0.2):0.05,(C:0.15,D:0.1):0.2,(E:0.3,Missing:0.2):0.1);")
    check_problem(tree, aln, 0.4, [1, 1, 1, 2, 1], 8, nodes_per_kind=6)


def test_branch_eval_refusals():
    from hyphy3.engine import HyError
    tree, aln, sim = sim_problem(8, 40, 62)
    bl = branch_lengths(tree)
    eng = make_engine(tree, aln)
    node = [n.id for n in tree.nodes if n.is_leaf][0]
    C = np.eye(61)
    with pytest.raises(HyError):  # no evaluation yet
        eng.branch_eval(node, C, precision=FP64)
    eng.eval(bl, [0.5], [1.0], sim["theta"], precision=FP64, grad=False)
    with pytest.raises(HyError):  # no outside pass
        eng.branch_eval(node, C, precision=FP64)
    eng.eval(bl, [0.5], [1.0], sim["theta"], precision=FP64)
    with pytest.raises(HyError):  # other precision
        eng.branch_eval(node, C, precision=FP32)
    with pytest.raises(HyError):  # the root has no branch
        eng.branch_eval(tree.root, C, precision=FP64)
    assert np.isfinite(eng.branch_eval(node, C, precision=FP64)[0])
    eng.close()
