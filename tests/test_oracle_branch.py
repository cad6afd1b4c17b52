"""Pins of the oracle's aBSREL local projection (ABSRELAnalyzer::
compute_branch_context_matrices and eval_branch_ll_from_matrices,
absrel.hpp:438-455, 723-780).  The likelihood is linear in the matrix on any
one branch, so the branch context scored with the branch's own matrix is the
whole-tree lnL, and scored with another matrix it is the whole-tree lnL with
that matrix substituted (an independent pruning), for internal and tip
branches, with gaps and ambiguity codes; a zero matrix gives every pattern
the -1e20 sentinel."""
import numpy as np

from helpers import sim_problem
from oracle import hyoracle as ho
from test_oracle_workspace import transition_stack


def test_branch_context_identities():
    tree, aln, sim = sim_problem(11, 90, 47, gaps=0.15)
    prob = ho.Problem.from_refs(tree, aln)
    P = transition_stack(tree, aln, sim)
    pi = aln.cf3x4
    full = ho.loglik_from_P(prob, P, pi)
    rng = np.random.default_rng(5)
    p2 = ho.params(1.0, 2.5, sim["theta"])
    Q2, mu2 = ho.mg94_Q(ho.Code(0), p2, aln.pos_freqs, aln.cf3x4)
    ev2, U2, V2 = ho.decompose(Q2, aln.cf3x4)
    internal = [n.id for n in tree.nodes if not n.is_leaf and n.id != tree.root]
    tips = [n.id for n in tree.nodes if n.is_leaf]
    assert len(internal) >= 3 and len(tips) >= 3
    for node in internal[:3] + tips[:3]:
        Vm, Dm, off = ho.branch_context(prob, P, pi, node, tile=64)
        own = ho.branch_ll(prob, P[node], Vm, Dm, off)
        assert abs(own - full) < 1e-10 * abs(full), (node, own, full)
        for P_alt in (ho.transition(ev2, U2, V2, 0.37), rng.uniform(0.0, 1.0, (61, 61))):
            Ps = P.copy()
            Ps[node] = P_alt
            ref = ho.loglik_from_P(prob, Ps, pi)
            got = ho.branch_ll(prob, P_alt, Vm, Dm, off)
            assert abs(got - ref) < 1e-10 * abs(ref), (node, got, ref)
        zero = ho.branch_ll(prob, np.zeros((61, 61)), Vm, Dm, off)
        assert abs(zero - (off - 1e20 * prob.weights.sum())) < 1e-12 * 1e20 * prob.weights.sum()


def test_branch_context_offset_is_tile_independent():
    tree, aln, sim = sim_problem(7, 140, 48)
    prob = ho.Problem.from_refs(tree, aln)
    P = transition_stack(tree, aln, sim)
    node = [n.id for n in tree.nodes if not n.is_leaf and n.id != tree.root][0]
    a = ho.branch_context(prob, P, aln.cf3x4, node, tile=256)
    b = ho.branch_context(prob, P, aln.cf3x4, node, tile=32)
    np.testing.assert_allclose(a[0], b[0], rtol=1e-14)
    np.testing.assert_allclose(a[1], b[1], rtol=1e-14)
    assert abs(a[2] - b[2]) < 1e-9 * max(1.0, abs(a[2]))
