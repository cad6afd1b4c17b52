"""Pins of the oracle's tile workspaces (compute_inside_batch_in_place /
compute_inside_outside_batch_in_place, likelihood.hpp:295-762): the
inside-outside invariant V_v . (P_v D_v) = L for every branch, U = P^T V,
P_D = P D, and per-pattern lnL equal to the independent pruning routine."""
import numpy as np

from helpers import branch_lengths, sim_problem
from oracle import hyoracle as ho


def transition_stack(tree, aln, sim, omega=0.5):
    p = ho.params(1.0, omega, sim["theta"])
    Q, mu = ho.mg94_Q(ho.Code(0), p, aln.pos_freqs, aln.cf3x4)
    ev, U, V = ho.decompose(Q, aln.cf3x4)
    bl = branch_lengths(tree) * 3.0 / mu
    P = np.stack([np.eye(61) if n.id == tree.root else ho.transition(ev, U, V, bl[n.id]) for n in tree.nodes])
    return P


def test_workspace_invariants():
    tree, aln, sim = sim_problem(9, 70, 41, gaps=0.1)
    prob = ho.Problem.from_refs(tree, aln)
    P = transition_stack(tree, aln, sim)
    pi = aln.cf3x4
    _, site = ho.loglik_from_P(prob, P, pi, site=True)
    for p0, p1 in ((0, 64), (64, prob.num_patterns)):
        w = ho.workspace(prob, P, pi, p0, p1, outside=True)
        np.testing.assert_allclose(w["logL"], site[p0:p1], rtol=1e-13)
        wi = ho.workspace(prob, P, pi, p0, p1, outside=False)
        np.testing.assert_allclose(wi["D"], w["D"], rtol=1e-14)
        np.testing.assert_allclose(wi["logL"], w["logL"], rtol=1e-14)
        for n in tree.nodes:
            if n.id == tree.root:
                np.testing.assert_allclose(w["U"][n.id], np.repeat(pi[:, None], p1 - p0, 1))
                continue
            np.testing.assert_allclose(w["P_D"][n.id], P[n.id] @ w["D"][n.id], rtol=1e-12, atol=1e-300)
            np.testing.assert_allclose(w["U"][n.id], P[n.id].T @ w["V"][n.id], rtol=1e-12, atol=1e-300)
            inv = np.einsum("sb,sb->b", w["V"][n.id], w["P_D"][n.id])
            np.testing.assert_allclose(inv, w["L"], rtol=1e-12)
