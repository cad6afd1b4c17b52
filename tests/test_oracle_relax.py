"""Pins of the oracle's RELAX restatement (oracle/relax_ref.py,
relax.hpp:546-568, 803-1166): central differences of its own lnL (computed
from the workspaces' log-likelihoods, a different path from its V.(dP D)
gradients) at both mixture levels and both normalisations; at k = 1 without
unclassified branches the branch-site lnL and gradients equal the C oracle's
one-class three-component mixture under the 3 / mu-bar convention; with all
weight on one class the site level equals the branch-site level."""
import numpy as np

from helpers import branch_lengths, sim_problem
from oracle import hyoracle as ho
from oracle import relax_ref as rr


def setup(n=8, L=40, seed=71, gaps=0.1):
    tree, aln, sim = sim_problem(n, L, seed, gaps=gaps)
    prob = ho.Problem.from_refs(tree, aln)
    rng = np.random.default_rng(seed)
    part = rng.integers(0, 2, len(tree.nodes)).astype(np.int32)
    return tree, aln, sim, prob, part


def relax(prob, aln, sim, part, bl, **kw):
    return rr.relax_branch_gradients(prob, ho.Code(0), sim["theta"], aln.pos_freqs, aln.cf3x4, part, bl, **kw)


def test_relax_central_differences():
    tree, aln, sim, prob, part = setup()
    part_u = part.copy()
    part_u[[n.id for n in tree.nodes if n.is_leaf][:2]] = 2
    bl = branch_lengths(tree)
    cases = [
        dict(omegas_R=[0.1, 0.7, 2.5], weights=[0.5, 0.3, 0.2], k=0.6),
        dict(omegas_R=[0.1, 0.7, 2.5], weights=[0.5, 0.3, 0.2], k=1.7, branch_site=False),
        dict(omegas_R=[0.2, 1.0, 4.0], weights=[0.6, 0.3, 0.1], k=0.8, branch_site=False, shared=False),
        dict(omegas_R=[0.1, 0.7, 2.5], weights=[0.5, 0.3, 0.2], k=1.3, omegas_U=[0.05, 0.9, 3.0],
             weights_U=[0.4, 0.4, 0.2], part=part_u),
    ]
    h = 1e-6
    for kw in cases:
        pt = kw.pop("part", part)
        lnl, g = relax(prob, aln, sim, pt, bl, **kw)
        for node in [n.id for n in tree.nodes if n.id != tree.root][:6]:
            bp, bm = bl.copy(), bl.copy()
            bp[node] += h
            bm[node] -= h
            fd = (relax(prob, aln, sim, pt, bp, **kw)[0] - relax(prob, aln, sim, pt, bm, **kw)[0]) / (2 * h)
            assert abs(fd - g[node]) < 2e-5 * max(1.0, abs(g[node])), (kw, node, fd, g[node])


def test_relax_k1_is_the_plain_mixture():
    tree, aln, sim, prob, part = setup(seed=72)
    bl = branch_lengths(tree)
    om, wts = [0.1, 0.7, 2.5], [0.5, 0.3, 0.2]
    lnl, g = relax(prob, aln, sim, part, bl, omegas_R=om, weights=wts, k=1.0)
    N = len(tree.nodes)
    mix = ho.Mixture(om, np.zeros(N, dtype=np.int32), [[(0, wts[0]), (1, wts[1]), (2, wts[2])]], convention=1)
    ref = ho.mixture_eval(prob, ho.Code(0), ho.params(1.0, 1.0, sim["theta"]), aln.pos_freqs, aln.cf3x4, mix, bl)
    assert abs(lnl - ref["log_likelihood"]) < 1e-11 * abs(lnl)
    nr = [n.id for n in tree.nodes if n.id != tree.root]
    np.testing.assert_allclose(g[nr], ref["grad_bl"][nr], rtol=1e-9, atol=1e-9 * np.abs(g).max())


def test_relax_site_level_with_one_class():
    tree, aln, sim, prob, part = setup(seed=73)
    bl = branch_lengths(tree)
    kw = dict(omegas_R=[0.3, 0.7, 2.5], weights=[1.0, 0.0, 0.0], k=1.4)
    a = relax(prob, aln, sim, part, bl, **kw)
    b = relax(prob, aln, sim, part, bl, branch_site=False, **kw)
    assert abs(a[0] - b[0]) < 1e-12 * abs(a[0])
    np.testing.assert_allclose(a[1], b[1], rtol=1e-10, atol=1e-12)
