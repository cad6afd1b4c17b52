"""Wheatstone-bridged trees on the GPU (hyphy3.network, hy_network_*)
against the oracle's bridged pruning (oracle/bridge_ref.py, restating
wheatstone_likelihood.hpp:21-148 / likelihood.hpp:376-425, pinned by the
reference's exact JC69 example): lnL in both precisions with one bridge and
with two (nine resolution configurations) under a two-class mixture and the
3 / mu-bar convention, and every gradient (branch lengths, omega, theta,
mixture weights, bridge lengths, bridge weights) against central differences
of the oracle's lnL."""
import numpy as np
import pytest

from helpers import branch_lengths, rel
from oracle import bridge_ref as br
from oracle import data_ref as dr
from oracle import hyoracle as ho

pytestmark = pytest.mark.gpu

FP64, FP32 = 0, 1
THETA = [1.2, 0.8, 1.0, 2.0, 0.7]


def alignment(seed, taxa):
    rng = np.random.default_rng(seed)
    codons = ["ATG", "AAA", "CCC", "GGG", "TTT", "CAT", "ACG", "AGA", "GAC", "TAC", "AAG", "CCA"]
    base = rng.choice(codons, size=45)
    seqs = {}
    for i, n in enumerate(taxa):
        s = base.copy()
        flip = rng.random(45) < 0.35
        s[flip] = rng.choice(codons, size=int(flip.sum()))
        seqs[n] = "".join(s)
    seqs[taxa[1]] = seqs[taxa[1]][:30] + "---" + seqs[taxa[1]][33:]
    return dr.AlignmentRef(list(seqs), list(seqs.values()))


class Model:
    """Oracle P for a mixture: classes[c] = [(matrix, weight)], convention 0 raw / 1 3-over-mu-bar."""

    def __init__(self, aln, omegas, classes, convention):
        self.spec = []
        for om in omegas:
            Q, mu = ho.mg94_Q(ho.Code(0), ho.params(1.0, om, THETA), aln.pos_freqs, aln.cf3x4)
            self.spec.append((ho.decompose(Q, aln.cf3x4), mu))
        self.classes, self.conv = classes, convention

    def P(self, cls, t, weights=None):
        comps = self.classes[cls] if weights is None else [(m, w) for (m, _), w in zip(self.classes[cls], weights)]
        scale = 1.0
        if self.conv == 1:
            scale = 3.0 / max(sum(w * self.spec[m][1] for m, w in comps), 1e-12)
        return sum(w * ho.transition(*self.spec[m][0], t * scale) for m, w in comps)


def oracle_lnl(prob, tree, model, bl, bclass, bridges, bw, blen):
    P = np.stack([np.eye(61) if n.id == tree.root else model.P(bclass[n.id], bl[n.id]) for n in tree.nodes])
    Pb = [[model.P(b["cls"], blen[k][r]) for r in range(3)] for k, b in enumerate(bridges)]
    bs = [dict(hub=b["hub"], children=b["children"], weights=list(bw[k])) for k, b in enumerate(bridges)]
    return br.network_loglik(prob, P, Pb, bs, prob_pi(prob))


def prob_pi(prob):
    return prob._pi


def setup(nwk, taxa, seed):
    aln = alignment(seed, taxa)
    tree = dr.TreeRef.from_newick(nwk)
    prob = ho.Problem.from_refs(tree, aln)
    prob._pi = aln.cf3x4
    hubs = [n.id for n in tree.nodes if len(n.children) == 3 and n.id != tree.root]
    return aln, tree, prob, hubs


def network_for(tree, aln, bridges):
    from hyphy3.network import Network
    net = Network(0)
    parent = [n.parent for n in tree.nodes]
    children = [list(n.children) for n in tree.nodes]
    net.set_problem(parent, children, tree.leaf_to_taxon(aln.names), aln.states, aln.pmasks, aln.weights,
                    [(b["hub"], b["children"]) for b in bridges])
    net.set_model(0, aln.pos_freqs, aln.cf3x4)
    return net


def check(nwk, taxa, seed, omegas, classes, convention, class_of, bw, blen, fd_nodes=4):
    aln, tree, prob, hubs = setup(nwk, taxa, seed)
    bridges = [dict(hub=u, children=tuple(tree.nodes[u].children), cls=class_of(u)) for u in hubs]
    bclass = np.array([class_of(n.id) for n in tree.nodes], dtype=np.int32)
    model = Model(aln, omegas, classes, convention)
    bl = branch_lengths(tree)
    net = network_for(tree, aln, bridges)
    net.set_mixture(len(omegas), bclass, [[m for m, _ in c] for c in classes], convention=convention,
                    bridge_class=[b["cls"] for b in bridges])
    comp_w = [w for c in classes for _, w in c]
    ref = oracle_lnl(prob, tree, model, bl, bclass, bridges, bw, blen)
    for prec, tl, tg in ((FP64, 1e-9, 2e-6), (FP32, 1e-5, 1e-4)):
        r = net.eval(bl, bw, blen, omegas, comp_w, THETA, precision=prec)
        assert rel(r["log_likelihood"], ref) < tl, (prec, r["log_likelihood"], ref)
        h = 1e-6
        scale = max(1.0, np.abs(r["grad_bl"]).max())

        def close(g, fd):
            assert abs(g - fd) < tg * max(scale, abs(fd)), (prec, g, fd)
        nodes = [n.id for n in tree.nodes if n.id != tree.root][:fd_nodes] + [c for u in hubs
                                                                                for c in tree.nodes[u].children]
        for v in nodes:
            bp, bm = bl.copy(), bl.copy()
            bp[v] += h
            bm[v] -= h
            fd = (oracle_lnl(prob, tree, model, bp, bclass, bridges, bw, blen)
                  - oracle_lnl(prob, tree, model, bm, bclass, bridges, bw, blen)) / (2 * h)
            close(r["grad_bl"][v], fd)
        for k in range(len(bridges)):
            for rr in range(3):
                if bw[k][rr] > 0.0:
                    lp, lm = [np.array(blen, dtype=float) for _ in range(2)]
                    lp[k][rr] += h
                    lm[k][rr] -= h
                    fd = (oracle_lnl(prob, tree, model, bl, bclass, bridges, bw, lp)
                          - oracle_lnl(prob, tree, model, bl, bclass, bridges, bw, lm)) / (2 * h)
                    close(r["grad_bridge_lengths"][k][rr], fd)
                    wp, wm = [np.array(bw, dtype=float) for _ in range(2)]
                    wp[k][rr] += h
                    wm[k][rr] -= h
                    fd = (oracle_lnl(prob, tree, model, bl, bclass, bridges, wp, blen)
                          - oracle_lnl(prob, tree, model, bl, bclass, bridges, wm, blen)) / (2 * h)
                    close(r["grad_bridge_weights"][k][rr], fd)
        # omega moves the eigendecomposition, whose rounding (about 1e-13 of
        # lnL) a step of 1e-6 would amplify to 1e-4: a wider step (truncation
        # error about 1e-8 relative)
        ho_ = 1e-4
        for m in range(len(omegas)):
            op, om_ = list(omegas), list(omegas)
            op[m] += ho_
            om_[m] -= ho_
            fd = (oracle_lnl(prob, tree, Model(aln, op, classes, convention), bl, bclass, bridges, bw, blen)
                  - oracle_lnl(prob, tree, Model(aln, om_, classes, convention), bl, bclass, bridges, bw, blen)) / (2 * ho_)
            close(r["grad_omega"][m], fd)
    net.close()


def test_one_bridge_single_matrix():
    nwk = "((A:0.1,B:0.2,C:0.15):0.05,(D:0.1,E:0.2):0.1,(F:0.3,G:0.05):0.2);"
    check(nwk, list("ABCDEFG"), 21, [0.6], [[(0, 1.0)]], 0, lambda v: 0,
          [[0.5, 0.3, 0.2]], [[0.07, 0.11, 0.13]])


def test_two_bridges_mixture_conv():
    nwk = "((A:0.1,B:0.2,C:0.15):0.05,(D:0.1,E:0.2,H:0.07):0.1,(F:0.3,G:0.05):0.2);"
    classes = [[(0, 0.7), (1, 0.3)], [(2, 1.0)]]
    check(nwk, list("ABCDEFGH"), 22, [0.2, 2.0, 0.9], classes, 1, lambda v: 1 if v % 3 == 0 else 0,
          [[0.6, 0.4, 0.0], [0.2, 0.5, 0.3]], [[0.05, 0.09, 0.0], [0.12, 0.04, 0.08]], fd_nodes=2)
