"""Pins of the oracle's bridged pruning (oracle/bridge_ref.py, restating
wheatstone_likelihood.hpp:21-148 and likelihood.hpp:376-425):
  * the exact JC69 log-likelihoods of the 5-taxon Wheatstone example in the
    reference's tests (tests/test_ohmic.cpp:176-265: per-site lnL of the
    invariant, majority and slow-shunt partitions for weights (1, 0, 0),
    (0, 1, 0) and (0.7, 0.3, 0)), JC69 transition matrices in closed form;
  * a bridge with all weight on resolution r equals the C oracle's pruning of
    the binary tree resolved that way (MG94 codon problem with gaps);
  * per pattern the bridged likelihood is the weighted sum of the three
    resolutions' likelihoods."""
import numpy as np

from helpers import branch_lengths
from oracle import bridge_ref as br
from oracle import data_ref as dr
from oracle import hyoracle as ho


def jc69(t):
    e = np.exp(-4.0 * t / 3.0)
    return np.full((4, 4), 0.25 - 0.25 * e) + np.eye(4) * e


def test_jc69_five_taxon_known_answers():
    # leaves A, B, C, D, O = 0..4; hub u = 5 (children A, B, C); root = 6 (children u, D, O)
    parent = [5, 5, 5, 6, 6, 6, -1]
    kids = [[], [], [], [], [], [0, 1, 2], [5, 3, 4]]
    bl = [0.005, 0.005, 0.005, 0.005, 0.0, 0.0, 0.0]
    post = [0, 1, 2, 5, 3, 4, 6]
    is_leaf = [1, 1, 1, 1, 1, 0, 0]
    l2t = [0, 1, 2, 3, 4, -1, -1]
    nuc = {"A": 0, "C": 1, "G": 2, "T": 3}
    # one pattern per partition: invariant (90 sites), AB|CD majority (7), AC|BD shunt (3)
    parts = {"inv": ("AAAAA", 90, 1.0), "maj": ("CCAAA", 7, 1.0), "shunt": ("GAGAA", 3, 0.05)}
    expected = {(1.0, 0.0, 0.0): (-1.485192, -5.083371, -20.175114),
                (0.0, 1.0, 0.0): (-1.485192, -14.248631, -8.010028),
                (0.7, 0.3, 0.0): (-1.485192, -5.440001, -9.213988)}
    for wts, exp in expected.items():
        for (name, (col, nsites, scale)), want in zip(parts.items(), exp):
            states = np.array([[nuc[ch] for ch in col]], dtype=np.int8)
            prob = ho.Problem(parent, kids, post, is_leaf, l2t, states, np.zeros((1, 5), dtype=np.uint64),
                              [float(nsites)], S=4)
            P = np.stack([jc69(b * scale) for b in bl])
            Pb = [[jc69(0.08 * scale)] * 3]
            lnl = br.network_loglik(prob, P, Pb, [dict(hub=5, children=(0, 1, 2), weights=wts)], np.full(4, 0.25))
            assert abs(lnl / nsites - want) < 1e-5 * abs(want), (wts, name, lnl / nsites, want)


def codon_setup():
    rng = np.random.default_rng(11)
    names = ["A", "B", "C", "D", "E", "F", "G"]
    codons = ["ATG", "AAA", "CCC", "GGG", "TTT", "CAT", "ACG", "AGA", "GAC", "TAC"]
    seqs = {n: "".join(rng.choice(codons, size=40)) for n in names}
    seqs["B"] = seqs["B"][:30] + "---" + seqs["B"][33:]
    aln = dr.AlignmentRef(list(seqs), list(seqs.values()))
    p = ho.params(1.0, 0.6, [1.2, 0.8, 1.0, 2.0, 0.7])
    Q, _ = ho.mg94_Q(ho.Code(0), p, aln.pos_freqs, aln.cf3x4)
    ev, U, V = ho.decompose(Q, aln.cf3x4)
    return aln, (ev, U, V)


def P_of(tree, spec, extra=None):
    ev, U, V = spec
    bl = branch_lengths(tree)
    return np.stack([np.eye(61) if n.id == tree.root else ho.transition(ev, U, V, bl[n.id]) for n in tree.nodes])


def test_single_resolution_is_the_resolved_tree():
    aln, spec = codon_setup()
    ev, U, V = spec
    hub_nwk = "((A:0.1,B:0.2,C:0.15):0.05,(D:0.1,E:0.2):0.1,(F:0.3,G:0.05):0.2);"
    tlen = (0.07, 0.11, 0.13)
    resolved = ("(((A:0.1,B:0.2):%g,C:0.15):0.05,(D:0.1,E:0.2):0.1,(F:0.3,G:0.05):0.2);" % tlen[0],
                "(((A:0.1,C:0.15):%g,B:0.2):0.05,(D:0.1,E:0.2):0.1,(F:0.3,G:0.05):0.2);" % tlen[1],
                "(((B:0.2,C:0.15):%g,A:0.1):0.05,(D:0.1,E:0.2):0.1,(F:0.3,G:0.05):0.2);" % tlen[2])
    tree = dr.TreeRef.from_newick(hub_nwk)
    prob = ho.Problem.from_refs(tree, aln)
    hub = [n.id for n in tree.nodes if len(n.children) == 3 and n.id != tree.root][0]
    a, b, c = tree.nodes[hub].children
    P = P_of(tree, spec)
    Pb = [[ho.transition(ev, U, V, t) for t in tlen]]
    site_r = []
    for r in range(3):
        w = [0.0, 0.0, 0.0]
        w[r] = 1.0
        lnl, site = br.network_loglik(prob, P, Pb, [dict(hub=hub, children=(a, b, c), weights=w)], aln.cf3x4,
                                      site=True)
        rt = dr.TreeRef.from_newick(resolved[r])
        rp = ho.Problem.from_refs(rt, aln)
        ref, rsite = ho.loglik_from_P(rp, P_of(rt, spec), aln.cf3x4, site=True)
        assert abs(lnl - ref) < 1e-11 * abs(ref), (r, lnl, ref)
        np.testing.assert_allclose(site, rsite, rtol=1e-11)
        site_r.append(site)
    w = (0.5, 0.3, 0.2)
    _, site = br.network_loglik(prob, P, Pb, [dict(hub=hub, children=(a, b, c), weights=w)], aln.cf3x4, site=True)
    mix = np.log(sum(w[r] * np.exp(site_r[r]) for r in range(3)))
    np.testing.assert_allclose(site, mix, rtol=1e-12)
