"""BUSTED's error sink in the CPU oracle (busted.hpp:193-206, 571-577).

Pins the restatement by what the definition fixes rather than by retyping it:
on a star tree whose branches are pure sink the likelihood is a brute-force
sum over the root state of products of P_sink rows (so the orientation,
parent state in the row, is checked: P_sink is not symmetric); a sink of
weight zero leaves lnL and every gradient of the plain mixture unchanged;
the sink weight's gradient and the branch-length gradients of a mixed class
match central differences of lnL."""
import numpy as np
import pytest

from oracle import data_ref as dr
from oracle import hyoracle as ho

CODONS = dr.Code().sense_codons


def _problem(rng, newick, n, L):
    tree = dr.TreeRef.from_newick(newick)
    names = ["T%d" % i for i in range(n)]
    seqs = ["".join(CODONS[int(c)] for c in rng.integers(0, len(CODONS), size=L)) for _ in range(n)]
    aln = dr.AlignmentRef(names, seqs)
    return tree, aln, seqs


def _bl(tree):
    return np.array([nd.bl if nd.parent >= 0 else 0.0 for nd in tree.nodes])


def _eval(tree, aln, mix, bl, theta=(1.0, 1.0, 1.0, 1.0, 1.0)):
    return ho.mixture_eval(ho.Problem.from_refs(tree, aln), ho.Code(0), ho.params(1.0, 1.0, list(theta)),
                           aln.pos_freqs, aln.cf3x4, mix, bl)


def test_pure_sink_star_tree_brute_force():
    rng = np.random.default_rng(5)
    tree, aln, seqs = _problem(rng, "(T0:0.1,T1:0.2,T2:0.3);", 3, 12)
    f = rng.dirichlet(np.ones(len(CODONS)))
    N = len(tree.nodes)
    mix = ho.Mixture([0.5], np.zeros(N, dtype=np.int32), [[(-1, 1.0)]], 0, sink_freqs=f)
    r = _eval(tree, aln, mix, _bl(tree))
    S = len(CODONS)
    Ps = np.array([[f[j] / (1.0 - f[i]) if i != j else 0.0 for j in range(S)] for i in range(S)])
    assert np.allclose(Ps.sum(axis=1), 1.0)
    pi = np.asarray(aln.cf3x4)
    idx = {c: k for k, c in enumerate(CODONS)}
    ll = 0.0
    for col in range(12):
        x = [idx[s[3 * col:3 * col + 3]] for s in seqs]
        ll += np.log(np.sum(pi * Ps[:, x[0]] * Ps[:, x[1]] * Ps[:, x[2]]))
    assert abs(r["log_likelihood"] - ll) < 1e-10 * abs(ll)
    assert np.all(r["grad_bl"] == 0.0)


@pytest.mark.parametrize("seed", [1, 2])
def test_zero_weight_sink_is_the_plain_mixture(seed):
    rng = np.random.default_rng(seed)
    tree, aln, _ = _problem(rng, "((T0:0.1,T1:0.05):0.07,(T2:0.2,T3:0.1):0.03,T4:0.15);", 5, 20)
    N = len(tree.nodes)
    f = rng.dirichlet(np.ones(len(CODONS)))
    plain = ho.Mixture([0.3, 2.0], np.zeros(N, dtype=np.int32), [[(0, 0.6), (1, 0.4)]], 0)
    sink = ho.Mixture([0.3, 2.0], np.zeros(N, dtype=np.int32), [[(0, 0.6), (1, 0.4), (-1, 0.0)]], 0, sink_freqs=f)
    a, b = _eval(tree, aln, plain, _bl(tree)), _eval(tree, aln, sink, _bl(tree))
    assert abs(a["log_likelihood"] - b["log_likelihood"]) <= 1e-13 * abs(a["log_likelihood"])
    for k in ("grad_bl", "grad_omega", "grad_theta"):
        assert np.allclose(a[k], b[k], rtol=1e-12, atol=1e-12)
    assert np.allclose(a["grad_w"], b["grad_w"][:2], rtol=1e-12)


def test_sink_gradients_match_central_differences():
    rng = np.random.default_rng(9)
    tree, aln, _ = _problem(rng, "((T0:0.1,T1:0.05):0.07,(T2:0.2,T3:0.1):0.03,T4:0.15);", 5, 20)
    N = len(tree.nodes)
    f = rng.dirichlet(np.ones(len(CODONS)))
    bc = np.zeros(N, dtype=np.int32)
    bc[[nd.id for nd in tree.nodes if nd.parent >= 0][::2]] = 1

    def mix(ws):
        return ho.Mixture([0.4, 0.2, 3.0], bc, [[(0, 1.0)], [(1, 0.55), (2, 0.35), (-1, ws)]], 0, sink_freqs=f)

    bl = _bl(tree)
    ws = 0.1
    r = _eval(tree, aln, mix(ws), bl)
    h = 1e-6
    fd = (_eval(tree, aln, mix(ws + h), bl)["log_likelihood"] - _eval(tree, aln, mix(ws - h), bl)["log_likelihood"]) / (2 * h)
    assert abs(r["grad_w"][3] - fd) < 1e-6 * max(1.0, abs(fd))
    for n in range(N):
        if tree.nodes[n].parent < 0:
            continue
        bp, bm = bl.copy(), bl.copy()
        bp[n] += h
        bm[n] -= h
        fd = (_eval(tree, aln, mix(ws), bp)["log_likelihood"] - _eval(tree, aln, mix(ws), bm)["log_likelihood"]) / (2 * h)
        assert abs(r["grad_bl"][n] - fd) < 1e-5 * max(1.0, abs(fd)), n


def test_sink_needs_raw_time():
    rng = np.random.default_rng(3)
    tree, aln, _ = _problem(rng, "(T0:0.1,T1:0.2,T2:0.3);", 3, 4)
    N = len(tree.nodes)
    mix = ho.Mixture([0.5], np.zeros(N, dtype=np.int32), [[(0, 0.9), (-1, 0.1)]], 1,
                     sink_freqs=np.full(len(CODONS), 1.0 / len(CODONS)))
    with pytest.raises(RuntimeError):
        _eval(tree, aln, mix, _bl(tree))
