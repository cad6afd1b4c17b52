"""GPU read-back of the reference's workspaces through the C++ facade
(compute_inside_batch_in_place / compute_inside_outside_batch_in_place)
against the oracle's tile workspaces."""
import os

import numpy as np
import pytest

from helpers import GOLDEN, branch_lengths, rel
from oracle import data_ref as dr
from oracle import hyoracle as ho
from test_oracle_workspace import transition_stack

pytestmark = pytest.mark.gpu


def close_cols(a, b, tol):
    """Column-wise relative agreement of [node][state][pattern] stacks."""
    scale = np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-300)
    return float(np.max(np.abs(a - b) / scale))


def problems():
    from hyphy3 import simulate
    out = []
    for n, L, seed, gaps in ((12, 150, 51, 0.1), (30, 70, 52, 0.0)):
        sim = simulate.simulate(n, L, seed, gap_fraction=gaps)
        out.append((sim["fasta"], sim["newick"], sim))
    return out


@pytest.mark.parametrize("k", [0, 1])
def test_workspaces_match_oracle(k):
    import _hyphy3
    fasta, newick, sim = problems()[k]
    aln = _hyphy3.Alignment.from_text(fasta)
    tree = _hyphy3.Tree.from_newick(newick, False)
    names, seqs = dr.parse_fasta(fasta)
    ar = dr.AlignmentRef(names, seqs)
    tr = dr.TreeRef.from_newick(newick, keep_rooted=False)
    prob = ho.Problem.from_refs(tr, ar)
    P = transition_stack(tr, ar, sim)
    pi = list(ar.cf3x4)
    np_ = prob.num_patterns
    for p0, p1 in ((0, min(64, np_)), (min(64, np_ - 1), np_), (3, 4)):
        ref = ho.workspace(prob, P, pi, p0, p1, outside=True)
        w = _hyphy3.compute_inside_outside_batch_in_place(tree, aln, p0, p1, P, pi)
        for f in ("D", "V", "P_D", "U"):
            assert close_cols(w[f], ref[f], 1e-10) < 1e-10, (f, p0, p1)
        assert np.array_equal(w["boost_counts"], ref["boost_counts"])
        np.testing.assert_allclose(w["node_log_scale"], ref["node_log_scale"], rtol=0, atol=0)
        np.testing.assert_allclose(w["likelihood"], ref["L"], rtol=1e-10)
        np.testing.assert_allclose(w["log_likelihood"], ref["logL"], rtol=1e-11)
        wi = _hyphy3.compute_inside_batch_in_place(tree, aln, p0, p1, P, pi)
        refi = ho.workspace(prob, P, pi, p0, p1, outside=False)
        assert close_cols(wi["D"], refi["D"], 1e-10) < 1e-10
        np.testing.assert_allclose(wi["root_log_L"], refi["logL"], rtol=1e-11)


def test_workspace_boosts_on_deep_tree():
    """A 200-taxon caterpillar underflows 1e-150: boosts, log scales and
    the invariant V . P_D = L carry through the read-back."""
    import _hyphy3
    n = 200
    rng = np.random.default_rng(7)
    nwk = "T0:2.0"
    for i in range(1, n - 1):
        nwk = "(%s,T%d:2.0):1.5" % (nwk, i)
    nwk = "(%s,T%d:2.0);" % (nwk, n - 1)
    codons = dr.Code().sense_codons
    seqs = ["".join(rng.choice(codons, size=20)) for _ in range(n)]
    fasta = "".join(">T%d\n%s\n" % (i, s) for i, s in enumerate(seqs))
    aln = _hyphy3.Alignment.from_text(fasta)
    tree = _hyphy3.Tree.from_newick(nwk, False)
    tr = dr.TreeRef.from_newick(nwk, keep_rooted=False)
    ar = dr.AlignmentRef(["T%d" % i for i in range(n)], seqs)
    prob = ho.Problem.from_refs(tr, ar)
    sim = {"theta": [1.0, 1.0, 1.0, 2.0, 1.0]}
    P = transition_stack(tr, ar, sim)
    pi = list(ar.cf3x4)
    ref = ho.workspace(prob, P, pi, 0, prob.num_patterns, outside=True)
    w = _hyphy3.compute_inside_outside_batch_in_place(tree, aln, 0, prob.num_patterns, P, pi)
    assert ref["boost_counts"].sum() > 0
    np.testing.assert_allclose(w["log_likelihood"], ref["logL"], rtol=1e-11)
    for v in range(len(tr.nodes)):
        if v == tr.root:
            continue
        inv = np.einsum("sb,sb->b", w["V"][v], w["P_D"][v])
        np.testing.assert_allclose(inv, w["likelihood"], rtol=1e-9)
    # where the boosts agree with the reference's, the stored values do too
    same = (w["boost_counts"] == ref["boost_counts"]).all(axis=1) & \
        (w["node_log_scale"] == ref["node_log_scale"]).all(axis=1)
    assert same.mean() > 0.9
    for v in np.where(same)[0]:
        assert close_cols(w["D"][v:v + 1], ref["D"][v:v + 1], 1e-9) < 1e-9
