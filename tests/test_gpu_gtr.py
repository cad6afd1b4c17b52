"""Nucleotide GTR stage on the MI355X (hy_gtr_eval) against the oracle's
restatement (hyoracle.gtr_eval; compute_gtr_log_likelihood,
likelihood.hpp:181-254): lnL, site lnL and the branch-length and
exchangeability gradients, FP64 (1e-9 / 1e-7) and FP32 (1e-5 / 1e-4), on the
reference's fixtures, random trees with polytomies, ambiguity codes and gaps,
and a 128-taxon simulated alignment; repeated evaluations are
bit-identical."""
import os

import numpy as np
import pytest

from helpers import GOLDEN, branch_lengths, grad_err, rel, sim_problem, tree_arrays
from oracle import data_ref as dr
from oracle import hyoracle as ho
from test_gpu_sweep import random_newick

pytestmark = pytest.mark.gpu

FP64, FP32 = 0, 1


def nuc_engine(tree, naln):
    from hyphy3.engine import Engine
    eng = Engine(0)
    parent, children, l2t = tree_arrays(tree, naln)
    eng.set_problem(parent, children, l2t, naln.states, naln.pmasks, naln.weights, S=4)
    return eng


def check(tree, naln, theta, freqs, bl, label):
    ref = ho.gtr_eval(ho.nuc_problem(tree, naln), theta, freqs, bl, threads=8)
    eng = nuc_engine(tree, naln)
    try:
        for prec, tl, tg in ((FP64, 1e-9, 1e-7), (FP32, 1e-5, 1e-4)):
            print(f"[gtr] {label} precision {prec}", flush=True)
            r = eng.gtr_eval(bl, theta, freqs, precision=prec, site=True)
            assert rel(r["log_likelihood"], ref["log_likelihood"]) < tl, (label, prec)
            np.testing.assert_allclose(r["site_ll"], ref["site_ll"], rtol=10 * tl)
            assert grad_err(r["grad_bl"], ref["grad_bl"]) < tg, (label, prec)
            assert grad_err(r["grad_theta"], ref["grad_theta"]) < tg, (label, prec)
        a = eng.gtr_eval(bl, theta, freqs, precision=FP64)
        b = eng.gtr_eval(bl, theta, freqs, precision=FP64)
        assert a["log_likelihood"] == b["log_likelihood"] and np.array_equal(a["grad_bl"], b["grad_bl"])
        f = eng.gtr_eval(bl, theta, freqs, precision=FP64, grad=False)
        assert f["log_likelihood"] == a["log_likelihood"]
    finally:
        eng.close()


@pytest.mark.parametrize("name", ["cd2_reduced", "cd2", "adh"])
def test_gtr_fixtures(name):
    names, seqs = dr.parse_fasta(open(os.path.join(GOLDEN, name + ".fna")).read())
    aln = dr.AlignmentRef(names, seqs)
    naln = dr.NucleotideRef(aln)
    tree = dr.TreeRef.from_newick(open(os.path.join(GOLDEN, name + ".nwk")).read(), keep_rooted=False)
    bl = branch_lengths(tree)
    if not bl.any():
        # the cd2 and adh trees carry no lengths: all-zero branches make every P
        # the identity up to rounding noise, where a pattern's likelihood is 0
        # or tiny depending on the last bit (so in the reference too); seeded
        # lengths instead
        bl = np.random.default_rng(7).uniform(0.02, 0.3, size=len(bl))
        bl[tree.root] = 0.0
    check(tree, naln, [0.5, 0.3, 0.5, 1.0, 0.3], naln.nuc_freqs, bl, name)


@pytest.mark.parametrize("seed,n,L", [(1, 5, 33), (2, 17, 120), (3, 40, 301), (4, 4, 1)])
def test_gtr_random_trees(seed, n, L):
    rng = np.random.default_rng(900 + seed)
    tree = dr.TreeRef.from_newick(random_newick(rng, n))
    letters = list("ACGTACGTACGTACGTRYN-")
    seqs = ["".join(rng.choice(letters, size=3 * L)) for _ in range(n)]
    naln = dr.NucleotideRef(dr.AlignmentRef(["T%d" % i for i in range(n)], seqs))
    theta = rng.uniform(0.3, 3.0, size=5)
    check(tree, naln, theta, rng.dirichlet(np.ones(4) * 4), branch_lengths(tree), f"random {n} x {L}")


def test_gtr_simulated_128_taxa():
    tree, aln, sim = sim_problem(128, 700, 515, gaps=0.05)
    naln = dr.NucleotideRef(aln)
    check(tree, naln, sim["theta"], naln.nuc_freqs, branch_lengths(tree) * sim["conv"] / 3.0, "128 taxa")


def test_gtr_needs_four_states():
    from hyphy3.engine import HyError
    from helpers import load_cd2_reduced, make_engine
    tree, aln = load_cd2_reduced()
    eng = make_engine(tree, aln)
    try:
        with pytest.raises(HyError):
            eng.gtr_eval(branch_lengths(tree), np.ones(5), np.full(4, 0.25))
    finally:
        eng.close()


def test_gtr_facade():
    """_hyphy3.compute_gtr_log_likelihood (the reference's name, over the
    product's Tree / Alignment and its nucleotide view) against the oracle."""
    import _hyphy3
    from hyphy3 import simulate
    sim = simulate.simulate(20, 150, 616)
    tree = _hyphy3.Tree.from_newick(sim["newick"], False)
    aln = _hyphy3.Alignment.from_text(sim["fasta"])
    names, seqs = dr.parse_fasta(sim["fasta"])
    naln = dr.NucleotideRef(dr.AlignmentRef(names, seqs))
    rtree = dr.TreeRef.from_newick(sim["newick"], keep_rooted=False)
    theta = [0.7, 0.4, 1.3, 2.2, 0.9]
    ref = ho.gtr_eval(ho.nuc_problem(rtree, naln), theta, naln.nuc_freqs, branch_lengths(rtree), want_grad=False)
    ll = _hyphy3.compute_gtr_log_likelihood(tree, aln, theta, list(aln.nuc_frequencies))
    assert rel(ll, ref["log_likelihood"]) < 1e-9
