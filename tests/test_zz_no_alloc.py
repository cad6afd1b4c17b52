"""hy_loglik_from_P_noalloc (GPU): the allocation-free variant of
hy_loglik_from_P, called through the C ABI, against hy_loglik_from_P on the
same caller-supplied transition matrices (FP64 1e-9 relative, FP32 1e-5;
site log-likelihoods alike), on two calls in a row of which the second has a
smaller problem (the fixed scratch keeps the first call's contents), with
polytomies (identity edges), gaps and partial ambiguities; a problem whose
partials exceed the fixed scratch is refused before any device work."""
import numpy as np
import pytest

from helpers import branch_lengths, make_engine, rel
from oracle import data_ref as dr
from oracle import hyoracle as ho
from test_gpu_sweep import random_alignment, random_newick
from hyphy3.engine import HY_ERR_ARG, HyError

pytestmark = pytest.mark.gpu

FP64, FP32 = 0, 1


def problem(seed, n, L, gaps):
    rng = np.random.default_rng(seed)
    tree = dr.TreeRef.from_newick(random_newick(rng, n))
    aln = dr.AlignmentRef(["T%d" % i for i in range(n)], random_alignment(rng, n, L, gaps=gaps))
    # transition matrices of a random MG94 model at the tree's branch lengths
    Q, _ = ho.mg94_Q(ho.Code(0), ho.params(1.0, rng.uniform(0.2, 2.0), rng.uniform(0.5, 2.0, size=5)),
                     aln.pos_freqs, aln.cf3x4)
    ev, U, V = ho.decompose(Q, aln.cf3x4)
    bl = branch_lengths(tree)
    P = np.stack([ho.transition(ev, U, V, b) for b in bl])
    return tree, aln, P


def check(eng, P, pi, label):
    for prec, tl in ((FP64, 1e-9), (FP32, 1e-5)):
        print(f"[no_alloc] {label} precision {prec}", flush=True)
        ll0, s0 = eng.loglik_from_P(P, pi, precision=prec, site=True)
        ll1, s1 = eng.loglik_from_P_noalloc(P, pi, precision=prec, site=True)
        assert rel(ll1, ll0) < tl, (label, prec, ll1, ll0)
        np.testing.assert_allclose(s1, s0, rtol=10 * tl, atol=1e-12)


def test_noalloc_matches_loglik_from_P_two_calls():
    big = problem(31, 70, 900, 0.1)
    small = problem(32, 11, 57, 0.2)
    eng = make_engine(big[0], big[1])
    eng2 = make_engine(small[0], small[1])
    try:
        check(eng, big[2], big[1].cf3x4, "first call, 70 taxa x 900 codons")
        check(eng2, small[2], small[1].cf3x4, "second call, 11 taxa x 57 codons")
        # the same handle reused for the smaller problem
        from helpers import tree_arrays
        parent, children, l2t = tree_arrays(small[0], small[1])
        eng.set_problem(parent, children, l2t, small[1].states, small[1].pmasks, small[1].weights, S=61)
        eng.set_model(0, small[1].pos_freqs, small[1].cf3x4)
        check(eng, small[2], small[1].cf3x4, "handle reused, 11 taxa x 57 codons")
        # against the oracle too
        ref = ho.loglik_from_P(ho.Problem.from_refs(small[0], small[1]), small[2], small[1].cf3x4)
        assert rel(eng.loglik_from_P_noalloc(small[2], small[1].cf3x4, precision=FP64), ref) < 1e-9
    finally:
        eng.close()
        eng2.close()


def test_noalloc_refuses_problems_beyond_the_scratch():
    """64 taxa x 5,000 distinct codon columns: the FP32 partials (about 80 MB)
    fit the 128 MiB scratch, the FP64 ones (about 160 MB) do not."""
    rng = np.random.default_rng(5)
    n, L = 64, 5000
    tree = dr.TreeRef.from_newick(random_newick(rng, n))
    codons = dr.Code().sense_codons
    seqs = ["".join(rng.choice(codons, size=L)) for _ in range(n)]
    aln = dr.AlignmentRef(["T%d" % i for i in range(n)], seqs)
    N = len(tree.nodes)
    P = np.broadcast_to(0.2 * np.eye(61) + 0.8 / 61, (N, 61, 61)).copy()
    pi = aln.cf3x4
    eng = make_engine(tree, aln)
    try:
        with pytest.raises(HyError) as ex:
            eng.loglik_from_P_noalloc(P, pi, precision=FP64)
        assert ex.value.code == HY_ERR_ARG and "scratch" in str(ex.value)
        ll0 = eng.loglik_from_P(P, pi, precision=FP32)
        ll1 = eng.loglik_from_P_noalloc(P, pi, precision=FP32)
        assert rel(ll1, ll0) < 1e-5
    finally:
        eng.close()
