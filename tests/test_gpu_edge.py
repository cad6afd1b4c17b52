"""MI355X engine edge cases against the FP64 CPU oracle (GPU): polytomies,
ambiguity codes and gaps, unmatched leaves, caller-supplied P matrices with
zero-likelihood patterns, ragged tile counts, deep trees, two-class
(test/background) mixtures, another genetic code, handle reuse, the C++
facade, the autograd functions and the C ABI's error paths."""
import numpy as np
import pytest

from helpers import branch_lengths, grad_err, make_engine, rel, sim_problem
from oracle import data_ref as dr
from oracle import hyoracle as ho

pytestmark = pytest.mark.gpu

FP64, FP32 = 0, 1


def engine_for(tree, aln, code_id=0):
    return make_engine(tree, aln, code_id=code_id)


def check_plain(tree, aln, omega, theta, bl, code_id=0, tl64=1e-9, tg64=1e-7):
    prob = ho.Problem.from_refs(tree, aln)
    ll, g = ho.branch_length_gradients(prob, ho.Code(code_id), ho.params(1.0, omega, theta), aln.pos_freqs,
                                       aln.cf3x4, bl)
    eng = engine_for(tree, aln, code_id)
    for prec, tl, tg in ((FP64, tl64, tg64), (FP32, 1e-5, 1e-4)):
        r = eng.eval(bl, [omega], [1.0], theta, precision=prec)
        print("prec", prec, r["log_likelihood"], ll)
        assert rel(r["log_likelihood"], ll) < tl
        assert grad_err(r["grad_bl"], g) < tg
    return ll


def alignment(seqs):
    return dr.AlignmentRef(list(seqs), list(seqs.values()))


def test_polytomies_resolved_with_identity_edges():
    rng = np.random.default_rng(3)
    names = ["T%d" % i for i in range(11)]
    seqs = {n: "".join(rng.choice(list("ACG"), size=60)) for n in names}
    nwk = "(T0:0.1,T1:0.2,T2:0.05,(T3:0.1,T4:0.2,T5:0.3,T6:0.1,T7:0.15):0.2,(T8:0.3,T9:0.1):0.05,T10:0.4);"
    tree = dr.TreeRef.from_newick(nwk)
    aln = alignment(seqs)
    check_plain(tree, aln, 0.6, [1.2, 0.8, 1.0, 2.0, 0.7], branch_lengths(tree))


def test_ambiguity_gaps_stops_and_unmatched_leaf():
    seqs = {
        "A": "ATGAAACCCGGGTTTCATNNNAGA",
        "B": "ATGAARCCYGGG---CATACGAGA",
        "C": "ATRAAACC?GGGTTTCA-ACGTAA",   # a stop codon in the data -> all-sense mask
        "D": "ATGAAACCCGGNTTTCATACGAGA",
        "E": "AYGAAACCCGGGTTKCATACGWGA",
    }
    aln = alignment(seqs)
    nwk = "((A:0.1,B:0.2):0.05,(C:0.15,D:0.1):0.2,(E:0.3,Missing:0.2):0.1);"
    tree = dr.TreeRef.from_newick(nwk)
    assert -1 in [x for i, x in enumerate(tree.leaf_to_taxon(aln.names)) if tree.nodes[i].is_leaf]
    check_plain(tree, aln, 0.4, [1, 1, 1, 2, 1], branch_lengths(tree))


def test_many_distinct_ambiguity_masks():
    """More distinct partially ambiguous codons than the adjoint stages in LDS
    (512): its ambiguity-mask table is then read from HBM."""
    rng = np.random.default_rng(11)
    iupac = list("ACGTRYKMSWBDHVN")
    names = ["T%d" % i for i in range(10)]
    seqs = {}
    for n in names:
        cods = []
        for _ in range(300):
            c = [str(rng.choice(list("ACGT"))) for _ in range(3)]
            if rng.uniform() < 0.7:
                for k in rng.choice(3, size=int(rng.integers(1, 3)), replace=False):
                    c[k] = str(rng.choice(iupac[4:]))
            cods.append("".join(c))
        seqs[n] = "".join(cods)
    aln = alignment(seqs)
    m = aln.pmasks
    full = int(np.max(m))
    single = (m & (m - np.uint64(1))) == 0
    distinct = {int(x) for x in m[~single].ravel() if int(x) != full}
    assert len(distinct) > 512
    rng2 = np.random.default_rng(12)
    nwk = "(" + ",".join("%s:%.3f" % (n, rng2.uniform(0.05, 0.4)) for n in names[:4]) + ",(" + \
          ",".join("%s:%.3f" % (n, rng2.uniform(0.05, 0.4)) for n in names[4:]) + "):0.1);"
    tree = dr.TreeRef.from_newick(nwk)
    check_plain(tree, aln, 0.7, [1.1, 0.9, 1.3, 2.2, 0.8], branch_lengths(tree))


@pytest.mark.parametrize("n_codons", [1, 15, 17, 33, 250])
def test_ragged_pattern_counts(n_codons):
    tree, aln, sim = sim_problem(9, n_codons, 40 + n_codons)
    check_plain(tree, aln, 0.5, sim["theta"], branch_lengths(tree) * sim["conv"])


def test_deep_caterpillar_rescaling():
    n = 260
    rng = np.random.default_rng(5)
    nwk = "T0:2.0"
    for i in range(1, n - 1):
        nwk = "(%s,T%d:2.0):1.5" % (nwk, i)
    nwk = "(%s,T%d:2.0);" % (nwk, n - 1)
    codons = dr.Code().sense_codons
    seqs = {"T%d" % i: "".join(rng.choice(codons, size=40)) for i in range(n)}
    tree = dr.TreeRef.from_newick(nwk)
    ll = check_plain(tree, alignment(seqs), 0.5, [1, 1, 1, 2, 1], branch_lengths(tree))
    assert ll / 40 < -745.0  # per-site likelihoods far below the FP64 range


def test_loglik_from_P_with_zero_likelihood_pattern():
    """compute_tree_log_likelihood_from_P (likelihood.hpp:766-806) on
    caller-supplied matrices; identity P makes differing tips impossible, the
    pattern contributes weight * -1e20."""
    seqs = {"A": "AAAAAAAAC", "B": "AAAAAAAAA", "C": "AAACAAAAA"}
    aln = alignment(seqs)
    tree = dr.TreeRef.from_newick("(A:0.1,B:0.2,C:0.3);")
    prob = ho.Problem.from_refs(tree, aln)
    eng = engine_for(tree, aln)
    N = len(tree.nodes)
    rng = np.random.default_rng(2)
    P = rng.uniform(0.0, 1.0, size=(N, 61, 61))
    pi = rng.dirichlet(np.ones(61))
    ref, sref = ho.loglik_from_P(prob, P, pi, site=True)
    for prec, tl in ((FP64, 1e-11), (FP32, 1e-5)):
        ll, s = eng.loglik_from_P(P, pi, precision=prec, site=True)
        assert rel(ll, ref) < tl
        np.testing.assert_allclose(s, sref, rtol=tl * 10)
    Pid = np.broadcast_to(np.eye(61), (N, 61, 61)).copy()
    ref, sref = ho.loglik_from_P(prob, Pid, pi, site=True)
    ll, s = eng.loglik_from_P(Pid, pi, precision=FP64, site=True)
    assert (sref == -1e20).sum() == 2 and (s == -1e20).sum() == 2
    assert ll == ref


def test_two_class_test_background_mixture():
    """BUSTED-style partition: test branches mix 3 omegas, background
    branches 2 other matrices (busted.hpp:561-581)."""
    tree, aln, sim = sim_problem(14, 300, 77)
    N = len(tree.nodes)
    rng = np.random.default_rng(1)
    bc = (rng.uniform(size=N) < 0.4).astype(np.int32)
    omegas = [0.05, 0.9, 4.0, 0.3, 1.0]
    classes = [[(3, 0.7), (4, 0.3)], [(0, 0.5), (1, 0.4), (2, 0.1)]]
    bl = branch_lengths(tree)
    for conv in (0, 1):
        mix = ho.Mixture(omegas, bc, classes, conv)
        ref = ho.mixture_eval(ho.Problem.from_refs(tree, aln), ho.Code(0), ho.params(1.0, 1.0, sim["theta"]),
                              aln.pos_freqs, aln.cf3x4, mix, bl * (1.0 if conv else sim["conv"]))
        eng = engine_for(tree, aln)
        eng.set_mixture(5, bc, [[m for m, _ in c] for c in classes], convention=conv)
        w = [x for c in classes for _, x in c]
        for prec, tl, tg in ((FP64, 1e-9, 1e-7), (FP32, 1e-5, 1e-4)):
            r = eng.eval(bl * (1.0 if conv else sim["conv"]), omegas, w, sim["theta"], precision=prec)
            assert rel(r["log_likelihood"], ref["log_likelihood"]) < tl
            for k in ("grad_bl", "grad_omega", "grad_w", "grad_theta"):
                assert grad_err(r[k], ref[k]) < tg, (conv, prec, k)


def test_vertebrate_mitochondrial_code():
    code = ho.Code(1)
    assert code.S == 60
    dcode = dr.Code(bytes(code.c.aa).decode())
    rng = np.random.default_rng(9)
    names = ["S%d" % i for i in range(8)]
    seqs = {n: "".join(rng.choice(dcode.sense_codons, size=50)) for n in names}
    aln = dr.AlignmentRef(list(seqs), list(seqs.values()), code=dcode)
    tree = dr.TreeRef.from_newick("((S0:0.1,S1:0.2):0.1,(S2:0.1,S3:0.3):0.2,((S4:0.1,S5:0.1):0.1,(S6:0.2,S7:0.1):0.1):0.1);")
    check_plain(tree, aln, 0.3, [1.1, 0.9, 1.3, 2.2, 0.8], branch_lengths(tree), code_id=1)


def test_handle_reuse_across_problems():
    eng = None
    for seed, n, L in ((1, 10, 100), (2, 30, 700), (3, 6, 20)):
        tree, aln, sim = sim_problem(n, L, seed)
        if eng is None:
            eng = engine_for(tree, aln)
        else:
            from helpers import tree_arrays
            parent, children, l2t = tree_arrays(tree, aln)
            eng.set_problem(parent, children, l2t, aln.states, aln.pmasks, aln.weights, S=61)
            eng.set_model(0, aln.pos_freqs, aln.cf3x4)
        bl = branch_lengths(tree) * sim["conv"]
        prob = ho.Problem.from_refs(tree, aln)
        ll, g = ho.branch_length_gradients(prob, ho.Code(0), ho.params(1.0, 0.5, sim["theta"]), aln.pos_freqs,
                                           aln.cf3x4, bl)
        r = eng.eval(bl, [0.5], [1.0], sim["theta"], precision=FP64)
        assert rel(r["log_likelihood"], ll) < 1e-9 and grad_err(r["grad_bl"], g) < 1e-7


def test_medium_problem_fp64_vs_oracle_and_fp32():
    omegas, weights = [0.1, 0.8, 3.0], [0.6, 0.35, 0.05]
    tree, aln, sim = sim_problem(96, 1500, 505, omegas=omegas, weights=weights, gaps=0.1)
    N = len(tree.nodes)
    bl = branch_lengths(tree)
    mix = ho.Mixture(omegas, np.zeros(N, dtype=np.int32), [[(k, w) for k, w in enumerate(weights)]], 1)
    ref = ho.mixture_eval(ho.Problem.from_refs(tree, aln), ho.Code(0), ho.params(1.0, 1.0, sim["theta"]),
                          aln.pos_freqs, aln.cf3x4, mix, bl, threads=8)
    eng = engine_for(tree, aln)
    eng.set_mixture(3, np.zeros(N, dtype=np.int32), [[0, 1, 2]], convention=1)
    for prec, tl, tg in ((FP64, 1e-9, 1e-7), (FP32, 1e-5, 1e-4)):
        r = eng.eval(bl, omegas, weights, sim["theta"], precision=prec)
        assert rel(r["log_likelihood"], ref["log_likelihood"]) < tl
        for k in ("grad_bl", "grad_omega", "grad_w", "grad_theta"):
            assert grad_err(r[k], ref[k]) < tg, (prec, k)


def test_facade_and_autograd():
    import torch
    import _hyphy3
    from hyphy3 import autograd
    from helpers import GOLDEN
    import os
    aln = _hyphy3.Alignment.from_text(open(os.path.join(GOLDEN, "cd2_reduced.fna")).read())
    tree = _hyphy3.Tree.from_newick(open(os.path.join(GOLDEN, "cd2_reduced.nwk")).read(), False)
    p = _hyphy3.MG94Parameters()
    p.beta = 0.5
    p.theta_CT = 2.0
    ll = _hyphy3.compute_mg94_log_likelihood(tree, aln, p)
    assert rel(ll, -92.77744887734814) < 1e-8
    ll2, g = _hyphy3.compute_branch_length_gradients(tree, aln, p)
    assert rel(ll2, ll) < 1e-12
    tr, al = dr.TreeRef.from_newick(open(os.path.join(GOLDEN, "cd2_reduced.nwk")).read()), None
    names, seqs = dr.parse_fasta(open(os.path.join(GOLDEN, "cd2_reduced.fna")).read())
    al = dr.AlignmentRef(names, seqs)
    _, gref = ho.branch_length_gradients(ho.Problem.from_refs(tr, al), ho.Code(0), ho.params(1.0, 0.5, (1, 1, 1, 2, 1)),
                                         al.pos_freqs, al.cf3x4, branch_lengths(tr))
    assert grad_err(g, gref) < 1e-7
    bl = torch.tensor(tree.get_branch_lengths(), dtype=torch.float64, requires_grad=True)
    out = autograd.tree_log_likelihood(bl, tree, aln, p)
    out.backward()
    assert rel(out.item(), ll) < 1e-12
    assert grad_err(bl.grad.numpy(), gref) < 1e-7
    _hyphy3.set_precision("fp32")
    try:
        ll32 = _hyphy3.compute_mg94_log_likelihood(tree, aln, p)
        assert rel(ll32, ll) < 1e-5
    finally:
        _hyphy3.set_precision("fp64")
    # joint Frechet facade (tests/test_core.cpp:149-223 point)
    q = _hyphy3.MG94Parameters()
    q.beta = 0.45
    q.theta_AC, q.theta_AT, q.theta_CG, q.theta_CT, q.theta_GT = 0.8, 1.2, 0.6, 1.8, 0.9
    blj = [max(1e-4, x) for x in tree.get_branch_lengths()]
    r = _hyphy3.compute_joint_frechet_gradients(tree, aln, q, blj)
    blr = np.array([0.0 if n.id == tr.root else max(1e-4, n.bl) for n in tr.nodes])
    ref = ho.joint_frechet(ho.Problem.from_refs(tr, al), ho.Code(0), ho.params(1.0, 0.45, (0.8, 1.2, 0.6, 1.8, 0.9)),
                           al.pos_freqs, al.cf3x4, blr)
    assert rel(r["log_likelihood"], ref["log_likelihood"]) < 1e-9
    assert rel(r["grad_omega"], ref["grad_omega"]) < 1e-7
    assert grad_err(r["grad_theta"], ref["grad_theta"]) < 1e-7
    # mixture autograd: all four parameter groups
    om = torch.tensor([0.2, 1.5], dtype=torch.float64, requires_grad=True)
    w = torch.tensor([0.7, 0.3], dtype=torch.float64, requires_grad=True)
    th = torch.tensor([1.0, 1.0, 1.0, 2.0, 1.0], dtype=torch.float64, requires_grad=True)
    b2 = torch.tensor(tree.get_branch_lengths(), dtype=torch.float64, requires_grad=True)
    lm = autograd.mixture_log_likelihood(b2, om, w, th, tree, aln, time_convention=1)
    lm.backward()
    mix = ho.Mixture([0.2, 1.5], np.zeros(len(tr.nodes), dtype=np.int32), [[(0, 0.7), (1, 0.3)]], 1)
    mref = ho.mixture_eval(ho.Problem.from_refs(tr, al), ho.Code(0), ho.params(1.0, 1.0, (1, 1, 1, 2, 1)),
                           al.pos_freqs, al.cf3x4, mix, branch_lengths(tr))
    assert rel(lm.item(), mref["log_likelihood"]) < 1e-9
    assert grad_err(om.grad.numpy(), mref["grad_omega"]) < 1e-7
    assert grad_err(w.grad.numpy(), mref["grad_w"]) < 1e-7
    assert grad_err(th.grad.numpy(), mref["grad_theta"]) < 1e-7


def test_c_abi_error_paths():
    from hyphy3.engine import Engine, HyError
    eng = Engine(0)
    with pytest.raises(HyError) as e:
        eng.eval(np.zeros(3), [0.5], [1.0], [1, 1, 1, 1, 1])
    assert e.value.code == -3  # HY_ERR_STATE: no problem set
    with pytest.raises(HyError) as e:  # two roots
        eng.set_problem([-1, -1, 0], [[2], [], []], [-1, 0, 0], np.zeros((1, 1), np.int8), np.ones((1, 1), np.uint64),
                        [1.0])
    assert e.value.code == -1
    tree, aln, sim = sim_problem(5, 10, 1)
    eng = engine_for(tree, aln)
    with pytest.raises(HyError):
        eng.set_tile(24)
    with pytest.raises(HyError):
        eng.set_mixture(2, None, [[0, 5]])


def test_graph_replays_match_eager():
    """Captured evaluation graphs against eager launches: repeated replays with
    new parameters, interleaved eager evaluations, precision and tile switches
    and a mixture change (which invalidates the captured graphs); results must
    be bitwise identical."""
    import os
    omegas, weights = [0.2, 1.5], [0.7, 0.3]
    tree, aln, sim = sim_problem(20, 300, 17, omegas=omegas, weights=weights)
    N = len(tree.nodes)
    os.environ["HY_NO_GRAPH"] = "1"
    try:
        eager = make_engine(tree, aln)
    finally:
        del os.environ["HY_NO_GRAPH"]
    graph = make_engine(tree, aln)
    for e in (eager, graph):
        e.set_mixture(2, np.zeros(N, dtype=np.int32), [[0, 1]], convention=1)
    bl = branch_lengths(tree)
    for step in range(8):
        f = 1.0 + 0.01 * step
        prec = FP32 if step % 3 else FP64
        if step == 5:
            for e in (eager, graph):
                e.set_tile(32)
        if step == 6:
            for e in (eager, graph):
                e.set_mixture(2, np.zeros(N, dtype=np.int32), [[0, 1]], convention=0)
        args = (bl * f, [o * f for o in omegas], weights, np.asarray(sim["theta"]) / f)
        re = eager.eval(*args, precision=prec, site=(step == 2))
        rg = graph.eval(*args, precision=prec, site=(step == 2))
        assert re["log_likelihood"] == rg["log_likelihood"], step
        for k in ("grad_bl", "grad_omega", "grad_w", "grad_theta"):
            np.testing.assert_array_equal(re[k], rg[k])
        if step == 2:
            np.testing.assert_array_equal(re["site_ll"], rg["site_ll"])


def test_balanced_tree_lean_outside_spills():
    """A perfectly balanced 256-taxon tree: its subtree parts are six levels
    deep, so the FP32 lean outside pass (four LDS slots for pending messages)
    spills messages to HBM; FP64 and FP32 against the oracle."""
    rng = np.random.default_rng(17)

    def clade(lo, hi):
        if hi - lo == 1:
            return "T%d:0.08" % lo
        mid = (lo + hi) // 2
        return "(%s,%s):0.05" % (clade(lo, mid), clade(mid, hi))

    n, L = 256, 200
    nwk = "(%s,%s);" % (clade(0, n // 2), clade(n // 2, n))
    codons = dr.Code().sense_codons
    anc = rng.choice(codons, size=L)
    seqs = {}
    for i in range(n):
        s = [c if rng.uniform() > 0.15 else rng.choice(codons) for c in anc]
        seqs["T%d" % i] = "".join(s)
    tree = dr.TreeRef.from_newick(nwk)
    check_plain(tree, alignment(seqs), 0.4, [1.2, 0.8, 1.0, 2.0, 0.9], branch_lengths(tree))
