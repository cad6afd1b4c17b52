"""Pins of the oracle's nucleotide GTR stage (hyoracle.gtr_eval; the
reference's GTRMatrix, rate_matrix.hpp:14-104, and compute_gtr_log_likelihood,
likelihood.hpp:181-254), CPU only:

  * JC69 (all exchangeabilities 1, uniform frequencies) and Kimura's K80
    (A<->G = C<->T, the transversions another rate) two-taxon closed forms,
    which fix the normalisation to one expected substitution per unit time
    and which exchangeability index is which;
  * a brute-force sum over all ancestral states of a four-taxon tree with
    ambiguous tips, against a generator built from the nucleotide names;
  * rooting the same unrooted tree elsewhere leaves lnL unchanged;
  * central differences of the branch-length and exchangeability gradients.

The reference's cd2_reduced point truth also lists a GTR lnL
(-90.8383219324723), computed by HyPhy 2.5 at parameters its batch file does
not record (benchmarks/compute_point_lf.bf only fixes the MG94 ones); it
cannot pin this stage."""
import math

import numpy as np

from helpers import branch_lengths
from oracle import data_ref as dr
from oracle import hyoracle as ho

NAMES = "ACGT"


def nuc_problem_from(seqs, nwk, keep_rooted=True):
    aln = dr.AlignmentRef(list(seqs), list(seqs.values()))
    tree = dr.TreeRef.from_newick(nwk, keep_rooted=keep_rooted)
    naln = dr.NucleotideRef(aln)
    return tree, naln, ho.nuc_problem(tree, naln)


def two_taxon(a, b, t1, t2):
    """Single-site two-taxon problem (root, A, B) with states a, b."""
    states = np.array([[a, b]], dtype=np.int8)
    masks = np.array([[1 << a, 1 << b]], dtype=np.uint64)
    return ho.Problem([-1, 0, 0], [[1, 2], [], []], [1, 2, 0], [0, 1, 1], [-1, 0, 1], states, masks, [1.0], S=4)


def test_jc69_closed_form():
    t1, t2 = 0.07, 0.31
    t = t1 + t2
    for a, b in ((0, 0), (1, 3), (2, 0)):
        r = ho.gtr_eval(two_taxon(a, b, t1, t2), np.ones(5), np.full(4, 0.25), [0.0, t1, t2], want_grad=False)
        p = 0.25 + 0.75 * math.exp(-4.0 * t / 3.0) if a == b else 0.25 - 0.25 * math.exp(-4.0 * t / 3.0)
        assert abs(r["log_likelihood"] - math.log(0.25 * p)) < 1e-13


def test_k80_closed_form():
    """A<->G = C<->T = 1 (theta_CT = 1), transversions beta: normalised rates
    alpha + 2 beta_n = 1; P_same = 1/4 + 1/4 e^{-4 b t} + 1/2 e^{-2 (a + b) t},
    P_transition = 1/4 + 1/4 e^{-4 b t} - 1/2 e^{-2 (a + b) t},
    P_transversion = 1/4 - 1/4 e^{-4 b t} (Kimura 1980)."""
    beta = 0.35
    theta = [beta, beta, beta, 1.0, beta]  # AC, AT, CG, CT, GT
    mu = (1.0 + 2.0 * beta) / 4.0
    al, be = 0.25 / mu, 0.25 * beta / mu
    t1, t2 = 0.2, 0.45
    t = t1 + t2
    e4, e2 = math.exp(-4.0 * be * t), math.exp(-2.0 * (al + be) * t)
    for a, b in ((0, 0), (0, 2), (1, 3), (0, 1), (2, 3), (3, 0)):
        if a == b:
            p = 0.25 + 0.25 * e4 + 0.5 * e2
        elif {a, b} in ({0, 2}, {1, 3}):
            p = 0.25 + 0.25 * e4 - 0.5 * e2
        else:
            p = 0.25 - 0.25 * e4
        r = ho.gtr_eval(two_taxon(a, b, t1, t2), theta, np.full(4, 0.25), [0.0, t1, t2], want_grad=False)
        assert abs(r["log_likelihood"] - math.log(0.25 * p)) < 1e-13, (a, b)


def named_generator(theta, pi):
    """GTR generator from the exchangeabilities' names (rate_matrix.hpp:41-46),
    normalised to one expected substitution per unit time."""
    rate = {"AC": theta[0], "AT": theta[1], "CG": theta[2], "CT": theta[3], "GT": theta[4], "AG": 1.0}
    Q = np.zeros((4, 4))
    for i in range(4):
        for j in range(4):
            if i != j:
                Q[i, j] = rate["".join(sorted(NAMES[i] + NAMES[j], key=NAMES.index))] * pi[j]
        Q[i, i] = -Q[i].sum()
    return Q / -np.sum(pi * np.diag(Q))


def test_brute_force_four_taxa():
    seqs = {"A": "ACGTRNAGG", "B": "ACTTAGCCY", "C": "GCGTACAG-", "D": "ACGAACWGT"}
    tree, naln, prob = nuc_problem_from(seqs, "((A:0.11,B:0.23):0.07,C:0.31,D:0.05);")
    theta = [0.7, 1.9, 0.4, 2.6, 0.8]
    pi = np.array([0.31, 0.17, 0.22, 0.30])
    r = ho.gtr_eval(prob, theta, pi, branch_lengths(tree), want_grad=False)
    Q = named_generator(theta, pi)
    P = {n.id: ho.expm_taylor(Q, n.bl) for n in tree.nodes if n.id != tree.root}
    l2t = tree.leaf_to_taxon(naln.names)
    root = tree.nodes[tree.root]
    u = [c for c in root.children if not tree.nodes[c].is_leaf][0]
    lu, lr = tree.nodes[u].children, [c for c in root.children if c != u]
    tot = 0.0
    for k in range(naln.num_patterns):
        x = {}
        for n in lu + lr:
            st, m = naln.states[k, l2t[n]], int(naln.pmasks[k, l2t[n]])
            x[n] = np.eye(4)[st] if st >= 0 else np.array([(m >> i) & 1 for i in range(4)], dtype=float)
        a_u = (P[lu[0]] @ x[lu[0]]) * (P[lu[1]] @ x[lu[1]])
        c_r = (P[lr[0]] @ x[lr[0]]) * (P[lr[1]] @ x[lr[1]])
        L = math.fsum((pi[:, None] * c_r[:, None] * P[u] * a_u[None, :]).ravel())
        assert abs(math.log(L) - r["site_ll"][k]) < 1e-12
        tot += naln.weights[k] * math.log(L)
    assert abs(tot - r["log_likelihood"]) < 1e-11


def test_rooting_invariance():
    seqs = {"A": "ACGTACGTAAGT", "B": "ACGTTCGAAAGT", "C": "ACCTACGTAAGC", "D": "TCGTACGGAAGT",
            "E": "ACGTACGTCAGA"}
    theta = [0.6, 1.4, 0.9, 3.0, 0.5]
    pi = np.array([0.27, 0.23, 0.26, 0.24])
    lls = []
    for nwk in ("((A:0.1,B:0.2):0.3,C:0.4,(D:0.15,E:0.25):0.05);", "(A:0.1,B:0.2,(C:0.4,(D:0.15,E:0.25):0.05):0.3);"):
        tree, naln, prob = nuc_problem_from(seqs, nwk)
        lls.append(ho.gtr_eval(prob, theta, pi, branch_lengths(tree), want_grad=False)["log_likelihood"])
    assert abs(lls[0] - lls[1]) < 1e-11


def test_gradients_central_differences():
    rng = np.random.default_rng(12)
    n = 12
    from test_gpu_sweep import random_newick
    names = ["T%d" % i for i in range(n)]
    seqs = {nm: "".join(rng.choice(list("ACGT"), size=60)) for nm in names}
    seqs["T3"] = "N" * 6 + seqs["T3"][6:]
    tree, naln, prob = nuc_problem_from(seqs, random_newick(rng, n))
    theta = rng.uniform(0.4, 2.5, size=5)
    pi = rng.dirichlet(np.ones(4) * 5)
    bl = branch_lengths(tree)
    r = ho.gtr_eval(prob, theta, pi, bl)
    f = lambda b, th: ho.gtr_eval(prob, th, pi, b, want_grad=False)["log_likelihood"]  # noqa: E731
    h = 1e-6
    for b in range(len(bl)):
        if b == tree.root:
            continue
        bp, bm = bl.copy(), bl.copy()
        bp[b] += h
        bm[b] -= h
        fd = (f(bp, theta) - f(bm, theta)) / (2 * h)
        assert abs(r["grad_bl"][b] - fd) / (abs(fd) + 1) < 1e-6, (b, r["grad_bl"][b], fd)
    for k in range(5):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (f(bl, tp) - f(bl, tm)) / (2 * h)
        assert abs(r["grad_theta"][k] - fd) / (abs(fd) + 1) < 1e-6, (k, r["grad_theta"][k], fd)
