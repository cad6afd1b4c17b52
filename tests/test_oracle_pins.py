"""Pins of the FP64 CPU oracle (CPU only).

Each test checks the oracle against something other than its own formula:
the reference's golden vectors and known answers (cd2_reduced point truth,
tests/test_core.cpp counts and invariants), closed forms, brute-force sums
over all ancestral states, an independent Taylor-series exponential and
central finite differences.  A plausible slip (dropped term, transposed
operand, wrong sign or index) breaks at least one of them.
"""
import json
import math
import os

import numpy as np
import pytest

from helpers import GOLDEN, branch_lengths, load_cd2_reduced, sim_problem
from oracle import data_ref as dr
from oracle import hyoracle as ho

CODE = ho.Code(0)
GOLDEN_MG94 = -92.77744887734814


def point_truth():
    d = json.load(open(os.path.join(GOLDEN, "cd2_reduced_point_truth.json")))
    codons = d["sense_codons"]
    d["sense_codons"] = codons[0] if codons and isinstance(codons[0], list) else codons
    return d


# ---------------------------------------------------------------------------
# a1 / a2: genetic code numbering, frequencies, compression
# ---------------------------------------------------------------------------
def test_sense_codon_order_matches_golden():
    assert dr.Code().sense_codons == point_truth()["sense_codons"]
    assert CODE.S == 61
    c = CODE.c
    order = ["".join("ACGT"[(c.sense_to_c64[i] >> s) & 3] for s in (4, 2, 0)) for i in range(61)]
    assert order == point_truth()["sense_codons"]


def test_other_codes_sense_counts():
    # NCBI table facts: vertebrate mitochondrial (id 1 here) has AGA/AGG as
    # stops and TGA as Trp -> 60 sense codons; ciliate nuclear code (TAA, TAG
    # = Gln) -> 63.
    assert ho.Code(1).S == 60
    c1 = ho.Code(1).c
    tga = 16 * 3 + 4 * 2 + 0
    aga = 16 * 0 + 4 * 2 + 0
    assert c1.c64_to_sense[tga] >= 0 and c1.c64_to_sense[aga] < 0
    counts = [ho.Code(i).S for i in range(22)]
    assert all(58 <= s <= 64 for s in counts)


def test_cd2_reduced_counts_and_frequencies():
    """tests/test_core.cpp:12-35."""
    _, aln = load_cd2_reduced()
    assert aln.num_taxa == 10 and aln.num_codons == 6
    assert aln.weights.sum() == 6 and aln.num_patterns <= 6
    np.testing.assert_allclose(aln.nuc_freqs, [0.456140, 0.114035, 0.254386, 0.175439], rtol=1e-4)


def test_cf3x4_fixed_point():
    """The corrected position frequencies reproduce the observed position
    marginals through the stop-free F3x4 product (alignment.hpp:962-1003)."""
    _, aln, _ = sim_problem(8, 300, 3)
    cur = aln.pos_freqs
    nucs = [((c >> 4) & 3, (c >> 2) & 3, c & 3) for c in aln.code.sense]
    f = np.array([cur[0, a] * cur[1, b] * cur[2, c] for a, b, c in nucs])
    f /= f.sum()
    np.testing.assert_allclose(f, aln.cf3x4, rtol=1e-12)
    obs = np.zeros((3, 4))
    for s in aln.seqs:
        for i, ch in enumerate(s):
            n = dr.char_to_nuc(ch)
            if n >= 0:
                obs[i % 3, n] += 1
    obs /= obs.sum(1, keepdims=True)
    marg = np.zeros((3, 4))
    for k, (a, b, c) in enumerate(nucs):
        marg[0, a] += f[k]
        marg[1, b] += f[k]
        marg[2, c] += f[k]
    np.testing.assert_allclose(marg, obs, atol=1e-10)


def test_pattern_compression_brute_force():
    _, aln, _ = sim_problem(6, 400, 5, gaps=0.2)
    cols = {}
    order = []
    for c in range(aln.num_codons):
        key = tuple(aln.codons[:, c]) + tuple(aln.masks[:, c])
        if key not in cols:
            cols[key] = 0
            order.append(key)
        cols[key] += 1
    assert aln.num_patterns == len(order)
    for p, key in enumerate(order):
        assert aln.weights[p] == cols[key]
        assert tuple(aln.states[p]) == key[:aln.num_taxa]
    for c in range(aln.num_codons):
        p = aln.site_to_pattern[c]
        assert tuple(aln.states[p]) == tuple(aln.codons[:, c])


# ---------------------------------------------------------------------------
# a4 / a5: rate matrix and transition matrices
# ---------------------------------------------------------------------------
PARAM_SETS = [(1.0, 0.5, (1, 1, 1, 2, 1), 0, 0), (0.7, 2.3, (0.8, 1.2, 0.6, 1.8, 0.9), 0, 0),
              (1.3, 0.1, (1.5, 0.5, 2.0, 0.7, 1.1), 0.05, 0.01)]


def model(pset, seed=0):
    rng = np.random.default_rng(seed)
    pos = rng.dirichlet([5, 5, 5, 5], size=3)
    nucs = [((c >> 4) & 3, (c >> 2) & 3, c & 3) for c in dr.Code().sense]
    pi = np.array([pos[0, a] * pos[1, b] * pos[2, c] for a, b, c in nucs])
    pi /= pi.sum()
    p = ho.params(*pset[:3], delta=pset[3], psi=pset[4])
    Q, mu = ho.mg94_Q(CODE, p, pos, pi)
    return Q, mu, pi, pos, p


@pytest.mark.parametrize("k", range(len(PARAM_SETS)))
def test_q_invariants(k):
    """tests/test_core.cpp:54-93 plus detailed balance and the unnormalised mu."""
    Q, mu, pi, pos, p = model(PARAM_SETS[k], k)
    assert np.abs(Q.sum(1)).max() < 1e-12
    F = pi[:, None] * Q
    assert np.abs(F - F.T).max() < 1e-14
    assert abs(mu + (pi * np.diag(Q)).sum()) < 1e-14
    off = Q - np.diag(np.diag(Q))
    assert (off >= 0).all()
    # without multi-hit rates only one-nucleotide neighbours exchange, with the
    # MG94 rate alpha|beta * theta(from, to) * pi_pos(to) (rate_matrix.hpp:165-220)
    cod = dr.Code().sense_codons
    aa = dr.Code().aa64
    alpha, beta, th = PARAM_SETS[k][0], PARAM_SETS[k][1], PARAM_SETS[k][2]
    pair = {(0, 1): th[0], (0, 3): th[1], (1, 2): th[2], (1, 3): th[3], (2, 3): th[4], (0, 2): 1.0}
    for i in range(61):
        for j in range(61):
            diff = [q for q in range(3) if cod[i][q] != cod[j][q]]
            if i == j or (len(diff) > 1 and PARAM_SETS[k][3] == 0 and PARAM_SETS[k][4] == 0):
                if i != j:
                    assert Q[i, j] == 0.0
                continue
            if len(diff) == 1:
                q = diff[0]
                a, b = "ACGT".index(cod[i][q]), "ACGT".index(cod[j][q])
                ci = 16 * "ACGT".index(cod[i][0]) + 4 * "ACGT".index(cod[i][1]) + "ACGT".index(cod[i][2])
                cj = 16 * "ACGT".index(cod[j][0]) + 4 * "ACGT".index(cod[j][1]) + "ACGT".index(cod[j][2])
                rate = alpha if aa[ci] == aa[cj] else beta
                expect = rate * pair[(min(a, b), max(a, b))] * pos[q, b]
                assert abs(Q[i, j] - expect) < 1e-14 * max(1.0, expect)


def q_from_strings(alpha, beta, th, delta, psi, pos):
    """MG94 generator built from the codon strings alone (rate_matrix.hpp:
    190-223): an entry for every pair differing at n positions is
    alpha|beta (synonymous or not, from the standard amino-acid table) times
    1, delta or psi for n = 1, 2, 3, times the product over the differing
    positions q of theta(from, to) * pi_q(to); theta_AG = 1; diagonal minus
    the row sum."""
    cod = dr.Code().sense_codons
    aa = dr.Code().aa64
    idx = lambda c: 16 * "ACGT".index(c[0]) + 4 * "ACGT".index(c[1]) + "ACGT".index(c[2])  # noqa: E731
    pair = {(0, 1): th[0], (0, 3): th[1], (1, 2): th[2], (1, 3): th[3], (2, 3): th[4], (0, 2): 1.0}
    hit = {1: 1.0, 2: delta, 3: psi}
    S = len(cod)
    Q = np.zeros((S, S))
    for i in range(S):
        for j in range(S):
            diff = [q for q in range(3) if cod[i][q] != cod[j][q]]
            if not diff:
                continue
            v = (alpha if aa[idx(cod[i])] == aa[idx(cod[j])] else beta) * hit[len(diff)]
            for q in diff:
                a, b = "ACGT".index(cod[i][q]), "ACGT".index(cod[j][q])
                v *= pair[(min(a, b), max(a, b))] * pos[q, b]
            Q[i, j] = v
        Q[i, i] = -Q[i].sum()
    return Q


@pytest.mark.parametrize("delta,psi", [(0.05, 0.0), (0.0, 0.02), (0.3, 0.07)])
def test_q_multihit_entries(delta, psi):
    """Two- and three-nucleotide entries (rate_matrix.hpp:205-219), the
    diagonal and mu against the string-built generator."""
    rng = np.random.default_rng(17)
    pos = rng.dirichlet([5, 5, 5, 5], size=3)
    nucs = [((c >> 4) & 3, (c >> 2) & 3, c & 3) for c in dr.Code().sense]
    pi = np.array([pos[0, a] * pos[1, b] * pos[2, c] for a, b, c in nucs])
    pi /= pi.sum()
    alpha, beta, th = 1.3, 0.4, (1.5, 0.5, 2.0, 0.7, 1.1)
    Q, mu = ho.mg94_Q(CODE, ho.params(alpha, beta, th, delta=delta, psi=psi), pos, pi)
    Qs = q_from_strings(alpha, beta, th, delta, psi, pos)
    np.testing.assert_allclose(Q, Qs, rtol=1e-14, atol=1e-17)
    assert abs(mu + (pi * np.diag(Qs)).sum()) < 1e-14
    cod = dr.Code().sense_codons
    n2 = sum(1 for i in range(61) for j in range(61) if sum(a != b for a, b in zip(cod[i], cod[j])) == 2)
    n3 = sum(1 for i in range(61) for j in range(61) if sum(a != b for a, b in zip(cod[i], cod[j])) == 3)
    assert (np.count_nonzero(Q) - 61 == 61 * 60 - (0 if delta else n2) - (0 if psi else n3))


def test_multihit_loglik_brute_force():
    """lnL with delta, psi > 0 on a four-taxon tree against the sum over all
    ancestral states with P = exp(Q t) of the string-built generator
    (independent Taylor exponential)."""
    nwk = "((A:0.21,B:0.13):0.17,C:0.31,D:0.45);"
    seqs = {"A": "ATGAAACCCGGGTTTCAT", "B": "CTGAAGCCAGGACTTGAT", "C": "ATAAGACGCGGTTTCGCT",
            "D": "TTGCAACCTGCATTTGCA"}
    aln = dr.AlignmentRef(list(seqs), list(seqs.values()))
    tree = dr.TreeRef.from_newick(nwk)
    prob = ho.Problem.from_refs(tree, aln)
    alpha, beta, th, delta, psi = 1.1, 0.6, (1.2, 0.8, 0.9, 2.5, 1.3), 0.25, 0.08
    p = ho.params(alpha, beta, th, delta=delta, psi=psi)
    bl = branch_lengths(tree)
    ll, site = ho.mg94_loglik(prob, CODE, p, aln.pos_freqs, aln.cf3x4, bl, site=True)
    Qs = q_from_strings(alpha, beta, th, delta, psi, aln.pos_freqs)
    P = {n.id: ho.expm_taylor(Qs, n.bl) for n in tree.nodes if n.id != tree.root}
    l2t = tree.leaf_to_taxon(aln.names)
    root = tree.nodes[tree.root]
    u = [c for c in root.children if not tree.nodes[c].is_leaf][0]
    lu, lr = tree.nodes[u].children, [c for c in root.children if c != u]
    tot = 0.0
    for k in range(aln.num_patterns):
        x = {n: tip_vec(aln.states[k, l2t[n]], aln.pmasks[k, l2t[n]]) for n in lu + lr}
        a_u = (P[lu[0]] @ x[lu[0]]) * (P[lu[1]] @ x[lu[1]])
        c_r = (P[lr[0]] @ x[lr[0]]) * (P[lr[1]] @ x[lr[1]])
        L = math.fsum((aln.cf3x4[:, None] * c_r[:, None] * P[u] * a_u[None, :]).ravel())
        assert abs(math.log(L) - site[k]) < 1e-11
        tot += aln.weights[k] * math.log(L)
    assert abs(tot - ll) < 1e-10
    # the multi-hit terms matter here: without them lnL differs
    ll0 = ho.mg94_loglik(prob, CODE, ho.params(alpha, beta, th), aln.pos_freqs, aln.cf3x4, bl)
    assert abs(ll0 - ll) > 1e-3


@pytest.mark.parametrize("k", range(len(PARAM_SETS)))
def test_transition_invariants(k):
    Q, mu, pi, pos, p = model(PARAM_SETS[k], k)
    ev, U, V = ho.decompose(Q, pi)
    assert (ev <= 0).all()
    P0 = ho.transition(ev, U, V, 0.0)
    assert np.abs(P0 - np.eye(61)).max() < 1e-11
    for t in (0.01, 0.3, 2.0):
        P = ho.transition(ev, U, V, t)
        assert np.abs(P.sum(1) - 1).max() < 1e-10
        assert np.abs(pi @ P - pi).max() < 1e-12
        # independent exponential (Taylor series with scaling and squaring)
        np.testing.assert_allclose(P, ho.expm_taylor(Q, t), atol=1e-11)
    Ps, Pt, Pst = (ho.transition(ev, U, V, x) for x in (0.2, 0.5, 0.7))
    assert np.abs(Ps @ Pt - Pst).max() < 1e-12


def test_sym_eigen_matches_lapack():
    rng = np.random.default_rng(1)
    A = rng.normal(size=(61, 61))
    A = A + A.T
    ev, Z = ho.sym_eigen(A)
    np.testing.assert_allclose(np.sort(ev), np.linalg.eigvalsh(A), atol=1e-11)
    np.testing.assert_allclose(Z @ np.diag(ev) @ Z.T, A, atol=1e-11)


# ---------------------------------------------------------------------------
# a7 - a10: pruning
# ---------------------------------------------------------------------------
def test_golden_point_mg94_logl():
    """benchmarks/ground_truth/cd2_reduced_point_truth.json via
    tests/test_core.cpp:95-121 (there at 1e-4)."""
    tree, aln = load_cd2_reduced()
    prob = ho.Problem.from_refs(tree, aln)
    ll = ho.mg94_loglik(prob, CODE, ho.params(1.0, 0.5, (1, 1, 1, 2, 1)), aln.pos_freqs, aln.cf3x4,
                        branch_lengths(tree))
    # HyPhy 2.5 computed the golden value; agreement is 3.2e-10 relative
    assert abs(ll - GOLDEN_MG94) / abs(GOLDEN_MG94) < 1e-9


def tip_vec(state, mask):
    if state >= 0:
        v = np.zeros(61)
        v[state] = 1.0
        return v
    return np.array([(int(mask) >> i) & 1 for i in range(61)], dtype=float)


def test_brute_force_four_taxa():
    """((A,B)u, C, D)r: L_p = sum over all ancestral states (r, u) of
    pi_r P_ru P_uA P_uB P_rC P_rD evaluated on the tip indicators."""
    nwk = "((A:0.11,B:0.23):0.07,C:0.31,D:0.05);"
    seqs = {"A": "ATGAAACCCGGGTTTNNN", "B": "ATGAAGCCAGGG---CAT", "C": "ATAAARCCCGGTTTCGCT",
            "D": "CTGAAACCTGGATTTGCA"}
    aln = dr.AlignmentRef(list(seqs), list(seqs.values()))
    tree = dr.TreeRef.from_newick(nwk)
    prob = ho.Problem.from_refs(tree, aln)
    p = ho.params(1.1, 0.6, (1.2, 0.8, 0.9, 2.5, 1.3))
    bl = branch_lengths(tree)
    ll, site = ho.mg94_loglik(prob, CODE, p, aln.pos_freqs, aln.cf3x4, bl, site=True)
    Q, _ = ho.mg94_Q(CODE, p, aln.pos_freqs, aln.cf3x4)
    P = {n.id: ho.expm_taylor(Q, n.bl) for n in tree.nodes if n.id != tree.root}
    l2t = tree.leaf_to_taxon(aln.names)
    root = tree.nodes[tree.root]
    u = [c for c in root.children if not tree.nodes[c].is_leaf][0]
    leaves_u = tree.nodes[u].children
    leaves_r = [c for c in root.children if c != u]
    tot = 0.0
    for k in range(aln.num_patterns):
        x = {n: tip_vec(aln.states[k, l2t[n]], aln.pmasks[k, l2t[n]]) for n in leaves_u + leaves_r}
        # tip factors are sums over the observed-compatible states only; the
        # ancestral pair (r, u) is enumerated explicitly (61 x 61 terms)
        a_u = (P[leaves_u[0]] @ x[leaves_u[0]]) * (P[leaves_u[1]] @ x[leaves_u[1]])
        c_r = (P[leaves_r[0]] @ x[leaves_r[0]]) * (P[leaves_r[1]] @ x[leaves_r[1]])
        terms = aln.cf3x4[:, None] * c_r[:, None] * P[u] * a_u[None, :]
        L = math.fsum(terms.ravel())
        assert abs(math.log(L) - site[k]) < 1e-11
        tot += aln.weights[k] * math.log(L)
    assert abs(tot - ll) < 1e-10


def two_taxon_problem(t1, t2, sa, sb):
    states = np.array([[sa, sb]], dtype=np.int8)
    masks = np.array([[1 << sa, 1 << sb]], dtype=np.uint64)
    # node 0 root, 1 = A, 2 = B
    return ho.Problem([-1, 0, 0], [[1, 2], [], []], [1, 2, 0], [0, 1, 1], [-1, 0, 1], states, masks, [1.0])


def test_two_taxon_closed_form():
    """Reversibility: sum_i pi_i P_ia(t1) P_ib(t2) = pi_a P_ab(t1 + t2)."""
    Q, mu, pi, pos, p = model(PARAM_SETS[1], 3)
    for sa, sb in ((0, 0), (5, 17), (60, 2)):
        prob = two_taxon_problem(0.13, 0.29, sa, sb)
        ll = ho.mg94_loglik(prob, CODE, p, pos, pi, [0.0, 0.13, 0.29])
        Pt = ho.expm_taylor(Q, 0.42)
        assert abs(ll - math.log(pi[sa] * Pt[sa, sb])) < 1e-11


def test_pulley_principle():
    """Rooting the same unrooted tree at different nodes leaves lnL unchanged."""
    seqs = {"A": "ATGAAACCCGGGTTTCAT", "B": "ATGAAGCCAGGGTTCCAT", "C": "ATAAAACCCGGTTTCGCT",
            "D": "CTGAAACCTGGATTTGCA", "E": "ATGCAACCTGGATTTGCA"}
    aln = dr.AlignmentRef(list(seqs), list(seqs.values()))
    lls = []
    for nwk in ("((A:0.1,B:0.2):0.3,C:0.4,(D:0.15,E:0.25):0.05);", "(A:0.1,B:0.2,(C:0.4,(D:0.15,E:0.25):0.05):0.3);",
                "((A:0.1,B:0.2):0.3,(C:0.4,(D:0.15,E:0.25):0.05):0.0);"):
        tree = dr.TreeRef.from_newick(nwk, keep_rooted=True)
        prob = ho.Problem.from_refs(tree, aln)
        lls.append(ho.mg94_loglik(prob, CODE, ho.params(1.0, 0.4, (1, 2, 1, 3, 1)), aln.pos_freqs, aln.cf3x4,
                                  branch_lengths(tree)))
    assert abs(lls[0] - lls[1]) < 1e-10 and abs(lls[0] - lls[2]) < 1e-10


def test_compressed_equals_uncompressed():
    tree, aln0, sim = sim_problem(10, 200, 9, gaps=0.1)
    # repeat a third of the columns so that compression merges them
    aln = dr.AlignmentRef(aln0.names, [s + s[:180] for s in aln0.seqs])
    prob = ho.Problem.from_refs(tree, aln)
    p = ho.params(1.0, 0.5, sim["theta"])
    bl = branch_lengths(tree) * sim["conv"]
    ll = ho.mg94_loglik(prob, CODE, p, aln.pos_freqs, aln.cf3x4, bl)
    full = ho.Problem([n.parent for n in tree.nodes], [list(n.children) for n in tree.nodes], tree.post_order,
                      [1 if n.is_leaf else 0 for n in tree.nodes], tree.leaf_to_taxon(aln.names),
                      np.ascontiguousarray(aln.codons.T), np.ascontiguousarray(aln.masks.T),
                      np.ones(aln.num_codons))
    ll_full = ho.mg94_loglik(full, CODE, p, aln.pos_freqs, aln.cf3x4, bl)
    assert abs(ll - ll_full) < 1e-9 * abs(ll)
    assert aln.num_patterns < aln.num_codons


def test_rescaling_on_deep_tree():
    """A 300-taxon caterpillar whose unscaled partials underflow FP64:
    matches an independent log-space pruning (logsumexp per node)."""
    n = 300
    rng = np.random.default_rng(4)
    nwk = "T0:0.9"
    for i in range(1, n - 1):
        nwk = "(%s,T%d:0.9):0.8" % (nwk, i)
    nwk = "(%s,T%d:0.9);" % (nwk, n - 1)
    seqs = ["".join(rng.choice(list("ACGT"), size=6)) for _ in range(n)]
    seqs = [s.replace("TAA", "TTA").replace("TAG", "TTG").replace("TGA", "TTA") for s in seqs]
    aln = dr.AlignmentRef(["T%d" % i for i in range(n)], seqs)
    tree = dr.TreeRef.from_newick(nwk, keep_rooted=True)
    prob = ho.Problem.from_refs(tree, aln)
    p = ho.params(1.0, 0.5, (1, 1, 1, 2, 1))
    bl = branch_lengths(tree)
    ll, site = ho.mg94_loglik(prob, CODE, p, aln.pos_freqs, aln.cf3x4, bl, site=True)
    Q, _ = ho.mg94_Q(CODE, p, aln.pos_freqs, aln.cf3x4)
    logP = {}
    l2t = tree.leaf_to_taxon(aln.names)
    for k in range(aln.num_patterns):
        logD = {}
        for u in tree.post_order:
            node = tree.nodes[u]
            if node.is_leaf:
                v = tip_vec(aln.states[k, l2t[u]], aln.pmasks[k, l2t[u]])
                with np.errstate(divide="ignore"):
                    logD[u] = np.log(v)
                continue
            acc = np.zeros(61)
            for c in node.children:
                b = tree.nodes[c].bl
                if b not in logP:
                    with np.errstate(divide="ignore"):
                        logP[b] = np.log(np.maximum(ho.expm_taylor(Q, b), 0.0))
                m = logP[b] + logD[c][None, :]
                mx = m.max(1)
                acc += mx + np.log(np.exp(m - mx[:, None]).sum(1))
            logD[u] = acc
        r = logD[tree.root] + np.log(aln.cf3x4)
        mx = r.max()
        lk = mx + math.log(np.exp(r - mx).sum())
        assert lk < -745  # unscaled FP64 would underflow
        assert abs(lk - site[k]) < 1e-9 * abs(lk)


def test_loglik_from_P_brute_force_random_matrices():
    """compute_tree_log_likelihood_from_P on arbitrary non-negative matrices
    (not from a model) against the full ancestral-state sum."""
    rng = np.random.default_rng(8)
    seqs = {"A": "ATGAAA", "B": "ATGAAG", "C": "CTAAAR"}
    aln = dr.AlignmentRef(list(seqs), list(seqs.values()))
    tree = dr.TreeRef.from_newick("(A:0.1,B:0.2,C:0.3);")
    prob = ho.Problem.from_refs(tree, aln)
    P = rng.uniform(0.0, 1.0, size=(len(tree.nodes), 61, 61))
    pi = rng.dirichlet(np.ones(61))
    ll = ho.loglik_from_P(prob, P, pi)
    l2t = tree.leaf_to_taxon(aln.names)
    ch = tree.nodes[tree.root].children
    tot = 0.0
    for k in range(aln.num_patterns):
        xs = [tip_vec(aln.states[k, l2t[c]], aln.pmasks[k, l2t[c]]) for c in ch]
        L = math.fsum(pi * (P[ch[0]] @ xs[0]) * (P[ch[1]] @ xs[1]) * (P[ch[2]] @ xs[2]))
        tot += aln.weights[k] * math.log(L)
    assert abs(ll - tot) < 1e-10 * abs(tot)


# ---------------------------------------------------------------------------
# a9, a11 - a13: gradients
# ---------------------------------------------------------------------------
def fd(f, x, h):
    return (f(x + h) - f(x - h)) / (2 * h)


def test_branch_gradients_fd_and_scale_identity():
    tree, aln, sim = sim_problem(8, 120, 21)
    prob = ho.Problem.from_refs(tree, aln)
    p = ho.params(1.0, 0.5, sim["theta"])
    bl = branch_lengths(tree) * sim["conv"]
    ll, g = ho.branch_length_gradients(prob, CODE, p, aln.pos_freqs, aln.cf3x4, bl)
    assert abs(ll - ho.mg94_loglik(prob, CODE, p, aln.pos_freqs, aln.cf3x4, bl)) < 1e-10 * abs(ll)
    for b in range(len(bl)):
        if b == tree.root:
            assert g[b] == 0.0
            continue
        e = np.zeros(len(bl))
        e[b] = 1.0
        num = fd(lambda s: ho.mg94_loglik(prob, CODE, p, aln.pos_freqs, aln.cf3x4, bl + s * e), 0.0, 1e-6)
        assert abs(g[b] - num) < 1e-5 * max(1.0, abs(num))
    # sum_b t_b dlnL/dt_b = d lnL(s t)/ds at s = 1
    num = fd(lambda s: ho.mg94_loglik(prob, CODE, p, aln.pos_freqs, aln.cf3x4, bl * s), 1.0, 1e-6)
    assert abs((bl * g).sum() - num) < 1e-5 * abs(num)


def test_joint_frechet_fd():
    """tests/test_core.cpp:149-223: omega 0.45, theta (0.8,1.2,0.6,1.8,0.9),
    times bl*3/mu; FD eps 1e-6."""
    tree, aln = load_cd2_reduced()
    prob = ho.Problem.from_refs(tree, aln)
    th0 = np.array([0.8, 1.2, 0.6, 1.8, 0.9])
    bl = np.array([0.0 if n.id == tree.root else max(1e-4, n.bl) for n in tree.nodes])
    res = ho.joint_frechet(prob, CODE, ho.params(1.0, 0.45, th0), aln.pos_freqs, aln.cf3x4, bl)

    def lnl(beta=0.45, th=th0, b=bl):
        p = ho.params(1.0, beta, th)
        _, mu = ho.mg94_Q(CODE, p, aln.pos_freqs, aln.cf3x4)
        return ho.mg94_loglik(prob, CODE, p, aln.pos_freqs, aln.cf3x4, b * 3.0 / mu)

    assert abs(res["log_likelihood"] - lnl()) < 1e-10 * abs(lnl())
    assert abs(res["grad_omega"] - fd(lambda x: lnl(beta=x), 0.45, 1e-6)) < 1e-5 * (1 + abs(res["grad_omega"]))
    for k in range(5):
        e = np.zeros(5)
        e[k] = 1
        num = fd(lambda x: lnl(th=th0 + x * e), 0.0, 1e-6)
        assert abs(res["grad_theta"][k] - num) < 1e-5 * (1 + abs(num))
    for b in range(len(bl)):
        if b == tree.root:
            continue
        e = np.zeros(len(bl))
        e[b] = 1
        num = fd(lambda x: lnl(b=bl + x * e), 0.0, 1e-7)
        assert abs(res["grad_branches"][b] - num) < 1e-4 * (1 + abs(num))


@pytest.mark.parametrize("convention", [0, 1])
def test_mixture_gradients_fd(convention):
    omegas, weights = [0.1, 0.8, 3.0], [0.6, 0.35, 0.05]
    tree, aln, sim = sim_problem(6, 80, 33, omegas=omegas, weights=weights)
    prob = ho.Problem.from_refs(tree, aln)
    N = len(tree.nodes)
    bl = branch_lengths(tree) * (1.0 if convention else sim["conv"])
    th0 = np.asarray(sim["theta"], dtype=float)

    def run(om=omegas, w=weights, th=th0, b=bl, grad=False):
        mix = ho.Mixture(om, np.zeros(N, dtype=np.int32), [[(k, x) for k, x in enumerate(w)]], convention)
        return ho.mixture_eval(prob, CODE, ho.params(1.0, 1.0, th), aln.pos_freqs, aln.cf3x4, mix, b, want_grad=grad)

    ref = run(grad=True)
    ll = lambda **kw: run(**kw)["log_likelihood"]  # noqa: E731
    for k in range(3):
        e = np.zeros(3)
        e[k] = 1
        num = fd(lambda x: ll(om=list(np.array(omegas) + x * e)), 0.0, 1e-6)
        assert abs(ref["grad_omega"][k] - num) < 1e-5 * (1 + abs(num))
        num = fd(lambda x: ll(w=list(np.array(weights) + x * e)), 0.0, 1e-6)
        assert abs(ref["grad_w"][k] - num) < 1e-5 * (1 + abs(num))
    for k in range(5):
        e = np.zeros(5)
        e[k] = 1
        num = fd(lambda x: ll(th=th0 + x * e), 0.0, 1e-6)
        assert abs(ref["grad_theta"][k] - num) < 1e-5 * (1 + abs(num))
    for b in range(N):
        if b == tree.root:
            continue
        e = np.zeros(N)
        e[b] = 1
        num = fd(lambda x: ll(b=bl + x * e), 0.0, 1e-7)
        assert abs(ref["grad_bl"][b] - num) < 1e-4 * (1 + abs(num))


def test_single_class_mixture_equals_plain_model():
    """tests/test_relax.cpp:234-298 analogue: K = 1 mixture == plain model;
    K = 3 with equal omegas == single omega."""
    tree, aln, sim = sim_problem(10, 200, 2)
    prob = ho.Problem.from_refs(tree, aln)
    N = len(tree.nodes)
    bl = branch_lengths(tree) * sim["conv"]
    p = ho.params(1.0, 0.7, sim["theta"])
    ll, g = ho.branch_length_gradients(prob, CODE, p, aln.pos_freqs, aln.cf3x4, bl)
    base = ho.params(1.0, 1.0, sim["theta"])
    m1 = ho.Mixture([0.7], np.zeros(N, dtype=np.int32), [[(0, 1.0)]], 0)
    r1 = ho.mixture_eval(prob, CODE, base, aln.pos_freqs, aln.cf3x4, m1, bl)
    assert abs(r1["log_likelihood"] - ll) < 1e-12 * abs(ll)
    np.testing.assert_allclose(r1["grad_bl"], g, rtol=1e-10, atol=1e-12)
    m3 = ho.Mixture([0.7, 0.7, 0.7], np.zeros(N, dtype=np.int32), [[(0, 0.2), (1, 0.5), (2, 0.3)]], 0)
    r3 = ho.mixture_eval(prob, CODE, base, aln.pos_freqs, aln.cf3x4, m3, bl)
    assert abs(r3["log_likelihood"] - ll) < 1e-11 * abs(ll)
    np.testing.assert_allclose(r3["grad_bl"], g, rtol=1e-9, atol=1e-11)


# ---------------------------------------------------------------------------
# a14: compensated summation
# ---------------------------------------------------------------------------
def neumaier(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return ho.lib().or_neumaier_sum(ho.ptr(x, ho._f64p), len(x))


def test_neumaier_exact_sums():
    assert neumaier([1e16, 1.0, -1e16]) == 1.0
    assert neumaier([1.0, 1e100, 1.0, -1e100]) == 2.0
    assert neumaier([0.1] * 10) == math.fsum([0.1] * 10)
    rng = np.random.default_rng(0)
    x = rng.normal(size=10000) * 10.0 ** rng.integers(-8, 8, size=10000)
    assert neumaier(x) == pytest.approx(math.fsum(x), rel=1e-15, abs=1e-300)
