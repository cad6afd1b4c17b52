"""Seeded synthetic codon data: Yule trees and branch-site omega-mixture evolution.

The reference ships no simulator; this one follows the recipe used for every
benchmark configuration:

* Yule (pure-birth) topology on ``n`` taxa, branch lengths iid Exp(mean 0.1)
  in expected substitutions, model time ``tau = bl * 3 / mubar`` with
  ``mubar = sum_k w_k mu_k`` of the generating mixture;
* GTR exchangeabilities theta_{AC,AT,CG,CT,GT} ~ U(0.5, 2), theta_AG = 1;
* three codon-position nucleotide frequency vectors ~ Dirichlet(5,5,5,5),
  root codon frequencies F3x4 of those (universal code, stops excluded);
* branch-site omega mixture: every (site, branch) pair draws its category
  with probability w_k and evolves under P_k(tau_b);
* optional gap injection: a fraction of tip codons replaced by ``---``.

Returns FASTA and Newick text so that the alignment/tree parsers and pattern
compression run on it exactly as on a user's files.
"""
from __future__ import annotations

import numpy as np

ACGT = "ACGT"
UNIVERSAL_ACGT = "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF"
SENSE = [c for c in range(64) if UNIVERSAL_ACGT[c] != "*"]
SENSE_TRIPLETS = ["".join(ACGT[(c >> s) & 3] for s in (4, 2, 0)) for c in SENSE]


def yule_newick(n: int, rng: np.random.Generator, mean_bl: float = 0.1, prefix: str = "T") -> str:
    """Rooted binary Yule topology with iid Exp(mean_bl) branch lengths."""
    # node: (children list); grow by splitting a uniformly chosen leaf
    children = {0: []}
    leaves = [0]
    nxt = 1
    while len(leaves) < n:
        k = int(rng.integers(len(leaves)))
        leaf = leaves.pop(k)
        a, b = nxt, nxt + 1
        nxt += 2
        children[leaf] = [a, b]
        children[a] = []
        children[b] = []
        leaves.extend([a, b])
    names = {leaf: f"{prefix}{i + 1}" for i, leaf in enumerate(sorted(leaves))}
    bl = {node: float(rng.exponential(mean_bl)) for node in children}

    def rec(node):
        ch = children[node]
        if not ch:
            return f"{names[node]}:{bl[node]:.10f}"
        inner = ",".join(rec(c) for c in ch)
        return f"({inner}):{bl[node]:.10f}" if node != 0 else f"({inner});"

    return rec(0)


def mg94_Q(theta, pos, omega, alpha=1.0):
    """Unnormalised MG94xREV generator on the 61 universal sense codons."""
    th = {(0, 1): theta[0], (0, 3): theta[1], (1, 2): theta[2], (1, 3): theta[3], (2, 3): theta[4], (0, 2): 1.0}
    S = len(SENSE)
    Q = np.zeros((S, S))
    for i, ci in enumerate(SENSE):
        ni = [(ci >> 4) & 3, (ci >> 2) & 3, ci & 3]
        for j, cj in enumerate(SENSE):
            if i == j:
                continue
            nj = [(cj >> 4) & 3, (cj >> 2) & 3, cj & 3]
            diff = [p for p in range(3) if ni[p] != nj[p]]
            if len(diff) != 1:
                continue
            p = diff[0]
            a, b = sorted((ni[p], nj[p]))
            syn = UNIVERSAL_ACGT[ci] == UNIVERSAL_ACGT[cj]
            Q[i, j] = (alpha if syn else omega) * th[(a, b)] * pos[p, nj[p]]
        Q[i, i] = -Q[i].sum()
    return Q


def f3x4(pos):
    pi = np.array([pos[0, (c >> 4) & 3] * pos[1, (c >> 2) & 3] * pos[2, c & 3] for c in SENSE])
    return pi / pi.sum()


def expm_sym(Q, pi, t):
    """exp(Q t) for a pi-reversible Q via the symmetric eigendecomposition."""
    d = np.sqrt(pi)
    Sm = (d[:, None] * Q) / d[None, :]
    Sm = 0.5 * (Sm + Sm.T)
    lam, Z = np.linalg.eigh(Sm)
    lam = np.minimum(lam, 0.0)
    P = (Z / d[:, None]) @ (np.exp(lam * t)[:, None] * (Z.T * d[None, :]))
    return np.maximum(P, 0.0)


def simulate(n_taxa: int, n_codons: int, seed: int, omegas=(0.5,), weights=(1.0,), gap_fraction: float = 0.0,
             mean_bl: float = 0.1):
    """Return dict(newick, fasta, theta, pos, omegas, weights, names)."""
    rng = np.random.default_rng(seed)
    newick = yule_newick(n_taxa, rng, mean_bl)
    theta = rng.uniform(0.5, 2.0, size=5)
    pos = np.vstack([rng.dirichlet([5.0] * 4) for _ in range(3)])
    pi = f3x4(pos)
    omegas = np.asarray(omegas, dtype=float)
    weights = np.asarray(weights, dtype=float)
    Qs = [mg94_Q(theta, pos, w) for w in omegas]
    mus = np.array([-(pi * np.diag(Q)).sum() for Q in Qs])
    mubar = float((weights * mus).sum())
    conv = 3.0 / mubar
    # parse the rooted newick into (children, bl) by a tiny recursive reader
    tree = _parse_simple(newick)
    S = len(SENSE)
    root_states = rng.choice(S, size=n_codons, p=pi)
    states = {tree["root"]: root_states}
    seqs = {}
    order = [tree["root"]]
    k = 0
    while k < len(order):
        node = order[k]
        k += 1
        for c in tree["children"][node]:
            tau = tree["bl"][c] * conv
            Ps = [np.cumsum(expm_sym(Q, pi, tau), axis=1) for Q in Qs]
            cat = rng.choice(len(omegas), size=n_codons, p=weights)
            parent_states = states[node]
            u = rng.random(n_codons)
            child = np.empty(n_codons, dtype=np.int64)
            for kk, Pc in enumerate(Ps):
                sel = np.where(cat == kk)[0]
                if sel.size == 0:
                    continue
                rows = Pc[parent_states[sel]]
                rows = rows / rows[:, -1:]
                child[sel] = np.minimum((rows < u[sel, None]).sum(axis=1), S - 1)
            states[c] = child
            order.append(c)
            if not tree["children"][c]:
                seqs[tree["name"][c]] = child
    names = sorted(seqs, key=lambda s: int(s[1:]) if s[1:].isdigit() else s)
    lines = []
    for nm in names:
        cod = [SENSE_TRIPLETS[s] for s in seqs[nm]]
        if gap_fraction > 0:
            mask = rng.random(n_codons) < gap_fraction
            cod = ["---" if m else c for c, m in zip(cod, mask)]
        lines.append(f">{nm}\n{''.join(cod)}")
    return {"newick": newick, "fasta": "\n".join(lines) + "\n", "theta": theta, "pos": pos, "omegas": omegas,
            "weights": weights, "names": names, "conv": conv, "mubar": mubar}


def _parse_simple(newick):
    """Reader for the simulator's own Newick output (names + lengths only)."""
    children, bl, name = {}, {}, {}
    stack = []
    i = 0
    nid = 0
    s = newick.strip().rstrip(";")
    cur = None
    root = None
    while i < len(s):
        c = s[i]
        if c == "(":
            node = nid
            nid += 1
            children[node] = []
            bl[node] = 0.0
            name[node] = ""
            if stack:
                children[stack[-1]].append(node)
            else:
                root = node
            stack.append(node)
            i += 1
        elif c in ",)":
            if c == ")":
                cur = stack.pop()
            i += 1
            if c == ")":
                j = i
                while j < len(s) and s[j] not in ",()":
                    j += 1
                tok = s[i:j]
                if ":" in tok:
                    bl[cur] = float(tok.split(":")[1])
                i = j
        else:
            j = i
            while j < len(s) and s[j] not in ",()":
                j += 1
            tok = s[i:j]
            node = nid
            nid += 1
            nm, ln = tok.split(":")
            children[node] = []
            bl[node] = float(ln)
            name[node] = nm
            children[stack[-1]].append(node)
            i = j
    return {"children": children, "bl": bl, "name": name, "root": root}
