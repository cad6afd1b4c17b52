"""FP64 restatement of pruning through Wheatstone bridges.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
never imports this module.

A bridge replaces the binary resolution of a hub u with three children
{A, B, C} by a weighted superposition of its three quartet resolutions:
  D_u = sum_r w_r (P_br(r) (M_x * M_y)) * M_z,   M_c = P_c D_c,
with (x, y | z) = (A, B | C), (A, C | B), (B, C | A).

Restated reference semantics (paths relative to the reference root):
  * init_leaf_likelihood                      include/hyphy/core/likelihood.hpp:53-87
  * bridged hub and ordinary nodes, 1e150
    boosts, per-pattern lnL, Neumaier sum      include/hyphy/core/wheatstone_likelihood.hpp:21-148
    (the batched engine form, likelihood.hpp:376-425, is the same algebra)
"""
from __future__ import annotations

import numpy as np

from oracle import hyoracle as ho

SCALER_THRESH = 1e-150
SCALER_UP = 1e150
LOG_SCALER = float(np.log(SCALER_UP))


def leaf_vector(prob, taxon, p):
    """init_leaf_likelihood (likelihood.hpp:53-87) for one pattern."""
    S = prob.t.S
    if taxon < 0:
        return np.ones(S)
    st = int(prob.states[p, taxon])
    mask = int(prob.masks[p, taxon])
    if 0 <= st < S:
        v = np.zeros(S)
        v[st] = 1.0
        return v
    if mask != 0:
        allm = (1 << S) - 1
        if mask & allm == allm:
            return np.ones(S)
        return np.array([1.0 if (mask >> s) & 1 else 0.0 for s in range(S)])
    return np.ones(S)


def children(prob, u):
    return [int(c) for c in prob.child_idx[prob.child_ptr[u]:prob.child_ptr[u + 1]]]


def network_loglik(prob, P, P_bridges, bridges, pi, site=False):
    """lnL of the bridged tree: P[node] (S x S) per branch, P_bridges[k][r]
    per bridge and resolution, bridges[k] = dict(hub, children (3 ids),
    weights (3)); wheatstone_likelihood.hpp:66-140 per pattern."""
    S, npat = prob.t.S, prob.num_patterns
    hub_of = {b["hub"]: k for k, b in enumerate(bridges)}
    pi = np.asarray(pi, dtype=np.float64)
    plog = np.zeros(npat)
    for p in range(npat):
        L = {}
        scale = {}
        for nid in prob.post_order:
            nid = int(nid)
            if prob.is_leaf[nid]:
                L[nid] = leaf_vector(prob, int(prob.leaf_taxon[nid]), p)
                scale[nid] = 0.0
                continue
            boost = 0
            if nid in hub_of:
                wb = bridges[hub_of[nid]]
                a, b, c = wb["children"]
                MA, MB, MC = P[a] @ L[a], P[b] @ L[b], P[c] @ L[c]
                child_scale = scale[a] + scale[b] + scale[c]
                Pb = P_bridges[hub_of[nid]]
                D = np.zeros(S)
                w = wb["weights"]
                if w[0] > 0.0:
                    D += w[0] * ((Pb[0] @ (MA * MB)) * MC)
                if w[1] > 0.0:
                    D += w[1] * ((Pb[1] @ (MA * MC)) * MB)
                if w[2] > 0.0:
                    D += w[2] * ((Pb[2] @ (MB * MC)) * MA)
            else:
                D = np.ones(S)
                child_scale = 0.0
                for cid in children(prob, nid):
                    D = D * (P[cid] @ L[cid])
                    child_scale += scale[cid]
                    mx = D.max()
                    while SCALER_THRESH > mx > 0.0:
                        D = D * SCALER_UP
                        mx *= SCALER_UP
                        boost += 1
            mx = D.max()
            while SCALER_THRESH > mx > 0.0:
                D = D * SCALER_UP
                mx *= SCALER_UP
                boost += 1
            L[nid] = D
            scale[nid] = child_scale - boost * LOG_SCALER
        lik = float(pi @ L[prob.root])
        plog[p] = np.log(lik) + scale[prob.root] if lik > 0.0 else -1e20
    s, cmp = 0.0, 0.0
    for p in range(npat):
        s, cmp = ho._neumaier_acc(s, cmp, prob.weights[p] * plog[p])
    return (s + cmp, plog) if site else s + cmp


def resolved_tree(parent, kids, bridges, resolution):
    """Parent / children arrays of the tree in which bridge k is resolved as
    resolution[k] (0: (A,B)|C, 1: (A,C)|B, 2: (B,C)|A): a new node v_k =
    num_nodes + k takes the pair, its branch to the hub carries the bridge's
    matrix.  Returns (parent, children)."""
    n = len(parent)
    par = list(parent) + [0] * len(bridges)
    ch = [list(c) for c in kids] + [[] for _ in bridges]
    for k, wb in enumerate(bridges):
        a, b, c = wb["children"]
        x, y, z = ((a, b, c), (a, c, b), (b, c, a))[resolution[k]]
        v = n + k
        u = wb["hub"]
        ch[u] = [v, z]
        ch[v] = [x, y]
        par[x] = v
        par[y] = v
        par[v] = u
    return par, ch
