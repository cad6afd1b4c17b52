"""FP64 restatement of RELAX's branch-length gradients.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
never imports this module.

Restated reference semantics (paths relative to the reference root), without
Wheatstone bridges:
  * RELAXAnalyzer::class_conversion_factors      include/hyphy/analyses/relax.hpp:546-568
  * RELAXAnalyzer::compute_relax_branch_gradients relax.hpp:803-1166
      test omegas omega_T = clamp(omega_R^k, 1e-6, 1e4)    :823-831
      matrices per class and partition, conversions        :838-866
      P and dP/dt per (class, branch), branch-site mixing  :872-922
      tiles of min(64, patterns) patterns                  :1023-1026
      branch-site level: one workspace, w_p / L V.(dP D)   :1039-1068
      site level: workspaces per class, log-sum-exp,
        posterior-weighted gradients                        :1069-1143
      Neumaier sums in pattern order                        :1146-1156
The inside-outside workspaces are the oracle's restatement of
compute_inside_outside_batch_in_place (hyoracle.workspace).
"""
from __future__ import annotations

import numpy as np

from oracle import hyoracle as ho

REFERENCE, TEST, UNCLASSIFIED = 0, 1, 2


def omegas_T(omegas_R, k):
    """omega_T,c = (omega_R,c)^k clamped to [1e-6, 1e4] (relax.hpp:823-831)."""
    out = []
    for w in omegas_R:
        if w <= 1e-12:
            out.append(0.0 if k > 0 else 1.0)
        else:
            out.append(min(max(w ** k, 1e-6), 10000.0))
    return out


def conversion_factors(mus, weights, shared):
    """relax.hpp:546-568: 3 / mixture mean rate (shared) or 3 / mu_c."""
    if shared:
        w_sum = 0.0
        mu_mix = 0.0
        for c in range(3):
            w = max(0.0, weights[c])
            w_sum += w
            mu_mix += w * mus[c]
        if w_sum > 0.0:
            mu_mix /= w_sum
        conv = 3.0 / mu_mix if mu_mix > 1e-12 else 1.0
        return [conv] * 3
    return [3.0 / m if m > 1e-12 else 1.0 for m in mus]


def _neumaier(values):
    s, c = 0.0, 0.0
    for x in values:
        s, c = ho._neumaier_acc(s, c, float(x))
    return s + c


def relax_branch_gradients(prob, code, theta, pos, pi, partition, bl, omegas_R, weights, k, omegas_U=None,
                           weights_U=None, branch_site=True, shared=True):
    """lnL and d lnL / d t_b of every non-root node under the three-class
    RELAX mixture (relax.hpp:803-1166); partition[node] is 0 (reference),
    1 (test) or 2 (unclassified)."""
    S = prob.t.S
    N, npat = prob.num_nodes, prob.num_patterns
    root = prob.root
    om_T = omegas_T(omegas_R, k)
    has_u = omegas_U is not None
    om_U = list(omegas_U) if has_u else [1.0, 1.0, 1.0]
    wt_U = list(weights_U) if has_u else [1.0, 0.0, 0.0]

    def mats(oms):
        out = []
        for om in oms:
            Q, mu = ho.mg94_Q(code, ho.params(1.0, om, theta), pos, pi)
            ev, U, V = ho.decompose(Q, pi)
            out.append((Q, mu, ev, U, V))
        return out

    mR, mT = mats(omegas_R), mats(om_T)
    mU = mats(om_U) if has_u else None
    convR = conversion_factors([m[1] for m in mR], weights, shared)
    convT = conversion_factors([m[1] for m in mT], weights, shared)
    convU = conversion_factors([m[1] for m in mU], wt_U, shared) if has_u else [1.0, 1.0, 1.0]

    # P and dP per (class, branch)
    Pc = np.zeros((3, N, S, S))
    dPc = np.zeros((3, N, S, S))
    for c in range(3):
        Pc[c, root] = np.eye(S)
    for nid in range(N):
        if nid == root:
            continue
        part = partition[nid]
        br_w = wt_U if part == 2 else weights
        sel = {0: (mR, convR), 1: (mT, convT), 2: (mU, convU)}[part]
        if branch_site or part == 2:
            Pm = np.zeros((S, S))
            dPm = np.zeros((S, S))
            for c in range(3):
                if br_w[c] <= 1e-12:
                    continue
                Q, mu, ev, U, V = sel[0][c]
                conv = sel[1][c]
                P = ho.transition(ev, U, V, bl[nid] * conv)
                Pm += br_w[c] * P
                dPm += br_w[c] * (conv * (Q @ P))
            for c in range(3):
                Pc[c, nid] = Pm
                dPc[c, nid] = dPm
        else:
            for c in range(3):
                if weights[c] <= 1e-12:
                    continue
                Q, mu, ev, U, V = sel[0][c]
                conv = sel[1][c]
                P = ho.transition(ev, U, V, bl[nid] * conv)
                Pc[c, nid] = P
                dPc[c, nid] = conv * (Q @ P)

    class_w = [1.0, 0.0, 0.0] if branch_site else list(weights)
    pattern_ll = np.zeros(npat)
    pattern_grad = np.zeros((N, npat))
    tile = min(64, max(1, npat))
    for p0 in range(0, npat, tile):
        p1 = min(npat, p0 + tile)
        B = p1 - p0
        w = prob.weights[p0:p1]
        ws = [ho.workspace(prob, Pc[c], pi, p0, p1, outside=True) if class_w[c] > 1e-12 else None for c in range(3)]
        post = np.zeros((3, B))
        if branch_site:
            L = ws[0]["L"]
            for b in range(B):
                if L[b] > 0.0:
                    pattern_ll[p0 + b] = w[b] * ws[0]["logL"][b]
                    post[0, b] = 1.0
        else:
            for b in range(B):
                terms = [(c, np.log(class_w[c]) + ws[c]["logL"][b]) for c in range(3)
                         if class_w[c] > 1e-12 and ws[c]["L"][b] > 0.0]
                if not terms:
                    continue
                mx = max(t for _, t in terms)
                if mx <= -1e19:
                    continue
                se = 0.0
                for c, t in terms:
                    post[c, b] = np.exp(t - mx)
                    se += post[c, b]
                if se > 0.0:
                    pattern_ll[p0 + b] = w[b] * (mx + np.log(se))
                    post[:, b] /= se
        for vid in range(N):
            if vid == root:
                continue
            for c in range(3):
                if class_w[c] <= 1e-12:
                    continue
                dPD = dPc[c, vid] @ ws[c]["D"][vid]
                V = ws[c]["V"][vid]
                for b in range(B):
                    pw = post[c, b]
                    Lb = ws[c]["L"][b]
                    if branch_site:
                        if Lb > 0.0:
                            pattern_grad[vid, p0 + b] = (w[b] / Lb) * float(V[:, b] @ dPD[:, b])
                    elif pw > 1e-12 and Lb > 0.0:
                        pattern_grad[vid, p0 + b] += w[b] * pw * (float(V[:, b] @ dPD[:, b]) / Lb)
    lnl = _neumaier(pattern_ll)
    grad = np.zeros(N)
    for i in range(N):
        if i != root:
            grad[i] = _neumaier(pattern_grad[i])
    return lnl, grad
