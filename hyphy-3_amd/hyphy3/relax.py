"""RELAX branch-length gradients on the MI355X engine.

RELAXAnalyzer::compute_relax_branch_gradients (include/hyphy/analyses/
relax.hpp:803-1166, without Wheatstone bridges) evaluates lnL and
d lnL / d t_b under RELAX's three-class omega mixture: reference branches use
omega_R,c, test branches omega_T,c = clamp(omega_R,c^k, 1e-6, 1e4) with the
same class weights, unclassified branches their own BS-REL mixture
(omega_U, w_U).

* Branch-site level (the reference's default): every branch mixes the
  classes, P_b = sum_c w_c exp(conv Q_c t_b).  With the shared normalisation
  (default) conv = 3 / sum_c w_c mu_c per partition, which is the engine's
  branch-class mixture under the 3 / mu-bar time convention: one evaluation,
  branch classes = partitions.
* Site level: a site draws one class for every branch; lnL_p =
  log sum_c w_c L_p(c) (relax.hpp:1069-1103).  Each class is one evaluation
  (reference and test branches one matrix each, unclassified branches their
  mixture) under the raw convention with the host's conversion of branch
  lengths; the gradients are the class gradients with pattern weights
  w_p post_c(p) (hy_set_pattern_weights), summed over classes.

The per-class normalisation (conv_c = 3 / mu_c) is supported at the site
level when no branch is unclassified; mixing per-class conversions on one
branch (branch-site level, or unclassified branches) has no counterpart in
the engine's time conventions and raises ValueError.

Side effect: the engine is left with the last configuration this function
set (mixture and convention); callers set their own again.
"""
from __future__ import annotations

import math

import numpy as np

from .engine import HY_PREC_FP64, HY_TIME_CONV, HY_TIME_RAW

REFERENCE, TEST, UNCLASSIFIED = 0, 1, 2


def omegas_T(omegas_R, k):
    """Test-branch omegas (relax.hpp:823-831)."""
    out = []
    for w in omegas_R:
        if w <= 1e-12:
            out.append(0.0 if k > 0 else 1.0)
        else:
            out.append(min(max(w ** k, 1e-6), 10000.0))
    return out


def _conv(mus, weights, shared):
    """class_conversion_factors (relax.hpp:546-568)."""
    if shared:
        w_sum = sum(max(0.0, w) for w in weights)
        mu_mix = sum(max(0.0, w) * m for w, m in zip(weights, mus))
        if w_sum > 0.0:
            mu_mix /= w_sum
        c = 3.0 / mu_mix if mu_mix > 1e-12 else 1.0
        return [c, c, c]
    return [3.0 / m if m > 1e-12 else 1.0 for m in mus]


def branch_gradients(eng, partition, branch_lengths, omegas_R, weights, k, theta, omegas_U=None, weights_U=None,
                     mixture_level="branch_site", normalization="shared", precision=HY_PREC_FP64):
    """(lnL, grad[num_nodes]) with grad the derivative with respect to each
    non-root node's branch length (root entry 0).  partition[node] is
    REFERENCE, TEST or UNCLASSIFIED (the root's entry is ignored)."""
    part = np.asarray(partition, dtype=np.int32)
    bl = np.asarray(branch_lengths, dtype=np.float64)
    N = eng.num_nodes
    if part.shape != (N,) or bl.shape != (N,):
        raise ValueError(f"partition and branch lengths need {N} entries")
    if len(omegas_R) != 3 or len(weights) != 3:
        raise ValueError("RELAX has three omega classes")
    has_u = bool(np.any(part == UNCLASSIFIED))
    if has_u and (omegas_U is None or weights_U is None):
        raise ValueError("unclassified branches need omegas_U and weights_U")
    shared = normalization == "shared"
    site = mixture_level == "site"
    if not shared and (not site or has_u):
        raise ValueError("the per-class normalisation mixes per-class time conversions on one branch "
                         "(branch-site level or unclassified branches): not supported")
    w = [float(x) for x in weights]
    om_T = omegas_T(omegas_R, k)
    wU = [float(x) for x in weights_U] if has_u else [1.0, 0.0, 0.0]
    oU = [float(x) for x in omegas_U] if has_u else [1.0, 1.0, 1.0]
    if not site:
        if abs(sum(w) - 1.0) > 1e-12 or (has_u and abs(sum(wU) - 1.0) > 1e-12):
            raise ValueError("the branch-site level needs class weights summing to one")
        # classes = partitions; components with weight <= 1e-12 are left out
        # as in relax.hpp:884; matrices R0..R2, T0..T2, U0..U2
        mats = [float(x) for x in omegas_R] + om_T + oU
        classes, comp_w = [], []
        for base, ws in ((0, w), (3, w), (6, wU)):
            comps = [base + c for c in range(3) if ws[c] > 1e-12]
            classes.append(comps)
            comp_w.extend(ws[c] for c in range(3) if ws[c] > 1e-12)
        eng.set_mixture(len(mats), part, classes, convention=HY_TIME_CONV)
        r = eng.eval(bl, mats, comp_w, theta, precision=precision)
        g = np.array(r["grad_bl"], dtype=np.float64)
        return r["log_likelihood"], g

    # site level: mean rates of every matrix for the conversions
    def mu(om):
        return eng.mg94_spectrum(1.0, om, theta)[3]

    muR = [mu(x) for x in omegas_R]
    muT = [mu(x) for x in om_T]
    convR, convT = _conv(muR, w, shared), _conv(muT, w, shared)
    convU = _conv([mu(x) for x in oU], wU, True) if has_u else [1.0, 1.0, 1.0]
    P = eng.num_patterns
    pw = np.asarray(eng.weights, dtype=np.float64)
    live = [c for c in range(3) if w[c] > 1e-12]
    u_comps = [c for c in range(3) if wU[c] > 1e-12]
    site_ll = np.full((3, P), -1e20)
    taus = {}
    for c in live:
        mats = [float(omegas_R[c]), om_T[c]] + [oU[j] for j in u_comps]
        classes = [[0], [1], [2 + j for j in range(len(u_comps))]]
        comp_w = [1.0, 1.0] + [wU[j] for j in u_comps]
        conv_b = np.where(part == TEST, convT[c], np.where(part == UNCLASSIFIED, convU[0], convR[c]))
        tau = bl * conv_b
        taus[c] = (mats, classes, comp_w, conv_b, tau)
        eng.set_mixture(len(mats), part, classes, convention=HY_TIME_RAW)
        r = eng.eval(tau, mats, comp_w, theta, precision=precision, grad=False, site=True)
        site_ll[c] = r["site_ll"]
    # per-pattern log-sum-exp over classes with a positive likelihood
    # (relax.hpp:1079-1103); patterns without one contribute 0
    post = np.zeros((3, P))
    pattern_ll = np.zeros(P)
    for p in range(P):
        terms = [(c, math.log(w[c]) + site_ll[c, p]) for c in live if site_ll[c, p] > -1e19]
        if not terms:
            continue
        mx = max(t for _, t in terms)
        se = 0.0
        for c, t in terms:
            post[c, p] = math.exp(t - mx)
            se += post[c, p]
        if se > 0.0:
            pattern_ll[p] = pw[p] * (mx + math.log(se))
            post[:, p] /= se
    s, cmp = 0.0, 0.0
    for x in pattern_ll:  # Neumaier in pattern order (relax.hpp:1146)
        t = s + x
        cmp += (s - t) + x if abs(s) >= abs(x) else (x - t) + s
        s = t
    lnl = s + cmp
    grad = np.zeros(N)
    try:
        for c in live:
            mats, classes, comp_w, conv_b, tau = taus[c]
            eng.set_mixture(len(mats), part, classes, convention=HY_TIME_RAW)
            eng.set_pattern_weights(np.where(post[c] > 1e-12, pw * post[c], 0.0))
            r = eng.eval(tau, mats, comp_w, theta, precision=precision)
            grad += conv_b * np.asarray(r["grad_bl"], dtype=np.float64)
    finally:
        eng.set_pattern_weights(None)
    grad[eng.root] = 0.0
    return lnl, grad
