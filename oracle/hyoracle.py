"""ctypes front end for the FP64 CPU oracle (libhyoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()``
and bench.py's CPU legs (``cpu_baseline``, the reference-fit check of
``--fit --cpu-fit`` and the CPU column of ``--estep``), as the checker or the
CPU baseline.  The product never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhyoracle.so")

_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_i8p = C.POINTER(C.c_int8)
_u64p = C.POINTER(C.c_uint64)
_f64p = C.POINTER(C.c_double)


class OrCode(C.Structure):
    _fields_ = [("S", C.c_int32), ("code_id", C.c_int32), ("c64_to_sense", C.c_int8 * 64),
                ("sense_to_c64", C.c_int8 * 64), ("aa", C.c_char * 64), ("nd", C.c_int8 * 4096),
                ("dpos", C.c_int8 * 12288), ("dfrom", C.c_int8 * 12288), ("dto", C.c_int8 * 12288),
                ("syn", C.c_uint8 * 4096)]


class OrTree(C.Structure):
    _fields_ = [("num_nodes", C.c_int32), ("root", C.c_int32), ("num_taxa", C.c_int32),
                ("num_patterns", C.c_int32), ("S", C.c_int32), ("pad0", C.c_int32),
                ("parent", _i32p), ("child_ptr", _i32p), ("child_idx", _i32p), ("post_order", _i32p),
                ("leaf_taxon", _i32p), ("is_leaf", _u8p), ("states", _i8p), ("masks", _u64p),
                ("weights", _f64p)]


class OrParams(C.Structure):
    _fields_ = [("alpha", C.c_double), ("beta", C.c_double), ("theta", C.c_double * 5),
                ("delta", C.c_double), ("psi", C.c_double)]


class OrMixture(C.Structure):
    _fields_ = [("n_mat", C.c_int32), ("n_class", C.c_int32), ("convention", C.c_int32), ("pad0", C.c_int32),
                ("omegas", _f64p), ("branch_class", _i32p), ("comp_ptr", _i32p), ("comp_mat", _i32p),
                ("comp_w", _f64p), ("sink_freqs", _f64p)]


class OrMixtureOut(C.Structure):
    _fields_ = [("log_likelihood", C.c_double), ("site_ll", _f64p), ("grad_bl", _f64p), ("grad_omega", _f64p),
                ("grad_w", _f64p), ("grad_theta", _f64p), ("dtau", _f64p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.or_code_build.argtypes = [C.c_int32, C.POINTER(OrCode)]
        L.or_mg94_Q.argtypes = [C.POINTER(OrCode), C.POINTER(OrParams), _f64p, _f64p, _f64p, _f64p]
        L.or_mg94_decompose.argtypes = [C.c_int32, _f64p, _f64p, _f64p, _f64p, _f64p]
        L.or_transition.argtypes = [C.c_int32, _f64p, _f64p, _f64p, C.c_double, _f64p]
        L.or_expm_taylor.argtypes = [C.c_int32, _f64p, C.c_double, _f64p]
        L.or_sym_eigen.argtypes = [C.c_int32, _f64p, _f64p, _f64p]
        L.or_loglik_from_P.argtypes = [C.POINTER(OrTree), _f64p, _f64p, C.c_int32, C.c_int32, _f64p]
        L.or_loglik_from_P.restype = C.c_double
        L.or_mg94_loglik.argtypes = [C.POINTER(OrTree), C.POINTER(OrCode), C.POINTER(OrParams), _f64p, _f64p,
                                     _f64p, C.c_double, C.c_int32, _f64p]
        L.or_mg94_loglik.restype = C.c_double
        L.or_branch_length_gradients.argtypes = [C.POINTER(OrTree), C.POINTER(OrCode), C.POINTER(OrParams),
                                                 _f64p, _f64p, _f64p, C.c_int32, _f64p]
        L.or_branch_length_gradients.restype = C.c_double
        L.or_joint_frechet.argtypes = [C.POINTER(OrTree), C.POINTER(OrCode), C.POINTER(OrParams), _f64p, _f64p,
                                       _f64p, _u8p, C.c_double, C.c_int32, _f64p, _f64p, _f64p, _f64p]
        L.or_joint_frechet.restype = C.c_double
        L.or_mixture_eval.argtypes = [C.POINTER(OrTree), C.POINTER(OrCode), C.POINTER(OrParams), _f64p, _f64p,
                                      C.POINTER(OrMixture), _f64p, C.c_int32, C.c_int32, C.c_int32,
                                      C.POINTER(OrMixtureOut)]
        L.or_mixture_eval.restype = C.c_int32
        L.or_srv_eval.argtypes = [C.POINTER(OrTree), C.POINTER(OrCode), C.POINTER(OrParams), _f64p, _f64p,
                                  C.POINTER(OrMixture), _f64p, C.c_int32, _f64p, _f64p, C.c_int32, C.c_int32,
                                  C.POINTER(OrMixtureOut), _f64p, _f64p]
        L.or_srv_eval.restype = C.c_int32
        L.or_workspace.argtypes = [C.POINTER(OrTree), _f64p, _f64p, C.c_int32, C.c_int32, C.c_int32, _f64p, _f64p,
                                   _f64p, _f64p, _i32p, _f64p, _f64p, _f64p]
        L.or_workspace.restype = C.c_int32
        L.or_neumaier_sum.argtypes = [_f64p, C.c_int64]
        L.or_neumaier_sum.restype = C.c_double
        _lib = L
    return _lib


def ptr(a, t):
    return a.ctypes.data_as(t)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Code:
    def __init__(self, code_id: int = 0):
        self.c = OrCode()
        if lib().or_code_build(code_id, C.byref(self.c)) != 0:
            raise ValueError(f"unknown genetic code id {code_id}")

    @property
    def S(self):
        return self.c.S


def params(alpha=1.0, beta=1.0, theta=(1, 1, 1, 1, 1), delta=0.0, psi=0.0):
    p = OrParams()
    p.alpha = alpha
    p.beta = beta
    for i in range(5):
        p.theta[i] = float(theta[i])
    p.delta = delta
    p.psi = psi
    return p


class Problem:
    """Flat tree + compressed-pattern arrays kept alive for the C oracle.

    ``parent``/``children``/``post_order``/``is_leaf`` describe the tree by
    node id; ``leaf_taxon`` maps node -> alignment row (-1 unmatched);
    ``states``/``masks`` are [num_patterns, num_taxa].
    """

    def __init__(self, parent, children, post_order, is_leaf, leaf_taxon, states, masks, weights, S=61):
        n = len(parent)
        self.num_nodes = n
        self.parent = np.ascontiguousarray(parent, dtype=np.int32)
        cp = [0]
        ci = []
        for k in range(n):
            ci.extend(children[k])
            cp.append(len(ci))
        self.child_ptr = np.array(cp, dtype=np.int32)
        self.child_idx = np.array(ci if ci else [0], dtype=np.int32)
        self.post_order = np.ascontiguousarray(post_order, dtype=np.int32)
        self.is_leaf = np.ascontiguousarray(is_leaf, dtype=np.uint8)
        self.leaf_taxon = np.ascontiguousarray(leaf_taxon, dtype=np.int32)
        self.states = np.ascontiguousarray(states, dtype=np.int8)
        self.masks = np.ascontiguousarray(masks, dtype=np.uint64)
        self.weights = f64(weights)
        self.root = int(np.where(self.parent < 0)[0][0])
        self.S = S
        t = OrTree()
        t.num_nodes = n
        t.root = self.root
        t.num_patterns = int(self.states.shape[0])
        t.num_taxa = int(self.states.shape[1]) if self.states.ndim == 2 else 0
        t.S = S
        t.parent = ptr(self.parent, _i32p)
        t.child_ptr = ptr(self.child_ptr, _i32p)
        t.child_idx = ptr(self.child_idx, _i32p)
        t.post_order = ptr(self.post_order, _i32p)
        t.leaf_taxon = ptr(self.leaf_taxon, _i32p)
        t.is_leaf = ptr(self.is_leaf, _u8p)
        t.states = ptr(self.states, _i8p)
        t.masks = ptr(self.masks, _u64p)
        t.weights = ptr(self.weights, _f64p)
        self.t = t
        self.num_patterns = t.num_patterns

    @staticmethod
    def from_refs(tree, aln):
        """Build from oracle.data_ref TreeRef + AlignmentRef."""
        parent = [n.parent for n in tree.nodes]
        children = [list(n.children) for n in tree.nodes]
        is_leaf = [1 if n.is_leaf else 0 for n in tree.nodes]
        l2t = tree.leaf_to_taxon(aln.names)
        return Problem(parent, children, tree.post_order, is_leaf, l2t, aln.states, aln.pmasks, aln.weights,
                       S=aln.code.S)


def mg94_Q(code, p, pos, pi):
    S = code.S
    Q = np.zeros((S, S))
    mu = C.c_double()
    pos = f64(pos).reshape(12)
    pi = f64(pi)
    lib().or_mg94_Q(C.byref(code.c), C.byref(p), ptr(pos, _f64p), ptr(pi, _f64p), ptr(Q, _f64p), C.byref(mu))
    return Q, mu.value


def decompose(Q, pi):
    S = Q.shape[0]
    ev = np.zeros(S)
    U = np.zeros((S, S))
    V = np.zeros((S, S))
    Qc = f64(Q)
    pi = f64(pi)
    rc = lib().or_mg94_decompose(S, ptr(Qc, _f64p), ptr(pi, _f64p), ptr(ev, _f64p), ptr(U, _f64p), ptr(V, _f64p))
    if rc != 0:
        raise RuntimeError("eigensolver did not converge")
    return ev, U, V


def transition(ev, U, V, t):
    S = len(ev)
    P = np.zeros((S, S))
    lib().or_transition(S, ptr(f64(ev), _f64p), ptr(f64(U), _f64p), ptr(f64(V), _f64p), float(t), ptr(P, _f64p))
    return P


def expm_taylor(A, t):
    n = A.shape[0]
    out = np.zeros((n, n))
    Ac = f64(A)
    lib().or_expm_taylor(n, ptr(Ac, _f64p), float(t), ptr(out, _f64p))
    return out


def sym_eigen(A):
    n = A.shape[0]
    Ac = f64(A).copy()
    ev = np.zeros(n)
    Z = np.zeros((n, n))
    rc = lib().or_sym_eigen(n, ptr(Ac, _f64p), ptr(ev, _f64p), ptr(Z, _f64p))
    if rc != 0:
        raise RuntimeError("eigensolver did not converge")
    return ev, Z


def loglik_from_P(prob, P, pi, tile=0, threads=1, site=False):
    Pc = f64(P)
    pi = f64(pi)
    s = np.zeros(prob.num_patterns) if site else None
    ll = lib().or_loglik_from_P(C.byref(prob.t), ptr(Pc, _f64p), ptr(pi, _f64p), tile, threads,
                                ptr(s, _f64p) if site else None)
    return (ll, s) if site else ll


def mg94_loglik(prob, code, p, pos, pi, bl, branch_scale=1.0, threads=1, site=False):
    pos = f64(pos).reshape(12)
    pi = f64(pi)
    blc = f64(bl)
    s = np.zeros(prob.num_patterns) if site else None
    ll = lib().or_mg94_loglik(C.byref(prob.t), C.byref(code.c), C.byref(p), ptr(pos, _f64p), ptr(pi, _f64p),
                              ptr(blc, _f64p), branch_scale, threads, ptr(s, _f64p) if site else None)
    return (ll, s) if site else ll


def branch_length_gradients(prob, code, p, pos, pi, bl, threads=1):
    pos = f64(pos).reshape(12)
    pi = f64(pi)
    blc = f64(bl)
    g = np.zeros(prob.num_nodes)
    ll = lib().or_branch_length_gradients(C.byref(prob.t), C.byref(code.c), C.byref(p), ptr(pos, _f64p),
                                          ptr(pi, _f64p), ptr(blc, _f64p), threads, ptr(g, _f64p))
    return ll, g


def joint_frechet(prob, code, p, pos, pi, bl, is_test=None, omega_bg=-1.0, threads=1):
    pos = f64(pos).reshape(12)
    pi = f64(pi)
    blc = f64(bl)
    g = np.zeros(prob.num_nodes)
    go = C.c_double()
    gob = C.c_double()
    gt = np.zeros(5)
    it = None if is_test is None else np.ascontiguousarray(is_test, dtype=np.uint8)
    ll = lib().or_joint_frechet(C.byref(prob.t), C.byref(code.c), C.byref(p), ptr(pos, _f64p), ptr(pi, _f64p),
                                ptr(blc, _f64p), ptr(it, _u8p) if it is not None else None, omega_bg, threads,
                                ptr(g, _f64p), C.byref(go), C.byref(gob), ptr(gt, _f64p))
    return {"log_likelihood": ll, "grad_branches": g, "grad_omega": go.value, "grad_omega_bg": gob.value,
            "grad_theta": gt}


class Mixture:
    """Branch-site mixture specification mirrored by the GPU library."""

    def __init__(self, omegas, branch_class, classes, convention=0, sink_freqs=None):
        # classes: list of lists of (matrix index, weight); matrix index -1 is
        # BUSTED's error sink built from sink_freqs (busted.hpp:193-206)
        self.omegas = f64(omegas)
        self.branch_class = np.ascontiguousarray(branch_class, dtype=np.int32)
        cp = [0]
        cm, cw = [], []
        for comps in classes:
            for m, w in comps:
                cm.append(m)
                cw.append(w)
            cp.append(len(cm))
        self.comp_ptr = np.array(cp, dtype=np.int32)
        self.comp_mat = np.array(cm, dtype=np.int32)
        self.comp_w = f64(cw)
        self.convention = convention
        m = OrMixture()
        m.n_mat = len(self.omegas)
        m.n_class = len(classes)
        m.convention = convention
        m.omegas = ptr(self.omegas, _f64p)
        m.branch_class = ptr(self.branch_class, _i32p)
        m.comp_ptr = ptr(self.comp_ptr, _i32p)
        m.comp_mat = ptr(self.comp_mat, _i32p)
        m.comp_w = ptr(self.comp_w, _f64p)
        self.sink_freqs = None if sink_freqs is None else f64(sink_freqs)
        m.sink_freqs = None if sink_freqs is None else ptr(self.sink_freqs, _f64p)
        self.m = m
        self.n_comp = len(cm)


def mixture_eval(prob, code, base, pos, pi, mix, bl, want_grad=True, tile=0, threads=1):
    pos = f64(pos).reshape(12)
    pi = f64(pi)
    blc = f64(bl)
    N = prob.num_nodes
    site = np.zeros(max(prob.num_patterns, 1))
    gbl = np.zeros(N)
    gom = np.zeros(len(mix.omegas))
    gw = np.zeros(max(mix.n_comp, 1))
    gth = np.zeros(5)
    dtau = np.zeros(N)
    out = OrMixtureOut()
    out.site_ll = ptr(site, _f64p)
    out.grad_bl = ptr(gbl, _f64p)
    out.grad_omega = ptr(gom, _f64p)
    out.grad_w = ptr(gw, _f64p)
    out.grad_theta = ptr(gth, _f64p)
    out.dtau = ptr(dtau, _f64p)
    rc = lib().or_mixture_eval(C.byref(prob.t), C.byref(code.c), C.byref(base), ptr(pos, _f64p), ptr(pi, _f64p),
                               C.byref(mix.m), ptr(blc, _f64p), 1 if want_grad else 0, tile, threads, C.byref(out))
    if rc != 0:
        raise RuntimeError(f"or_mixture_eval failed ({rc})")
    return {"log_likelihood": out.log_likelihood, "site_ll": site[:prob.num_patterns], "grad_bl": gbl,
            "grad_omega": gom, "grad_w": gw[:mix.n_comp], "grad_theta": gth, "dtau": dtau}


def srv_eval(prob, code, base, pos, pi, mix, bl, syn_rates, syn_weights, want_grad=True, threads=1):
    """BUSTED synonymous-rate classes over the branch-site mixture
    (busted.hpp:509-659, 1138-1429); see or_srv_eval."""
    pos = f64(pos).reshape(12)
    pi = f64(pi)
    blc = f64(bl)
    sr = f64(np.atleast_1d(syn_rates))
    sw = f64(np.atleast_1d(syn_weights))
    M = len(sr)
    N = prob.num_nodes
    site = np.zeros(max(prob.num_patterns, 1))
    gbl, gom, gw, gth = np.zeros(N), np.zeros(len(mix.omegas)), np.zeros(max(mix.n_comp, 1)), np.zeros(5)
    gsr, gsw = np.zeros(M), np.zeros(M)
    out = OrMixtureOut()
    out.site_ll = ptr(site, _f64p)
    out.grad_bl = ptr(gbl, _f64p)
    out.grad_omega = ptr(gom, _f64p)
    out.grad_w = ptr(gw, _f64p)
    out.grad_theta = ptr(gth, _f64p)
    rc = lib().or_srv_eval(C.byref(prob.t), C.byref(code.c), C.byref(base), ptr(pos, _f64p), ptr(pi, _f64p),
                           C.byref(mix.m), ptr(blc, _f64p), M, ptr(sr, _f64p), ptr(sw, _f64p),
                           1 if want_grad else 0, threads, C.byref(out), ptr(gsr, _f64p), ptr(gsw, _f64p))
    if rc != 0:
        raise RuntimeError(f"or_srv_eval failed ({rc})")
    return {"log_likelihood": out.log_likelihood, "site_ll": site[:prob.num_patterns], "grad_bl": gbl,
            "grad_omega": gom, "grad_w": gw[:mix.n_comp], "grad_theta": gth, "grad_syn_rates": gsr,
            "grad_syn_weights": gsw}


def workspace(prob, P, pi, p0, p1, outside=True):
    """The reference's tile workspace (compute_inside[_outside]_batch_in_place)
    for patterns [p0, p1): arrays [node][state][pattern] and [node][pattern]."""
    N, S, B = prob.num_nodes, prob.t.S, p1 - p0
    D = np.zeros((N, S, B))
    V, PD, U = (np.zeros((N, S, B)) for _ in range(3))
    boost = np.zeros((N, B), dtype=np.int32)
    logsc = np.zeros((N, B))
    L, logL = np.zeros(B), np.zeros(B)
    Pc = f64(P)
    rc = lib().or_workspace(C.byref(prob.t), ptr(Pc, _f64p), ptr(f64(pi), _f64p), p0, p1, 1 if outside else 0,
                            ptr(D, _f64p), ptr(V, _f64p), ptr(PD, _f64p), ptr(U, _f64p), ptr(boost, _i32p),
                            ptr(logsc, _f64p), ptr(L, _f64p), ptr(logL, _f64p))
    if rc != 0:
        raise RuntimeError("or_workspace failed")
    out = {"D": D, "node_log_scale": logsc, "L": L, "logL": logL}
    if outside:
        out.update(V=V, P_D=PD, U=U, boost_counts=boost)
    return out


def busted_expected_weights(prob, code, base, pos, pi, weights_p, omegas, weights, omega_bg, is_test, tau,
                            syn_rates=(1.0,), syn_weights=(1.0,), eps_sink=None, max_eps_sink=0.01):
    """BUSTEDAnalyzer::compute_expected_weights (busted.hpp:662-976) and
    compute_expected_syn_weights (busted.hpp:978-1135) without bridges,
    restated literally over the reference's inside-outside workspaces: per
    synonymous class m the mixed P(r_m tau) (test branches sum_k b w_k P_k
    with b = 1 - eps_sink, plus eps_sink P_sink with the error sink;
    background P_bg, :715-733), the workspace of the whole alignment
    (:791-799), log-sum-exp posteriors over classes (:802-823), and per test
    branch, component and class the share b w_k post_m V_b . (P_k D_b) / L
    (:826-845), with the sink the share eps post_m V_b . P_sink D_b / L
    (eval_sink_bilinear, :208-216, :846-860); totals over patterns with a
    positive likelihood (:847-866, :942-975), the new eps clamped to
    [1e-7, max_eps_sink] (:964-966).  tau: raw branch times per node;
    weights_p: pattern weights; the sink matrix comes from pi
    (build_sink_kernel, :193-206).  Returns (new omega weights, new
    synonymous weights), and the new eps when eps_sink is given."""
    N, S = prob.num_nodes, code.S
    sink = eps_sink is not None
    bio = (1.0 - eps_sink) if sink else 1.0
    if sink:
        inv = 1.0 / np.maximum(1e-12, 1.0 - np.asarray(pi, dtype=float))
        P_sink = np.asarray(pi, dtype=float)[None, :] * inv[:, None]
        np.fill_diagonal(P_sink, 0.0)
    K, M = len(omegas), len(syn_rates)
    P = prob.num_patterns
    dec = []
    for om in list(omegas) + [omega_bg]:
        b = params(1.0, om, list(base.theta), base.delta, base.psi)
        Q, _ = mg94_Q(code, b, pos, pi)
        dec.append(decompose(Q, pi))
    wsm, Pc = [], []
    for m in range(M):
        comp = np.zeros((K, N, S, S))
        mix = np.zeros((N, S, S))
        for n in range(N):
            for k in range(K):
                comp[k, n] = transition(*dec[k], syn_rates[m] * tau[n])
            if is_test[n]:
                mix[n] = sum(bio * weights[k] * comp[k, n] for k in range(K))
                if sink:
                    mix[n] = mix[n] + eps_sink * P_sink
            else:
                mix[n] = transition(*dec[K], syn_rates[m] * tau[n])
        Pc.append(comp)
        wsm.append(workspace(prob, mix, pi, 0, P, outside=True) if syn_weights[m] > 1e-12 else None)
    # posteriors over synonymous classes
    terms = np.full((M, P), -np.inf)
    for m in range(M):
        if wsm[m] is None:
            continue
        ok = wsm[m]["L"] > 0.0
        terms[m, ok] = np.log(syn_weights[m]) + wsm[m]["logL"][ok]
    mx = terms.max(axis=0)
    valid = mx > -1e19
    post = np.zeros((M, P))
    e = np.where(np.isfinite(terms), np.exp(terms - np.where(valid, mx, 0.0)), 0.0)
    post[:, valid] = e[:, valid] / e[:, valid].sum(axis=0)
    acc = np.zeros(K)
    sink_acc = 0.0
    tot = 0.0
    counted = post.sum(axis=0) > 0.0
    for n in range(N):
        if not is_test[n]:
            continue
        for k in range(K):
            for m in range(M):
                if wsm[m] is None:
                    continue
                ws = wsm[m]
                PD = Pc[m][k, n] @ ws["D"][n]
                ratio = np.einsum("sb,sb->b", ws["V"][n], PD)
                use = (post[m] > 1e-12) & (ws["L"] > 0.0)
                acc[k] += np.sum(weights_p[use] * bio * weights[k] * post[m, use] * ratio[use] / ws["L"][use])
        if sink:
            sink_ratio = np.zeros(P)
            for m in range(M):
                if wsm[m] is None:
                    continue
                ws = wsm[m]
                # V^T max(0, P_sink D) per pattern
                bil = np.einsum("sb,sb->b", ws["V"][n], np.maximum(0.0, P_sink @ ws["D"][n]))
                use = counted & (post[m] > 1e-12) & (ws["L"] > 0.0)
                sink_ratio[use] += post[m, use] * bil[use] / ws["L"][use]
            sink_acc += np.sum(weights_p[counted] * eps_sink * sink_ratio[counted])
        tot += np.sum(weights_p[counted])

    def norm(num, den, prev):
        if den <= 0.0:
            return np.array(prev, dtype=float)
        x = np.clip(num / den, 1e-6, 1.0 - 1e-6)
        return x / x.sum()

    new_w = np.ones(1) if K <= 1 else norm(acc, tot, weights)
    if M <= 1:
        new_v = np.ones(1)
    else:
        syn_acc = (weights_p[None, :] * post).sum(axis=1)
        new_v = norm(syn_acc, np.sum(weights_p[valid]), syn_weights)
    if not sink:
        return new_w, new_v
    new_eps = float(np.clip(sink_acc / max(1.0, tot), 1e-7, max_eps_sink)) if tot > 0.0 else eps_sink
    return new_w, new_v, new_eps


# ---------------------------------------------------------------------------
# aBSREL local projection (one branch, candidate transition matrices)
# ---------------------------------------------------------------------------
def _neumaier_acc(s, c, x):
    """LikelihoodEngine::neumaier_acc (likelihood.hpp:34-43)."""
    t = s + x
    if abs(s) >= abs(x):
        c += (s - t) + x
    else:
        c += (x - t) + s
    return t, c


def branch_context(prob, P, pi, node, tile=256):
    """ABSRELAnalyzer::compute_branch_context_matrices (absrel.hpp:723-780)
    for the branch above ``node`` (no bridges): the reference's inside-outside
    workspace per tile of 256 patterns, V_mat = V[node], D_mat = D[node]
    ([S, P], boosted representation) and log_scale_offset, the sum in pattern
    order of w_p (logL_p - ln L_p) over patterns with L_p > 0."""
    S, npat = prob.t.S, prob.num_patterns
    Vm = np.zeros((S, npat))
    Dm = np.zeros((S, npat))
    off = np.zeros(npat)
    for p0 in range(0, npat, tile):
        p1 = min(p0 + tile, npat)
        ws = workspace(prob, P, pi, p0, p1, outside=True)
        Vm[:, p0:p1] = ws["V"][node]
        Dm[:, p0:p1] = ws["D"][node]
        for b in range(p1 - p0):
            if ws["L"][b] > 0.0:
                off[p0 + b] = prob.weights[p0 + b] * (ws["logL"][b] - np.log(ws["L"][b]))
    total = 0.0
    for p in range(npat):
        total += off[p]
    return Vm, Dm, total


def branch_ll(prob, P_b, Vm, Dm, log_scale_offset):
    """ABSRELAnalyzer::eval_branch_ll_from_matrices (absrel.hpp:438-455):
    offset + sum_p w_p ln(V_p . (P_b D)_p), w_p (-1e20) where the product is
    not positive, Neumaier-accumulated in pattern order."""
    PD = np.asarray(P_b, dtype=np.float64) @ Dm
    ll, c = log_scale_offset, 0.0
    for p in range(prob.num_patterns):
        Lp = float(Vm[:, p] @ PD[:, p])
        term = prob.weights[p] * np.log(Lp) if Lp > 0.0 else prob.weights[p] * -1e20
        ll, c = _neumaier_acc(ll, c, term)
    return ll + c


# ---------------------------------------------------------------------------
# Nucleotide GTR stage (S = 4)
# ---------------------------------------------------------------------------
GTR_PAIRS = ((0, 1), (0, 3), (1, 2), (1, 3), (2, 3))  # theta order AC, AT, CG, CT, GT; AG = 1


def gtr_Q(theta, pi):
    """GTRMatrix::update (rate_matrix.hpp:29-90): Q_ij = r_ij pi_j with r_AG = 1,
    diagonal minus the row sum, divided by mu = -sum_i pi_i Q_ii when mu > 0.
    Returns (normalised Q, mu)."""
    r = np.ones((4, 4))
    for k, (a, b) in enumerate(GTR_PAIRS):
        r[a, b] = r[b, a] = theta[k]
    np.fill_diagonal(r, 0.0)
    pi = np.asarray(pi, dtype=float)
    Q = r * pi[None, :]
    Q -= np.diag(Q.sum(axis=1))
    mu = -float(np.sum(pi * np.diag(Q)))
    if mu > 0.0:
        Q = Q / mu
    return Q, mu


def gtr_eval(prob, theta, pi, bl, want_grad=True, threads=1):
    """compute_gtr_log_likelihood (likelihood.hpp:181-254): P_b = U diag(e^{l t_b})
    V of the normalised generator (transition_matrix, rate_matrix.hpp:93-99,
    clamped >= 0), pruning with the alignment's nucleotide frequencies at the
    root, Neumaier sum of w_p lnL_p.  With gradients (no counterpart in the
    reference, whose GTRFitter is derivative-free; the same adjoint as the
    codon joint gradient, likelihood.hpp:1435-1520): from the reference's
    inside-outside workspace, A_b = sum_p w_p / L_p V_b D_b^T over patterns with
    L_p > 0; d lnL / d t_b = <A_b, Q P_b>; d lnL / d theta_k = sum_b <U^T A_b V^T,
    (V dQ_k U) o E(t_b)> with E_ij = (e^{l_i t} - e^{l_j t}) / (l_i - l_j) (t e^{l_i t}
    when |l_i - l_j| < 1e-9) and dQ_k = (dQ/dtheta_k - Q dmu/dtheta_k) / mu the
    derivative of the normalised generator."""
    pi = np.asarray(pi, dtype=float)
    Q, mu = gtr_Q(theta, pi)
    ev, U, V = decompose(Q, pi)
    N = prob.num_nodes
    bl = np.asarray(bl, dtype=float)
    P = np.stack([np.eye(4) if n == prob.root else transition(ev, U, V, bl[n]) for n in range(N)])
    ll, site = loglik_from_P(prob, P, pi, threads=threads, site=True)
    out = {"log_likelihood": ll, "site_ll": site}
    if not want_grad:
        return out
    ws = workspace(prob, P, pi, 0, prob.num_patterns, outside=True)
    ok = ws["L"] > 0.0
    c = np.where(ok, prob.weights / np.where(ok, ws["L"], 1.0), 0.0)
    A = np.einsum("nip,njp,p->nij", ws["V"], ws["D"], c)
    gbl = np.zeros(N)
    gth = np.zeros(5)
    for k, (a, b) in enumerate(GTR_PAIRS):
        dQ = np.zeros((4, 4))
        dQ[a, b], dQ[b, a] = pi[b], pi[a]
        dQ[a, a], dQ[b, b] = -pi[b], -pi[a]
        dmu = 2.0 * pi[a] * pi[b]
        M = V @ ((dQ - Q * dmu) / mu) @ U
        for n in range(N):
            if n == prob.root:
                continue
            t = bl[n]
            e = np.exp(ev * t)
            d = ev[:, None] - ev[None, :]
            close = np.abs(d) < 1e-9
            E = np.where(close, t * e[:, None] * np.ones(4)[None, :], (e[:, None] - e[None, :]) / np.where(close, 1.0, d))
            gth[k] += np.sum((U.T @ A[n] @ V.T) * M * E)
    for n in range(N):
        if n != prob.root:
            gbl[n] = np.sum(A[n] * (Q @ P[n]))
    out.update(grad_bl=gbl, grad_theta=gth, A=A)
    return out


def nuc_problem(tree, naln):
    """Problem of the oracle over a nucleotide view (data_ref.NucleotideRef)."""
    return Problem.from_refs(tree, naln)
