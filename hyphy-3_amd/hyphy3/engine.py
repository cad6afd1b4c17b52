"""ctypes binding of the MI355X engine's C ABI (include/hyphy_mi355x.h).

The shared library ``libhyphy_mi355x.so`` is built in-tree by
``make -C hyphy-3_amd`` (or ``__graft_entry__.build()``).  Importing this
module never falls back to a CPU implementation: a missing library or a
failing HIP call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "libhyphy_mi355x.so")

HY_OK = 0
HY_ERR_ARG = -1
HY_ERR_STATE = -3
HY_ERR_NOMEM = -4
HY_ERR_UNSUPPORTED = -5
HY_PREC_FP64 = 0
HY_PREC_FP32 = 1
HY_SINK = -1  # matrix index naming BUSTED's error sink in set_mixture classes
HY_TIME_RAW = 0
HY_TIME_CONV = 1
HY_WANT_LOGL = 1
HY_WANT_SITE_LL = 2
HY_WANT_GRAD = 4

EXPORTS = ("hy_create", "hy_destroy", "hy_last_error", "hy_set_problem", "hy_set_model", "hy_set_mixture",
           "hy_eval", "hy_loglik_from_P", "hy_set_tile", "hy_last_timings", "hy_device_bytes",
           "hy_set_warm_start", "hy_last_eigen_status", "hy_last_tile", "hy_last_parts", "hy_eval_srv", "hy_eval_batch", "hy_create_group",
           "hy_inside_outside_from_P", "hy_read_partials", "hy_autotune", "hy_set_tuning", "hy_busted_expected_weights",
           "hy_autotune_batch", "hy_set_sink", "hy_loglik_from_P_noalloc", "hy_graph_count",
           "hy_set_memory_budget", "hy_last_checkpoint", "hy_gtr_eval", "hy_fel_eval", "hy_fel_drive", "hy_fel_run",
           "hy_fel_counters", "hy_mg94_spectrum", "hy_branch_eval", "hy_set_pattern_weights", "hy_network_create",
           "hy_network_destroy", "hy_network_last_error", "hy_network_set_problem", "hy_network_set_model",
           "hy_network_set_mixture", "hy_network_set_sink", "hy_network_eval", "hy_meme_eval", "hy_meme_drive",
           "hy_meme_run", "hy_lbfgs_exp", "hy_lbfgs_state_doubles", "hy_lbfgs_init", "hy_lbfgs_step", "hy_fit_lbfgs")

# optimizer status (hy_lbfgs_head.status)
HY_LBFGS_RUNNING, HY_LBFGS_CONV_GRAD, HY_LBFGS_CONV_F, HY_LBFGS_MAX_ITER, HY_LBFGS_LS_FAIL, HY_LBFGS_NUMERIC = range(6)

HY_FEL_FIELDS = 10
FEL_FIELD_NAMES = ("alpha", "beta", "beta_nuisance", "alpha_null", "lrt", "p_value", "total_branch_length", "log_l_alt",
                   "log_l_null", "evaluations")
# int (*)(void *user, int n, const int32_t *pattern, const double *alpha, const double *beta, double *out_ll)
FEL_EVAL_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_double),
                          C.POINTER(C.c_double), C.POINTER(C.c_double))

HY_MEME_FIELDS = 16
MEME_FIELD_NAMES = ("alpha", "beta1", "p1", "beta_plus", "p_plus", "lrt", "p_value", "branches_under_selection",
                    "total_branch_length", "log_l_meme", "log_l_fel", "p_value_meme_vs_fel", "fel_alpha", "fel_beta",
                    "beta_bg", "evaluations")
# int (*)(void *user, int n, const int32_t *pattern, const double *alpha, const double *omega1, const double *p1,
#         const double *beta_plus, const int32_t *cond_node, const int32_t *cond_class, double *out_ll)
MEME_EVAL_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_double),
                           C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                           C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_double))

_i32p = C.POINTER(C.c_int32)
_i8p = C.POINTER(C.c_int8)
_u8p = C.POINTER(C.c_uint8)
_u64p = C.POINTER(C.c_uint64)
_f64p = C.POINTER(C.c_double)

_lib = None


class BatchItem(C.Structure):
    """hy_batch_item of the C ABI."""
    _fields_ = [("h", C.c_void_p), ("bl", _f64p), ("alpha", C.c_double), ("omegas", _f64p), ("comp_w", _f64p),
                ("theta", _f64p), ("delta", C.c_double), ("psi", C.c_double), ("out_logL", _f64p),
                ("out_site_ll", _f64p), ("out_grad_bl", _f64p), ("out_grad_omega", _f64p), ("out_grad_w", _f64p),
                ("out_grad_theta", _f64p), ("status", C.c_int)]


class LbfgsHead(C.Structure):
    """hy_lbfgs_head of the C ABI."""
    _fields_ = [("n", C.c_int32), ("m", C.c_int32), ("status", C.c_int32), ("iter", C.c_int32),
                ("n_eval", C.c_int32), ("k", C.c_int32), ("head", C.c_int32), ("n_back", C.c_int32),
                ("max_iter", C.c_int32), ("max_back", C.c_int32), ("pad", C.c_int32 * 2), ("f", C.c_double),
                ("f_prev", C.c_double), ("alpha", C.c_double), ("gd", C.c_double), ("tol_grad", C.c_double),
                ("tol_f", C.c_double), ("c1", C.c_double), ("pgnorm", C.c_double)]


class HyError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hyphy_mi355x error {code}: {msg}")
        self.code = code


def load_library(path: str = LIB_PATH):
    """Load libhyphy_mi355x.so and declare its prototypes (no device work)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"MI355X engine library not built: {path} (run make -C hyphy-3_amd)")
    L = C.CDLL(path)
    L.hy_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    L.hy_destroy.argtypes = [C.c_void_p]
    L.hy_destroy.restype = None
    L.hy_last_error.argtypes = [C.c_void_p]
    L.hy_last_error.restype = C.c_char_p
    L.hy_set_problem.argtypes = [C.c_void_p, C.c_int, _i32p, _i32p, _i32p, _i32p, C.c_int, C.c_int, _i8p, _u64p,
                                 _f64p, C.c_int]
    L.hy_set_model.argtypes = [C.c_void_p, C.c_int, _f64p, _f64p]
    L.hy_set_mixture.argtypes = [C.c_void_p, C.c_int, C.c_int, _i32p, _i32p, _i32p, C.c_int]
    L.hy_eval.argtypes = [C.c_void_p, C.c_int, C.c_uint, _f64p, C.c_double, _f64p, _f64p, _f64p, C.c_double,
                          C.c_double, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p]
    L.hy_loglik_from_P.argtypes = [C.c_void_p, C.c_int, _f64p, _f64p, _f64p, _f64p]
    L.hy_inside_outside_from_P.argtypes = [C.c_void_p, C.c_int, _f64p, _f64p, _f64p, _f64p]
    L.hy_loglik_from_P_noalloc.argtypes = [C.c_void_p, C.c_int, _f64p, _f64p, _f64p, _f64p]
    L.hy_graph_count.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.hy_set_memory_budget.argtypes = [C.c_void_p, C.c_size_t]
    L.hy_gtr_eval.argtypes = [C.c_void_p, C.c_int, C.c_uint, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p]
    L.hy_fel_eval.argtypes = [C.c_void_p, C.c_int, _i32p, _f64p, _f64p, _f64p, C.c_double, C.c_double, _f64p, _f64p,
                              _f64p]
    L.hy_fel_drive.argtypes = [C.c_int, _u8p, C.c_double, FEL_EVAL_FN, C.c_void_p, _f64p, C.POINTER(C.c_int),
                               C.POINTER(C.c_int64)]
    L.hy_fel_run.argtypes = [C.c_void_p, _f64p, C.c_double, C.c_double, C.c_double, C.c_double, _f64p, _f64p,
                             C.POINTER(C.c_int), C.POINTER(C.c_int64), _f64p]
    L.hy_fel_counters.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.c_int]
    L.hy_set_pattern_weights.argtypes = [C.c_void_p, _f64p]
    L.hy_meme_eval.argtypes = [C.c_void_p, C.c_int, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p, _u8p, _f64p,
                               C.c_double, C.c_double, _f64p, _i32p, _i32p, _f64p]
    L.hy_meme_drive.argtypes = [C.c_int, _u8p, _f64p, C.c_int, C.c_int, MEME_EVAL_FN, C.c_void_p, _f64p, _f64p,
                                C.POINTER(C.c_int), C.POINTER(C.c_int64)]
    L.hy_meme_run.argtypes = [C.c_void_p, _f64p, C.c_double, C.c_double, C.c_double, C.c_double, _f64p, _f64p,
                              _f64p, _f64p, C.POINTER(C.c_int), C.POINTER(C.c_int64)]
    L.hy_branch_eval.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _f64p, _f64p, _f64p, _i32p]
    L.hy_mg94_spectrum.argtypes = [C.c_void_p, C.c_double, C.c_double, _f64p, C.c_double, C.c_double, _f64p, _f64p,
                                   _f64p, _f64p]
    L.hy_last_checkpoint.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_size_t)]
    L.hy_set_tile.argtypes = [C.c_void_p, C.c_int]
    L.hy_set_sink.argtypes = [C.c_void_p, _f64p]
    L.hy_last_timings.argtypes = [C.c_void_p, _f64p]
    L.hy_device_bytes.argtypes = [C.c_void_p, C.POINTER(C.c_size_t)]
    L.hy_set_warm_start.argtypes = [C.c_void_p, C.c_int]
    L.hy_last_eigen_status.argtypes = [C.c_void_p, _i32p, C.c_int]
    L.hy_last_tile.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    L.hy_last_parts.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    L.hy_eval_srv.argtypes = [C.c_void_p, C.c_int, C.c_uint, _f64p, C.c_double, _f64p, _f64p, _f64p, C.c_double,
                              C.c_double, C.c_int, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p,
                              _f64p]
    L.hy_eval_batch.argtypes = [C.c_int, C.POINTER(BatchItem), C.c_int, C.c_uint]
    L.hy_create_group.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
    L.hy_autotune.argtypes = [C.c_void_p, C.c_int, C.c_uint, _f64p, C.c_double, _f64p, _f64p, _f64p, C.c_double,
                              C.c_double, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _f64p]
    L.hy_set_tuning.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    L.hy_busted_expected_weights.argtypes = [C.c_void_p, C.c_int, _f64p, C.c_double, _f64p, _f64p, _f64p, C.c_double,
                                             C.c_double, C.c_int, _f64p, _f64p, C.c_int, C.c_double, _f64p, _f64p]
    L.hy_autotune_batch.argtypes = [C.c_int, C.POINTER(BatchItem), C.c_int, C.c_uint, C.c_int, C.POINTER(C.c_int),
                                    _f64p]
    L.hy_lbfgs_exp.restype = C.c_double
    L.hy_lbfgs_exp.argtypes = [C.c_double]
    L.hy_lbfgs_state_doubles.restype = C.c_size_t
    L.hy_lbfgs_state_doubles.argtypes = [C.c_int, C.c_int]
    L.hy_lbfgs_init.argtypes = [C.c_int, C.c_int, _f64p, C.POINTER(LbfgsHead), _f64p, _f64p, _f64p, C.c_int,
                                C.c_double, C.c_double]
    L.hy_lbfgs_step.argtypes = [_f64p, C.POINTER(LbfgsHead), C.c_double, _f64p, C.c_int, _f64p]
    L.hy_fit_lbfgs.argtypes = [C.c_void_p, C.c_int, C.c_int, _f64p, C.c_double, _f64p, _f64p, _f64p, C.c_double,
                               C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _f64p,
                               _i32p]
    _lib = L
    return L


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Engine:
    """One MI355X likelihood engine (one device, one stream)."""

    def __init__(self, device: int = 0, handle=None):
        self.L = load_library()
        if handle is None:
            h = C.c_void_p()
            rc = self.L.hy_create(device, C.byref(h))
            if rc != HY_OK:
                raise HyError(rc, self.L.hy_last_error(None).decode())
        else:
            h = C.c_void_p(handle)
        self.h = h
        self.num_nodes = 0
        self.num_patterns = 0
        self.n_mat = 1
        self.n_comp = 1

    def _check(self, rc):
        if rc != HY_OK:
            raise HyError(rc, self.L.hy_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.hy_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_problem(self, parent, children, leaf_taxon, states, masks, weights, S=61):
        n = len(parent)
        self._parent = np.ascontiguousarray(parent, dtype=np.int32)
        cp, ci = [0], []
        for k in range(n):
            ci.extend(children[k])
            cp.append(len(ci))
        self._cp = np.array(cp, dtype=np.int32)
        self._ci = np.array(ci if ci else [0], dtype=np.int32)
        self._l2t = np.ascontiguousarray(leaf_taxon, dtype=np.int32)
        st = np.ascontiguousarray(states, dtype=np.int8)
        mk = np.ascontiguousarray(masks, dtype=np.uint64)
        w = _f64(weights)
        P = int(st.shape[0])
        ntax = int(st.shape[1]) if st.ndim == 2 else 0
        self._check(self.L.hy_set_problem(self.h, n, _p(self._parent, _i32p), _p(self._cp, _i32p),
                                          _p(self._ci, _i32p), _p(self._l2t, _i32p), ntax, P, _p(st, _i8p),
                                          _p(mk, _u64p), _p(w, _f64p), S))
        self.num_nodes, self.num_patterns = n, P
        self.n_mat, self.n_comp = 1, 1
        self._class_sizes = [1]  # hy_set_problem restores the one-component mixture
        self.weights = w.copy()  # the alignment's pattern weights
        self.root = int(np.where(self._parent < 0)[0][0])
        self.S = int(S)

    def _check_states(self, name, a):
        # the C side reads exactly S values (one per sense codon)
        S = getattr(self, "S", None)
        if S is not None and a.shape != (S,):
            raise ValueError("%s: expected %d codon frequencies (one per sense codon), got shape %s"
                             % (name, S, a.shape))

    def set_model(self, code_id, pos_freqs, pi):
        pos = _f64(pos_freqs)
        if pos.size != 12:
            raise ValueError("set_model: pos_freqs must hold 3 x 4 position frequencies, got %d" % pos.size)
        pos = pos.reshape(12)
        pi = _f64(pi)
        self._check_states("set_model", pi)
        self._check(self.L.hy_set_model(self.h, code_id, _p(pos, _f64p), _p(pi, _f64p)))
        if getattr(self, "S", None) is None:  # a model without a problem (spectra only)
            self.S = int(pi.size)

    def set_mixture(self, n_mat, branch_class, classes, convention=HY_TIME_RAW):
        """classes: list (per class) of lists of matrix indices; HY_SINK names
        the error sink (set_sink), whose weight multiplies P_sink."""
        cp, cm = [0], []
        for comps in classes:
            cm.extend(int(m) for m in comps)
            cp.append(len(cm))
        cp = np.array(cp, dtype=np.int32)
        cm = np.array(cm, dtype=np.int32)
        bc = None if branch_class is None else np.ascontiguousarray(branch_class, dtype=np.int32)
        self._check(self.L.hy_set_mixture(self.h, n_mat, len(classes), _p(bc, _i32p), _p(cp, _i32p),
                                          _p(cm, _i32p), convention))
        self.n_mat, self.n_comp = n_mat, len(cm)
        self._class_sizes = [len(c) for c in classes]

    def set_sink(self, freqs):
        """Error-sink matrix from the codon frequencies (busted.hpp:193-206)."""
        f = np.ascontiguousarray(freqs, dtype=np.float64)
        self._check_states("set_sink", f)
        self._check(self.L.hy_set_sink(self.h, _p(f, _f64p)))

    def set_tile(self, T):
        self._check(self.L.hy_set_tile(self.h, int(T)))

    def _staging(self):
        """Persistent argument and result buffers of hy_eval with their
        addresses, rebuilt when the problem or mixture changes shape; saves
        the per-call ctypes pointer conversions of eval (about 2 us each)."""
        key = (self.num_nodes, self.n_mat, self.n_comp, self.num_patterns)
        st = getattr(self, "_stage", None)
        if st is None or st["key"] != key:
            if getattr(self, "_fast_eval", None) is None:
                f = self.L["hy_eval"]
                vp = C.c_void_p
                f.argtypes = [vp, C.c_int, C.c_uint, vp, C.c_double, vp, vp, vp, C.c_double, C.c_double,
                              vp, vp, vp, vp, vp, vp]
                f.restype = C.c_int
                self._fast_eval = f
            arr = {k: np.zeros(max(n, 1)) for k, n in (("bl", self.num_nodes), ("om", self.n_mat),
                                                       ("cw", self.n_comp), ("th", 5), ("ll", 1),
                                                       ("site", self.num_patterns), ("gbl", self.num_nodes),
                                                       ("gom", self.n_mat), ("gw", self.n_comp), ("gth", 5))}
            st = dict(key=key, arr=arr, adr={k: a.ctypes.data for k, a in arr.items()})
            self._stage = st
        return st

    def eval(self, bl, omegas, comp_w, theta, alpha=1.0, delta=0.0, psi=0.0, precision=HY_PREC_FP64, grad=True,
             site=False):
        st = self._staging()
        a, d = st["arr"], st["adr"]
        try:
            a["bl"][:self.num_nodes] = bl
            a["om"][:self.n_mat] = omegas
            a["cw"][:self.n_comp] = comp_w
            a["th"][:] = theta
        except ValueError:
            return self._eval_general(bl, omegas, comp_w, theta, alpha, delta, psi, precision, grad, site)
        flags = HY_WANT_LOGL | (HY_WANT_SITE_LL if site else 0) | (HY_WANT_GRAD if grad else 0)
        self._check(self._fast_eval(self.h, precision, flags, d["bl"], alpha, d["om"], d["cw"], d["th"], delta,
                                    psi, d["ll"], d["site"] if site else None, *(
                                        (d["gbl"], d["gom"], d["gw"], d["gth"]) if grad else (None,) * 4)))
        out = {"log_likelihood": float(a["ll"][0])}
        if site:
            out["site_ll"] = a["site"][:self.num_patterns].copy()
        if grad:
            out.update(grad_bl=a["gbl"][:self.num_nodes].copy(), grad_omega=a["gom"][:self.n_mat].copy(),
                       grad_w=a["gw"][:self.n_comp].copy(), grad_theta=a["gth"].copy())
        return out

    def _eval_general(self, bl, omegas, comp_w, theta, alpha, delta, psi, precision, grad, site):
        bl = _f64(bl)
        om = _f64(np.atleast_1d(omegas))
        cw = _f64(np.atleast_1d(comp_w))
        th = _f64(theta)
        ll = C.c_double()
        out = {}
        s = np.zeros(max(self.num_patterns, 1)) if site else None
        gbl = np.zeros(self.num_nodes) if grad else None
        gom = np.zeros(self.n_mat) if grad else None
        gw = np.zeros(self.n_comp) if grad else None
        gth = np.zeros(5) if grad else None
        flags = HY_WANT_LOGL | (HY_WANT_SITE_LL if site else 0) | (HY_WANT_GRAD if grad else 0)
        self._check(self.L.hy_eval(self.h, precision, flags, _p(bl, _f64p), alpha, _p(om, _f64p), _p(cw, _f64p),
                                   _p(th, _f64p), delta, psi, C.byref(ll), _p(s, _f64p), _p(gbl, _f64p),
                                   _p(gom, _f64p), _p(gw, _f64p), _p(gth, _f64p)))
        out["log_likelihood"] = ll.value
        if site:
            out["site_ll"] = s[:self.num_patterns]
        if grad:
            out.update(grad_bl=gbl, grad_omega=gom, grad_w=gw, grad_theta=gth)
        return out

    def eval_srv(self, bl, omegas, comp_w, theta, syn_rates, syn_weights, alpha=1.0, delta=0.0, psi=0.0,
                 precision=HY_PREC_FP64, grad=True, site=False):
        """Synonymous-rate mixture (M = len(syn_rates) classes) over the
        branch-site mixture; adds grad_syn_rates / grad_syn_weights."""
        bl = _f64(bl)
        om = _f64(np.atleast_1d(omegas))
        cw = _f64(np.atleast_1d(comp_w))
        th = _f64(theta)
        sr = _f64(np.atleast_1d(syn_rates))
        sw = _f64(np.atleast_1d(syn_weights))
        M = len(sr)
        if len(sw) != M:
            raise ValueError("syn_rates and syn_weights differ in length")
        ll = C.c_double()
        s = np.zeros(max(self.num_patterns, 1)) if site else None
        g = [np.zeros(k) if grad else None for k in (self.num_nodes, self.n_mat, self.n_comp, 5, M, M)]
        flags = HY_WANT_LOGL | (HY_WANT_SITE_LL if site else 0) | (HY_WANT_GRAD if grad else 0)
        self._check(self.L.hy_eval_srv(self.h, precision, flags, _p(bl, _f64p), alpha, _p(om, _f64p), _p(cw, _f64p),
                                       _p(th, _f64p), delta, psi, M, _p(sr, _f64p), _p(sw, _f64p), C.byref(ll),
                                       _p(s, _f64p), *[_p(x, _f64p) for x in g]))
        out = {"log_likelihood": ll.value}
        if site:
            out["site_ll"] = s[:self.num_patterns]
        if grad:
            out.update(grad_bl=g[0], grad_omega=g[1], grad_w=g[2], grad_theta=g[3], grad_syn_rates=g[4],
                       grad_syn_weights=g[5])
        return out

    def busted_expected_weights(self, bl, omegas, comp_w, theta, syn_rates=(1.0,), syn_weights=(1.0,), test_class=0,
                                alpha=1.0, delta=0.0, psi=0.0, precision=HY_PREC_FP64, max_eps_sink=0.01):
        """BUSTED's expectation step (busted.hpp:662-1135): updated weights of
        the components of ``test_class`` (omega components renormalised; an
        error-sink component gets the new eps, at most max_eps_sink) and
        synonymous-rate weights.  Needs raw branch times (HY_TIME_RAW
        mixture)."""
        bl = _f64(bl)
        om = _f64(np.atleast_1d(omegas))
        cw = _f64(np.atleast_1d(comp_w))
        th = _f64(theta)
        sr = _f64(np.atleast_1d(syn_rates))
        sw = _f64(np.atleast_1d(syn_weights))
        if len(sr) != len(sw):
            raise ValueError("syn_rates and syn_weights differ in length")
        K = int(self._class_sizes[test_class]) if hasattr(self, "_class_sizes") else len(cw)
        nw = np.zeros(max(K, 1))
        nv = np.zeros(len(sr))
        self._check(self.L.hy_busted_expected_weights(self.h, precision, _p(bl, _f64p), alpha, _p(om, _f64p),
                                                      _p(cw, _f64p), _p(th, _f64p), delta, psi, len(sr),
                                                      _p(sr, _f64p), _p(sw, _f64p), int(test_class),
                                                      float(max_eps_sink), _p(nw, _f64p), _p(nv, _f64p)))
        return nw[:K], nv

    def loglik_from_P(self, P, pi, precision=HY_PREC_FP64, site=False):
        P = _f64(P)
        pi = _f64(pi)
        ll = C.c_double()
        s = np.zeros(max(self.num_patterns, 1)) if site else None
        self._check(self.L.hy_loglik_from_P(self.h, precision, _p(P, _f64p), _p(pi, _f64p), C.byref(ll),
                                            _p(s, _f64p)))
        return (ll.value, s[:self.num_patterns]) if site else ll.value

    def gtr_eval(self, bl, theta, freqs, precision=HY_PREC_FP64, grad=True, site=False):
        """Nucleotide GTR stage (hy_gtr_eval) on a problem set with S = 4:
        lnL and, with grad, d lnL / d bl and d lnL / d theta."""
        bl = _f64(bl)
        th = _f64(theta)
        fr = _f64(freqs)
        if bl.size != self.num_nodes or th.size != 5 or fr.size != 4:
            raise ValueError("gtr_eval: need num_nodes branch lengths, 5 exchangeabilities and 4 frequencies")
        ll = C.c_double()
        s = np.zeros(max(self.num_patterns, 1)) if site else None
        gb = np.zeros(self.num_nodes) if grad else None
        gt = np.zeros(5) if grad else None
        flags = HY_WANT_LOGL | (HY_WANT_SITE_LL if site else 0) | (HY_WANT_GRAD if grad else 0)
        self._check(self.L.hy_gtr_eval(self.h, precision, flags, _p(bl, _f64p), _p(th, _f64p), _p(fr, _f64p),
                                       C.byref(ll), _p(s, _f64p), _p(gb, _f64p), _p(gt, _f64p)))
        out = {"log_likelihood": ll.value}
        if site:
            out["site_ll"] = s[:self.num_patterns]
        if grad:
            out.update(grad_bl=gb, grad_theta=gt)
        return out

    def fel_eval(self, pattern, alpha, beta, theta, bl, delta=0.0, psi=0.0, scale=False):
        """FEL candidate log-likelihoods (hy_fel_eval): pattern[k] at
        (alpha[k], beta[k]), bl the converted branch lengths.  Returns the
        array of lnL, and with scale also each candidate generator's mean
        rate."""
        pat = np.ascontiguousarray(pattern, dtype=np.int32).ravel()
        a = _f64(alpha).ravel()
        b = _f64(beta).ravel()
        th = _f64(theta)
        blc = _f64(bl)
        n = pat.size
        if a.size != n or b.size != n or th.size != 5 or blc.size != self.num_nodes:
            raise ValueError("fel_eval: need one alpha and beta per pattern, 5 exchangeabilities and num_nodes lengths")
        out = np.zeros(max(n, 1))
        sc = np.zeros(max(n, 1)) if scale else None
        self._check(self.L.hy_fel_eval(self.h, n, _p(pat, _i32p), _p(a, _f64p), _p(b, _f64p), _p(th, _f64p),
                                       float(delta), float(psi), _p(blc, _f64p), _p(out, _f64p), _p(sc, _f64p)))
        return (out[:n], sc[:n]) if scale else out[:n]

    def fel_run(self, theta, alpha, beta, bl, delta=0.0, psi=0.0):
        """FEL refits of every pattern (hy_fel_run): bl the fitted tree's
        lengths.  Returns (array [num_patterns, HY_FEL_FIELDS], stats)."""
        th = _f64(theta)
        blc = _f64(bl)
        if th.size != 5 or blc.size != self.num_nodes:
            raise ValueError("fel_run: need 5 exchangeabilities and num_nodes branch lengths")
        out = np.zeros((max(self.num_patterns, 1), HY_FEL_FIELDS))
        calls, cands, qs = C.c_int(), C.c_int64(), C.c_double()
        self._check(self.L.hy_fel_run(self.h, _p(th, _f64p), float(alpha), float(beta), float(delta), float(psi),
                                      _p(blc, _f64p), _p(out, _f64p), C.byref(calls), C.byref(cands), C.byref(qs)))
        return out[:self.num_patterns], {"calls": calls.value, "candidates": cands.value, "q_scale": qs.value}

    def mg94_spectrum(self, alpha, beta, theta, delta=0.0, psi=0.0):
        """Eigenvalues (increasing, clamped to <= 0), U_pi, V_pi and the mean
        rate of one MG94 generator on the engine's model (hy_mg94_spectrum)."""
        S = self.S
        th = _f64(theta)
        lam = np.zeros(S)
        U = np.zeros((S, S))
        V = np.zeros((S, S))
        mu = C.c_double()
        self._check(self.L.hy_mg94_spectrum(self.h, float(alpha), float(beta), _p(th, _f64p), float(delta), float(psi),
                                            _p(lam, _f64p), _p(U, _f64p), _p(V, _f64p), C.byref(mu)))
        return lam, U, V, mu.value

    def set_pattern_weights(self, weights=None):
        """hy_set_pattern_weights: per-pattern weights of every later
        evaluation (None restores the alignment's)."""
        if weights is None:
            self._check(self.L.hy_set_pattern_weights(self.h, None))
            return
        w = _f64(weights)
        if w.shape != (self.num_patterns,):
            raise ValueError(f"need {self.num_patterns} pattern weights")
        self._check(self.L.hy_set_pattern_weights(self.h, _p(w, _f64p)))

    def branch_eval(self, node, P_cand, precision=HY_PREC_FP64, site=False):
        """aBSREL's local projection of the branch above ``node`` (hy_branch_eval):
        lnL for each candidate transition matrix in ``P_cand`` ([n, S, S] or
        [S, S]) against the partials of the last gradient evaluation in
        ``precision``.  With ``site``, also the per-pattern products L[n, P]
        and exponents e[P] (L_c,p = L[c, p] * 2**e[p])."""
        P = _f64(P_cand)
        if P.ndim == 2:
            P = P[None]
        S = self.S
        if P.shape[1:] != (S, S):
            raise ValueError(f"candidate matrices must be {S} x {S}")
        n = P.shape[0]
        lnl = np.zeros(n)
        L = np.zeros((n, self.num_patterns)) if site else None
        e = np.zeros(self.num_patterns, dtype=np.int32) if site else None
        self._check(self.L.hy_branch_eval(self.h, precision, int(node), n, _p(P, _f64p), _p(lnl, _f64p),
                                          _p(L, _f64p) if site else None, _p(e, _i32p) if site else None))
        return (lnl, L, e) if site else lnl

    def meme_eval(self, pattern, alpha, omega1, p1, beta_plus, theta, bl, beta_bg=None, is_test=None, delta=0.0,
                  psi=0.0, cond_node=None, cond_class=None):
        """hy_meme_eval: MEME's per-site lnL for each candidate
        (pattern[k], alpha[k], omega1[k], p1[k], beta_plus[k][, beta_bg[k]]);
        with cond_node[k] >= 0 that branch takes class cond_class[k] alone
        (0: omega1, 1: beta_plus)."""
        pat = np.ascontiguousarray(pattern, dtype=np.int32)
        n = pat.size
        arrs = [_f64(np.broadcast_to(np.asarray(x, dtype=np.float64), (n,))) for x in (alpha, omega1, p1, beta_plus)]
        bb = None if beta_bg is None else _f64(np.broadcast_to(np.asarray(beta_bg, dtype=np.float64), (n,)))
        it = None if is_test is None else np.ascontiguousarray(is_test, dtype=np.uint8)
        cn = cc = None
        if cond_node is not None:
            cn = np.ascontiguousarray(np.broadcast_to(np.asarray(cond_node), (n,)), dtype=np.int32)
            cc = np.ascontiguousarray(np.broadcast_to(np.asarray(cond_class if cond_class is not None else 0), (n,)),
                                      dtype=np.int32)
        out = np.zeros(n)
        self._check(self.L.hy_meme_eval(self.h, n, _p(pat, _i32p), *[_p(a, _f64p) for a in arrs], _p(bb, _f64p),
                                        _p(it, _u8p), _p(_f64(theta), _f64p), float(delta), float(psi),
                                        _p(_f64(bl), _f64p), _p(cn, _i32p), _p(cc, _i32p), _p(out, _f64p)))
        return out

    def meme_run(self, theta, alpha, beta, bl, delta=0.0, psi=0.0):
        """MEME on every pattern, every branch tested (hy_meme_run): bl the
        fitted tree's (unconverted) lengths, (alpha, beta, theta, delta, psi)
        the global MG94 fit.  Returns a dict: "sites" [P, HY_MEME_FIELDS]
        (columns MEME_FIELD_NAMES), "ebf" [P, num_nodes], "fel"
        [P, HY_FEL_FIELDS], "calls", "candidates"."""
        P, N = self.num_patterns, self.num_nodes
        out = np.zeros((max(P, 1), HY_MEME_FIELDS))
        ebf = np.ones((max(P, 1), N))
        fel = np.zeros((max(P, 1), HY_FEL_FIELDS))
        calls, cands = C.c_int(), C.c_int64()
        self._check(self.L.hy_meme_run(self.h, _p(_f64(theta), _f64p), float(alpha), float(beta), float(delta),
                                       float(psi), _p(_f64(bl), _f64p), _p(out, _f64p), _p(ebf, _f64p),
                                       _p(fel, _f64p), C.byref(calls), C.byref(cands)))
        return {"sites": out[:P], "ebf": ebf[:P], "fel": fel[:P], "calls": calls.value, "candidates": cands.value}

    def fel_counters(self, reset=False):
        """hy_fel_eval counters: candidates, pool hits, warm and cold solves."""
        out = (C.c_int64 * 4)()
        self._check(self.L.hy_fel_counters(self.h, out, 1 if reset else 0))
        return {"candidates": out[0], "pool_hits": out[1], "warm_solves": out[2], "cold_solves": out[3]}

    def inside_outside_from_P(self, P, pi, precision=HY_PREC_FP64):
        """hy_inside_outside_from_P: lnL from caller-supplied matrices, with
        the outside pass run so that the partials can be read back."""
        P = _f64(P)
        pi = _f64(pi)
        ll = C.c_double()
        self._check(self.L.hy_inside_outside_from_P(self.h, precision, _p(P, _f64p), _p(pi, _f64p), C.byref(ll),
                                                    None))
        return ll.value

    def loglik_from_P_noalloc(self, P, pi, precision=HY_PREC_FP64, site=False):
        """loglik_from_P through hy_loglik_from_P_noalloc: no allocation on
        the call (fixed per-device scratch; calls on one device must not
        overlap)."""
        P = _f64(P)
        pi = _f64(pi)
        if P.size != self.num_nodes * self.S * self.S or pi.size < self.S:
            raise ValueError("loglik_from_P_noalloc: P must hold num_nodes x S x S values and pi S values")
        ll = C.c_double()
        s = np.zeros(max(self.num_patterns, 1)) if site else None
        self._check(self.L.hy_loglik_from_P_noalloc(self.h, precision, _p(P, _f64p), _p(pi, _f64p), C.byref(ll),
                                                    _p(s, _f64p)))
        return (ll.value, s[:self.num_patterns]) if site else ll.value

    def autotune(self, bl, omegas, comp_w, theta, alpha=1.0, delta=0.0, psi=0.0, precision=HY_PREC_FP64, grad=True,
                 reps=5):
        """Times every tile width with and without subtree parts at these
        parameters and keeps the fastest for this precision.  Returns
        (tile, parts, ms per evaluation)."""
        bl = _f64(bl)
        om = _f64(np.atleast_1d(omegas))
        cw = _f64(np.atleast_1d(comp_w))
        th = _f64(theta)
        tile, parts, ms = C.c_int(), C.c_int(), C.c_double()
        flags = HY_WANT_LOGL | (HY_WANT_GRAD if grad else 0)
        self._check(self.L.hy_autotune(self.h, precision, flags, _p(bl, _f64p), alpha, _p(om, _f64p), _p(cw, _f64p),
                                       _p(th, _f64p), delta, psi, int(reps), C.byref(tile), C.byref(parts),
                                       C.byref(ms)))
        return tile.value, parts.value, ms.value

    def set_tuning(self, precision, tile, parts):
        """Applies an autotune result (tile 0/16/32/64, parts -1/0/1)."""
        self._check(self.L.hy_set_tuning(self.h, precision, int(tile), int(parts)))

    def set_warm_start(self, enable):
        self._check(self.L.hy_set_warm_start(self.h, 1 if enable else 0))

    def eigen_status(self):
        st = np.zeros(self.n_mat, dtype=np.int32)
        self._check(self.L.hy_last_eigen_status(self.h, _p(st, _i32p), self.n_mat))
        return st

    def last_timings(self):
        t = np.zeros(8)
        self._check(self.L.hy_last_timings(self.h, _p(t, _f64p)))
        return t

    def last_tile(self):
        t = C.c_int()
        self._check(self.L.hy_last_tile(self.h, C.byref(t)))
        return t.value

    def last_parts(self):
        """Subtree parts of the last evaluation's pruning passes (0: none)."""
        t = C.c_int()
        self._check(self.L.hy_last_parts(self.h, C.byref(t)))
        return t.value

    def set_memory_budget(self, nbytes):
        """Bytes allowed for the partials (0: the device's free memory); above
        it evaluations checkpoint and recompute inside partials."""
        self._check(self.L.hy_set_memory_budget(self.h, int(nbytes)))

    def last_checkpoint(self):
        """(parts recomputed by the last evaluation, 0 = all partials
        resident; bytes of its partial buffers)."""
        n, b = C.c_int(), C.c_size_t()
        self._check(self.L.hy_last_checkpoint(self.h, C.byref(n), C.byref(b)))
        return n.value, b.value

    def fit_lbfgs(self, bl, omegas, comp_w, theta, alpha=1.0, delta=0.0, psi=0.0, precision=HY_PREC_FP64,
                  captured=True, scale=1.0, m=8, max_iter=500, tol_grad=1e-8, tol_f=1e-12, check_every=4):
        """Maximum-likelihood fit of log bl (nodes with a parent), log omegas
        and log theta by hy_fit_lbfgs (objective -scale * lnL).  captured:
        every trial point after the start is one replay of a graph holding
        the evaluation and the optimizer step.  Returns a dict with the final
        bl, omegas, theta, logL, status, iterations, evaluations, replays."""
        b = _f64(bl).copy()
        om = _f64(omegas).copy()
        th = _f64(theta).copy()
        cw = _f64(comp_w)
        if b.shape != (self.num_nodes,) or om.shape != (self.n_mat,) or th.shape != (5,) or cw.shape != (self.n_comp,):
            raise ValueError("fit_lbfgs: bl, omegas, comp_w, theta must have num_nodes, n_mat, n_comp, 5 entries")
        ll = np.zeros(1)
        info = np.zeros(4, dtype=np.int32)
        self._check(self.L.hy_fit_lbfgs(self.h, int(precision), int(bool(captured)), _p(b, _f64p), float(alpha),
                                        _p(om, _f64p), _p(cw, _f64p), _p(th, _f64p), float(delta), float(psi),
                                        float(scale), int(m), int(max_iter), float(tol_grad), float(tol_f),
                                        int(check_every), _p(ll, _f64p), _p(info, _i32p)))
        return {"bl": b, "omegas": om, "theta": th, "logL": float(ll[0]), "status": int(info[0]),
                "iterations": int(info[1]), "evaluations": int(info[2]), "replays": int(info[3])}

    def device_bytes(self):
        v = C.c_size_t()
        self._check(self.L.hy_device_bytes(self.h, C.byref(v)))
        return v.value


class Lbfgs:
    """The engine's projected L-BFGS step on the host (hy_lbfgs_init /
    hy_lbfgs_step) for any objective: ``minimize(fun, x0, lb, ub)`` with
    ``fun(x) -> (f, g)``; the same step the captured fit takes on the GPU."""

    def __init__(self, x0, lb, ub, m=8, max_iter=500, tol_grad=1e-7, tol_f=0.0):
        self.L = load_library()
        x0, lb, ub = _f64(x0), _f64(lb), _f64(ub)
        self.n = x0.size
        self.state = np.zeros(self.L.hy_lbfgs_state_doubles(self.n, m))
        self.head = LbfgsHead()
        rc = self.L.hy_lbfgs_init(self.n, m, _p(self.state, _f64p), C.byref(self.head), _p(x0, _f64p),
                                  _p(lb, _f64p), _p(ub, _f64p), int(max_iter), float(tol_grad), float(tol_f))
        if rc != HY_OK:
            raise HyError(rc, "hy_lbfgs_init: invalid arguments")
        self.xt = np.clip(x0, lb, ub)

    def step(self, f, g, ok=True):
        g = _f64(g)
        xn = np.zeros(self.n)
        rc = self.L.hy_lbfgs_step(_p(self.state, _f64p), C.byref(self.head), float(f), _p(g, _f64p), int(bool(ok)),
                                  _p(xn, _f64p))
        if rc != HY_OK:
            raise HyError(rc, "hy_lbfgs_step: invalid state")
        self.xt = xn
        return xn

    @property
    def x(self):
        return self.state[:self.n].copy()

    def minimize(self, fun, max_evals=100000):
        while self.head.status == HY_LBFGS_RUNNING and self.head.n_eval < max_evals:
            f, g = fun(self.xt)
            self.step(f, g, np.isfinite(f))
        return self.x, self.head.f


def fel_drive(invariant, tested_length, evaluate):
    """FEL refits (hy_fel_drive) over a Python candidate evaluator:
    ``evaluate(patterns, alphas, betas)`` returns the candidates'
    log-likelihoods.  Returns (array [n_patterns, HY_FEL_FIELDS], calls,
    candidates)."""
    L = load_library()
    inv = np.ascontiguousarray(invariant, dtype=np.uint8)
    n = inv.size
    out = np.zeros((max(n, 1), HY_FEL_FIELDS))
    failure = []

    def cb(user, k, pat, a, b, ll):
        try:
            ps = np.ctypeslib.as_array(pat, shape=(k,)).copy()
            av = np.ctypeslib.as_array(a, shape=(k,)).copy()
            bv = np.ctypeslib.as_array(b, shape=(k,)).copy()
            np.ctypeslib.as_array(ll, shape=(k,))[:] = evaluate(ps, av, bv)
            return 0
        except Exception as ex:  # reported after the drive returns
            failure.append(ex)
            return -100

    fn = FEL_EVAL_FN(cb)
    calls, cands = C.c_int(), C.c_int64()
    rc = L.hy_fel_drive(n, _p(inv, _u8p), float(tested_length), fn, None, _p(out, _f64p), C.byref(calls),
                        C.byref(cands))
    if failure:
        raise failure[0]
    if rc != 0:
        raise HyError(rc, "hy_fel_drive failed")
    return out[:n], calls.value, cands.value


def meme_drive(invariant, fel_rows, num_nodes, root, evaluate):
    """MEME refits (hy_meme_drive) over a Python candidate evaluator:
    ``evaluate(patterns, alphas, omega1s, p1s, beta_pluses, cond_nodes,
    cond_classes)`` returns the candidates' log-likelihoods.  Returns
    (sites [n_patterns, HY_MEME_FIELDS], ebf [n_patterns, num_nodes], calls,
    candidates)."""
    L = load_library()
    inv = np.ascontiguousarray(invariant, dtype=np.uint8)
    n = inv.size
    fel = np.ascontiguousarray(fel_rows, dtype=np.float64).reshape(n, HY_FEL_FIELDS) if n else np.zeros((1, HY_FEL_FIELDS))
    out = np.zeros((max(n, 1), HY_MEME_FIELDS))
    ebf = np.ones((max(n, 1), int(num_nodes)))
    failure = []

    def cb(user, k, pat, a, w1, p1, bp, cn, cc, ll):
        try:
            cols = [np.ctypeslib.as_array(x, shape=(k,)).copy() for x in (pat, a, w1, p1, bp, cn, cc)]
            np.ctypeslib.as_array(ll, shape=(k,))[:] = evaluate(*cols)
            return 0
        except Exception as ex:  # reported after the drive returns
            failure.append(ex)
            return -100

    fn = MEME_EVAL_FN(cb)
    calls, cands = C.c_int(), C.c_int64()
    rc = L.hy_meme_drive(n, _p(inv, _u8p), _p(fel, _f64p), int(num_nodes), int(root), fn, None, _p(out, _f64p),
                         _p(ebf, _f64p), C.byref(calls), C.byref(cands))
    if failure:
        raise failure[0]
    if rc != 0:
        raise HyError(rc, "hy_meme_drive failed")
    return out[:n], ebf[:n], calls.value, cands.value


def graph_count():
    """(instantiated evaluation graphs alive in the process, their cap)."""
    L = load_library()
    live, cap = C.c_int(), C.c_int()
    L.hy_graph_count(C.byref(live), C.byref(cap))
    return live.value, cap.value


def create_group(n, device=0, n_stream_sets=8):
    """n engines sharing a few stream sets (hy_create_group), for batches of
    many small problems evaluated with eval_batch."""
    L = load_library()
    hs = (C.c_void_p * n)()
    rc = L.hy_create_group(device, n, n_stream_sets, hs)
    if rc != HY_OK:
        raise HyError(rc, L.hy_last_error(None).decode())
    return [Engine(device, handle=hs[i]) for i in range(n)]


class Batch:
    """Reusable hy_eval_batch call over a fixed list of engines.  Parameters
    and outputs live in flat arrays (gene g's slice at the offsets in
    ``off_bl``, ``off_om``, ``off_cw``; theta rows ``th[g]``) allocated once;
    ``eval`` takes per-gene tuples, ``eval_flat`` whole flat arrays.  Result
    dicts are views into the output arrays, overwritten by the next call."""

    def __init__(self, engines):
        self.L = load_library()
        self.engines = list(engines)
        n = len(self.engines)
        nn = [e.num_nodes for e in self.engines]
        nm = [e.n_mat for e in self.engines]
        nc = [e.n_comp for e in self.engines]
        self.off_bl = np.concatenate([[0], np.cumsum(nn)]).astype(np.int64)
        self.off_om = np.concatenate([[0], np.cumsum(nm)]).astype(np.int64)
        self.off_cw = np.concatenate([[0], np.cumsum(nc)]).astype(np.int64)
        self.bl = np.zeros(max(int(self.off_bl[-1]), 1))
        self.om = np.zeros(max(int(self.off_om[-1]), 1))
        self.cw = np.zeros(max(int(self.off_cw[-1]), 1))
        self.th = np.zeros((max(n, 1), 5))
        self.alpha = np.ones(max(n, 1))
        self.ll = np.zeros(max(n, 1))
        self.gbl = np.zeros_like(self.bl)
        self.gom = np.zeros_like(self.om)
        self.gw = np.zeros_like(self.cw)
        self.gth = np.zeros_like(self.th)
        self.site = [np.zeros(max(e.num_patterns, 1)) for e in self.engines]
        self.items = (BatchItem * max(n, 1))()
        d = C.sizeof(C.c_double)
        base = lambda arr: arr.ctypes.data  # noqa: E731
        self._ptr = lambda arr, off: C.cast(base(arr) + int(off) * d, _f64p)  # noqa: E731
        for g, e in enumerate(self.engines):
            it = self.items[g]
            it.h = e.h
            it.bl = self._ptr(self.bl, self.off_bl[g])
            it.omegas = self._ptr(self.om, self.off_om[g])
            it.comp_w = self._ptr(self.cw, self.off_cw[g])
            it.theta = self._ptr(self.th, 5 * g)
            it.delta, it.psi = 0.0, 0.0
            it.out_logL = self._ptr(self.ll, g)

    def _run(self, precision, grad, site):
        n = len(self.engines)
        for g in range(n):
            it = self.items[g]
            it.alpha = float(self.alpha[g])
            it.out_site_ll = _p(self.site[g], _f64p) if site else None
            if grad:
                it.out_grad_bl = self._ptr(self.gbl, self.off_bl[g])
                it.out_grad_omega = self._ptr(self.gom, self.off_om[g])
                it.out_grad_w = self._ptr(self.gw, self.off_cw[g])
                it.out_grad_theta = self._ptr(self.gth, 5 * g)
            else:
                it.out_grad_bl = it.out_grad_omega = it.out_grad_w = it.out_grad_theta = None
        flags = HY_WANT_LOGL | (HY_WANT_SITE_LL if site else 0) | (HY_WANT_GRAD if grad else 0)
        rc = self.L.hy_eval_batch(n, self.items, precision, flags)
        if rc != HY_OK:
            bad = next((i for i in range(n) if self.items[i].status != HY_OK), 0)
            raise HyError(rc, f"item {bad}: " + self.L.hy_last_error(self.engines[bad].h).decode())

    def autotune(self, precision=HY_PREC_FP64, grad=True, reps=3):
        """Times the fused batch at every tile width with the current flat
        parameters (set by ``eval``/``eval_flat`` or directly) and keeps the
        fastest for this precision.  Returns (tile, ms per batch)."""
        self._run(precision, grad, False)  # output pointers for the tuning calls
        n = len(self.engines)
        flags = HY_WANT_LOGL | (HY_WANT_GRAD if grad else 0)
        tile, ms = C.c_int(), C.c_double()
        rc = self.L.hy_autotune_batch(n, self.items, precision, flags, int(reps), C.byref(tile), C.byref(ms))
        if rc != HY_OK:
            raise HyError(rc, self.L.hy_last_error(self.engines[0].h).decode())
        return tile.value, ms.value

    def results(self, grad=True, site=False):
        out = []
        for g, e in enumerate(self.engines):
            r = {"log_likelihood": float(self.ll[g])}
            if site:
                r["site_ll"] = self.site[g][:e.num_patterns]
            if grad:
                r.update(grad_bl=self.gbl[self.off_bl[g]:self.off_bl[g + 1]],
                         grad_omega=self.gom[self.off_om[g]:self.off_om[g + 1]],
                         grad_w=self.gw[self.off_cw[g]:self.off_cw[g + 1]], grad_theta=self.gth[g])
            out.append(r)
        return out

    def eval(self, params, precision=HY_PREC_FP64, grad=True, site=False):
        """params[g] = (bl, omegas, comp_w, theta[, alpha])."""
        for g, prm in enumerate(params):
            self.bl[self.off_bl[g]:self.off_bl[g + 1]] = prm[0]
            self.om[self.off_om[g]:self.off_om[g + 1]] = prm[1]
            self.cw[self.off_cw[g]:self.off_cw[g + 1]] = prm[2]
            self.th[g] = prm[3]
            self.alpha[g] = float(prm[4]) if len(prm) > 4 else 1.0
        self._run(precision, grad, site)
        return self.results(grad, site)

    def eval_flat(self, bl, om, cw, th, precision=HY_PREC_FP64, grad=True, site=False):
        """Flat parameter arrays in this object's layout (alpha from
        ``self.alpha``)."""
        np.copyto(self.bl, bl)
        np.copyto(self.om, om)
        np.copyto(self.cw, cw)
        np.copyto(self.th, th)
        self._run(precision, grad, site)
        return self.results(grad, site)


def eval_batch(engines, params, precision=HY_PREC_FP64, grad=True, site=False):
    """Evaluates many engines (one gene each) in one hy_eval_batch call.
    params[i] = (bl, omegas, comp_w, theta[, alpha]); returns a list of result
    dicts shaped like Engine.eval's."""
    L = load_library()
    n = len(engines)
    items = (BatchItem * max(n, 1))()
    keep, outs = [], []
    flags = HY_WANT_LOGL | (HY_WANT_SITE_LL if site else 0) | (HY_WANT_GRAD if grad else 0)
    for i, (e, prm) in enumerate(zip(engines, params)):
        bl, om, cw, th = (_f64(np.atleast_1d(x)) for x in prm[:4])
        alpha = float(prm[4]) if len(prm) > 4 else 1.0
        ll = np.zeros(1)
        s = np.zeros(max(e.num_patterns, 1)) if site else None
        g = [np.zeros(k) if grad else None for k in (e.num_nodes, e.n_mat, e.n_comp, 5)]
        keep.append((bl, om, cw, th))
        outs.append((ll, s, g))
        it = items[i]
        it.h = e.h
        it.bl, it.omegas, it.comp_w, it.theta = (_p(x, _f64p) for x in (bl, om, cw, th))
        it.alpha, it.delta, it.psi = alpha, 0.0, 0.0
        it.out_logL = _p(ll, _f64p)
        it.out_site_ll = _p(s, _f64p)
        it.out_grad_bl, it.out_grad_omega, it.out_grad_w, it.out_grad_theta = (_p(x, _f64p) for x in g)
    rc = L.hy_eval_batch(n, items, precision, flags)
    if rc != HY_OK:
        bad = next(i for i in range(n) if items[i].status != HY_OK)
        raise HyError(rc, f"item {bad}: " + L.hy_last_error(engines[bad].h).decode())
    res = []
    for e, (ll, s, g) in zip(engines, outs):
        r = {"log_likelihood": float(ll[0])}
        if site:
            r["site_ll"] = s[:e.num_patterns]
        if grad:
            r.update(grad_bl=g[0], grad_omega=g[1], grad_w=g[2], grad_theta=g[3])
        res.append(r)
    return res
