"""Wheatstone-bridged trees on the MI355X engine (the hy_network_* C ABI).

A bridge at hub u with children {A, B, C} superposes u's three quartet
resolutions with weights w_r and bridge branch lengths t_r (tree.hpp:32-69,
likelihood.hpp:376-425, wheatstone_likelihood.hpp).  ``Network`` mirrors
``Engine``'s set-up calls and evaluates the network exactly as the
per-pattern mixture of the resolved binary trees, with gradients for the
branch lengths, the model parameters, and every bridge's lengths and
weights.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .engine import HY_PREC_FP64, HY_TIME_RAW, HY_WANT_GRAD, HY_WANT_LOGL, HY_WANT_SITE_LL, HyError, _f64, _p, \
    load_library

_i32p = C.POINTER(C.c_int32)
_i8p = C.POINTER(C.c_int8)
_u64p = C.POINTER(C.c_uint64)
_f64p = C.POINTER(C.c_double)

MAX_BRIDGES = 5


def _declare(L):
    if getattr(L, "_network_declared", False):
        return
    L.hy_network_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    L.hy_network_destroy.argtypes = [C.c_void_p]
    L.hy_network_destroy.restype = None
    L.hy_network_last_error.argtypes = [C.c_void_p]
    L.hy_network_last_error.restype = C.c_char_p
    L.hy_network_set_problem.argtypes = [C.c_void_p, C.c_int, _i32p, _i32p, _i32p, _i32p, C.c_int, C.c_int, _i8p,
                                         _u64p, _f64p, C.c_int, C.c_int, _i32p, _i32p]
    L.hy_network_set_model.argtypes = [C.c_void_p, C.c_int, _f64p, _f64p]
    L.hy_network_set_mixture.argtypes = [C.c_void_p, C.c_int, C.c_int, _i32p, _i32p, _i32p, C.c_int, _i32p]
    L.hy_network_set_sink.argtypes = [C.c_void_p, _f64p]
    L.hy_network_eval.argtypes = [C.c_void_p, C.c_int, C.c_uint, _f64p, _f64p, _f64p, C.c_double, _f64p, _f64p,
                                  _f64p, C.c_double, C.c_double, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p,
                                  _f64p]
    L._network_declared = True


class Network:
    """A tree with Wheatstone bridges.  bridges: list of (hub, (A, B, C))."""

    def __init__(self, device=0):
        self.L = load_library()
        _declare(self.L)
        h = C.c_void_p()
        rc = self.L.hy_network_create(device, C.byref(h))
        if rc != 0:
            raise HyError(rc, "hy_network_create failed")
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise HyError(rc, self.L.hy_network_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.hy_network_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_problem(self, parent, children, leaf_taxon, states, masks, weights, bridges, S=61):
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
        if len(bridges) > MAX_BRIDGES:
            raise ValueError(f"at most {MAX_BRIDGES} bridges")
        hub = np.array([b[0] for b in bridges] or [0], dtype=np.int32)
        kids = np.array([c for b in bridges for c in b[1]] or [0, 0, 0], dtype=np.int32)
        P = int(st.shape[0])
        ntax = int(st.shape[1]) if st.ndim == 2 else 0
        self._check(self.L.hy_network_set_problem(self.h, n, _p(self._parent, _i32p), _p(self._cp, _i32p),
                                                  _p(self._ci, _i32p), _p(self._l2t, _i32p), ntax, P, _p(st, _i8p),
                                                  _p(mk, _u64p), _p(w, _f64p), S, len(bridges), _p(hub, _i32p),
                                                  _p(kids, _i32p)))
        self.num_nodes, self.num_patterns, self.n_bridges = n, P, len(bridges)
        self.n_mat, self.n_comp = 1, 1

    def set_model(self, code_id, pos_freqs, pi):
        self._check(self.L.hy_network_set_model(self.h, int(code_id), _p(_f64(pos_freqs), _f64p),
                                                _p(_f64(pi), _f64p)))

    def set_mixture(self, n_mat, branch_class, classes, convention=HY_TIME_RAW, bridge_class=None):
        cp, cm = [0], []
        for comps in classes:
            cm.extend(int(m) for m in comps)
            cp.append(len(cm))
        cp = np.array(cp, dtype=np.int32)
        cm = np.array(cm, dtype=np.int32)
        bc = None if branch_class is None else np.ascontiguousarray(branch_class, dtype=np.int32)
        bbc = None if bridge_class is None else np.ascontiguousarray(bridge_class, dtype=np.int32)
        self._check(self.L.hy_network_set_mixture(self.h, n_mat, len(classes), _p(bc, _i32p), _p(cp, _i32p),
                                                  _p(cm, _i32p), convention, _p(bbc, _i32p)))
        self.n_mat, self.n_comp = n_mat, len(cm)

    def set_sink(self, freqs):
        self._check(self.L.hy_network_set_sink(self.h, _p(_f64(freqs), _f64p)))

    def eval(self, bl, bridge_weights, bridge_lengths, omegas, comp_w, theta, alpha=1.0, delta=0.0, psi=0.0,
             precision=HY_PREC_FP64, grad=True, site=False):
        """bridge_weights / bridge_lengths: [n_bridges, 3]."""
        bw = _f64(np.asarray(bridge_weights, dtype=np.float64).reshape(-1)) if self.n_bridges else _f64([0.0])
        blen = _f64(np.asarray(bridge_lengths, dtype=np.float64).reshape(-1)) if self.n_bridges else _f64([0.0])
        if self.n_bridges and (bw.size != 3 * self.n_bridges or blen.size != 3 * self.n_bridges):
            raise ValueError("three weights and three lengths per bridge")
        ll = C.c_double()
        s = np.zeros(max(self.num_patterns, 1)) if site else None
        outs = [np.zeros(self.num_nodes), np.zeros(self.n_mat), np.zeros(self.n_comp), np.zeros(5),
                np.zeros(max(3 * self.n_bridges, 1)), np.zeros(max(3 * self.n_bridges, 1))] if grad else [None] * 6
        flags = HY_WANT_LOGL | (HY_WANT_SITE_LL if site else 0) | (HY_WANT_GRAD if grad else 0)
        self._check(self.L.hy_network_eval(self.h, precision, flags, _p(_f64(bl), _f64p), _p(bw, _f64p),
                                           _p(blen, _f64p), float(alpha), _p(_f64(np.atleast_1d(omegas)), _f64p),
                                           _p(_f64(np.atleast_1d(comp_w)), _f64p), _p(_f64(theta), _f64p),
                                           float(delta), float(psi), C.byref(ll), _p(s, _f64p),
                                           *[_p(o, _f64p) for o in outs]))
        out = {"log_likelihood": ll.value}
        if site:
            out["site_ll"] = s[:self.num_patterns]
        if grad:
            nb = self.n_bridges
            out.update(grad_bl=outs[0], grad_omega=outs[1], grad_w=outs[2], grad_theta=outs[3],
                       grad_bridge_lengths=outs[4][:3 * nb].reshape(nb, 3),
                       grad_bridge_weights=outs[5][:3 * nb].reshape(nb, 3))
        return out
