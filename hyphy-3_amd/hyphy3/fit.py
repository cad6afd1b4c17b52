"""Maximum-likelihood fits of the branch-site omega mixture with torch's L-BFGS
over the MI355X engine.

Follows the joint phase of the reference's MG94Fitter
(include/hyphy/opt/optimizer.hpp:934-1010, :1070-1110): branch lengths, the
omegas and the five GTR exchangeabilities are optimised in log space with the
reference's box bounds (log bl in [-18.420680744, 6.91], log omega and
log theta in [-10, 6.91]) by a limited-memory quasi-Newton method with eight
correction pairs.  Here the objective is ``EngineLogLikelihood`` (a
``torch.autograd.Function`` whose backward returns the engine's adjoint
gradient) and the optimiser is ``torch.optim.LBFGS`` with a strong-Wolfe line
search; the bounds are enforced by clamping the log parameters (zero gradient
beyond a bound).  Branch lengths are in expected substitutions per codon
(tau = bl * 3 / mu_bar, likelihood.hpp:1246-1256).

The objective is the mean log-likelihood per site (lnL / n_sites), so the
gradient tolerance means the same at every alignment size.  A fit runs in two
phases: FP32-MFMA evaluations until L-BFGS stalls at the FP32 noise level of
the objective (changes below 1e-8 of its start value), then FP64 evaluations
from that point until the largest gradient component falls below the
tolerance, so the returned maximum is an FP64 one.  Mixture weights stay fixed
unless ``fit_weights`` (then they are a softmax of free logits).
"""
import time

import numpy as np
import torch

import hyphy3
from hyphy3.autograd import engine_log_likelihood
from hyphy3.engine import HY_PREC_FP32, HY_PREC_FP64, HY_TIME_CONV, Engine

LOG_BL_BOUNDS = (-18.420680744, 6.91)  # optimizer.hpp:966-967
LOG_RATE_BOUNDS = (-10.0, 6.91)  # optimizer.hpp:971-972, 986-987


def make_engine(tree, aln, n_omega, device=0, engine=None):
    """Engine holding ``tree``/``aln`` (hyphy3.Tree, hyphy3.Alignment) with one
    branch class mixing ``n_omega`` omega components on every branch, branch
    lengths in expected substitutions (3 / mu_bar convention)."""
    arrays = hyphy3.problem_arrays(tree, aln)
    eng = engine or Engine(device)
    n_nodes = len(arrays["parent"])
    cp, ci = arrays["child_ptr"], arrays["child_idx"]
    children = [list(ci[cp[i]:cp[i + 1]]) for i in range(n_nodes)]
    eng.set_problem(arrays["parent"], children, arrays["leaf_taxon"], arrays["states"], arrays["masks"],
                    arrays["weights"], S=int(arrays["S"]))
    eng.set_model(int(arrays["code_id"]), arrays["pos_freqs"], arrays["pi"])
    eng.set_mixture(n_omega, np.zeros(n_nodes, dtype=np.int32), [list(range(n_omega))], convention=HY_TIME_CONV)
    return eng, arrays


class MixtureFit:
    """L-BFGS fit of (branch lengths, omegas, theta[, weights]) on one engine.

    ``engine`` is any object with the ``Engine.eval`` contract: set up for
    ``len(omegas)`` components in one class on every branch, time convention
    3 / mu_bar.  ``parent`` (per node, -1 at the root) fixes which entries
    of the per-node branch-length vector are free.
    """

    def __init__(self, engine, parent, bl, omegas, weights, theta, n_sites=1.0, fit_weights=False, history_size=8):
        self.engine = engine
        self.scale = 1.0 / float(n_sites)
        parent = np.asarray(parent)
        self.n_nodes = len(parent)
        self.branches = torch.as_tensor(np.nonzero(parent >= 0)[0], dtype=torch.long)
        self.B = len(self.branches)
        self.K = len(omegas)
        self.fit_weights = bool(fit_weights)
        self.history_size = history_size
        bl = np.asarray(bl, dtype=np.float64)
        x = np.concatenate([np.log(np.clip(bl[self.branches.numpy()], 1e-7, 100.0)),
                            np.log(np.clip(np.asarray(omegas, dtype=np.float64), 1e-4, 100.0)),
                            np.log(np.clip(np.asarray(theta, dtype=np.float64), 1e-4, 100.0))])
        self.x = torch.tensor(x, dtype=torch.float64, requires_grad=True)
        n_rate = self.K + 5
        self.lb = torch.tensor([LOG_BL_BOUNDS[0]] * self.B + [LOG_RATE_BOUNDS[0]] * n_rate, dtype=torch.float64)
        self.ub = torch.tensor([LOG_BL_BOUNDS[1]] * self.B + [LOG_RATE_BOUNDS[1]] * n_rate, dtype=torch.float64)
        w = np.asarray(weights, dtype=np.float64)
        self.w_fixed = torch.tensor(w / w.sum(), dtype=torch.float64)
        self.logits = torch.tensor(np.log(np.maximum(w / w.sum(), 1e-300)), dtype=torch.float64,
                                   requires_grad=self.fit_weights)
        self.evaluations = 0
        self.trace = []

    def parameters(self):
        """Current (bl per node, omegas, weights, theta) as numpy arrays."""
        with torch.no_grad():
            bl, om, w, th = self._unpack()
        return bl.numpy(), om.numpy(), w.numpy(), th.numpy()

    def _unpack(self):
        y = torch.minimum(torch.maximum(self.x, self.lb), self.ub)
        bl = torch.zeros(self.n_nodes, dtype=torch.float64).index_put((self.branches,), torch.exp(y[:self.B]))
        om = torch.exp(y[self.B:self.B + self.K])
        th = torch.exp(y[self.B + self.K:])
        w = torch.softmax(self.logits, 0) if self.fit_weights else self.w_fixed
        return bl, om, w, th

    def _phase(self, precision, max_iter, tol_grad, tol_change):
        params = [self.x] + ([self.logits] if self.fit_weights else [])
        opt = torch.optim.LBFGS(params, lr=1.0, max_iter=max_iter, max_eval=int(max_iter * 1.5) + 10,
                                history_size=self.history_size, tolerance_grad=tol_grad, tolerance_change=tol_change,
                                line_search_fn="strong_wolfe")
        state = {"ll": None}

        def closure():
            opt.zero_grad()
            bl, om, w, th = self._unpack()
            ll = engine_log_likelihood(bl, om, w, th, self.engine, precision)
            loss = -ll * self.scale
            loss.backward()
            self.evaluations += 1
            state["ll"] = float(ll.detach())
            self.trace.append((precision, state["ll"]))
            return loss

        opt.step(closure)
        st = opt.state[opt._params[0]]
        return int(st.get("n_iter", 0)), state["ll"]

    def run(self, max_iter=500, tol_grad=1e-7, fp32_first=True):
        """Fits and returns a dict: log_likelihood (FP64, at the returned
        parameters), bl, omegas, weights, theta, iterations per phase,
        evaluations, seconds.  ``tol_grad`` bounds the largest gradient
        component of lnL / n_sites with respect to the log parameters."""
        t0 = time.perf_counter()
        with torch.no_grad():
            bl, om, w, th = self._unpack()
        f0 = abs(self.engine.eval(bl.numpy(), om.numpy(), w.numpy(), th.numpy(), precision=HY_PREC_FP64,
                                  grad=False)["log_likelihood"]) * self.scale
        it32 = 0
        if fp32_first:
            it32, _ = self._phase(HY_PREC_FP32, max_iter, tol_grad, 1e-8 * f0)
        it64, _ = self._phase(HY_PREC_FP64, max_iter, tol_grad, 1e-14 * f0)
        bl, om, w, th = self.parameters()
        final = self.engine.eval(bl, om, w, th, precision=HY_PREC_FP64)
        g = np.concatenate([final["grad_bl"][self.branches.numpy()] * bl[self.branches.numpy()],
                            final["grad_omega"] * om, final["grad_theta"] * th])
        # projected on the box: components pushing a parameter beyond the bound it sits at count as zero
        x = self.x.detach().numpy()
        lb, ub = self.lb.numpy(), self.ub.numpy()
        g = np.where(((x <= lb) & (g < 0)) | ((x >= ub) & (g > 0)), 0.0, g)
        return {"log_likelihood": final["log_likelihood"], "bl": bl, "omegas": om, "weights": w, "theta": th,
                "iterations_fp32": it32, "iterations_fp64": it64, "evaluations": self.evaluations,
                "max_abs_log_gradient": float(np.max(np.abs(g))) if len(g) else 0.0,
                "max_abs_log_gradient_per_site": float(np.max(np.abs(g))) * self.scale if len(g) else 0.0,
                "seconds": time.perf_counter() - t0}


def fit_mixture(tree, aln, omegas, weights, theta, bl=None, fit_weights=False, max_iter=500, tol_grad=1e-7,
                device=0):
    """Fits the branch-site omega mixture to ``aln`` on ``tree`` (branch
    lengths from the tree unless ``bl`` is given, one per node)."""
    eng, arrays = make_engine(tree, aln, len(omegas), device)
    try:
        if bl is None:
            bl = np.array([n.branch_length for n in tree.nodes], dtype=np.float64)
        fit = MixtureFit(eng, arrays["parent"], bl, omegas, weights, theta, n_sites=float(np.sum(arrays["weights"])),
                         fit_weights=fit_weights)
        return fit.run(max_iter=max_iter, tol_grad=tol_grad)
    finally:
        eng.close()
