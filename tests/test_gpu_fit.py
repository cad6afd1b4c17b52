"""Differentiable use from Python on the MI355X engine: the engine-backed
torch.autograd.Function against central finite differences (h = 1e-6, FP64),
and the torch L-BFGS maximum-likelihood fit (hyphy3.fit) against a reference
fit (scipy L-BFGS-B over the FP64 CPU oracle, the reference fitter's log
parameterisation and bounds, optimizer.hpp:934-1010)."""
import numpy as np
import pytest
import torch

from helpers import branch_lengths, oracle_fit, sim_problem

pytestmark = pytest.mark.gpu


def _engine_problem(sim, K):
    import hyphy3
    from hyphy3.fit import make_engine
    aln = hyphy3.Alignment.from_text(sim["fasta"])
    tree = hyphy3.Tree.from_newick(sim["newick"], keep_rooted=False)
    eng, arrays = make_engine(tree, aln, K)
    return eng, arrays


def test_autograd_matches_central_differences():
    from hyphy3.autograd import engine_log_likelihood
    from hyphy3.engine import HY_PREC_FP64
    omegas, weights = [0.3, 1.8], [0.65, 0.35]
    tree, aln, sim = sim_problem(10, 200, 41, omegas=omegas, weights=weights)
    eng, arrays = _engine_problem(sim, 2)
    bl = torch.tensor(branch_lengths(tree), dtype=torch.float64, requires_grad=True)
    om = torch.tensor(omegas, dtype=torch.float64, requires_grad=True)
    w = torch.tensor(weights, dtype=torch.float64, requires_grad=True)
    th = torch.tensor(sim["theta"], dtype=torch.float64, requires_grad=True)
    ll = engine_log_likelihood(bl, om, w, th, eng, HY_PREC_FP64)
    ll.backward()
    h = 1e-6

    def f(b, o, ww, t):
        return eng.eval(b, o, ww, t, precision=HY_PREC_FP64, grad=False)["log_likelihood"]

    base = [x.detach().numpy().copy() for x in (bl, om, w, th)]
    grads = [bl.grad.numpy(), om.grad.numpy(), w.grad.numpy(), th.grad.numpy()]
    parent = np.asarray(arrays["parent"])
    checked = 0
    for a, g in enumerate(grads):
        idx = range(len(g)) if a else [i for i in np.nonzero(parent >= 0)[0][:12]]
        for i in idx:
            up = [x.copy() for x in base]
            dn = [x.copy() for x in base]
            up[a][i] += h
            dn[a][i] -= h
            fd = (f(*up) - f(*dn)) / (2 * h)
            assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (a, i, fd, g[i])
            checked += 1
    assert checked >= 12 + 2 + 2 + 5
    eng.close()


def test_lbfgs_fit_matches_reference_fit():
    from hyphy3.fit import MixtureFit
    omegas, weights = [0.2, 1.5], [0.7, 0.3]
    tree, aln, sim = sim_problem(12, 400, 42, omegas=omegas, weights=weights)
    bl0 = branch_lengths(tree) * 1.5
    om0, th0 = [0.5, 2.0], np.ones(5)
    ll_ref, b_ref, om_ref, th_ref = oracle_fit(tree, aln, om0, weights, th0, bl0, threads=16)
    eng, arrays = _engine_problem(sim, 2)
    assert list(arrays["parent"]) == [n.parent for n in tree.nodes]
    fit = MixtureFit(eng, arrays["parent"], bl0, om0, weights, th0, n_sites=float(np.sum(arrays["weights"])))
    r = fit.run(max_iter=500, tol_grad=1e-9)
    print("GPU fit %.10f (FP32 its %d, FP64 its %d, %d evals, %.2f s); reference fit %.10f"
          % (r["log_likelihood"], r["iterations_fp32"], r["iterations_fp64"], r["evaluations"], r["seconds"], ll_ref))
    assert abs(r["log_likelihood"] - ll_ref) < 1e-4
    assert np.allclose(np.sort(r["omegas"]), np.sort(om_ref), rtol=2e-2)
    assert np.allclose(r["theta"], th_ref, rtol=2e-2)
    assert r["max_abs_log_gradient"] < 1e-2
    eng.close()
