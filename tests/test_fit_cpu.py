"""Host logic of the L-BFGS mixture fit (hyphy3.fit.MixtureFit: log
parameterisation, bounds, autograd chain through the per-node branch-length
vector, FP32 then FP64 phases) driven by the FP64 CPU oracle in place of the
engine, against scipy's L-BFGS-B fit of the same objective (the reference
fitter's parameterisation, optimizer.hpp:934-1010)."""
import numpy as np

from helpers import OracleMixtureEngine, branch_lengths, oracle_fit, sim_problem


def test_fit_logic_matches_reference_fit():
    from hyphy3.fit import MixtureFit
    omegas, weights = [0.2, 1.5], [0.7, 0.3]
    tree, aln, sim = sim_problem(8, 150, 31, omegas=omegas, weights=weights)
    bl0 = branch_lengths(tree) * 1.5  # start away from the generating lengths
    om0, th0 = [0.5, 2.0], np.ones(5)
    ll_ref, b_ref, om_ref, th_ref = oracle_fit(tree, aln, om0, weights, th0, bl0)
    eng = OracleMixtureEngine(tree, aln)
    fit = MixtureFit(eng, [n.parent for n in tree.nodes], bl0, om0, weights, th0, n_sites=float(np.sum(aln.weights)))
    r = fit.run(max_iter=400, tol_grad=1e-9)
    print("fit", r["log_likelihood"], "reference fit", ll_ref, "evals", r["evaluations"])
    assert abs(r["log_likelihood"] - ll_ref) < 1e-4
    assert r["log_likelihood"] > eng.eval(bl0, om0, weights, th0)["log_likelihood"] + 1.0
    assert np.allclose(np.sort(r["omegas"]), np.sort(om_ref), rtol=2e-2)
    assert np.allclose(r["theta"], th_ref, rtol=2e-2)
    parent = np.array([n.parent for n in tree.nodes])
    assert r["bl"][parent < 0].sum() == 0.0
    assert r["max_abs_log_gradient"] < 1e-2


def test_fit_respects_bounds():
    """A parameter pushed against its bound stays at the bound (clamped log
    parameters have zero gradient beyond it)."""
    from hyphy3.fit import LOG_RATE_BOUNDS, MixtureFit
    tree, aln, sim = sim_problem(6, 60, 32)
    eng = OracleMixtureEngine(tree, aln)
    fit = MixtureFit(eng, [n.parent for n in tree.nodes], branch_lengths(tree), [1e-6], [1.0], np.ones(5))
    bl, om, w, th = fit.parameters()
    assert om[0] >= np.exp(LOG_RATE_BOUNDS[0]) * (1 - 1e-12)
    r = fit.run(max_iter=100, tol_grad=1e-8)
    assert np.all(np.log(r["omegas"]) >= LOG_RATE_BOUNDS[0] - 1e-9)
    assert np.all(np.log(r["theta"]) <= LOG_RATE_BOUNDS[1] + 1e-9)
