"""Maximum-likelihood fits whose optimizer iterations are captured graphs
(hy_fit_lbfgs): every trial point after the start is one replay of a graph
holding the evaluation (logL + full gradient) and the optimizer step on the
GPU.  Checked against the same fit stepped on the host (hy_eval +
hy_lbfgs_step: bit-identical iterates, the step being one source compiled for
both), against the FP64 oracle at the fitted point, and against a reference
fit (scipy L-BFGS-B over the FP64 oracle with the reference fitter's log
parameterisation and bounds, optimizer.hpp:934-1010)."""
import numpy as np
import pytest

from helpers import OracleMixtureEngine, branch_lengths, oracle_fit, sim_problem

pytestmark = pytest.mark.gpu


def _engine(sim, K):
    import hyphy3
    from hyphy3.fit import make_engine
    aln = hyphy3.Alignment.from_text(sim["fasta"])
    tree = hyphy3.Tree.from_newick(sim["newick"], keep_rooted=False)
    return make_engine(tree, aln, K)


@pytest.mark.parametrize("precision", [0, 1])
def test_captured_fit_equals_host_stepped_fit(precision):
    omegas, weights = [0.2, 1.5], [0.7, 0.3]
    tree, aln, sim = sim_problem(12, 400, 42, omegas=omegas, weights=weights)
    eng, arrays = _engine(sim, 2)
    bl0 = branch_lengths(tree) * 1.5
    scale = 1.0 / float(np.sum(arrays["weights"]))
    kw = dict(scale=scale, max_iter=300, tol_grad=1e-8, tol_f=1e-12 if precision == 0 else 1e-9, precision=precision)
    print("case: precision", precision, "12 taxa x 400 codons, K=2", flush=True)
    host = eng.fit_lbfgs(bl0, [0.5, 2.0], weights, np.ones(5), captured=False, **kw)
    cap = eng.fit_lbfgs(bl0, [0.5, 2.0], weights, np.ones(5), captured=True, check_every=4, **kw)
    print("host", host["logL"], host["iterations"], host["evaluations"], host["status"])
    print("captured", cap["logL"], cap["iterations"], cap["evaluations"], cap["status"], cap["replays"])
    assert cap["status"] == host["status"] and cap["status"] in (1, 2)
    assert cap["iterations"] == host["iterations"] and cap["evaluations"] == host["evaluations"]
    assert cap["replays"] >= cap["evaluations"] - 1
    for k in ("bl", "omegas", "theta"):
        assert np.array_equal(cap[k], host[k]), k
    assert cap["logL"] == host["logL"]
    # replays past the end leave the result unchanged
    cap8 = eng.fit_lbfgs(bl0, [0.5, 2.0], weights, np.ones(5), captured=True, check_every=8, **kw)
    assert np.array_equal(cap8["bl"], cap["bl"]) and cap8["logL"] == cap["logL"]
    eng.close()


def test_captured_fit_matches_reference_fit():
    omegas, weights = [0.2, 1.5], [0.7, 0.3]
    tree, aln, sim = sim_problem(12, 400, 42, omegas=omegas, weights=weights)
    bl0 = branch_lengths(tree) * 1.5
    om0, th0 = [0.5, 2.0], np.ones(5)
    ll_ref, b_ref, om_ref, th_ref = oracle_fit(tree, aln, om0, weights, th0, bl0, threads=16)
    eng, arrays = _engine(sim, 2)
    assert list(arrays["parent"]) == [n.parent for n in tree.nodes]
    scale = 1.0 / float(np.sum(arrays["weights"]))
    # FP32 iterations until the objective stalls, then FP64 from there
    r32 = eng.fit_lbfgs(bl0, om0, weights, th0, precision=1, scale=scale, tol_grad=1e-8, tol_f=1e-9)
    r = eng.fit_lbfgs(r32["bl"], r32["omegas"], weights, r32["theta"], precision=0, scale=scale, tol_grad=1e-8,
                      tol_f=1e-12)
    print("captured fit %.10f (FP32 its %d, FP64 its %d); reference fit %.10f"
          % (r["logL"], r32["iterations"], r["iterations"], ll_ref))
    assert abs(r["logL"] - ll_ref) < 1e-4
    assert np.allclose(np.sort(r["omegas"]), np.sort(om_ref), rtol=2e-2)
    assert np.allclose(r["theta"], th_ref, rtol=2e-2)
    # the reported maximum is the FP64 oracle's lnL at the returned point
    ref = OracleMixtureEngine(tree, aln, threads=16).eval(r["bl"], r["omegas"], weights, r["theta"], grad=False)
    assert abs(r["logL"] - ref["log_likelihood"]) <= 1e-9 * abs(ref["log_likelihood"])
    parent = np.asarray(arrays["parent"])
    assert np.array_equal(r["bl"][parent < 0], bl0[parent < 0])
    eng.close()


def test_captured_fit_argument_checks():
    from hyphy3.engine import HyError
    tree, aln, sim = sim_problem(6, 50, 43)
    eng, arrays = _engine(sim, 1)
    bl = branch_lengths(tree)
    with pytest.raises(HyError):
        eng.fit_lbfgs(bl, [0.5], [1.0], np.ones(5), m=0)
    bad = bl.copy()
    bad[np.nonzero(np.asarray(arrays["parent"]) >= 0)[0][0]] = -1.0
    with pytest.raises(HyError):
        eng.fit_lbfgs(bad, [0.5], [1.0], np.ones(5))
    eng.set_pattern_weights(np.ones(len(arrays["weights"])))
    with pytest.raises(HyError):
        eng.fit_lbfgs(bl, [0.5], [1.0], np.ones(5))
    eng.set_pattern_weights(None)
    r = eng.fit_lbfgs(bl, [0.5], [1.0], np.ones(5), max_iter=3, scale=1e-2)
    assert r["status"] in (1, 2, 3) and r["iterations"] <= 3
    eng.close()


def test_group_engine_fit_steps_on_the_host():
    """Members of an engine group launch eagerly: hy_fit_lbfgs steps them on
    the host (no replays) with the iterates of a standalone engine."""
    from hyphy3.engine import create_group
    from hyphy3.fit import make_engine
    import hyphy3
    omegas, weights = [0.2, 1.5], [0.7, 0.3]
    tree, aln, sim = sim_problem(10, 200, 44, omegas=omegas, weights=weights)
    a = hyphy3.Alignment.from_text(sim["fasta"])
    t = hyphy3.Tree.from_newick(sim["newick"], keep_rooted=False)
    grp = create_group(2)
    eng_g, arrays = make_engine(t, a, 2, engine=grp[0])
    eng_s, _ = make_engine(t, a, 2)
    bl0 = branch_lengths(tree) * 1.5
    kw = dict(scale=1.0 / float(np.sum(arrays["weights"])), max_iter=100, tol_grad=1e-8, tol_f=1e-12, precision=0)
    rg = eng_g.fit_lbfgs(bl0, [0.5, 2.0], weights, np.ones(5), captured=True, **kw)
    rs = eng_s.fit_lbfgs(bl0, [0.5, 2.0], weights, np.ones(5), captured=True, **kw)
    print("group", rg["logL"], rg["iterations"], rg["replays"], "standalone", rs["logL"], rs["iterations"], rs["replays"])
    assert rg["replays"] == 0 and rs["replays"] > 0
    assert rg["status"] == rs["status"] and rg["iterations"] == rs["iterations"]
    assert abs(rg["logL"] - rs["logL"]) <= 1e-9 * abs(rs["logL"])
    for e in list(grp) + [eng_s]:
        e.close()
