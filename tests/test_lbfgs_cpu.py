"""The projected L-BFGS step shared by the host loop and the captured GPU fit
(hy_lbfgs_init / hy_lbfgs_step; the solver the reference's joint fit drives is
LBFGSpp's L-BFGS-B with m = 8, optimizer.hpp:1074-1087), run on the host
through the C ABI: known minima of quadratics (closed form, and the clamp of
the unconstrained minimiser for separable quadratics with bounds), the
Rosenbrock valley, scipy's L-BFGS-B on a coupled bounded quadratic, every
termination status, and the maximum-likelihood fit of the FP64 oracle's
likelihood against scipy's L-BFGS-B fit of the same objective."""
import numpy as np

from helpers import OracleMixtureEngine, branch_lengths, oracle_fit, sim_problem


def _lbfgs(*a, **k):
    from hyphy3.engine import Lbfgs
    return Lbfgs(*a, **k)


def test_quadratic_closed_form():
    rng = np.random.default_rng(1)
    n = 40
    Q = rng.normal(size=(n, n))
    A = Q @ Q.T + n * np.eye(n)
    c = rng.normal(size=n)
    opt = _lbfgs(np.zeros(n), np.full(n, -1e9), np.full(n, 1e9), tol_grad=1e-10, max_iter=500)
    x, f = opt.minimize(lambda x: (0.5 * (x - c) @ A @ (x - c), A @ (x - c)))
    from hyphy3.engine import HY_LBFGS_CONV_GRAD
    assert opt.head.status == HY_LBFGS_CONV_GRAD
    assert np.max(np.abs(x - c)) < 1e-10 * 1e2
    assert f < 1e-18 * n


def test_separable_quadratic_with_bounds_is_the_clamp():
    """f = sum a_i (x_i - c_i)^2 on a box: the minimiser is clip(c, lb, ub)
    exactly, with the bound-held coordinates on their bounds."""
    rng = np.random.default_rng(2)
    n = 300  # more than the 256 lanes of a reduction
    a = rng.uniform(0.5, 20.0, n)
    c = rng.uniform(-3, 3, n)
    lb, ub = np.full(n, -1.0), np.full(n, 1.5)
    opt = _lbfgs(rng.uniform(-1, 1.5, n), lb, ub, tol_grad=1e-12)
    x, _ = opt.minimize(lambda x: (np.sum(a * (x - c) ** 2), 2 * a * (x - c)))
    want = np.clip(c, lb, ub)
    assert np.max(np.abs(x - want)) < 1e-9
    held = (c < lb) | (c > ub)
    assert held.sum() > 50 and np.array_equal(x[held], want[held])


def test_rosenbrock():
    def rosen(x):
        f = np.sum(100.0 * (x[1:] - x[:-1] ** 2) ** 2 + (1 - x[:-1]) ** 2)
        g = np.zeros_like(x)
        g[:-1] = -400 * x[:-1] * (x[1:] - x[:-1] ** 2) - 2 * (1 - x[:-1])
        g[1:] += 200 * (x[1:] - x[:-1] ** 2)
        return f, g
    for n in (2, 10):
        x0 = np.full(n, -1.2)
        x0[1::2] = 1.0
        opt = _lbfgs(x0, np.full(n, -5.0), np.full(n, 5.0), tol_grad=1e-9, max_iter=2000)
        x, f = opt.minimize(rosen)
        assert np.max(np.abs(x - 1.0)) < 1e-6, (n, x)
        assert f < 1e-12


def test_coupled_bounded_quadratic_matches_scipy():
    from scipy.optimize import minimize
    rng = np.random.default_rng(3)
    n = 25
    Q = rng.normal(size=(n, n))
    A = Q @ Q.T / n + np.eye(n)
    b = rng.normal(size=n) * 3
    lb, ub = np.full(n, -0.5), np.full(n, 0.8)

    def fun(x):
        return 0.5 * x @ A @ x - b @ x, A @ x - b
    opt = _lbfgs(np.zeros(n), lb, ub, tol_grad=1e-11, max_iter=1000)
    x, f = opt.minimize(fun)
    ref = minimize(fun, np.zeros(n), jac=True, method="L-BFGS-B", bounds=list(zip(lb, ub)),
                   options={"ftol": 1e-16, "gtol": 1e-12, "maxiter": 5000})
    assert f <= ref.fun + 1e-12
    assert np.max(np.abs(x - ref.x)) < 1e-6
    # KKT: free coordinates have zero gradient, bound ones push outward
    g = A @ x - b
    free = (x > lb) & (x < ub)
    assert np.max(np.abs(g[free])) < 1e-9
    assert np.all(g[x <= lb] >= -1e-9) and np.all(g[x >= ub] <= 1e-9)


def test_termination_statuses():
    from hyphy3.engine import (HY_LBFGS_CONV_F, HY_LBFGS_LS_FAIL, HY_LBFGS_MAX_ITER, HY_LBFGS_NUMERIC,
                               HY_LBFGS_RUNNING)
    n = 5
    lb, ub = np.full(n, -10.0), np.full(n, 10.0)
    # non-finite start
    opt = _lbfgs(np.ones(n), lb, ub)
    opt.step(np.nan, np.zeros(n), ok=False)
    assert opt.head.status == HY_LBFGS_NUMERIC and opt.head.n_eval == 1
    # iteration cap
    opt = _lbfgs(np.full(n, 3.0), lb, ub, max_iter=2, tol_grad=0.0)
    opt.minimize(lambda x: (np.sum(np.cosh(x)), np.sinh(x)))
    assert opt.head.status == HY_LBFGS_MAX_ITER and opt.head.iter == 2
    # relative change of the objective
    opt = _lbfgs(np.full(n, 3.0), lb, ub, tol_f=1e-3, tol_grad=0.0)
    opt.minimize(lambda x: (np.sum(np.cosh(x)), np.sinh(x)))
    assert opt.head.status == HY_LBFGS_CONV_F
    # every trial fails: backtracking gives up after max_back halvings
    opt = _lbfgs(np.ones(n), lb, ub)
    opt.step(5.0, np.ones(n))
    assert opt.head.status == HY_LBFGS_RUNNING
    while opt.head.status == HY_LBFGS_RUNNING:
        opt.step(np.inf, np.ones(n), ok=False)
    assert opt.head.status == HY_LBFGS_LS_FAIL and opt.head.n_back == opt.head.max_back + 1
    assert np.array_equal(opt.x, np.ones(n))
    # after the end, further steps change nothing and are not counted
    ne = opt.head.n_eval
    xt = opt.step(0.0, np.zeros(n))
    assert opt.head.n_eval == ne and np.array_equal(xt, opt.x)


def test_exp_map_accuracy():
    """The log-parameter map's own exp (bit-identical on host and GPU) is
    within 2 ulp of the C library's over the box the fits use."""
    from hyphy3.engine import Lbfgs
    xs = np.linspace(-18.420680744, 6.91, 20001)
    # through the step: a first trial at x0 is evaluated as exp(x0) by the fits;
    # here the map is checked directly on the state vector's trial point
    opt = Lbfgs(xs, np.full(xs.size, -20.0), np.full(xs.size, 7.0))
    assert np.array_equal(opt.xt, xs)
    import hyphy3.engine as E
    L = E.load_library()
    got = np.array([L.hy_lbfgs_exp(float(x)) for x in xs])
    ulp = np.spacing(np.exp(xs))
    assert np.max(np.abs(got - np.exp(xs)) / ulp) <= 2.0


def test_oracle_likelihood_fit_matches_reference_fit():
    """The step driven by the FP64 oracle's lnL and gradient in the fit's log
    parameterisation and box (what hy_fit_lbfgs does with the engine)
    reaches scipy L-BFGS-B's maximum of the same objective."""
    omegas, weights = [0.2, 1.5], [0.7, 0.3]
    tree, aln, sim = sim_problem(6, 100, 33, omegas=omegas, weights=weights)
    bl0 = branch_lengths(tree) * 1.5
    om0, th0 = [0.5, 2.0], np.ones(5)
    ll_ref, b_ref, om_ref, th_ref = oracle_fit(tree, aln, om0, weights, th0, bl0)
    eng = OracleMixtureEngine(tree, aln)
    parent = np.array([n.parent for n in tree.nodes])
    br = np.nonzero(parent >= 0)[0]
    K = len(om0)
    scale = 1.0 / float(np.sum(aln.weights))
    x0 = np.concatenate([np.log(bl0[br]), np.log(om0), np.log(th0)])
    lb = np.array([-18.420680744] * len(br) + [-10.0] * (K + 5))
    ub = np.full(x0.size, 6.91)

    def fun(x):
        b = np.zeros(len(parent))
        b[br] = np.exp(x[:len(br)])
        om, th = np.exp(x[len(br):len(br) + K]), np.exp(x[len(br) + K:])
        r = eng.eval(b, om, weights, th)
        g = np.concatenate([r["grad_bl"][br] * b[br], r["grad_omega"] * om, r["grad_theta"] * th])
        return -scale * r["log_likelihood"], -scale * g
    opt = _lbfgs(x0, lb, ub, tol_grad=1e-8, tol_f=1e-12, max_iter=1000)
    x, f = opt.minimize(fun)
    ll = -f / scale
    print("host step fit %.10f in %d evaluations; reference fit %.10f" % (ll, opt.head.n_eval, ll_ref))
    assert abs(ll - ll_ref) < 1e-4
    assert np.allclose(np.sort(np.exp(x[len(br):len(br) + K])), np.sort(om_ref), rtol=2e-2)
