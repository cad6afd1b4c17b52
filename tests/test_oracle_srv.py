"""Pins of the oracle's synonymous-rate (SRV) classes, BUSTED's M > 1 path
(busted.hpp:509-659 forward, 1138-1429 gradients): reductions to the plain
mixture, a two-taxon closed form and central finite differences of every
returned gradient."""
import math

import numpy as np
import pytest

from helpers import branch_lengths, sim_problem
from oracle import hyoracle as ho

CODE = ho.Code(0)
OMEGAS, WEIGHTS = [0.1, 0.8, 3.0], [0.6, 0.35, 0.05]


def setup(n=6, L=60, seed=33, convention=1):
    tree, aln, sim = sim_problem(n, L, seed, omegas=OMEGAS, weights=WEIGHTS)
    prob = ho.Problem.from_refs(tree, aln)
    N = len(tree.nodes)
    bl = branch_lengths(tree) * (1.0 if convention else sim["conv"])
    th = np.asarray(sim["theta"], dtype=float)
    return tree, aln, prob, N, bl, th


def mix(N, om=OMEGAS, w=WEIGHTS, convention=1):
    return ho.Mixture(om, np.zeros(N, dtype=np.int32), [[(k, x) for k, x in enumerate(w)]], convention)


def test_single_class_equals_mixture():
    tree, aln, prob, N, bl, th = setup()
    base = ho.params(1.0, 1.0, th)
    ref = ho.mixture_eval(prob, CODE, base, aln.pos_freqs, aln.cf3x4, mix(N), bl)
    r = ho.srv_eval(prob, CODE, base, aln.pos_freqs, aln.cf3x4, mix(N), bl, [1.0], [1.0])
    assert abs(r["log_likelihood"] - ref["log_likelihood"]) < 1e-12 * abs(ref["log_likelihood"])
    for k in ("grad_bl", "grad_omega", "grad_w", "grad_theta"):
        assert np.allclose(r[k], ref[k], rtol=1e-11, atol=1e-11), k
    # d lnL / d syn_rate at rate 1 = sum_b bl_b dlnL/dbl_b (time scale identity)
    assert abs(r["grad_syn_rates"][0] - (bl * ref["grad_bl"]).sum()) < 1e-9 * (1 + abs(r["grad_syn_rates"][0]))
    # a zero-weight class is skipped and contributes nothing
    r2 = ho.srv_eval(prob, CODE, base, aln.pos_freqs, aln.cf3x4, mix(N), bl, [1.0, 2.0], [1.0, 0.0])
    assert r2["log_likelihood"] == r["log_likelihood"]
    assert r2["grad_syn_weights"][1] == 0.0 and r2["grad_syn_rates"][1] == 0.0


def test_equal_rates_reduce_to_scaled_branch_lengths():
    """All classes at the same rate s (weights summing to one): the mixture at
    branch lengths s * bl, with d/dbl scaled by s."""
    tree, aln, prob, N, bl, th = setup(convention=0)
    base = ho.params(1.0, 1.0, th)
    s = 1.3
    ref = ho.mixture_eval(prob, CODE, base, aln.pos_freqs, aln.cf3x4, mix(N, convention=0), bl * s)
    r = ho.srv_eval(prob, CODE, base, aln.pos_freqs, aln.cf3x4, mix(N, convention=0), bl, [s, s, s], [0.2, 0.5, 0.3])
    assert abs(r["log_likelihood"] - ref["log_likelihood"]) < 1e-12 * abs(ref["log_likelihood"])
    assert np.allclose(r["grad_bl"], s * ref["grad_bl"], rtol=1e-10, atol=1e-10)
    for k in ("grad_omega", "grad_w", "grad_theta"):
        assert np.allclose(r[k], ref[k], rtol=1e-10, atol=1e-10), k
    # posteriors equal the prior weights, so d lnL/d v_m = sum_p w_p
    assert np.allclose(r["grad_syn_weights"], np.sum(aln.weights), rtol=1e-12)


def test_two_taxon_closed_form():
    """lnL = log sum_m v_m pi_a P_ab(r_m (t1 + t2)) for one site."""
    _, _, _, _, _, th = setup()
    pos = np.full((3, 4), 0.25)
    pi = np.full(61, 1.0 / 61)
    p = ho.params(1.0, 0.6, th)
    Q, _ = ho.mg94_Q(CODE, p, pos, pi)
    rates, vw = [0.2, 1.0, 3.5], [0.25, 0.5, 0.25]
    m1 = ho.Mixture([0.6], np.zeros(3, dtype=np.int32), [[(0, 1.0)]], 0)
    for sa, sb in ((0, 0), (5, 17), (60, 2)):
        states = np.array([[sa, sb]], dtype=np.int8)
        masks = np.array([[1 << sa, 1 << sb]], dtype=np.uint64)
        prob = ho.Problem([-1, 0, 0], [[1, 2], [], []], [1, 2, 0], [0, 1, 1], [-1, 0, 1], states, masks, [1.0])
        r = ho.srv_eval(prob, CODE, ho.params(1.0, 1.0, th), pos, pi, m1, [0.0, 0.13, 0.29], rates, vw, want_grad=False)
        want = math.log(sum(v * pi[sa] * ho.expm_taylor(Q, s * 0.42)[sa, sb] for s, v in zip(rates, vw)))
        assert abs(r["log_likelihood"] - want) < 1e-11 * abs(want)


def fd(f, x, h):
    return (f(x + h) - f(x - h)) / (2 * h)


@pytest.mark.parametrize("convention", [0, 1])
def test_srv_gradients_fd(convention):
    tree, aln, prob, N, bl, th0 = setup(convention=convention)
    rates0, vw0 = np.array([0.3, 1.0, 2.5]), np.array([0.3, 0.5, 0.2])

    def run(om=OMEGAS, w=WEIGHTS, th=th0, b=bl, sr=rates0, sw=vw0, grad=False):
        return ho.srv_eval(prob, CODE, ho.params(1.0, 1.0, th), aln.pos_freqs, aln.cf3x4, mix(N, om, w, convention), b,
                           sr, sw, want_grad=grad)

    ref = run(grad=True)
    ll = lambda **kw: run(**kw)["log_likelihood"]  # noqa: E731
    for k in range(3):
        e = np.zeros(3)
        e[k] = 1
        num = fd(lambda x: ll(om=list(np.array(OMEGAS) + x * e)), 0.0, 1e-6)
        assert abs(ref["grad_omega"][k] - num) < 1e-5 * (1 + abs(num)), ("omega", k)
        num = fd(lambda x: ll(w=list(np.array(WEIGHTS) + x * e)), 0.0, 1e-6)
        assert abs(ref["grad_w"][k] - num) < 1e-5 * (1 + abs(num)), ("w", k)
        num = fd(lambda x: ll(sr=rates0 + x * e), 0.0, 1e-6)
        assert abs(ref["grad_syn_rates"][k] - num) < 1e-5 * (1 + abs(num)), ("syn_rate", k)
        num = fd(lambda x: ll(sw=vw0 + x * e), 0.0, 1e-6)
        assert abs(ref["grad_syn_weights"][k] - num) < 1e-5 * (1 + abs(num)), ("syn_weight", k)
    for k in range(5):
        e = np.zeros(5)
        e[k] = 1
        num = fd(lambda x: ll(th=th0 + x * e), 0.0, 1e-6)
        assert abs(ref["grad_theta"][k] - num) < 1e-5 * (1 + abs(num)), ("theta", k)
    for b in range(N):
        if b == tree.root:
            continue
        e = np.zeros(N)
        e[b] = 1
        num = fd(lambda x: ll(b=bl + x * e), 0.0, 1e-7)
        assert abs(ref["grad_bl"][b] - num) < 1e-4 * (1 + abs(num)), ("bl", b)
