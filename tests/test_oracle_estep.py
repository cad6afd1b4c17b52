"""Pins of the oracle's literal BUSTED expectation step
(busted.hpp:662-976, 978-1135; hyoracle.busted_expected_weights, built from
the reference's inside-outside workspaces): the accumulated component shares
equal w_k dlnL/dw_k of the oracle's mixture gradient and the synonymous
shares v_m dlnL/dv_m of its SRV gradient (two independent computations), the
unclamped shares sum to one (EM), and one component reduces to {1}."""
import numpy as np

from helpers import branch_lengths, sim_problem
from oracle import hyoracle as ho

OMEGAS, W, OMEGA_BG = [0.2, 1.0, 3.0], [0.5, 0.3, 0.2], 0.5


def _setup(seed):
    tree, aln, sim = sim_problem(8, 120, seed, omegas=OMEGAS, weights=W)
    N = len(tree.nodes)
    is_test = np.array([n.id != tree.root and n.id % 2 == 0 for n in tree.nodes])
    bc = np.where(is_test, 0, 1).astype(np.int32)
    tau = branch_lengths(tree) * 2.0
    prob = ho.Problem.from_refs(tree, aln)
    base = ho.params(1.0, 1.0, sim["theta"])
    mix = ho.Mixture(OMEGAS + [OMEGA_BG], bc, [[(k, W[k]) for k in range(3)], [(3, 1.0)]], 0)
    return tree, aln, prob, base, mix, is_test, tau, N


def _expected_from_gradients(aln, r, is_test, syn_weights=None):
    ok = r["site_ll"] > -1e19
    total = np.sum(aln.weights[ok])
    n_test = int(np.sum(is_test))
    num = np.array(W) * r["grad_w"][:3]
    x = np.clip(num / (n_test * total), 1e-6, 1 - 1e-6)
    out = [x / x.sum(), num.sum() / (n_test * total)]
    if syn_weights is not None:
        v = np.clip(np.array(syn_weights) * r["grad_syn_weights"] / total, 1e-6, 1 - 1e-6)
        out.append(v / v.sum())
    return out


def test_estep_equals_weight_gradient_identity():
    tree, aln, prob, base, mix, is_test, tau, N = _setup(71)
    code = ho.Code(0)
    new_w, new_v = ho.busted_expected_weights(prob, code, base, aln.pos_freqs, aln.cf3x4, aln.weights, OMEGAS, W,
                                              OMEGA_BG, is_test, tau)
    r = ho.mixture_eval(prob, code, base, aln.pos_freqs, aln.cf3x4, mix, tau)
    ref_w, em_sum = _expected_from_gradients(aln, r, is_test)
    np.testing.assert_allclose(new_w, ref_w, rtol=1e-10)
    assert abs(em_sum - 1.0) < 1e-10  # unclamped shares sum to one
    assert np.array_equal(new_v, [1.0])
    assert not np.allclose(new_w, W, rtol=1e-3)  # an actual update


def test_estep_with_synonymous_classes():
    tree, aln, prob, base, mix, is_test, tau, N = _setup(72)
    code = ho.Code(0)
    sr, sv = [0.5, 1.8], [0.6, 0.4]
    new_w, new_v = ho.busted_expected_weights(prob, code, base, aln.pos_freqs, aln.cf3x4, aln.weights, OMEGAS, W,
                                              OMEGA_BG, is_test, tau, sr, sv)
    r = ho.srv_eval(prob, code, base, aln.pos_freqs, aln.cf3x4, mix, tau, sr, sv)
    ref_w, em_sum, ref_v = _expected_from_gradients(aln, r, is_test, sv)
    np.testing.assert_allclose(new_w, ref_w, rtol=1e-10)
    np.testing.assert_allclose(new_v, ref_v, rtol=1e-10)
    assert abs(em_sum - 1.0) < 1e-10


def test_estep_single_component():
    tree, aln, prob, base, mix, is_test, tau, N = _setup(73)
    new_w, _ = ho.busted_expected_weights(prob, ho.Code(0), base, aln.pos_freqs, aln.cf3x4, aln.weights, [0.7], [1.0],
                                          OMEGA_BG, is_test, tau)
    assert np.array_equal(new_w, [1.0])


def test_estep_error_sink_update():
    """With the error sink (busted.hpp:846-860, 964-966): the literal sink
    share eps * sum post V . P_sink D / L equals eps dlnL/d eps of the
    oracle's mixture gradient (the sink component's weight gradient), the
    omega shares (1 - eps) w_k dlnL/dw_k, all unclamped shares sum to one,
    and the new eps is clamped to [1e-7, max_eps_sink]."""
    tree, aln, prob, base, _, is_test, tau, N = _setup(74)
    code = ho.Code(0)
    eps = 0.04
    bc = np.where(is_test, 0, 1).astype(np.int32)
    mix = ho.Mixture(OMEGAS + [OMEGA_BG], bc, [[(k, (1 - eps) * W[k]) for k in range(3)] + [(-1, eps)], [(3, 1.0)]],
                     0, sink_freqs=aln.cf3x4)
    for sr, sv in (([1.0], [1.0]), ([0.5, 1.8], [0.6, 0.4])):
        new_w, new_v, new_eps = ho.busted_expected_weights(prob, code, base, aln.pos_freqs, aln.cf3x4, aln.weights,
                                                           OMEGAS, W, OMEGA_BG, is_test, tau, sr, sv, eps_sink=eps,
                                                           max_eps_sink=0.5)
        r = ho.srv_eval(prob, code, base, aln.pos_freqs, aln.cf3x4, mix, tau, sr, sv)
        ok = r["site_ll"] > -1e19
        tot = int(np.sum(is_test)) * np.sum(aln.weights[ok])
        num = (1 - eps) * np.array(W) * r["grad_w"][:3]
        sink_share = eps * r["grad_w"][3]
        x = np.clip(num / tot, 1e-6, 1 - 1e-6)
        np.testing.assert_allclose(new_w, x / x.sum(), rtol=1e-10)
        assert abs(new_eps - sink_share / max(1.0, tot)) < 1e-10 * max(new_eps, 1e-12)
        assert abs((num.sum() + sink_share) / tot - 1.0) < 1e-10  # EM: shares sum to one
    _, _, capped = ho.busted_expected_weights(prob, code, base, aln.pos_freqs, aln.cf3x4, aln.weights, OMEGAS, W,
                                              OMEGA_BG, is_test, tau, eps_sink=eps, max_eps_sink=1e-5)
    assert capped == 1e-5
