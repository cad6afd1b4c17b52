"""BUSTED expectation step on the MI355X (hy_busted_expected_weights) against
the oracle's literal restatement over the reference's workspaces
(busted.hpp:662-976, 978-1135): test/background branch classes, with and
without synonymous-rate classes, FP64 and FP32."""
import numpy as np
import pytest

from helpers import branch_lengths, make_engine, sim_problem
from oracle import hyoracle as ho

pytestmark = pytest.mark.gpu

OMEGAS, W, OMEGA_BG = [0.2, 1.0, 3.0], [0.5, 0.3, 0.2], 0.5


@pytest.mark.parametrize("seed,syn", [(81, None), (82, ([0.5, 1.8], [0.6, 0.4])), (83, ([0.3, 1.0, 2.5], [0.2, 0.5, 0.3]))])
def test_expected_weights_match_oracle(seed, syn):
    from hyphy3.engine import HY_PREC_FP32, HY_PREC_FP64, HY_TIME_RAW
    tree, aln, sim = sim_problem(14, 300, seed, omegas=OMEGAS, weights=W)
    N = len(tree.nodes)
    is_test = np.array([n.id != tree.root and n.id % 3 != 0 for n in tree.nodes])
    bc = np.where(is_test, 0, 1).astype(np.int32)
    tau = branch_lengths(tree) * 2.0
    sr, sv = syn if syn else ([1.0], [1.0])
    ref_w, ref_v = ho.busted_expected_weights(ho.Problem.from_refs(tree, aln), ho.Code(0),
                                              ho.params(1.0, 1.0, sim["theta"]), aln.pos_freqs, aln.cf3x4,
                                              aln.weights, OMEGAS, W, OMEGA_BG, is_test, tau, sr, sv)
    eng = make_engine(tree, aln)
    eng.set_mixture(4, bc, [[0, 1, 2], [3]], convention=HY_TIME_RAW)
    for prec, tol in ((HY_PREC_FP64, 1e-9), (HY_PREC_FP32, 1e-5)):
        nw, nv = eng.busted_expected_weights(tau, OMEGAS + [OMEGA_BG], W + [1.0], sim["theta"], sr, sv,
                                             test_class=0, precision=prec)
        print(seed, prec, nw, ref_w, nv, ref_v)
        np.testing.assert_allclose(nw, ref_w, rtol=tol)
        np.testing.assert_allclose(nv, ref_v, rtol=tol)
    eng.close()


def test_expected_weights_needs_raw_times():
    from hyphy3.engine import HY_TIME_CONV, HyError
    tree, aln, sim = sim_problem(6, 40, 84)
    N = len(tree.nodes)
    eng = make_engine(tree, aln)
    eng.set_mixture(2, np.zeros(N, dtype=np.int32), [[0, 1]], convention=HY_TIME_CONV)
    with pytest.raises(HyError):
        eng.busted_expected_weights(branch_lengths(tree), [0.3, 2.0], [0.5, 0.5], sim["theta"])
    eng.close()


@pytest.mark.parametrize("seed,syn", [(85, None), (86, ([0.5, 1.8], [0.6, 0.4]))])
def test_expected_weights_with_error_sink(seed, syn):
    """Test branches mix the omega components (weights (1 - eps) w_k) with
    the error sink (weight eps); the new eps (busted.hpp:846-860, 964-966)
    and omega weights against the oracle's literal restatement."""
    from hyphy3.engine import HY_PREC_FP32, HY_PREC_FP64, HY_SINK, HY_TIME_RAW
    tree, aln, sim = sim_problem(14, 300, seed, omegas=OMEGAS, weights=W)
    is_test = np.array([n.id != tree.root and n.id % 3 != 0 for n in tree.nodes])
    bc = np.where(is_test, 0, 1).astype(np.int32)
    tau = branch_lengths(tree) * 2.0
    sr, sv = syn if syn else ([1.0], [1.0])
    eps = 0.03
    ref_w, ref_v, ref_eps = ho.busted_expected_weights(ho.Problem.from_refs(tree, aln), ho.Code(0),
                                                       ho.params(1.0, 1.0, sim["theta"]), aln.pos_freqs, aln.cf3x4,
                                                       aln.weights, OMEGAS, W, OMEGA_BG, is_test, tau, sr, sv,
                                                       eps_sink=eps, max_eps_sink=0.05)
    eng = make_engine(tree, aln)
    eng.set_mixture(4, bc, [[0, 1, 2, HY_SINK], [3]], convention=HY_TIME_RAW)
    eng.set_sink(aln.cf3x4)
    comp_w = [(1 - eps) * x for x in W] + [eps, 1.0]
    for prec, tol in ((HY_PREC_FP64, 1e-9), (HY_PREC_FP32, 1e-5)):
        nw, nv = eng.busted_expected_weights(tau, OMEGAS + [OMEGA_BG], comp_w, sim["theta"], sr, sv, test_class=0,
                                             precision=prec, max_eps_sink=0.05)
        np.testing.assert_allclose(nw[:3], ref_w, rtol=tol)
        np.testing.assert_allclose(nw[3], ref_eps, rtol=tol)
        np.testing.assert_allclose(nv, ref_v, rtol=tol)
    eng.close()
