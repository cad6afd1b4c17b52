"""The C++ facade's MG94Matrix spectrum (solved on the GPU through
hy_mg94_spectrum) and its process-wide engine:

  * eigenvalues (increasing, clamped to <= 0) against the oracle's
    eigensolver, U_pi V_pi = I, transition_matrix(t) against scipy's expm
    (at least as close as the oracle's own P(t), and within 1e-10 of it) over
    short and long times, also with multi-hit rates;
  * hyphy3.engine.Engine.mg94_spectrum on a model-only engine gives the same;
  * compute_inside_batch_in_place called from several host threads at once
    (the reference's OpenMP tile loop) returns what sequential calls return.
"""
import threading

import numpy as np
import pytest

from helpers import load_cd2_reduced
from oracle import hyoracle as ho

pytestmark = pytest.mark.gpu


def cd2_objects():
    import _hyphy3
    import os
    from helpers import GOLDEN
    text = open(os.path.join(GOLDEN, "cd2_reduced.fna")).read()
    aln = _hyphy3.Alignment.from_text(text)
    tree = _hyphy3.Tree.from_newick(open(os.path.join(GOLDEN, "cd2_reduced.nwk")).read(), True)
    return _hyphy3, tree, aln


@pytest.mark.parametrize("case", [(1.0, 0.4, 0.0, 0.0), (0.7, 3.0, 0.15, 0.03), (1.0, 1e-4, 0.0, 0.0)])
def test_facade_spectrum_against_oracle(case):
    h3, tree, aln = cd2_objects()
    rtree, raln = load_cd2_reduced()
    alpha, beta, delta, psi = case
    p = h3.MG94Parameters()
    p.alpha, p.beta, p.delta, p.psi = alpha, beta, delta, psi
    p.theta_AC, p.theta_AT, p.theta_CG, p.theta_CT, p.theta_GT = 0.6, 0.3, 0.5, 1.3, 0.4
    theta = [0.6, 0.3, 0.5, 1.3, 0.4]
    print(f"[spectrum] alpha {alpha} beta {beta} delta {delta} psi {psi}", flush=True)
    d = h3.mg94_matrix(aln, p)
    Q, mu = ho.mg94_Q(ho.Code(0), ho.params(alpha, beta, theta, delta, psi), raln.pos_freqs, raln.cf3x4)
    ev, U, V = ho.decompose(Q, raln.cf3x4)
    lam = np.asarray(d["eigenvalues"])
    assert np.all(np.diff(lam) >= 0) and lam.max() <= 0.0
    assert np.abs(lam - np.sort(ev)).max() < 1e-12 * np.abs(ev).max()
    assert np.abs(np.asarray(d["Q"]) - Q).max() < 1e-15 * np.abs(Q).max()
    assert abs(d["scale_factor"] - mu) < 1e-14 * mu
    Up, Vp = np.asarray(d["U_pi"]), np.asarray(d["V_pi"])
    assert np.abs(Up @ Vp - np.eye(len(lam))).max() < 1e-13
    import scipy.linalg as sl
    for t in (1e-5, 0.05, 0.7, 4.0):
        # against expm: at least as close as the oracle's own eigen route
        # (whose Jacobi vectors, scaled by 1/sqrt(pi) for rare codons, are off
        # by up to 8e-12 here at t = 1e-5)
        P = np.asarray(h3.mg94_transition_matrix(aln, p, t))
        Pr = ho.transition(ev, U, V, t)
        Pe = np.maximum(sl.expm(Q * t), 0.0)
        err_f, err_o = np.abs(P - Pe).max(), np.abs(Pr - Pe).max()
        print(f"[spectrum] t {t}: |P - expm| facade {err_f:.1e}, oracle {err_o:.1e}", flush=True)
        assert err_f <= max(1e-13, 2.0 * err_o), t
        assert np.abs(P - Pr).max() < 1e-10, t
    # the engine-level binding, on an engine that holds only the model
    from hyphy3.engine import Engine
    e = Engine(0)
    try:
        e.set_model(0, raln.pos_freqs, raln.cf3x4)
        l2, U2, V2, mu2 = e.mg94_spectrum(alpha, beta, theta, delta, psi)
        assert np.abs(l2 - lam).max() < 1e-12 * np.abs(lam).max()
        assert abs(mu2 - mu) < 1e-14 * mu
        assert np.abs(U2 @ V2 - np.eye(len(l2))).max() < 1e-13
    finally:
        e.close()


def test_facade_shared_engine_threads():
    h3, tree, aln = cd2_objects()
    rtree, raln = load_cd2_reduced()
    S = raln.code.S
    rng = np.random.default_rng(8)
    N = len(rtree.nodes)
    Q, _ = ho.mg94_Q(ho.Code(0), ho.params(1.0, 0.6, [0.8, 0.5, 0.7, 1.6, 0.6]), raln.pos_freqs, raln.cf3x4)
    ev, U, V = ho.decompose(Q, raln.cf3x4)
    P = np.stack([ho.transition(ev, U, V, float(t)) for t in rng.uniform(0.02, 0.4, size=N)])
    P_list = P.reshape(N, S * S)
    P_pass = P_list.reshape(N, S, S)
    n_pat = aln.num_patterns
    ranges = [(a, min(n_pat, a + 2)) for a in range(0, n_pat, 2)]
    seq = [h3.compute_inside_batch_in_place(tree, aln, a, b, P_pass, list(raln.cf3x4)) for a, b in ranges]
    out = [None] * len(ranges)

    def work(i):
        a, b = ranges[i]
        out[i] = h3.compute_inside_batch_in_place(tree, aln, a, b, P_pass, list(raln.cf3x4))

    print(f"[threads] {len(ranges)} tile ranges on 4 threads", flush=True)
    threads = [threading.Thread(target=lambda k=k: [work(i) for i in range(k, len(ranges), 4)]) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for s_, o in zip(seq, out):
        assert np.array_equal(np.asarray(s_["root_log_L"]), np.asarray(o["root_log_L"]))
        assert np.array_equal(np.asarray(s_["D"]), np.asarray(o["D"]))
