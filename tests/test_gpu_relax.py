"""RELAX branch gradients on the GPU (hyphy3.relax.branch_gradients) against
the oracle's restatement of compute_relax_branch_gradients (relax.hpp:
803-1166): branch-site level (the reference's default: one mixture
evaluation, partitions as branch classes, 3 / mu-bar convention) with and
without unclassified branches, site level with shared and per-class
normalisation (per-class evaluations, posterior-weighted gradients through
caller pattern weights), both precisions; the caller pattern weights are
cleared afterwards."""
import numpy as np
import pytest

from helpers import branch_lengths, grad_err, make_engine, rel, sim_problem
from oracle import hyoracle as ho
from oracle import relax_ref as rr

pytestmark = pytest.mark.gpu

FP64, FP32 = 0, 1


@pytest.mark.parametrize("case", ["branch_site", "branch_site_u", "site_shared", "site_per_class"])
def test_relax_matches_oracle(case):
    from hyphy3 import relax
    tree, aln, sim = sim_problem(16, 90, 81, gaps=0.1)
    prob = ho.Problem.from_refs(tree, aln)
    rng = np.random.default_rng(9)
    part = rng.integers(0, 2, len(tree.nodes)).astype(np.int32)
    bl = branch_lengths(tree)
    kw = dict(omegas_R=[0.1, 0.7, 2.5], weights=[0.5, 0.3, 0.2], k=0.6)
    okw = {}
    if case == "branch_site_u":
        part[[n.id for n in tree.nodes if n.is_leaf][:4]] = 2
        kw.update(omegas_U=[0.05, 0.9, 3.0], weights_U=[0.4, 0.4, 0.2], k=1.4)
    if case.startswith("site"):
        okw["branch_site"] = False
        kw["mixture_level"] = "site"
    if case == "site_per_class":
        okw["shared"] = False
        kw["normalization"] = "per_class"
    rkw = {k: v for k, v in kw.items() if k not in ("mixture_level", "normalization")}
    ref_l, ref_g = rr.relax_branch_gradients(prob, ho.Code(0), sim["theta"], aln.pos_freqs, aln.cf3x4, part, bl,
                                             **rkw, **okw)
    eng = make_engine(tree, aln)
    nr = [n.id for n in tree.nodes if n.id != tree.root]
    for prec, tl, tg in ((FP64, 1e-9, 1e-7), (FP32, 1e-5, 1e-4)):
        lnl, g = relax.branch_gradients(eng, part, bl, theta=sim["theta"], precision=prec, **kw)
        assert rel(lnl, ref_l) < tl, (case, prec, lnl, ref_l)
        assert grad_err(g[nr], ref_g[nr]) < tg, (case, prec)
    # the engine's own weights are back: a plain evaluation matches the oracle
    eng.set_mixture(1, None, [[0]])
    r = eng.eval(bl, [0.5], [1.0], sim["theta"], precision=FP64, grad=False)
    ll = ho.branch_length_gradients(prob, ho.Code(0), ho.params(1.0, 0.5, sim["theta"]), aln.pos_freqs, aln.cf3x4,
                                    bl)[0]
    assert rel(r["log_likelihood"], ll) < 1e-9
    eng.close()


def test_pattern_weights_scale_the_evaluation():
    tree, aln, sim = sim_problem(10, 60, 82)
    bl = branch_lengths(tree)
    eng = make_engine(tree, aln)
    base = eng.eval(bl, [0.5], [1.0], sim["theta"], precision=FP64, site=True)
    w = np.random.default_rng(3).uniform(0.0, 3.0, eng.num_patterns)
    eng.set_pattern_weights(w)
    r = eng.eval(bl, [0.5], [1.0], sim["theta"], precision=FP64)
    assert rel(r["log_likelihood"], float(np.dot(w, base["site_ll"]))) < 1e-12
    # gradient linear in the weights: weights w give sum_p w_p dlnL_p; check against doubling
    eng.set_pattern_weights(2 * w)
    r2 = eng.eval(bl, [0.5], [1.0], sim["theta"], precision=FP64)
    np.testing.assert_allclose(r2["grad_bl"], 2 * r["grad_bl"], rtol=1e-12, atol=1e-12 * np.abs(r["grad_bl"]).max())
    eng.set_pattern_weights(None)
    r3 = eng.eval(bl, [0.5], [1.0], sim["theta"], precision=FP64)
    assert rel(r3["log_likelihood"], base["log_likelihood"]) < 1e-14
    eng.close()
