"""Parity at the benchmark workloads themselves (GPU): every configuration the
throughput runs use, built exactly as bench.py builds it, evaluated by the
MI355X engine in FP64 and FP32 and by the FP64 CPU oracle on the same seeded
tree and alignment.

  c1  16 taxa x 500 codons, one omega
  c2  64 taxa x 5,000 codons, 3 omega components (both time conventions)
  c4  512 taxa x 20,000 codons, 4 components, 25 % gaps
  c5  256 taxa x 10,000 codons, 3 components, plus central differences
      (h = 1e-6, FP64) of the largest gradient components
  c3  1,024 genes in one fused batch: every 32nd gene against the oracle,
      every gene's FP32 result against its FP64 one

The product's data layer (pattern compression, masks, CF3x4) is compared
bit-exactly with the oracle's restatement on each workload first.

Tolerances (stated per test): FP64 logL 1e-9 relative, gradients 1e-7;
FP32 logL 1e-5, gradients 1e-4 (component-wise, floor 1e-2 * max |g|)."""
import functools
import os
import sys

import numpy as np
import pytest

from helpers import grad_err, rel, ROOT

pytestmark = pytest.mark.gpu
sys.path.insert(0, ROOT)
import bench  # noqa: E402

FP64, FP32 = 0, 1
KEYS = ("grad_bl", "grad_omega", "grad_w", "grad_theta")


def host_threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        n = min(n, int(os.environ["OMP_NUM_THREADS"]))
    return max(1, min(n, 16))


@functools.lru_cache(maxsize=None)
def workload(name):
    """(cfg, sim, arrays, tau, oracle problem, oracle alignment, oracle tree)
    of a benchmark configuration; the product's problem arrays must equal the
    restatement's bit for bit."""
    from oracle import data_ref as dr
    from oracle import hyoracle as ho
    cfg = bench.CONFIGS[name]
    print(f"[configs] building {name}: {cfg['taxa']} taxa x {cfg['codons']} codons", flush=True)
    sim, _, _, arrays, tau = bench.build_problem(cfg)
    names, seqs = dr.parse_fasta(sim["fasta"])
    raln = dr.AlignmentRef(names, seqs)
    rtree = dr.TreeRef.from_newick(sim["newick"], keep_rooted=False)
    np.testing.assert_array_equal(arrays["states"], raln.states)
    np.testing.assert_array_equal(arrays["masks"], raln.pmasks)
    np.testing.assert_array_equal(arrays["weights"], raln.weights)
    np.testing.assert_allclose(arrays["pi"], raln.cf3x4, rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(arrays["pos_freqs"], raln.pos_freqs, rtol=1e-13, atol=1e-16)
    assert list(arrays["parent"]) == [n.parent for n in rtree.nodes]
    return cfg, sim, arrays, tau, ho.Problem.from_refs(rtree, raln), raln, rtree


def oracle_eval(name, convention):
    from oracle import hyoracle as ho
    cfg, sim, arrays, tau, prob, raln, rtree = workload(name)
    N = len(rtree.nodes)
    mix = ho.Mixture(cfg["omegas"], np.zeros(N, dtype=np.int32), [[(k, w) for k, w in enumerate(cfg["weights"])]],
                     convention)
    bl = tau if convention == 0 else tau / sim["conv"]
    P = prob.num_patterns
    th = host_threads()
    tile = int(min(64, max(16, -(-P // th))))
    return bl, ho.mixture_eval(prob, ho.Code(0), ho.params(1.0, 1.0, sim["theta"]), raln.pos_freqs, raln.cf3x4,
                               mix, bl, want_grad=True, tile=tile, threads=th)


def engine_for(name, convention):
    cfg, sim, arrays, tau, *_ = workload(name)
    K = len(cfg["omegas"])
    eng = bench.make_engine(arrays, K)
    if convention != 0:
        eng.set_mixture(K, np.zeros(len(arrays["parent"]), dtype=np.int32), [list(range(K))], convention=convention)
    return eng


def check_parity(name, convention):
    cfg, sim, *_ = workload(name)
    bl, ref = oracle_eval(name, convention)
    eng = engine_for(name, convention)
    try:
        out = {}
        for prec, tl, tg in ((FP64, 1e-9, 1e-7), (FP32, 1e-5, 1e-4)):
            print(f"[configs] {name} convention {convention} precision {prec}", flush=True)
            r = eng.eval(bl, cfg["omegas"], cfg["weights"], sim["theta"], precision=prec)
            e_ll = rel(r["log_likelihood"], ref["log_likelihood"])
            e_g = {k: grad_err(r[k], ref[k]) for k in KEYS}
            print(f"[configs] {name} prec {prec}: logL rel err {e_ll:.2e}, grad errs "
                  + ", ".join(f"{k} {v:.2e}" for k, v in e_g.items()), flush=True)
            assert e_ll < tl, (name, prec, r["log_likelihood"], ref["log_likelihood"])
            for k in KEYS:
                assert e_g[k] < tg, (name, prec, k, e_g[k])
            out[prec] = r
        return eng, bl, out
    except BaseException:
        eng.close()
        raise


@pytest.mark.timeout(600)
@pytest.mark.parametrize("name,convention", [("c1", 0), ("c2", 0), ("c2", 1), ("c5", 0), ("c4", 0)])
def test_config_parity(name, convention):
    eng, _, _ = check_parity(name, convention)
    eng.close()


@pytest.mark.timeout(600)
def test_c5_central_differences():
    """The C5 gradient against central differences of the FP64 engine's logL
    (h = 1e-6, parameters perturbed one at a time, cold spectra), on the
    largest branch-length components and on every omega and theta; the
    reference's criterion |g - fd| / (|g| + 1) < 1e-4 (test_core.cpp:195)."""
    cfg, sim, arrays, tau, *_ = workload("c5")
    eng = engine_for("c5", 0)
    try:
        eng.set_warm_start(False)
        om, w, th = list(cfg["omegas"]), list(cfg["weights"]), np.asarray(sim["theta"], dtype=float)
        g = eng.eval(tau, om, w, th, precision=FP64)
        f = lambda b, o, t: eng.eval(b, o, w, t, precision=FP64, grad=False)["log_likelihood"]  # noqa: E731
        h = 1e-6
        order = np.argsort(-np.abs(g["grad_bl"]))[:4]
        for b in order:
            bp, bm = tau.copy(), tau.copy()
            bp[b] += h
            bm[b] -= h
            fd = (f(bp, om, th) - f(bm, om, th)) / (2 * h)
            a = g["grad_bl"][b]
            assert abs(a - fd) / (abs(a) + 1) < 1e-4, ("bl", b, a, fd)
        for k in range(len(om)):
            op, omm = list(om), list(om)
            op[k] += h
            omm[k] -= h
            fd = (f(tau, op, th) - f(tau, omm, th)) / (2 * h)
            a = g["grad_omega"][k]
            assert abs(a - fd) / (abs(a) + 1) < 1e-4, ("omega", k, a, fd)
        for k in range(5):
            tp, tm = th.copy(), th.copy()
            tp[k] += h
            tm[k] -= h
            fd = (f(tau, om, tp) - f(tau, om, tm)) / (2 * h)
            a = g["grad_theta"][k]
            assert abs(a - fd) / (abs(a) + 1) < 1e-4, ("theta", k, a, fd)
    finally:
        eng.close()


@pytest.mark.timeout(900)
def test_c3_fused_batch():
    """C3: 1,024 genes (taxa log-U[8,128], codons log-U[100,3000], 1-4 omega
    components, seed 303) evaluated in one fused hy_eval_batch per precision;
    every 32nd gene against the oracle (FP64 1e-9 / 1e-7, FP32 1e-5 / 1e-4),
    every gene's FP32 result against its FP64 one (1e-5 / 1e-4)."""
    from multiprocessing import get_context
    from oracle import data_ref as dr
    from oracle import hyoracle as ho
    from hyphy3.engine import Batch, create_group
    specs = bench.gene_specs(bench.C3["genes"], bench.C3["seed"])
    # spawned workers (no fork of a process whose HIP runtime may be live)
    with get_context("spawn").Pool(host_threads()) as pool:
        genes = pool.map(bench.build_gene, specs, chunksize=16)
    print(f"[configs] c3: {len(genes)} genes built", flush=True)
    engines = create_group(len(genes))
    try:
        params = []
        for (cfg, sim, arrays, tau), eng in zip(genes, engines):
            bench.make_engine(arrays, len(cfg["omegas"]), eng=eng)
            params.append((tau, list(cfg["omegas"]), list(cfg["weights"]), np.asarray(sim["theta"], dtype=float)))
        batch = Batch(engines)
        res = {}
        for prec in (FP64, FP32):
            print(f"[configs] c3 fused batch precision {prec}", flush=True)
            # result dicts are views of the batch's output arrays: copied
            res[prec] = [{k: np.array(v, copy=True) if isinstance(v, np.ndarray) else v for k, v in r.items()}
                         for r in batch.eval(params, precision=prec)]
        for i in range(0, len(genes), 32):
            cfg, sim, arrays, tau = genes[i]
            names, seqs = dr.parse_fasta(sim["fasta"])
            raln = dr.AlignmentRef(names, seqs)
            rtree = dr.TreeRef.from_newick(sim["newick"], keep_rooted=False)
            np.testing.assert_array_equal(arrays["states"], raln.states)
            np.testing.assert_array_equal(arrays["weights"], raln.weights)
            N = len(rtree.nodes)
            mix = ho.Mixture(cfg["omegas"], np.zeros(N, dtype=np.int32),
                             [[(k, w) for k, w in enumerate(cfg["weights"])]], 0)
            ref = ho.mixture_eval(ho.Problem.from_refs(rtree, raln), ho.Code(0), ho.params(1.0, 1.0, sim["theta"]),
                                  raln.pos_freqs, raln.cf3x4, mix, tau, want_grad=True, threads=host_threads())
            for prec, tl, tg in ((FP64, 1e-9, 1e-7), (FP32, 1e-5, 1e-4)):
                r = res[prec][i]
                assert rel(r["log_likelihood"], ref["log_likelihood"]) < tl, (i, prec)
                for k in KEYS:
                    assert grad_err(r[k], ref[k]) < tg, (i, prec, k)
        worst_ll, worst_g = 0.0, 0.0
        for i, (a, b) in enumerate(zip(res[FP32], res[FP64])):
            e = rel(a["log_likelihood"], b["log_likelihood"])
            ga = np.concatenate([a[k] for k in KEYS])
            gb = np.concatenate([b[k] for k in KEYS])
            eg = grad_err(ga, gb)
            worst_ll, worst_g = max(worst_ll, e), max(worst_g, eg)
            assert e < 1e-5 and eg < 1e-4, (i, e, eg)
        print(f"[configs] c3 FP32 vs FP64 over all genes: logL {worst_ll:.2e}, gradient {worst_g:.2e}", flush=True)
    finally:
        for e in engines:
            e.close()
