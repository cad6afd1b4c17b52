#!/usr/bin/env python3
"""MG94xREV codon likelihood + full gradient throughput on one MI355X.

Workload (default, "c2"): 64 taxa x 5,000 codons simulated under a 3-category
branch-site omega mixture (omega = 0.1, 0.8, 3.0; weights 0.6, 0.35, 0.05),
Yule tree from seed 202.  One step = one evaluation of lnL plus the gradient
with respect to all 125 branch lengths, the 3 omegas, the 3 mixture weights
and the 5 GTR exchangeabilities, including the per-evaluation eigensolves and
P(t) builds, with the compressed patterns already resident in HBM.

Prints one JSON line (rank 0).  ``value`` is FP32-MFMA evaluations/s; the FP64
path, the FP32-vs-FP64 errors, per-stage times, the roofline of the dominant
kernel and the CPU baseline (FP64 restatement of the reference algorithm on
the host cores) are reported beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hyphy-3_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "LL+gradient evals/sec (64 taxa, 5k codons) FP32-MFMA vs FP64; |ΔlogL| vs FP64"
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
S_ALG = 61

CONFIGS = {
    "c1": dict(taxa=16, codons=500, omegas=[0.5], weights=[1.0], seed=101, gaps=0.0),
    "c2": dict(taxa=64, codons=5000, omegas=[0.1, 0.8, 3.0], weights=[0.6, 0.35, 0.05], seed=202, gaps=0.0),
    "c4": dict(taxa=512, codons=20000, omegas=[0.05, 0.4, 1.0, 3.0], weights=[0.4, 0.3, 0.2, 0.1], seed=404,
               gaps=0.25),
    "c5": dict(taxa=256, codons=10000, omegas=[0.1, 0.8, 3.0], weights=[0.6, 0.35, 0.05], seed=505, gaps=0.0),
}
# Many-genes batch: sizes log-uniform, categories uniform on 1..4.
C3 = dict(genes=1024, seed=303, taxa=(8, 128), codons=(100, 3000), omegas=[0.1, 0.8, 3.0, 0.4])
STAGES = ["spectra", "pmat", "inside", "outside", "adjoint", "frechet", "final"]


def build_problem(cfg):
    import hyphy3
    from hyphy3 import simulate
    sim = simulate.simulate(cfg["taxa"], cfg["codons"], cfg["seed"], omegas=cfg["omegas"], weights=cfg["weights"],
                            gap_fraction=cfg["gaps"])
    aln = hyphy3.Alignment.from_text(sim["fasta"])
    tree = hyphy3.Tree.from_newick(sim["newick"], keep_rooted=False)
    arrays = hyphy3.problem_arrays(tree, aln)
    tau = np.array([0.0 if n.id == tree.root_id else n.branch_length * sim["conv"] for n in tree.nodes])
    return sim, aln, tree, arrays, tau


def stage_work(arrays, K, b):
    """Algorithmic flops and HBM bytes per launch of each pruning kernel
    (b = bytes per partial entry; 64 padded states per pattern; tip edges are
    gathers, 0 flop; flops count the 61 real states)."""
    parent = arrays["parent"]
    cp = arrays["child_ptr"]
    n_nodes = len(parent)
    is_leaf = np.array([cp[i + 1] == cp[i] for i in range(n_nodes)])
    n_tax = int(is_leaf.sum())
    n_int = n_nodes - n_tax
    root = int(np.where(parent < 0)[0][0])
    e_int = n_int - 1  # internal nodes other than the root
    B = n_nodes - 1
    P = int(len(arrays["weights"]))
    S = S_ALG
    blk = b * 64 * P  # one node's partials (64 padded states x P patterns)
    tips = 2 * (n_tax + 1) * P  # uint16 tip codes
    return {
        # writes D and E of every internal node, reads the tip codes
        "inside": dict(flops=2.0 * S * S * P * e_int, bytes=blk * n_int + 4 * n_int * P + tips),
        # reads D and E of internal children, writes V and G of every branch;
        # X_c = P_c D_c is recomputed (2 S^2 P per internal edge, counted) instead
        # of being stored by the inside pass
        "outside": dict(flops=4.0 * S * S * P * e_int, bytes=blk * e_int + blk * B + 4 * (e_int + B) * P + tips),
        # A_b = sum_p c_p V_b D_b^T: reads V of every branch and D of internal
        # children, exponents and root likelihoods; writes 64x64 FP64 per branch
        "adjoint": dict(flops=2.0 * S * S * P * e_int + 2.0 * S * P * n_tax,
                        bytes=blk * B + blk * e_int + 4 * (B + e_int + 1) * P + 8 * P + tips + 8 * 4096 * B),
        "pmat": dict(flops=2.0 * S ** 3 * B * K, bytes=3 * b * 4096 * B),
        "_meta": dict(P=P, n_int=n_int, e_int=e_int, B=B, n_tax=n_tax, root=root),
    }


def make_engine(arrays, K, T=0, eng=None):
    from hyphy3.engine import Engine
    eng = eng or Engine(0)
    children = [list(arrays["child_idx"][arrays["child_ptr"][i]:arrays["child_ptr"][i + 1]])
                for i in range(len(arrays["parent"]))]
    eng.set_problem(arrays["parent"], children, arrays["leaf_taxon"], arrays["states"], arrays["masks"],
                    arrays["weights"], S=int(arrays["S"]))
    eng.set_model(int(arrays["code_id"]), arrays["pos_freqs"], arrays["pi"])
    eng.set_mixture(K, np.zeros(len(arrays["parent"]), dtype=np.int32), [list(range(K))], convention=0)
    if T:
        eng.set_tile(T)
    return eng


def step_params(base, i):
    """Parameters of step i: the generating values with a small deterministic
    perturbation (relative 1e-3 .. 2e-3), as consecutive optimizer iterates."""
    tau, om, w, th = base
    f = 1.0 + 1e-3 * (1.0 + np.sin(0.7 * i))
    g = 1.0 + 1e-3 * (1.0 + np.cos(0.9 * i))
    return (tau * g, [o * f for o in om], list(w), np.asarray(th) * (2.0 - f))


def run_steps(eng, base, precision, steps, warmup, profile):
    for i in range(warmup):
        eng.eval(*step_params(base, i), precision=precision)
    sync()
    stage = np.zeros(8)
    t0 = time.perf_counter()
    out = None
    for i in range(steps):
        out = eng.eval(*step_params(base, warmup + i), precision=precision)
        if profile:
            stage += eng.last_timings()
    sync()
    dt = time.perf_counter() - t0
    return dt / steps, out, stage / max(steps, 1)


def sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def cpu_baseline(cfg, sim, seconds):
    """FP64 restatement of the reference algorithm (oracle) on the host cores."""
    from oracle import data_ref as dr
    from oracle import hyoracle as ho
    names, seqs = dr.parse_fasta(sim["fasta"])
    aln = dr.AlignmentRef(names, seqs)
    tree = dr.TreeRef.from_newick(sim["newick"], keep_rooted=False)
    prob = ho.Problem.from_refs(tree, aln)
    K = len(cfg["omegas"])
    N = len(tree.nodes)
    mix = ho.Mixture(cfg["omegas"], np.zeros(N, dtype=np.int32), [[(k, w) for k, w in enumerate(cfg["weights"])]], 0)
    bl = np.array([0.0 if n.id == tree.root else n.bl * sim["conv"] for n in tree.nodes])
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    # BUSTED tile rule (busted.hpp:385-401): clamp(ceil(P/threads), 16, 64)
    P = prob.num_patterns
    tile = 64 if threads <= 1 else int(min(64, max(16, -(-P // threads))))
    base = ho.params(1.0, 1.0, sim["theta"])
    n = 0
    t0 = time.perf_counter()
    while True:
        ho.mixture_eval(prob, ho.Code(0), base, aln.pos_freqs, aln.cf3x4, mix, bl, want_grad=True, tile=tile,
                        threads=threads)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 200:
            break
    return {"value": n / el, "unit": "evals/s", "cores": threads, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"full {cfg['taxa']} taxa x {cfg['codons']} codons x K={K} problem, {n} logL+gradient evals "
                      f"in {el:.1f} s (FP64 restatement of the reference algorithm, tile {tile}, OpenMP; "
                      f"register-blocked S x S by S x tile products, -O3 without -march as the reference builds)"}


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform's."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS) + ["c3"])
    ap.add_argument("--genes", type=int, default=C3["genes"], help="c3 only: number of genes in the batch")
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the cost-model tile choice instead of timing the candidates first")
    ap.add_argument("--fit", action="store_true",
                    help="instead of the throughput run: torch L-BFGS maximum-likelihood fit of branch lengths, "
                         "omegas and GTR rates from a perturbed start (FP32 phase, then FP64)")
    ap.add_argument("--cpu-fit", action="store_true", help="with --fit: also the reference fit on the host (slow)")
    ap.add_argument("--estep", action="store_true",
                    help="instead of the throughput run: BUSTED expectation steps (omega and synonymous-rate weights)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of known size (scales rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--gtr", action="store_true",
                    help="instead of the codon run: the nucleotide GTR stage (lnL + branch and exchangeability "
                         "gradients) on the workload's nucleotide view, against the oracle on the host")
    ap.add_argument("--fel", action="store_true",
                    help="instead of the codon run: FEL per-site refits (grid, Nelder-Mead, Brent per pattern)")
    ap.add_argument("--meme", action="store_true",
                    help="instead of the codon run: MEME per-site refits (FEL baseline, 4-D/3-D Nelder-Mead, EBFs)")
    ap.add_argument("--fel-data", default="cd2", choices=["cd2", "sim"], help="workload of --fel and --meme")
    ap.add_argument("--memory-budget", type=float, default=0.0,
                    help="GB allowed for the partials; below what every node's partials need the engine "
                         "checkpoints inside partials (about sqrt(internal nodes) parts recomputed)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build measures one MI355X; multi-GPU runs are not part of it "
                                   "(the path would shard patterns across GPUs)"}))
        sys.exit(2)
    # the timed regions are bracketed by torch.cuda.synchronize(): initialise
    # torch's device context now, so that its start-up work (threads, device
    # queries) is over before anything is timed
    sync()
    if args.config == "c3":
        bench_genes(args)
        return
    cfg = CONFIGS[args.config]
    K = len(cfg["omegas"])
    if args.copy_calibration:
        copy_calibration()
        return
    if args.fit:
        bench_fit(args, cfg)
        return
    if args.estep:
        bench_estep(args, cfg)
        return
    if args.gtr:
        bench_gtr(args, cfg)
        return
    if args.fel:
        bench_fel(args)
        return
    if args.meme:
        bench_meme(args)
        return
    sim, aln, tree, arrays, tau = build_problem(cfg)
    print(f"[bench] {args.config}: taxa={cfg['taxa']} codons={cfg['codons']} patterns={len(arrays['weights'])} "
          f"nodes={len(arrays['parent'])} K={K}", file=sys.stderr, flush=True)
    eval_args = (tau, cfg["omegas"], cfg["weights"], sim["theta"])
    # timed engine: no per-stage events
    os.environ.pop("HY_PROFILE", None)
    eng = make_engine(arrays, K, args.tile)
    budget = int(args.memory_budget * 1e9)
    if budget:
        eng.set_memory_budget(budget)
    tuning = {}
    if not args.no_autotune and not args.tile:
        # measured tile/parts choice per precision, before (and outside) the timed steps
        t0 = time.perf_counter()
        for prec, name in ((1, "fp32"), (0, "fp64")):
            if prec == 0 and args.no_fp64:
                continue
            tile, parts, ms = eng.autotune(*eval_args, precision=prec, reps=5)
            tuning[name] = {"tile": tile, "parts": parts, "ms": ms}
        tuning["seconds"] = time.perf_counter() - t0
        print(f"[bench] autotune {tuning}", file=sys.stderr, flush=True)
    ms32, out32, _ = run_steps(eng, eval_args, 1, args.steps, args.warmup, False)
    tile32 = eng.last_tile()
    parts32 = eng.last_parts()
    ck32 = eng.last_checkpoint()
    print(f"[bench] fp32 {ms32 * 1e3:.3f} ms/eval", file=sys.stderr, flush=True)
    res = {}
    if not args.no_fp64:
        ms64, out64, _ = run_steps(eng, eval_args, 0, args.steps, args.warmup, False)
        res["fp64_tile"] = eng.last_tile()
        res["partial_bytes_fp64"] = eng.last_checkpoint()[1]
        print(f"[bench] fp64 {ms64 * 1e3:.3f} ms/eval", file=sys.stderr, flush=True)
        g32 = np.concatenate([out32["grad_bl"], out32["grad_omega"], out32["grad_w"], out32["grad_theta"]])
        g64 = np.concatenate([out64["grad_bl"], out64["grad_omega"], out64["grad_w"], out64["grad_theta"]])
        scale = np.abs(g64).max()
        res.update({
            "fp64_evals_per_s": 1.0 / ms64, "fp64_ms_per_step": ms64 * 1e3, "fp32_over_fp64": ms64 / ms32,
            "rel_err_logl_vs_fp64": abs(out32["log_likelihood"] - out64["log_likelihood"]) / abs(out64["log_likelihood"]),
            "abs_err_logl_vs_fp64": abs(out32["log_likelihood"] - out64["log_likelihood"]),
            "rel_err_grad_vs_fp64_max_component": float(np.max(np.abs(g32 - g64) / np.maximum(np.abs(g64), 1e-2 * scale))),
            "rel_err_grad_vs_fp64_norm": float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64)),
            "logL": out64["log_likelihood"],
        })
    # device memory the engine holds (both precisions' buffers once both have run)
    res["device_bytes"] = int(eng.device_bytes())
    eng.close()
    # profiled engine: HIP events around every stage on the engine's stream
    os.environ["HY_PROFILE"] = "1"
    peng = make_engine(arrays, K, args.tile)
    if budget:
        peng.set_memory_budget(budget)
    for prec, name in ((1, "fp32"), (0, "fp64")):
        if name in tuning:
            peng.set_tuning(prec, tuning[name]["tile"], tuning[name]["parts"])
    _, _, st32 = run_steps(peng, eval_args, 1, args.steps, args.warmup, True)
    stage_ms = {s: float(st32[i]) for i, s in enumerate(STAGES)}
    if not args.no_fp64:
        _, _, st64 = run_steps(peng, eval_args, 0, max(3, args.steps // 2), 1, True)
        res["stage_ms_fp64"] = {s: float(st64[i]) for i, s in enumerate(STAGES)}
    peng.set_warm_start(False)
    ms32c, _, st32c = run_steps(peng, eval_args, 1, max(3, args.steps // 4), 1, True)
    peng.close()
    work = stage_work(arrays, K, 4)
    tile = tile32
    cand = {s: stage_ms[s] for s in ("inside", "outside", "adjoint")}
    dom = max(cand, key=cand.get)
    roof = roofline(dom, work[dom], stage_ms[dom] * 1e-3)
    # with subtree parts the pruning passes are two launches per evaluation
    # (subtree parts, top part): the kernel time, bytes and flops are per
    # evaluation, summed over both
    roof["launches_per_eval"] = 2 if (parts32 > 0 and dom in ("inside", "outside")) else 1
    roof["traffic"] = committed_traffic(KERNEL_NAMES[dom], args.config, tile)
    meta = work["_meta"]
    # whole-evaluation bound (algorithmic work of one logL + full gradient, FP32)
    S = S_ALG
    f_alg = 6.0 * S * S * meta["P"] * meta["e_int"] + 8.0 * S ** 3 * meta["B"] * K
    b_alg = 4 * 64 * meta["P"] * 4 * meta["n_int"] + meta["n_tax"] * meta["P"] + 8 * meta["P"] * meta["n_int"] + \
        2 * 4 * 4096 * meta["B"] * K
    t_bound = max(f_alg / (PEAK_FP32_TFLOPS * 1e12), b_alg / (PEAK_HBM_GBS * 1e9))
    line = {
        "metric": METRIC, "value": 1.0 / ms32, "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms32 * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.config}: {cfg['taxa']} taxa x {cfg['codons']} codons x {K} omega categories, "
                               "logL + full gradient (branches, omegas, weights, GTR)",
                   "taxa": cfg["taxa"], "codons": cfg["codons"], "patterns": int(len(arrays["weights"])),
                   "categories": K, "seed": cfg["seed"], "tile": tile, "subtree_parts": parts32},
        "roofline": roof,
        "eval_roofline": {"flops": f_alg, "bytes": b_alg, "t_bound_ms": t_bound * 1e3,
                          "frac": t_bound / ms32, "note": "max(F/157.3 TFLOP/s, B/8 TB/s) over the measured eval"},
        "stage_ms_fp32": stage_ms,
        "ms_per_step_cold_spectra": ms32c * 1e3,
        "stage_ms_fp32_cold_spectra": {s: float(st32c[i]) for i, s in enumerate(STAGES)},
        "steps_note": "consecutive steps evaluate parameters perturbed by 1e-3 relative (optimizer iterates); spectra are "
                      "warm-started from the previous step (cold-start figures reported beside); stage times come from "
                      "a second engine with HIP events around each stage",
    }
    line.update(res)
    line["partial_bytes_fp32"] = ck32[1]
    line["checkpoint"] = {"budget_bytes": budget or None, "parts_recomputed": ck32[0],
                          "note": "parts 0: every node's partials resident; otherwise inside partials of the top "
                                  "part and part roots resident, each part's interior recomputed"}
    line["autotune"] = tuning or None
    if not args.no_cpu:
        try:
            line["cpu_baseline"] = cpu_baseline(cfg, sim, args.cpu_seconds)
            line["gpu_over_cpu"] = line["value"] / line["cpu_baseline"]["value"]
        except Exception as ex:  # the oracle library may not be built on a bare box
            line["cpu_baseline"] = {"value": None, "error": str(ex)}
    print(json.dumps(line), flush=True)


def bench_fit(args, cfg):
    """Maximum-likelihood fit through hyphy3.fit (torch L-BFGS over the
    engine-backed autograd function): start from the generating values
    perturbed (branch lengths x1.3, omegas x1.5, theta 1), report the fitted
    lnL, iterations and time.  With --cpu-fit, the reference fitter (scipy
    L-BFGS-B over the FP64 oracle, the reference's parameterisation and
    bounds, on the host cores) is started at the GPU optimum for at most 40
    evaluations: the oracle's lnL there and the improvement it still finds (a
    full reference fit from the start would take about 1,300 evaluations of
    over a second each)."""
    import hyphy3
    from hyphy3 import simulate
    from hyphy3.fit import MixtureFit, make_engine as fit_engine
    sim = simulate.simulate(cfg["taxa"], cfg["codons"], cfg["seed"], omegas=cfg["omegas"], weights=cfg["weights"],
                            gap_fraction=cfg["gaps"])
    aln = hyphy3.Alignment.from_text(sim["fasta"])
    tree = hyphy3.Tree.from_newick(sim["newick"], keep_rooted=False)
    K = len(cfg["omegas"])
    eng, arrays = fit_engine(tree, aln, K)
    bl0 = np.array([0.0 if n.id == tree.root_id else n.branch_length * 1.3 for n in tree.nodes])
    om0 = [o * 1.5 for o in cfg["omegas"]]
    th0 = np.ones(5)
    n_sites = float(np.sum(arrays["weights"]))
    # every evaluation of the fit: its wall time and each rate matrix's
    # eigensolver status (>= 100: warm start refined; below: cold solve)
    calls = []
    plain_eval = eng.eval

    def recorded_eval(*a, **kw):
        t0 = time.perf_counter()
        out = plain_eval(*a, **kw)
        calls.append((time.perf_counter() - t0, eng.eigen_status().copy(), kw.get("precision", 0)))
        return out

    eng.eval = recorded_eval
    fit = MixtureFit(eng, arrays["parent"], bl0, om0, cfg["weights"], th0, n_sites=n_sites)
    r = fit.run(max_iter=1000, tol_grad=1e-7)
    line = {"workload": f"{args.config} fit: {cfg['taxa']} taxa x {cfg['codons']} codons x {K} omega categories, "
                        "branch lengths + omegas + GTR by torch L-BFGS (FP32 phase, then FP64)",
            "log_likelihood": r["log_likelihood"], "iterations_fp32": r["iterations_fp32"],
            "iterations_fp64": r["iterations_fp64"], "evaluations": r["evaluations"], "seconds": r["seconds"],
            "max_abs_log_gradient": r["max_abs_log_gradient"],
            "max_abs_log_gradient_per_site": r["max_abs_log_gradient_per_site"],
            "omegas": list(r["omegas"]), "theta": list(r["theta"])}
    if calls:
        st = np.concatenate([c[1] for c in calls])
        cold_evals = sum(1 for c in calls if (c[1] < 100).any())
        line["spectra"] = {
            "evaluations_recorded": len(calls), "matrices_solved": int(st.size),
            "fraction_matrices_cold": float(np.mean(st < 100)),
            "fraction_evaluations_with_a_cold_solve": cold_evals / len(calls),
            "mean_ms_per_evaluation": 1e3 * float(np.mean([c[0] for c in calls])),
            "mean_ms_per_evaluation_fp32": 1e3 * float(np.mean([c[0] for c in calls if c[2] == 1] or [0.0])),
            "mean_ms_per_evaluation_fp64": 1e3 * float(np.mean([c[0] for c in calls if c[2] != 1] or [0.0])),
            "note": "eigensolver status per rate matrix: >= 100 warm start refined, 50+k tridiagonal cold solve, "
                    "1..40 Jacobi"}
    eng.eval = plain_eval
    # the same fit with the optimizer on the GPU (hy_fit_lbfgs): FP32 phase,
    # then FP64, each trial point one replay of a captured evaluation + step
    # graph; then the FP64 phase again stepped on the host (hy_eval +
    # hy_lbfgs_step) from the same FP32 result, for the per-iteration cost
    legs = {}
    for name, captured in (("captured", True), ("host_stepped", False)):
        t0 = time.perf_counter()
        r32 = eng.fit_lbfgs(bl0, om0, cfg["weights"], th0, precision=1, captured=captured, scale=1.0 / n_sites,
                            max_iter=1000, tol_grad=1e-7, tol_f=1e-9)
        t1 = time.perf_counter()
        r64 = eng.fit_lbfgs(r32["bl"], r32["omegas"], cfg["weights"], r32["theta"], precision=0, captured=captured,
                            scale=1.0 / n_sites, max_iter=1000, tol_grad=1e-7, tol_f=1e-12)
        t2 = time.perf_counter()
        legs[name] = {"log_likelihood": r64["logL"], "iterations_fp32": r32["iterations"],
                      "evaluations_fp32": r32["evaluations"], "seconds_fp32": t1 - t0,
                      "ms_per_evaluation_fp32": 1e3 * (t1 - t0) / max(1, r32["evaluations"]),
                      "iterations_fp64": r64["iterations"], "evaluations_fp64": r64["evaluations"],
                      "seconds_fp64": t2 - t1, "ms_per_evaluation_fp64": 1e3 * (t2 - t1) / max(1, r64["evaluations"]),
                      "replays": r32["replays"] + r64["replays"], "status": [r32["status"], r64["status"]]}
        print(f"[bench] fit {name}: {legs[name]}", file=sys.stderr, flush=True)
    legs["identical_iterates"] = bool(legs["captured"]["log_likelihood"] == legs["host_stepped"]["log_likelihood"]
                                      and legs["captured"]["evaluations_fp64"] == legs["host_stepped"]["evaluations_fp64"])
    line["gpu_optimizer"] = legs
    eng.close()
    print(json.dumps(line), flush=True)
    if args.cpu_fit:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle import data_ref as dr
        from helpers import OracleMixtureEngine, oracle_fit
        names, seqs = dr.parse_fasta(sim["fasta"])
        raln = dr.AlignmentRef(names, seqs)
        rtree = dr.TreeRef.from_newick(sim["newick"], keep_rooted=False)
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        if os.environ.get("OMP_NUM_THREADS", "").isdigit():
            threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
        t0 = time.perf_counter()
        ll_at = OracleMixtureEngine(rtree, raln, threads).eval(r["bl"], r["omegas"], r["weights"], r["theta"],
                                                               grad=False)["log_likelihood"]
        ll_ref, _, om_ref, th_ref = oracle_fit(rtree, raln, r["omegas"], r["weights"], r["theta"], r["bl"],
                                               threads=threads, max_evals=40)
        line["cpu_reference_check"] = {
            "oracle_log_likelihood_at_gpu_optimum": ll_at, "rel_diff_gpu_vs_oracle": abs(ll_at - r["log_likelihood"]) / abs(ll_at),
            "reference_fitter_log_likelihood": ll_ref, "improvement_found": ll_ref - ll_at,
            "seconds": time.perf_counter() - t0, "cores": threads,
            "kind": "scipy L-BFGS-B over the FP64 restatement of the reference algorithm, started at the GPU optimum, "
                    "at most 40 evaluations"}
        print(json.dumps(line), flush=True)


def bench_estep(args, cfg):
    """BUSTED expectation step (hy_busted_expected_weights) on the workload:
    every branch a test branch with the K omega components, raw branch times,
    M = 1 and M = 3 synonymous-rate classes; per-call time and the updated
    weights.  The CPU column is the oracle's literal restatement over the
    reference workspaces for M = 1 (one call)."""
    sim, aln, tree, arrays, tau = build_problem(cfg)
    K = len(cfg["omegas"])
    eng = make_engine(arrays, K)
    res = {"workload": f"{args.config} E-step: {cfg['taxa']} taxa x {cfg['codons']} codons x {K} omega components "
                       "on every branch", "calls": args.steps}
    for M, sr, sv in ((1, [1.0], [1.0]), (3, [0.4, 1.0, 2.2], [0.3, 0.5, 0.2])):
        for prec, name in ((1, "fp32"), (0, "fp64")):
            for _ in range(args.warmup):
                eng.busted_expected_weights(tau, cfg["omegas"], cfg["weights"], sim["theta"], sr, sv, precision=prec)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                nw, nv = eng.busted_expected_weights(tau, cfg["omegas"], cfg["weights"], sim["theta"], sr, sv,
                                                     precision=prec)
            res[f"ms_per_call_M{M}_{name}"] = (time.perf_counter() - t0) / args.steps * 1e3
            res[f"weights_M{M}_{name}"] = list(nw)
            res[f"syn_weights_M{M}_{name}"] = list(nv)
    eng.close()
    if not args.no_cpu:
        from oracle import data_ref as dr
        from oracle import hyoracle as ho
        names, seqs = dr.parse_fasta(sim["fasta"])
        raln = dr.AlignmentRef(names, seqs)
        rtree = dr.TreeRef.from_newick(sim["newick"], keep_rooted=False)
        is_test = np.array([n.id != rtree.root for n in rtree.nodes])
        t0 = time.perf_counter()
        rw, _ = ho.busted_expected_weights(ho.Problem.from_refs(rtree, raln), ho.Code(0),
                                           ho.params(1.0, 1.0, sim["theta"]), raln.pos_freqs, raln.cf3x4,
                                           raln.weights, cfg["omegas"], cfg["weights"], 1.0, is_test, tau)
        res["cpu_literal_M1"] = {"seconds": time.perf_counter() - t0, "weights": list(rw), "cores": 1,
                                 "kind": "oracle: literal restatement over the reference workspaces (numpy)"}
        res["max_rel_diff_fp64_vs_literal"] = float(np.max(np.abs(np.array(res["weights_M1_fp64"]) - rw) / rw))
    print(json.dumps(res), flush=True)


def bench_gtr(args, cfg):
    """Nucleotide GTR stage (hy_gtr_eval; compute_gtr_log_likelihood,
    likelihood.hpp:181-254) on the workload's nucleotide view: lnL plus the
    branch-length and exchangeability gradients per step, FP64 and FP32; the
    oracle (FP64, host cores) evaluates the same lnL + gradient for the CPU
    column."""
    import hyphy3
    from hyphy3 import simulate
    from hyphy3.engine import Engine
    sim = simulate.simulate(cfg["taxa"], cfg["codons"], cfg["seed"], omegas=cfg["omegas"], weights=cfg["weights"],
                            gap_fraction=cfg["gaps"])
    aln = hyphy3.Alignment.from_text(sim["fasta"])
    nuc = aln.to_nucleotide_alignment()
    tree = hyphy3.Tree.from_newick(sim["newick"], keep_rooted=False)
    arrays = hyphy3.problem_arrays(tree, nuc)
    children = [list(arrays["child_idx"][arrays["child_ptr"][i]:arrays["child_ptr"][i + 1]])
                for i in range(len(arrays["parent"]))]
    eng = Engine(0)
    eng.set_problem(arrays["parent"], children, arrays["leaf_taxon"], arrays["states"], arrays["masks"],
                    arrays["weights"], S=4)
    bl = np.array([0.0 if n.id == tree.root_id else n.branch_length for n in tree.nodes])
    theta, freqs = np.asarray(sim["theta"], dtype=float), np.asarray(arrays["pi"], dtype=float)
    res = {"workload": f"{args.config} nucleotide view: {cfg['taxa']} taxa x {3 * cfg['codons']} sites "
                       f"({len(arrays['weights'])} patterns), GTR lnL + gradients (branches, 5 exchangeabilities)",
           "steps": args.steps, "warmup": args.warmup}
    outs = {}
    for prec, name in ((1, "fp32"), (0, "fp64")):
        for i in range(args.warmup):
            eng.gtr_eval(bl * (1.0 + 1e-3 * i), theta, freqs, precision=prec)
        t0 = time.perf_counter()
        for i in range(args.steps):
            out = eng.gtr_eval(bl * (1.0 + 1e-3 * (i % 7)), theta, freqs, precision=prec)
        dt = (time.perf_counter() - t0) / args.steps
        res[f"{name}_ms_per_eval"] = dt * 1e3
        res[f"{name}_evals_per_s"] = 1.0 / dt
        outs[name] = eng.gtr_eval(bl, theta, freqs, precision=prec)
    g32 = np.concatenate([outs["fp32"]["grad_bl"], outs["fp32"]["grad_theta"]])
    g64 = np.concatenate([outs["fp64"]["grad_bl"], outs["fp64"]["grad_theta"]])
    res["rel_err_logl_fp32_vs_fp64"] = abs(outs["fp32"]["log_likelihood"] - outs["fp64"]["log_likelihood"]) / \
        abs(outs["fp64"]["log_likelihood"])
    res["rel_err_grad_fp32_vs_fp64_max_component"] = float(np.max(np.abs(g32 - g64) / np.maximum(
        np.abs(g64), 1e-2 * np.abs(g64).max())))
    res["logL"] = outs["fp64"]["log_likelihood"]
    eng.close()
    if not args.no_cpu:
        from oracle import data_ref as dr
        from oracle import hyoracle as ho
        names, seqs = dr.parse_fasta(sim["fasta"])
        naln = dr.NucleotideRef(dr.AlignmentRef(names, seqs))
        rtree = dr.TreeRef.from_newick(sim["newick"], keep_rooted=False)
        prob = ho.nuc_problem(rtree, naln)
        rbl = np.array([0.0 if n.id == rtree.root else n.bl for n in rtree.nodes])
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        if os.environ.get("OMP_NUM_THREADS", "").isdigit():
            threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
        n, t0 = 0, time.perf_counter()
        while True:
            ref = ho.gtr_eval(prob, theta, naln.nuc_freqs, rbl, threads=threads)
            n += 1
            if time.perf_counter() - t0 >= args.cpu_seconds or n >= 50:
                break
        el = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": n / el, "unit": "evals/s", "cores": threads, "kind": "port",
                               "sample": f"full problem, {n} lnL + gradient evaluations in {el:.1f} s (oracle: "
                                         "OpenMP pruning and inside-outside workspace, numpy gradient assembly)"}
        res["rel_err_logl_fp64_vs_oracle"] = abs(res["logL"] - ref["log_likelihood"]) / abs(ref["log_likelihood"])
    print(json.dumps(res), flush=True)


def bench_fel(args):
    """FEL per-site refits (FELAnalyzer's all-branches-tested path,
    fel.hpp:326-445,778-834) on CD2 with the fitted tree and global MG94 rates
    of the reference's matched-tree parity test (tests/golden), or on a
    simulated alignment (``--fel-data sim``).  One step = the whole refit
    stage over every pattern; the oracle's sequential FEL (one host core) on a
    bounded sample of patterns is the CPU column, in candidates per second."""
    from hyphy3 import fel as pf
    fasta, newick, theta, alpha, beta, label, truth, aln, tree = refit_workload(args)
    fel = pf.FEL(tree, aln, theta, alpha, beta)
    n_pat = int(fel.states.shape[0])
    res = {"metric": "FEL per-site refit stage", "unit": "ms", "workload": label, "patterns": n_pat,
           "steps": args.steps, "warmup": args.warmup}
    for _ in range(args.warmup):
        fel.run()
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        out = fel.run()
        times.append(time.perf_counter() - t0)
    st = fel.stats
    fel.engine.fel_counters(reset=True)
    fel.run()
    res.update(ms_per_stage=1e3 * float(np.median(times)), ms_min=1e3 * float(np.min(times)), calls=st["calls"],
               candidates=st["candidates"], candidates_per_s=st["candidates"] / float(np.median(times)),
               spectra=fel.engine.fel_counters())
    if args.fel_data == "cd2":
        from scipy.stats import pearsonr
        lrt = [r["lrt"] for r in out]
        res["lrt_pearson_vs_hyphy25"] = float(pearsonr(lrt, [m[3] for m in truth["mle"]])[0])
    fel.engine.close()
    if not args.no_cpu:
        from oracle import data_ref as dr
        from oracle import fel_ref as fr
        from oracle import hyoracle as ho
        names, seqs = dr.parse_fasta(fasta)
        raln = dr.AlignmentRef(names, seqs)
        rtree = dr.TreeRef.from_newick(newick)
        prob = ho.Problem.from_refs(rtree, raln)
        rbl = np.array([0.0 if n.id == rtree.root else n.bl for n in rtree.nodes])
        n_done, n_cand, t0 = 0, 0, time.perf_counter()
        for p in range(prob.num_patterns):
            r = fr.fel(prob, ho.Code(0), theta, raln.pos_freqs, raln.cf3x4, rbl, alpha, beta, patterns=[p])[0]
            n_done += 1
            n_cand += r["evaluations"]
            if time.perf_counter() - t0 >= args.cpu_seconds:
                break
        el = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": n_cand / el, "unit": "candidates/s", "cores": 1, "kind": "port",
                               "sample": f"first {n_done} of {prob.num_patterns} patterns refitted sequentially "
                                         f"({n_cand} candidates in {el:.1f} s; FP64 oracle: eigendecomposition "
                                         "per candidate omega, single-pattern pruning)"}
    print(json.dumps(res), flush=True)


def refit_workload(args):
    """The per-site refit workload: CD2 with the fitted tree and global MG94
    rates of the reference's matched-tree FEL parity test (tests/golden), or
    a simulated alignment (``--fel-data sim``)."""
    import hyphy3
    from hyphy3 import simulate
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden")
    truth = json.load(open(os.path.join(golden, "cd2_fel_mle.json")))
    if args.fel_data == "cd2":
        g = truth["global_mg94"]
        fasta = open(os.path.join(golden, "cd2.fna")).read()
        newick = open(os.path.join(golden, "cd2_fitted.nwk")).read()
        theta, alpha, beta = g["theta"], g["alpha"], g["beta"]
        label = "CD2 (10 taxa x 187 codons), fitted tree and global MG94 rates of the reference's FEL parity test"
    else:
        sim = simulate.simulate(32, 1000, 11, omegas=[0.1, 1.0, 4.0], weights=[0.6, 0.3, 0.1])
        fasta, newick = sim["fasta"], sim["newick"]
        theta, alpha, beta = list(sim["theta"]), 1.0, 0.5
        label = "simulated 32 taxa x 1000 codons (omega mixture 0.1/1/4), generating tree"
    aln = hyphy3.Alignment.from_text(fasta)
    tree = hyphy3.Tree.from_newick(newick, keep_rooted=True)
    return fasta, newick, theta, alpha, beta, label, truth, aln, tree


def bench_meme(args):
    """MEME per-site refits (MEMEAnalyzer::analyze_pattern with every branch
    tested, meme.hpp:166-648, after FEL's refits) on the FEL workload.  One
    step = FEL's and MEME's refit stages over every pattern (at most 10 timed
    steps); the oracle's sequential MEME (one host core) on a bounded sample
    of patterns is the CPU column, in candidates per second."""
    from hyphy3 import meme as pm
    fasta, newick, theta, alpha, beta, label, truth, aln, tree = refit_workload(args)
    meme = pm.MEME(tree, aln, theta, alpha, beta)
    n_pat = int(meme.states.shape[0])
    steps = max(1, min(args.steps, 10))
    res = {"metric": "MEME per-site refit stage (FEL baseline included)", "unit": "ms", "workload": label,
           "patterns": n_pat, "steps": steps, "warmup": min(args.warmup, 2)}
    for _ in range(res["warmup"]):
        meme.run()
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        out = meme.run()
        times.append(time.perf_counter() - t0)
    st = meme.stats
    res.update(ms_per_stage=1e3 * float(np.median(times)), ms_min=1e3 * float(np.min(times)),
               meme_calls=st["calls"], meme_candidates=st["candidates"],
               meme_candidates_per_s=st["candidates"] / float(np.median(times)),
               sites_p_below_0_1=int(sum(r["p_value"] <= 0.1 for r in out)))
    if args.fel_data == "cd2":
        from scipy.stats import pearsonr, spearmanr
        mt = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                         "cd2_meme_mle.json")))
        lrt = [r["lrt"] for r in out]
        g = [m[5] for m in mt["mle"]]
        res["lrt_pearson_vs_hyphy25"] = float(pearsonr(lrt, g)[0])
        res["lrt_spearman_vs_hyphy25"] = float(spearmanr(lrt, g)[0])
    meme.engine.close()
    if not args.no_cpu:
        from oracle import data_ref as dr
        from oracle import hyoracle as ho
        from oracle import meme_ref as mr
        names, seqs = dr.parse_fasta(fasta)
        raln = dr.AlignmentRef(names, seqs)
        rtree = dr.TreeRef.from_newick(newick)
        prob = ho.Problem.from_refs(rtree, raln)
        rbl = np.array([0.0 if n.id == rtree.root else n.bl for n in rtree.nodes])
        n_done, n_cand, t0 = 0, 0, time.perf_counter()
        for p in range(prob.num_patterns):
            r, _ = mr.meme(prob, ho.Code(0), theta, raln.pos_freqs, raln.cf3x4, rbl, alpha, beta, patterns=[p])[0]
            n_done += 1
            n_cand += r.get("evaluations", 0)
            if time.perf_counter() - t0 >= args.cpu_seconds:
                break
        el = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": n_cand / el, "unit": "MEME candidates/s", "cores": 1, "kind": "port",
                               "sample": f"first {n_done} of {prob.num_patterns} patterns (FEL then MEME) refitted "
                                         f"sequentially ({n_cand} MEME candidates in {el:.1f} s; FP64 oracle: "
                                         "eigendecompositions per candidate, single-pattern pruning)"}
    print(json.dumps(res), flush=True)


def gene_specs(n_genes, seed):
    """(taxa, codons, categories, seed) of every gene of the batch."""
    rng = np.random.default_rng(seed)
    out = []
    for g in range(n_genes):
        n = int(round(np.exp(rng.uniform(*np.log(C3["taxa"])))))
        L = int(round(np.exp(rng.uniform(*np.log(C3["codons"])))))
        K = int(rng.integers(1, 5))
        out.append((n, L, K, seed * 100000 + g))
    return out


def build_gene(spec):
    n, L, K, seed = spec
    w = list(np.full(K, 1.0 / K))
    cfg = dict(taxa=n, codons=L, omegas=C3["omegas"][:K], weights=w, seed=seed, gaps=0.0)
    sim, aln, tree, arrays, tau = build_problem(cfg)
    return cfg, {"newick": sim["newick"], "fasta": sim["fasta"], "theta": sim["theta"], "conv": sim["conv"]}, arrays, tau


def bench_genes(args):
    """Batched many-genes workload: one step = logL + full gradient of every
    gene of the batch through one hy_eval_batch call."""
    from multiprocessing import get_context
    from hyphy3.engine import Batch, create_group
    specs = gene_specs(args.genes, C3["seed"])
    nproc = min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1))
    t0 = time.perf_counter()
    with get_context("fork").Pool(nproc) as pool:  # before any device work
        genes = pool.map(build_gene, specs, chunksize=8)
    print(f"[bench] c3: {len(genes)} genes built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    engines, bases = create_group(len(genes)), []
    S = S_ALG
    f_alg = 0.0
    pats = 0
    for (cfg, sim, arrays, tau), eng in zip(genes, engines):
        K = len(cfg["omegas"])
        make_engine(arrays, K, eng=eng)
        bases.append((tau, list(cfg["omegas"]), list(cfg["weights"]), sim["theta"]))
        meta = stage_work(arrays, K, 4)["_meta"]
        f_alg += 6.0 * S * S * meta["P"] * meta["e_int"] + 8.0 * S ** 3 * meta["B"] * K
        pats += meta["P"]
    dev_bytes = sum(e.device_bytes() for e in engines)
    print(f"[bench] c3: {len(engines)} engines set up", file=sys.stderr, flush=True)

    batch = Batch(engines)
    # flat base parameters; step i perturbs them as step_params does, vectorised
    bl0 = np.concatenate([b[0] for b in bases])
    om0 = np.concatenate([b[1] for b in bases])
    w0 = np.concatenate([b[2] for b in bases])
    th0 = np.stack([np.asarray(b[3], dtype=float) for b in bases])

    def step(i, prec):
        f = 1.0 + 1e-3 * (1.0 + np.sin(0.7 * i))
        g = 1.0 + 1e-3 * (1.0 + np.cos(0.9 * i))
        return batch.eval_flat(bl0 * g, om0 * f, w0, th0 * (2.0 - f), precision=prec)

    res = {}
    tuning = {}
    for prec, name in ((1, "fp32"), (0, "fp64")):
        if prec == 0 and args.no_fp64:
            continue
        if not args.no_autotune:
            # measured tile of the fused batch, before (and outside) the timed steps
            step(0, prec)
            t0 = time.perf_counter()
            tile, ms = batch.autotune(precision=prec, reps=2)
            tuning[name] = {"tile": tile, "ms": ms, "seconds": time.perf_counter() - t0}
            print(f"[bench] c3 {name} autotune {tuning[name]}", file=sys.stderr, flush=True)
        for i in range(args.warmup):
            step(i, prec)
            print(f"[bench] c3 {name} warmup step {i} done", file=sys.stderr, flush=True)
        sync()
        t0 = time.perf_counter()
        for i in range(args.steps):
            out = step(args.warmup + i, prec)
        sync()
        res[name] = ((time.perf_counter() - t0) / args.steps, [dict(r) for r in out])
        print(f"[bench] c3 {name} {res[name][0] * 1e3:.2f} ms/batch", file=sys.stderr, flush=True)
    ms32, out32 = res["fp32"]
    t_bound = f_alg / (PEAK_FP32_TFLOPS * 1e12)
    line = {
        "metric": "batched genes: LL+gradient evals/sec of the whole batch (FP32-MFMA)", "value": 1.0 / ms32,
        "unit": "batch evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms32 * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"c3: {len(genes)} genes, taxa log-U{C3['taxa']}, codons log-U{C3['codons']}, "
                               "K ~ U{1..4}; logL + full gradient per gene via hy_eval_batch",
                   "genes": len(genes), "patterns_total": int(pats), "seed": C3["seed"]},
        "gene_evals_per_s": len(genes) / ms32,
        "eval_roofline": {"flops": f_alg, "t_bound_ms": t_bound * 1e3, "frac": t_bound / ms32,
                          "note": "sum over genes of 6 S^2 P E_int + 8 S^3 B K at 157.3 TFLOP/s"},
        "device_bytes": int(dev_bytes),
        "autotune": tuning or None,
    }
    if "fp64" in res:
        ms64, out64 = res["fp64"]
        line["fp64_ms_per_step"] = ms64 * 1e3
        line["fp32_over_fp64"] = ms64 / ms32
        line["max_rel_err_logl_vs_fp64"] = max(abs(a["log_likelihood"] - b["log_likelihood"]) / abs(b["log_likelihood"])
                                               for a, b in zip(out32, out64))
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_genes(genes, args.cpu_seconds, f_alg)
        if line["cpu_baseline"].get("value"):
            line["gpu_over_cpu"] = line["value"] / line["cpu_baseline"]["value"]
    for e in engines:
        e.close()
    print(json.dumps(line), flush=True)


def cpu_baseline_genes(genes, seconds, f_total):
    """Oracle on the first genes of the batch until `seconds` pass, scaled to
    the whole batch by algorithmic flops."""
    from oracle import data_ref as dr
    from oracle import hyoracle as ho
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    S = S_ALG
    t0 = time.perf_counter()
    f_done, n = 0.0, 0
    for cfg, sim, arrays, tau in genes:
        names, seqs = dr.parse_fasta(sim["fasta"])
        aln = dr.AlignmentRef(names, seqs)
        tree = dr.TreeRef.from_newick(sim["newick"], keep_rooted=False)
        prob = ho.Problem.from_refs(tree, aln)
        K = len(cfg["omegas"])
        N = len(tree.nodes)
        mix = ho.Mixture(cfg["omegas"], np.zeros(N, dtype=np.int32), [[(k, w) for k, w in enumerate(cfg["weights"])]], 0)
        bl = np.array([0.0 if nd.id == tree.root else nd.bl * sim["conv"] for nd in tree.nodes])
        P = prob.num_patterns
        tile = 64 if threads <= 1 else int(min(64, max(16, -(-P // threads))))
        ho.mixture_eval(prob, ho.Code(0), ho.params(1.0, 1.0, sim["theta"]), aln.pos_freqs, aln.cf3x4, mix, bl,
                        want_grad=True, tile=tile, threads=threads)
        meta = stage_work(arrays, K, 4)["_meta"]
        f_done += 6.0 * S * S * meta["P"] * meta["e_int"] + 8.0 * S ** 3 * meta["B"] * K
        n += 1
        if time.perf_counter() - t0 >= seconds:
            break
    el = time.perf_counter() - t0
    frac = f_done / f_total
    return {"value": frac / el, "unit": "batch evals/s", "cores": threads, "kind": "port",
            "sample": f"first {n} of {len(genes)} genes ({100 * frac:.1f}% of the batch's flops) in {el:.1f} s, "
                      "scaled to the batch by algorithmic flops (FP64 restatement of the reference algorithm, OpenMP)"}


KERNEL_NAMES = {"inside": "k_inside", "outside": "k_outside", "adjoint": "k_adjoint"}


def roofline(stage, w, t):
    """Roofline entry of one kernel from its algorithmic flops and HBM bytes per
    launch and its measured average launch time t (s); the binding resource is
    the one with the larger ideal time."""
    fl, by = w["flops"], w["bytes"]
    if fl / (PEAK_FP32_TFLOPS * 1e12) >= by / (PEAK_HBM_GBS * 1e9):
        r = {"bound": "mfma", "achieved": fl / t / 1e12, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s"}
    else:
        r = {"bound": "hbm", "achieved": by / t / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s"}
    r["frac"] = r["achieved"] / r["peak"]
    r.update({"kernel": KERNEL_NAMES[stage] + "<float>", "algorithmic_flops": fl, "algorithmic_bytes": by,
              "kernel_ms": t * 1e3})
    return r


def committed_traffic(kernel, config, tile):
    """HBM bytes of `kernel` per evaluation (all its launches in one
    evaluation) measured with rocprofv3 PMC counters (FETCH_SIZE + WRITE_SIZE,
    scaled by a copy kernel of known size), read from the summary committed
    under profiles/; None when no measurement matches."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        data = json.load(open(path))
    except (OSError, ValueError):
        return None
    e = data.get(f"{kernel}/{config}/T{tile}")
    return None if e is None else e.get("bytes_per_eval", e.get("bytes_per_launch"))


def copy_calibration():
    """A plain device copy of known size (256 MiB read + 256 MiB written), run
    under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to scale the counters."""
    import torch
    n = 64 * 1024 * 1024
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}))



if __name__ == "__main__":
    main()
