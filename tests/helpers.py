"""Shared builders for the parity tests: problems for the FP64 CPU oracle and
for the MI355X engine from the same seeded inputs."""
import os

import numpy as np

from oracle import data_ref as dr
from oracle import hyoracle as ho

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_cd2_reduced():
    names, seqs = dr.parse_fasta(open(os.path.join(GOLDEN, "cd2_reduced.fna")).read())
    aln = dr.AlignmentRef(names, seqs)
    tree = dr.TreeRef.from_newick(open(os.path.join(GOLDEN, "cd2_reduced.nwk")).read())
    return tree, aln


def sim_problem(n_taxa, n_codons, seed, omegas=(0.5,), weights=(1.0,), gaps=0.0, keep_rooted=False):
    from hyphy3 import simulate
    sim = simulate.simulate(n_taxa, n_codons, seed, omegas=omegas, weights=weights, gap_fraction=gaps)
    names, seqs = dr.parse_fasta(sim["fasta"])
    aln = dr.AlignmentRef(names, seqs)
    tree = dr.TreeRef.from_newick(sim["newick"], keep_rooted=keep_rooted)
    return tree, aln, sim


def tree_arrays(tree, aln):
    parent = [n.parent for n in tree.nodes]
    children = [list(n.children) for n in tree.nodes]
    l2t = tree.leaf_to_taxon(aln.names)
    return parent, children, l2t


def make_engine(tree, aln, code_id=0, eng=None):
    from hyphy3.engine import Engine
    eng = eng or Engine(0)
    parent, children, l2t = tree_arrays(tree, aln)
    eng.set_problem(parent, children, l2t, aln.states, aln.pmasks, aln.weights, S=aln.code.S)
    eng.set_model(code_id, aln.pos_freqs, aln.cf3x4)
    return eng


def branch_lengths(tree):
    return np.array([0.0 if n.id == tree.root else n.bl for n in tree.nodes])


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def grad_err(g, ref, floor_frac=1e-2):
    """Max component-wise relative error, each component's denominator
    floored at floor_frac * max|ref| (near-zero components are compared on
    the scale of the whole gradient)."""
    g = np.asarray(g, dtype=float)
    ref = np.asarray(ref, dtype=float)
    if ref.size == 0:
        return 0.0
    scale = max(np.abs(ref).max(), 1e-300)
    den = np.maximum(np.abs(ref), floor_frac * scale)
    return float(np.max(np.abs(g - ref) / den))


class OracleMixtureEngine:
    """Stand-in with the ``Engine.eval`` contract backed by the FP64 CPU oracle
    (one class mixing every omega component on all branches, 3 / mu_bar time
    convention); drives host-side logic such as the L-BFGS fit on the CPU."""

    def __init__(self, tree, aln, threads=8):
        self.tree, self.aln, self.threads = tree, aln, threads
        self.prob = ho.Problem.from_refs(tree, aln)
        self.code = ho.Code(0)
        self.calls = 0

    def eval(self, bl, omegas, comp_w, theta, precision=0, grad=True):
        N = len(self.tree.nodes)
        mix = ho.Mixture(list(omegas), np.zeros(N, dtype=np.int32), [[(k, float(w)) for k, w in enumerate(comp_w)]], 1)
        self.calls += 1
        return ho.mixture_eval(self.prob, self.code, ho.params(1.0, 1.0, theta), self.aln.pos_freqs, self.aln.cf3x4,
                               mix, np.asarray(bl, dtype=np.float64), want_grad=grad, threads=self.threads)


def oracle_fit(tree, aln, omegas, weights, theta, bl, threads=8, max_evals=15000):
    """Reference fit on the CPU: scipy's L-BFGS-B (8 correction pairs) over the
    FP64 oracle with the reference fitter's log parameterisation and bounds
    (optimizer.hpp:964-989), at most max_evals evaluations.  Returns
    (lnL, bl, omegas, theta)."""
    from scipy.optimize import minimize
    eng = OracleMixtureEngine(tree, aln, threads)
    parent = np.array([n.parent for n in tree.nodes])
    br = np.nonzero(parent >= 0)[0]
    K = len(omegas)
    x0 = np.concatenate([np.log(np.clip(np.asarray(bl)[br], 1e-7, 100.0)), np.log(omegas), np.log(theta)])
    bounds = [(-18.420680744, 6.91)] * len(br) + [(-10.0, 6.91)] * (K + 5)

    def unpack(x):
        b = np.zeros(len(parent))
        b[br] = np.exp(x[:len(br)])
        return b, np.exp(x[len(br):len(br) + K]), np.exp(x[len(br) + K:])

    def f(x):
        b, om, th = unpack(x)
        r = eng.eval(b, om, weights, th)
        g = np.concatenate([r["grad_bl"][br] * b[br], r["grad_omega"] * om, r["grad_theta"] * th])
        return -r["log_likelihood"], -g

    res = minimize(f, x0, jac=True, method="L-BFGS-B", bounds=bounds,
                   options={"maxcor": 8, "maxiter": 2000, "maxfun": max_evals, "ftol": 1e-15, "gtol": 1e-8})
    b, om, th = unpack(res.x)
    return -res.fun, b, om, th
