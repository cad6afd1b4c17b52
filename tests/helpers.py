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
