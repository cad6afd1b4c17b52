"""Host data layer (C++ hyphy::core Alignment / Tree / GeneticCode behind
_hyphy3) against the Python restatement in oracle/data_ref.py, bit-exact on
the reference's own data fixtures and on hand-built Newick edge cases (CPU)."""
import os

import numpy as np
import pytest

from helpers import GOLDEN
from oracle import data_ref as dr

h3 = pytest.importorskip("_hyphy3")

FASTA = ["cd2_reduced.fna", "cd2.fna", "adh.fna"]
NEWICK = ["cd2_reduced.nwk", "cd2.nwk", "adh.nwk"]


def read(name):
    return open(os.path.join(GOLDEN, name)).read()


@pytest.mark.parametrize("name", FASTA)
def test_alignment_matches_restatement(name):
    text = read(name)
    a = h3.Alignment.from_text(text)
    names, seqs = dr.parse_fasta(text)
    r = dr.AlignmentRef(names, seqs)
    assert list(a.taxon_names) == r.names
    assert a.num_codons == r.num_codons and a.num_taxa == r.num_taxa
    assert a.num_patterns == r.num_patterns
    tree = h3.Tree.from_newick("(%s);" % ",".join(names), True)
    arr = h3.problem_arrays(tree, a)
    np.testing.assert_array_equal(arr["states"], r.states)
    np.testing.assert_array_equal(arr["masks"], r.pmasks)
    np.testing.assert_array_equal(arr["weights"], r.weights)
    np.testing.assert_array_equal(np.asarray(a.site_to_pattern), r.site_to_pattern)
    np.testing.assert_allclose(np.asarray(a.nuc_frequencies), r.nuc_freqs, rtol=1e-15, atol=0)
    np.testing.assert_allclose(arr["pos_freqs"], r.pos_freqs, rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(arr["pi"], r.cf3x4, rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(np.asarray(a.codon_frequencies_f1x4), r.f1x4, rtol=1e-13, atol=1e-16)


def assert_same_tree(t, r):
    assert t.num_nodes() == len(r.nodes)
    assert t.root_id == r.root
    assert list(t.post_order) == list(r.post_order)
    for n, m in zip(t.nodes, r.nodes):
        assert n.id == m.id
        assert n.name == m.name
        assert n.parent_id == m.parent
        assert list(n.children) == list(m.children)
        assert n.is_leaf == m.is_leaf
        assert n.branch_length == m.bl
        assert n.model_tag == (m.tag or "")


EDGE_CASES = [
    "(A,B,(C,D));",
    "((A:0.1,B:0.2):0.3,(C:1e-3,D:2.5E-1):0.05);",
    "((A:0.1,B:0.2)inner:0.3,C:0.4,D:0.5,E:0.6);",
    "(('Homo sapiens':0.1,'Pan troglodytes':0.2):0.3,[comment]Gorilla:0.4,Pongo:0.5);",
    "((A{FG}:0.1,B:0.2){FG}:0.3,C:0.4,D:0.5);",
    "(A:0.1,B:0.2,C:0.3,D:0.4,E:0.5,F:0.6);",
    "((((A:1,B:1):1,C:1):1,D:1):1,E:1);\n",
    "(A:0.1,(B:0.2,(C:0.3,(D:0.4,E:0.5):0.6):0.7):0.8);",
]


@pytest.mark.parametrize("keep_rooted", [True, False])
@pytest.mark.parametrize("nwk", [read(n) for n in NEWICK] + EDGE_CASES)
def test_tree_matches_restatement(nwk, keep_rooted):
    assert_same_tree(h3.Tree.from_newick(nwk, keep_rooted), dr.TreeRef.from_newick(nwk, keep_rooted=keep_rooted))


def test_unrooting_counts():
    """Bifurcating root merged away: N = 2n - 2 nodes, B = 2n - 3 branches."""
    nwk = "(((A:1,B:2):3,(C:4,D:5):6):7,((E:8,F:9):10,G:11):12);"
    t = h3.Tree.from_newick(nwk, False)
    n = t.num_leaves()
    assert n == 7 and t.num_nodes() == 2 * n - 2
    assert sum(t.get_branch_lengths()) == 78.0
    t = h3.Tree.from_newick("((A:0.1,B:0.2):0.3,(C:0.4,D:0.5):0.6);", False)
    assert t.num_nodes() == 6 and len(t.nodes[t.root_id].children) == 3
    assert abs(sum(t.get_branch_lengths()) - 2.1) < 1e-15


def test_leaf_to_taxon_matching():
    """Exact, then case-insensitive, then unique prefix (>= 3 chars)
    (likelihood.hpp:106-178); unmatched leaves map to -1."""
    fasta = ">Human_sapiens\nATGAAA\n>chimp\nATGAAG\n>Gorilla\nATGAGA\n>Pongo\nATGCGA\n"
    a = h3.Alignment.from_text(fasta)
    t = h3.Tree.from_newick("(Human:0.1,CHIMP:0.2,Gorilla:0.3,Macaca:0.4);", True)
    arr = h3.problem_arrays(t, a)
    names, seqs = dr.parse_fasta(fasta)
    r = dr.TreeRef.from_newick("(Human:0.1,CHIMP:0.2,Gorilla:0.3,Macaca:0.4);", keep_rooted=True)
    ref = r.leaf_to_taxon(names)
    got = [int(x) for x in arr["leaf_taxon"]]
    assert got == [(-1 if (x is None or x < 0) else x) if r.nodes[i].is_leaf else -1 for i, x in enumerate(ref)]
    by_name = {t.nodes[i].name: got[i] for i in range(t.num_nodes()) if t.nodes[i].is_leaf}
    assert by_name == {"Human": 0, "CHIMP": 1, "Gorilla": 2, "Macaca": -1}


def test_nexus_and_fasta_agree():
    a = h3.Alignment.from_text(read("adh.fna"))
    b = h3.Alignment.load(os.path.join(GOLDEN, "adh.nex"))
    assert [n.strip("'") for n in b.taxon_names] == list(a.taxon_names)
    ta = h3.Tree.from_newick("(%s);" % ",".join(a.taxon_names), True)
    pa, pb = h3.problem_arrays(ta, a), h3.problem_arrays(ta, b)
    np.testing.assert_array_equal(pa["states"], pb["states"])
    np.testing.assert_array_equal(pa["weights"], pb["weights"])
    np.testing.assert_array_equal(pa["pi"], pb["pi"])


def test_genetic_codes():
    u = h3.GeneticCode.universal()
    assert u.num_sense_codons == 61 and len(u.sense_codons) == 61
    assert list(u.sense_codons) == dr.Code().sense_codons
    assert sorted(u.stop_codons) == ["TAA", "TAG", "TGA"]
    m = h3.GeneticCode.from_name("Vertebrate-mtDNA")
    assert m.num_sense_codons == 60 and "TGA" in m.sense_codons and "AGA" in m.stop_codons
    from oracle import hyoracle as ho
    assert m.num_sense_codons == ho.Code(m.id).S
