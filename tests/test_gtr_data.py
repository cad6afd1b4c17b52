"""Nucleotide view of an alignment for the GTR stage
(Alignment::to_nucleotide_alignment, alignment.hpp:832-890): the product's
C++ (_hyphy3) against the oracle's restatement, bit-exact on the reference's
fixtures and on a column with every IUPAC letter, gaps and unknown
characters; both against a brute-force column count (CPU)."""
import os

import numpy as np
import pytest

from helpers import GOLDEN
from oracle import data_ref as dr

h3 = pytest.importorskip("_hyphy3")


def check(text):
    a = h3.Alignment.from_text(text)
    n = a.to_nucleotide_alignment()
    names, seqs = dr.parse_fasta(text)
    r = dr.NucleotideRef(dr.AlignmentRef(names, seqs))
    tree = h3.Tree.from_newick("(%s);" % ",".join(names), True)
    arr = h3.problem_arrays(tree, n)
    assert arr["S"] == 4 and n.num_codons == 0 and n.num_patterns == r.num_patterns
    np.testing.assert_array_equal(arr["states"], r.states)
    np.testing.assert_array_equal(arr["masks"], r.pmasks)
    np.testing.assert_array_equal(arr["weights"], r.weights)
    np.testing.assert_array_equal(np.asarray(n.site_to_pattern), r.site_to_pattern)
    np.testing.assert_array_equal(arr["pi"], np.asarray(a.nuc_frequencies))
    np.testing.assert_allclose(arr["pi"], r.nuc_freqs, rtol=1e-15, atol=0)
    # brute force: distinct mask columns in first-occurrence order with their counts
    cols = {}
    for site in range(len(seqs[0])):
        key = tuple(dr.iupac_mask(s[site]) for s in seqs)
        cols[key] = cols.get(key, 0) + 1
    assert list(cols.values()) == list(r.weights)
    assert [list(k) for k in cols] == r.pmasks.astype(int).tolist()
    return r


@pytest.mark.parametrize("name", ["cd2_reduced.fna", "cd2.fna", "adh.fna"])
def test_nucleotide_view_matches_restatement(name):
    r = check(open(os.path.join(GOLDEN, name)).read())
    if name == "cd2_reduced.fna":
        # tests/test_core.cpp:12-35 frequencies; 18 nucleotide sites
        assert r.num_sites == 18 and r.weights.sum() == 18
        np.testing.assert_allclose(r.nuc_freqs, [0.456140, 0.114035, 0.254386, 0.175439], rtol=1e-4)


def test_nucleotide_view_ambiguity_codes():
    seqs = ["ACGTRYSWKMBDHVN-?.AAA", "ACGTACGTACGTACG?-NAAA", "UCGTRYSWKMBDHVN-?.AAA"]
    text = "".join(">s%d\n%s\n" % (i, s) for i, s in enumerate(seqs))
    r = check(text)
    assert r.states[0].tolist() == [0, 0, 3]  # A, A, U
    # '-', '?', '.' and 'N' are all-ones masks with state -1: columns 15-17 are one pattern
    p = r.site_to_pattern[15]
    assert r.site_to_pattern[16] == p and r.site_to_pattern[17] == p
    assert r.pmasks[p].tolist() == [15, 15, 15] and r.states[p].tolist() == [-1, -1, -1]
    # an ambiguity letter keeps its mask and state -1 (R = A|G)
    q = r.site_to_pattern[4]
    assert r.pmasks[q].tolist() == [0b0101, 0b0001, 0b0101] and r.states[q].tolist() == [-1, 0, -1]
