"""Every genetic code of the registry (ids 0..21, genetic_code.hpp:274-370)
against the NCBI translation tables written out here by hand (CPU).

The standard table is NCBI's 64-letter string in TCAG codon order; every other
table is the standard one plus that table's published reassignments (NCBI
"The Genetic Codes", transl_table 2-33).  From these the test derives the
sense codons (ACGT lexicographic order, stops skipped), the stops and, for
every sense pair, the number of differing positions, their positions and
nucleotides and whether the pair is synonymous; the product's tables
(hyphy::core::GeneticCode behind _hyphy3) and the oracle's (or_code_build)
must equal them."""
import itertools

import pytest

from oracle import hyoracle as ho

h3 = pytest.importorskip("_hyphy3")

# NCBI transl_table=1, codons in TCAG order (first base slowest)
STANDARD_TCAG = "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG"

# registry id -> (NCBI table number, reassignments codon -> amino acid)
NCBI = [
    (1, {}),
    (2, {"AGA": "*", "AGG": "*", "ATA": "M", "TGA": "W"}),
    (3, {"ATA": "M", "CTT": "T", "CTC": "T", "CTA": "T", "CTG": "T", "TGA": "W"}),
    (4, {"TGA": "W"}),
    (5, {"AGA": "S", "AGG": "S", "ATA": "M", "TGA": "W"}),
    (6, {"TAA": "Q", "TAG": "Q"}),
    (9, {"AAA": "N", "AGA": "S", "AGG": "S", "TGA": "W"}),
    (10, {"TGA": "C"}),
    (12, {"CTG": "S"}),
    (13, {"AGA": "G", "AGG": "G", "ATA": "M", "TGA": "W"}),
    (14, {"AAA": "N", "AGA": "S", "AGG": "S", "TAA": "Y", "TGA": "W"}),
    (15, {"TAG": "Q"}),
    (16, {"TAG": "L"}),
    (21, {"TGA": "W", "ATA": "M", "AGA": "S", "AGG": "S", "AAA": "N"}),
    (22, {"TCA": "*", "TAG": "L"}),
    (23, {"TTA": "*"}),
    (24, {"AGA": "S", "AGG": "K", "TGA": "W"}),
    (25, {"TGA": "G"}),
    (26, {"CTG": "A"}),
    (29, {"TAA": "Y", "TAG": "Y"}),
    (30, {"TAA": "E", "TAG": "E"}),
    (33, {"AGA": "S", "AGG": "K", "TAA": "Y", "TGA": "W"}),
]


def table(code_id):
    """codon (ACGT string) -> amino-acid letter of registry code `code_id`."""
    tcag = "TCAG"
    aa = {a + b + c: STANDARD_TCAG[16 * tcag.index(a) + 4 * tcag.index(b) + tcag.index(c)]
          for a, b, c in itertools.product(tcag, repeat=3)}
    aa.update(NCBI[code_id][1])
    return aa


def expected(code_id):
    aa = table(code_id)
    codons = ["".join(t) for t in itertools.product("ACGT", repeat=3)]
    sense = [c for c in codons if aa[c] != "*"]
    stops = [c for c in codons if aa[c] == "*"]
    return aa, codons, sense, stops


def pair_diff(a, b):
    pos = [q for q in range(3) if a[q] != b[q]]
    pad = lambda x: tuple(x + [-1] * (3 - len(x)))  # noqa: E731
    return (len(pos), pad(pos), pad(["ACGT".index(a[q]) for q in pos]), pad(["ACGT".index(b[q]) for q in pos]))


def test_standard_table_facts():
    aa = table(0)
    assert [c for c in aa if aa[c] == "*"] == ["TAA", "TAG", "TGA"]
    assert aa["ATG"] == "M" and aa["TGG"] == "W" and aa["TTT"] == "F" and aa["GGG"] == "G"
    assert sum(1 for c in aa if aa[c] == "L") == 6 and sum(1 for c in aa if aa[c] == "S") == 6


@pytest.mark.parametrize("code_id", range(len(NCBI)))
def test_product_code_matches_ncbi(code_id):
    aa, codons, sense, stops = expected(code_id)
    g = h3.GeneticCode.from_id(code_id)
    assert g.id == code_id
    assert h3.GeneticCode.from_name("transl_table=%d" % NCBI[code_id][0]).id == code_id
    assert g.codon_aa64 == "".join(aa[c] for c in codons)
    assert list(g.sense_codons) == sense and g.num_sense_codons == len(sense)
    assert sorted(g.stop_codons) == sorted(stops)
    for u, v in itertools.product(range(len(sense)), repeat=2):
        if u == v:
            continue
        nd, syn, dpos, dfrom, dto = g.diff(u, v)
        assert (nd, dpos, dfrom, dto) == pair_diff(sense[u], sense[v])
        assert syn == (aa[sense[u]] == aa[sense[v]])
        assert g.is_synonymous(u, v) == syn


@pytest.mark.parametrize("code_id", range(len(NCBI)))
def test_oracle_code_matches_ncbi(code_id):
    aa, codons, sense, stops = expected(code_id)
    c = ho.Code(code_id).c
    assert c.S == len(sense)
    assert bytes(c.aa).decode() == "".join(aa[x] for x in codons)
    for k, cod in enumerate(codons):
        assert c.c64_to_sense[k] == (sense.index(cod) if cod in sense else -1)
    for u, v in itertools.product(range(len(sense)), repeat=2):
        if u == v:
            continue
        idx = u * 64 + v
        nd, dpos, dfrom, dto = pair_diff(sense[u], sense[v])
        assert c.nd[idx] == nd
        assert tuple(c.dpos[3 * idx:3 * idx + 3]) == dpos
        assert tuple(c.dfrom[3 * idx:3 * idx + 3]) == dfrom
        assert tuple(c.dto[3 * idx:3 * idx + 3]) == dto
        assert c.syn[idx] == (1 if aa[sense[u]] == aa[sense[v]] else 0)
