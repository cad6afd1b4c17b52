"""Pure-Python restatement of HyPhy 3's host data layer (alignment + tree).

TEST INFRASTRUCTURE ONLY: the checker for the product's C++ data layer and
the input builder for the FP64 CPU oracle.  Only tests/, ``smoke()`` and
bench.py's CPU legs (``cpu_baseline``, the reference-fit and E-step checks)
import it.

Restated reference semantics (paths relative to the reference root):
  * FASTA parsing               include/hyphy/core/alignment.hpp:118-190
  * codon encoding / masks      alignment.hpp:665-723, genetic_code.hpp:82-109
  * nucleotide, F1x4, CF3x4     alignment.hpp:893-1019
  * pattern compression         alignment.hpp:1021-1078
  * Newick parsing / naming     include/hyphy/core/tree.hpp:273-585
  * unrooting                   tree.hpp:591-690
  * post-order                  tree.hpp:881-892
  * leaf -> taxon matching      include/hyphy/core/likelihood.hpp:106-178
"""
from __future__ import annotations

import numpy as np

NUC = {"A": 0, "C": 1, "G": 2, "T": 3, "U": 3}
IUPAC = {
    "A": 0b0001, "C": 0b0010, "G": 0b0100, "T": 0b1000, "U": 0b1000,
    "R": 0b0101, "Y": 0b1010, "S": 0b0110, "W": 0b1001, "K": 0b1100,
    "M": 0b0011, "B": 0b1110, "D": 0b1101, "H": 0b1011, "V": 0b0111,
}

UNIVERSAL_ACGT = "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF"


def char_to_nuc(c: str) -> int:
    return NUC.get(c.upper(), -1)


def iupac_mask(c: str) -> int:
    return IUPAC.get(c.upper(), 0b1111)


class Code:
    """Sense-codon numbering of genetic_code.hpp:127-172 (universal code)."""

    def __init__(self, aa64: str = UNIVERSAL_ACGT):
        self.aa64 = aa64
        self.c64_to_sense = [-1] * 64
        self.sense = []
        for c in range(64):
            if aa64[c] != "*":
                self.c64_to_sense[c] = len(self.sense)
                self.sense.append(c)
        self.S = len(self.sense)
        acgt = "ACGT"
        self.sense_codons = ["".join(acgt[(c >> s) & 3] for s in (4, 2, 0)) for c in self.sense]

    def triplet_to_sense(self, a, b, c):
        n = [char_to_nuc(x) for x in (a, b, c)]
        if min(n) < 0:
            return -1
        return self.c64_to_sense[16 * n[0] + 4 * n[1] + n[2]]

    def triplet_mask(self, a, b, c):
        all_sense = (1 << self.S) - 1 if self.S < 64 else (1 << 64) - 1
        if "!" in (a, b, c):
            return all_sense
        m = [iupac_mask(x) for x in (a, b, c)]
        if m == [15, 15, 15]:
            return all_sense
        mask = 0
        for i in range(4):
            if not m[0] >> i & 1:
                continue
            for j in range(4):
                if not m[1] >> j & 1:
                    continue
                for k in range(4):
                    if not m[2] >> k & 1:
                        continue
                    s = self.c64_to_sense[16 * i + 4 * j + k]
                    if s >= 0:
                        mask |= 1 << s
        return mask if mask else all_sense


def parse_fasta(text: str):
    names, seqs = [], []
    cur_name, cur = None, []
    for line in text.splitlines():
        line = line.rstrip(" \t\r")
        if not line:
            continue
        if line[0] == "#" and cur_name is None and not cur:
            continue
        if line[0] == ">":
            if cur_name is not None:
                names.append(cur_name)
                seqs.append("".join(cur))
                cur = []
            nm = line[1:].lstrip(" \t")
            cur_name = nm.split()[0] if nm.strip() else "unnamed"
        elif line[0] == "(":
            break
        else:
            cur.append(line)
    if cur_name is not None:
        names.append(cur_name)
        seqs.append("".join(cur))
    return names, seqs


class AlignmentRef:
    def __init__(self, names, seqs, code: Code | None = None):
        self.code = code or Code()
        self.names = list(names)
        self.seqs = list(seqs)
        self.num_taxa = len(seqs)
        L = len(seqs[0])
        if any(len(s) != L for s in seqs):
            raise ValueError("Sequences have unequal lengths in alignment")
        if L % 3:
            raise ValueError("Nucleotide alignment length is not a multiple of 3")
        self.num_codons = L // 3
        S = self.code.S
        self.codons = np.full((self.num_taxa, self.num_codons), -1, dtype=np.int8)
        self.masks = np.zeros((self.num_taxa, self.num_codons), dtype=np.uint64)
        for t, s in enumerate(seqs):
            for c in range(self.num_codons):
                a, b, d = s[3 * c], s[3 * c + 1], s[3 * c + 2]
                st = self.code.triplet_to_sense(a, b, d)
                self.codons[t, c] = st
                self.masks[t, c] = (1 << st) if st >= 0 else self.code.triplet_mask(a, b, d)
        self._frequencies(S)
        self._compress()

    def _frequencies(self, S):
        nuc = np.zeros(4)
        pos = np.zeros((3, 4))
        for s in self.seqs:
            for c in range(self.num_codons):
                for p in range(3):
                    n = char_to_nuc(s[3 * c + p])
                    if n >= 0:
                        nuc[n] += 1.0
                        pos[p, n] += 1.0
        self.nuc_freqs = nuc / nuc.sum() if nuc.sum() > 0 else np.zeros(4)
        if (self.nuc_freqs < 1e-8).any():
            self.nuc_freqs = np.maximum(self.nuc_freqs, 1e-8)
            self.nuc_freqs = self.nuc_freqs / self.nuc_freqs.sum()
        posf = np.zeros((3, 4))
        for p in range(3):
            tot = pos[p].sum()
            if tot > 0:
                posf[p] = pos[p] / tot
            if (posf[p] < 1e-8).any():
                posf[p] = np.maximum(posf[p], 1e-8)
                posf[p] = posf[p] / posf[p].sum()
        nucs = [[(c >> 4) & 3, (c >> 2) & 3, c & 3] for c in self.code.sense]
        f1 = np.array([self.nuc_freqs[a] * self.nuc_freqs[b] * self.nuc_freqs[c] for a, b, c in nucs])
        self.f1x4 = f1 / f1.sum()
        # CF3x4 fixed point (alignment.hpp:962-1003)
        cur = posf.copy()
        for _ in range(200):
            prev = cur.copy()
            f = np.array([cur[0, a] * cur[1, b] * cur[2, c] for a, b, c in nucs])
            if f.sum() > 0:
                f = f / f.sum()
            q = np.zeros((3, 4))
            for k, (a, b, c) in enumerate(nucs):
                q[0, a] += f[k]
                q[1, b] += f[k]
                q[2, c] += f[k]
            for p in range(3):
                for b in range(4):
                    cur[p, b] *= posf[p, b] / max(q[p, b], 1e-12)
                s = cur[p].sum()
                if s > 0:
                    cur[p] = cur[p] / s
            if np.abs(cur - prev).max() < 1e-14:
                break
        f = np.array([cur[0, a] * cur[1, b] * cur[2, c] for a, b, c in nucs])
        self.cf3x4 = f / f.sum()
        self.pos_freqs = cur

    def _compress(self):
        seen = {}
        self.pattern_states, self.pattern_masks, self.weights, self.site_to_pattern = [], [], [], []
        for c in range(self.num_codons):
            key = (tuple(int(x) for x in self.codons[:, c]), tuple(int(x) for x in self.masks[:, c]))
            if key not in seen:
                seen[key] = len(self.weights)
                self.pattern_states.append(self.codons[:, c].copy())
                self.pattern_masks.append(self.masks[:, c].copy())
                self.weights.append(0)
            idx = seen[key]
            self.weights[idx] += 1
            self.site_to_pattern.append(idx)
        self.num_patterns = len(self.weights)
        self.states = np.ascontiguousarray(np.array(self.pattern_states, dtype=np.int8).reshape(self.num_patterns, self.num_taxa))
        self.pmasks = np.ascontiguousarray(np.array(self.pattern_masks, dtype=np.uint64).reshape(self.num_patterns, self.num_taxa))
        self.weights = np.array(self.weights, dtype=np.float64)


class NucCode:
    """State space of the nucleotide view: A, C, G, T."""
    S = 4


class NucleotideRef:
    """Alignment::to_nucleotide_alignment (alignment.hpp:832-890): one site per
    nucleotide column; state = char_to_nuc (A/C/G/T/U -> 0..3, else -1), mask
    = IUPAC 4-bit mask (unknown characters: 0b1111); patterns keyed on the
    column's masks, first-occurrence order; weight = count; the codon
    alignment's nucleotide frequencies (alignment.hpp:845)."""

    def __init__(self, aln: "AlignmentRef"):
        self.code = NucCode()
        self.names = list(aln.names)
        self.num_taxa = aln.num_taxa
        self.num_sites = len(aln.seqs[0]) if aln.seqs else 0
        self.nuc_freqs = np.array(aln.nuc_freqs, dtype=float)
        seen = {}
        states, masks, weights, s2p = [], [], [], []
        for site in range(self.num_sites):
            col = tuple(iupac_mask(s[site]) for s in aln.seqs)
            if col not in seen:
                seen[col] = len(weights)
                states.append([char_to_nuc(s[site]) for s in aln.seqs])
                masks.append(list(col))
                weights.append(0)
            k = seen[col]
            weights[k] += 1
            s2p.append(k)
        self.num_patterns = len(weights)
        self.states = np.array(states, dtype=np.int8).reshape(self.num_patterns, self.num_taxa)
        self.pmasks = np.array(masks, dtype=np.uint64).reshape(self.num_patterns, self.num_taxa)
        self.weights = np.array(weights, dtype=np.float64)
        self.site_to_pattern = np.array(s2p, dtype=np.int64)


# --------------------------------------------------------------------------
# Trees
# --------------------------------------------------------------------------
class NodeRef:
    __slots__ = ("id", "name", "parent", "children", "bl", "is_leaf", "tag")

    def __init__(self, nid):
        self.id = nid
        self.name = ""
        self.parent = -1
        self.children = []
        self.bl = 0.0
        self.is_leaf = False
        self.tag = ""


def _stod_prefix(s: str) -> float:
    """std::stod semantics: longest numeric prefix after leading spaces; 0 on failure."""
    s = s.lstrip()
    best = None
    for end in range(len(s), 0, -1):
        try:
            best = float(s[:end])
            break
        except ValueError:
            continue
    if best is None:
        return 0.0
    return best


class TreeRef:
    def __init__(self):
        self.nodes: list[NodeRef] = []
        self.root = -1
        self.post_order: list[int] = []

    @staticmethod
    def from_newick(text: str, keep_rooted: bool = True) -> "TreeRef":
        t = TreeRef()
        s = text
        while s and s[-1] in "; \t\n\r":
            s = s[:-1]
        clean = []
        in_comment = False
        q = ""
        k = 0
        while k < len(s):
            c = s[k]
            if q:
                clean.append(c)
                if c == q:
                    if q == "'" and k + 1 < len(s) and s[k + 1] == "'":
                        k += 1
                        clean.append(s[k])
                    else:
                        q = ""
                k += 1
                continue
            if in_comment:
                if c == "]":
                    in_comment = False
                k += 1
                continue
            if c == "[":
                in_comment = True
            elif c in "'\"":
                q = c
                clean.append(c)
            elif c not in " \t\n\r":
                clean.append(c)
            k += 1
        clean = "".join(clean)

        def create():
            n = NodeRef(len(t.nodes))
            t.nodes.append(n)
            return n.id

        def process(target, tok):
            if not tok:
                return
            tag = ""
            ob = cb = -1
            qq = ""
            i = 0
            while i < len(tok):
                ch = tok[i]
                if qq:
                    if ch == qq:
                        if qq == "'" and i + 1 < len(tok) and tok[i + 1] == "'":
                            i += 1
                        else:
                            qq = ""
                    i += 1
                    continue
                if ch in "'\"":
                    qq = ch
                elif ch == "{" and ob < 0:
                    ob = i
                elif ch == "}" and ob >= 0 and cb < 0:
                    cb = i
                    break
                i += 1
            if ob >= 0 and cb > ob:
                tag = tok[ob + 1:cb]
                tok = tok[:ob] + tok[cb + 1:]
            colon = -1
            qq = ""
            i = 0
            while i < len(tok):
                ch = tok[i]
                if qq:
                    if ch == qq:
                        if qq == "'" and i + 1 < len(tok) and tok[i + 1] == "'":
                            i += 1
                        else:
                            qq = ""
                    i += 1
                    continue
                if ch in "'\"":
                    qq = ch
                elif ch == ":":
                    colon = i
                    break
                i += 1
            if colon >= 0:
                name = tok[:colon]
                length = _stod_prefix(tok[colon + 1:])
            else:
                name, length = tok, 0.0
            if len(name) >= 2 and name[0] == name[-1] and name[0] in "'\"":
                inner = name[1:-1]
                name = inner.replace("''", "'") if name[0] == "'" else inner
            t.nodes[target].name = name
            t.nodes[target].bl = length
            t.nodes[target].tag = tag

        stack = []
        tok = ""
        q = ""
        i = 0
        while i < len(clean):
            c = clean[i]
            if q:
                tok += c
                if c == q:
                    if q == "'" and i + 1 < len(clean) and clean[i + 1] == "'":
                        i += 1
                        tok += clean[i]
                    else:
                        q = ""
                i += 1
                continue
            if c in "'\"":
                q = c
                tok += c
                i += 1
                continue
            if c == ";":
                break
            if c == "(":
                nid = create()
                if stack:
                    p = stack[-1]
                    t.nodes[p].children.append(nid)
                    t.nodes[nid].parent = p
                stack.append(nid)
            elif c in ",)":
                if tok:
                    lid = create()
                    t.nodes[lid].is_leaf = True
                    process(lid, tok)
                    tok = ""
                    if stack:
                        p = stack[-1]
                        t.nodes[p].children.append(lid)
                        t.nodes[lid].parent = p
                if c == ")":
                    if not stack:
                        raise ValueError("Invalid Newick tree: unbalanced parentheses")
                    cur = stack.pop()
                    j = i + 1
                    itok = ""
                    iq = ""
                    while j < len(clean):
                        cj = clean[j]
                        if iq:
                            itok += cj
                            if cj == iq:
                                if iq == "'" and j + 1 < len(clean) and clean[j + 1] == "'":
                                    j += 1
                                    itok += clean[j]
                                else:
                                    iq = ""
                            j += 1
                            continue
                        if cj in "'\"":
                            iq = cj
                            itok += cj
                            j += 1
                            continue
                        if cj in ",)(;":
                            break
                        itok += cj
                        j += 1
                    if iq:
                        raise ValueError("Invalid Newick tree: unterminated quoted label")
                    if itok:
                        process(cur, itok)
                        i = j - 1
            else:
                tok += c
            i += 1
        if q:
            raise ValueError("Invalid Newick tree: unterminated quoted label")
        if stack:
            raise ValueError("Invalid Newick tree: unbalanced parentheses")
        if not t.nodes:
            raise ValueError("Failed to parse Newick tree: empty tree")
        used = {n.name for n in t.nodes if n.is_leaf and n.name}
        counter = 1
        for n in t.nodes:
            if n.parent == -1:
                t.root = n.id
            if not n.is_leaf:
                if not n.name or n.name in used:
                    while True:
                        cand = f"Node{counter}"
                        counter += 1
                        if cand not in used:
                            break
                    n.name = cand
                used.add(n.name)
            elif not n.name:
                while True:
                    cand = f"Leaf{counter}"
                    counter += 1
                    if cand not in used:
                        break
                n.name = cand
                used.add(n.name)
        if not keep_rooted:
            t.unroot_if_bifurcating()
        t.compute_post_order()
        return t

    def compute_post_order(self):
        out = []
        stack = [(self.root, 0)]
        while stack:
            nid, k = stack.pop()
            ch = self.nodes[nid].children
            if k < len(ch):
                stack.append((nid, k + 1))
                stack.append((ch[k], 0))
            else:
                out.append(nid)
        self.post_order = out

    def unroot_if_bifurcating(self):
        if self.root < 0 or len(self.nodes) <= 3:
            return
        r = self.nodes[self.root]
        if len(r.children) != 2:
            return
        c0, c1 = r.children
        n0, n1 = self.nodes[c0], self.nodes[c1]
        if not n0.is_leaf and not n1.is_leaf:
            if not n1.tag and n0.tag:
                absorb, other = c1, c0
            else:
                absorb, other = c0, c1
        elif not n0.is_leaf:
            absorb, other = c0, c1
        elif not n1.is_leaf:
            absorb, other = c1, c0
        else:
            return
        self.nodes[other].bl += self.nodes[absorb].bl
        if not self.nodes[other].tag and self.nodes[absorb].tag:
            self.nodes[other].tag = self.nodes[absorb].tag
        if absorb == c0:
            new_children = list(self.nodes[absorb].children) + [other]
        else:
            new_children = [other] + list(self.nodes[absorb].children)
        r.children = new_children
        old_to_new = {}
        new_nodes = []
        for i, n in enumerate(self.nodes):
            if i == absorb:
                continue
            old_to_new[i] = len(new_nodes)
            n.id = len(new_nodes)
            new_nodes.append(n)
        self.root = old_to_new[self.root]
        for n in new_nodes:
            if n.id == self.root:
                n.parent = -1
                n.bl = 0.0
            n.children = [old_to_new[c] for c in n.children]
            for c in n.children:
                new_nodes[c].parent = n.id
        self.nodes = new_nodes
        self.compute_post_order()

    def leaf_to_taxon(self, names):
        exact = {n: i for i, n in enumerate(names)}
        lower = {n.lower(): i for i, n in enumerate(names)}
        out = [-1] * len(self.nodes)
        leaves = matched = 0
        for n in self.nodes:
            if not n.is_leaf:
                continue
            leaves += 1
            idx = find_taxon_index(n.name, exact, lower)
            out[n.id] = idx
            if idx >= 0:
                matched += 1
        if leaves and not matched:
            raise ValueError("Phylogenetic tree leaves do not match any sequences in alignment!")
        return out


def find_taxon_index(name, exact, lower):
    nm = name
    if len(nm) >= 2 and nm[0] == nm[-1] and nm[0] in "'\"":
        nm = nm[1:-1]
    if nm in exact:
        return exact[nm]
    lo = nm.lower()
    if lo in lower:
        return lower[lo]
    if len(lo) >= 3:
        hits = [i for s, i in lower.items()
                if (len(s) > len(lo) and s.startswith(lo)) or (len(s) >= 3 and len(lo) > len(s) and lo.startswith(s))]
        if len(hits) == 1:
            return hits[0]
    return -1
