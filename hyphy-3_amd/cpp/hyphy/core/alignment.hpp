// Codon alignments with HyPhy 3 semantics: parsing, codon encoding, CF3x4
// frequencies and first-occurrence pattern compression.
//
// Mirrors hyphy::core::Alignment (include/hyphy/core/alignment.hpp:19-1079 of
// the reference): same public fields and member names; Eigen vectors are
// replaced by std::array / std::vector of double.
//   parsing           alignment.hpp:118-640 (FASTA, NEXUS MATRIX + TREES, PHYLIP)
//   finalize          alignment.hpp:665-723
//   compute_frequencies alignment.hpp:893-1019 (1e-8 floor, CF3x4 fixed point)
//   compress_patterns alignment.hpp:1021-1078 (first-occurrence order)
#pragma once

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "hyphy/core/genetic_code.hpp"
#include "hyphy/core/nexus.hpp"

namespace hyphy::core {

using Scalar = double;

struct Sequence {
    std::string name;
    std::string raw_seq;
    std::vector<int8_t> codons;
    std::vector<uint64_t> codon_masks;
};

struct SitePattern {
    std::vector<int8_t> states;
    std::vector<uint64_t> ambiguous_masks;
    size_t weight = 0;
    std::vector<size_t> sites;
};

class Alignment {
public:
    std::vector<std::string> taxon_names;
    std::unordered_map<std::string, size_t> taxon_to_index;
    std::unordered_map<std::string, size_t> lower_taxon_to_index;
    std::vector<Sequence> sequences;
    size_t num_taxa = 0;
    size_t num_codons = 0;
    size_t num_nucleotides = 0;
    std::shared_ptr<const GeneticCode> code;
    std::string embedded_tree_newick;
    std::vector<SitePattern> patterns;
    std::vector<size_t> site_to_pattern;
    std::array<Scalar, 4> nuc_frequencies{0, 0, 0, 0};
    std::array<std::array<Scalar, 4>, 3> pos_nuc_frequencies{};  // CF3x4-corrected after finalize
    std::vector<Scalar> codon_frequencies_f1x4;
    std::vector<Scalar> codon_frequencies_f3x4;

    static std::string to_lower(std::string s) {
        std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return char(std::tolower(c)); });
        return s;
    }

    static std::string read_file(const std::string &path) {
        std::ifstream f(path, std::ios::binary);
        if (!f.is_open()) throw std::runtime_error("Could not open alignment file: " + path);
        std::stringstream b;
        b << f.rdbuf();
        return b.str();
    }

    void add_sequence(const std::string &name, const std::string &seq) {
        Sequence s;
        s.name = name;
        s.raw_seq = seq;
        taxon_to_index[name] = taxon_names.size();
        lower_taxon_to_index[to_lower(name)] = taxon_names.size();
        taxon_names.push_back(name);
        sequences.push_back(std::move(s));
    }

    // ---- FASTA (alignment.hpp:118-190) ----
    static Alignment from_fasta_text(const std::string &raw, std::shared_ptr<const GeneticCode> gcode = nullptr,
                                     bool do_finalize = true) {
        Alignment aln;
        std::istringstream in(raw);
        std::string line, cur_name, cur_seq;
        bool have_name = false;
        while (std::getline(in, line)) {
            while (!line.empty() && (line.back() == '\r' || line.back() == ' ' || line.back() == '\t')) line.pop_back();
            if (line.empty()) continue;
            if (line[0] == '#' && !have_name && cur_seq.empty()) continue;
            if (line[0] == '>') {
                if (have_name) {
                    aln.add_sequence(cur_name, cur_seq);
                    cur_seq.clear();
                }
                std::string nm = line.substr(1);
                const size_t a = nm.find_first_not_of(" \t");
                if (a == std::string::npos) {
                    nm = "unnamed";
                } else {
                    nm = nm.substr(a);
                    const size_t b = nm.find_first_of(" \t");
                    if (b != std::string::npos) nm = nm.substr(0, b);
                }
                cur_name = nm;
                have_name = true;
            } else if (line[0] == '(') {
                aln.embedded_tree_newick = line;
                while (aln.embedded_tree_newick.find(';') == std::string::npos && std::getline(in, line)) {
                    while (!line.empty() && (line.back() == '\r' || line.back() == ' ' || line.back() == '\t'))
                        line.pop_back();
                    const size_t a = line.find_first_not_of(" \t");
                    if (a != std::string::npos) aln.embedded_tree_newick += line.substr(a);
                }
                const size_t semi = aln.embedded_tree_newick.find(';');
                if (semi != std::string::npos)
                    aln.embedded_tree_newick.resize(semi + 1);
                else if (!aln.embedded_tree_newick.empty())
                    aln.embedded_tree_newick += ";";
                break;
            } else {
                cur_seq += line;
            }
        }
        if (have_name) aln.add_sequence(cur_name, cur_seq);
        if (do_finalize) aln.finalize(gcode);
        return aln;
    }

    static Alignment from_fasta(const std::string &path, std::shared_ptr<const GeneticCode> gcode = nullptr) {
        return from_fasta_text(read_file(path), gcode);
    }

    // ---- NEXUS MATRIX + embedded TREES (alignment.hpp:192-472) ----
    static Alignment from_nexus_text(const std::string &raw, std::shared_ptr<const GeneticCode> gcode = nullptr,
                                     bool do_finalize = true) {
        Alignment aln;
        const std::string text = nexus::strip_comments(raw);
        size_t mat = nexus::find_ci(text, "MATRIX");
        if (mat == std::string::npos) throw std::runtime_error("NEXUS file does not contain a MATRIX block");
        size_t expected_nchar = 0;
        for (size_t dp = nexus::find_ci(text, "DIMENSIONS"); dp != std::string::npos && dp < mat;
             dp = nexus::find_ci(text, "DIMENSIONS", dp + 10)) {
            const size_t de = text.find(';', dp);
            if (de == std::string::npos || de > mat) continue;
            const std::string blk = text.substr(dp, de - dp);
            const size_t nc = nexus::find_ci(blk, "NCHAR");
            if (nc == std::string::npos) continue;
            size_t eq = blk.find('=', nc + 5);
            if (eq == std::string::npos) continue;
            size_t i = eq + 1, v = 0;
            bool any = false;
            while (i < blk.size() && std::isspace((unsigned char)blk[i])) ++i;
            while (i < blk.size() && std::isdigit((unsigned char)blk[i])) {
                v = v * 10 + size_t(blk[i] - '0');
                any = true;
                ++i;
            }
            if (any && v > 0) expected_nchar = v;
        }
        std::vector<std::string> taxlabels;
        const size_t tl = nexus::find_ci(text, "TAXLABELS");
        if (tl != std::string::npos && tl < mat) {
            const size_t te = text.find(';', tl + 9);
            if (te != std::string::npos && te < mat) {
                const std::string body = text.substr(tl + 9, te - tl - 9);
                size_t i = 0;
                while (i < body.size()) {
                    while (i < body.size() && std::isspace((unsigned char)body[i])) ++i;
                    if (i >= body.size()) break;
                    std::string nm = nexus::read_label(body, i);
                    if (!nm.empty()) taxlabels.push_back(nm);
                }
            }
        }
        bool nolabels = false, interleaved = false;
        for (size_t fp = nexus::find_ci(text, "FORMAT"); fp != std::string::npos && fp < mat;
             fp = nexus::find_ci(text, "FORMAT", fp + 6)) {
            const size_t fe = text.find(';', fp);
            if (fe == std::string::npos || fe > mat) continue;
            const std::string blk = text.substr(fp, fe - fp);
            if (nexus::find_ci(blk, "NOLABELS") != std::string::npos) nolabels = true;
            const size_t il = nexus::find_ci(blk, "INTERLEAVE");
            if (il != std::string::npos) {
                bool no = false;
                const size_t eq = blk.find('=', il + 10);
                if (eq != std::string::npos && eq - il <= 14) {
                    size_t v = eq + 1;
                    while (v < blk.size() && std::isspace((unsigned char)blk[v])) ++v;
                    if (v < blk.size() && std::strchr("NnFf0", blk[v])) no = true;
                }
                if (!no) interleaved = true;
            }
        }
        mat += 6;
        size_t mat_end = std::string::npos;
        {
            char q = 0;
            for (size_t i = mat; i < text.size(); ++i) {
                const char c = text[i];
                if (q) {
                    if (c == q) {
                        if (c == '\'' && i + 1 < text.size() && text[i + 1] == '\'')
                            ++i;
                        else
                            q = 0;
                    } else if (c == '\n' || c == '\r') {
                        q = 0;
                    }
                } else if (c == '\'' || c == '"') {
                    q = c;
                } else if (c == ';') {
                    mat_end = i;
                    break;
                }
            }
        }
        if (mat_end == std::string::npos) throw std::runtime_error("NEXUS MATRIX block is not terminated with ';'");
        const std::string m = text.substr(mat, mat_end - mat);
        size_t i = 0, row = 0;
        const size_t n = m.size();
        while (i < n) {
            while (i < n && std::isspace((unsigned char)m[i])) ++i;
            if (i >= n) break;
            std::string taxon;
            if (nolabels) {
                if (taxlabels.empty()) throw std::runtime_error("NEXUS NOLABELS without TAXLABELS");
                taxon = taxlabels[row % taxlabels.size()];
                ++row;
            } else {
                if (m[i] == '\'' || m[i] == '"') {
                    taxon = nexus::read_label(m, i);
                } else {
                    while (i < n && !std::isspace((unsigned char)m[i])) taxon += m[i++];
                }
                while (i < n && std::isspace((unsigned char)m[i])) ++i;
                if (i >= n) break;
            }
            size_t existing = 0;
            auto ex = aln.taxon_to_index.find(taxon);
            if (ex != aln.taxon_to_index.end() && interleaved) existing = aln.sequences[ex->second].raw_seq.size();
            std::string seq;
            while (i < n) {
                const char c = m[i];
                if (std::isalpha((unsigned char)c) || c == '-' || c == '?' || c == '.' || c == '!' || c == '*') {
                    seq += (c == '.' ? '-' : c);
                    ++i;
                } else if (c == '\n' || c == '\r') {
                    ++i;
                    break;
                } else if (std::isspace((unsigned char)c)) {
                    if (nolabels) {
                        ++i;
                        continue;
                    }
                    size_t la = i;
                    while (la < n && (m[la] == ' ' || m[la] == '\t')) ++la;
                    if (la < n && (m[la] == '\'' || m[la] == '"')) break;
                    if (expected_nchar > 0 && existing + seq.size() >= expected_nchar) break;
                    i = la;
                } else {
                    ++i;
                }
            }
            if (!taxon.empty() && !seq.empty()) {
                auto it = aln.taxon_to_index.find(taxon);
                const bool append = it != aln.taxon_to_index.end() &&
                                    (interleaved || (expected_nchar > 0 &&
                                                     aln.sequences[it->second].raw_seq.size() + seq.size() <= expected_nchar));
                if (!append)
                    aln.add_sequence(taxon, seq);
                else
                    aln.sequences[it->second].raw_seq += seq;
            }
        }
        aln.embedded_tree_newick = nexus::extract_tree_newick(text, mat);
        if (do_finalize) aln.finalize(gcode);
        return aln;
    }

    static Alignment from_nexus(const std::string &path, std::shared_ptr<const GeneticCode> gcode = nullptr) {
        return from_nexus_text(read_file(path), gcode);
    }

    // ---- PHYLIP (alignment.hpp:474-596), sequential and interleaved ----
    static Alignment from_phylip_text(const std::string &raw, std::shared_ptr<const GeneticCode> gcode = nullptr,
                                      bool do_finalize = true) {
        Alignment aln;
        std::istringstream in(raw);
        std::string line;
        size_t ntax = 0, nchar = 0;
        auto header = [&](const std::string &l) {
            std::istringstream is(l);
            long long a = 0, b = 0;
            if (!(is >> a >> b) || a <= 0 || b <= 0) return false;
            std::string rem;
            if (is >> rem && rem.size() > 2) return false;
            ntax = size_t(a);
            nchar = size_t(b);
            return true;
        };
        auto trim = [](std::string &l) {
            while (!l.empty() && std::isspace((unsigned char)l.back())) l.pop_back();
            size_t a = 0;
            while (a < l.size() && std::isspace((unsigned char)l[a])) ++a;
            l = l.substr(a);
        };
        auto seqchars = [](const std::string &l, size_t from) {
            std::string s;
            for (size_t k = from; k < l.size(); ++k) {
                const char c = l[k];
                if (std::isalpha((unsigned char)c) || c == '-' || c == '?' || c == '.' || c == '!' || c == '*')
                    s += (c == '.' ? '-' : c);
            }
            return s;
        };
        while (std::getline(in, line)) {
            trim(line);
            if (!line.empty() && header(line)) break;
        }
        if (!ntax) throw std::runtime_error("Invalid PHYLIP header");
        bool paml = false;
        size_t il = 0;
        while (std::getline(in, line)) {
            trim(line);
            if (line.empty()) continue;
            if (aln.sequences.size() == ntax) {
                const size_t lp = line.find('(');
                if (lp != std::string::npos) {
                    aln.embedded_tree_newick = line.substr(lp);
                    break;
                }
                bool digits = true;
                for (char c : line)
                    if (!std::isdigit((unsigned char)c) && !std::isspace((unsigned char)c)) digits = false;
                if (digits) continue;
                bool done = true;
                for (const auto &s : aln.sequences)
                    if (s.raw_seq.size() < nchar) done = false;
                if (done) continue;
            }
            if (paml && !aln.sequences.empty() && aln.sequences.back().raw_seq.size() < nchar) {
                aln.sequences.back().raw_seq += seqchars(line, 0);
                continue;
            }
            if (aln.sequences.size() < ntax) {
                size_t sp = 0;
                while (sp < line.size() && !std::isspace((unsigned char)line[sp])) ++sp;
                const std::string name = line.substr(0, sp), seq = seqchars(line, sp);
                if (!name.empty() && seq.empty()) {
                    paml = true;
                    aln.add_sequence(name, "");
                } else if (!name.empty()) {
                    aln.add_sequence(name, seq);
                }
            } else {
                const std::string seq = seqchars(line, 0);
                if (seq.empty()) continue;
                size_t guard = 0;
                while (aln.sequences[il].raw_seq.size() >= nchar && guard++ < ntax) il = (il + 1) % ntax;
                aln.sequences[il].raw_seq += seq;
                il = (il + 1) % ntax;
            }
        }
        if (do_finalize) aln.finalize(gcode);
        return aln;
    }

    static Alignment load_text(const std::string &raw, std::shared_ptr<const GeneticCode> gcode = nullptr) {
        std::istringstream in(raw);
        std::string line;
        while (std::getline(in, line)) {
            if (!line.empty() && line.back() == '\r') line.pop_back();
            size_t a = 0;
            while (a < line.size() && std::isspace((unsigned char)line[a])) ++a;
            line = line.substr(a);
            if (line.empty()) continue;
            std::string up = line;
            for (auto &c : up) c = char(std::toupper((unsigned char)c));
            if (line[0] == '#') {
                if (up.rfind("#NEXUS", 0) == 0) return from_nexus_text(raw, gcode);
                continue;
            }
            if (line[0] == '>') return from_fasta_text(raw, gcode);
            std::istringstream is(line);
            long long x = 0, y = 0;
            std::string rem;
            if ((is >> x >> y) && x > 0 && y > 0 && (!(is >> rem) || rem.size() <= 2)) return from_phylip_text(raw, gcode);
            if (up.rfind("BEGIN", 0) == 0) return from_nexus_text(raw, gcode);
            if (line[0] == '(' && line.find(':') != std::string::npos)
                throw std::runtime_error("Input appears to be a Newick tree rather than an alignment");
        }
        return from_fasta_text(raw, gcode);
    }

    static Alignment load(const std::string &path, std::shared_ptr<const GeneticCode> gcode = nullptr) {
        return load_text(read_file(path), gcode);
    }

    // ---- finalize (alignment.hpp:665-723) ----
    void finalize(std::shared_ptr<const GeneticCode> gcode = nullptr) {
        if (gcode)
            code = gcode;
        else if (!code)
            code = GeneticCode::universal();
        if (sequences.empty()) throw std::runtime_error("Empty alignment");
        num_taxa = sequences.size();
        num_nucleotides = sequences[0].raw_seq.size();
        if (taxon_to_index.size() != num_taxa || taxon_names.size() != num_taxa) {
            taxon_names.resize(num_taxa);
            taxon_to_index.clear();
            lower_taxon_to_index.clear();
            for (size_t i = 0; i < num_taxa; ++i) {
                taxon_names[i] = sequences[i].name;
                taxon_to_index[sequences[i].name] = i;
                lower_taxon_to_index[to_lower(sequences[i].name)] = i;
            }
        }
        for (const auto &s : sequences)
            if (s.raw_seq.size() != num_nucleotides) throw std::runtime_error("Sequences have unequal lengths in alignment");
        if (num_nucleotides % 3 != 0)
            throw std::runtime_error("Nucleotide alignment length is not a multiple of 3 (in-frame codons required)");
        num_codons = num_nucleotides / 3;
        for (auto &s : sequences) {
            s.codons.assign(num_codons, -1);
            s.codon_masks.assign(num_codons, 0);
            for (size_t c = 0; c < num_codons; ++c) {
                const char a = s.raw_seq[3 * c], b = s.raw_seq[3 * c + 1], d = s.raw_seq[3 * c + 2];
                const int8_t st = code->triplet_to_sense(a, b, d);
                s.codons[c] = st;
                s.codon_masks[c] = st >= 0 ? (uint64_t(1) << st) : code->resolve_triplet_mask(a, b, d);
            }
        }
        compute_frequencies();
        compress_patterns();
    }

    // Flat pattern-major arrays for the GPU engine.
    // Nucleotide view for the GTR stage (Alignment::to_nucleotide_alignment,
    // alignment.hpp:832-890 of the reference): one site per nucleotide
    // column, state = A/C/G/T index or -1, mask = the IUPAC 4-bit mask (any
    // other character: all four), patterns keyed on the column of masks in
    // first-occurrence order; no codons (num_codons = 0), the nucleotide
    // frequencies of this alignment.  An alignment without codons is its own
    // nucleotide view.
    Alignment to_nucleotide_alignment() const {
        if (num_codons == 0 && !patterns.empty()) return *this;
        Alignment n;
        n.taxon_names = taxon_names;
        n.taxon_to_index = taxon_to_index;
        n.lower_taxon_to_index = lower_taxon_to_index;
        n.num_taxa = num_taxa;
        n.num_nucleotides = num_nucleotides;
        n.num_codons = 0;
        n.code = code;
        n.embedded_tree_newick = embedded_tree_newick;
        n.nuc_frequencies = nuc_frequencies;
        n.sequences.resize(num_taxa);
        for (size_t t = 0; t < num_taxa; ++t) {
            n.sequences[t].name = sequences[t].name;
            n.sequences[t].raw_seq = sequences[t].raw_seq;
        }
        n.site_to_pattern.assign(num_nucleotides, 0);
        // the column's masks identify the pattern (states follow from them)
        std::unordered_map<std::string, size_t> seen;
        std::string col(num_taxa, '\0');
        for (size_t site = 0; site < num_nucleotides; ++site) {
            for (size_t t = 0; t < num_taxa; ++t) col[t] = char('A' + char_to_iupac_mask(sequences[t].raw_seq[site]));
            auto it = seen.find(col);
            if (it != seen.end()) {
                n.patterns[it->second].weight += 1;
                n.patterns[it->second].sites.push_back(site);
                n.site_to_pattern[site] = it->second;
                continue;
            }
            const size_t idx = n.patterns.size();
            seen.emplace(col, idx);
            SitePattern sp;
            sp.states.resize(num_taxa);
            sp.ambiguous_masks.resize(num_taxa);
            for (size_t t = 0; t < num_taxa; ++t) {
                const char ch = sequences[t].raw_seq[site];
                sp.states[t] = char_to_nuc(ch);
                sp.ambiguous_masks[t] = char_to_iupac_mask(ch);
            }
            sp.weight = 1;
            sp.sites.push_back(site);
            n.patterns.push_back(std::move(sp));
            n.site_to_pattern[site] = idx;
        }
        return n;
    }

    void pattern_arrays(std::vector<int8_t> &states, std::vector<uint64_t> &masks, std::vector<double> &weights) const {
        const size_t P = patterns.size();
        states.resize(P * num_taxa);
        masks.resize(P * num_taxa);
        weights.resize(P);
        for (size_t p = 0; p < P; ++p) {
            for (size_t t = 0; t < num_taxa; ++t) {
                states[p * num_taxa + t] = patterns[p].states[t];
                masks[p * num_taxa + t] = patterns[p].ambiguous_masks[t];
            }
            weights[p] = double(patterns[p].weight);
        }
    }

private:
    void compute_frequencies() {
        std::array<Scalar, 4> nuc{0, 0, 0, 0};
        std::array<std::array<Scalar, 4>, 3> pos{};
        for (const auto &s : sequences)
            for (size_t c = 0; c < num_codons; ++c)
                for (int p = 0; p < 3; ++p) {
                    const int8_t n = char_to_nuc(s.raw_seq[3 * c + p]);
                    if (n >= 0) {
                        nuc[n] += 1.0;
                        pos[p][n] += 1.0;
                    }
                }
        const Scalar tot = nuc[0] + nuc[1] + nuc[2] + nuc[3];
        if (tot > 0.0)
            for (int b = 0; b < 4; ++b) nuc_frequencies[b] = nuc[b] / tot;
        floor_renorm(nuc_frequencies);
        for (int p = 0; p < 3; ++p) {
            const Scalar pt = pos[p][0] + pos[p][1] + pos[p][2] + pos[p][3];
            if (pt > 0.0)
                for (int b = 0; b < 4; ++b) pos_nuc_frequencies[p][b] = pos[p][b] / pt;
            floor_renorm(pos_nuc_frequencies[p]);
        }
        const int S = code->num_sense_codons;
        std::vector<std::array<int, 3>> nucs(S);
        for (int i = 0; i < S; ++i)
            for (int p = 0; p < 3; ++p) nucs[i][p] = char_to_nuc(code->sense_codons[i][p]);
        codon_frequencies_f1x4.assign(S, 0.0);
        Scalar s1 = 0.0;
        for (int i = 0; i < S; ++i) {
            codon_frequencies_f1x4[i] = nuc_frequencies[nucs[i][0]] * nuc_frequencies[nucs[i][1]] * nuc_frequencies[nucs[i][2]];
            s1 += codon_frequencies_f1x4[i];
        }
        if (s1 > 0.0)
            for (auto &v : codon_frequencies_f1x4) v /= s1;
        // CF3x4: rescale position frequencies until the sense-codon marginals
        // reproduce the observed ones (alignment.hpp:962-1003).
        auto cur = pos_nuc_frequencies;
        codon_frequencies_f3x4.assign(S, 0.0);
        auto build = [&](const std::array<std::array<Scalar, 4>, 3> &pf) {
            Scalar sum = 0.0;
            for (int i = 0; i < S; ++i) {
                codon_frequencies_f3x4[i] = pf[0][nucs[i][0]] * pf[1][nucs[i][1]] * pf[2][nucs[i][2]];
                sum += codon_frequencies_f3x4[i];
            }
            if (sum > 0.0)
                for (auto &v : codon_frequencies_f3x4) v /= sum;
        };
        for (int it = 0; it < 200; ++it) {
            const auto prev = cur;
            build(cur);
            std::array<std::array<Scalar, 4>, 3> q{};
            for (int i = 0; i < S; ++i)
                for (int p = 0; p < 3; ++p) q[p][nucs[i][p]] += codon_frequencies_f3x4[i];
            Scalar change = 0.0;
            for (int p = 0; p < 3; ++p) {
                for (int b = 0; b < 4; ++b) cur[p][b] *= pos_nuc_frequencies[p][b] / std::max(q[p][b], 1e-12);
                const Scalar s = cur[p][0] + cur[p][1] + cur[p][2] + cur[p][3];
                if (s > 0.0)
                    for (int b = 0; b < 4; ++b) cur[p][b] /= s;
                for (int b = 0; b < 4; ++b) change = std::max(change, std::fabs(cur[p][b] - prev[p][b]));
            }
            if (change < 1e-14) break;
        }
        build(cur);
        pos_nuc_frequencies = cur;
    }

    static void floor_renorm(std::array<Scalar, 4> &f) {
        bool pad = false;
        for (auto &v : f)
            if (v < 1e-8) {
                v = 1e-8;
                pad = true;
            }
        if (pad) {
            const Scalar s = f[0] + f[1] + f[2] + f[3];
            for (auto &v : f) v /= s;
        }
    }

    struct Key {
        std::vector<int8_t> states;
        std::vector<uint64_t> masks;
        bool operator==(const Key &o) const { return states == o.states && masks == o.masks; }
    };
    struct KeyHash {
        size_t operator()(const Key &k) const {
            uint64_t h = 1469598103934665603ull ^ k.states.size();
            for (size_t i = 0; i < k.states.size(); ++i) {
                const uint64_t v = k.states[i] >= 0 ? uint64_t(k.states[i] + 1) : (k.masks[i] * 0x9E3779B97F4A7C15ull);
                h = (h ^ v) * 1099511628211ull;
            }
            return size_t(h);
        }
    };

    void compress_patterns() {
        patterns.clear();
        site_to_pattern.assign(num_codons, 0);
        std::unordered_map<Key, size_t, KeyHash> seen;
        Key key;
        key.states.resize(num_taxa);
        key.masks.resize(num_taxa);
        for (size_t c = 0; c < num_codons; ++c) {
            for (size_t t = 0; t < num_taxa; ++t) {
                key.states[t] = sequences[t].codons[c];
                key.masks[t] = sequences[t].codon_masks[c];
            }
            auto it = seen.find(key);
            if (it == seen.end()) {
                const size_t idx = patterns.size();
                seen.emplace(key, idx);
                SitePattern sp;
                sp.states = key.states;
                sp.ambiguous_masks = key.masks;
                sp.weight = 1;
                sp.sites.push_back(c);
                patterns.push_back(std::move(sp));
                site_to_pattern[c] = idx;
            } else {
                patterns[it->second].weight += 1;
                patterns[it->second].sites.push_back(c);
                site_to_pattern[c] = it->second;
            }
        }
    }
};

}  // namespace hyphy::core
