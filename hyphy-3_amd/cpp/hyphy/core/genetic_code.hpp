// Genetic codes with HyPhy 3's sense-codon numbering.
//
// Mirrors hyphy::core::GeneticCode / GeneticCodeRegistry
// (include/hyphy/core/genetic_code.hpp:14-384 of the reference): sense codons
// are the non-stop codons in lexicographic ACGT order; the registry holds the
// same 22 NCBI tables under the same ids, names and aliases.  Tables here are
// stored as amino-acid strings in ACGT codon order plus per-code overrides.
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace hyphy::core {

inline constexpr int8_t char_to_nuc(char c) {
    switch (c) {
        case 'A': case 'a': return 0;
        case 'C': case 'c': return 1;
        case 'G': case 'g': return 2;
        case 'T': case 't': case 'U': case 'u': return 3;
        default: return -1;
    }
}

// IUPAC 4-bit masks, bit0 = A, bit1 = C, bit2 = G, bit3 = T (types.hpp:49-69).
inline constexpr uint8_t char_to_iupac_mask(char c) {
    switch (c) {
        case 'A': case 'a': return 0x1;
        case 'C': case 'c': return 0x2;
        case 'G': case 'g': return 0x4;
        case 'T': case 't': case 'U': case 'u': return 0x8;
        case 'R': case 'r': return 0x5;
        case 'Y': case 'y': return 0xA;
        case 'S': case 's': return 0x6;
        case 'W': case 'w': return 0x9;
        case 'K': case 'k': return 0xC;
        case 'M': case 'm': return 0x3;
        case 'B': case 'b': return 0xE;
        case 'D': case 'd': return 0xD;
        case 'H': case 'h': return 0xB;
        case 'V': case 'v': return 0x7;
        default: return 0xF;
    }
}

class GeneticCode {
public:
    struct StepDiff {
        int8_t pos = -1;
        int8_t nuc_from = -1;
        int8_t nuc_to = -1;
        bool synonymous = false;
        int8_t num_diffs = 0;
        int8_t diff_pos[3] = {-1, -1, -1};
        int8_t diff_from[3] = {-1, -1, -1};
        int8_t diff_to[3] = {-1, -1, -1};
    };

    std::string name;
    int id = 0;
    int num_sense_codons = 0;
    std::vector<std::string> sense_codons;
    std::vector<char> sense_codon_aa;
    std::vector<std::string> stop_codons;
    int8_t codon64_to_sense[64];
    int8_t sense_to_codon64[64];
    char codon_aa64[64];
    std::vector<std::vector<StepDiff>> diff_matrix;

    GeneticCode(std::string code_name, int code_id, const std::string &aa64) : name(std::move(code_name)), id(code_id) {
        if (aa64.size() != 64) throw std::invalid_argument("genetic code table must have 64 entries");
        static const char nucs[4] = {'A', 'C', 'G', 'T'};
        std::fill(std::begin(codon64_to_sense), std::end(codon64_to_sense), int8_t(-1));
        std::fill(std::begin(sense_to_codon64), std::end(sense_to_codon64), int8_t(-1));
        for (int c = 0; c < 64; ++c) {
            codon_aa64[c] = aa64[c];
            std::string trip = {nucs[c >> 4], nucs[(c >> 2) & 3], nucs[c & 3]};
            if (aa64[c] == '*') {
                stop_codons.push_back(trip);
                continue;
            }
            codon64_to_sense[c] = int8_t(sense_codons.size());
            sense_to_codon64[sense_codons.size()] = int8_t(c);
            sense_codons.push_back(trip);
            sense_codon_aa.push_back(aa64[c]);
        }
        num_sense_codons = int(sense_codons.size());
        const int S = num_sense_codons;
        diff_matrix.assign(S, std::vector<StepDiff>(S));
        for (int u = 0; u < S; ++u) {
            const int cu = sense_to_codon64[u];
            for (int v = 0; v < S; ++v) {
                if (u == v) continue;
                const int cv = sense_to_codon64[v];
                StepDiff &d = diff_matrix[u][v];
                int n = 0;
                for (int p = 0; p < 3; ++p) {
                    const int sh = 4 - 2 * p;
                    const int a = (cu >> sh) & 3, b = (cv >> sh) & 3;
                    if (a != b) {
                        d.diff_pos[n] = int8_t(p);
                        d.diff_from[n] = int8_t(a);
                        d.diff_to[n] = int8_t(b);
                        ++n;
                    }
                }
                d.num_diffs = int8_t(n);
                d.synonymous = sense_codon_aa[u] == sense_codon_aa[v];
                if (n == 1) {
                    d.pos = d.diff_pos[0];
                    d.nuc_from = d.diff_from[0];
                    d.nuc_to = d.diff_to[0];
                }
            }
        }
    }

    int8_t triplet_to_sense(char c1, char c2, char c3) const {
        const int8_t a = char_to_nuc(c1), b = char_to_nuc(c2), c = char_to_nuc(c3);
        if (a < 0 || b < 0 || c < 0) return -1;
        return codon64_to_sense[16 * a + 4 * b + c];
    }

    bool is_stop_triplet(char c1, char c2, char c3) const {
        const int8_t a = char_to_nuc(c1), b = char_to_nuc(c2), c = char_to_nuc(c3);
        if (a < 0 || b < 0 || c < 0) return false;
        return codon_aa64[16 * a + 4 * b + c] == '*';
    }

    bool is_synonymous(int8_t from, int8_t to) const {
        if (from < 0 || to < 0 || from >= num_sense_codons || to >= num_sense_codons) return false;
        return sense_codon_aa[from] == sense_codon_aa[to];
    }

    uint64_t all_sense_mask() const {
        return num_sense_codons >= 64 ? ~uint64_t(0) : ((uint64_t(1) << num_sense_codons) - 1);
    }

    // Compatible sense codons of an (ambiguous) triplet; all sense codons when
    // nothing resolves (genetic_code.hpp:82-109 semantics).
    uint64_t resolve_triplet_mask(char c1, char c2, char c3) const {
        const uint64_t all = all_sense_mask();
        if (c1 == '!' || c2 == '!' || c3 == '!') return all;
        const uint8_t m1 = char_to_iupac_mask(c1), m2 = char_to_iupac_mask(c2), m3 = char_to_iupac_mask(c3);
        if (m1 == 0xF && m2 == 0xF && m3 == 0xF) return all;
        uint64_t mask = 0;
        for (int i = 0; i < 4; ++i) {
            if (!(m1 >> i & 1)) continue;
            for (int j = 0; j < 4; ++j) {
                if (!(m2 >> j & 1)) continue;
                for (int k = 0; k < 4; ++k) {
                    if (!(m3 >> k & 1)) continue;
                    const int8_t s = codon64_to_sense[16 * i + 4 * j + k];
                    if (s >= 0) mask |= uint64_t(1) << s;
                }
            }
        }
        return mask ? mask : all;
    }

    std::vector<int> resolve_ambiguous_triplet(char c1, char c2, char c3) const {
        const uint64_t m = resolve_triplet_mask(c1, c2, c3);
        std::vector<int> out;
        for (int s = 0; s < num_sense_codons; ++s)
            if (m >> s & 1) out.push_back(s);
        return out;
    }

    static std::shared_ptr<const GeneticCode> from_name(std::string name);
    static std::shared_ptr<const GeneticCode> from_id(int id);
    static std::shared_ptr<const GeneticCode> universal() { return from_id(0); }
    static std::shared_ptr<const GeneticCode> vertebrate_mtdna() { return from_id(1); }
};

class GeneticCodeRegistry {
public:
    static GeneticCodeRegistry &instance() {
        static GeneticCodeRegistry reg;
        return reg;
    }
    std::shared_ptr<const GeneticCode> get(const std::string &name) const {
        auto it = by_name_.find(lower(name));
        if (it == by_name_.end()) throw std::runtime_error("Unknown genetic code: '" + name + "'");
        return it->second;
    }
    std::shared_ptr<const GeneticCode> get(int id) const {
        if (id < 0 || id >= int(by_id_.size())) throw std::runtime_error("Unknown genetic code id " + std::to_string(id));
        return by_id_[id];
    }

private:
    std::unordered_map<std::string, std::shared_ptr<const GeneticCode>> by_name_;
    std::vector<std::shared_ptr<const GeneticCode>> by_id_;

    static std::string lower(std::string s) {
        std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return char(std::tolower(c)); });
        return s;
    }

    struct Entry {
        const char *name;
        const char *overrides;  // "c64:aa,c64:aa,..."
        std::vector<std::string> aliases;
    };

    GeneticCodeRegistry() {
        // Standard code, codons in ACGT-lexicographic order.
        const std::string universal = "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF";
        const std::vector<Entry> entries = {
            {"Universal", "", {"standard", "default", "transl_table=1", "1"}},
            {"Vertebrate-mtDNA", "8:*,10:*,12:M,56:W", {"vertebrate_mtdna", "vertebrate-mtdna", "transl_table=2", "2"}},
            {"Yeast-mtDNA", "12:M,28:T,29:T,30:T,31:T,56:W", {"yeast_mtdna", "transl_table=3", "3"}},
            {"Mold-Protozoan-mtDNA", "56:W", {"mold_protozoan_mtdna", "mold-mtdna", "transl_table=4", "4"}},
            {"Invertebrate-mtDNA", "8:S,10:S,12:M,56:W", {"invertebrate_mtdna", "transl_table=5", "5"}},
            {"Ciliate-Nuclear", "48:Q,50:Q", {"ciliate_nuclear", "ciliate", "transl_table=6", "6"}},
            {"Echinoderm-mtDNA", "0:N,8:S,10:S,56:W", {"echinoderm_mtdna", "transl_table=9", "9"}},
            {"Euplotid-Nuclear", "56:C", {"euplotid_nuclear", "euplotid", "transl_table=10", "10"}},
            {"Alt-Yeast-Nuclear", "30:S", {"alt_yeast_nuclear", "transl_table=12", "12"}},
            {"Ascidian-mtDNA", "8:G,10:G,12:M,56:W", {"ascidian_mtdna", "transl_table=13", "13"}},
            {"Flatworm-mtDNA", "0:N,8:S,10:S,48:Y,56:W", {"flatworm_mtdna", "transl_table=14", "14"}},
            {"Blepharisma-Nuclear", "50:Q", {"blepharisma_nuclear", "transl_table=15", "15"}},
            {"Chlorophycean-mtDNA", "50:L", {"chlorophycean_mtdna", "chlorophycean", "transl_table=16", "16"}},
            {"Trematode-mtDNA", "0:N,8:S,10:S,12:M,56:W", {"trematode_mtdna", "trematode", "transl_table=21", "21"}},
            {"Scenedesmus-obliquus-mtDNA", "50:L,52:*",
             {"scenedesmus_obliquus_mtdna", "scenedesmus-mtdna", "scenedesmus_mtdna", "transl_table=22", "22"}},
            {"Thraustochytrium-mtDNA", "60:*", {"thraustochytrium_mtdna", "thraustochytrium", "transl_table=23", "23"}},
            {"Pterobranchia-mtDNA", "8:S,10:K,56:W", {"pterobranchia_mtdna", "pterobranchia", "transl_table=24", "24"}},
            {"SR1-and-Gracilibacteria", "56:G",
             {"sr1_and_gracilibacteria", "sr1", "gracilibacteria", "transl_table=25", "25"}},
            {"Pachysolen-Nuclear", "30:A", {"pachysolen_nuclear", "pachysolen", "transl_table=26", "26"}},
            {"Mesodinium-Nuclear", "48:Y,50:Y", {"mesodinium_nuclear", "mesodinium", "transl_table=29", "29"}},
            {"Peritrich-Nuclear", "48:E,50:E", {"peritrich_nuclear", "peritrich", "transl_table=30", "30"}},
            {"Cephalodiscidae-mtDNA", "8:S,10:K,48:Y,56:W",
             {"cephalodiscidae_mtdna", "cephalodiscidae", "transl_table=33", "33"}},
        };
        for (size_t id = 0; id < entries.size(); ++id) {
            std::string tbl = universal;
            const std::string ov = entries[id].overrides;
            size_t pos = 0;
            while (pos < ov.size()) {
                size_t comma = ov.find(',', pos);
                if (comma == std::string::npos) comma = ov.size();
                const std::string item = ov.substr(pos, comma - pos);
                const size_t colon = item.find(':');
                tbl[std::stoi(item.substr(0, colon))] = item[colon + 1];
                pos = comma + 1;
            }
            auto gc = std::make_shared<GeneticCode>(entries[id].name, int(id), tbl);
            by_id_.push_back(gc);
            by_name_[lower(entries[id].name)] = gc;
            for (const auto &al : entries[id].aliases) by_name_[lower(al)] = gc;
        }
    }
};

inline std::shared_ptr<const GeneticCode> GeneticCode::from_name(std::string name) {
    return GeneticCodeRegistry::instance().get(name);
}
inline std::shared_ptr<const GeneticCode> GeneticCode::from_id(int id) { return GeneticCodeRegistry::instance().get(id); }

}  // namespace hyphy::core
