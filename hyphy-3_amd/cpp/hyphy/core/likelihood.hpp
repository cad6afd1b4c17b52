// C++20 facade with the reference's LikelihoodEngine entry names, backed by
// the MI355X C ABI (include/hyphy_mi355x.h).
//
// Reference interface restated (include/hyphy/core/likelihood.hpp of the
// reference; rate matrix parameters from include/hyphy/core/rate_matrix.hpp):
//   compute_mg94_log_likelihood         likelihood.hpp:257-292
//   compute_tree_log_likelihood_from_P  likelihood.hpp:766-806
//   compute_branch_length_gradients     likelihood.hpp:1102-1211
//   compute_joint_frechet_gradients     likelihood.hpp:1213-1594
//   build_leaf_to_taxon_map / find_taxon_index / neumaier_sum  :34-178
// plus compute_mixture_gradients for branch-site omega mixtures (BUSTED's
// P_mix, include/hyphy/analyses/busted.hpp:509-659, 1138-1429) with full
// omega / weight / GTR gradients.
//
// All likelihood arithmetic runs on the GPU; one engine per host thread is
// created lazily on device 0 and keeps the last problem resident.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "hyphy/core/alignment.hpp"
#include "hyphy/core/tree.hpp"
#include "hyphy_mi355x.h"

namespace hyphy::core {

struct MG94Parameters {
    Scalar alpha = 1.0;
    Scalar beta = 1.0;
    Scalar theta_AC = 1.0;
    Scalar theta_AT = 1.0;
    Scalar theta_CG = 1.0;
    Scalar theta_CT = 1.0;
    Scalar theta_GT = 1.0;
    Scalar delta = 0.0;
    Scalar psi = 0.0;
    std::array<double, 5> theta() const { return {theta_AC, theta_AT, theta_CG, theta_CT, theta_GT}; }
};

// Frequencies + parameters of one MG94 generator.  The host builds Q and mu
// (rate_matrix.hpp:165-234); spectra and transition matrices live on the GPU.
class MG94Matrix {
public:
    MG94Parameters params;
    std::vector<Scalar> Q;  // row-major S x S
    std::vector<Scalar> pi;
    std::array<std::array<Scalar, 4>, 3> pos{};
    std::shared_ptr<const GeneticCode> code;
    Scalar scale_factor = 1.0;

    void update(const MG94Parameters &p, const std::array<std::array<Scalar, 4>, 3> &pos_nuc_freqs,
                const std::vector<Scalar> &codon_freqs, const GeneticCode &gc = *GeneticCode::universal()) {
        params = p;
        pos = pos_nuc_freqs;
        pi = codon_freqs;
        code = GeneticCode::from_id(gc.id);
        const int S = gc.num_sense_codons;
        const double r[4][4] = {{0, p.theta_AC, 1, p.theta_AT},
                                {p.theta_AC, 0, p.theta_CG, p.theta_CT},
                                {1, p.theta_CG, 0, p.theta_GT},
                                {p.theta_AT, p.theta_CT, p.theta_GT, 0}};
        Q.assign(size_t(S) * S, 0.0);
        for (int i = 0; i < S; ++i) {
            double row = 0.0;
            for (int j = 0; j < S; ++j) {
                if (i == j) continue;
                const auto &d = gc.diff_matrix[i][j];
                const double rm = d.synonymous ? p.alpha : p.beta;
                double q = 0.0;
                if (d.num_diffs == 1) {
                    q = rm * r[d.diff_from[0]][d.diff_to[0]] * pos[d.diff_pos[0]][d.diff_to[0]];
                } else if (d.num_diffs == 2 && p.delta > 0.0) {
                    q = rm * p.delta * (r[d.diff_from[0]][d.diff_to[0]] * r[d.diff_from[1]][d.diff_to[1]]) *
                        (pos[d.diff_pos[0]][d.diff_to[0]] * pos[d.diff_pos[1]][d.diff_to[1]]);
                } else if (d.num_diffs == 3 && p.psi > 0.0) {
                    q = rm * p.psi *
                        (r[d.diff_from[0]][d.diff_to[0]] * r[d.diff_from[1]][d.diff_to[1]] * r[d.diff_from[2]][d.diff_to[2]]) *
                        (pos[d.diff_pos[0]][d.diff_to[0]] * pos[d.diff_pos[1]][d.diff_to[1]] * pos[d.diff_pos[2]][d.diff_to[2]]);
                }
                Q[size_t(i) * S + j] = q;
                row += q;
            }
            Q[size_t(i) * S + i] = -row;
        }
        double mu = 0.0;
        for (int i = 0; i < S; ++i) mu -= pi[i] * Q[size_t(i) * S + i];
        scale_factor = mu;
    }
};

struct JointMG94Gradients {
    Scalar log_likelihood = 0.0;
    std::vector<Scalar> grad_branches;
    Scalar grad_omega = 0.0;
    Scalar grad_omega_bg = 0.0;
    Scalar grad_theta_AC = 0.0;
    Scalar grad_theta_AT = 0.0;
    Scalar grad_theta_CG = 0.0;
    Scalar grad_theta_CT = 0.0;
    Scalar grad_theta_GT = 0.0;
};

// Branch-site mixture (one class of branches or a test/background split).
struct MixtureGradients {
    Scalar log_likelihood = 0.0;
    std::vector<Scalar> site_log_likelihoods;
    std::vector<Scalar> grad_branches;
    std::vector<Scalar> grad_omegas;
    std::vector<Scalar> grad_weights;
    std::array<Scalar, 5> grad_theta{0, 0, 0, 0, 0};
};

class LikelihoodEngine {
public:
    enum Precision { FP64 = HY_PREC_FP64, FP32 = HY_PREC_FP32 };

    static int &precision_storage() {
        static int p = FP64;
        return p;
    }
    static void set_precision(Precision p) { precision_storage() = p; }
    static Precision get_precision() { return Precision(precision_storage()); }

    static void neumaier_acc(Scalar &sum, Scalar &c, Scalar x) {
        const Scalar t = sum + x;
        if (std::fabs(sum) >= std::fabs(x))
            c += (sum - t) + x;
        else
            c += (x - t) + sum;
        sum = t;
    }
    static Scalar neumaier_sum(const std::vector<Scalar> &v) {
        Scalar s = 0.0, c = 0.0;
        for (Scalar x : v) neumaier_acc(s, c, x);
        return s + c;
    }

    static size_t find_taxon_index(const Alignment &aln, const std::string &name) {
        std::string nm = name;
        if (nm.size() >= 2 && ((nm.front() == '\'' && nm.back() == '\'') || (nm.front() == '"' && nm.back() == '"')))
            nm = nm.substr(1, nm.size() - 2);
        if (auto it = aln.taxon_to_index.find(nm); it != aln.taxon_to_index.end()) return it->second;
        const std::string lo = Alignment::to_lower(nm);
        if (auto it = aln.lower_taxon_to_index.find(lo); it != aln.lower_taxon_to_index.end()) return it->second;
        if (lo.size() >= 3) {
            size_t hit = size_t(-1), count = 0;
            for (const auto &[s, idx] : aln.lower_taxon_to_index) {
                if ((s.size() > lo.size() && s.compare(0, lo.size(), lo) == 0) ||
                    (s.size() >= 3 && lo.size() > s.size() && lo.compare(0, s.size(), s) == 0)) {
                    hit = idx;
                    ++count;
                }
            }
            if (count == 1) return hit;
        }
        return size_t(-1);
    }

    static std::vector<size_t> build_leaf_to_taxon_map(const Tree &tree, const Alignment &aln) {
        std::vector<size_t> m(tree.num_nodes(), size_t(-1));
        size_t leaves = 0, matched = 0;
        for (const auto &n : tree.nodes) {
            if (!n.is_leaf) continue;
            ++leaves;
            m[n.id] = find_taxon_index(aln, n.name);
            if (m[n.id] != size_t(-1)) ++matched;
        }
        if (leaves > 0 && matched == 0)
            throw std::runtime_error("Phylogenetic tree leaves do not match any sequences in alignment! (0/" +
                                     std::to_string(leaves) + " matched)");
        return m;
    }

    // ---- reference entry points ----
    static Scalar compute_mg94_log_likelihood(const Tree &tree, const Alignment &aln, const MG94Matrix &mg,
                                              const std::vector<size_t> *leaf_to_taxon = nullptr,
                                              Scalar branch_scale = 1.0) {
        Handle &h = bind(tree, aln, leaf_to_taxon, mg.code ? mg.code->id : code_id(aln));
        set_frame(h, aln, mg.pos, mg.pi);
        std::vector<double> bl(tree.num_nodes(), 0.0);
        for (const auto &n : tree.nodes) bl[n.id] = n.branch_length * branch_scale;
        single_mixture(h, HY_TIME_RAW);
        double ll = 0.0;
        const double om = mg.params.beta, w = 1.0;
        const auto th = mg.params.theta();
        check(h, hy_eval(h.e, get_precision(), HY_WANT_LOGL, bl.data(), mg.params.alpha, &om, &w, th.data(),
                         mg.params.delta, mg.params.psi, &ll, nullptr, nullptr, nullptr, nullptr, nullptr));
        return ll;
    }

    static Scalar compute_tree_log_likelihood_from_P(const Tree &tree, const Alignment &aln,
                                                     const std::vector<std::vector<Scalar>> &P_branches,
                                                     const std::vector<Scalar> &pi, int S,
                                                     const std::vector<size_t> *leaf_to_taxon = nullptr) {
        Handle &h = bind(tree, aln, leaf_to_taxon, code_id(aln));
        std::vector<double> flat(tree.num_nodes() * size_t(S) * S, 0.0);
        for (size_t n = 0; n < tree.num_nodes() && n < P_branches.size(); ++n)
            if (P_branches[n].size() == size_t(S) * S) std::copy(P_branches[n].begin(), P_branches[n].end(), flat.begin() + n * S * S);
        double ll = 0.0;
        check(h, hy_loglik_from_P(h.e, get_precision(), flat.data(), pi.data(), &ll, nullptr));
        return ll;
    }

    static std::pair<Scalar, std::vector<Scalar>> compute_branch_length_gradients(
        const Tree &tree, const Alignment &aln, const MG94Parameters &params,
        const std::vector<Scalar> &branch_lengths = {}, const std::vector<size_t> *leaf_to_taxon = nullptr) {
        Handle &h = bind(tree, aln, leaf_to_taxon, code_id(aln));
        set_frame(h, aln, aln.pos_nuc_frequencies, aln.codon_frequencies_f3x4);
        single_mixture(h, HY_TIME_RAW);
        const auto bl = lengths(tree, branch_lengths);
        double ll = 0.0, gom = 0.0, gw = 0.0, gth[5];
        std::vector<double> g(tree.num_nodes(), 0.0);
        const double om = params.beta, w = 1.0;
        const auto th = params.theta();
        check(h, hy_eval(h.e, get_precision(), HY_WANT_LOGL | HY_WANT_GRAD, bl.data(), params.alpha, &om, &w,
                         th.data(), params.delta, params.psi, &ll, nullptr, g.data(), &gom, &gw, gth));
        return {ll, g};
    }

    static JointMG94Gradients compute_joint_frechet_gradients(const Tree &tree, const Alignment &aln,
                                                              const MG94Parameters &params,
                                                              const std::vector<Scalar> &branch_lengths = {},
                                                              const std::vector<size_t> *leaf_to_taxon = nullptr,
                                                              const std::vector<bool> *is_test_branch = nullptr,
                                                              Scalar omega_bg = -1.0) {
        Handle &h = bind(tree, aln, leaf_to_taxon, code_id(aln));
        set_frame(h, aln, aln.pos_nuc_frequencies, aln.codon_frequencies_f3x4);
        const size_t N = tree.num_nodes();
        const bool part = is_test_branch && is_test_branch->size() == N && omega_bg >= 0.0;
        std::vector<int32_t> bc(N, 0);
        if (part)
            for (size_t n = 0; n < N; ++n) bc[n] = (*is_test_branch)[n] ? 0 : 1;
        const int32_t cp2[3] = {0, 1, 2}, cm2[2] = {0, 1};
        check(h, hy_set_mixture(h.e, part ? 2 : 1, part ? 2 : 1, bc.data(), cp2, cm2, HY_TIME_CONV));
        h.mixture_key = -1;
        const auto bl = lengths(tree, branch_lengths);
        const double om[2] = {params.beta, part ? omega_bg : params.beta}, w[2] = {1.0, 1.0};
        const auto th = params.theta();
        double ll = 0.0, gom[2] = {0, 0}, gw[2], gth[5];
        JointMG94Gradients r;
        r.grad_branches.assign(N, 0.0);
        check(h, hy_eval(h.e, get_precision(), HY_WANT_LOGL | HY_WANT_GRAD, bl.data(), params.alpha, om, w, th.data(),
                         params.delta, params.psi, &ll, nullptr, r.grad_branches.data(), gom, gw, gth));
        r.log_likelihood = ll;
        r.grad_omega = gom[0];
        r.grad_omega_bg = part ? gom[1] : 0.0;
        r.grad_theta_AC = gth[0];
        r.grad_theta_AT = gth[1];
        r.grad_theta_CG = gth[2];
        r.grad_theta_CT = gth[3];
        r.grad_theta_GT = gth[4];
        return r;
    }

    // Branch-site omega mixture over all branches (one class), BUSTED's
    // M = 1 evaluation with times given directly (time_convention RAW) or
    // as expected substitutions (CONV, mubar = sum_k w_k mu_k).
    static MixtureGradients compute_mixture_gradients(const Tree &tree, const Alignment &aln,
                                                      const MG94Parameters &base, const std::vector<Scalar> &omegas,
                                                      const std::vector<Scalar> &weights,
                                                      const std::vector<Scalar> &branch_lengths = {},
                                                      int time_convention = HY_TIME_RAW, bool want_grad = true,
                                                      const std::vector<size_t> *leaf_to_taxon = nullptr) {
        if (omegas.empty() || omegas.size() != weights.size())
            throw std::invalid_argument("compute_mixture_gradients: omegas and weights must be non-empty and equal length");
        Handle &h = bind(tree, aln, leaf_to_taxon, code_id(aln));
        set_frame(h, aln, aln.pos_nuc_frequencies, aln.codon_frequencies_f3x4);
        const int K = int(omegas.size());
        std::vector<int32_t> cm(K), bc(tree.num_nodes(), 0);
        for (int k = 0; k < K; ++k) cm[k] = k;
        const int32_t cp[2] = {0, K};
        check(h, hy_set_mixture(h.e, K, 1, bc.data(), cp, cm.data(), time_convention));
        h.mixture_key = -1;
        const auto bl = lengths(tree, branch_lengths);
        const auto th = base.theta();
        MixtureGradients r;
        r.site_log_likelihoods.assign(aln.patterns.size(), 0.0);
        r.grad_branches.assign(tree.num_nodes(), 0.0);
        r.grad_omegas.assign(K, 0.0);
        r.grad_weights.assign(K, 0.0);
        double ll = 0.0;
        const unsigned flags = HY_WANT_LOGL | HY_WANT_SITE_LL | (want_grad ? HY_WANT_GRAD : 0u);
        check(h, hy_eval(h.e, get_precision(), flags, bl.data(), base.alpha, omegas.data(), weights.data(), th.data(),
                         base.delta, base.psi, &ll, r.site_log_likelihoods.data(), r.grad_branches.data(),
                         r.grad_omegas.data(), r.grad_weights.data(), r.grad_theta.data()));
        r.log_likelihood = ll;
        return r;
    }

private:
    struct Handle {
        hy_engine e = nullptr;
        const void *tree_ptr = nullptr;
        const void *aln_ptr = nullptr;
        uint64_t key = 0;
        long mixture_key = -1;
        std::vector<double> pos, pi;
        int code = -1;
        ~Handle() {
            if (e) hy_destroy(e);
        }
    };

    static void check(Handle &h, int rc) {
        if (rc != HY_OK) throw std::runtime_error(std::string("MI355X engine: ") + hy_last_error(h.e));
    }

    static int code_id(const Alignment &aln) { return aln.code ? aln.code->id : 0; }

    static std::vector<double> lengths(const Tree &tree, const std::vector<Scalar> &bl) {
        std::vector<double> out(tree.num_nodes(), 0.0);
        for (const auto &n : tree.nodes)
            out[n.id] = (!bl.empty() && size_t(n.id) < bl.size()) ? bl[n.id] : n.branch_length;
        return out;
    }

    static uint64_t fingerprint(const Tree &tree, const Alignment &aln, const std::vector<size_t> &l2t) {
        uint64_t h = 1469598103934665603ull;
        auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
        mix(tree.num_nodes());
        for (const auto &n : tree.nodes) {
            mix(uint64_t(n.parent_id + 1));
            mix(n.children.size());
            for (auto c : n.children) mix(uint64_t(c));
            mix(l2t[n.id]);
        }
        mix(aln.patterns.size());
        mix(aln.num_taxa);
        for (const auto &p : aln.patterns) {
            mix(p.weight);
            for (size_t t = 0; t < p.states.size(); ++t) {
                mix(uint64_t(uint8_t(p.states[t])));
                mix(p.ambiguous_masks[t]);
            }
        }
        return h;
    }

    static Handle &bind(const Tree &tree, const Alignment &aln, const std::vector<size_t> *leaf_to_taxon, int code) {
        thread_local Handle h;
        if (!h.e) {
            if (hy_create(0, &h.e) != HY_OK)
                throw std::runtime_error(std::string("MI355X engine unavailable: ") + hy_last_error(nullptr));
        }
        std::vector<size_t> local;
        const std::vector<size_t> &l2t = leaf_to_taxon ? *leaf_to_taxon : (local = build_leaf_to_taxon_map(tree, aln));
        const uint64_t key = fingerprint(tree, aln, l2t);
        if (key != h.key || h.code != code) {
            const int N = int(tree.num_nodes());
            std::vector<int32_t> parent(N), cp(N + 1, 0), ci, lt(N, -1);
            for (const auto &n : tree.nodes) {
                parent[n.id] = n.parent_id;
                if (n.is_leaf && l2t[n.id] != size_t(-1)) lt[n.id] = int32_t(l2t[n.id]);
            }
            for (int n = 0; n < N; ++n) {
                for (auto c : tree.nodes[n].children) ci.push_back(c);
                cp[n + 1] = int32_t(ci.size());
            }
            std::vector<int8_t> st;
            std::vector<uint64_t> mk;
            std::vector<double> w;
            aln.pattern_arrays(st, mk, w);
            const int S = GeneticCode::from_id(code)->num_sense_codons;
            h.pos.clear();
            h.code = -1;
            check(h, hy_set_problem(h.e, N, parent.data(), cp.data(), ci.data(), lt.data(), int(aln.num_taxa),
                                    int(aln.patterns.size()), st.data(), mk.data(), w.data(), S));
            h.key = key;
            h.code = code;
            h.mixture_key = -1;
        }
        return h;
    }

    static void set_frame(Handle &h, const Alignment &aln, const std::array<std::array<Scalar, 4>, 3> &pos,
                          const std::vector<Scalar> &pi) {
        std::vector<double> p12(12);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 4; ++b) p12[4 * a + b] = pos[a][b];
        if (p12 != h.pos || pi != h.pi) {
            check(h, hy_set_model(h.e, h.code, p12.data(), pi.data()));
            h.pos = p12;
            h.pi = pi;
        }
    }

    static void single_mixture(Handle &h, int conv) {
        if (h.mixture_key == conv) return;
        const int32_t cp[2] = {0, 1}, cm[1] = {0};
        check(h, hy_set_mixture(h.e, 1, 1, nullptr, cp, cm, conv));
        h.mixture_key = conv;
    }
};

}  // namespace hyphy::core
