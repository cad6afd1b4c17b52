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
//   compute_inside_batch_in_place / compute_inside_outside_batch_in_place
//       with their workspaces                                   :295-762
// plus compute_mixture_gradients for branch-site omega mixtures (BUSTED's
// P_mix, include/hyphy/analyses/busted.hpp:509-659, 1138-1429) with full
// omega / weight / GTR gradients.
//
// All likelihood arithmetic runs on the GPU.  One engine per process is
// created lazily on device 0 and keeps the last problem resident; entries
// from several host threads (the reference's OpenMP tile loops) are
// serialised on it, so a tile loop costs one device evaluation and one
// read-back cache, not one per thread.  MG94Matrix spectra come from a
// second process-wide engine that holds only the model.
#pragma once

#include <array>
#include <cmath>
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "hyphy/core/alignment.hpp"
#include "hyphy/core/tree.hpp"
#include "hyphy_mi355x.h"

namespace hyphy::core {

// Column-major dense matrices with the storage order of the reference's
// Eigen::MatrixXd / MatrixXi workspace fields (element (i, j) at i + j * rows).
// The subset of Eigen's dense interface that the reference's callers of
// these entry points use (busted.hpp:555-620: Zero, +=, scalar products,
// noalias, element access).
template <typename T>
struct DenseMatrix {
    std::vector<T> v;
    long r = 0, c = 0;
    DenseMatrix() = default;
    DenseMatrix(long rows, long cols, T fill = T(0)) : v(size_t(rows) * cols, fill), r(rows), c(cols) {}
    static DenseMatrix Zero(long rows, long cols) { return DenseMatrix(rows, cols); }
    static DenseMatrix Constant(long rows, long cols, T x) { return DenseMatrix(rows, cols, x); }
    static DenseMatrix Identity(long n, long m = -1) {
        DenseMatrix a(n, m < 0 ? n : m);
        for (long i = 0; i < std::min(a.r, a.c); ++i) a(i, i) = T(1);
        return a;
    }
    long rows() const { return r; }
    long cols() const { return c; }
    long size() const { return r * c; }
    void resize(long rows, long cols) {
        if (rows * cols != r * c) v.assign(size_t(rows) * cols, T(0));
        r = rows;
        c = cols;
    }
    DenseMatrix &setZero() {
        std::fill(v.begin(), v.end(), T(0));
        return *this;
    }
    DenseMatrix &setConstant(T x) {
        std::fill(v.begin(), v.end(), x);
        return *this;
    }
    T *data() { return v.data(); }
    const T *data() const { return v.data(); }
    T &operator()(long i, long j) { return v[size_t(i) + size_t(j) * r]; }
    T operator()(long i, long j) const { return v[size_t(i) + size_t(j) * r]; }
    T *col(long j) { return v.data() + size_t(j) * r; }
    const T *col(long j) const { return v.data() + size_t(j) * r; }
    DenseMatrix &noalias() { return *this; }
    DenseMatrix &operator+=(const DenseMatrix &o) {
        if (o.r != r || o.c != c) throw std::invalid_argument("DenseMatrix +=: size mismatch");
        for (size_t k = 0; k < v.size(); ++k) v[k] += o.v[k];
        return *this;
    }
    DenseMatrix &operator-=(const DenseMatrix &o) {
        if (o.r != r || o.c != c) throw std::invalid_argument("DenseMatrix -=: size mismatch");
        for (size_t k = 0; k < v.size(); ++k) v[k] -= o.v[k];
        return *this;
    }
    DenseMatrix &operator*=(T x) {
        for (auto &e : v) e *= x;
        return *this;
    }
    friend DenseMatrix operator*(T x, DenseMatrix a) { return a *= x; }
    friend DenseMatrix operator*(DenseMatrix a, T x) { return a *= x; }
    friend DenseMatrix operator+(DenseMatrix a, const DenseMatrix &b) { return a += b; }
    friend DenseMatrix operator-(DenseMatrix a, const DenseMatrix &b) { return a -= b; }
};
template <typename T>
struct DenseVector {
    std::vector<T> v;
    DenseVector() = default;
    explicit DenseVector(long n, T fill = T(0)) : v(size_t(n), fill) {}
    DenseVector(std::initializer_list<T> l) : v(l) {}
    static DenseVector Zero(long n) { return DenseVector(n); }
    static DenseVector Constant(long n, T x) { return DenseVector(n, x); }
    long size() const { return long(v.size()); }
    long rows() const { return long(v.size()); }
    void resize(long n) { v.resize(size_t(n)); }
    DenseVector &setZero() {
        std::fill(v.begin(), v.end(), T(0));
        return *this;
    }
    DenseVector &setConstant(T x) {
        std::fill(v.begin(), v.end(), x);
        return *this;
    }
    T *data() { return v.data(); }
    const T *data() const { return v.data(); }
    T &operator()(long i) { return v[size_t(i)]; }
    T operator()(long i) const { return v[size_t(i)]; }
    T &operator[](long i) { return v[size_t(i)]; }
    T operator[](long i) const { return v[size_t(i)]; }
};
using Matrix = DenseMatrix<Scalar>;
using IntMatrix = DenseMatrix<int>;
using Vector = DenseVector<Scalar>;

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

// Process-wide engine holding only an MG94 model, for MG94Matrix spectra.
class SpectrumEngine {
public:
    static SpectrumEngine &get() {
        static SpectrumEngine e;
        return e;
    }
    // hy_mg94_spectrum on (code, pos, pi): eigenvalues increasing, U_pi and
    // V_pi row-major S x S
    void solve(int code_id, const std::vector<double> &pos12, const std::vector<double> &pi, double alpha, double beta,
               const std::array<double, 5> &theta, double delta, double psi, std::vector<double> &lam,
               std::vector<double> &U, std::vector<double> &V) {
        std::lock_guard<std::mutex> lock(mu_);
        if (!e_ && hy_create(0, &e_) != HY_OK)
            throw std::runtime_error(std::string("MI355X engine unavailable: ") + hy_last_error(nullptr));
        if (code_id != code_ || pos12 != pos_ || pi != pi_) {
            if (hy_set_model(e_, code_id, pos12.data(), pi.data()) != HY_OK)
                throw std::runtime_error(std::string("MI355X engine: ") + hy_last_error(e_));
            code_ = code_id;
            pos_ = pos12;
            pi_ = pi;
        }
        const size_t S = pi.size();
        lam.assign(S, 0.0);
        U.assign(S * S, 0.0);
        V.assign(S * S, 0.0);
        if (hy_mg94_spectrum(e_, alpha, beta, theta.data(), delta, psi, lam.data(), U.data(), V.data(), nullptr) != HY_OK)
            throw std::runtime_error(std::string("MI355X engine: ") + hy_last_error(e_));
    }
    ~SpectrumEngine() {
        if (e_) hy_destroy(e_);
    }

private:
    std::mutex mu_;
    hy_engine e_ = nullptr;
    int code_ = -1;
    std::vector<double> pos_, pi_;
};

// One MG94 generator (rate_matrix.hpp:120-258) with the reference's fields:
// the host builds Q and scale_factor (rate_matrix.hpp:165-234); the
// spectrum (eigenvalues clamped to <= 0 in increasing order, eigenvectors of
// the symmetrised generator, U_pi, V_pi) is solved on the GPU, and
// transition_matrix(t) = max(U_pi diag(e^{lambda t}) V_pi, 0) is formed on
// the host.  update(..., false) skips the spectrum (the evaluation entries
// build their own on the device).
class MG94Matrix {
public:
    MG94Parameters params;
    Matrix Q;
    Vector pi, sqrt_pi, inv_sqrt_pi;
    Vector eigenvalues;
    Matrix eigenvectors, inv_eigenvectors, U_pi, V_pi;
    std::array<std::array<Scalar, 4>, 3> pos{};
    std::shared_ptr<const GeneticCode> code;
    Scalar scale_factor = 1.0;

    struct SubstitutionRates {
        Scalar rate_single = 0.0;
        Scalar rate_double = 0.0;
        Scalar rate_triple = 0.0;
        Scalar total_rate = 0.0;
        Scalar frac_double = 0.0;
        Scalar frac_triple = 0.0;
    };

    // Flux of single, double and triple nucleotide substitutions
    // (compute_substitution_rates, rate_matrix.hpp:137-157).
    SubstitutionRates compute_substitution_rates(const GeneticCode &gc) const {
        SubstitutionRates sr;
        const int S = gc.num_sense_codons;
        for (int i = 0; i < S; ++i)
            for (int j = 0; j < S; ++j) {
                if (i == j) continue;
                const Scalar flux = pi(i) * Q(i, j);
                const int nd = gc.diff_matrix[i][j].num_diffs;
                if (nd == 1)
                    sr.rate_single += flux;
                else if (nd == 2)
                    sr.rate_double += flux;
                else if (nd == 3)
                    sr.rate_triple += flux;
            }
        sr.total_rate = sr.rate_single + sr.rate_double + sr.rate_triple;
        if (sr.total_rate > 1e-12) {
            sr.frac_double = sr.rate_double / sr.total_rate;
            sr.frac_triple = sr.rate_triple / sr.total_rate;
        }
        return sr;
    }

    void update(const MG94Parameters &p, const std::array<std::array<Scalar, 4>, 3> &pos_nuc_freqs,
                const std::vector<Scalar> &codon_freqs, const GeneticCode &gc = *GeneticCode::universal(),
                bool with_spectrum = true) {
        params = p;
        pos = pos_nuc_freqs;
        code = GeneticCode::from_id(gc.id);
        const int S = gc.num_sense_codons;
        if (int(codon_freqs.size()) != S)
            throw std::invalid_argument("MG94Matrix::update: codon frequencies must have one entry per sense codon");
        pi = Vector(S);
        sqrt_pi = Vector(S);
        inv_sqrt_pi = Vector(S);
        for (int i = 0; i < S; ++i) {
            pi(i) = codon_freqs[i];
            sqrt_pi(i) = std::sqrt(std::max(pi(i), 1e-12));
            inv_sqrt_pi(i) = 1.0 / sqrt_pi(i);
        }
        const double r[4][4] = {{0, p.theta_AC, 1, p.theta_AT},
                                {p.theta_AC, 0, p.theta_CG, p.theta_CT},
                                {1, p.theta_CG, 0, p.theta_GT},
                                {p.theta_AT, p.theta_CT, p.theta_GT, 0}};
        Q = Matrix(S, S);
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
                Q(i, j) = q;
                row += q;
            }
            Q(i, i) = -row;
        }
        double mu = 0.0;
        for (int i = 0; i < S; ++i) mu -= pi(i) * Q(i, i);
        scale_factor = mu;
        if (!with_spectrum) return;
        std::vector<double> pos12(12), lam, U, V;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 4; ++b) pos12[4 * a + b] = pos[a][b];
        SpectrumEngine::get().solve(gc.id, pos12, codon_freqs, p.alpha, p.beta, p.theta(), p.delta, p.psi, lam, U, V);
        eigenvalues = Vector(S);
        eigenvectors = Matrix(S, S);
        inv_eigenvectors = Matrix(S, S);
        U_pi = Matrix(S, S);
        V_pi = Matrix(S, S);
        for (int k = 0; k < S; ++k) eigenvalues(k) = lam[k];
        for (int i = 0; i < S; ++i)
            for (int k = 0; k < S; ++k) {
                U_pi(i, k) = U[size_t(i) * S + k];
                V_pi(k, i) = V[size_t(k) * S + i];
                eigenvectors(i, k) = U_pi(i, k) * sqrt_pi(i);
                inv_eigenvectors(k, i) = eigenvectors(i, k);
            }
    }

    // transition_matrix (rate_matrix.hpp:251-258)
    Matrix transition_matrix(Scalar branch_length) const {
        const long S = eigenvalues.size();
        if (S == 0) throw std::logic_error("MG94Matrix::transition_matrix: update() with the spectrum first");
        std::vector<double> ex(S);
        for (long k = 0; k < S; ++k) ex[k] = std::exp(eigenvalues(k) * branch_length);
        Matrix P(S, S);
        for (long j = 0; j < S; ++j)
            for (long i = 0; i < S; ++i) {
                double acc = 0.0;
                for (long k = 0; k < S; ++k) acc += U_pi(i, k) * ex[k] * V_pi(k, j);
                P(i, j) = acc > 0.0 ? acc : 0.0;
            }
        return P;
    }
};

// Nucleotide GTR model (rate_matrix.hpp:14-104): exchangeabilities with
// A<->G fixed to 1 and the frequencies; the engine builds the normalised
// generator, its spectrum and the transition matrices.
struct GTRParameters {
    Scalar theta_AC = 1.0;
    Scalar theta_AT = 1.0;
    Scalar theta_CG = 1.0;
    Scalar theta_CT = 1.0;
    Scalar theta_GT = 1.0;
    std::array<double, 5> theta() const { return {theta_AC, theta_AT, theta_CG, theta_CT, theta_GT}; }
};

class GTRMatrix {
public:
    GTRParameters params;
    std::array<Scalar, 4> pi{0, 0, 0, 0};
    Scalar scale_factor = 1.0;  // mean rate of the unnormalised generator

    void update(const GTRParameters &p, const std::array<Scalar, 4> &frequencies) {
        params = p;
        pi = frequencies;
        const auto th = p.theta();
        const double r[4][4] = {{0, th[0], 1, th[1]}, {th[0], 0, th[2], th[3]}, {1, th[2], 0, th[4]},
                                {th[1], th[3], th[4], 0}};
        double mu = 0.0;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                if (i != j) mu += pi[i] * r[i][j] * pi[j];
        scale_factor = mu > 0.0 ? mu : 1.0;
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
        std::lock_guard<std::recursive_mutex> lock_(engine_mutex());
        Handle &h = bind(tree, aln, leaf_to_taxon, mg.code ? mg.code->id : code_id(aln));
        ++h.epoch;  // this evaluation replaces the device partials
        set_frame(h, aln, mg.pos, mg.pi.v);
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

    // compute_gtr_log_likelihood (likelihood.hpp:181-254): aln is the
    // nucleotide view (Alignment::to_nucleotide_alignment)
    static Scalar compute_gtr_log_likelihood(const Tree &tree, const Alignment &aln, const GTRMatrix &gtr,
                                             const std::vector<size_t> *leaf_to_taxon = nullptr) {
        std::lock_guard<std::recursive_mutex> lock_(engine_mutex());
        if (aln.num_codons != 0)
            throw std::invalid_argument("compute_gtr_log_likelihood needs the nucleotide view of the alignment");
        Handle &h = bind(tree, aln, leaf_to_taxon, code_id(aln));
        ++h.epoch;
        std::vector<double> bl(tree.num_nodes(), 0.0);
        for (const auto &n : tree.nodes) bl[n.id] = n.branch_length;
        const auto th = gtr.params.theta();
        double ll = 0.0;
        check(h, hy_gtr_eval(h.e, get_precision(), HY_WANT_LOGL, bl.data(), th.data(), gtr.pi.data(), &ll, nullptr,
                             nullptr, nullptr));
        return ll;
    }

    static Scalar compute_tree_log_likelihood_from_P(const Tree &tree, const Alignment &aln,
                                                     const std::vector<std::vector<Scalar>> &P_branches,
                                                     const std::vector<Scalar> &pi, int S,
                                                     const std::vector<size_t> *leaf_to_taxon = nullptr) {
        std::lock_guard<std::recursive_mutex> lock_(engine_mutex());
        Handle &h = bind(tree, aln, leaf_to_taxon, code_id(aln));
        ++h.epoch;  // this evaluation replaces the device partials
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
        std::lock_guard<std::recursive_mutex> lock_(engine_mutex());
        Handle &h = bind(tree, aln, leaf_to_taxon, code_id(aln));
        ++h.epoch;  // this evaluation replaces the device partials
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
        std::lock_guard<std::recursive_mutex> lock_(engine_mutex());
        Handle &h = bind(tree, aln, leaf_to_taxon, code_id(aln));
        ++h.epoch;  // this evaluation replaces the device partials
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
        std::lock_guard<std::recursive_mutex> lock_(engine_mutex());
        if (omegas.empty() || omegas.size() != weights.size())
            throw std::invalid_argument("compute_mixture_gradients: omegas and weights must be non-empty and equal length");
        Handle &h = bind(tree, aln, leaf_to_taxon, code_id(aln));
        ++h.epoch;  // this evaluation replaces the device partials
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

    // ---- workspace entries (read back from the device) ----
    // Field names and meaning of the reference's workspaces
    // (likelihood.hpp:295-340): D/V/P_D/U[node] are S x B column-major,
    // scaled by whole powers of 1e150 recorded in node_log_scale
    // (-345.38776394910684 per boost) and boost_counts (the node's own
    // boosts), so that V_v . P_D[v] = likelihood(b) for every branch and
    // log_likelihood(b) = log(likelihood(b)) + node_log_scale(root, b).
    // Boosts are assigned by the reference's end-of-node rule; a node whose
    // partial child product already fell below 1e-150 may carry its boosts
    // one level differently (same true values).
    struct BatchedInsideWorkspace {
        std::vector<Matrix> D;  // [num_nodes] each S x B_cap
        Matrix M_tmp;           // S x B_cap scratch
        Matrix node_log_scale;  // num_nodes x B_cap
        Vector root_L;          // B_cap
        Vector root_log_L;      // B_cap

        void init(size_t num_nodes, int S, size_t B_cap) {
            if (D.size() != num_nodes || (num_nodes > 0 && (D[0].rows() != S || size_t(D[0].cols()) < B_cap))) {
                D.assign(num_nodes, Matrix::Zero(S, long(B_cap)));
                M_tmp.resize(S, long(B_cap));
                node_log_scale.resize(long(num_nodes), long(B_cap));
                root_L.resize(long(B_cap));
                root_log_L.resize(long(B_cap));
            }
        }
    };
    struct BatchedInsideOutsideWorkspace {
        std::vector<Matrix> D, V, P_D, U;  // [num_nodes] each S x B_cap
        IntMatrix boost_counts;            // num_nodes x B_cap
        Matrix node_log_scale;             // num_nodes x B_cap
        std::vector<Matrix> pref, suff;    // scratch of the reference's CPU pass (unused here)
        IntMatrix pref_boost, suff_boost;
        Vector likelihood;                 // B_cap
        Vector log_likelihood;             // B_cap

        void init(size_t num_nodes, int S, size_t B_cap) {
            if (D.size() != num_nodes || (num_nodes > 0 && (D[0].rows() != S || size_t(D[0].cols()) < B_cap))) {
                D.assign(num_nodes, Matrix::Zero(S, long(B_cap)));
                V.assign(num_nodes, Matrix::Zero(S, long(B_cap)));
                P_D.assign(num_nodes, Matrix::Zero(S, long(B_cap)));
                U.assign(num_nodes, Matrix::Zero(S, long(B_cap)));
                boost_counts.resize(long(num_nodes), long(B_cap));
                node_log_scale.resize(long(num_nodes), long(B_cap));
                likelihood.resize(long(B_cap));
                log_likelihood.resize(long(B_cap));
            }
        }
    };

    // compute_inside_batch_in_place (likelihood.hpp:343-476) for patterns
    // [p_start, p_end): P_branches[node] row-major S x S (root ignored).
    static void compute_inside_batch_in_place(const Tree &tree, const Alignment &aln, size_t p_start, size_t p_end,
                                              const std::vector<size_t> &leaf_to_taxon,
                                              const std::vector<std::vector<Scalar>> &P_branches,
                                              const std::vector<Scalar> &pi, int S, BatchedInsideWorkspace &ws) {
        std::lock_guard<std::recursive_mutex> lock_(engine_mutex());
        Readback rb;
        read_back(tree, aln, p_start, p_end, leaf_to_taxon, P_branches, pi, S, false, rb);
        fill(ws.D, rb.D);
        fill(ws.node_log_scale, rb.log_scale);
        fill(ws.root_L, rb.L);
        fill(ws.root_log_L, rb.logL);
    }

    // The reference's form: P_branches[node] an S x S Matrix, bridges unsupported.
    static void compute_inside_batch_in_place(const Tree &tree, const Alignment &aln, size_t p_start, size_t p_end,
                                              const std::vector<size_t> &leaf_to_taxon,
                                              const std::vector<Matrix> &P_branches, const std::vector<Scalar> &pi,
                                              int S, BatchedInsideWorkspace &ws,
                                              const std::vector<std::array<Matrix, 3>> *P_bridges = nullptr) {
        no_bridges(P_bridges);
        compute_inside_batch_in_place(tree, aln, p_start, p_end, leaf_to_taxon, row_major(P_branches, S), pi, S, ws);
    }

    // compute_inside_outside_batch_in_place (likelihood.hpp:479-762).
    static void compute_inside_outside_batch_in_place(const Tree &tree, const Alignment &aln, size_t p_start,
                                                      size_t p_end, const std::vector<size_t> &leaf_to_taxon,
                                                      const std::vector<std::vector<Scalar>> &P_branches,
                                                      const std::vector<Scalar> &pi, int S,
                                                      BatchedInsideOutsideWorkspace &ws) {
        std::lock_guard<std::recursive_mutex> lock_(engine_mutex());
        Readback rb;
        read_back(tree, aln, p_start, p_end, leaf_to_taxon, P_branches, pi, S, true, rb);
        fill(ws.D, rb.D);
        fill(ws.V, rb.V);
        fill(ws.P_D, rb.PD);
        fill(ws.U, rb.U);
        fill(ws.boost_counts, rb.boost);
        fill(ws.node_log_scale, rb.log_scale);
        fill(ws.likelihood, rb.L);
        fill(ws.log_likelihood, rb.logL);
    }

    static void compute_inside_outside_batch_in_place(const Tree &tree, const Alignment &aln, size_t p_start,
                                                      size_t p_end, const std::vector<size_t> &leaf_to_taxon,
                                                      const std::vector<Matrix> &P_branches,
                                                      const std::vector<Scalar> &pi, int S,
                                                      BatchedInsideOutsideWorkspace &ws,
                                                      const std::vector<std::array<Matrix, 3>> *P_bridges = nullptr) {
        no_bridges(P_bridges);
        compute_inside_outside_batch_in_place(tree, aln, p_start, p_end, leaf_to_taxon, row_major(P_branches, S), pi,
                                              S, ws);
    }

    static Scalar compute_tree_log_likelihood_from_P(const Tree &tree, const Alignment &aln,
                                                     const std::vector<Matrix> &P_branches, const Vector &pi, int S,
                                                     const std::vector<size_t> *leaf_to_taxon = nullptr,
                                                     const std::vector<std::array<Matrix, 3>> *P_bridges = nullptr) {
        no_bridges(P_bridges);
        return compute_tree_log_likelihood_from_P(tree, aln, row_major(P_branches, S), pi.v, S, leaf_to_taxon);
    }

private:
    static std::recursive_mutex &engine_mutex() {
        static std::recursive_mutex m;
        return m;
    }
    static void no_bridges(const std::vector<std::array<Matrix, 3>> *P_bridges) {
        if (P_bridges && !P_bridges->empty())
            throw std::invalid_argument("Wheatstone-bridged trees are not supported by the MI355X engine");
    }
    static std::vector<std::vector<Scalar>> row_major(const std::vector<Matrix> &P, int S) {
        std::vector<std::vector<Scalar>> out(P.size());
        for (size_t n = 0; n < P.size(); ++n) {
            if (P[n].rows() != S || P[n].cols() != S) continue;  // the root's entry may be empty
            out[n].resize(size_t(S) * S);
            for (int i = 0; i < S; ++i)
                for (int j = 0; j < S; ++j) out[n][size_t(i) * S + j] = P[n](i, j);
        }
        return out;
    }
    // Copies a read-back into a workspace field, into its first columns when
    // init() gave it a larger capacity.
    template <typename T>
    static void fill(DenseMatrix<T> &dst, const DenseMatrix<T> &src) {
        if (dst.rows() == src.rows() && dst.cols() >= src.cols()) {
            for (long j = 0; j < src.cols(); ++j) std::copy(src.col(j), src.col(j) + src.rows(), dst.col(j));
        } else {
            dst = src;
        }
    }
    static void fill(std::vector<Matrix> &dst, const std::vector<Matrix> &src) {
        if (dst.size() != src.size()) dst.resize(src.size());
        for (size_t n = 0; n < src.size(); ++n) fill(dst[n], src[n]);
    }
    static void fill(Vector &dst, const std::vector<Scalar> &src) {
        if (dst.size() < long(src.size())) dst.resize(long(src.size()));
        std::copy(src.begin(), src.end(), dst.v.begin());
    }

    struct Readback {
        std::vector<Matrix> D, V, PD, U;
        IntMatrix boost;
        Matrix log_scale;
        std::vector<Scalar> L, logL;
    };

    // One device evaluation per distinct (problem, P, pi, pass) serves every
    // pattern range the caller's tile loop asks for.
    static void read_back(const Tree &tree, const Alignment &aln, size_t p_start, size_t p_end,
                          const std::vector<size_t> &leaf_to_taxon, const std::vector<std::vector<Scalar>> &P_branches,
                          const std::vector<Scalar> &pi, int S, bool outside, Readback &rb) {
        constexpr double LOG_SCALER = 345.38776394910684, LOG2_1E150 = 498.28921423310435;
        if (tree.has_wheatstone_bridges())
            throw std::invalid_argument("Wheatstone-bridged trees are not supported by the MI355X engine");
        Handle &h = bind(tree, aln, &leaf_to_taxon, code_id(aln));
        const size_t N = tree.num_nodes();
        if (p_end <= p_start || p_end > aln.patterns.size())
            throw std::out_of_range("compute_inside_*_batch_in_place: pattern range outside the alignment");
        if (size_t(S) != size_t(h.S)) throw std::invalid_argument("compute_inside_*_batch_in_place: S mismatch");
        std::vector<double> flat(N * size_t(S) * S, 0.0);
        for (size_t n = 0; n < N && n < P_branches.size(); ++n)
            if (P_branches[n].size() == size_t(S) * S) std::copy(P_branches[n].begin(), P_branches[n].end(), flat.begin() + n * S * S);
        uint64_t key = 1469598103934665603ull ^ (outside ? 0x9e37ull : 0x51ull) ^ h.key;
        auto mixd = [&](double x) {
            uint64_t u;
            std::memcpy(&u, &x, 8);
            key = (key ^ u) * 1099511628211ull;
        };
        for (double x : flat) mixd(x);
        for (double x : pi) mixd(x);
        mixd(double(get_precision()));
        if (key != h.part_key || h.part_epoch != h.epoch) {
            double ll = 0.0;
            check(h, outside ? hy_inside_outside_from_P(h.e, get_precision(), flat.data(), pi.data(), &ll, nullptr)
                             : hy_loglik_from_P(h.e, get_precision(), flat.data(), pi.data(), &ll, nullptr));
            h.part_key = key;
            h.part_epoch = h.epoch;
        }
        const long B = long(p_end - p_start);
        std::vector<double> Dm(N * B * S, 0.0), Vm(outside ? N * B * S : 0, 0.0);
        std::vector<int32_t> Em(N * B, 0), Gm(outside ? N * B : 0, 0);
        check(h, hy_read_partials(h.e, int(p_start), int(p_end), Dm.data(), Em.data(), outside ? Vm.data() : nullptr,
                                  outside ? Gm.data() : nullptr, nullptr, nullptr));
        // true D of internal nodes = Dm * 2^Em; leaves from the alignment
        // (init_leaf_likelihood, likelihood.hpp:53-87)
        rb.D.assign(N, Matrix(S, B));
        rb.log_scale = Matrix(long(N), B);
        rb.boost = IntMatrix(long(N), B);
        std::vector<long> cum(N * B, 0);  // total boosts of each subtree
        const uint64_t all = (S >= 64) ? ~uint64_t(0) : ((uint64_t(1) << S) - 1);
        auto scaled = [&](double m, double log2f) { return std::ldexp(m * std::exp2(log2f - std::floor(log2f)), int(std::floor(log2f))); };
        for (int nid : tree.post_order) {
            const auto &node = tree.nodes[nid];
            Matrix &Dn = rb.D[nid];
            for (long b = 0; b < B; ++b) {
                const size_t p = p_start + b;
                if (node.is_leaf) {
                    const size_t t = leaf_to_taxon[nid];
                    double *c = Dn.col(b);
                    const auto &pat = aln.patterns[p];
                    const int st = t == size_t(-1) ? -1 : pat.states[t];
                    const uint64_t mk = t == size_t(-1) ? 0 : pat.ambiguous_masks[t];
                    for (int s2 = 0; s2 < S; ++s2)
                        c[s2] = (t == size_t(-1)) ? 1.0
                                : (st >= 0 && st < S) ? (s2 == st ? 1.0 : 0.0)
                                : (mk != 0 && (mk & all) != all) ? ((mk >> s2) & 1 ? 1.0 : 0.0)
                                                                 : 1.0;
                    continue;
                }
                long csum = 0;
                for (auto c : node.children) csum += cum[size_t(c) * B + b];
                const double *m = Dm.data() + (size_t(nid) * B + b) * S;
                double mx = 0.0;
                for (int s2 = 0; s2 < S; ++s2) mx = std::max(mx, m[s2]);
                const int e = Em[size_t(nid) * B + b];
                long boost = 0;
                if (mx > 0.0) {
                    // reference rule: multiply by 1e150 while max < 1e-150
                    const double l2 = std::log2(mx) + e + double(csum) * LOG2_1E150;
                    while (l2 + double(boost) * LOG2_1E150 < -LOG2_1E150) ++boost;
                }
                cum[size_t(nid) * B + b] = csum + boost;
                rb.boost(nid, b) = int(boost);
                rb.log_scale(nid, b) = -double(csum + boost) * LOG_SCALER;
                const double f = double(e) + double(csum + boost) * LOG2_1E150;
                double *c = Dn.col(b);
                for (int s2 = 0; s2 < S; ++s2) c[s2] = scaled(m[s2], f);
            }
        }
        const int root = tree.root_id;
        rb.L.assign(B, 0.0);
        rb.logL.assign(B, -1e20);
        for (long b = 0; b < B; ++b) {
            double L = 0.0;
            for (int s2 = 0; s2 < S; ++s2) L += rb.D[root](s2, b) * pi[s2];
            rb.L[b] = L;
            if (L > 0.0) rb.logL[b] = std::log(L) + rb.log_scale(root, b);
        }
        if (!outside) return;
        // P_D = P D, V from the outside messages (V_v . P_D[v] = L), U = P^T V
        rb.PD.assign(N, Matrix(S, B));
        rb.V.assign(N, Matrix(S, B));
        rb.U.assign(N, Matrix(S, B));
        for (size_t n = 0; n < N; ++n) {
            if (int(n) == root) {
                for (long b = 0; b < B; ++b)
                    for (int s2 = 0; s2 < S; ++s2) rb.U[n](s2, b) = pi[s2];
                continue;
            }
            const double *P = flat.data() + n * S * S;
            for (long b = 0; b < B; ++b) {
                const double *d = rb.D[n].col(b);
                double *pd = rb.PD[n].col(b);
                for (int i = 0; i < S; ++i) {
                    double acc = 0.0;
                    for (int k = 0; k < S; ++k) acc += P[i * S + k] * d[k];
                    pd[i] = acc;
                }
                const double f = double(Gm[n * B + b]) + double(cum[size_t(root) * B + b] - cum[n * B + b]) * LOG2_1E150;
                const double *vm = Vm.data() + (n * B + b) * S;
                double *vv = rb.V[n].col(b), *uu = rb.U[n].col(b);
                for (int s2 = 0; s2 < S; ++s2) vv[s2] = scaled(vm[s2], f);
                for (int j = 0; j < S; ++j) uu[j] = 0.0;
                for (int i = 0; i < S; ++i)
                    for (int j = 0; j < S; ++j) uu[j] += P[i * S + j] * vv[i];
            }
        }
    }

    struct Handle {
        hy_engine e = nullptr;
        const void *tree_ptr = nullptr;
        const void *aln_ptr = nullptr;
        uint64_t key = 0;
        long mixture_key = -1;
        std::vector<double> pos, pi;
        int code = -1, S = 61;
        uint64_t epoch = 0, part_key = 0, part_epoch = ~uint64_t(0);  // read-back cache
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
        static Handle h;  // one per process; callers hold engine_mutex()
        if (!h.e) {
            if (hy_create(0, &h.e) != HY_OK)
                throw std::runtime_error(std::string("MI355X engine unavailable: ") + hy_last_error(nullptr));
        }
        std::vector<size_t> local;
        const std::vector<size_t> &l2t = leaf_to_taxon ? *leaf_to_taxon : (local = build_leaf_to_taxon_map(tree, aln));
        const uint64_t key = fingerprint(tree, aln, l2t);
        // a nucleotide view (no codons) is a 4-state problem
        const int S = aln.num_codons == 0 ? 4 : GeneticCode::from_id(code)->num_sense_codons;
        if (key != h.key || h.code != code || h.S != S) {
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
            h.pos.clear();
            h.code = -1;
            check(h, hy_set_problem(h.e, N, parent.data(), cp.data(), ci.data(), lt.data(), int(aln.num_taxa),
                                    int(aln.patterns.size()), st.data(), mk.data(), w.data(), S));
            h.key = key;
            h.code = code;
            h.S = S;
            h.mixture_key = -1;
            ++h.epoch;
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
