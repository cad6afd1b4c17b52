// MEME per-site refit driver (host): MEMEAnalyzer::analyze_pattern with
// every branch tested (meme.hpp:166-648, no bridges) and NelderMeadND
// (nelder_mead.hpp:208-360), written as coroutines that ask for candidate
// log-likelihoods and driven in lockstep, so that one evaluator call serves
// the pending candidates of every pattern.
//
// Each pattern sees the reference's sequence of evaluations.  Points the
// reference evaluates independently of each other's values are asked for as
// one batch: a simplex's start points (its initial point and grid), its
// first Dim perturbed vertices, the Dim points of a shrink step, and the two
// conditional likelihoods of every branch for the EBFs.
#include "hyphy_mi355x.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <coroutine>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

#pragma clang fp contract(off)

namespace {

// A batched minimiser coroutine: yields vectors of points, receives the
// objective at each.
template <typename Point, typename Result> struct BatchMin {
    struct promise_type {
        std::vector<Point> ask;
        std::vector<double> reply;
        Result result{};
        BatchMin get_return_object() { return BatchMin(std::coroutine_handle<promise_type>::from_promise(*this)); }
        std::suspend_always initial_suspend() noexcept { return {}; }
        std::suspend_always final_suspend() noexcept { return {}; }
        void return_value(Result r) { result = r; }
        void unhandled_exception() { std::terminate(); }
        struct Ask {
            promise_type *p;
            bool await_ready() const noexcept { return false; }
            void await_suspend(std::coroutine_handle<>) const noexcept {}
            std::vector<double> await_resume() const { return p->reply; }
        };
        Ask yield_value(std::vector<Point> v) {
            ask = std::move(v);
            return Ask{this};
        }
    };
    std::coroutine_handle<promise_type> h;
    explicit BatchMin(std::coroutine_handle<promise_type> hh) : h(hh) {}
    BatchMin(BatchMin &&o) noexcept : h(std::exchange(o.h, {})) {}
    BatchMin(const BatchMin &) = delete;
    ~BatchMin() {
        if (h) h.destroy();
    }
    bool start() {
        h.resume();
        return !h.done();
    }
    bool step(std::vector<double> f) {
        h.promise().reply = std::move(f);
        h.resume();
        return !h.done();
    }
    const std::vector<Point> &points() const { return h.promise().ask; }
    Result result() const { return h.promise().result; }
};

template <int D> struct NmVertex {
    std::array<double, D> x{};
    double f = 0.0;
};
template <int D> struct NmResult {
    NmVertex<D> best;
    int iterations = 0;
    bool converged = false;
};

// NelderMeadND::minimize (nelder_mead.hpp:219-360): every point is clamped to
// [lb, ub] before it is evaluated; the best of init and the grid (the first
// strictly lower) starts a simplex whose vertex d + 1 moves coordinate d by
// max(0.05, 0.2 |x_d|) (up when that stays within ub, else down); reflection
// 1, expansion 2, outside / inside contraction 0.5, shrink 0.5 toward the
// best; converged when the best and worst vertices differ by less than tol in
// f and in every coordinate.  The vertices are ordered by a stable insertion
// sort, which is what the reference's std::sort of at most six points does.
template <int D>
BatchMin<std::array<double, D>, NmResult<D>> nelder_mead_nd(std::array<double, D> init, std::array<double, D> lb,
                                                           std::array<double, D> ub, double tol, int max_iter,
                                                           std::vector<std::array<double, D>> grid) {
    using X = std::array<double, D>;
    using V = NmVertex<D>;
    auto clamp = [&](X v) {
        for (int d = 0; d < D; ++d) v[d] = std::clamp(v[d], lb[d], ub[d]);
        return v;
    };
    auto sort = [](std::array<V, D + 1> &s) {
        for (int i = 1; i <= D; ++i)
            for (int j = i; j > 0 && s[j].f < s[j - 1].f; --j) std::swap(s[j], s[j - 1]);
    };
    // start: init, then the grid points in order
    std::vector<X> pts{clamp(init)};
    for (const X &g : grid) pts.push_back(clamp(g));
    std::vector<double> f = co_yield pts;
    V best{pts[0], f[0]};
    for (size_t i = 1; i < pts.size(); ++i)
        if (f[i] < best.f) best = V{pts[i], f[i]};
    std::array<V, D + 1> s;
    s[0] = best;
    pts.clear();
    for (int d = 0; d < D; ++d) {
        X p = best.x;
        const double step = std::max(0.05, 0.2 * std::fabs(best.x[d]));
        if (best.x[d] + step <= ub[d])
            p[d] += step;
        else
            p[d] -= step;
        pts.push_back(clamp(p));
    }
    f = co_yield pts;
    for (int d = 0; d < D; ++d) s[d + 1] = V{pts[d], f[d]};
    int iters = 0;
    bool converged = false;
    for (int it = 0; it < max_iter; ++it) {
        iters = it + 1;
        sort(s);
        double dx = 0.0;
        for (int d = 0; d < D; ++d) dx = std::max(dx, std::fabs(s[D].x[d] - s[0].x[d]));
        if (std::fabs(s[D].f - s[0].f) < tol && dx < tol) {
            converged = true;
            break;
        }
        X c{};
        for (int i = 0; i < D; ++i)
            for (int d = 0; d < D; ++d) c[d] += s[i].x[d];
        for (int d = 0; d < D; ++d) c[d] /= (double)D;
        X xr;
        for (int d = 0; d < D; ++d) xr[d] = c[d] + 1.0 * (c[d] - s[D].x[d]);
        xr = clamp(xr);
        const V r{xr, (co_yield std::vector<X>{xr})[0]};
        if (r.f < s[0].f) {
            X xe;
            for (int d = 0; d < D; ++d) xe[d] = c[d] + 2.0 * (r.x[d] - c[d]);
            xe = clamp(xe);
            const V e{xe, (co_yield std::vector<X>{xe})[0]};
            s[D] = (e.f < r.f) ? e : r;
        } else if (r.f < s[D - 1].f) {
            s[D] = r;
        } else {
            const bool outside = r.f < s[D].f;
            const X &toward = outside ? r.x : s[D].x;
            X xc;
            for (int d = 0; d < D; ++d) xc[d] = c[d] + 0.5 * (toward[d] - c[d]);
            xc = clamp(xc);
            const V cv{xc, (co_yield std::vector<X>{xc})[0]};
            const double ref_f = outside ? r.f : s[D].f;
            if (cv.f < ref_f) {
                s[D] = cv;
            } else {
                pts.clear();
                for (int i = 1; i <= D; ++i) {
                    X xs;
                    for (int d = 0; d < D; ++d) xs[d] = s[0].x[d] + 0.5 * (s[i].x[d] - s[0].x[d]);
                    pts.push_back(clamp(xs));
                }
                f = co_yield pts;
                for (int i = 1; i <= D; ++i) s[i] = V{pts[i - 1], f[i - 1]};
            }
        }
    }
    sort(s);
    co_return NmResult<D>{s[0], iters, converged};
}

struct MemeCand {
    double a, w1, p1, bp;
    int32_t node, cls;
};

// One pattern's MEME refit: yields batches of candidates, receives their
// log-likelihoods; the result is the site's row of HY_MEME_FIELDS values.
struct Refit {
    struct promise_type {
        std::vector<MemeCand> ask;
        std::vector<double> reply;
        std::array<double, HY_MEME_FIELDS> result{};
        Refit get_return_object() { return Refit(std::coroutine_handle<promise_type>::from_promise(*this)); }
        std::suspend_always initial_suspend() noexcept { return {}; }
        std::suspend_always final_suspend() noexcept { return {}; }
        void return_value(const std::array<double, HY_MEME_FIELDS> &r) { result = r; }
        void unhandled_exception() { std::terminate(); }
        struct Ask {
            promise_type *p;
            bool await_ready() const noexcept { return false; }
            void await_suspend(std::coroutine_handle<>) const noexcept {}
            std::vector<double> await_resume() const { return p->reply; }
        };
        Ask yield_value(std::vector<MemeCand> v) {
            ask = std::move(v);
            return Ask{this};
        }
    };
    std::coroutine_handle<promise_type> h;
    explicit Refit(std::coroutine_handle<promise_type> hh) : h(hh) {}
    Refit(Refit &&o) noexcept : h(std::exchange(o.h, {})) {}
    Refit(const Refit &) = delete;
    ~Refit() {
        if (h) h.destroy();
    }
};

using X4 = std::array<double, 4>;
using X3 = std::array<double, 3>;
constexpr X4 kLbAlt{1e-4, 0.0, 1e-6, 1e-4};
constexpr X4 kUbAlt{10000.0, 1.0, 1.0 - 1e-6, 100000.0};
constexpr X3 kLbNull{1e-4, 0.0, 1e-6};
constexpr X3 kUbNull{10000.0, 1.0, 1.0 - 1e-6};
constexpr double kNmTol = 1e-5;
constexpr int kNmIter = 800;

// Runs a NelderMeadND coroutine to its end inside a pattern's refit: `to`
// maps a simplex point to a MEME candidate.
#define HY_RUN_NM(nm, to, evals)                                             \
    for (bool more_ = (nm).start(); more_;) {                                \
        std::vector<MemeCand> c_;                                            \
        for (const auto &x_ : (nm).points()) c_.push_back(to(x_));           \
        std::vector<double> r_ = co_yield std::move(c_);                     \
        (evals) += (int)r_.size();                                           \
        for (double &v_ : r_) v_ = -v_;                                      \
        more_ = (nm).step(std::move(r_));                                    \
    }

// MEMEAnalyzer::analyze_pattern, every branch tested (meme.hpp:166-648).
Refit refit_pattern(bool invariant, const double *fel, int num_nodes, int root, double *ebf) {
    std::array<double, HY_MEME_FIELDS> res{};
    if (ebf)
        for (int n = 0; n < num_nodes; ++n) ebf[n] = 1.0;
    res[HY_MEME_FEL_ALPHA] = fel[HY_FEL_ALPHA];
    res[HY_MEME_FEL_BETA] = fel[HY_FEL_BETA];
    if (invariant) {  // meme.hpp:186-214
        res[HY_MEME_P1] = 1.0;
        res[HY_MEME_P_VALUE] = 1.0;
        res[HY_MEME_P_VS_FEL] = 1.0;
        co_return res;
    }
    res[HY_MEME_LOGL_FEL] = fel[HY_FEL_LOGL_ALT];
    int evals = 0;
    // start grid (meme.hpp:280-320)
    const double alpha_fel = std::clamp(fel[HY_FEL_ALPHA], 1e-4, 10000.0);
    const double beta_fel = std::clamp(fel[HY_FEL_BETA], 1e-4, 100000.0);
    const double w1_0 = std::clamp(beta_fel / std::max(alpha_fel, 1e-4), 0.0, 1.0);
    std::vector<X4> grid;
    if (alpha_fel >= beta_fel) {
        grid.push_back({alpha_fel, w1_0, 0.75, beta_fel});
        grid.push_back({alpha_fel, 0.0, 0.75, beta_fel});
        grid.push_back({alpha_fel, 0.0, 0.95, alpha_fel * 0.5});
        grid.push_back({alpha_fel, 0.25, 0.85, alpha_fel * 0.9});
        const double bp1 = std::max(alpha_fel * 1.5, beta_fel * 2.0);
        const double bp2 = std::max(alpha_fel * 2.5, beta_fel * 4.0);
        const double bp3 = std::max(alpha_fel * 4.0, beta_fel * 8.0);
        for (double pp : {0.05, 0.10, 0.20}) {
            const double p1p = 1.0 - pp;
            grid.push_back({alpha_fel, w1_0, p1p, bp1});
            grid.push_back({alpha_fel, w1_0, p1p, bp2});
            grid.push_back({alpha_fel, w1_0, p1p, bp3});
        }
    } else {
        grid.push_back({alpha_fel, w1_0, 0.75, beta_fel});
        grid.push_back({alpha_fel, 0.0, 0.75, beta_fel});
        grid.push_back({alpha_fel, 0.25, 0.50, beta_fel * 2.0});
        grid.push_back({alpha_fel, 0.25, 0.75, beta_fel * 5.0});
        grid.push_back({alpha_fel, 0.10, 0.90, beta_fel * 15.0});
        grid.push_back({alpha_fel, 0.00, 0.95, beta_fel * 30.0});
        grid.push_back({alpha_fel, 1.00, 0.01, std::max(beta_fel, alpha_fel * 1.5)});
    }
    auto alt_cand = [](const X4 &x) { return MemeCand{x[0], x[1], x[2], x[3], -1, 0}; };
    auto null_cand = [](const X3 &x) { return MemeCand{x[0], x[1], x[2], x[0], -1, 0}; };
    // alternative: 4-D Nelder-Mead (meme.hpp:351-367)
    double alpha, omega1, p1, bp;
    {
        auto nm = nelder_mead_nd<4>(grid[0], kLbAlt, kUbAlt, kNmTol, kNmIter, grid);
        HY_RUN_NM(nm, alt_cand, evals);
        const NmResult<4> r = nm.result();
        alpha = r.best.x[0];
        omega1 = r.best.x[1];
        p1 = r.best.x[2];
        bp = r.best.x[3];
        res[HY_MEME_LOGL] = -r.best.f;
    }
    double p_plus = 1.0 - p1;
    if (bp > alpha && p_plus > 1e-6) {
        // null beta_plus = alpha: 3-D Nelder-Mead from its grid (meme.hpp:425-453)
        const double an = std::clamp(fel[HY_FEL_ALPHA_NULL], 1e-4, 50.0);
        std::vector<X3> ng{{alpha, omega1, p1}, {an, 1.0, 0.5}, {an, omega1, p1}, {std::clamp(alpha, 1e-4, 50.0), 1.0, 0.5}};
        for (double ac : {0.01, 0.1, 0.5, 1.0, 2.0, 5.0, 10.0, 20.0}) {
            ng.push_back({ac, 0.5, 0.5});
            ng.push_back({ac, 0.0, 0.9});
        }
        double lnl_null, a_null, w_null, p_null;
        {
            auto nm = nelder_mead_nd<3>(ng[0], kLbNull, kUbNull, kNmTol, kNmIter, ng);
            HY_RUN_NM(nm, null_cand, evals);
            const NmResult<3> r = nm.result();
            lnl_null = -r.best.f;
            a_null = r.best.x[0];
            w_null = r.best.x[1];
            p_null = r.best.x[2];
        }
        if (lnl_null > res[HY_MEME_LOGL] && std::isfinite(lnl_null)) {
            // warm alternative from the null MLE (meme.hpp:432-452)
            std::vector<X4> wg{{a_null, w_null, p_null, a_null}, {a_null, w_null, 0.99, a_null}};
            auto nm = nelder_mead_nd<4>(wg[0], kLbAlt, kUbAlt, kNmTol, kNmIter, wg);
            HY_RUN_NM(nm, alt_cand, evals);
            const NmResult<4> r = nm.result();
            if (-r.best.f > res[HY_MEME_LOGL]) {
                alpha = r.best.x[0];
                omega1 = r.best.x[1];
                p1 = r.best.x[2];
                bp = r.best.x[3];
                p_plus = 1.0 - p1;
                res[HY_MEME_LOGL] = -r.best.f;
            }
        }
        // the test (meme.hpp:455-482)
        const double ll = res[HY_MEME_LOGL];
        if (lnl_null <= -1e10 || ll <= -1e10 || !std::isfinite(lnl_null) || !std::isfinite(ll)) {
            res[HY_MEME_LRT] = 0.0;
            res[HY_MEME_P_VALUE] = 1.0;
        } else if (lnl_null > ll) {
            alpha = a_null;
            omega1 = w_null;
            p1 = p_null;
            bp = a_null;
            p_plus = 1.0 - p_null;
            res[HY_MEME_LOGL] = lnl_null;
            res[HY_MEME_LRT] = 0.0;
            res[HY_MEME_P_VALUE] = 2.0 / 3.0;
        } else if (std::fabs(ll - lnl_null) <= 1e-12) {
            res[HY_MEME_LRT] = 0.0;
            res[HY_MEME_P_VALUE] = 2.0 / 3.0;
        } else {
            const double lrt = 2.0 * (ll - lnl_null);
            res[HY_MEME_LRT] = lrt;
            res[HY_MEME_P_VALUE] =
                (2.0 / 3.0) * (0.45 * std::erfc(std::sqrt(lrt / 2.0)) + 0.55 * std::exp(-lrt / 2.0));
        }
        // branch EBFs at the final rates: each branch alone in class 1 and in
        // class + (meme.hpp:484-611)
        std::vector<MemeCand> ec;
        for (int n = 0; n < num_nodes; ++n) {
            if (n == root) continue;
            ec.push_back(MemeCand{alpha, omega1, p1, bp, n, 0});
            ec.push_back(MemeCand{alpha, omega1, p1, bp, n, 1});
        }
        const std::vector<double> el = co_yield std::move(ec);
        evals += (int)el.size();
        int under = 0;
        size_t q = 0;
        for (int n = 0; n < num_nodes; ++n) {
            if (n == root) continue;
            const double l1 = el[q], lp = el[q + 1];
            q += 2;
            double e;
            if (l1 > -1e10 && lp > -1e10)
                e = std::exp(std::clamp(lp - l1, -50.0, 50.0));
            else if (l1 <= -1e10 && lp > -1e10)
                e = 1e6;
            else
                e = 0.0;
            if (ebf) ebf[n] = e;
            if (e >= 100.0) ++under;
        }
        res[HY_MEME_BRANCHES] = under;
    } else {
        res[HY_MEME_LRT] = 0.0;
        res[HY_MEME_P_VALUE] = 2.0 / 3.0;
    }
    res[HY_MEME_ALPHA] = alpha;
    res[HY_MEME_BETA1] = alpha * omega1;
    res[HY_MEME_P1] = p1;
    res[HY_MEME_BETA_PLUS] = bp;
    res[HY_MEME_P_PLUS] = p_plus;
    // MEME against FEL, two degrees of freedom (meme.hpp:604-611)
    const double dfel = 2.0 * (res[HY_MEME_LOGL] - res[HY_MEME_LOGL_FEL]);
    res[HY_MEME_P_VS_FEL] = dfel > 0.0 ? std::exp(-dfel / 2.0) : 1.0;
    res[HY_MEME_EVALS] = evals;
    co_return res;
}

#undef HY_RUN_NM

}  // namespace

extern "C" int hy_meme_drive(int n_patterns, const uint8_t *invariant, const double *fel, int num_nodes, int root,
                             hy_meme_eval_fn evaluate, void *user, double *out, double *ebf, int *out_calls,
                             int64_t *out_candidates) {
    if (n_patterns < 0 || (n_patterns > 0 && (!invariant || !fel || !out)) || !evaluate || num_nodes < 1 ||
        root < 0 || root >= num_nodes)
        return HY_ERR_ARG;
    std::vector<Refit> co;
    co.reserve(n_patterns);
    std::vector<int> live;
    auto finish = [&](int p) {
        std::memcpy(out + (size_t)p * HY_MEME_FIELDS, co[p].h.promise().result.data(), sizeof(double) * HY_MEME_FIELDS);
    };
    for (int p = 0; p < n_patterns; ++p) {
        co.push_back(refit_pattern(invariant[p] != 0, fel + (size_t)p * HY_FEL_FIELDS, num_nodes, root,
                                   ebf ? ebf + (size_t)p * num_nodes : nullptr));
        co[p].h.resume();
        if (co[p].h.done())
            finish(p);
        else
            live.push_back(p);
    }
    int calls = 0;
    int64_t cands = 0;
    std::vector<int32_t> pat, node, cls;
    std::vector<double> a, w1, p1, bp, ll;
    while (!live.empty()) {
        pat.clear();
        node.clear();
        cls.clear();
        a.clear();
        w1.clear();
        p1.clear();
        bp.clear();
        for (int p : live)
            for (const MemeCand &c : co[p].h.promise().ask) {
                pat.push_back(p);
                a.push_back(c.a);
                w1.push_back(c.w1);
                p1.push_back(c.p1);
                bp.push_back(c.bp);
                node.push_back(c.node);
                cls.push_back(c.cls);
            }
        ll.assign(pat.size(), 0.0);
        const int rc = evaluate(user, (int)pat.size(), pat.data(), a.data(), w1.data(), p1.data(), bp.data(),
                                node.data(), cls.data(), ll.data());
        if (rc != 0) return rc;
        ++calls;
        cands += (int64_t)pat.size();
        size_t off = 0;
        std::vector<int> next;
        next.reserve(live.size());
        for (int p : live) {
            auto &pr = co[p].h.promise();
            const size_t n = pr.ask.size();
            pr.reply.assign(ll.begin() + off, ll.begin() + off + n);
            off += n;
            co[p].h.resume();
            if (co[p].h.done())
                finish(p);
            else
                next.push_back(p);
        }
        live.swap(next);
    }
    if (out_calls) *out_calls = calls;
    if (out_candidates) *out_candidates = cands;
    return HY_OK;
}
