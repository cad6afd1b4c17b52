// FEL per-site refit driver (host): the reference's per-pattern optimisation
// (FELAnalyzer::analyze_pattern, all branches tested, fel.hpp:326-445, with
// NelderMead2D and Brent1D, nelder_mead.hpp:15-205) written as coroutines
// that ask for candidate log-likelihoods, driven in lockstep so that one
// evaluator call serves the pending candidates of every pattern.
//
// Each pattern sees exactly the reference's sequence of evaluations; the
// start grid and the null grid are asked for as one batch each (their points
// are independent), every simplex and Brent step one candidate at a time.
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

struct Cand {
    double a, b;
};

// A minimiser coroutine: yields points, receives the objective at each.
template <typename Point, typename Result> struct Minimiser {
    struct promise_type {
        Point point{};
        double reply = 0.0;
        Result result{};
        Minimiser get_return_object() {
            return Minimiser(std::coroutine_handle<promise_type>::from_promise(*this));
        }
        std::suspend_always initial_suspend() noexcept { return {}; }
        std::suspend_always final_suspend() noexcept { return {}; }
        void return_value(Result r) { result = r; }
        void unhandled_exception() { std::terminate(); }
        struct Ask {
            promise_type *p;
            bool await_ready() const noexcept { return false; }
            void await_suspend(std::coroutine_handle<>) const noexcept {}
            double await_resume() const noexcept { return p->reply; }
        };
        Ask yield_value(Point pt) {
            point = pt;
            return Ask{this};
        }
    };
    std::coroutine_handle<promise_type> h;
    explicit Minimiser(std::coroutine_handle<promise_type> hh) : h(hh) {}
    Minimiser(Minimiser &&o) noexcept : h(std::exchange(o.h, {})) {}
    Minimiser(const Minimiser &) = delete;
    ~Minimiser() {
        if (h) h.destroy();
    }
    bool start() {
        h.resume();
        return !h.done();
    }
    bool step(double f) {
        h.promise().reply = f;
        h.resume();
        return !h.done();
    }
    Point point() const { return h.promise().point; }
    Result result() const { return h.promise().result; }
};

struct BrentResult {
    double x, f;
    int iterations;
    bool converged;
};
struct Vertex {
    double a, b, f;
};
struct SimplexResult {
    Vertex best;
    int iterations;
    bool converged;
};

// Brent1D::minimize_point (nelder_mead.hpp:24-101): bracketed parabolic
// interpolation with golden-section fallback on [min(ax, cx), max(ax, cx)].
Minimiser<double, BrentResult> brent_1d(double ax, double bx, double cx, double tol, int max_iter) {
    const double cgold = 0.3819660, zeps = 1e-10;
    double a = std::min(ax, cx), b = std::max(ax, cx);
    double v = bx, w = bx, x = bx, e = 0.0;
    double fx = co_yield x;
    double fv = fx, fw = fx;
    int iters = 0;
    bool converged = false;
    for (int it = 0; it < max_iter; ++it) {
        iters = it + 1;
        const double xm = 0.5 * (a + b);
        const double tol1 = tol * std::fabs(x) + zeps, tol2 = 2.0 * tol1;
        if (std::fabs(x - xm) <= tol2 - 0.5 * (b - a)) {
            converged = true;
            break;
        }
        double d = 0.0;
        if (std::fabs(e) > tol1) {
            // parabola through (v, fv), (w, fw), (x, fx)
            const double r = (x - w) * (fx - fv);
            double q = (x - v) * (fx - fw);
            double p = (x - v) * q - (x - w) * r;
            q = 2.0 * (q - r);
            if (q > 0.0) p = -p;
            q = std::fabs(q);
            const double etemp = e;
            e = d;
            if (std::fabs(p) >= std::fabs(0.5 * q * etemp) || p <= q * (a - x) || p >= q * (b - x)) {
                e = (x >= xm) ? a - x : b - x;
                d = cgold * e;
            } else {
                d = p / q;
                const double u = x + d;
                if (u - a < tol2 || b - u < tol2) d = (xm - x >= 0.0) ? std::fabs(tol1) : -std::fabs(tol1);
            }
        } else {
            e = (x >= xm) ? a - x : b - x;
            d = cgold * e;
        }
        const double u = (std::fabs(d) >= tol1) ? x + d : x + ((d >= 0.0) ? std::fabs(tol1) : -std::fabs(tol1));
        const double fu = co_yield u;
        if (fu <= fx) {
            if (u >= x)
                a = x;
            else
                b = x;
            v = w;
            fv = fw;
            w = x;
            fw = fx;
            x = u;
            fx = fu;
        } else {
            if (u < x)
                a = u;
            else
                b = u;
            if (fu <= fw || w == x) {
                v = w;
                fv = fw;
                w = u;
                fw = fu;
            } else if (fu <= fv || v == x || v == w) {
                v = u;
                fv = fu;
            }
        }
    }
    co_return BrentResult{x, fx, iters, converged};
}

// NelderMead2D::minimize (nelder_mead.hpp:121-204): a simplex of three
// vertices clamped to [lb, ub]^2; reflection 1, expansion 2, contraction and
// shrink 0.5.  The vertices are ordered by a stable insertion sort, as the
// reference's sort of three points is.
Minimiser<Cand, SimplexResult> nelder_mead_2d(double init_a, double init_b, double lb, double ub, double tol,
                                              int max_iter) {
    auto clamp = [&](double v) { return std::clamp(v, lb, ub); };
    auto sort3 = [](std::array<Vertex, 3> &s) {
        for (int i = 1; i < 3; ++i)
            for (int j = i; j > 0 && s[j].f < s[j - 1].f; --j) std::swap(s[j], s[j - 1]);
    };
    std::array<Vertex, 3> s;
    const double step_a = std::max(0.1, 0.2 * init_a), step_b = std::max(0.1, 0.2 * init_b);
    const Cand start[3] = {{init_a, init_b}, {init_a + step_a, init_b}, {init_a, init_b + step_b}};
    for (int i = 0; i < 3; ++i) {
        const Cand c{clamp(start[i].a), clamp(start[i].b)};
        const double f = co_yield c;
        s[i] = Vertex{c.a, c.b, f};
    }
    int iters = 0;
    bool converged = false;
    for (int it = 0; it < max_iter; ++it) {
        iters = it + 1;
        sort3(s);
        if (std::fabs(s[2].f - s[0].f) < tol && std::fabs(s[2].a - s[0].a) < tol && std::fabs(s[2].b - s[0].b) < tol) {
            converged = true;
            break;
        }
        const double ca = 0.5 * (s[0].a + s[1].a), cb = 0.5 * (s[0].b + s[1].b);
        Cand c{clamp(ca + 1.0 * (ca - s[2].a)), clamp(cb + 1.0 * (cb - s[2].b))};
        const Vertex r{c.a, c.b, co_yield c};
        if (r.f < s[0].f) {
            c = Cand{clamp(ca + 2.0 * (r.a - ca)), clamp(cb + 2.0 * (r.b - cb))};
            const Vertex e{c.a, c.b, co_yield c};
            s[2] = (e.f < r.f) ? e : r;
        } else if (r.f < s[1].f) {
            s[2] = r;
        } else {
            const bool outside = r.f < s[2].f;
            const Vertex toward = outside ? r : s[2];
            c = Cand{clamp(ca + 0.5 * (toward.a - ca)), clamp(cb + 0.5 * (toward.b - cb))};
            const Vertex cv{c.a, c.b, co_yield c};
            if (cv.f < toward.f) {
                s[2] = cv;
            } else {
                c = Cand{clamp(s[0].a + 0.5 * (s[1].a - s[0].a)), clamp(s[0].b + 0.5 * (s[1].b - s[0].b))};
                s[1] = Vertex{c.a, c.b, co_yield c};
                c = Cand{clamp(s[0].a + 0.5 * (s[2].a - s[0].a)), clamp(s[0].b + 0.5 * (s[2].b - s[0].b))};
                s[2] = Vertex{c.a, c.b, co_yield c};
            }
        }
    }
    sort3(s);
    co_return SimplexResult{s[0], iters, converged};
}

// One pattern's refit: yields batches of candidates (a, b), receives their
// log-likelihoods; the result is the site's row of HY_FEL_FIELDS values.
struct Refit {
    struct promise_type {
        std::vector<Cand> ask;
        std::vector<double> reply;
        std::array<double, HY_FEL_FIELDS> result{};
        Refit get_return_object() { return Refit(std::coroutine_handle<promise_type>::from_promise(*this)); }
        std::suspend_always initial_suspend() noexcept { return {}; }
        std::suspend_always final_suspend() noexcept { return {}; }
        void return_value(const std::array<double, HY_FEL_FIELDS> &r) { result = r; }
        void unhandled_exception() { std::terminate(); }
        struct Ask {
            promise_type *p;
            bool await_ready() const noexcept { return false; }
            void await_suspend(std::coroutine_handle<>) const noexcept {}
            std::vector<double> await_resume() const { return p->reply; }
        };
        Ask yield_value(std::vector<Cand> v) {
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

constexpr Cand kStartGrid[12] = {{0.01, 0.1}, {1.0, 0.1},  {1.0, 0.5},  {1.0, 1.0},   {1.0, 5.0},   {10.0, 0.1},
                                 {0.01, 0.5}, {0.01, 5.0}, {10.0, 0.5}, {10.0, 1.0}, {10.0, 50.0}, {100.0, 1.0}};
constexpr double kNullGrid[7] = {0.01, 0.1, 0.5, 1.0, 2.0, 5.0, 10.0};

// FELAnalyzer::analyze_pattern with every branch tested (fel.hpp:326-445).
Refit refit_pattern(bool invariant, double tested_length) {
    std::array<double, HY_FEL_FIELDS> res{};  // alpha, beta, beta_nuisance, alpha_null, lrt, p, length, lnL alt/null, n
    res[HY_FEL_P_VALUE] = 1.0;
    if (invariant) co_return res;
    int evals = 0;
    // start grid (fel.hpp:387-402): the first strictly best point
    std::vector<double> g = co_yield std::vector<Cand>(std::begin(kStartGrid), std::end(kStartGrid));
    evals += 12;
    double best_lnl = -1e20, best_a = 1.0, best_b = 1.0;
    for (int i = 0; i < 12; ++i)
        if (g[i] > best_lnl) {
            best_lnl = g[i];
            best_a = kStartGrid[i].a;
            best_b = kStartGrid[i].b;
        }
    // alternative: Nelder-Mead on -lnL (fel.hpp:404-411)
    auto nm = nelder_mead_2d(best_a, best_b, 1e-4, 1000.0, 1e-4, 100);
    for (bool more = nm.start(); more;) {
        const std::vector<double> r = co_yield std::vector<Cand>{nm.point()};
        ++evals;
        more = nm.step(-r[0]);
    }
    const Vertex alt = nm.result().best;
    res[HY_FEL_ALPHA] = alt.a;
    res[HY_FEL_BETA] = alt.b;
    res[HY_FEL_LOGL_ALT] = -alt.f;
    // null alpha = beta: start value and grid (fel.hpp:413-425), then Brent (:427-431)
    const double init_null = (std::min(alt.a, 100.0) + 3.0 * std::min(alt.b, 100.0)) / 4.0;
    std::vector<Cand> ng{{init_null, init_null}};
    for (double c : kNullGrid) ng.push_back({c, c});
    g = co_yield std::move(ng);
    evals += 8;
    double best_null_lnl = g[0], best_null_a = init_null;
    for (int i = 0; i < 7; ++i)
        if (g[1 + i] > best_null_lnl) {
            best_null_lnl = g[1 + i];
            best_null_a = kNullGrid[i];
        }
    auto br = brent_1d(1e-4, std::clamp(best_null_a, 0.01, 45.0), 50.0, 1e-4, 60);
    for (bool more = br.start(); more;) {
        const double x = br.point();
        const std::vector<double> r = co_yield std::vector<Cand>{{x, x}};
        ++evals;
        more = br.step(-r[0]);
    }
    const BrentResult nul = br.result();
    res[HY_FEL_ALPHA_NULL] = nul.x;
    res[HY_FEL_LOGL_NULL] = -nul.f;
    double lrt = 0.0;
    if (res[HY_FEL_LOGL_ALT] > -1e10 && res[HY_FEL_LOGL_NULL] > -1e10) {
        lrt = 2.0 * (res[HY_FEL_LOGL_ALT] - res[HY_FEL_LOGL_NULL]);
        if (lrt < 0.0) lrt = 0.0;
    }
    res[HY_FEL_LRT] = lrt;
    res[HY_FEL_P_VALUE] = std::erfc(std::sqrt(lrt / 2.0));
    res[HY_FEL_LENGTH] = nul.x * tested_length;
    res[HY_FEL_EVALS] = evals;
    co_return res;
}

}  // namespace

extern "C" int hy_fel_drive(int n_patterns, const uint8_t *invariant, double tested_length, hy_fel_eval_fn evaluate,
                            void *user, double *out, int *out_calls, int64_t *out_candidates) {
    if (n_patterns < 0 || (n_patterns > 0 && (!invariant || !out)) || !evaluate) return HY_ERR_ARG;
    std::vector<Refit> co;
    co.reserve(n_patterns);
    std::vector<int> live;
    for (int p = 0; p < n_patterns; ++p) {
        co.push_back(refit_pattern(invariant[p] != 0, tested_length));
        co[p].h.resume();
        if (co[p].h.done())
            std::memcpy(out + (size_t)p * HY_FEL_FIELDS, co[p].h.promise().result.data(), sizeof(double) * HY_FEL_FIELDS);
        else
            live.push_back(p);
    }
    int calls = 0;
    int64_t cands = 0;
    std::vector<int32_t> pat;
    std::vector<double> a, b, ll;
    while (!live.empty()) {
        pat.clear();
        a.clear();
        b.clear();
        for (int p : live)
            for (const Cand &c : co[p].h.promise().ask) {
                pat.push_back(p);
                a.push_back(c.a);
                b.push_back(c.b);
            }
        ll.assign(pat.size(), 0.0);
        const int rc = evaluate(user, (int)pat.size(), pat.data(), a.data(), b.data(), ll.data());
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
                std::memcpy(out + (size_t)p * HY_FEL_FIELDS, pr.result.data(), sizeof(double) * HY_FEL_FIELDS);
            else
                next.push_back(p);
        }
        live.swap(next);
    }
    if (out_calls) *out_calls = calls;
    if (out_candidates) *out_candidates = cands;
    return HY_OK;
}
