// Projected limited-memory quasi-Newton step over a box, shared by the host
// loop (hy_lbfgs_step, hy_fit_lbfgs with captured = 0) and the k_lbfgs kernel
// that closes each captured optimizer iteration on the GPU.
//
// Replaces the solver the reference's joint fit drives
// (include/hyphy/opt/optimizer.hpp:1074-1087: LBFGSpp::LBFGSBSolver with
// m = 8 correction pairs, epsilon = the gradient tolerance, max_iterations,
// box bounds on the log parameters).  The method here is the two-loop
// recursion restricted to the free variables (those not held at a bound by
// the sign of their gradient), a projected trial point x + alpha d clamped to
// the box, and a backtracking line search with the sufficient-decrease test
// f(xt) <= f(x) + c1 g.(xt - x) and quadratic interpolation of the step.
//
// One call of lb_step consumes one evaluated trial point (objective ft and
// gradient gt at xt) and leaves the next trial point in xt: the whole
// decision (accept or backtrack, history update, new direction, convergence)
// needs no host round trip, so an optimizer iteration replays as one graph.
//
// Every reduction goes through the policy's sum(), which adds in one fixed
// order (64 strided partial sums, then a pairwise tree) on the host and on
// the GPU, and floating-point contraction is off in this file: the host loop
// and the captured kernel take bit-identical steps from identical
// evaluations.
#pragma once
#include <cstdint>

#include "hyphy_mi355x.h"

#define HY_HD __host__ __device__
#define LB_NOCONTRACT _Pragma("clang fp contract(off)")
// one out-of-line copy of the larger steps keeps the kernel's code small (it
// runs right after the pruning kernels, with a cold instruction cache)
#define LB_OUTLINE __attribute__((noinline))

constexpr int LB_MMAX = 8;  // correction pairs at most (the reference's m)
constexpr int LB_LANES = 64;  // partial sums of a reduction (= k_lbfgs threads, one wave)

// Vectors of one optimizer state, carved from one array of
// lb_state_doubles(n, m) doubles.  Pair b (ring slot) is (S_b, Y_b) =
// (x_new - x_old, g_new - g_old); SY[a][b] = S_a.Y_b and YY[a][b] = Y_a.Y_b
// are kept for the pairs held, so that the two-loop recursion runs on these
// inner products instead of one dependent pass over the vectors per pair.
struct LbVec {
    double *x, *g, *d, *xt, *gt, *lb, *ub;  // n each
    double *S, *Y;                          // m x n correction pairs
    double *SY, *YY;                        // m x m inner products of the pairs
    int m;
};

HY_HD inline size_t lb_state_doubles(int n, int m) { return (size_t)7 * n + (size_t)2 * m * n + (size_t)2 * m * m; }

HY_HD inline LbVec lb_vec(double *base, int n, int m) {
    LbVec v;
    v.x = base;
    v.g = v.x + n;
    v.d = v.g + n;
    v.xt = v.d + n;
    v.gt = v.xt + n;
    v.lb = v.gt + n;
    v.ub = v.lb + n;
    v.S = v.ub + n;
    v.Y = v.S + (size_t)m * n;
    v.SY = v.Y + (size_t)m * n;
    v.YY = v.SY + (size_t)m * m;
    v.m = m;
    return v;
}

// exp(x) for the log-parameter map, written out so that host and GPU give the
// same bits: x = k ln2 + r with a two-part ln2 (Cody and Waite), |r| <= ln2/2,
// then the Taylor series of e^r to degree 13 by Horner (truncation below
// 2e-17 relative) and an exact scaling by 2^k.
HY_HD inline double lb_exp(double x) {
    LB_NOCONTRACT
    const double LOG2E = 1.4426950408889634, LN2_HI = 6.93147180369123816490e-01,
                 LN2_LO = 1.90821492927058770002e-10;
    const double k = __builtin_rint(x * LOG2E);
    const double r = (x - k * LN2_HI) - k * LN2_LO;
    double s = 1.0 / 6227020800.0;  // 1/13!
    const double inv[13] = {1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0, 1.0 / 362880.0, 1.0 / 40320.0,
                            1.0 / 5040.0, 1.0 / 720.0, 1.0 / 120.0, 1.0 / 24.0, 1.0 / 6.0, 0.5, 1.0, 1.0};
    for (int i = 0; i < 13; ++i) s = s * r + inv[i];
    return __builtin_ldexp(s, (int)k);
}

// Ring slot of the pair j steps old (0 = newest), 0 <= j < m.
HY_HD inline int lb_slot(int head, int j, int m) {
    const int s = head - 1 - j;
    return s < 0 ? s + m : s;
}

// Component of the projected gradient: zero where the bound holds the
// variable (at the lower bound with g > 0, at the upper bound with g < 0).
HY_HD inline bool lb_active(double x, double g, double lo, double hi) {
    return (x <= lo && g > 0.0) || (x >= hi && g < 0.0);
}

template <class P>
HY_HD void lb_set_trial(P &p, const LbVec &v, int n, double alpha) {
    LB_NOCONTRACT
    if (alpha == 0.0) {
        p.loop(n, [&](int i) { v.xt[i] = v.x[i]; });
        return;
    }
    p.loop(n, [&](int i) {
        double t = v.x[i] + alpha * v.d[i];
        t = t < v.lb[i] ? v.lb[i] : t;
        t = t > v.ub[i] ? v.ub[i] : t;
        v.xt[i] = t;
    });
}

// d = -H q for the projected gradient q (two-loop recursion, newest pair
// first), zero on the active set; steepest descent when the history gives no
// descent.  With a_j = S_j.q and e_j = Y_j.q (one batched pass), the first
// loop's S_j.q_j = a_j - sum_{i<j} al_i S_j.Y_i and the second loop's
// Y_j.r_j = gamma (e_j - sum_i al_i Y_j.Y_i) + sum_{i>j} c_i Y_j.S_i are
// scalar recurrences, and d = -(gamma (q - sum_i al_i Y_i) + sum_i c_i S_i)
// is one more pass.
template <class P>
HY_HD LB_OUTLINE void lb_direction(P &p, const LbVec &v, hy_lbfgs_head &hd) {
    LB_NOCONTRACT
    const int n = hd.n, m = hd.m, k = hd.k;
    // slots 0..k-1 hold pairs (the ring fills from slot 0)
    double aev[2 * LB_MMAX];
    p.template sums<2 * LB_MMAX>(n, aev, [&](int i, double *acc) {
        const double q = lb_active(v.x[i], v.g[i], v.lb[i], v.ub[i]) ? 0.0 : v.g[i];
#pragma unroll
        for (int b = 0; b < LB_MMAX; ++b)
            if (b < k) {
                acc[2 * b] = acc[2 * b] + v.S[(size_t)b * n + i] * q;
                acc[2 * b + 1] = acc[2 * b + 1] + v.Y[(size_t)b * n + i] * q;
            }
    });
    // the recurrences run on a work area shared by the lanes (LDS on the GPU):
    // the pairs' inner products, a and e, and the coefficients by slot (al of
    // the first loop, c = al - b of the second)
    double *W = p.work();
    double *wsy = W, *wyy = W + LB_MMAX * LB_MMAX, *ae = wyy + LB_MMAX * LB_MMAX, *al = ae + 2 * LB_MMAX,
           *cc = al + LB_MMAX;
    p.loop(m * m, [&](int i) {
        wsy[i] = v.SY[i];
        wyy[i] = v.YY[i];
    });
    p.single([&]() {
#pragma unroll
        for (int q = 0; q < 2 * LB_MMAX; ++q) ae[q] = aev[q];
#pragma unroll
        for (int b = 0; b < LB_MMAX; ++b) al[b] = cc[b] = 0.0;
    });
    p.sync();
    double gamma = 1.0;
    if (k > 0) {
        for (int j = 0; j < k; ++j) {  // newest first
            const int sj = lb_slot(hd.head, j, m);
            double t = ae[2 * sj];
            for (int i = 0; i < j; ++i) {
                const int si = lb_slot(hd.head, i, m);
                t = t - al[si] * wsy[sj * m + si];
            }
            t = t / wsy[sj * m + sj];
            p.single([&]() { al[sj] = t; });
            p.sync();
        }
        const int s0 = lb_slot(hd.head, 0, m);
        gamma = wsy[s0 * m + s0] / wyy[s0 * m + s0];
        for (int j = k - 1; j >= 0; --j) {  // oldest first
            const int sj = lb_slot(hd.head, j, m);
            double t = ae[2 * sj + 1];
            for (int i = 0; i < k; ++i) {
                const int si = lb_slot(hd.head, i, m);
                t = t - al[si] * wyy[sj * m + si];
            }
            t = gamma * t;
            for (int i = k - 1; i > j; --i) {
                const int si = lb_slot(hd.head, i, m);
                t = t + cc[si] * wsy[si * m + sj];
            }
            const double c = al[sj] - t / wsy[sj * m + sj];
            p.single([&]() { cc[sj] = c; });
            p.sync();
        }
    }
    p.loop(n, [&](int i) {
        if (lb_active(v.x[i], v.g[i], v.lb[i], v.ub[i])) {
            v.d[i] = 0.0;
            return;
        }
        double t = v.g[i];
#pragma unroll
        for (int b = 0; b < LB_MMAX; ++b)
            if (b < k) t = t - al[b] * v.Y[(size_t)b * n + i];
        t = gamma * t;
#pragma unroll
        for (int b = 0; b < LB_MMAX; ++b)
            if (b < k) t = t + cc[b] * v.S[(size_t)b * n + i];
        v.d[i] = -t;
    });
    hd.gd = p.sum(n, [&](int i) { return v.g[i] * v.d[i]; });
    if (!(hd.gd < 0.0)) {  // not a descent direction: drop the history
        hd.k = 0;
        hd.head = 0;
        p.loop(n, [&](int i) { v.d[i] = lb_active(v.x[i], v.g[i], v.lb[i], v.ub[i]) ? 0.0 : -v.g[i]; });
        hd.gd = p.sum(n, [&](int i) { return v.g[i] * v.d[i]; });
    }
}

// x, g, f hold a new point: convergence tests, else the next direction and
// the first trial point along it.
template <class P>
HY_HD LB_OUTLINE void lb_new_point(P &p, const LbVec &v, hy_lbfgs_head &hd) {
    LB_NOCONTRACT
    const int n = hd.n;
    hd.pgnorm = p.max(n, [&](int i) {
        const double g = lb_active(v.x[i], v.g[i], v.lb[i], v.ub[i]) ? 0.0 : v.g[i];
        return g < 0.0 ? -g : g;
    });
    const double af = hd.f < 0.0 ? -hd.f : hd.f;
    const double df = hd.f_prev - hd.f;
    if (hd.pgnorm <= hd.tol_grad)
        hd.status = HY_LBFGS_CONV_GRAD;
    else if (hd.iter > 0 && hd.tol_f > 0.0 && (df < 0.0 ? -df : df) <= hd.tol_f * (af > 1.0 ? af : 1.0))
        hd.status = HY_LBFGS_CONV_F;
    else if (hd.iter >= hd.max_iter)
        hd.status = HY_LBFGS_MAX_ITER;
    if (hd.status != HY_LBFGS_RUNNING) {
        lb_set_trial(p, v, n, 0.0);
        return;
    }
    lb_direction(p, v, hd);
    if (hd.k == 0) {  // first step along a steepest-descent direction: at most 1 in every coordinate
        const double dm = p.max(n, [&](int i) { return v.d[i] < 0.0 ? -v.d[i] : v.d[i]; });
        hd.alpha = dm > 1.0 ? 1.0 / dm : 1.0;
    } else {
        hd.alpha = 1.0;
    }
    hd.n_back = 0;
    lb_set_trial(p, v, n, hd.alpha);
}

// Consumes the evaluation of the trial point xt (objective ft, gradient in
// v.gt; ok = 0 when the evaluation failed) and leaves the next trial in xt.
template <class P>
HY_HD void lb_step(P &p, const LbVec &v, hy_lbfgs_head &hd, double ft, int ok) {
    LB_NOCONTRACT
    const int n = hd.n, m = hd.m;
    if (hd.status != HY_LBFGS_RUNNING) {  // finished: evaluations of x change nothing
        lb_set_trial(p, v, n, 0.0);
        return;
    }
    hd.n_eval += 1;
    // non-finite objective or gradient components count as a failed
    // evaluation; dec = g.(xt - x), the slope term of the decrease test
    double bd[2];
    p.template sums<2>(n, bd, [&](int i, double *acc) {
        acc[0] = acc[0] + ((v.gt[i] - v.gt[i] == 0.0) ? 0.0 : 1.0);
        acc[1] = acc[1] + v.g[i] * (v.xt[i] - v.x[i]);
    });
    const bool good = ok && bd[0] == 0.0 && (ft - ft == 0.0);
    if (hd.n_eval == 1) {  // the start point
        p.loop(n, [&](int i) {
            v.x[i] = v.xt[i];
            v.g[i] = v.gt[i];
        });
        if (!good) {
            hd.status = HY_LBFGS_NUMERIC;
            lb_set_trial(p, v, n, 0.0);
            return;
        }
        hd.f = ft;
        hd.f_prev = ft;
        lb_new_point(p, v, hd);
        return;
    }
    const double dec = bd[1];
    if (good && ft <= hd.f + hd.c1 * dec) {
        // the new pair (s, y) = (xt - x, gt - g) and its inner products with
        // itself and with the pairs held (slots 0..k-1)
        const int k = hd.k;
        double pr[2 + 3 * LB_MMAX];
        p.template sums<2 + 3 * LB_MMAX>(n, pr, [&](int i, double *acc) {
            const double sv = v.xt[i] - v.x[i], yv = v.gt[i] - v.g[i];
            acc[0] = acc[0] + sv * yv;
            acc[1] = acc[1] + yv * yv;
#pragma unroll
            for (int b = 0; b < LB_MMAX; ++b)
                if (b < k) {
                    const double sb = v.S[(size_t)b * n + i], yb = v.Y[(size_t)b * n + i];
                    acc[2 + 3 * b] = acc[2 + 3 * b] + sv * yb;
                    acc[3 + 3 * b] = acc[3 + 3 * b] + sb * yv;
                    acc[4 + 3 * b] = acc[4 + 3 * b] + yv * yb;
                }
        });
        const double sy = pr[0], yy = pr[1];
        const bool keep = yy > 0.0 && sy > 1e-10 * yy;  // curvature pair kept (overwrites the oldest when full)
        const int h = hd.head;
        p.loop(n, [&](int i) {
            if (keep) {
                v.S[(size_t)h * n + i] = v.xt[i] - v.x[i];
                v.Y[(size_t)h * n + i] = v.gt[i] - v.g[i];
            }
            v.x[i] = v.xt[i];
            v.g[i] = v.gt[i];
        });
        if (keep) {
            p.single([&]() {
#pragma unroll
                for (int b = 0; b < LB_MMAX; ++b)
                    if (b < k && b != h) {
                        v.SY[h * m + b] = pr[2 + 3 * b];
                        v.SY[b * m + h] = pr[3 + 3 * b];
                        v.YY[h * m + b] = pr[4 + 3 * b];
                        v.YY[b * m + h] = pr[4 + 3 * b];
                    }
                v.SY[h * m + h] = sy;
                v.YY[h * m + h] = yy;
            });
            p.sync();
            hd.head = (h + 1) % m;
            hd.k = k + 1 < m ? k + 1 : m;
        }
        hd.f_prev = hd.f;
        hd.f = ft;
        hd.iter += 1;
        lb_new_point(p, v, hd);
        return;
    }
    hd.n_back += 1;
    if (hd.n_back > hd.max_back) {
        hd.status = HY_LBFGS_LS_FAIL;
        lb_set_trial(p, v, n, 0.0);
        return;
    }
    // minimiser of the quadratic through f, the slope dec and ft along the
    // trial segment, kept within [0.1, 0.5] of the current step
    double t = 0.25;
    if (good) {
        const double c = ft - hd.f - dec;
        t = c > 0.0 ? -dec / (2.0 * c) : 0.5;
        t = t < 0.1 ? 0.1 : (t > 0.5 ? 0.5 : t);
    }
    hd.alpha = hd.alpha * t;
    lb_set_trial(p, v, n, hd.alpha);
}

// Host policy: the same strided partial sums and pairwise tree as the GPU's
// one-wave reduction (lane t adds lane t + st for st = 32, 16, ..., 1).
struct LbHostPar {
    double w[2 * LB_MMAX * LB_MMAX + 4 * LB_MMAX];
    double *work() { return w; }
    template <class F>
    void single(F f) {
        f();
    }
    void sync() {}
    template <class F>
    void loop(int n, F f) {
        for (int i = 0; i < n; ++i) f(i);
    }
    template <class F>
    double sum(int n, F f) {
        LB_NOCONTRACT
        double red[LB_LANES];
        for (int t = 0; t < LB_LANES; ++t) {
            double s = 0.0;
            for (int i = t; i < n; i += LB_LANES) s = s + f(i);
            red[t] = s;
        }
        for (int st = LB_LANES / 2; st > 0; st >>= 1)
            for (int t = 0; t < st; ++t) red[t] = red[t] + red[t + st];
        return red[0];
    }
    // NQ sums in one pass: f(i, acc) adds element i's terms to acc[0..NQ-1]
    template <int NQ, class F>
    void sums(int n, double *out, F f) {
        LB_NOCONTRACT
        double red[LB_LANES][NQ];
        for (int t = 0; t < LB_LANES; ++t) {
            for (int q = 0; q < NQ; ++q) red[t][q] = 0.0;
            for (int i = t; i < n; i += LB_LANES) f(i, red[t]);
        }
        for (int st = LB_LANES / 2; st > 0; st >>= 1)
            for (int t = 0; t < st; ++t)
                for (int q = 0; q < NQ; ++q) red[t][q] = red[t][q] + red[t + st][q];
        for (int q = 0; q < NQ; ++q) out[q] = red[0][q];
    }
    template <class F>
    double max(int n, F f) {
        double r = 0.0;
        for (int i = 0; i < n; ++i) {
            const double x = f(i);
            r = x > r ? x : r;
        }
        return r;
    }
};

// Objective and gradient of the trial point from an evaluation's results:
// F = -scale lnL, dF/dx_i = -scale dlnL/dparam_i * param_i for param_i =
// exp(x_i).  res_map / par_map: index of variable i in the result vector
// (lnL first) and in the packed parameter vector.
template <class P>
HY_HD double lb_gather(P &p, const LbVec &v, int n, double scale, const double *res, const double *params,
                       const int32_t *res_map, const int32_t *par_map) {
    LB_NOCONTRACT
    p.loop(n, [&](int i) { v.gt[i] = -(scale * res[res_map[i]]) * params[par_map[i]]; });
    return -(scale * res[0]);
}

template <class P>
HY_HD void lb_scatter(P &p, const LbVec &v, int n, double *params, const int32_t *par_map) {
    p.loop(n, [&](int i) { params[par_map[i]] = lb_exp(v.xt[i]); });
}

// Arguments of k_lbfgs: the optimizer state in device memory, the
// evaluation's results and packed parameters (pinned host memory read and
// written through their device addresses), the eigensolver status words.
struct LbArgs {
    double *state;
    hy_lbfgs_head *hd;
    const double *res;
    double *params;
    const int *eig_status;
    int n_mat;
    double scale;
    const int32_t *res_map, *par_map;
};
void launch_lbfgs(const LbArgs &a, hipStream_t s);
