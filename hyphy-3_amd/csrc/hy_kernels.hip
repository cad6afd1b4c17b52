// HIP kernels of the MI355X MG94xREV codon likelihood engine (gfx950).
//
// Pipeline per evaluation (one stream):
//   k_model    one workgroup per rate matrix: MG94 Q, mu, symmetric spectrum
//              (warm-started Ogita-Aishima refinement on FP64 MFMA, or
//              parallel cyclic Jacobi in LDS when cold)
//   k_pmat     one workgroup per branch: P_b = sum_j w_j exp(Q_j tau_b) in FP64,
//              written as MFMA A-operand tiles of P and P^T in the working type;
//              with gradients, further (matrix, parameter) workgroups build
//              M = V_pi dQ U_pi (the projections below) in the same launch
//   k_inside   one workgroup per pattern tile: post-order pruning, 64x64 by
//              64xT child products on MFMA, per-pattern power-of-two rescale
//   k_outside  one workgroup per pattern tile: pre-order outside messages
//   k_adjoint  (branch, pattern chunk) workgroups: A_b = sum_p c_p V_b D_b^T
//   k_frechet  (branch, component) workgroups: U^T A V^T, spectral Frechet
//              projections (FP64)
//   k_final    one workgroup of 16 waves: deterministic compensated reductions
//
// MFMA lane layout (v_mfma_{f32,f64}_16x16x4, measured on gfx950): lane l
// supplies A[l%16][l/16] and B[l/16][l%16]; accumulator element r of lane l
// holds C[row][l%16] with row = 4*(l/16)+r for f32 and row = 4*r+(l/16) for
// f64 (MF<R>::crow).
#include "hy_internal.h"
#include "hy_lbfgs.h"

#include <cmath>
#include <cstdio>
#include <mutex>

namespace hy {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename R> struct MF;
template <> struct MF<float> {
    typedef f32x4 acc;
    static __device__ __forceinline__ acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int crow(int lg, int r) { return 4 * lg + r; }
};
template <> struct MF<double> {
    typedef f64x4 acc;
    static __device__ __forceinline__ acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int crow(int lg, int r) { return 4 * r + lg; }
};

template <typename A> __device__ __forceinline__ A splat(double v) {
    A x;
    x[0] = v; x[1] = v; x[2] = v; x[3] = v;
    return x;
}

__device__ __forceinline__ float ldexp_r(float x, int e) { return ldexpf(x, e); }
__device__ __forceinline__ double ldexp_r(double x, int e) { return ldexp(x, e); }
__device__ __forceinline__ int frexp_e(float m) { int e; frexpf(m, &e); return e; }
__device__ __forceinline__ int frexp_e(double m) { int e; frexp(m, &e); return e; }

// Nucleotide exchangeabilities with theta_AG == 1.
__device__ __forceinline__ double nuc_rate(const double *th, int a, int b) {
    // index pairs: AC=0, AT=1, CG=2, CT=3, GT=4, AG=1.0
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    if (lo == 0 && hi == 1) return th[0];
    if (lo == 0 && hi == 3) return th[1];
    if (lo == 1 && hi == 2) return th[2];
    if (lo == 1 && hi == 3) return th[3];
    if (lo == 2 && hi == 3) return th[4];
    return 1.0;  // A<->G
}

// ---------------------------------------------------------------------------
// k_model: Q, mu and the symmetric spectrum of D^1/2 Q D^-1/2 (FP64).
//
// Warm start (a.warm): the previous call's eigenvectors X are refined by the
// Ogita-Aishima iteration (Ogita & Aishima, "Iterative refinement for symmetric
// eigenvalue decomposition", JJIAM 35, 2018): R = I - X^T X, S = X^T A X,
// l_i = s_ii / (1 - r_ii), E_ij = (s_ij + l_j r_ij) / (l_j - l_i) for separated
// pairs and r_ij / 2 inside clusters, X <- X + X E; its four 64x64x64 products
// run on FP64 MFMA.  Accepted once X^T A X is diagonal to 1e-9 ||A|| and
// the update is below 1e-9 (the remaining error is then second order); after
// eight steps, or when the iteration stops contracting, the matrix is solved
// cold.
// Cold: parallel cyclic Jacobi (round-robin pairing, 32 disjoint rotations per
// round), rotations skipped below 1e-17 ||A||_F, converged after a sweep with
// no rotation.
// ---------------------------------------------------------------------------
constexpr int LDA = 65;  // padded LDS row stride (doubles) for 64x64 FP64 tiles

// C = op(X) * op(Y) for 64x64 FP64 tiles in LDS (row stride LDA) on
// v_mfma_f64_16x16x4, in two halves: products into registers (wave w owns row
// block w of four waves), then the store, so the result may overwrite an
// operand once every wave has finished reading.
template <bool TX, bool TY>
__device__ __forceinline__ void gemm64_acc(const double *X, const double *Y, f64x4 *acc, int rb, int lr, int lg) {
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = splat<f64x4>(0.0);
#pragma unroll 4
    for (int kk = 0; kk < 16; ++kk) {
        const int i = 16 * rb + lr, k = 4 * kk + lg;
        const double av = TX ? X[k * LDA + i] : X[i * LDA + k];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = 16 * q + lr;
            acc[q] = MF<double>::mma(av, TY ? Y[j * LDA + k] : Y[k * LDA + j], acc[q]);
        }
    }
}

__device__ __forceinline__ void gemm64_store(const f64x4 *acc, double *C, int rb, int lr, int lg) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) C[(16 * rb + MF<double>::crow(lg, r)) * LDA + 16 * q + lr] = acc[q][r];
}

// Deterministic block sum of one double per thread (NT threads); red has NT
// entries; every thread gets the result.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *red, int tid) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < NT / 64; ++k) s += red[k];
    __syncthreads();
    return s;
}

// Two block sums at once (same contract as block_sum; red has 2 NT / 64 entries).
template <int NT>
__device__ __forceinline__ void block_sum2(double &a, double &b, double *red, int tid) {
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off);
        b += __shfl_down(b, off);
    }
    __syncthreads();
    if ((tid & 63) == 0) {
        red[2 * (tid >> 6)] = a;
        red[2 * (tid >> 6) + 1] = b;
    }
    __syncthreads();
    double sa = 0.0, sb = 0.0;
    for (int k = 0; k < NT / 64; ++k) {
        sa += red[2 * k];
        sb += red[2 * k + 1];
    }
    __syncthreads();
    a = sa;
    b = sb;
}

// Two block sums with one barrier: red2 (2 NT / 64 entries) must not be the
// buffer of the previous call (callers alternate two buffers; a thread that
// reaches this call's barrier has finished reading the other one).
template <int NT>
__device__ __forceinline__ void block_sum2_once(double &a, double &b, double *red2, int tid) {
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off);
        b += __shfl_down(b, off);
    }
    if ((tid & 63) == 0) {
        red2[2 * (tid >> 6)] = a;
        red2[2 * (tid >> 6) + 1] = b;
    }
    __syncthreads();
    double sa = 0.0, sb = 0.0;
    for (int k = 0; k < NT / 64; ++k) {
        sa += red2[2 * k];
        sb += red2[2 * k + 1];
    }
    a = sa;
    b = sb;
}

constexpr int MODEL_NT = 1024;

// Reciprocal to about one ulp: the hardware estimate and one Newton step.
__device__ __forceinline__ double rcp_nr(double q) {
    const double r = __builtin_amdgcn_rcp(q);
    return fma(r, fma(-q, r, 1.0), r);
}

// DPP lane exchange of a double (both halves with the same control word).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// Sum over each 16-lane row, in every lane of the row (quad swaps, then the
// half-row and row mirrors); every lane of the row must be active.
__device__ __forceinline__ double row_sum16(double v) {
    v += dpp_f64<0xB1>(v);   // quad_perm [1, 0, 3, 2]
    v += dpp_f64<0x4E>(v);   // quad_perm [2, 3, 0, 1]
    v += dpp_f64<0x141>(v);  // row_half_mirror
    v += dpp_f64<0x140>(v);  // row_mirror
    return v;
}

// Value of lane l (uniform) of a double.
__device__ __forceinline__ double readlane_f64(double v, int l) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Sum over the 64 lanes of a wave, in every lane (fixed order); all lanes active.
__device__ __forceinline__ double wave_sum64(double v) {
    v = row_sum16(v);
    return ((readlane_f64(v, 0) + readlane_f64(v, 16)) + readlane_f64(v, 32)) + readlane_f64(v, 48);
}

// Makes LDS stores of a wave visible to its other lanes.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Scratch of the tridiagonal eigensolver (LDS, one per workgroup).
constexpr int TRI_WAVES = 8;  // waves of the tridiagonalisation
constexpr int TRI_BIS = 8;    // lanes (points) per eigenvalue in the multisection
struct TriScratch {
    double d[64], e[64], e2[64];     // diagonal, off-diagonal, squares (split entries zeroed)
    double v[TRI_WAVES][64];         // reflector of the current step (one copy per wave)
    double w[TRI_WAVES][64];         // its rank-2 partner (one copy per wave)
    double part[TRI_WAVES][64];      // partial products M v per wave
    double beta[64];                 // reflector scales (0: no reflection)
    int blo[64], bhi[64];            // unreduced block (first, last row) of each row
    double pivmin, tnorm, scale;
};

// Cold symmetric eigensolver for the n x n leading block of A (n >= 3, zero
// outside it).
//  * Householder tridiagonalisation of a copy in M by TRI_WAVES waves: lane i
//    of every wave owns row i, wave w a share of the trailing columns; each
//    wave forms the reflector itself (same arithmetic, same result), so a
//    step needs two workgroup barriers.  Reflector k is kept in column k of
//    W, then moved below the diagonal of M.
//  * Eigenvalues by (TRI_BIS + 1)-way multisection on Sturm counts of each
//    unreduced block (TRI_BIS lanes per eigenvalue), on T scaled by a power of two to norm
//    about one: the Sturm sequence p_r = (d_r - x) p_r-1 - e_r-1^2 p_r-2 needs
//    no division, and a power-of-two rescale every eight steps keeps it in
//    range.
//  * Eigenvectors by twisted factorisations of T - lam I (one lane each;
//    the lane's scratch in column k of W and of Z).
//  * Back-transformation by the reflectors, 16 lanes per column with the
//    column in registers.
//  * Two Loewdin steps Z <- Z (I + (I - Z^T Z) / 2), which restore
//    orthonormality where close eigenvalues left their vectors tilted.
// On exit Z holds the eigenvectors (columns; identity outside n) and lam the
// eigenvalues (0 outside n).  A is not modified.
template <int NT>
__device__ __attribute__((noinline)) void tridiag_eigen(const double *A, double *Z, double *M, double *W, double *lam,
                                                        int n, int tid, TriScratch &ts) {
    const int lane = tid & 63, w = tid >> 6;
    for (int idx = tid; idx < 4096; idx += NT) M[(idx >> 6) * LDA + (idx & 63)] = A[(idx >> 6) * LDA + (idx & 63)];
    __syncthreads();
    // ---- Householder tridiagonalisation (waves 0 .. TRI_WAVES - 1) ----
    for (int k = 0; k + 2 < n; ++k) {
        const int i = lane;
        const bool mine = i > k && i < n;
        double vi = 0.0, beta = 0.0;
        int j0 = 0, j1 = 0;
        if (w < TRI_WAVES) {
            const double xi = mine ? M[i * LDA + k] : 0.0;
            const double s = wave_sum64(xi * xi);
            const double x0 = readlane_f64(xi, k + 1);
            const double alpha = s > 0.0 ? -copysign(sqrt(s), x0) : 0.0;
            const double v0 = x0 - alpha;
            const double vn2 = s - x0 * x0 + v0 * v0;
            beta = vn2 > 0.0 ? 2.0 / vn2 : 0.0;
            vi = mine ? (i == k + 1 ? v0 : xi) : 0.0;
            ts.v[w][i] = vi;
            if (w == 0) {
                if (mine) W[i * LDA + k] = vi;
                if (i == 0) {
                    ts.d[k] = M[k * LDA + k];
                    ts.e[k] = alpha;
                    ts.beta[k] = beta;
                }
            }
            // this wave's columns of the trailing block
            const int m = n - k - 1, c = (m + TRI_WAVES - 1) / TRI_WAVES;
            j0 = k + 1 + w * c;
            j1 = min(n, j0 + c);
            wave_sync();
            double pa[4] = {0.0, 0.0, 0.0, 0.0};
            if (mine && beta != 0.0) {
                int j = j0;
                for (; j + 7 < j1; j += 8) {
                    double mr[8], vr[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        mr[u] = M[i * LDA + j + u];
                        vr[u] = ts.v[w][j + u];
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) pa[u & 3] += mr[u] * vr[u];
                }
                for (; j < j1; ++j) pa[0] += M[i * LDA + j] * ts.v[w][j];
            }
            ts.part[w][i] = (pa[0] + pa[1]) + (pa[2] + pa[3]);
        }
        __syncthreads();
        if (w < TRI_WAVES && beta != 0.0) {
            double psum = 0.0;
#pragma unroll
            for (int u = 0; u < TRI_WAVES; ++u) psum += ts.part[u][i];
            const double pi_ = beta * psum;
            const double t = wave_sum64(vi * pi_);
            const double wi = pi_ - 0.5 * beta * t * vi;
            ts.w[w][i] = mine ? wi : 0.0;
            wave_sync();
            if (mine) {  // row i, this wave's columns: M22 -= v w^T + w v^T
                int j = j0;
                for (; j + 7 < j1; j += 8) {
                    double mr[8], wr[8], vr[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        mr[u] = M[i * LDA + j + u];
                        wr[u] = ts.w[w][j + u];
                        vr[u] = ts.v[w][j + u];
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) M[i * LDA + j + u] = mr[u] - (vi * wr[u] + wi * vr[u]);
                }
                for (; j < j1; ++j) M[i * LDA + j] -= vi * ts.w[w][j] + wi * ts.v[w][j];
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        ts.d[n - 2] = M[(n - 2) * LDA + n - 2];
        ts.e[n - 2] = M[(n - 1) * LDA + n - 2];
        ts.d[n - 1] = M[(n - 1) * LDA + n - 1];
        ts.e[n - 1] = 0.0;
        ts.beta[n - 2] = 0.0;
        ts.beta[n - 1] = 0.0;
        // unreduced blocks: off-diagonal entries below eps (|d_i| + |d_i+1|) are dropped
        double tn = 0.0;
        for (int r = 0; r + 1 < n; ++r)
            if (fabs(ts.e[r]) <= 2.220446049250313e-16 * (fabs(ts.d[r]) + fabs(ts.d[r + 1]))) ts.e[r] = 0.0;
        for (int r = 0; r < n; ++r)
            tn = fmax(tn, fabs(ts.d[r]) + (r > 0 ? fabs(ts.e[r - 1]) : 0.0) + fabs(ts.e[r]));
        // the Sturm counts run on T / 2^ceil(log2 tnorm)
        const double sc = tn > 0.0 ? ldexp(1.0, -frexp_e(tn)) : 1.0;
        double emax = 0.0;
        for (int r = 0; r < n; ++r) {
            ts.e2[r] = (ts.e[r] * sc) * (ts.e[r] * sc);
            emax = fmax(emax, ts.e2[r]);
        }
        ts.scale = sc;
        ts.tnorm = tn;
        ts.pivmin = 2.2250738585072014e-308 * fmax(1.0, emax) * 4.0;
        int s0 = 0;
        for (int r = 0; r < n; ++r) {
            if (r + 1 == n || ts.e[r] == 0.0) {
                for (int c = s0; c <= r; ++c) {
                    ts.blo[c] = s0;
                    ts.bhi[c] = r;
                }
                s0 = r + 1;
            }
        }
    }
    __syncthreads();
    // reflectors below the diagonal of M (column k, rows k+1 .. n-1)
    for (int idx = tid; idx < 4096; idx += NT) {
        const int r = idx >> 6, c = idx & 63;
        if (r > c && r < n && c + 2 < n) M[r * LDA + c] = W[r * LDA + c];
    }
    __syncthreads();
    // ---- eigenvalues: eigenvalue j - blo of block [blo, bhi] for row j, TRI_BIS lanes each ----
    {
        constexpr int L = TRI_BIS;
        const int j = tid / L, q = tid % L;
        double lo = 0.0, hi = 0.0;
        const double sc = ts.scale;
        if (j < n) {
            const int s = ts.blo[j], e = ts.bhi[j], want = j - s;
            const double pivmin = ts.pivmin;
            lo = 1e300;
            hi = -1e300;
            for (int r = s; r <= e; ++r) {  // Gershgorin interval of the (scaled) block
                const double rad = ((r > s ? fabs(ts.e[r - 1]) : 0.0) + (r < e ? fabs(ts.e[r]) : 0.0)) * sc;
                lo = fmin(lo, ts.d[r] * sc - rad);
                hi = fmax(hi, ts.d[r] * sc + rad);
            }
            const double pad = 2.220446049250313e-16 * 4.0 * ts.tnorm * sc + 2.0 * pivmin;
            lo -= pad;
            hi += pad;
            for (int it = 0; it < 80; ++it) {
                // to full precision: a refinement step would only add
                // rounding noise divided by the eigenvalue gaps
                if (hi - lo <= 2.220446049250313e-16 * 2.0 * fmax(fabs(lo), fabs(hi)) ||
                    hi - lo <= 2.220446049250313e-16 * 0.125 * ts.tnorm * sc)
                    break;
                const double x = lo + (hi - lo) * (double)(q + 1) / (double)(L + 1);
                // eigenvalues of the block below x: sign changes of the Sturm
                // sequence (a zero p_r and the next term, of opposite sign to
                // p_r-1, count one change together, as a perturbed zero would)
                double pm = 1.0, pc = ts.d[s] * sc - x;
                int cnt = pc < 0.0;
                int r = s + 1;
                for (; r + 7 <= e; r += 8) {
                    double dr[8], er[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        dr[u] = ts.d[r + u];
                        er[u] = ts.e2[r + u - 1];
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const double pn = fma(dr[u] * sc - x, pc, -er[u] * pm);
                        cnt += (pn < 0.0) != (pc < 0.0);
                        pm = pc;
                        pc = pn;
                    }
                    const int ex = frexp_e(pc);
                    if (ex > 400 || ex < -400) {
                        pm = ldexp(pm, -ex);
                        pc = ldexp(pc, -ex);
                    }
                }
                for (; r <= e; ++r) {
                    const double pn = fma(ts.d[r] * sc - x, pc, -ts.e2[r - 1] * pm);
                    cnt += (pn < 0.0) != (pc < 0.0);
                    pm = pc;
                    pc = pn;
                }
                const unsigned long long bal = __ballot(cnt > want);
                const unsigned bits = (unsigned)((bal >> (L * ((tid & 63) / L))) & ((1ull << L) - 1ull));
                const int f = bits ? __ffs(bits) - 1 : L;  // first point with more than `want` below
                const double nlo = f > 0 ? lo + (hi - lo) * (double)f / (double)(L + 1) : lo;
                const double nhi = f < L ? lo + (hi - lo) * (double)(f + 1) / (double)(L + 1) : hi;
                lo = nlo;
                hi = nhi;
            }
        }
        if (q == 0 && j < 64) lam[j] = j < n ? 0.5 * (lo + hi) / sc : 0.0;
    }
    __syncthreads();
    // ---- eigenvectors: twisted factorisation of (T - lam I) / 2^.. on the block, one lane each;
    //      L D+ L^T from the top by wave 0 (in W), U D- U^T from the bottom by wave 1 (in Z) ----
    if (tid < 128) {
        const int k = lane;
        if (k < n) {
            const int s = ts.blo[k], e = ts.bhi[k];
            const double sc = ts.scale, l = lam[k] * sc, pivmin = ts.pivmin;
            if (w == 0) {
                double dp = ts.d[s] * sc - l;
                if (fabs(dp) < pivmin) dp = -pivmin;
                W[s * LDA + k] = dp;
                for (int r = s + 1; r <= e; ++r) {
                    dp = (ts.d[r] * sc - l) - ts.e2[r - 1] * rcp_nr(dp);
                    if (fabs(dp) < pivmin) dp = -pivmin;
                    W[r * LDA + k] = dp;
                }
            } else {
                double dm = ts.d[e] * sc - l;
                if (fabs(dm) < pivmin) dm = -pivmin;
                Z[e * LDA + k] = dm;
                for (int r = e - 1; r >= s; --r) {
                    dm = (ts.d[r] * sc - l) - ts.e2[r] * rcp_nr(dm);
                    if (fabs(dm) < pivmin) dm = -pivmin;
                    Z[r * LDA + k] = dm;
                }
            }
        }
    }
    __syncthreads();
    if (tid < 64) {
        const int k = tid;
        if (k < n) {
            const int s = ts.blo[k], e = ts.bhi[k];
            const double sc = ts.scale, l = lam[k] * sc;
            int tw = e;  // twist at the smallest |gamma_r| = |D+_r + D-_r - (d_r - l)|
            double gbest = 1e308;
            for (int r = e; r >= s; --r) {
                const double g = fabs(W[r * LDA + k] + Z[r * LDA + k] - (ts.d[r] * sc - l));
                if (g < gbest) {
                    gbest = g;
                    tw = r;
                }
            }
            double nrm = 1.0, x = 1.0;
            for (int r = tw; r < e; ++r) {  // below the twist: x_r+1 = -e_r x_r / D-_r+1
                x = -(ts.e[r] * sc) * x * rcp_nr(Z[(r + 1) * LDA + k]);
                Z[(r + 1) * LDA + k] = x;
                nrm += x * x;
            }
            Z[tw * LDA + k] = 1.0;
            x = 1.0;
            for (int r = tw - 1; r >= s; --r) {  // above: x_r = -e_r x_r+1 / D+_r
                x = -(ts.e[r] * sc) * x * rcp_nr(W[r * LDA + k]);
                Z[r * LDA + k] = x;
                nrm += x * x;
            }
            const double inv = 1.0 / sqrt(nrm);
#pragma unroll 4
            for (int r = 0; r < 64; ++r) Z[r * LDA + k] = (r >= s && r <= e) ? Z[r * LDA + k] * inv : 0.0;
        } else {
#pragma unroll 4
            for (int r = 0; r < 64; ++r) Z[r * LDA + k] = r == k ? 1.0 : 0.0;
        }
    }
    __syncthreads();
    // ---- back-transformation Z <- H_0 ... H_{n-3} Z: column j by 16 lanes, rows q, q + 16, .. in registers ----
    {
        const int j = tid >> 4, q = tid & 15;
        double z[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) z[c] = Z[(q + 16 * c) * LDA + j];
        if (j < n) {
#pragma unroll 2
            for (int k = n - 3; k >= 0; --k) {
                const double beta = ts.beta[k];  // 0: no reflection, the update vanishes
                double vr[4], dot = 0.0;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int r = q + 16 * c;
                    vr[c] = (r > k && r < n) ? M[r * LDA + k] : 0.0;
                    dot += vr[c] * z[c];
                }
                dot = row_sum16(dot) * beta;
#pragma unroll
                for (int c = 0; c < 4; ++c) z[c] -= vr[c] * dot;
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 4; ++c) Z[(q + 16 * c) * LDA + j] = z[c];
    }
    __syncthreads();
    // ---- two Loewdin steps: Z <- Z + Z (I - Z^T Z) / 2 ----
    const int lr = tid & 15, lg = (tid & 63) >> 4, rb = w & 3, cb = w >> 2;
    for (int step = 0; step < 2; ++step) {
        f64x4 aR = splat<f64x4>(0.0);
#pragma unroll 2
        for (int kk = 0; kk < 16; ++kk) {
            const int i = 16 * rb + lr, k = 4 * kk + lg, j = 16 * cb + lr;
            aR = MF<double>::mma(Z[k * LDA + i], Z[k * LDA + j], aR);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = 16 * rb + MF<double>::crow(lg, r), j = 16 * cb + lr;
            M[i * LDA + j] = (i == j ? 1.0 : 0.0) - aR[r];
        }
        __syncthreads();
        f64x4 aX = splat<f64x4>(0.0);
#pragma unroll 2
        for (int kk = 0; kk < 16; ++kk) {
            const int i = 16 * rb + lr, k = 4 * kk + lg, j = 16 * cb + lr;
            aX = MF<double>::mma(Z[i * LDA + k], M[k * LDA + j], aX);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 4; ++r) Z[(16 * rb + MF<double>::crow(lg, r)) * LDA + 16 * cb + lr] += 0.5 * aX[r];
        __syncthreads();
    }
}

// Batched evaluations: matrix m belongs to gene mat_gene[m], whose
// frequencies and scalar parameters (alpha, theta, delta, psi) it uses.
template <typename Args>
__device__ __forceinline__ void matrix_gene_view(Args &a, int m) {
    if (!a.mat_gene) return;
    const int g = a.mat_gene[m];
    a.pi += 64 * g;
    a.pos += 12 * g;
    a.pl.off_alpha += g;
    a.pl.off_theta += 5 * g;
    a.pl.off_delta += g;
    a.pl.off_psi += g;
}

__global__ __launch_bounds__(MODEL_NT) void k_model(ModelArgs args) {
    ModelArgs a = args;
    matrix_gene_view(a, blockIdx.x);
    constexpr int NT = MODEL_NT;
    const int m = blockIdx.x, tid = threadIdx.x, w = tid >> 6, lr = tid & 15, lg = (tid & 63) >> 4;
    const int o = a.oslot ? a.oslot[m] : m;  // output slot
    __shared__ double A[64 * LDA];
    __shared__ double Z[64 * LDA];
    __shared__ double T1[64 * LDA];
    __shared__ double T2[64 * LDA];
    __shared__ double sq[64], isq[64], lam[64];
    __shared__ double rc[32], rs[32], rt[32];
    __shared__ int rp[32], rq[32];
    __shared__ double red[4 * NT / 64];  // block_sum; two alternating buffers of block_sum2_once
    __shared__ int any_rot;
    __shared__ double thl[8], posl[12];
    __shared__ TriScratch tri;

    if (a.params_host)
        for (int i = m * NT + tid; i < a.n_params; i += gridDim.x * NT) a.params_dev[i] = a.params_host[i];
    const double *prm = a.params_host ? a.params_host : a.params;
    const double alpha = prm[a.pl.off_alpha], beta = prm[a.pl.off_omega + m];
    const double delta = prm[a.pl.off_delta], psi = prm[a.pl.off_psi];
    const int S = a.S;
    if (tid < 8) thl[tid] = tid < 5 ? prm[a.pl.off_theta + tid] : 1.0;  // index 5: A<->G
    if (tid < 12) posl[tid] = a.pos[tid];
    __syncthreads();

    // Q_ij from the packed pair descriptor (see QDESC_*): one coalesced load
    // per entry, exchangeabilities and position frequencies from LDS.
    for (int idx = tid; idx < 4096; idx += NT) {
        const int i = idx >> 6, j = idx & 63;
        const int d = a.qdesc[idx], nd = d & 3;
        double q = 0.0;
        if (i < S && j < S && i != j && nd > 0) {
            const double rm = (d & QDESC_SYN) ? alpha : beta;
            const double r0 = thl[(d >> 3) & 7], p0 = posl[(d >> 6) & 15];
            if (nd == 1) {
                q = rm * r0 * p0;
            } else {
                const double r1 = thl[(d >> 10) & 7], p1 = posl[(d >> 13) & 15];
                if (nd == 2) {
                    if (delta > 0.0) q = rm * delta * (r0 * r1) * (p0 * p1);
                } else if (psi > 0.0) {
                    const double r2 = thl[(d >> 17) & 7], p2 = posl[(d >> 20) & 15];
                    q = rm * psi * (r0 * r1 * r2) * (p0 * p1 * p2);
                }
            }
        }
        A[i * LDA + j] = q;
        // a wave holds one whole row (NT is a multiple of 64): Q_ii = -sum_j Q_ij
        double rsum = q;
        for (int off = 32; off > 0; off >>= 1) rsum += __shfl_xor(rsum, off);
        if (j == i) A[i * LDA + i] = -rsum;
    }
    __syncthreads();
    if (tid < 64) {
        const double p = a.pi[tid];
        if (tid < S) {
            const double v = sqrt(p > 1e-12 ? p : 1e-12);
            sq[tid] = v;
            isq[tid] = 1.0 / v;
        } else {
            sq[tid] = 0.0;
            isq[tid] = 0.0;
        }
    }
    __syncthreads();
    if (tid < 64) {
        double t = tid < S ? -a.pi[tid] * A[tid * LDA + tid] : 0.0;
        for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off);
        if (tid == 0) a.mu[o] = t;
    }
    for (int idx = tid; idx < 4096; idx += NT) {
        const int i = idx >> 6, j = idx & 63;
        Z[i * LDA + j] = sq[i] * A[i * LDA + j] * isq[j];
    }
    __syncthreads();
    double fro = 0.0;
    for (int idx = tid; idx < 4096; idx += NT) {
        const int i = idx >> 6, j = idx & 63;
        const double v = 0.5 * (Z[i * LDA + j] + Z[j * LDA + i]);
        T1[i * LDA + j] = v;
        fro += v * v;
    }
    const double anorm = sqrt(block_sum<NT>(fro, red, tid));
    for (int idx = tid; idx < 4096; idx += NT) {
        const int i = idx >> 6, j = idx & 63;
        A[i * LDA + j] = T1[i * LDA + j];
    }
    __syncthreads();

    int status = 0;
    bool done = false;
    double *Zg = a.Z + (size_t)o * 4096;
    const int zs = a.zsrc ? a.zsrc[m] : o;  // warm-start source slot (negative: cold)
    // Starting eigenvectors: the previous ones (warm start), else those of a
    // tridiagonal solve; either is refined by Ogita-Aishima steps and, when
    // that does not converge, the cold Jacobi solve below takes over.
    for (int att = (a.warm && zs >= 0) ? 0 : 1; att < 2 && !done; ++att) {
        if (att == 0) {
            const double *Zs = a.Z + (size_t)zs * 4096;
            for (int idx = tid; idx < 4096; idx += NT) Z[(idx >> 6) * LDA + (idx & 63)] = Zs[idx];
        } else {
            if (S < 3) break;
            tridiag_eigen<NT>(A, Z, T1, T2, lam, S, tid, tri);
        }
        __syncthreads();
        // ---- Ogita-Aishima refinement ----
        double prev_en = 1e300;
        // Wave w owns the 16x16 block (rows 16 rb.., columns 16 cb..) of
        // every product; S = X^T A X and X E stay in registers.
        const int rb = w & 3, cb = w >> 2;
        for (int it = 0; it < 8 && !done; ++it) {
            // X^T X and A X in one pass (shared X operand loads)
            f64x4 aR = splat<f64x4>(0.0), aAX = splat<f64x4>(0.0);
#pragma unroll 4
            for (int kk = 0; kk < 16; ++kk) {
                const int i = 16 * rb + lr, k = 4 * kk + lg, j = 16 * cb + lr;
                const double xb = Z[k * LDA + j];
                aR = MF<double>::mma(Z[k * LDA + i], xb, aR);
                aAX = MF<double>::mma(A[i * LDA + k], xb, aAX);
            }
            // R = I - X^T X into T1, A X into T2
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * rb + MF<double>::crow(lg, r), j = 16 * cb + lr;
                T1[i * LDA + j] = (i == j ? 1.0 : 0.0) - aR[r];
                T2[i * LDA + j] = aAX[r];
            }
            __syncthreads();
            // S = X^T (A X), kept in registers; l_i = s_ii / (1 - r_ii)
            f64x4 aS = splat<f64x4>(0.0);
#pragma unroll 4
            for (int kk = 0; kk < 16; ++kk) {
                const int i = 16 * rb + lr, k = 4 * kk + lg, j = 16 * cb + lr;
                aS = MF<double>::mma(Z[k * LDA + i], T2[k * LDA + j], aS);
            }
            double so = 0.0, rq = 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * rb + MF<double>::crow(lg, r), j = 16 * cb + lr;
                const double rr = T1[i * LDA + j];
                rq += rr * rr;
                if (i != j)
                    so += aS[r] * aS[r];
                else
                    lam[i] = aS[r] / (1.0 - rr);
            }
            // (the reduction's barrier also publishes lam and ends every read of T2)
            block_sum2_once<NT>(so, rq, red, tid);
            const double soff = sqrt(so);
            // Eigenvectors that already diagonalise A to 1e-12 (and are
            // orthonormal to 1e-12) are kept as they are: a step from there
            // only adds rounding noise divided by the eigenvalue gaps (and
            // repeated evaluations at the same parameters would accumulate it).
            if (it == 0 && soff <= 1e-12 * anorm && sqrt(rq) <= 1e-12) {
                done = true;
                status = 100;
                break;
            }
            // Pairs closer than the off-diagonal mass are treated as a cluster
            // (orthogonalised only).  The threshold leaves out Ogita-Aishima's
            // ||A|| ||R|| term: after an update R is second order, and that term
            // froze close pairs for an extra step; a step that does not
            // contract still falls back to the cold solve.
            const double dlt = 2.0 * soff;
            // E into T2, from S in registers
            double en = 0.0, fz = 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * rb + MF<double>::crow(lg, r), j = 16 * cb + lr;
                const double rr = T1[i * LDA + j];
                double e;
                if (i == j) {
                    e = 0.5 * rr;
                } else {
                    const double d = lam[j] - lam[i];
                    if (fabs(d) > dlt) {
                        e = (aS[r] + lam[j] * rr) / d;
                    } else {
                        e = 0.5 * rr;
                        fz += aS[r] * aS[r];
                    }
                }
                T2[i * LDA + j] = e;
                en += e * e;
            }
            block_sum2_once<NT>(en, fz, red + 2 * NT / 64, tid);
            const double enorm = sqrt(en), frozen = sqrt(fz);
            // X <- X + X E
            f64x4 aX = splat<f64x4>(0.0);
#pragma unroll 4
            for (int kk = 0; kk < 16; ++kk) {
                const int i = 16 * rb + lr, k = 4 * kk + lg, j = 16 * cb + lr;
                aX = MF<double>::mma(Z[i * LDA + k], T2[k * LDA + j], aX);
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < 4; ++r) Z[(16 * rb + MF<double>::crow(lg, r)) * LDA + 16 * cb + lr] += aX[r];
            __syncthreads();
            if (!(enorm < 0.5) || (it >= 3 && enorm > 0.5 * prev_en)) break;  // not converging: solve cold
            // Accept X + X E when X already diagonalised A (off-diagonal of
            // X^T A X below 1e-9 ||A||) and the correction is second-order
            // small, or when the steps contract quadratically (the next
            // correction, predicted as |E|^3 / |E_prev|^2, is below tol)
            // and no cluster hides an off-diagonal entry above tol ||A||.
            // tol = 1e-14 for FP64 evaluations, 1e-10 for FP32 ones (whose
            // transition matrices are rounded to FP32); the prediction is
            // trusted for corrections below 1e-5 (from a correction of 1e-4
            // FP32 gradients came out 6e-4 off their FP64 values at 64 taxa).
            const bool quad = it >= 1 && enorm < 1e-5 && enorm * enorm * enorm <= a.tol * prev_en * prev_en &&
                              frozen <= a.tol * anorm;
            prev_en = enorm;
            if ((soff <= 1e-9 * anorm && enorm < 1e-9) || quad) {
                done = true;
                status = 100 + it + 1;
            }
        }
        if (done) {
            if (att == 1) status -= 50;  // 50 + k: tridiagonal start refined k steps
            if (tid < 64) {
                const double l = lam[tid];
                lam[tid] = l > 0.0 ? 0.0 : l;
            }
            __syncthreads();
        }
    }
    if (!done) {
        // ---- last resort: parallel cyclic Jacobi ----
        for (int idx = tid; idx < 4096; idx += NT) Z[(idx >> 6) * LDA + (idx & 63)] = ((idx >> 6) == (idx & 63)) ? 1.0 : 0.0;
        const double thresh = 1e-17 * anorm;
        int sweep = 0;
        bool conv = false;
        __syncthreads();
        for (; sweep < 40 && !conv; ++sweep) {
            if (tid == 0) any_rot = 0;
            __syncthreads();
            for (int round = 0; round < 63; ++round) {
                int active = 0;
                if (tid < 32) {
                    int p, q;
                    if (tid == 0) {
                        p = 63;
                        q = round;
                    } else {
                        p = (round + tid) % 63;
                        q = (round - tid + 63) % 63;
                    }
                    const double apq = A[p * LDA + q];
                    double c = 1.0, s = 0.0, t = 0.0;
                    if (fabs(apq) > thresh) {
                        const double app = A[p * LDA + p], aqq = A[q * LDA + q];
                        const double theta = (aqq - app) / (2.0 * apq);
                        if (fabs(theta) > 1e150)
                            t = 0.5 / theta;
                        else
                            t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                        c = 1.0 / sqrt(t * t + 1.0);
                        s = t * c;
                        active = 1;
                        any_rot = 1;
                    }
                    rp[tid] = p;
                    rq[tid] = q;
                    rc[tid] = c;
                    rs[tid] = s;
                    rt[tid] = t;
                }
                if (!__syncthreads_or(active)) continue;
                // 2x2 blocks (I, J), I <= J, diagonal-major thread mapping.
                for (int b = tid; b < 1024; b += NT) {
                    const int I = b & 31, J = ((b >> 5) + I) & 31;
                    if (I > J) continue;
                    const double sI = rs[I], sJ = rs[J];
                    if (sI == 0.0 && sJ == 0.0) continue;
                    const double cI = rc[I], cJ = rc[J];
                    const int p = rp[I], q = rq[I], r = rp[J], s = rq[J];
                    if (I == J) {
                        const double apq = A[p * LDA + q], t = rt[I];
                        A[p * LDA + p] -= t * apq;
                        A[q * LDA + q] += t * apq;
                        A[p * LDA + q] = 0.0;
                        A[q * LDA + p] = 0.0;
                    } else {
                        const double xpr = A[p * LDA + r], xps = A[p * LDA + s], xqr = A[q * LDA + r], xqs = A[q * LDA + s];
                        const double ypr = cI * xpr - sI * xqr, yps = cI * xps - sI * xqs;
                        const double yqr = sI * xpr + cI * xqr, yqs = sI * xps + cI * xqs;
                        const double zpr = cJ * ypr - sJ * yps, zps = sJ * ypr + cJ * yps;
                        const double zqr = cJ * yqr - sJ * yqs, zqs = sJ * yqr + cJ * yqs;
                        A[p * LDA + r] = zpr; A[r * LDA + p] = zpr;
                        A[p * LDA + s] = zps; A[s * LDA + p] = zps;
                        A[q * LDA + r] = zqr; A[r * LDA + q] = zqr;
                        A[q * LDA + s] = zqs; A[s * LDA + q] = zqs;
                    }
                }
                for (int idx = tid; idx < 2048; idx += NT) {
                    const int k = idx & 31, i = idx >> 5;
                    const double s = rs[k];
                    if (s == 0.0) continue;
                    const double c = rc[k];
                    const int p = rp[k], q = rq[k];
                    const double zp = Z[i * LDA + p], zq = Z[i * LDA + q];
                    Z[i * LDA + p] = c * zp - s * zq;
                    Z[i * LDA + q] = s * zp + c * zq;
                }
                __syncthreads();
            }
            __syncthreads();
            conv = (any_rot == 0);
            __syncthreads();
        }
        status = conv ? sweep : -1;
        if (tid < 64) {
            const double l = A[tid * LDA + tid];
            lam[tid] = l > 0.0 ? 0.0 : l;
        }
        __syncthreads();
    }
    double *lo = a.lam + (size_t)o * 64, *U = a.U + (size_t)o * 4096, *V = a.V + (size_t)o * 4096;
    if (tid < 64) lo[tid] = lam[tid];
    for (int idx = tid; idx < 4096; idx += NT) {
        const int i = idx >> 6, k = idx & 63;
        const double z = Z[i * LDA + k];
        Zg[idx] = z;
        U[idx] = isq[i] * z;                       // U[i][k]
        V[idx] = Z[k * LDA + i] * sq[k];           // V[i][k] = Z[k][i] sqrt(pi_k)
    }
    if (tid == 0) a.status[o] = status;
}

// ---------------------------------------------------------------------------
// k_pmat: mixed transition matrices per branch, four workgroups per branch,
// each building one block of 16 rows of P.
// ---------------------------------------------------------------------------
// Rows 16 qr .. 16 qr + 15 of one branch's P (FP64 in LDS, row stride 64) in
// the layouts the pruning passes read: row-major rows with their row sums, the
// A-operand swizzle entries of P whose row lies in the block (PA[qr]) and of
// P^T whose column does (PTA entries with kk in [4 qr, 4 qr + 4)).
template <typename R>
__device__ __forceinline__ void emit_p_rows(const double *Pq, int qr, int node, int S, R *Prow_all, R *PA_all,
                                            R *PTA_all, int tid) {
    R *Prow = Prow_all + (size_t)node * PRS, *PA = PA_all + (size_t)node * 4096, *PTA = PTA_all + (size_t)node * 4096;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int e = tid + 256 * k;  // 0..1023
        Prow[16 * 64 * qr + e] = (R)Pq[e];
        // PA[qr][l][kk]: row 16 qr + (l & 15), column 4 kk + (l >> 4)
        const int l = e >> 4, kk = e & 15;
        PA[1024 * qr + e] = (R)Pq[(l & 15) * 64 + 4 * kk + (l >> 4)];
        // PTA[w][l][kk] with kk = 4 qr + (e & 3): P[col][row], col = 4 kk + (l >> 4)
        // in this block, row = 16 w + (l & 15)
        const int w = e >> 8, l2 = (e >> 2) & 63, kk2 = 4 * qr + (e & 3);
        const int col = 4 * kk2 + (l2 >> 4), row = 16 * w + (l2 & 15);
        PTA[1024 * w + 16 * l2 + kk2] = (R)Pq[(col - 16 * qr) * 64 + row];
    }
    if (tid < 16) {
        double s = 0.0;
        for (int j = 0; j < S; ++j) s += Pq[tid * 64 + j];
        Prow[4096 + 16 * qr + tid] = (R)s;
    }
}

// Row block qr of P_b = sum_j w_j max(U_j diag(e^{lambda_j tau}) V_j, 0):
// wave w forms the 16 x 16 block in columns 16 w .. 16 w + 15 on FP64 MFMA.
// Us: the block's 16 rows of U diag(e^{lambda tau}), then the finished rows
// (16 x LDA); Vs: V (64 x LDA).
template <typename R>
__device__ __forceinline__ void pmat_body(const PmatArgs<R> &a, int bi, int qr, double *Us, double *Vs) {
    const int tid = threadIdx.x, w = tid >> 6, lr = tid & 15, lg = (tid & 63) >> 4;
    const int node = a.blist[bi * BR_STRIDE], cls = a.blist[bi * BR_STRIDE + 3];
    __shared__ double ex[64];
    const double *prm = a.params;
    const int j0 = a.comp_ptr[cls], j1 = a.comp_ptr[cls + 1];
    double conv = 1.0;
    if (a.convention == 1) {
        double mb = 0.0;
        for (int j = j0; j < j1; ++j)
            if (a.comp_mat[j] >= 0) mb += prm[a.pl.off_w + j] * a.mu[a.comp_mat[j]];  // sink: no mu
        if (mb < 1e-12) mb = 1e-12;
        conv = 3.0 / mb;
    }
    const double tau = prm[a.pl.off_bl + node] * conv;
    if (tid == 0 && qr == 0) a.tau[node] = tau;
    f64x4 Pacc = splat<f64x4>(0.0);
    // U rows, V and lambda of the next component are requested before the
    // current component's product, so only the first component waits on them
    double ru[4], rv[16], rl = 0.0;
    auto fetch = [&](int j) {
        const int m = a.comp_mat[j];
#pragma unroll
        for (int k = 0; k < 4; ++k) ru[k] = a.U[(size_t)m * 4096 + 1024 * qr + tid + 256 * k];
#pragma unroll
        for (int k = 0; k < 16; ++k) rv[k] = a.V[(size_t)m * 4096 + tid + 256 * k];
        if (tid < 64) rl = a.lam[m * 64 + tid];
    };
    // error-sink components (comp_mat < 0) have no spectrum: skipped here,
    // added as w_j P_sink after the products
    auto next_mat = [&](int j) {
        while (j < j1 && a.comp_mat[j] < 0) ++j;
        return j;
    };
    int j = next_mat(j0);
    if (j < j1) fetch(j);
    while (j < j1) {
        const double wj = prm[a.pl.off_w + j];
        const int jn = next_mat(j + 1);
        __syncthreads();  // the previous component's operand reads are done
        if (tid < 64) ex[tid] = exp(rl * tau);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int e = tid + 256 * k;
            Us[(e >> 6) * LDA + (e & 63)] = ru[k] * ex[e & 63];
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = tid + 256 * k;
            Vs[(e >> 6) * LDA + (e & 63)] = rv[k];
        }
        if (jn < j1) fetch(jn);
        __syncthreads();
        f64x4 acc = splat<f64x4>(0.0);
#pragma unroll 4
        for (int kk = 0; kk < 16; ++kk) {
            const int k = 4 * kk + lg;
            acc = MF<double>::mma(Us[lr * LDA + k], Vs[k * LDA + 16 * w + lr], acc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) Pacc[r] += wj * (acc[r] > 0.0 ? acc[r] : 0.0);
        j = jn;
    }
    for (int js = j0; js < j1; ++js)
        if (a.comp_mat[js] < 0) {
            const double ws = prm[a.pl.off_w + js];
#pragma unroll
            for (int r = 0; r < 4; ++r)
                Pacc[r] += ws * a.sink[(16 * qr + MF<double>::crow(lg, r)) * 64 + 16 * w + lr];
        }
    __syncthreads();
    double *Pq = Us;  // the block's rows, row stride 64 (16 x 64 fits in 16 x LDA)
#pragma unroll
    for (int r = 0; r < 4; ++r) Pq[MF<double>::crow(lg, r) * 64 + 16 * w + lr] = Pacc[r];
    __syncthreads();
    emit_p_rows<R>(Pq, qr, node, a.S, a.Prow, a.PA, a.PTA, tid);
}

// Write one branch's P (FP64 in LDS) as row-major with row sums and the MFMA
// A-operand swizzles of P and P^T in the working type.
template <typename R>
__device__ __forceinline__ void emit_p(const double *P, int node, int S, R *Prow_all, R *PA_all, R *PTA_all,
                                       int tid) {
    R *Prow = Prow_all + (size_t)node * PRS, *PA = PA_all + (size_t)node * 4096, *PTA = PTA_all + (size_t)node * 4096;
    for (int idx = tid; idx < 4096; idx += 256) {
        Prow[idx] = (R)P[idx];
        const int w = idx >> 10, l = (idx >> 4) & 63, kk = idx & 15;  // [w][l][kk]
        const int row = 16 * w + (l & 15), col = 4 * kk + (l >> 4);
        PA[idx] = (R)P[row * 64 + col];
        PTA[idx] = (R)P[col * 64 + row];
    }
    if (tid < 64) {
        double s = 0.0;
        for (int j = 0; j < S; ++j) s += P[tid * 64 + j];
        Prow[4096 + tid] = (R)s;
    }
}

// Caller-supplied P matrices (row-major S x S per node, FP64).
template <typename R>
__global__ __launch_bounds__(256) void k_pmat_host(PmatArgs<R> a, const double *Phost) {
    const int bi = blockIdx.x, tid = threadIdx.x;
    const int node = a.blist[bi * BR_STRIDE];
    __shared__ double P[64 * 64];
    const int S = a.S;
    for (int idx = tid; idx < 4096; idx += 256) {
        const int i = idx >> 6, j = idx & 63;
        P[idx] = (i < S && j < S) ? Phost[((size_t)node * S + i) * S + j] : 0.0;
    }
    __syncthreads();
    emit_p<R>(P, node, S, a.Prow, a.PA, a.PTA, tid);
}

// ---------------------------------------------------------------------------
// Pruning helpers
// ---------------------------------------------------------------------------
template <typename R, int T>
__device__ __forceinline__ void load_tile(R *dst, const R *__restrict__ src, int tid) {
    constexpr int N = 64 * T;
    if constexpr (sizeof(R) == 4) {
        const float4 *s4 = reinterpret_cast<const float4 *>(src);
        float4 *d4 = reinterpret_cast<float4 *>(dst);
        for (int i = tid; i < N / 4; i += 256) d4[i] = s4[i];
    } else {
        const double2 *s2 = reinterpret_cast<const double2 *>(src);
        double2 *d2 = reinterpret_cast<double2 *>(dst);
        for (int i = tid; i < N / 2; i += 256) d2[i] = s2[i];
    }
}

// Schedule entries are read through the scalar cache (constant address
// space: the schedule is immutable while a kernel runs).
#define HY_CONST __attribute__((address_space(4)))
template <int N> struct Entry { int v[N]; };
template <int N> __device__ __forceinline__ Entry<N> load_entry(const int32_t *sched, int k) {
    const int32_t HY_CONST *p = (const int32_t HY_CONST *)(sched + (size_t)k * N);
    Entry<N> e;
#pragma unroll
    for (int i = 0; i < N; ++i) e.v[i] = p[i];
    return e;
}
// Vector-memory variant for prefetching: its completion is tracked apart
// from LDS traffic; uniform() moves an arrived entry to scalar registers.
template <int N> __device__ __forceinline__ Entry<N> load_entry_v(const int32_t *sched, int k) {
    const int4 *p = reinterpret_cast<const int4 *>(sched + (size_t)k * N);
    Entry<N> e;
#pragma unroll
    for (int i = 0; i < N / 4; ++i) {
        const int4 q = p[i];
        e.v[4 * i] = q.x;
        e.v[4 * i + 1] = q.y;
        e.v[4 * i + 2] = q.z;
        e.v[4 * i + 3] = q.w;
    }
    return e;
}
template <int N> __device__ __forceinline__ Entry<N> uniform(const Entry<N> &x) {
    Entry<N> e;
#pragma unroll
    for (int i = 0; i < N; ++i) e.v[i] = __builtin_amdgcn_readfirstlane(x.v[i]);
    return e;
}
// Materialises a uniform entry in scalar registers at this point of the
// program: the wait for its (vector) load is placed here, before loads issued
// later, instead of where the entry is first used.
template <int N> __device__ __forceinline__ void pin(const Entry<N> &e) {
#pragma unroll
    for (int i = 0; i < N; ++i) asm volatile("" ::"s"(e.v[i]));
}

// Column-wise power-of-two normalisation of a 64xT tile held in C layout by
// 4 waves (wave w holds rows 16w..16w+15).  ex[nb] receives the exponent e
// with max in [0.5, 1) after scaling by 2^-e (0 for an all-zero column).
// One barrier: consecutive calls must alternate between two cmax buffers.
template <typename R, int T>
__device__ __forceinline__ void normalize_cols(typename MF<R>::acc *x, int *ex, R *cmax, int w, int lr, int lg) {
    constexpr int NB = T / 16;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        R m = x[nb][0];
        m = m > x[nb][1] ? m : x[nb][1];
        m = m > x[nb][2] ? m : x[nb][2];
        m = m > x[nb][3] ? m : x[nb][3];
        R o = __shfl_xor(m, 16);
        m = m > o ? m : o;
        o = __shfl_xor(m, 32);
        m = m > o ? m : o;
        if (lg == 0) cmax[w * T + 16 * nb + lr] = m;
    }
    __syncthreads();
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        const int t = 16 * nb + lr;
        R m = cmax[t];
        m = m > cmax[T + t] ? m : cmax[T + t];
        m = m > cmax[2 * T + t] ? m : cmax[2 * T + t];
        m = m > cmax[3 * T + t] ? m : cmax[3 * T + t];
        int e = 0;
        if (m > (R)0) e = frexp_e(m);
        ex[nb] = e;
#pragma unroll
        for (int r = 0; r < 4; ++r) x[nb][r] = ldexp_r(x[nb][r], -e);
    }
}

// Edge tile of node pn in MFMA A-operand order: lane l of wave w reads its 16
// values as one contiguous run (PA/PTA layout [node][w][l][kk]).
template <typename R>
__device__ __forceinline__ void fetch_edge(R *dst, const R *__restrict__ tab, int pn, int w, int l) {
    const R *p = tab + (((size_t)pn * 4 + w) * 64 + l) * 16;
    if constexpr (sizeof(R) == 4) {
        const float4 *q = reinterpret_cast<const float4 *>(p);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float4 v = q[i];
            dst[4 * i] = v.x;
            dst[4 * i + 1] = v.y;
            dst[4 * i + 2] = v.z;
            dst[4 * i + 3] = v.w;
        }
    } else {
        const double2 *q = reinterpret_cast<const double2 *>(p);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const double2 v = q[i];
            dst[2 * i] = v.x;
            dst[2 * i + 1] = v.y;
        }
    }
}

// X = P * B for the wave's 16 rows: A operand in registers, B a 64xT tile
// (row stride T) in LDS or global memory.
template <typename R, int T>
__device__ __forceinline__ void mfma_tile(const R *av, const R *Bs, typename MF<R>::acc *X, int lr, int lg) {
    constexpr int NB = T / 16;
    typedef typename MF<R>::acc acc_t;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) X[nb] = splat<acc_t>(0.0);
#pragma unroll
    for (int kk = 0; kk < 16; ++kk)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) X[nb] = MF<R>::mma(av[kk], Bs[(4 * kk + lg) * T + 16 * nb + lr], X[nb]);
}

// Waits here for n loaded values (one wait for all of them), so that loads
// issued together are not split into a load-and-wait per register group.
template <typename R, int N>
__device__ __forceinline__ void arrive(const R *v) {
#pragma unroll
    for (int i = 0; i < N; ++i) asm volatile("" ::"v"(v[i]));
}

// Makes the MFMA results in x available before the next instruction, so
// that following loads may reuse the registers those MFMAs read.
template <typename R, int T>
__device__ __forceinline__ void consume(const typename MF<R>::acc *x) {
#pragma unroll
    for (int nb = 0; nb < T / 16; ++nb)
#pragma unroll
        for (int r = 0; r < 4; ++r) asm volatile("" ::"v"(x[nb][r]));
}

// X = P * B with the lane's B operands already in registers (bv[kk * NB + nb]).
template <typename R, int T>
__device__ __forceinline__ void mfma_regs(const R *av, const R *bv, typename MF<R>::acc *X) {
    constexpr int NB = T / 16;
    typedef typename MF<R>::acc acc_t;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) X[nb] = splat<acc_t>(0.0);
#pragma unroll
    for (int kk = 0; kk < 16; ++kk)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) X[nb] = MF<R>::mma(av[kk], bv[kk * NB + nb], X[nb]);
}

// Tip child: X = P_pn * indicator(code) in C layout, gathered from the
// row-major P (one column per definite codon) or its row sums (all-ones
// tips).  Only issues loads: lanes whose code is partially ambiguous get bit
// (slot * NB + nb) set in *amb_bits and are completed by fix_tip() when the
// operand is consumed, so no wait is placed on the gathers here.
template <typename R, int T, bool TL>
__device__ __forceinline__ unsigned tip_code(const PruneArgs<R> &a, const uint16_t *tcodes, int row, int p0, int t) {
    return TL ? tcodes[row * T + t] : a.tips[(size_t)row * a.P_pad + p0 + t];
}

// Tip codes of the lane's patterns for every child slot of a schedule entry
// (TL: from LDS, all slots read before the first use so that one wait covers
// them; non-tip slots read row 0).  gather_tip reads the code itself when
// given no precomputed codes.
template <typename R, int T, int NSLOT, int STRIDE>
__device__ __forceinline__ void slot_codes(unsigned (*cds)[T / 16], const int *ev, int nc, const uint16_t *tcodes,
                                           int lr) {
#pragma unroll
    for (int c = 0; c < NSLOT; ++c) {
        const int ref = ev[7 + STRIDE * c];
        const int row = (c < nc && (ref >> 28) == KIND_TIP) ? (ref & REF_MASK) : 0;
#pragma unroll
        for (int nb = 0; nb < T / 16; ++nb) cds[c][nb] = tcodes[row * T + 16 * nb + lr];
    }
}

template <typename R, int T, bool TL>
__device__ __forceinline__ void gather_tip(typename MF<R>::acc *x, unsigned &amb_bits, int slot, bool is_tip,
                                           const PruneArgs<R> &a, const uint16_t *tcodes, const unsigned *cds, int pn,
                                           int row, int p0, int w, int lr, int lg) {
    // Branch-free in the loaded values (every slot issues its loads, non-tip
    // slots from a harmless address) so that no register merge forces a wait
    // on the gathers before they are consumed.  One base per node (row-major
    // P followed by its row sums), 32-bit offsets.
    const bool gather = is_tip && pn >= 0;
    const R *P = a.Prow + (size_t)(gather ? pn : 0) * PRS;
#pragma unroll
    for (int nb = 0; nb < T / 16; ++nb) {
        const unsigned code = is_tip ? ((TL && cds) ? cds[nb] : tip_code<R, T, TL>(a, tcodes, row, p0, 16 * nb + lr)) : 0u;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int s = 16 * w + MF<R>::crow(lg, r);
            x[nb][r] = P[(gather && code < 64u) ? s * 64 + (int)code : 4096 + s];
        }
        if (is_tip && (pn < 0 || (code >= 64u && code != TIP_ONES))) amb_bits |= 1u << (slot * (T / 16) + nb);
    }
}

// Completes the partially ambiguous lanes of a gathered tip product:
// X[s][t] = sum of P[s][j] over the compatible codons j.
// pn < 0 (single-node tree): X is the tip's indicator itself.
template <typename R, int T, bool TL>
__device__ __forceinline__ void fix_tip(typename MF<R>::acc *x, unsigned amb_bits, int slot, const PruneArgs<R> &a,
                                        const uint16_t *tcodes, int pn, int row, int p0, int w, int lr, int lg,
                                        uint64_t all) {
#pragma unroll
    for (int nb = 0; nb < T / 16; ++nb) {
        if (amb_bits & (1u << (slot * (T / 16) + nb))) {
            const unsigned code = tip_code<R, T, TL>(a, tcodes, row, p0, 16 * nb + lr);
            if (pn < 0) {
                const uint64_t m = code < 64u ? (1ull << code) : (code == TIP_ONES ? all : a.amb[code & 0x7FFFu]);
#pragma unroll
                for (int r = 0; r < 4; ++r) x[nb][r] = (R)((m >> (16 * w + MF<R>::crow(lg, r))) & 1ull);
                continue;
            }
            const uint64_t m0 = a.amb[code & 0x7FFFu] & all;
            const R *P = a.Prow + (size_t)pn * PRS;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int s = 16 * w + MF<R>::crow(lg, r);
                R acc = 0;
                for (uint64_t m = m0; m; m &= m - 1) acc += P[s * 64 + __builtin_ctzll(m)];
                x[nb][r] = acc;
            }
        }
    }
}

template <typename R, int T>
__device__ __forceinline__ void store_c(R *dst, const typename MF<R>::acc *x, int w, int lr, int lg) {
#pragma unroll
    for (int nb = 0; nb < T / 16; ++nb)
#pragma unroll
        for (int r = 0; r < 4; ++r) dst[(16 * w + MF<R>::crow(lg, r)) * T + 16 * nb + lr] = x[nb][r];
}

template <typename R, int T>
__device__ __forceinline__ void load_c(typename MF<R>::acc *x, const R *src, int w, int lr, int lg) {
#pragma unroll
    for (int nb = 0; nb < T / 16; ++nb)
#pragma unroll
        for (int r = 0; r < 4; ++r) x[nb][r] = src[(16 * w + MF<R>::crow(lg, r)) * T + 16 * nb + lr];
}

// Inside partials in HBM use a "B-operand-major" tile layout: element
// (state s, pattern t) of a 64 x T tile sits at
//   ((t / 16 * 4 + s % 4) * 16 + t % 16) * 16 + s / 4,
// so the 16 MFMA B operands of lane (lg, lr) for pattern block nb
// (states 4kk + lg, kk = 0..15) are one contiguous run.
__device__ __forceinline__ int bmaj(int s, int t) { return bmaj_index(s, t); }

// B operands of lane (lg, lr) from a B-major tile: bv[kk * NB + nb].
template <typename R, int T>
__device__ __forceinline__ void fetch_b(R *bv, const R *__restrict__ tile, int lr, int lg) {
    constexpr int NB = T / 16;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        const R *p = tile + ((nb * 4 + lg) * 16 + lr) * 16;
        if constexpr (sizeof(R) == 4) {
            const float4 *q = reinterpret_cast<const float4 *>(p);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float4 v = q[i];
                bv[(4 * i) * NB + nb] = v.x;
                bv[(4 * i + 1) * NB + nb] = v.y;
                bv[(4 * i + 2) * NB + nb] = v.z;
                bv[(4 * i + 3) * NB + nb] = v.w;
            }
        } else {
            const double2 *q = reinterpret_cast<const double2 *>(p);
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const double2 v = q[i];
                bv[(2 * i) * NB + nb] = v.x;
                bv[(2 * i + 1) * NB + nb] = v.y;
            }
        }
    }
}

// C-layout values (wave w's rows) into a B-major tile.
template <typename R, int T>
__device__ __forceinline__ void store_bmaj(R *tile, const typename MF<R>::acc *x, int w, int lr, int lg) {
#pragma unroll
    for (int nb = 0; nb < T / 16; ++nb)
#pragma unroll
        for (int r = 0; r < 4; ++r) tile[bmaj(16 * w + MF<R>::crow(lg, r), 16 * nb + lr)] = x[nb][r];
}

template <int T, bool TL>
__device__ __forceinline__ void stage_tip_codes(uint16_t *tcodes, const uint16_t *tips, int rows, int P_pad, int p0,
                                                int tid) {
    if constexpr (TL)
        for (int idx = tid; idx < rows * T; idx += 256) tcodes[idx] = tips[(size_t)(idx / T) * P_pad + p0 + idx % T];
}

// Batched evaluations: the workgroup's tile belongs to one gene; shift every
// per-problem pointer and size of the arguments to that gene's slice and
// return the gene-local tile index.  Single problems pass no table.
template <typename R>
__device__ __forceinline__ int gene_view(PruneArgs<R> &a, int tile) {
    if (a.parts) {  // subtree part blockIdx.y of a single problem
        const int4 pp = a.parts[a.part0 + blockIdx.y];
        a.sched += (size_t)pp.x * SCHED_STRIDE;
        a.n_sched = pp.y;
        a.osched += (size_t)pp.z * OSCHED_STRIDE;
        a.n_osched = pp.w;
    }
    if (!a.genes) return tile;
    const int gi = a.tile_gene[tile];
    const GeneInfo g = a.genes[gi];
    a.sched += (size_t)g.sched_off * SCHED_STRIDE;
    a.n_sched = g.n_sched;
    a.osched += (size_t)g.osched_off * OSCHED_STRIDE;
    a.n_osched = g.n_osched;
    a.tip_rows = g.tip_rows;
    a.tips += g.tip_base;
    a.D += g.d_base;
    a.E += g.e_base;
    a.V += g.v_base;
    a.G += g.g_base;
    a.Uspill += g.u_base;
    a.Lroot += g.pbase;
    a.Eroot += g.pbase;
    a.site_ll += g.pbase;
    a.pi += 64 * gi;
    a.P_pad = g.P_pad;
    a.n_tiles = g.ntiles;
    a.n_patterns = g.P;
    return tile - g.tile0;
}

// ---------------------------------------------------------------------------
// k_inside: post-order pruning of one pattern tile.  Child partials stay in
// LDS slots assigned by the host schedule (read straight from HBM for nodes
// that did not get a slot); every node's partials are also written to HBM
// for the outside pass and the adjoint.  The next node's operands (edge
// tiles of internal children, gathered tip products) are fetched while the
// current node is computed; schedule entries arrive two nodes ahead.
// TL: the tile's tip codes are staged in LDS.
// ---------------------------------------------------------------------------
template <typename R, int T, bool TL, bool LEAN = false>
__global__ __launch_bounds__(256, LEAN ? 4 : ((sizeof(R) == 4 && T == 16) ? 2 : 1)) void k_inside(PruneArgs<R> args) {
    PruneArgs<R> a = args;
    const int tile = gene_view(a, blockIdx.x);
    constexpr int NB = T / 16, NS = inside_slots(T, sizeof(R), LEAN), TIPC = TL ? tip_lds_codes(T, sizeof(R)) : 1;
    constexpr bool PIPE = !LEAN && !(sizeof(R) == 8 && T == 64);  // register budget of double-buffered operands
    typedef Entry<SCHED_STRIDE> Ent;
    typedef typename MF<R>::acc acc_t;
    __shared__ R Ds[NS * 64 * T];
    __shared__ int Es[NS * T];
    __shared__ R cmax[2][4 * T];
    __shared__ double red[4 * T];
    __shared__ uint16_t tcodes[TIPC];
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, lr = l & 15, lg = l >> 4;
    const int p0 = tile * T, n = a.n_sched;
    stage_tip_codes<T, TL>(tcodes, a.tips, a.tip_rows, a.P_pad, p0, tid);
    const uint64_t all = (a.S >= 64) ? ~0ull : ((1ull << a.S) - 1ull);
    __syncthreads();
    auto operands = [&](const Ent &e, R (*pa)[16], acc_t (*xt)[NB], unsigned &ab) {
        const int nc = e.v[2];
        ab = 0;
        unsigned cds[MAXC][NB];
        if constexpr (TL) slot_codes<R, T, MAXC, 3>(cds, e.v, nc, tcodes, lr);
#pragma unroll
        for (int c = 0; c < MAXC; ++c) {
            const int pn = e.v[6 + 3 * c], ref = e.v[7 + 3 * c], kind = ref >> 28;
            if (c < nc && kind == KIND_INT) fetch_edge<R>(pa[c], a.PA, pn, w, l);
            if (c < nc && kind == KIND_TIP)
                gather_tip<R, T, TL>(xt[c], ab, c, true, a, tcodes, cds[c], pn, ref & REF_MASK, p0, w, lr, lg);
        }
    };
    Ent ec = load_entry<SCHED_STRIDE>(a.sched, 0);
    Ent en = load_entry<SCHED_STRIDE>(a.sched, n > 1 ? 1 : 0);
    R pa[MAXC][16], pa2[MAXC][16];
    acc_t xt[MAXC][NB], xt2[MAXC][NB];
    unsigned ab = 0, ab2 = 0;  // partially ambiguous tip lanes
    operands(ec, pa, xt, ab);
    int cb = 0;
    acc_t Dc[NB];
    int Ec[NB];
    for (int k = 0; k < n; ++k) {
        const Ent enn = load_entry_v<SCHED_STRIDE>(a.sched, k + 2 < n ? k + 2 : n - 1);
        if constexpr (PIPE) {
            if (k + 1 < n) operands(en, pa2, xt2, ab2);
        }
        const int gidx = ec.v[1], nc = ec.v[2], out = ec.v[3];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            Dc[nb] = splat<acc_t>(1.0);
            Ec[nb] = 0;
        }
#pragma unroll
        for (int c = 0; c < MAXC; ++c) {
            if (c < nc) {
                const int ref = ec.v[7 + 3 * c], src = ec.v[8 + 3 * c];
                const int kind = ref >> 28, idx = ref & REF_MASK;
                acc_t X[NB];
                if (kind == KIND_TIP) {
                    if (ab) fix_tip<R, T, TL>(xt[c], ab, c, a, tcodes, ec.v[6 + 3 * c], idx, p0, w, lr, lg, all);
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb) X[nb] = xt[c][nb];
                } else if (src >= 0) {  // partials in an LDS slot
                    mfma_tile<R, T>(pa[c], Ds + src * 64 * T, X, lr, lg);
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb) Ec[nb] += Es[src * T + 16 * nb + lr];
                } else {  // partials only in HBM
                    R bv[16 * NB];
                    fetch_b<R, T>(bv, a.D + ((size_t)idx * a.n_tiles + tile) * (64 * T), lr, lg);
                    if constexpr (sizeof(R) == 4) arrive<R, 16 * NB>(bv);  // (FP64: interleaved is faster)
                    mfma_regs<R, T>(pa[c], bv, X);
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb) Ec[nb] += a.E[(size_t)idx * a.P_pad + p0 + 16 * nb + lr];
                }
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) Dc[nb] *= X[nb];
                if (c >= 1 && c + 1 < nc) {
                    int ex[NB];
                    normalize_cols<R, T>(Dc, ex, cmax[cb], w, lr, lg);
                    cb ^= 1;
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb) Ec[nb] += ex[nb];
                }
            }
        }
        int ex[NB];
        normalize_cols<R, T>(Dc, ex, cmax[cb], w, lr, lg);
        cb ^= 1;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) Ec[nb] += ex[nb];
        store_bmaj<R, T>(a.D + ((size_t)gidx * a.n_tiles + tile) * (64 * T), Dc, w, lr, lg);
        if (out >= 0) store_c<R, T>(Ds + out * 64 * T, Dc, w, lr, lg);
        if (w == 0 && lg == 0) {
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                a.E[(size_t)gidx * a.P_pad + p0 + 16 * nb + lr] = Ec[nb];
                if (out >= 0) Es[out * T + 16 * nb + lr] = Ec[nb];
            }
        }
        __syncthreads();
        ec = en;
        en = uniform(enn);
        if constexpr (PIPE) {
#pragma unroll
            for (int c = 0; c < MAXC; ++c) {
#pragma unroll
                for (int kk = 0; kk < 16; ++kk) pa[c][kk] = pa2[c][kk];
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) xt[c][nb] = xt2[c][nb];
            }
            ab = ab2;
        } else {
            if (k + 1 < n) operands(ec, pa, xt, ab);
        }
    }
    if (!a.root_out) return;  // a subtree part: its root's partials are in HBM
    // Root likelihood: L_t = sum_s pi_s D_root[s][t] (FP64 accumulation).
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r) s += a.pi[16 * w + MF<R>::crow(lg, r)] * (double)Dc[nb][r];
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        if (lg == 0) red[w * T + 16 * nb + lr] = s;
    }
    __syncthreads();
    if (tid < T) {
        int Et = 0;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
            if (nb == lg) Et = Ec[nb];
        const double L = ((red[tid] + red[T + tid]) + red[2 * T + tid]) + red[3 * T + tid];
        const int p = p0 + tid;
        a.Lroot[p] = L;
        a.Eroot[p] = Et;
        if (p < a.n_patterns) a.site_ll[p] = (L > 0.0) ? log(L) + (double)Et * 0.69314718055994530942 : -1e20;
    }
}

// ---------------------------------------------------------------------------
// k_outside: pre-order outside messages V_c = U_u (.) prod_{siblings} X_s with
// X_s = P_s D_s, U_c = P_c^T V_c.  The message of the node processed next
// stays in registers, other pending messages in LDS slots (HBM when none is
// free).  The next node's operands are fetched once the current node's
// products are formed; with DREG the children's partials come along as MFMA
// B operands, otherwise they are read from HBM when used.
// REV: the edge matrices are reversible with respect to the root
// frequencies (every matrix built from the model), P^T = diag(pi) P
// diag(pi)^-1, so U_c = pi (.) (P_c (V_c / pi)) reuses the P_c operands of the
// X product; no transposed tiles are loaded, and the next node's P tiles go
// to a second register set.  Otherwise (caller-supplied matrices) U_c is
// formed from the transposed tiles.
// ---------------------------------------------------------------------------
template <typename R, int T, bool TL, bool REV, bool LEAN = false>
__global__ __launch_bounds__(256, LEAN ? 3 : 1) void k_outside(PruneArgs<R> args) {
    PruneArgs<R> a = args;
    const int tile = gene_view(a, blockIdx.x);
    constexpr int NB = T / 16, NS = outside_slots(T, sizeof(R), LEAN), TIPC = TL ? tip_lds_codes(T, sizeof(R)) : 1;
    constexpr bool DREG = !LEAN && NB * sizeof(R) <= 8;  // f32 T <= 32, f64 T == 16
    constexpr int DV = DREG ? 16 * NB : 1;
    typedef Entry<OSCHED_STRIDE> Ent;
    typedef typename MF<R>::acc acc_t;
    __shared__ R Us[NS * 64 * T];
    __shared__ int Hs[NS * T];
    __shared__ R Xc[MAXC * 64 * T];
    __shared__ R cmax[2][MAXC * 4 * T];
    __shared__ uint16_t tcodes[TIPC];
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, lr = l & 15, lg = l >> 4;
    const int p0 = tile * T, n = a.n_osched;
    stage_tip_codes<T, TL>(tcodes, a.tips, a.tip_rows, a.P_pad, p0, tid);
    const uint64_t all = (a.S >= 64) ? ~0ull : ((1ull << a.S) - 1ull);
    __syncthreads();
    // Register-prefetched operands cover the first RS children (one in the
    // lean variant); the host orders internal children first and tips last,
    // so further internal children are loaded when used.
    constexpr int RS = LEAN ? 1 : 2;
    // pa: P tiles of the current node's children; pt: their transposes
    // (!REV) or the next node's P tiles (REV; not in the lean variant, which
    // requests the next node's tiles into pa once the current ones are used)
    constexpr int PT = (REV && LEAN) ? 1 : RS;
    R pa[RS][16], pt[PT][16], dv[RS][DV];
    int de[MAXC][NB];
    acc_t xt[MAXC][NB];
    unsigned ab = 0;  // partially ambiguous tip lanes of the prefetched operands
    acc_t upi[NB], us[NB];  // root message; prefetched message of a node whose U was spilled
    int hs[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int r = 0; r < 4; ++r) upi[nb][r] = (R)a.pi[16 * w + MF<R>::crow(lg, r)];
    R ipi[4];  // 1 / pi of the lane's rows (0 where pi = 0)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const double pv = a.pi[16 * w + MF<R>::crow(lg, r)];
        ipi[r] = pv > 0.0 ? (R)(1.0 / pv) : (R)0;
    }
    auto operands = [&](const Ent &e) {
        const int nc = e.v[2];
        ab = 0;
        if (e.v[5] == U_SPILL) {  // message spilled to HBM by the parent
            load_c<R, T>(us, a.Uspill + ((size_t)e.v[1] * a.n_tiles + tile) * (64 * T), w, lr, lg);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) hs[nb] = a.G[(size_t)e.v[3] * a.P_pad + p0 + 16 * nb + lr];
        }
#pragma unroll
        for (int c = 0; c < MAXC; ++c) {
            const int pn = e.v[6 + 4 * c], ref = e.v[7 + 4 * c], kind = ref >> 28, idx = ref & REF_MASK;
            if (c < nc && kind == KIND_TIP)
                gather_tip<R, T, TL>(xt[c], ab, c, true, a, tcodes, nullptr, pn, idx, p0, w, lr, lg);
            if (c < nc) {
                if (kind != KIND_TIP) {
                    if (c < RS) {
                        fetch_edge<R>((REV && !LEAN) ? pt[c < PT ? c : 0] : pa[c], a.PA, pn, w, l);
                        if constexpr (DREG)
                            fetch_b<R, T>(dv[c], a.D + ((size_t)idx * a.n_tiles + tile) * (64 * T), lr, lg);
                    }
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb) de[c][nb] = a.E[(size_t)idx * a.P_pad + p0 + 16 * nb + lr];
                }
            }
        }
    };
    Ent ec = load_entry<OSCHED_STRIDE>(a.osched, 0);
    Ent en = load_entry<OSCHED_STRIDE>(a.osched, n > 1 ? 1 : 0);
    operands(ec);
    acc_t Un[NB];
    int Hn[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        Un[nb] = splat<acc_t>(0.0);
        Hn[nb] = 0;
    }
    int cb = 0;
    for (int k = 0; k < n; ++k) {
        if constexpr (REV && !LEAN)
#pragma unroll
            for (int c = 0; c < RS; ++c)
#pragma unroll
                for (int i = 0; i < 16; ++i) pa[c][i] = pt[c < PT ? c : 0][i];
        const Ent enn = load_entry_v<OSCHED_STRIDE>(a.osched, k + 2 < n ? k + 2 : n - 1);
        const int nc = ec.v[2], usrc = ec.v[5];
        bool any_int = false;
#pragma unroll
        for (int c = 0; c < MAXC; ++c)
            if (c < nc && (ec.v[7 + 4 * c] >> 28) == KIND_INT) any_int = true;
        acc_t U[NB];
        int H[NB];
        if (usrc == U_ROOT) {
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                U[nb] = upi[nb];
                H[nb] = 0;
            }
        } else if (usrc == U_NEXT) {
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                U[nb] = Un[nb];
                H[nb] = Hn[nb];
            }
        } else if (usrc >= 0) {
            load_c<R, T>(U, Us + usrc * 64 * T, w, lr, lg);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) H[nb] = Hs[usrc * T + 16 * nb + lr];
        } else {
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                U[nb] = us[nb];
                H[nb] = hs[nb];
            }
        }
        acc_t X[MAXC][NB];
        int Ex[MAXC][NB];
#pragma unroll
        for (int c = 0; c < MAXC; ++c) {
            if (c < nc) {
                const int ref = ec.v[7 + 4 * c];
                const int kind = ref >> 28, idx = ref & REF_MASK;
                if (kind == KIND_TIP) {
                    if (ab) fix_tip<R, T, TL>(xt[c], ab, c, a, tcodes, ec.v[6 + 4 * c], idx, p0, w, lr, lg, all);
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb) {
                        X[c][nb] = xt[c][nb];
                        Ex[c][nb] = 0;
                    }
                } else {
                    if (c < RS && DREG) {
                        mfma_regs<R, T>(pa[c < RS ? c : 0], dv[c < RS ? c : 0], X[c]);
                    } else {
                        // operands loaded at use (the next child's loads may
                        // reuse these registers once X is consumed)
                        R bv[16 * NB], av[16];
                        fetch_b<R, T>(bv, a.D + ((size_t)idx * a.n_tiles + tile) * (64 * T), lr, lg);
                        if (c < RS) {
                            mfma_regs<R, T>(pa[c < RS ? c : 0], bv, X[c]);
                        } else {
                            fetch_edge<R>(av, a.PA, ec.v[6 + 4 * c], w, l);
                            if constexpr (DREG) {  // rare path here; FP64 wide tiles: interleaved is faster
                                arrive<R, 16 * NB>(bv);
                                arrive<R, 16>(av);
                            }
                            mfma_regs<R, T>(av, bv, X[c]);
                        }
                        consume<R, T>(X[c]);
                    }
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb) Ex[c][nb] = de[c][nb];
                }
            }
        }
        acc_t V[MAXC][NB];
        int G[MAXC][NB];
        R *cm = cmax[cb];
#pragma unroll
        for (int c = 0; c < MAXC; ++c) {
            if (c < nc) {
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) {
                    V[c][nb] = U[nb];
                    G[c][nb] = H[nb];
                }
#pragma unroll
                for (int c2 = 0; c2 < MAXC; ++c2) {
                    if (c2 != c && c2 < nc) {
#pragma unroll
                        for (int nb = 0; nb < NB; ++nb) {
                            V[c][nb] *= X[c2][nb];
                            G[c][nb] += Ex[c2][nb];
                        }
                    }
                }
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) {
                    R mx = V[c][nb][0];
                    mx = mx > V[c][nb][1] ? mx : V[c][nb][1];
                    mx = mx > V[c][nb][2] ? mx : V[c][nb][2];
                    mx = mx > V[c][nb][3] ? mx : V[c][nb][3];
                    R o = __shfl_xor(mx, 16);
                    mx = mx > o ? mx : o;
                    o = __shfl_xor(mx, 32);
                    mx = mx > o ? mx : o;
                    if (lg == 0) cm[(c * 4 + w) * T + 16 * nb + lr] = mx;
                }
            }
        }
        // Loads into registers that earlier MFMAs read as operands are issued
        // only after those MFMAs' results have been consumed (X by the
        // products above, the previous U_c by U): transposed edge tiles of this
        // node and all operands of the next one.
        // entry k + 2 to scalar registers before the next operands are
        // requested, so that this wait does not cover them
        const Ent enu = uniform(enn);
        pin(enu);
        if constexpr (!REV)
#pragma unroll
            for (int c = 0; c < RS; ++c)
                if (c < nc && (ec.v[7 + 4 * c] >> 28) == KIND_INT) fetch_edge<R>(pt[c < PT ? c : 0], a.PTA, ec.v[6 + 4 * c], w, l);
        if constexpr (!LEAN)
            if (k + 1 < n) operands(en);
        __syncthreads();
#pragma unroll
        for (int c = 0; c < MAXC; ++c) {
            if (c < nc) {
                const int ref = ec.v[7 + 4 * c], vsc = ec.v[8 + 4 * c];
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) {
                    const int t = 16 * nb + lr;
                    const R *cc = cm + c * 4 * T;
                    R mx = cc[t];
                    mx = mx > cc[T + t] ? mx : cc[T + t];
                    mx = mx > cc[2 * T + t] ? mx : cc[2 * T + t];
                    mx = mx > cc[3 * T + t] ? mx : cc[3 * T + t];
                    const int ex = (mx > (R)0) ? frexp_e(mx) : 0;
#pragma unroll
                    for (int r = 0; r < 4; ++r) V[c][nb][r] = ldexp_r(V[c][nb][r], -ex);
                    G[c][nb] += ex;
                }
                store_c<R, T>(a.V + ((size_t)vsc * a.n_tiles + tile) * (64 * T), V[c], w, lr, lg);
                if ((ref >> 28) == KIND_INT) {
                    if constexpr (REV) {
                        acc_t Vs_[NB];
#pragma unroll
                        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                            for (int r = 0; r < 4; ++r) Vs_[nb][r] = V[c][nb][r] * ipi[r];
                        store_c<R, T>(Xc + c * 64 * T, Vs_, w, lr, lg);
                    } else {
                        store_c<R, T>(Xc + c * 64 * T, V[c], w, lr, lg);
                    }
                }
                if (w == 0 && lg == 0) {
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb) a.G[(size_t)vsc * a.P_pad + p0 + 16 * nb + lr] = G[c][nb];
                }
            }
        }
        cb ^= 1;
        if (any_int) __syncthreads();
#pragma unroll
        for (int c = 0; c < MAXC; ++c) {
            if (c < nc) {
                const int ref = ec.v[7 + 4 * c], dst = ec.v[9 + 4 * c];
                const int kind = ref >> 28, idx = ref & REF_MASK;
                if (kind != KIND_TIP) {
                    acc_t Uc[NB];
                    if (c < RS) {
                        mfma_tile<R, T>(REV ? pa[c < RS ? c : 0] : pt[c < PT ? c : 0], Xc + c * 64 * T, Uc, lr, lg);
                    } else {
                        R av[16];
                        fetch_edge<R>(av, REV ? a.PA : a.PTA, ec.v[6 + 4 * c], w, l);
                        mfma_tile<R, T>(av, Xc + c * 64 * T, Uc, lr, lg);
                    }
                    if constexpr (REV)
#pragma unroll
                        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                            for (int r = 0; r < 4; ++r) Uc[nb][r] *= upi[nb][r];
                    if (dst == U_NEXT) {
#pragma unroll
                        for (int nb = 0; nb < NB; ++nb) {
                            Un[nb] = Uc[nb];
                            Hn[nb] = G[c][nb];
                        }
                    } else if (dst >= 0) {
                        store_c<R, T>(Us + dst * 64 * T, Uc, w, lr, lg);
                        if (w == 0 && lg == 0) {
#pragma unroll
                            for (int nb = 0; nb < NB; ++nb) Hs[dst * T + 16 * nb + lr] = G[c][nb];
                        }
                    } else {
                        store_c<R, T>(a.Uspill + ((size_t)idx * a.n_tiles + tile) * (64 * T), Uc, w, lr, lg);
                    }
                }
            }
        }
        // lean: the next node's operands once this node's U products have
        // consumed the P tiles in pa
        if constexpr (LEAN)
            if (k + 1 < n) operands(en);
        ec = en;
        en = enu;
    }
}

// ---------------------------------------------------------------------------
// k_adjoint: A_b partial = sum_{p in chunk} c_p V_b[:,p] D_b[:,p]^T with
// c_p = w_p / L_p * 2^(G_b + E_b - E_root).  MFMA in the working type within
// a tile, FP64 accumulation across tiles.
// ---------------------------------------------------------------------------
// Adjoint operand tiles in LDS: (state s, pattern p of a 64-pattern group) at
// s * ADJ_LD + (p % 4) * 16 + p / 4, so the 16 patterns 4 kk + lg a lane feeds
// to the MFMA are contiguous; the 68-element row stride staggers rows over
// the LDS banks.
constexpr int ADJ_LD = 68;
__device__ __forceinline__ int adj_idx(int s, int p) { return s * ADJ_LD + (p & 3) * 16 + (p >> 2); }

// 16 contiguous, 16-byte aligned elements (LDS or global) into registers
// (128-bit loads).
template <typename R>
__device__ __forceinline__ void lds_run16(R *dst, const R *src) {
    if constexpr (sizeof(R) == 4) {
        const float4 *q = reinterpret_cast<const float4 *>(src);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float4 v = q[i];
            dst[4 * i] = v.x;
            dst[4 * i + 1] = v.y;
            dst[4 * i + 2] = v.z;
            dst[4 * i + 3] = v.w;
        }
    } else {
        const double2 *q = reinterpret_cast<const double2 *>(src);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const double2 v = q[i];
            dst[2 * i] = v.x;
            dst[2 * i + 1] = v.y;
        }
    }
}

template <typename R>
__device__ __forceinline__ R tip_onehot(unsigned code, int j, const uint64_t *amb, int S) {
    if (code < 64u) return (R)(code == (unsigned)j ? 1 : 0);
    if (code == TIP_ONES) return (R)(j < S ? 1 : 0);
    return (R)((amb[code & 0x7FFFu] >> j) & 1ull);
}

// Ambiguity masks staged in LDS by the adjoint (larger tables: AMBG variant,
// read from HBM).
constexpr int AMB_LDS = 512;

template <typename R, int T, bool AMBG>
__global__ __launch_bounds__(256) void k_adjoint(AdjArgs<R> args) {
    AdjArgs<R> a = args;
    // Works on groups of 64 patterns (G pruning tiles, contiguous in HBM):
    // the next group is fetched into registers while the current one is
    // multiplied from LDS (double-buffered).  In LDS both operands are stored
    // transposed to lane-contiguous runs (adj_idx): each lane reads its 16
    // MFMA operands per block with 128-bit loads, bank-conflict free.
    constexpr int G = 64 / T, PER = 4096 / 256;
    constexpr int CBI = sizeof(R) == 4 ? 4 : 2;  // column blocks whose B operands are held at once
    typedef typename MF<R>::acc acc_t;
    __shared__ __align__(16) R Vs[2][64 * ADJ_LD];
    __shared__ __align__(16) R Ds[2][64 * ADJ_LD];
    __shared__ __align__(16) R cs[2][64];
    __shared__ uint64_t ambl[AMBG ? 1 : AMB_LDS];
    const int chunk = blockIdx.x, bi = a.bsel ? a.bsel[blockIdx.y] : (int)blockIdx.y;
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, lr = l & 15, lg = l >> 4;
    const int32_t *br = a.blist + bi * BR_STRIDE;
    if (a.genes) {  // batched evaluation: this branch's gene slice
        const GeneInfo g = a.genes[br[4]];
        a.D += g.d_base;
        a.V += g.v_base;
        a.E += g.e_base;
        a.G += g.g_base;
        a.Eroot += g.pbase;
        a.Lroot += g.pbase;
        a.weights += g.pbase;
        a.tips += g.tip_base;
        a.P_pad = g.P_pad;
        a.n_tiles = g.ntiles;
        a.tiles_per_chunk = g.tiles_per_chunk;
    }
    const int vslot = br[1], ref = br[2];
    const int kind = ref >> 28, idx = ref & REF_MASK;
    const bool tip = kind == KIND_TIP;
    const uint64_t all = (a.S >= 64) ? ~0ull : ((1ull << a.S) - 1ull);
    const int t0 = chunk * a.tiles_per_chunk;
    const int t1 = min(t0 + a.tiles_per_chunk, a.n_tiles);
    const int ng = (t1 - t0 + G - 1) / G;
    // Two register sets of fetched groups: group g + 2 is requested while
    // group g is multiplied, so two groups' loads are in flight.  A fetch only
    // issues loads, all from valid addresses (runs past the chunk re-read its
    // first tile and get c_p = 0), so that no use or register merge of a
    // loaded value makes the wave wait for them before the group is stashed.
    struct Regs {
        R v[PER], d[PER];
        double L, wt;
        int gv, er, ev, pv;
        unsigned tcode;  // tips: code of the pattern of this thread's D elements
    };
    const R *Vbr = a.V + (size_t)vslot * a.n_tiles * (64 * T);
    const R *Dbr = tip ? Vbr : a.D + (size_t)idx * a.n_tiles * (64 * T);  // tips: D unused
    const int32_t *Gr = a.G + (size_t)vslot * a.P_pad;
    const int32_t *Er = a.E + (size_t)(tip ? 0 : idx) * a.P_pad;
    const uint16_t *Cr = a.tips + (size_t)(tip ? idx : 0) * a.P_pad;
    auto fetch = [&](Regs &x, int g) {
        const int tg0 = t0 + g * G;
        {
            // every wave issues the same loads (waves 1..3 duplicate wave 0's)
            const int p = tg0 * T + l;
            x.pv = p < t1 * T;
            const int pc = x.pv ? p : tg0 * T;
            x.L = a.Lroot[pc];
            x.wt = a.weights[pc];
            x.gv = Gr[pc];
            x.er = a.Eroot[pc];
            x.ev = Er[pc];
        }
        // each thread loads one contiguous run of PER elements (128-bit loads;
        // a run never crosses a tile)
        const int e0 = tid * PER;
        const int ec = (tg0 + e0 / (64 * T) < t1) ? e0 : e0 % (64 * T);
        lds_run16<R>(x.v, Vbr + (size_t)tg0 * (64 * T) + ec);
        lds_run16<R>(x.d, Dbr + (size_t)tg0 * (64 * T) + ec);
        // the 16 D elements of a thread share one pattern (B-operand-major tile)
        const int el = ec % (64 * T);
        x.tcode = Cr[(tg0 + ec / (64 * T)) * T + ((el >> 10) << 4) + ((el >> 4) & 15)];
    };
    auto stash = [&](const Regs &x, int buf) {
        // tips: D is the indicator of the pattern's code (all sense codons for
        // gaps, the ambiguity mask for partially ambiguous codes)
        uint64_t tmask = 0;
        if (tip) {
            const unsigned cd = x.tcode;
            tmask = cd < 64u ? (1ull << cd) : (cd == TIP_ONES ? all : 0ull);
            if (cd >= 64u && cd != TIP_ONES) tmask = AMBG ? a.amb[cd & 0x7FFFu] : ambl[cd & 0x7FFFu];
        }
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = tid * PER + q, ti = e / (64 * T), el = e % (64 * T);
            // V tiles are [state][pattern]; D tiles B-operand-major (bmaj_index)
            Vs[buf][adj_idx(el / T, ti * T + el % T)] = x.v[q];
            const int sd = ((el & 15) << 2) | ((el >> 8) & 3), td = ((el >> 10) << 4) | ((el >> 4) & 15);
            Ds[buf][adj_idx(sd, ti * T + td)] = tip ? (R)((tmask >> sd) & 1ull) : x.d[q];
        }
        if (tid < 64) {
            double c = 0.0;
            if (x.pv && x.wt > 0.0 && x.L > 0.0) c = ldexp(x.wt / x.L, x.gv - x.er + (tip ? 0 : x.ev));
            cs[buf][adj_idx(0, tid)] = (R)c;
        }
    };
    f64x4 acc[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[cb] = splat<f64x4>(0.0);
    auto multiply = [&](int buf) {
        acc_t loc[4];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) loc[cb] = splat<acc_t>(0.0);
        // lane (lr, lg) of wave w: patterns 4 kk + lg, kk = 0..15
        R av[16], cv[16];
        lds_run16<R>(av, &Vs[buf][adj_idx(16 * w + lr, lg)]);
        lds_run16<R>(cv, &cs[buf][adj_idx(0, lg)]);
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) av[kk] *= cv[kk];
#pragma unroll
        for (int c0 = 0; c0 < 4; c0 += CBI) {
            R bv[CBI][16];
#pragma unroll
            for (int c = 0; c < CBI; ++c) lds_run16<R>(bv[c], &Ds[buf][adj_idx(16 * (c0 + c) + lr, lg)]);
#pragma unroll
            for (int kk = 0; kk < 16; ++kk)
#pragma unroll
                for (int c = 0; c < CBI; ++c) loc[c0 + c] = MF<R>::mma(av[kk], bv[c][kk], loc[c0 + c]);
            // the next block's operands may reuse bv's registers: let these
            // MFMAs complete first
            if (c0 + CBI < 4)
#pragma unroll
                for (int c = 0; c < CBI; ++c)
#pragma unroll
                    for (int r = 0; r < 4; ++r) asm volatile("" ::"v"(loc[c0 + c][r]));
        }
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[cb][r] += (double)loc[cb][r];
    };
    if (!AMBG && tip) {
        for (int i = tid; i < a.n_amb; i += 256) ambl[i] = a.amb[i];
        __syncthreads();
    }
    // Group k travels in register set k % 2 and LDS buffer k % 2.  Fetches
    // past the last group re-read it (from L2) so that every iteration issues
    // the same loads: the compiler then waits for exactly the group a stash
    // needs and leaves the newer group's loads in flight.
    Regs x0, x1;
    if (ng > 0) {
        fetch(x0, 0);
        stash(x0, 0);
        fetch(x1, min(1, ng - 1));
        __syncthreads();
        for (int g = 0; g < ng; g += 2) {
            fetch(x0, min(g + 2, ng - 1));
            multiply(0);
            stash(x1, 1);
            __syncthreads();
            if (g + 1 >= ng) break;
            fetch(x1, min(g + 3, ng - 1));
            multiply(1);
            stash(x0, 0);
            __syncthreads();
        }
    }
    double *out = a.Apart + ((size_t)bi * a.n_chunks + chunk) * 4096;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int r = 0; r < 4; ++r) out[(16 * w + MF<R>::crow(lg, r)) * 64 + 16 * cb + lr] = acc[cb][r];
}

// ---------------------------------------------------------------------------
// Projections: M_phi = V_pi dQ_phi U_pi and dmu_phi = -sum_i pi_i dQ_phi[i][i]
// (likelihood.hpp:1497-1562).  phi 0 = omega, 1..5 = theta AC, AT, CG, CT, GT.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void proj_body(const ProjArgs &args, int m, int phi, int qc, double *X, double *Y) {
    // Column block qc of M_phi = V (dQ_phi U): Y_c = dQ_phi U[:, c] (64 x 16),
    // then V Y_c; no product is formed twice across the four blocks.
    // X: dQ_phi, then V (64 x LDA); Y: U[:, c], then Y_c (16 x LDA, transposed:
    // Y[j * LDA + k] = entry (k, j)).
    ProjArgs a = args;
    matrix_gene_view(a, m);
    const int tid = threadIdx.x, w = tid >> 6, lr = tid & 15, lg = (tid & 63) >> 4;
    __shared__ double thl[8], posl[12];
    const double *prm = a.params;
    const double alpha = prm[a.pl.off_alpha], beta = prm[a.pl.off_omega + m];
    const int S = a.S;
    // descriptors, U columns and V: every load issued before the first use
    int rd[16];
    double ru[4], rv[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        rd[k] = a.qdesc[tid + 256 * k];
        rv[k] = a.V[(size_t)m * 4096 + tid + 256 * k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int e = tid + 256 * k, row = e >> 4, col = 16 * qc + (e & 15);
        ru[k] = a.U[(size_t)m * 4096 + row * 64 + col];
    }
    if (tid < 8) thl[tid] = tid < 5 ? prm[a.pl.off_theta + tid] : 1.0;
    if (tid < 12) posl[tid] = a.pos[tid];
    __syncthreads();
    // dQ_phi: single-nucleotide changes only (likelihood.hpp:1529-1562);
    // phi = 0: omega (nonsynonymous pairs), phi = 1..5: theta_AC..theta_GT
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int idx = tid + 256 * k;
        const int i = idx >> 6, j = idx & 63, d = rd[k];
        double q = 0.0;
        if (i < S && j < S && i != j && (d & 3) == 1) {
            const int ti = (d >> 3) & 7;
            const double pf = posl[(d >> 6) & 15];
            if (phi == 0)
                q = (d & QDESC_SYN) ? 0.0 : thl[ti] * pf;
            else if (ti == phi - 1)
                q = ((d & QDESC_SYN) ? alpha : beta) * pf;
        }
        X[i * LDA + j] = q;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int e = tid + 256 * k;
        Y[(e & 15) * LDA + (e >> 4)] = ru[k];  // U[row][16 qc + j] at Y[j][row]
    }
    __syncthreads();
    if (tid < 64) {
        double s = 0.0;
        for (int j = 0; j < 64; ++j)
            if (j != tid) s += X[tid * LDA + j];
        X[tid * LDA + tid] = -s;
        double t = tid < S ? a.pi[tid] * s : 0.0;  // d mu / d phi = -sum_i pi_i dQ_ii
        for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off);
        if (tid == 0 && qc == 0) a.dmu[m * 6 + phi] = t;
    }
    __syncthreads();
    // Y_c = dQ U[:, c]: wave w forms rows 16 w .. 16 w + 15
    f64x4 acc = splat<f64x4>(0.0);
#pragma unroll 4
    for (int kk = 0; kk < 16; ++kk) {
        const int k = 4 * kk + lg;
        acc = MF<double>::mma(X[(16 * w + lr) * LDA + k], Y[lr * LDA + k], acc);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) Y[lr * LDA + 16 * w + MF<double>::crow(lg, r)] = acc[r];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = tid + 256 * k;
        X[(e >> 6) * LDA + (e & 63)] = rv[k];
    }
    __syncthreads();
    // M[:, c] = V Y_c: wave w forms rows 16 w .. 16 w + 15
    acc = splat<f64x4>(0.0);
#pragma unroll 4
    for (int kk = 0; kk < 16; ++kk) {
        const int k = 4 * kk + lg;
        acc = MF<double>::mma(X[(16 * w + lr) * LDA + k], Y[lr * LDA + k], acc);
    }
    double *M = a.M + ((size_t)m * 6 + phi) * 4096;
#pragma unroll
    for (int r = 0; r < 4; ++r) M[(16 * w + MF<double>::crow(lg, r)) * 64 + 16 * qc + lr] = acc[r];
}

// P builds and (with gradients) the dQ projections depend only on the
// spectra: one launch, four workgroups per branch build P (one row block
// each), then four per (matrix, parameter) project (one column block each).
template <typename R>
__global__ __launch_bounds__(256, 3) void k_pmat(PmatArgs<R> a, ProjArgs pj, int n_proj) {
    __shared__ double L0[64 * LDA], L1[16 * LDA];  // shared by both roles
    const int b = blockIdx.x >> 2, q = blockIdx.x & 3;
    if (b < a.n_branch)
        pmat_body<R>(a, b, q, L1, L0);
    else if (b - a.n_branch < n_proj)
        proj_body(pj, (b - a.n_branch) / 6, (b - a.n_branch) % 6, q, L0, L1);
}

// ---------------------------------------------------------------------------
// Fréchet stage (likelihood.hpp:1435-1520), in three kernels:
//   k_asum     A_b = sum over pattern chunks of the adjoint partials (fixed
//              order, deterministic)
//   k_frechet  per (matrix m, group of the (branch, component) pairs using m,
//              in fixed branch order): U_m, V_m stay in LDS while the
//              workgroup walks its pairs (the next A_b prefetched into
//              registers); for each, A~ = U_m^T A_b V_m^T
//              (two FP64 MFMA GEMMs), the component's share of
//              d lnL / d tau_b = w_j sum_i lambda_i e^{lambda_i tau} A~_ii and
//              <A_b, P_m(tau_b)> = sum_i e^{lambda_i tau} A~_ii, and the group's
//              partial H_m += w_j E o A~ with E_ij = (e^{l_i t} - e^{l_j t}) /
//              (l_i - l_j) (t e^{l_i t} when |l_i - l_j| < 1e-9), in registers
//   k_hproj    per (matrix m, 256-entry chunk): H_m = sum of the group
//              partials (fixed order), projected on M_phi = V dQ_phi U for
//              phi = omega, theta_AC..theta_GT
// ---------------------------------------------------------------------------
// GROUPED: the workgroup walks several pairs of one matrix with U and V
// resident in LDS (three 64 x 64 buffers, one workgroup per CU); otherwise it
// transforms one pair with two buffers (V, then U, reloaded), two per CU.
// Grid numbering of the Fréchet stage: workgroups go to the eight XCDs in
// turn (linear id mod 8), and each XCD has its own L2.  Linear id
// L = 8 ((g / 8) n_mat + m) + (g + r_m) % 8 puts group g of matrix m on XCD
// (g + r_m) % 8, with r_m the matrix's gene in a gene batch (0 otherwise).
// With a mixture whose classes use every matrix, group g of each matrix's
// pair list is the same branch, so the A_b those workgroups read (summed
// over its pattern chunks) comes from HBM once and from L2 after; the
// rotation by gene spreads batches, whose small genes have one group per
// matrix, over all eight XCDs (without it every gene's group 0 ran on one).
constexpr int FRECHET_XCD = 8;
inline int frechet_grid(int n_mat, int n_grp) {
    return (n_grp + FRECHET_XCD - 1) / FRECHET_XCD * FRECHET_XCD * n_mat;
}

template <bool GROUPED>
__global__ __launch_bounds__(256, GROUPED ? 1 : 2) void k_frechet(FrechetArgs a) {
    const int slot = blockIdx.x / FRECHET_XCD, x = blockIdx.x % FRECHET_XCD;
    const int blk = slot / a.n_mat, m = slot - blk * a.n_mat;
    const int rot = a.mat_gene ? a.mat_gene[m] % FRECHET_XCD : 0;
    const int grp = blk * FRECHET_XCD + (x + FRECHET_XCD - rot) % FRECHET_XCD;
    if (grp >= a.n_grp) return;  // padding of the last block of groups
    const int tid = threadIdx.x, w = tid >> 6, lr = tid & 15, lg = (tid & 63) >> 4;
    __shared__ double Bs[64 * LDA];                  // V (and U unless GROUPED)
    __shared__ double Us[GROUPED ? 64 * LDA : 1];    // U (GROUPED)
    __shared__ double As[64 * LDA];
    __shared__ double lam[64], ex[64];
    __shared__ double red[24];
    const int e0 = a.wl_ptr[m] + grp * a.grp_size, e1 = min(a.wl_ptr[m + 1], e0 + a.grp_size);
    if (e0 >= e1) {  // a matrix with fewer pairs than groups: zero projections
        if (tid < 6) a.proj[((size_t)m * 6 + tid) * a.n_grp + grp] = 0.0;
        return;
    }
    f64x4 H[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) H[q] = splat<f64x4>(0.0);
    // A_b of a pair, summed over its pattern chunks in chunk order: the loads
    // of up to four chunks are issued together (clamped to the last chunk, so
    // every load is valid and the code has no branches), then summed
    double ra[16];
    auto fetch_a = [&](int bi) {
        const double *src = a.Apart + (size_t)bi * a.n_chunks * 4096 + tid;
        const int nc = a.n_chunks;
        if (nc == 1) {  // one chunk per branch (gene batches): no clamped re-reads
#pragma unroll
            for (int k = 0; k < 16; ++k) ra[k] = src[256 * k];
            return;
        }
        double t[4][16];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int cc = c < nc ? c : nc - 1;
#pragma unroll
            for (int k = 0; k < 16; ++k) t[c][k] = src[(size_t)cc * 4096 + 256 * k];
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) ra[k] = t[0][k];
#pragma unroll
        for (int c = 1; c < 4; ++c)
            if (c < nc)
#pragma unroll
                for (int k = 0; k < 16; ++k) ra[k] += t[c][k];
        for (int c = 4; c < nc; ++c)
#pragma unroll
            for (int k = 0; k < 16; ++k) ra[k] += src[(size_t)c * 4096 + 256 * k];
    };
    // pair metadata; the chain slot -> class -> weight is scalar loads that
    // do not wait for the operand loads issued before it
    struct Pair {
        int bi, jj;
        double wj, tau;
    };
    auto pair_info = [&](int e, int bi) {
        Pair q;
        q.bi = bi;
        q.jj = a.wl_idx[e] - bi * a.max_comp;
        const int node = a.blist[bi * BR_STRIDE], cls = a.blist[bi * BR_STRIDE + 3];
        q.wj = a.params[a.pl.off_w + a.comp_ptr[cls] + q.jj];
        q.tau = a.tau[node];
        return q;
    };
    // operands of the first pair: V (and U when it stays in LDS) of the
    // matrix and the pair's A_b, all requested before any is used
    const int bi0 = a.wl_idx[e0] / a.max_comp;
    double rv[16], ru[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        rv[k] = a.V[(size_t)m * 4096 + tid + 256 * k];
        if constexpr (GROUPED) ru[k] = a.U[(size_t)m * 4096 + tid + 256 * k];
    }
    fetch_a(bi0);
    Pair cur = pair_info(e0, bi0);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int idx = tid + 256 * k;
        if constexpr (GROUPED) Us[(idx >> 6) * LDA + (idx & 63)] = ru[k];
        Bs[(idx >> 6) * LDA + (idx & 63)] = rv[k];
    }
    if (tid < 64) lam[tid] = a.lam[m * 64 + tid];
    for (int e = e0; e < e1; ++e) {
        const int bi = cur.bi, jj = cur.jj;
        const double wj = cur.wj, tau = cur.tau;
        __syncthreads();  // the previous pair is done with As and ex
        if (tid < 64) ex[tid] = exp(lam[tid] * tau);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int idx = tid + 256 * k;
            As[(idx >> 6) * LDA + (idx & 63)] = ra[k];
        }
        if (GROUPED && e + 1 < e1) {  // (one pair per workgroup otherwise)
            const int bn = a.wl_idx[e + 1] / a.max_comp;
            fetch_a(bn);
            cur = pair_info(e + 1, bn);
        }
        __syncthreads();
        // U over V in the second buffer once A V^T is formed: its loads are
        // in flight during that product
        if constexpr (!GROUPED)
#pragma unroll
            for (int k = 0; k < 16; ++k) ru[k] = a.U[(size_t)m * 4096 + tid + 256 * k];
        f64x4 acc[4];
        gemm64_acc<false, true>(As, Bs, acc, w, lr, lg);  // A V^T
        __syncthreads();
        gemm64_store(acc, As, w, lr, lg);
        if constexpr (!GROUPED)
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const int idx = tid + 256 * k;
                Bs[(idx >> 6) * LDA + (idx & 63)] = ru[k];
            }
        __syncthreads();
        gemm64_acc<true, false>(GROUPED ? Us : Bs, As, acc, w, lr, lg);  // A~ = U^T (A V^T), rows of this wave
        double d0 = 0.0, d1 = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * w + MF<double>::crow(lg, r), j = 16 * q + lr;
                const double at = acc[q][r], ei = ex[i], d = lam[i] - lam[j];
                if (i == j) {
                    d0 += ei * at;
                    d1 += lam[i] * ei * at;
                }
                const double E = (fabs(d) < 1e-9) ? tau * ei : (ei - ex[j]) / d;
                H[q][r] += wj * E * at;
            }
        block_sum2<256>(d0, d1, red, tid);
        if (tid == 0) {
            double *o = a.out + ((size_t)bi * a.max_comp + jj) * 8;
            o[0] = d0;
            o[1] = wj * d1;
        }
    }
    // <M_phi, H> of this group's partial H (the projection is linear, so
    // k_final sums the groups' projections in group order)
    const double *Mm = a.M + (size_t)m * 6 * 4096;
    double v[6];
#pragma unroll
    for (int phi = 0; phi < 6; ++phi) {
        v[phi] = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                v[phi] += Mm[(size_t)phi * 4096 + (16 * w + MF<double>::crow(lg, r)) * 64 + 16 * q + lr] * H[q][r];
    }
#pragma unroll
    for (int phi = 0; phi < 6; ++phi)
        for (int off = 32; off > 0; off >>= 1) v[phi] += __shfl_down(v[phi], off);
    __syncthreads();
    if ((tid & 63) == 0)
#pragma unroll
        for (int phi = 0; phi < 6; ++phi) red[6 * w + phi] = v[phi];
    __syncthreads();
    if (tid < 6) a.proj[((size_t)m * 6 + tid) * a.n_grp + grp] = ((red[tid] + red[6 + tid]) + red[12 + tid]) + red[18 + tid];
}

// ---------------------------------------------------------------------------
// k_final: deterministic reductions into the result vector, one workgroup of
// 16 waves.  lnL: per-thread compensated sums over contiguous pattern ranges,
// combined inside each wave by a butterfly of error-free additions and across
// waves in wave order.  With gradients the upper eight waves assemble them
// meanwhile: sums over branches with lanes striding branches and a fixed
// shuffle tree.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void neum(double &s, double &c, double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x))
        c += (s - t) + x;
    else
        c += (x - t) + s;
    s = t;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    return __shfl(v, 0);
}

// (s, c) += (s', c') with the rounding error of s + s' kept exactly (TwoSum);
// symmetric in its operands, so every lane of a butterfly gets the same sum.
__device__ __forceinline__ void comp_add(double &s, double &c, double s2, double c2) {
    const double t = s + s2, bp = t - s, ap = t - bp;
    c += c2 + ((s - ap) + (s2 - bp));
    s = t;
}

constexpr int FINAL_NT = 1024;
constexpr int FINAL_LDS_BR = 2048;  // branches whose per-branch terms k_final stages in LDS
__global__ __launch_bounds__(FINAL_NT) void k_final(FinalArgs args) {
    FinalArgs a = args;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // batched evaluation: workgroup g reduces gene g (its patterns, branches,
    // classes, components and matrices); global indices in the branch list and
    // component tables are offset by node0 / cls0 / comp0 / mat0
    int node0 = 0, cls0 = 0, comp0 = 0, mat0 = 0;
    if (a.genes) {
        const GeneInfo g = a.genes[blockIdx.x];
        a.site_ll += g.pbase;
        a.weights += g.pbase;
        a.Lroot += g.pbase;
        a.n_patterns = g.P;
        a.res += g.res_off;
        a.blist += (size_t)g.br0 * BR_STRIDE;
        a.fr += (size_t)g.br0 * a.max_comp * 8;
        a.n_branch = g.n_br;
        a.n_real_nodes = g.n_real_nodes;
        a.comp_ptr += g.cls0;
        a.n_class = g.n_cls;
        a.proj += (size_t)g.mat0 * 6 * a.proj_stride;
        a.pl.n_mat = g.n_mat;
        a.pl.n_comp = g.n_comp;
        node0 = g.node0;
        cls0 = g.cls0;
        comp0 = g.comp0;
        mat0 = g.mat0;
    }
    __shared__ double wsum[16], wcmp[16];
    __shared__ double sum_tau[64], mubar[64];
    // per branch (up to FINAL_LDS_BR): class, node, sum_j dlnL/dtau share and tau times it
    __shared__ int s_cls[FINAL_LDS_BR], s_node[FINAL_LDS_BR];
    __shared__ double s_dt[FINAL_LDS_BR], s_tdt[FINAL_LDS_BR];
    const bool grad = a.want_grad != 0;
    const bool staged = a.n_branch <= FINAL_LDS_BR;
    const int NL = grad ? FINAL_NT / 2 : FINAL_NT;  // threads on the lnL sum
    const int P = a.n_patterns;
    const double *prm = a.params;
    const int NR = a.n_real_nodes, NM = a.pl.n_mat, NCP = a.pl.n_comp, MC = a.max_comp;
    double *gbl = a.res + 1, *gom = gbl + NR, *gw = gom + NM, *gth = gw + NCP;
    const int gt = tid - NL, gwave = wave - NL / 64, NGW = (FINAL_NT - NL) / 64;  // gradient threads / waves
    if (tid < NL) {
        const int per = (P + NL - 1) / NL;
        double s = 0.0, c = 0.0;
        const int pe = min((tid + 1) * per, P);
        for (int p0 = tid * per; p0 < pe; p0 += 8) {
            // eight patterns' loads in flight (unconditional, clamped to the
            // range), then the compensated sums in order
            double x[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int p = p0 + u, pc = p < pe ? p : pe - 1;
                const double L = a.Lroot[pc], wv = a.weights[pc], sl = a.site_ll[pc];
                x[u] = (p < pe && !(grad && !(L > 0.0))) ? wv * sl : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (p0 + u < pe) neum(s, c, x[u]);
        }
        for (int off = 1; off < 64; off <<= 1) comp_add(s, c, __shfl_xor(s, off), __shfl_xor(c, off));
        if (lane == 0) {
            wsum[wave] = s;
            wcmp[wave] = c;
        }
    } else {
        if (gt < a.n_class) {
            double mb = 0.0;
            // error-sink components (comp_mat < 0) have no rate matrix and no mu
            for (int j = a.comp_ptr[gt]; j < a.comp_ptr[gt + 1]; ++j)
                if (a.comp_mat[j] >= 0) mb += prm[a.pl.off_w + j] * a.mu[a.comp_mat[j]];
            mubar[gt] = mb < 1e-12 ? 1e-12 : mb;
        }
        for (int n = gt; n < NR; n += FINAL_NT - NL) gbl[n] = 0.0;
        if (grad && staged)
            for (int bi = gt; bi < a.n_branch; bi += FINAL_NT - NL) {
                const int node = a.blist[bi * BR_STRIDE], cls = a.blist[bi * BR_STRIDE + 3] - cls0;
                // the first four components' shares are requested with the
                // branch's entry (they do not wait on its class), then summed
                // in component order
                double f4[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) f4[j] = j < MC ? a.fr[((size_t)bi * MC + j) * 8 + 1] : 0.0;
                double dt = 0.0;
                const int nj = a.comp_ptr[cls + 1] - a.comp_ptr[cls];  // components of this branch's class
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (j < nj) dt += f4[j];
                for (int j = 4; j < nj; ++j) dt += a.fr[((size_t)bi * MC + j) * 8 + 1];
                s_cls[bi] = cls;
                s_node[bi] = node - node0;
                s_dt[bi] = dt;
                s_tdt[bi] = a.tau[node] * dt;
            }
    }
    __syncthreads();
    if (grad && tid >= NL) {
        for (int cl = gwave; cl < a.n_class; cl += NGW) {
            double v = 0.0;
            for (int bi = lane; bi < a.n_branch; bi += 64) {
                if (staged) {
                    if (s_cls[bi] == cl) v += s_tdt[bi];
                    continue;
                }
                if (a.blist[bi * BR_STRIDE + 3] != cls0 + cl) continue;
                double dt = 0.0;
                const int nj = a.comp_ptr[cl + 1] - a.comp_ptr[cl];  // components of this branch's class
                for (int j = 0; j < nj; ++j) dt += a.fr[((size_t)bi * MC + j) * 8 + 1];
                v += a.tau[a.blist[bi * BR_STRIDE]] * dt;
            }
            v = wave_sum(v);
            if (lane == 0) sum_tau[cl] = v;
        }
        for (int bi = gt; bi < a.n_branch; bi += FINAL_NT - NL) {
            int node, cls;
            double dt = 0.0;
            if (staged) {
                node = s_node[bi];
                cls = s_cls[bi];
                dt = s_dt[bi];
            } else {
                node = a.blist[bi * BR_STRIDE] - node0;
                cls = a.blist[bi * BR_STRIDE + 3] - cls0;
                const int nj = a.comp_ptr[cls + 1] - a.comp_ptr[cls];
                for (int j = 0; j < nj; ++j) dt += a.fr[((size_t)bi * MC + j) * 8 + 1];
            }
            const double conv = (a.convention == 1) ? 3.0 / mubar[cls] : 1.0;
            if (node >= 0 && node < NR) gbl[node] = conv * dt;
        }
    }
    __syncthreads();
    if (tid == 0) {
        double S0 = 0.0, C0 = 0.0;
        for (int w = 0; w < NL / 64; ++w) comp_add(S0, C0, wsum[w], wcmp[w]);
        a.res[0] = S0 + C0;
    }
    if (grad && tid >= NL) {
        const int ntask = NM + 5 + NCP;
        for (int t = gwave; t < ntask; t += NGW) {
            double v = 0.0;
            if (t < NM + 5) {
                const bool is_om = t < NM;
                const int phi = is_om ? 0 : (t - NM + 1);
                // <M_{m,phi}, H_m> partials: one matrix for omega_t, all matrices for theta
                const int m0 = is_om ? t : 0, m1 = is_om ? t + 1 : NM;
                const int PS = a.proj_stride;
                for (int q = lane; q < (m1 - m0) * PS; q += 64) v += a.proj[((size_t)(m0 + q / PS) * 6 + phi) * PS + q % PS];
                v = wave_sum(v);
                if (lane == 0) {
                    if (a.convention == 1) {
                        for (int cl = 0; cl < a.n_class; ++cl) {
                            const double f = sum_tau[cl] / mubar[cl];
                            for (int j = a.comp_ptr[cl]; j < a.comp_ptr[cl + 1]; ++j) {
                                const int mm = a.comp_mat[j];
                                if (mm < 0 || (is_om && mm != mat0 + t)) continue;
                                v -= f * prm[a.pl.off_w + j] * a.dmu[mm * 6 + phi];
                            }
                        }
                    }
                    if (is_om)
                        gom[t] = v;
                    else
                        gth[phi - 1] = v;
                }
            } else {
                const int J = t - NM - 5, Jg = comp0 + J;
                int cls = 0;
                while (cls + 1 < a.n_class && a.comp_ptr[cls + 1] <= Jg) ++cls;
                const int jj = Jg - a.comp_ptr[cls];
                for (int bi = lane; bi < a.n_branch; bi += 64)
                    if (a.blist[bi * BR_STRIDE + 3] == cls0 + cls) v += a.fr[((size_t)bi * MC + jj) * 8 + 0];
                v = wave_sum(v);
                if (lane == 0) {
                    if (a.convention == 1 && a.comp_mat[Jg] >= 0) v -= sum_tau[cls] / mubar[cls] * a.mu[a.comp_mat[Jg]];
                    gw[J] = v;
                }
            }
        }
    }
    // single evaluation: every thread's result stores reach the host before
    // the call number (params[off_seq]) is published in the completion word
    if (a.done) {
        __threadfence_system();
        __syncthreads();
        if (tid == 0)
            __hip_atomic_store(a.done, (unsigned long long)prm[a.pl.off_seq], __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
void launch_model(const ModelArgs &a, int n_mat, hipStream_t s) {
    hipLaunchKernelGGL(k_model, dim3(n_mat), dim3(MODEL_NT), 0, s, a);
}

template <typename R> void launch_pmat(const PmatArgs<R> &a, const ProjArgs *pj, int n_mat, hipStream_t s) {
    const int n_proj = pj ? 6 * n_mat : 0;
    if (a.n_branch + n_proj > 0)
        hipLaunchKernelGGL(k_pmat<R>, dim3(4 * (a.n_branch + n_proj)), dim3(256), 0, s, a, pj ? *pj : ProjArgs{}, n_proj);
}

template <typename R, int T> void launch_inside_t(const PruneArgs<R> &a, int ny, hipStream_t s) {
    if (a.tip_rows * T <= tip_lds_codes(T, sizeof(R)))
        hipLaunchKernelGGL((k_inside<R, T, true>), dim3(a.n_tiles, ny), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_inside<R, T, false>), dim3(a.n_tiles, ny), dim3(256), 0, s, a);
}

// FP64 tiles stop at 32 patterns (the host never asks for more): the
// 64-pattern FP64 variants are not instantiated.
template <typename R> void launch_inside(const PruneArgs<R> &a, int T, bool lean, int ny, hipStream_t s) {
    if constexpr (sizeof(R) == 4) {
        if (lean && T == 16) {
            hipLaunchKernelGGL((k_inside<R, 16, false, true>), dim3(a.n_tiles, ny), dim3(256), 0, s, a);
            return;
        }
    }
    switch (T) {
        case 16: launch_inside_t<R, 16>(a, ny, s); break;
        case 32: launch_inside_t<R, 32>(a, ny, s); break;
        default:
            if constexpr (sizeof(R) == 4) launch_inside_t<R, 64>(a, ny, s);
            break;
    }
}

template <typename R, int T, bool REV> void launch_outside_t(const PruneArgs<R> &a, int ny, hipStream_t s) {
    if (a.tip_rows * T <= tip_lds_codes(T, sizeof(R)))
        hipLaunchKernelGGL((k_outside<R, T, true, REV>), dim3(a.n_tiles, ny), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_outside<R, T, false, REV>), dim3(a.n_tiles, ny), dim3(256), 0, s, a);
}

template <typename R> void launch_outside(const PruneArgs<R> &a, int T, bool rev, bool lean, int ny, hipStream_t s) {
    if constexpr (sizeof(R) == 4) {
        if (lean && T == 16) {
            if (rev)
                hipLaunchKernelGGL((k_outside<R, 16, false, true, true>), dim3(a.n_tiles, ny), dim3(256), 0, s, a);
            else
                hipLaunchKernelGGL((k_outside<R, 16, false, false, true>), dim3(a.n_tiles, ny), dim3(256), 0, s, a);
            return;
        }
    }
    switch (T) {
        case 16: rev ? launch_outside_t<R, 16, true>(a, ny, s) : launch_outside_t<R, 16, false>(a, ny, s); break;
        case 32: rev ? launch_outside_t<R, 32, true>(a, ny, s) : launch_outside_t<R, 32, false>(a, ny, s); break;
        default:
            if constexpr (sizeof(R) == 4)
                rev ? launch_outside_t<R, 64, true>(a, ny, s) : launch_outside_t<R, 64, false>(a, ny, s);
            break;
    }
}

// a.n_branch workgroup rows (branch list entries, or a.bsel's selection)
template <typename R> void launch_adjoint(const AdjArgs<R> &a, int T, hipStream_t s) {
    const dim3 grid(a.n_chunks, a.n_branch);
    if (a.n_amb <= AMB_LDS) switch (T) {
        case 16: hipLaunchKernelGGL((k_adjoint<R, 16, false>), grid, dim3(256), 0, s, a); break;
        case 32: hipLaunchKernelGGL((k_adjoint<R, 32, false>), grid, dim3(256), 0, s, a); break;
        default:
            if constexpr (sizeof(R) == 4) hipLaunchKernelGGL((k_adjoint<R, 64, false>), grid, dim3(256), 0, s, a);
            break;
    }
    else switch (T) {
        case 16: hipLaunchKernelGGL((k_adjoint<R, 16, true>), grid, dim3(256), 0, s, a); break;
        case 32: hipLaunchKernelGGL((k_adjoint<R, 32, true>), grid, dim3(256), 0, s, a); break;
        default:
            if constexpr (sizeof(R) == 4) hipLaunchKernelGGL((k_adjoint<R, 64, true>), grid, dim3(256), 0, s, a);
            break;
    }
}

template <typename R> void launch_pmat_host(const PmatArgs<R> &a, const double *Phost, hipStream_t s) {
    hipLaunchKernelGGL(k_pmat_host<R>, dim3(a.n_branch), dim3(256), 0, s, a, Phost);
}


// BUSTED's error sink (busted.hpp:208-216 evaluates V^T P_sink D per
// pattern): the weight gradient of a sink component is <A_b, P_sink> with the
// branch's adjoint A_b summed over its pattern chunks in chunk order.
__global__ __launch_bounds__(256) void k_sink(SinkArgs a) {
    const int tid = threadIdx.x, slot = a.slots[blockIdx.x], bi = slot / a.max_comp;
    __shared__ double red[24];
    const double *src = a.Apart + (size_t)bi * a.n_chunks * 4096 + tid;
    double d0 = 0.0, d1 = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        double v = src[256 * k];
        for (int c = 1; c < a.n_chunks; ++c) v += src[(size_t)c * 4096 + 256 * k];
        d0 += v * a.sink[tid + 256 * k];
    }
    block_sum2<256>(d0, d1, red, tid);
    if (tid == 0) {
        a.out[(size_t)slot * 8] = d0;
        a.out[(size_t)slot * 8 + 1] = 0.0;
    }
}

void launch_sink(const SinkArgs &a, int n_slots, hipStream_t s) {
    hipLaunchKernelGGL(k_sink, dim3(n_slots), dim3(256), 0, s, a);
}

void launch_frechet(const FrechetArgs &a, hipStream_t s) {
    const int n = frechet_grid(a.n_mat, a.n_grp);
    if (a.grp_size > 1)
        hipLaunchKernelGGL(k_frechet<true>, dim3(n), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_frechet<false>, dim3(n), dim3(256), 0, s, a);
}

void launch_final(const FinalArgs &a, int n_genes, hipStream_t s) {
    hipLaunchKernelGGL(k_final, dim3(n_genes), dim3(FINAL_NT), 0, s, a);
}

// ---------------------------------------------------------------------------
// Nucleotide GTR stage (S = 4; compute_gtr_log_likelihood,
// likelihood.hpp:181-254).  Four states are no MFMA shape: one thread per
// pattern walks the tree with 4x4 products in registers; partials and
// messages go to HBM pattern-minor ([slot][state][pattern], coalesced).  Every
// node's column is normalised to max in [0.5, 1) by an exact power of two
// (exponent kept per node and pattern), as in the codon passes.
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ void gtr_tip(unsigned code, const uint64_t *amb, R *d) {
    if (code < 4u) {
#pragma unroll
        for (int i = 0; i < 4; ++i) d[i] = (unsigned)i == code ? R(1) : R(0);
    } else if (code == TIP_ONES) {
#pragma unroll
        for (int i = 0; i < 4; ++i) d[i] = R(1);
    } else {
        const uint64_t m = amb[code & 0x7FFFu];
#pragma unroll
        for (int i = 0; i < 4; ++i) d[i] = ((m >> i) & 1ull) ? R(1) : R(0);
    }
}

// The child's partial (mantissa d, exponent e) and X = P_c d.
template <typename R>
__device__ __forceinline__ void gtr_child(const GtrArgs<R> &a, int cn, int p, R *d, int &e, R *X) {
    const int ref = a.nref[cn], kind = ref >> 28, idx = ref & REF_MASK;
    if (kind == KIND_TIP) {
        gtr_tip<R>(a.tips[(size_t)idx * a.P_pad + p], a.amb, d);
        e = 0;
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) d[i] = a.D[((size_t)idx * 4 + i) * a.P_pad + p];
        e = a.E[(size_t)idx * a.P_pad + p];
    }
    const R *P = a.Pm + (size_t)cn * 16;
#pragma unroll
    for (int i = 0; i < 4; ++i) X[i] = P[4 * i] * d[0] + P[4 * i + 1] * d[1] + P[4 * i + 2] * d[2] + P[4 * i + 3] * d[3];
}

template <typename R>
__device__ __forceinline__ void gtr_normalise(R *x, int &e) {
    const R m = fmax(fmax(x[0], x[1]), fmax(x[2], x[3]));
    if (m > R(0)) {
        const int k = frexp_e(m);
#pragma unroll
        for (int i = 0; i < 4; ++i) x[i] = ldexp_r(x[i], -k);
        e += k;
    }
}

template <typename R>
__global__ __launch_bounds__(GTR_BLOCK) void k_gtr_inside(GtrArgs<R> a) {
    const int p = blockIdx.x * GTR_BLOCK + threadIdx.x;  // < P_pad (grid covers P_pad exactly)
    R acc[4];
    int ex = 0;
    for (int k = 0; k < a.n_sched; ++k) {
        const int32_t *e = a.sched + (size_t)k * GTR_STRIDE;
        const int slot = e[1], nc = e[2];
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = R(1);
        ex = 0;
        for (int c = 0; c < nc; ++c) {
            R d[4], X[4];
            int ec;
            gtr_child<R>(a, e[3 + c], p, d, ec, X);
#pragma unroll
            for (int i = 0; i < 4; ++i) acc[i] *= X[i];
            ex += ec;
        }
        gtr_normalise<R>(acc, ex);
#pragma unroll
        for (int i = 0; i < 4; ++i) a.D[((size_t)slot * 4 + i) * a.P_pad + p] = acc[i];
        a.E[(size_t)slot * a.P_pad + p] = ex;
    }
    // the last entry is the root
    double L = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) L += a.pi[i] * (double)acc[i];
    a.Lroot[p] = L;
    a.Eroot[p] = ex;
    if (p < a.P) a.site_ll[p] = L > 0.0 ? log(L) + (double)ex * 0.69314718055994530942 : -1e20;
}

template <typename R>
__global__ __launch_bounds__(GTR_BLOCK) void k_gtr_outside(GtrArgs<R> a) {
    const int p = blockIdx.x * GTR_BLOCK + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = p >> 6;
    const double L = a.Lroot[p];
    const int Er = a.Eroot[p];
    // c_p = w_p / L_p (zero for padding and for patterns with L <= 0, which
    // contribute no gradient), the exponents applied per branch
    const double cw = (p < a.P && L > 0.0) ? a.weights[p] / L : 0.0;
    for (int k = a.n_sched - 1; k >= 0; --k) {
        const int32_t *e = a.sched + (size_t)k * GTR_STRIDE;
        const int slot = e[1], nc = e[2];
        R Uu[4];
        int Hu;
        if (k == a.n_sched - 1) {  // root: the frequencies
#pragma unroll
            for (int i = 0; i < 4; ++i) Uu[i] = (R)a.pi[i];
            Hu = 0;
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) Uu[i] = a.U[((size_t)slot * 4 + i) * a.P_pad + p];
            Hu = a.H[(size_t)slot * a.P_pad + p];
        }
        R d[MAXC][4], X[MAXC][4];
        int ec[MAXC];
        // (loops over children unrolled with guards: register arrays stay in registers)
#pragma unroll
        for (int c = 0; c < MAXC; ++c) {
            ec[c] = 0;
#pragma unroll
            for (int i = 0; i < 4; ++i) X[c][i] = d[c][i] = R(1);
            if (c < nc) gtr_child<R>(a, e[3 + c], p, d[c], ec[c], X[c]);
        }
#pragma unroll
        for (int c = 0; c < MAXC; ++c) {
            if (c >= nc) continue;
            R V[4];
            int G = Hu;
#pragma unroll
            for (int i = 0; i < 4; ++i) V[i] = Uu[i];
#pragma unroll
            for (int s = 0; s < MAXC; ++s) {
                if (s == c || s >= nc) continue;
#pragma unroll
                for (int i = 0; i < 4; ++i) V[i] *= X[s][i];
                G += ec[s];
            }
            gtr_normalise<R>(V, G);
            const int cn = e[3 + c], b = a.nbr[cn];
            if (b >= 0) {
                // A_b += c_p 2^(G + E_c - E_root) V d_c^T, summed over the wave's 64 patterns
                const double f = cw != 0.0 ? ldexp(cw, G + ec[c] - Er) : 0.0;
                double acc16[16];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc16[4 * i + j] = f * (double)V[i] * (double)d[c][j];
#pragma unroll
                for (int q = 0; q < 16; ++q)
                    for (int off = 32; off > 0; off >>= 1) acc16[q] += __shfl_xor(acc16[q], off);
                if (lane < 16) {
                    double v = 0.0;
#pragma unroll
                    for (int q = 0; q < 16; ++q)
                        if (q == lane) v = acc16[q];
                    a.Apart[((size_t)b * a.n_waves + wave) * 16 + lane] = v;
                }
            }
            const int ref = a.nref[cn];
            if ((ref >> 28) == KIND_INT) {
                // message into the child: U_c = P_c^T V
                const R *P = a.Pm + (size_t)cn * 16;
                R Uc[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) Uc[j] = P[j] * V[0] + P[4 + j] * V[1] + P[8 + j] * V[2] + P[12 + j] * V[3];
                gtr_normalise<R>(Uc, G);
                const int cs = ref & REF_MASK;
#pragma unroll
                for (int j = 0; j < 4; ++j) a.U[((size_t)cs * 4 + j) * a.P_pad + p] = Uc[j];
                a.H[(size_t)cs * a.P_pad + p] = G;
            }
        }
    }
}

// Fixed-order reductions: A_b over the waves (16 interleaved slices, then the
// slices in order), and the compensated lnL sum (last workgroup).
__global__ __launch_bounds__(256) void k_gtr_reduce(GtrReduceArgs a) {
    const int tid = threadIdx.x;
    __shared__ double part[256], comp[256];
    if ((int)blockIdx.x < a.n_branch) {
        const int q = tid & 15, sl = tid >> 4;
        const double *src = a.Apart + (size_t)blockIdx.x * a.n_waves * 16;
        double v = 0.0;
        for (int w = sl; w < a.n_waves; w += 16) v += src[(size_t)w * 16 + q];
        part[tid] = v;
        __syncthreads();
        if (tid < 16) {
            double t = 0.0;
            for (int k = 0; k < 16; ++k) t += part[k * 16 + tid];
            a.A[(size_t)blockIdx.x * 16 + tid] = t;
        }
        return;
    }
    // lnL: per-thread Neumaier sums over contiguous pattern ranges, then in thread order
    const int per = (a.P + 255) / 256;
    double s = 0.0, c = 0.0;
    for (int p = tid * per; p < min((tid + 1) * per, a.P); ++p) neum(s, c, a.weights[p] * a.site_ll[p]);
    part[tid] = s;
    comp[tid] = c;
    __syncthreads();
    if (tid == 0) {
        double S0 = 0.0, C0 = 0.0;
        for (int k = 0; k < 256; ++k) comp_add(S0, C0, part[k], comp[k]);
        a.res[0] = S0 + C0;
    }
}

template <typename R> void launch_gtr(const GtrArgs<R> &a, bool grad, hipStream_t s) {
    const dim3 grid(a.P_pad / GTR_BLOCK);
    hipLaunchKernelGGL(k_gtr_inside<R>, grid, dim3(GTR_BLOCK), 0, s, a);
    if (grad) hipLaunchKernelGGL(k_gtr_outside<R>, grid, dim3(GTR_BLOCK), 0, s, a);
}

// workgroups 0 .. n_branch - 1 reduce A_b (with gradients), the last one lnL
void launch_gtr_reduce(const GtrReduceArgs &a, bool grad, hipStream_t s) {
    GtrReduceArgs r = a;
    if (!grad) r.n_branch = 0;
    hipLaunchKernelGGL(k_gtr_reduce, dim3(r.n_branch + 1), dim3(256), 0, s, r);
}

template void launch_gtr<float>(const GtrArgs<float> &, bool, hipStream_t);
template void launch_gtr<double>(const GtrArgs<double> &, bool, hipStream_t);

// ---------------------------------------------------------------------------
// k_fel: per-site refit candidates (FELAnalyzer::compute_pattern_log_likelihood,
// fel.hpp:276-308, at eval_pattern_lnl's alpha and omega, fel.hpp:367-385).
// One workgroup per candidate walks its pattern through the tree.  A child
// edge contributes x = max(P, 0) d with P = U diag(e^{l a t}) V: for a tip
// with one or a few possible codons only those columns of P are formed
// (four threads per row, each column's entries clamped before they are
// summed); otherwise (internal child, gap, wide ambiguity) the whole P is
// built in LDS on FP64 MFMA and applied.  Partials are normalised to max in
// [0.5, 1) by an exact power of two.  FP64 throughout.
// ---------------------------------------------------------------------------
constexpr int FEL_MAX_COLUMNS = 12;  // tip columns formed one by one up to this many

__global__ __launch_bounds__(256) void k_fel(FelArgs a) {
    const int k = blockIdx.x, tid = threadIdx.x, w = tid >> 6, lr = tid & 15, lg = (tid & 63) >> 4;
    __shared__ double Us[64 * LDA], Vs[64 * LDA], Pt[64 * LDA];
    __shared__ double ex[64], dv[64], acc[64], red[4];
    __shared__ int cols[64];
    __shared__ int ncol;
    const int p = a.pat[k], sl = a.spec[k];
    const double at = fmax(a.alpha[k], 1e-12);
    const double *U = a.U + (size_t)sl * 4096, *V = a.V + (size_t)sl * 4096, *lam = a.lam + (size_t)sl * 64;
    for (int idx = tid; idx < 4096; idx += 256) {
        Us[(idx >> 6) * LDA + (idx & 63)] = U[idx];
        Vs[(idx >> 6) * LDA + (idx & 63)] = V[idx];
    }
    const uint64_t all = (a.S >= 64) ? ~0ull : ((1ull << a.S) - 1ull);
    double *Dk = a.D + (size_t)k * a.n_int * 64;
    int32_t *Ek = a.E + (size_t)k * a.n_int;
    const int row = tid >> 2, part = tid & 3;
    int eacc = 0;
    for (int q = 0; q < a.n_sched; ++q) {
        const int32_t *e = a.sched + (size_t)q * GTR_STRIDE;
        const int slot = e[1], nc = e[2];
        if (tid < 64) acc[tid] = 1.0;
        eacc = 0;
        for (int c = 0; c < nc; ++c) {
            const int cn = e[3 + c], ref = a.nref[cn], kind = ref >> 28, idx = ref & REF_MASK;
            __syncthreads();  // previous child's reads of dv / Pt / cols are done
            uint64_t mask = 0;
            if (kind == KIND_TIP) {
                const unsigned code = a.tips[(size_t)idx * a.tip_stride + p];
                mask = code == TIP_ONES ? all : (code & TIP_AMB) ? (a.amb[code & 0x7FFFu] & all) : (1ull << code);
            }
            const int pc = kind == KIND_TIP ? __popcll(mask) : 64;
            if (tid < 64) {
                dv[tid] = kind == KIND_TIP ? (double)((mask >> tid) & 1ull) : a.D[((size_t)k * a.n_int + idx) * 64 + tid];
                ex[tid] = exp(lam[tid] * at * a.bl[cn]);
            }
            if (tid == 0) {
                if (kind != KIND_TIP) eacc += Ek[idx];
                int n = 0;
                for (uint64_t mm = mask; mm && n < FEL_MAX_COLUMNS; mm &= mm - 1) cols[n++] = __ffsll((long long)mm) - 1;
                ncol = n;
            }
            __syncthreads();
            if (a.ident[cn]) {
                if (tid < 64) acc[tid] *= dv[tid];
                continue;
            }
            if (kind == KIND_TIP && pc <= FEL_MAX_COLUMNS) {
                // x_i = sum over the tip's codons s of max(P_is, 0)
                double x = 0.0;
                for (int j = 0; j < ncol; ++j) {
                    const int sc = cols[j];
                    double v = 0.0;
#pragma unroll
                    for (int kk = 0; kk < 16; ++kk) {
                        const int kx = 16 * part + kk;
                        v += Us[row * LDA + kx] * ex[kx] * Vs[kx * LDA + sc];
                    }
                    v += __shfl_xor(v, 1);
                    v += __shfl_xor(v, 2);
                    x += v > 0.0 ? v : 0.0;
                }
                if (part == 0) acc[row] *= x;
                continue;
            }
            // P = max(U diag(e) V, 0) in Pt (rows of wave w: 16w .. 16w + 15)
            f64x4 pacc[4];
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) pacc[qq] = splat<f64x4>(0.0);
#pragma unroll 4
            for (int kk = 0; kk < 16; ++kk) {
                const int i = 16 * w + lr, kx = 4 * kk + lg;
                const double av = Us[i * LDA + kx] * ex[kx];
#pragma unroll
                for (int qq = 0; qq < 4; ++qq) pacc[qq] = MF<double>::mma(av, Vs[kx * LDA + 16 * qq + lr], pacc[qq]);
            }
#pragma unroll
            for (int qq = 0; qq < 4; ++qq)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const double v = pacc[qq][r];
                    Pt[(16 * w + MF<double>::crow(lg, r)) * LDA + 16 * qq + lr] = v > 0.0 ? v : 0.0;
                }
            __syncthreads();
            // x = P d (four threads per row), acc *= x
            {
                double sacc = 0.0;
#pragma unroll
                for (int j = 0; j < 16; ++j) sacc += Pt[row * LDA + 16 * part + j] * dv[16 * part + j];
                sacc += __shfl_xor(sacc, 1);
                sacc += __shfl_xor(sacc, 2);
                if (part == 0) acc[row] *= sacc;
            }
        }
        __syncthreads();
        // normalise the node's partial by a power of two (wave 0)
        if (tid < 64) {
            double m = acc[tid];
            for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
            double v = acc[tid];
            int ke = 0;
            if (m > 0.0) {
                ke = frexp_e(m);
                v = ldexp(v, -ke);
            }
            acc[tid] = v;
            Dk[(size_t)slot * 64 + tid] = v;
            if (tid == 0) Ek[slot] = eacc + ke;
            if (tid == 0) red[0] = (double)(eacc + ke);
        }
        __syncthreads();
    }
    // root (last entry): L = pi . D_root
    if (tid < 64) {
        double l = a.pi[tid] * acc[tid];
        for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off);
        if (tid == 0) {
            a.out[k] = l > 0.0 ? log(l) + red[0] * 0.69314718055994530942 : -1e20;
            a.out[a.n_cand + k] = a.mu[sl];
            a.out[2 * a.n_cand + k] = (double)a.status[sl];
        }
    }
}

void launch_fel(const FelArgs &a, int n_cand, hipStream_t s) {
    if (n_cand > 0) hipLaunchKernelGGL(k_fel, dim3(n_cand), dim3(256), 0, s, a);
}

// ---------------------------------------------------------------------------
// k_meme: MEME's per-site candidates (eval_pattern_meme_lnl, meme.hpp:219-277;
// compute_pattern_log_likelihood, fel.hpp:276-308).  One workgroup per
// candidate walks its pattern through the tree like k_fel, with up to three
// spectra: a test branch contributes x = (p1 max(P_A, 0) + (1 - p1)
// max(P_B, 0)) d, a background branch max(P_C, 0) d, at time a t; a
// candidate may condition one branch on class A or B alone.  Each
// spectrum is held in LDS by its eigenvectors Z of the symmetrised generator
// (P_ij = pi_i^-1/2 pi_j^1/2 sum_k Z_ik e^{l_k a t} Z_jk, k_model's U and V
// factored), so three fit beside the P tile.  FP64 throughout.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_meme(MemeArgs a) {
    const int k = blockIdx.x, tid = threadIdx.x, w = tid >> 6, lr = tid & 15, lg = (tid & 63) >> 4;
    __shared__ double Zs[3][64 * LDA];
    __shared__ double Pt[64 * LDA];
    __shared__ double ex[2][64], lamc[3][64], dv[64], acc[64], sq[64], isq[64], red[4];
    __shared__ int cols[64];
    __shared__ int ncol;
    const int p = a.pat[k];
    const int sl[3] = {a.specA[k], a.specB[k], a.specC[k]};
    const double at = a.alpha[k], w1 = a.p1[k], w2 = 1.0 - w1;
    const int cnd = a.cond[k], cnode = cnd >= 0 ? cnd >> 1 : -1, ccls = cnd & 1;
    for (int c = 0; c < 3; ++c) {
        const double *Z = a.Z + (size_t)sl[c] * 4096;
#pragma unroll 4
        for (int q = 0; q < 16; ++q) {
            const int idx = tid + 256 * q;
            Zs[c][(idx >> 6) * LDA + (idx & 63)] = Z[idx];
        }
        if (tid < 64) lamc[c][tid] = a.lam[(size_t)sl[c] * 64 + tid];
    }
    if (tid < 64) {
        // the square roots k_model factored into U and V (rate_matrix.hpp:211-246)
        const double pv = a.pi[tid];
        const double v = sqrt(pv > 1e-12 ? pv : 1e-12);
        sq[tid] = tid < a.S ? v : 0.0;
        isq[tid] = tid < a.S ? 1.0 / v : 0.0;
    }
    const uint64_t all = (a.S >= 64) ? ~0ull : ((1ull << a.S) - 1ull);
    double *Dk = a.D + (size_t)k * a.n_int * 64;
    int32_t *Ek = a.E + (size_t)k * a.n_int;
    const int row = tid >> 2, part = tid & 3;
    int eacc = 0;
    for (int q = 0; q < a.n_sched; ++q) {
        const int32_t *e = a.sched + (size_t)q * GTR_STRIDE;
        const int slot = e[1], nc = e[2];
        if (tid < 64) acc[tid] = 1.0;
        eacc = 0;
        for (int c = 0; c < nc; ++c) {
            const int cn = e[3 + c], ref = a.nref[cn], kind = ref >> 28, idx = ref & REF_MASK;
            const bool test = a.is_test[cn] != 0;
            // components of this edge: (spectrum, weight); the conditioned
            // branch takes one class alone (the branch EBFs, meme.hpp:552-580)
            const bool cbr = cn == cnode;
            const int n_comp = cbr ? 1 : test ? 2 : 1;
            const int cs0 = cbr ? ccls : test ? 0 : 2;
            const double cw0 = cbr ? 1.0 : test ? w1 : 1.0;
            __syncthreads();  // previous child's reads of dv / Pt / cols / ex are done
            uint64_t mask = 0;
            if (kind == KIND_TIP) {
                const unsigned code = a.tips[(size_t)idx * a.tip_stride + p];
                mask = code == TIP_ONES ? all : (code & TIP_AMB) ? (a.amb[code & 0x7FFFu] & all) : (1ull << code);
            }
            const int pc = kind == KIND_TIP ? __popcll(mask) : 64;
            if (tid < 64) {
                dv[tid] = kind == KIND_TIP ? (double)((mask >> tid) & 1ull) : a.D[((size_t)k * a.n_int + idx) * 64 + tid];
                ex[0][tid] = exp(lamc[cs0][tid] * at * a.bl[cn]);
                ex[1][tid] = exp(lamc[1][tid] * at * a.bl[cn]);
            }
            if (tid == 0) {
                if (kind != KIND_TIP) eacc += Ek[idx];
                int n = 0;
                for (uint64_t mm = mask; mm && n < FEL_MAX_COLUMNS; mm &= mm - 1) cols[n++] = __ffsll((long long)mm) - 1;
                ncol = n;
            }
            __syncthreads();
            if (a.ident[cn]) {
                if (tid < 64) acc[tid] *= dv[tid];
                continue;
            }
            if (kind == KIND_TIP && pc <= FEL_MAX_COLUMNS) {
                // x_i = sum over the tip's codons s of sum_comp w max(P_is, 0)
                double x = 0.0;
                for (int j = 0; j < ncol; ++j) {
                    const int sc = cols[j];
                    for (int m = 0; m < n_comp; ++m) {
                        const int cs = m == 0 ? cs0 : 1;
                        const double *Zc = Zs[cs];
                        double v = 0.0;
#pragma unroll
                        for (int kk = 0; kk < 16; ++kk) {
                            const int kx = 16 * part + kk;
                            v += Zc[row * LDA + kx] * ex[m][kx] * Zc[sc * LDA + kx];
                        }
                        v += __shfl_xor(v, 1);
                        v += __shfl_xor(v, 2);
                        v *= isq[row] * sq[sc];
                        x += (m == 0 ? cw0 : w2) * (v > 0.0 ? v : 0.0);
                    }
                }
                if (part == 0) acc[row] *= x;
                continue;
            }
            // P = sum_comp w max(P_comp, 0) in Pt (rows of wave w: 16w .. 16w + 15)
            f64x4 pmix[4];
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) pmix[qq] = splat<f64x4>(0.0);
            for (int m = 0; m < n_comp; ++m) {
                const int cs = m == 0 ? cs0 : 1;
                const double cw = m == 0 ? cw0 : w2;
                const double *Zc = Zs[cs];
                f64x4 pacc[4];
#pragma unroll
                for (int qq = 0; qq < 4; ++qq) pacc[qq] = splat<f64x4>(0.0);
#pragma unroll 4
                for (int kk = 0; kk < 16; ++kk) {
                    const int i = 16 * w + lr, kx = 4 * kk + lg;
                    const double av = Zc[i * LDA + kx] * ex[m][kx];
#pragma unroll
                    for (int qq = 0; qq < 4; ++qq) pacc[qq] = MF<double>::mma(av, Zc[(16 * qq + lr) * LDA + kx], pacc[qq]);
                }
#pragma unroll
                for (int qq = 0; qq < 4; ++qq)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int i = 16 * w + MF<double>::crow(lg, r), j = 16 * qq + lr;
                        const double v = pacc[qq][r] * isq[i] * sq[j];
                        pmix[qq][r] += cw * (v > 0.0 ? v : 0.0);
                    }
            }
#pragma unroll
            for (int qq = 0; qq < 4; ++qq)
#pragma unroll
                for (int r = 0; r < 4; ++r) Pt[(16 * w + MF<double>::crow(lg, r)) * LDA + 16 * qq + lr] = pmix[qq][r];
            __syncthreads();
            {
                double sacc = 0.0;
#pragma unroll
                for (int j = 0; j < 16; ++j) sacc += Pt[row * LDA + 16 * part + j] * dv[16 * part + j];
                sacc += __shfl_xor(sacc, 1);
                sacc += __shfl_xor(sacc, 2);
                if (part == 0) acc[row] *= sacc;
            }
        }
        __syncthreads();
        // normalise the node's partial by a power of two (wave 0)
        if (tid < 64) {
            double m = acc[tid];
            for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
            double v = acc[tid];
            int ke = 0;
            if (m > 0.0) {
                ke = frexp_e(m);
                v = ldexp(v, -ke);
            }
            acc[tid] = v;
            Dk[(size_t)slot * 64 + tid] = v;
            if (tid == 0) Ek[slot] = eacc + ke;
            if (tid == 0) red[0] = (double)(eacc + ke);
        }
        __syncthreads();
    }
    if (tid < 64) {
        double l = a.pi[tid] * acc[tid];
        for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off);
        if (tid == 0) a.out[k] = l > 0.0 ? log(l) + red[0] * 0.69314718055994530942 : -1e20;
    }
}

void launch_meme(const MemeArgs &a, int n_cand, hipStream_t s) {
    if (n_cand > 0) hipLaunchKernelGGL(k_meme, dim3(n_cand), dim3(256), 0, s, a);
}

// ---------------------------------------------------------------------------
// Fixed scratch of the allocation-free entry point (hy_loglik_from_P_noalloc):
// one array of NOALLOC_SCRATCH_BYTES per device that loads this library, held
// until the process ends.  Host code never passes the array's name to a
// kernel (on the host it is only a stand-in for the device symbol): the
// device address comes from hipGetSymbolAddress, once per device, with that
// device current.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

hipError_t noalloc_scratch(int device, void **out) {
    static std::mutex mu;
    static void *addr[NOALLOC_MAX_DEVICES] = {};
    if (!out || device < 0 || device >= NOALLOC_MAX_DEVICES) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) {
            const hipError_t e2 = hipSetDevice(cur);
            if (e == hipSuccess) e = e2;
        }
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        addr[device] = p;
    }
    *out = addr[device];
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// Local projection of one branch (aBSREL, absrel.hpp:438-455 over the branch
// context of :723-780): L_c,p = V_b,p^T P_c D_b,p for candidate matrices P_c,
// from the outside partials V_b and the inside partials D_b the last
// evaluation left resident (a tip child's D is the indicator of the pattern's
// code).  One workgroup per (pattern tile, candidate); wave w forms the
// products of states 16 w .. 16 w + 15 (P_c D_b on MFMA in the working type,
// the column dot in FP64), the four waves' shares are summed in wave order.
// ---------------------------------------------------------------------------
template <typename R, int T>
__global__ __launch_bounds__(256) void k_branch_ctx(BranchCtxArgs<R> a) {
    constexpr int NB = T / 16;
    typedef typename MF<R>::acc acc_t;
    __shared__ double part[4][T];
    const int tile = blockIdx.x, c = blockIdx.y;
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, lr = l & 15, lg = l >> 4;
    const double *P = a.P + (size_t)c * 4096;
    const int p0 = tile * T;
    acc_t X[NB];
    if (a.tip_row < 0) {
        R av[16], bv[16 * NB];
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) av[kk] = (R)P[(16 * w + lr) * 64 + 4 * kk + lg];
        fetch_b<R, T>(bv, a.D + ((size_t)a.d_slot * a.n_tiles + tile) * (64 * T), lr, lg);
        mfma_regs<R, T>(av, bv, X);
    } else {
        // X[row][p]: P summed over the codons compatible with the code (in
        // codon order), rounded to the working type like the pruning passes'
        // tip columns
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const unsigned code = a.tips[(size_t)a.tip_row * a.P_pad + p0 + 16 * nb + lr];
            const uint64_t all = (a.S >= 64) ? ~0ull : ((1ull << a.S) - 1ull);
            const uint64_t mask = code < 64u ? 0ull : (code == TIP_ONES ? all : a.amb[code & 0x7FFFu]);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double *row = P + (16 * w + MF<R>::crow(lg, r)) * 64;
                double x = 0.0;
                if (code < 64u) {
                    x = row[code];
                } else {
                    for (int j = 0; j < 64; ++j)
                        if ((mask >> j) & 1ull) x += row[j];
                }
                X[nb][r] = (R)x;
            }
        }
    }
    acc_t Vt[NB];
    load_c<R, T>(Vt, a.V + ((size_t)a.v_slot * a.n_tiles + tile) * (64 * T), w, lr, lg);
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r) s += (double)Vt[nb][r] * (double)X[nb][r];
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        if (lg == 0) part[w][16 * nb + lr] = s;
    }
    __syncthreads();
    if (tid < T) {
        const int p = p0 + tid;
        a.out_L[(size_t)c * a.P_pad + p] = ((part[0][tid] + part[1][tid]) + part[2][tid]) + part[3][tid];
        if (c == 0)
            a.out_e[p] = a.G[(size_t)a.v_slot * a.P_pad + p] + (a.tip_row < 0 ? a.E[(size_t)a.d_slot * a.P_pad + p] : 0);
    }
}

// sum_p w_p ln L_p per candidate (L_p = L 2^e; w_p (-1e20) where L_p <= 0,
// absrel.hpp:449-452), one workgroup per candidate: compensated per-thread
// sums over contiguous pattern ranges, combined by error-free additions in a
// fixed order.
__global__ __launch_bounds__(1024) void k_branch_sum(BranchSumArgs a) {
    const int c = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ double ws[16], wc[16];
    const double *L = a.L + (size_t)c * a.P_pad;
    const int P = a.n_patterns, per = (P + 1023) / 1024;
    const int pe = min((tid + 1) * per, P);
    double s = 0.0, cmp = 0.0;
    for (int p = tid * per; p < pe; ++p) {
        const double v = L[p];
        const double term = v > 0.0 ? a.weights[p] * (log(v) + (double)a.e[p] * 0.69314718055994530942)
                                    : a.weights[p] * -1e20;
        neum(s, cmp, term);
    }
    for (int off = 1; off < 64; off <<= 1) comp_add(s, cmp, __shfl_xor(s, off), __shfl_xor(cmp, off));
    if (lane == 0) {
        ws[wave] = s;
        wc[wave] = cmp;
    }
    __syncthreads();
    if (tid == 0) {
        double S0 = 0.0, C0 = 0.0;
        for (int k = 0; k < 16; ++k) comp_add(S0, C0, ws[k], wc[k]);
        a.out[c] = S0 + C0;
    }
}

template <typename R> void launch_branch_ctx(const BranchCtxArgs<R> &a, int T, int n_cand, hipStream_t s) {
    const dim3 g(a.n_tiles, n_cand);
    if (T == 16) {
        hipLaunchKernelGGL((k_branch_ctx<R, 16>), g, dim3(256), 0, s, a);
    } else if (T == 32) {
        hipLaunchKernelGGL((k_branch_ctx<R, 32>), g, dim3(256), 0, s, a);
    } else if constexpr (sizeof(R) == 4) {  // FP64 tiles stop at 32 patterns
        hipLaunchKernelGGL((k_branch_ctx<R, 64>), g, dim3(256), 0, s, a);
    }
}

void launch_branch_sum(const BranchSumArgs &a, int n_cand, hipStream_t s) {
    hipLaunchKernelGGL(k_branch_sum, dim3(n_cand), dim3(1024), 0, s, a);
}

template void launch_branch_ctx<float>(const BranchCtxArgs<float> &, int, int, hipStream_t);
template void launch_branch_ctx<double>(const BranchCtxArgs<double> &, int, int, hipStream_t);
template void launch_pmat<float>(const PmatArgs<float> &, const ProjArgs *, int, hipStream_t);
template void launch_pmat<double>(const PmatArgs<double> &, const ProjArgs *, int, hipStream_t);
template void launch_pmat_host<float>(const PmatArgs<float> &, const double *, hipStream_t);
template void launch_pmat_host<double>(const PmatArgs<double> &, const double *, hipStream_t);
template void launch_inside<float>(const PruneArgs<float> &, int, bool, int, hipStream_t);
template void launch_inside<double>(const PruneArgs<double> &, int, bool, int, hipStream_t);
template void launch_outside<float>(const PruneArgs<float> &, int, bool, bool, int, hipStream_t);
template void launch_outside<double>(const PruneArgs<double> &, int, bool, bool, int, hipStream_t);
template void launch_adjoint<float>(const AdjArgs<float> &, int, hipStream_t);
template void launch_adjoint<double>(const AdjArgs<double> &, int, hipStream_t);

}  // namespace hy

// ---------------------------------------------------------------------------
// k_lbfgs: the optimizer step that closes a captured fit iteration (one wave;
// the shared step logic with wave-wide reductions).
// ---------------------------------------------------------------------------
struct LbDevPar {
    int tid;
    double *w;  // LDS work area, 2 LB_MMAX^2 + 4 LB_MMAX doubles
    __device__ double *work() { return w; }
    template <class F>
    __device__ void single(F f) {
        if (tid == 0) f();
    }
    __device__ void sync() { __syncthreads(); }
    template <class F>
    __device__ void loop(int n, F f) {
        for (int i = tid; i < n; i += LB_LANES) f(i);
    }
    template <class F>
    __device__ double sum(int n, F f) {
        LB_NOCONTRACT
        double s = 0.0;
        for (int i = tid; i < n; i += LB_LANES) s = s + f(i);
        // lane t adds lane t + st: the host policy's pairwise tree
        for (int st = LB_LANES / 2; st > 0; st >>= 1) s = s + __shfl_down(s, st);
        return __shfl(s, 0);
    }
    template <int NQ, class F>
    __device__ void sums(int n, double *out, F f) {
        LB_NOCONTRACT
        double acc[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
        for (int i = tid; i < n; i += LB_LANES) f(i, acc);
        for (int st = LB_LANES / 2; st > 0; st >>= 1)
#pragma unroll
            for (int q = 0; q < NQ; ++q) acc[q] = acc[q] + __shfl_down(acc[q], st);
#pragma unroll
        for (int q = 0; q < NQ; ++q) out[q] = __shfl(acc[q], 0);
    }
    template <class F>
    __device__ double max(int n, F f) {
        double s = 0.0;
        for (int i = tid; i < n; i += LB_LANES) {
            const double x = f(i);
            s = x > s ? x : s;
        }
        for (int st = LB_LANES / 2; st > 0; st >>= 1) {
            const double o = __shfl_down(s, st);
            s = o > s ? o : s;
        }
        return __shfl(s, 0);
    }
};

__global__ __launch_bounds__(LB_LANES) void k_lbfgs(LbArgs a) {
    __shared__ double lb_work[2 * LB_MMAX * LB_MMAX + 4 * LB_MMAX];
    LbDevPar p{(int)threadIdx.x, lb_work};
    hy_lbfgs_head hd = *a.hd;  // every lane steps its own copy of the (uniform) scalars
    const LbVec v = lb_vec(a.state, hd.n, hd.m);
    int ok = 1;
    for (int m = 0; m < a.n_mat; ++m)
        if (a.eig_status[m] < 0) ok = 0;
    const double ft = lb_gather(p, v, hd.n, a.scale, a.res, a.params, a.res_map, a.par_map);
    lb_step(p, v, hd, ft, ok);
    lb_scatter(p, v, hd.n, a.params, a.par_map);
    __syncthreads();
    if (threadIdx.x == 0) *a.hd = hd;
}

void launch_lbfgs(const LbArgs &a, hipStream_t s) { hipLaunchKernelGGL(k_lbfgs, dim3(1), dim3(LB_LANES), 0, s, a); }
