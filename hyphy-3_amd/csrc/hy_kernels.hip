// HIP kernels of the MI355X MG94xREV codon likelihood engine (gfx950).
//
// Pipeline per evaluation (one stream):
//   k_model    one workgroup per rate matrix: MG94 Q, mu, symmetric spectrum by
//              parallel cyclic Jacobi in LDS (FP64)
//   k_pmat     one workgroup per branch: P_b = sum_j w_j exp(Q_j tau_b) in FP64,
//              written as MFMA A-operand tiles of P and P^T in the working type
//   k_inside   one workgroup per pattern tile: post-order pruning, 64x64 by
//              64xT child products on MFMA, per-pattern power-of-two rescale
//   k_outside  one workgroup per pattern tile: pre-order outside messages
//   k_adjoint  (branch, pattern chunk) workgroups: A_b = sum_p c_p V_b D_b^T
//   k_proj     (matrix, parameter) workgroups: M = V_pi dQ U_pi
//   k_frechet  (branch, component) workgroups: U^T A V^T, spectral Frechet
//              projections (FP64)
//   k_final    one workgroup: deterministic compensated reductions
//
// MFMA lane layout (v_mfma_{f32,f64}_16x16x4, measured on gfx950): lane l
// supplies A[l%16][l/16] and B[l/16][l%16]; accumulator element r of lane l
// holds C[row][l%16] with row = 4*(l/16)+r for f32 and row = 4*r+(l/16) for
// f64 (MF<R>::crow).
#include "hy_internal.h"

#include <cmath>
#include <cstdio>

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
// k_model: Q, mu and the symmetric spectrum of D^1/2 Q D^-1/2.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_model(ModelArgs a) {
    const int m = blockIdx.x, tid = threadIdx.x;
    __shared__ double A[64 * 64];
    __shared__ double Z[64 * 64];
    __shared__ double sq[64], isq[64];
    __shared__ double rc[32], rs[32], rt[32];
    __shared__ int rp[32], rq[32];
    __shared__ double red[256];
    __shared__ int nrot_sweep;

    const double *prm = a.params;
    const double alpha = prm[a.pl.off_alpha], beta = prm[a.pl.off_omega + m];
    const double *th = prm + a.pl.off_theta;
    const double delta = prm[a.pl.off_delta], psi = prm[a.pl.off_psi];
    const int S = a.S;

    for (int idx = tid; idx < 4096; idx += 256) {
        const int i = idx >> 6, j = idx & 63;
        double q = 0.0;
        if (i < S && j < S && i != j) {
            const int k = i * 64 + j;
            const double rm = a.syn[k] ? alpha : beta;
            const int nd = a.nd[k];
            const int8_t *dp = a.dpos + 3 * k, *df = a.dfrom + 3 * k, *dt = a.dto + 3 * k;
            if (nd == 1) {
                q = rm * nuc_rate(th, df[0], dt[0]) * a.pos[dp[0] * 4 + dt[0]];
            } else if (nd == 2 && delta > 0.0) {
                q = rm * delta * (nuc_rate(th, df[0], dt[0]) * nuc_rate(th, df[1], dt[1])) *
                    (a.pos[dp[0] * 4 + dt[0]] * a.pos[dp[1] * 4 + dt[1]]);
            } else if (nd == 3 && psi > 0.0) {
                q = rm * psi * (nuc_rate(th, df[0], dt[0]) * nuc_rate(th, df[1], dt[1]) * nuc_rate(th, df[2], dt[2])) *
                    (a.pos[dp[0] * 4 + dt[0]] * a.pos[dp[1] * 4 + dt[1]] * a.pos[dp[2] * 4 + dt[2]]);
            }
        }
        A[idx] = q;
    }
    __syncthreads();
    if (tid < 64) {
        double s = 0.0;
        for (int j = 0; j < 64; ++j)
            if (j != tid) s += A[tid * 64 + j];
        A[tid * 64 + tid] = -s;
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
    if (tid == 0) {
        double mu = 0.0;
        for (int i = 0; i < S; ++i) mu -= a.pi[i] * A[i * 64 + i];
        a.mu[m] = mu;
    }
    for (int idx = tid; idx < 4096; idx += 256) {
        const int i = idx >> 6, j = idx & 63;
        Z[idx] = sq[i] * A[idx] * isq[j];
    }
    __syncthreads();
    double fro = 0.0;
    for (int idx = tid; idx < 4096; idx += 256) {
        const int i = idx >> 6, j = idx & 63;
        const double v = 0.5 * (Z[i * 64 + j] + Z[j * 64 + i]);
        A[idx] = v;
        fro += v * v;
    }
    red[tid] = fro;
    __syncthreads();
    for (int idx = tid; idx < 4096; idx += 256) Z[idx] = ((idx >> 6) == (idx & 63)) ? 1.0 : 0.0;
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    const double thresh = 1e-17 * sqrt(red[0]);

    int sweep = 0, converged = 0;
    for (; sweep < 40 && !converged; ++sweep) {
        if (tid == 0) nrot_sweep = 0;
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
                const double apq = A[p * 64 + q];
                double c = 1.0, s = 0.0, t = 0.0;
                if (fabs(apq) > thresh) {
                    const double app = A[p * 64 + p], aqq = A[q * 64 + q];
                    const double theta = (aqq - app) / (2.0 * apq);
                    if (fabs(theta) > 1e150) {
                        t = 0.5 / theta;
                    } else {
                        t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                    }
                    c = 1.0 / sqrt(t * t + 1.0);
                    s = t * c;
                    active = 1;
                }
                rp[tid] = p;
                rq[tid] = q;
                rc[tid] = c;
                rs[tid] = s;
                rt[tid] = t;
                if (active) atomicAdd(&nrot_sweep, 1);
            }
            if (!__syncthreads_or(active)) continue;
            for (int b = tid; b < 1024; b += 256) {
                const int I = b >> 5, J = b & 31;
                if (I > J) continue;
                const double sI = rs[I], sJ = rs[J];
                if (sI == 0.0 && sJ == 0.0) continue;
                const double cI = rc[I], cJ = rc[J];
                const int p = rp[I], q = rq[I], r = rp[J], s = rq[J];
                if (I == J) {
                    const double apq = A[p * 64 + q], t = rt[I];
                    A[p * 64 + p] -= t * apq;
                    A[q * 64 + q] += t * apq;
                    A[p * 64 + q] = 0.0;
                    A[q * 64 + p] = 0.0;
                } else {
                    const double xpr = A[p * 64 + r], xps = A[p * 64 + s], xqr = A[q * 64 + r], xqs = A[q * 64 + s];
                    const double ypr = cI * xpr - sI * xqr, yps = cI * xps - sI * xqs;
                    const double yqr = sI * xpr + cI * xqr, yqs = sI * xps + cI * xqs;
                    const double zpr = cJ * ypr - sJ * yps, zps = sJ * ypr + cJ * yps;
                    const double zqr = cJ * yqr - sJ * yqs, zqs = sJ * yqr + cJ * yqs;
                    A[p * 64 + r] = zpr; A[r * 64 + p] = zpr;
                    A[p * 64 + s] = zps; A[s * 64 + p] = zps;
                    A[q * 64 + r] = zqr; A[r * 64 + q] = zqr;
                    A[q * 64 + s] = zqs; A[s * 64 + q] = zqs;
                }
            }
            for (int idx = tid; idx < 2048; idx += 256) {
                const int k = idx & 31, i = idx >> 5;
                const double s = rs[k];
                if (s == 0.0) continue;
                const double c = rc[k];
                const int p = rp[k], q = rq[k];
                const double zp = Z[i * 64 + p], zq = Z[i * 64 + q];
                Z[i * 64 + p] = c * zp - s * zq;
                Z[i * 64 + q] = s * zp + c * zq;
            }
            __syncthreads();
        }
        __syncthreads();
        converged = (nrot_sweep == 0);
        __syncthreads();
    }
    double *lam = a.lam + m * 64, *U = a.U + m * 4096, *V = a.V + m * 4096;
    if (tid < 64) {
        const double l = A[tid * 64 + tid];
        lam[tid] = l > 0.0 ? 0.0 : l;
    }
    for (int idx = tid; idx < 4096; idx += 256) {
        const int i = idx >> 6, k = idx & 63;
        U[idx] = isq[i] * Z[idx];               // U[i][k]
        V[k * 64 + i] = Z[idx] * sq[i];         // V[k][i]
    }
    if (tid == 0) a.status[m] = converged ? sweep : -1;
}

// Write one branch's P (FP64 in LDS) as row-major, row sums and the MFMA
// A-operand swizzles of P and P^T in the working type.
template <typename R>
__device__ __forceinline__ void emit_p(const double *P, int node, int S, R *Prow_all, R *Rsum_all, R *PA_all,
                                       R *PTA_all, int tid) {
    R *Prow = Prow_all + (size_t)node * 4096, *PA = PA_all + (size_t)node * 4096, *PTA = PTA_all + (size_t)node * 4096;
    for (int idx = tid; idx < 4096; idx += 256) {
        Prow[idx] = (R)P[idx];
        const int w = idx >> 10, kk = (idx >> 6) & 15, l = idx & 63;
        const int row = 16 * w + (l & 15), col = 4 * kk + (l >> 4);
        PA[idx] = (R)P[row * 64 + col];
        PTA[idx] = (R)P[col * 64 + row];
    }
    if (tid < 64) {
        double s = 0.0;
        for (int j = 0; j < S; ++j) s += P[tid * 64 + j];
        Rsum_all[(size_t)node * 64 + tid] = (R)s;
    }
}

// ---------------------------------------------------------------------------
// k_pmat: mixed transition matrices per branch.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_pmat(PmatArgs<R> a) {
    const int bi = blockIdx.x, tid = threadIdx.x;
    const int node = a.blist[bi * BR_STRIDE], cls = a.blist[bi * BR_STRIDE + 3];
    __shared__ double P[64 * 64];
    __shared__ double Us[64 * 64];
    __shared__ double Vs[64 * 64];
    __shared__ double ex[64];
    const double *prm = a.params;
    const int j0 = a.comp_ptr[cls], j1 = a.comp_ptr[cls + 1];
    double conv = 1.0;
    if (a.convention == 1) {
        double mb = 0.0;
        for (int j = j0; j < j1; ++j) mb += prm[a.pl.off_w + j] * a.mu[a.comp_mat[j]];
        if (mb < 1e-12) mb = 1e-12;
        conv = 3.0 / mb;
    }
    const double tau = prm[a.pl.off_bl + node] * conv;
    if (tid == 0) a.tau[node] = tau;
    const int i0 = (tid >> 4) * 4, c0 = (tid & 15) * 4;
    double acc[4][4];
    for (int ii = 0; ii < 4; ++ii)
        for (int jj = 0; jj < 4; ++jj) acc[ii][jj] = 0.0;
    for (int j = j0; j < j1; ++j) {
        const int m = a.comp_mat[j];
        const double w = prm[a.pl.off_w + j];
        __syncthreads();
        if (tid < 64) ex[tid] = exp(a.lam[m * 64 + tid] * tau);
        __syncthreads();
        for (int idx = tid; idx < 4096; idx += 256) {
            Us[idx] = a.U[m * 4096 + idx] * ex[idx & 63];
            Vs[idx] = a.V[m * 4096 + idx];
        }
        __syncthreads();
        double t[4][4];
        for (int ii = 0; ii < 4; ++ii)
            for (int jj = 0; jj < 4; ++jj) t[ii][jj] = 0.0;
        for (int k = 0; k < 64; ++k) {
            double u[4], v[4];
            for (int ii = 0; ii < 4; ++ii) u[ii] = Us[(i0 + ii) * 64 + k];
            for (int jj = 0; jj < 4; ++jj) v[jj] = Vs[k * 64 + c0 + jj];
            for (int ii = 0; ii < 4; ++ii)
                for (int jj = 0; jj < 4; ++jj) t[ii][jj] = fma(u[ii], v[jj], t[ii][jj]);
        }
        for (int ii = 0; ii < 4; ++ii)
            for (int jj = 0; jj < 4; ++jj) acc[ii][jj] += w * (t[ii][jj] > 0.0 ? t[ii][jj] : 0.0);
    }
    for (int ii = 0; ii < 4; ++ii)
        for (int jj = 0; jj < 4; ++jj) P[(i0 + ii) * 64 + c0 + jj] = acc[ii][jj];
    __syncthreads();
    emit_p<R>(P, node, a.S, a.Prow, a.Rsum, a.PA, a.PTA, tid);
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
    emit_p<R>(P, node, S, a.Prow, a.Rsum, a.PA, a.PTA, tid);
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

template <typename R>
__device__ __forceinline__ R tip_val(const R *Pn, const R *Rn, const uint64_t *amb, unsigned code, int s) {
    if (code < 64u) return Pn[s * 64 + code];
    if (code == TIP_ONES) return Rn[s];
    uint64_t msk = amb[code & 0x7FFFu];
    R acc = 0;
    while (msk) {
        const int j = __builtin_ctzll(msk);
        acc += Pn[s * 64 + j];
        msk &= msk - 1;
    }
    return acc;
}

// Column-wise power-of-two normalisation of a 64xT tile held in C layout by
// 4 waves (wave w holds rows 16w..16w+15).  ex[nb] receives the exponent e
// with max in [0.5, 1) after scaling by 2^-e (0 for an all-zero column).
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
        x[nb][0] = ldexp_r(x[nb][0], -e);
        x[nb][1] = ldexp_r(x[nb][1], -e);
        x[nb][2] = ldexp_r(x[nb][2], -e);
        x[nb][3] = ldexp_r(x[nb][3], -e);
    }
    __syncthreads();
}

// Child message X = P_child * D_child (or the tip gather) in C layout.
template <typename R, int T>
__device__ __forceinline__ void child_message(const PruneArgs<R> &a, int pn, int ref, int tile, int p0, R *Xs,
                                              typename MF<R>::acc *X, int *Ex, int tid, int w, int l, int lr, int lg) {
    constexpr int NB = T / 16;
    typedef typename MF<R>::acc acc_t;
    const int kind = ref >> 28, idx = ref & REF_MASK;
    if (kind == KIND_TIP) {
        const R *Pn = a.Prow + (size_t)pn * 4096;
        const R *Rn = a.Rsum + (size_t)pn * 64;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const unsigned code = a.tips[(size_t)idx * a.P_pad + p0 + 16 * nb + lr];
#pragma unroll
            for (int r = 0; r < 4; ++r) X[nb][r] = tip_val(Pn, Rn, a.amb, code, 16 * w + MF<R>::crow(lg, r));
            Ex[nb] = 0;
        }
        return;
    }
    const R *src = a.D + ((size_t)idx * a.n_tiles + tile) * (64 * T);
    load_tile<R, T>(Xs, src, tid);
    __syncthreads();
    if (kind == KIND_IDENT) {
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int r = 0; r < 4; ++r) X[nb][r] = Xs[(16 * w + MF<R>::crow(lg, r)) * T + 16 * nb + lr];
    } else {
        const R *pa = a.PA + ((size_t)pn * 4 + w) * 1024;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) X[nb] = splat<acc_t>(0.0);
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            const R av = pa[kk * 64 + l];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) X[nb] = MF<R>::mma(av, Xs[(4 * kk + lg) * T + 16 * nb + lr], X[nb]);
        }
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) Ex[nb] = a.E[(size_t)idx * a.P_pad + p0 + 16 * nb + lr];
    __syncthreads();
}

// ---------------------------------------------------------------------------
// k_inside: post-order pruning of one pattern tile.
// ---------------------------------------------------------------------------
template <typename R, int T>
__global__ __launch_bounds__(256) void k_inside(PruneArgs<R> a) {
    constexpr int NB = T / 16;
    typedef typename MF<R>::acc acc_t;
    __shared__ R Xs[64 * T];
    __shared__ R cmax[4 * T];
    __shared__ double red[4 * T];
    const int tile = blockIdx.x, tid = threadIdx.x, w = tid >> 6, l = tid & 63, lr = l & 15, lg = l >> 4;
    const int p0 = tile * T;
    acc_t Dc[NB];
    int Ec[NB];
    for (int k = 0; k < a.n_sched; ++k) {
        const int32_t *e = a.sched + k * SCHED_STRIDE;
        const int slot = e[1], nc = e[2];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            Dc[nb] = splat<acc_t>(1.0);
            Ec[nb] = 0;
        }
        for (int c = 0; c < nc; ++c) {
            acc_t X[NB];
            int Ex[NB];
            child_message<R, T>(a, e[6 + 3 * c], e[7 + 3 * c], tile, p0, Xs, X, Ex, tid, w, l, lr, lg);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                Dc[nb] *= X[nb];
                Ec[nb] += Ex[nb];
            }
            if (c >= 1 && c + 1 < nc) {
                int ex[NB];
                normalize_cols<R, T>(Dc, ex, cmax, w, lr, lg);
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) Ec[nb] += ex[nb];
            }
        }
        int ex[NB];
        normalize_cols<R, T>(Dc, ex, cmax, w, lr, lg);
        R *dst = a.D + ((size_t)slot * a.n_tiles + tile) * (64 * T);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            Ec[nb] += ex[nb];
#pragma unroll
            for (int r = 0; r < 4; ++r) dst[(16 * w + MF<R>::crow(lg, r)) * T + 16 * nb + lr] = Dc[nb][r];
            if (w == 0 && lg == 0) a.E[(size_t)slot * a.P_pad + p0 + 16 * nb + lr] = Ec[nb];
        }
        __syncthreads();
    }
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
// k_outside: pre-order outside messages V_c = U_u (.) prod_{siblings} X_s.
// ---------------------------------------------------------------------------
template <typename R, int T>
__global__ __launch_bounds__(256) void k_outside(PruneArgs<R> a) {
    constexpr int NB = T / 16;
    typedef typename MF<R>::acc acc_t;
    __shared__ R Xs[64 * T];
    __shared__ R cmax[4 * T];
    const int tile = blockIdx.x, tid = threadIdx.x, w = tid >> 6, l = tid & 63, lr = l & 15, lg = l >> 4;
    const int p0 = tile * T;
    for (int k = a.n_sched - 1; k >= 0; --k) {
        const int32_t *e = a.sched + k * SCHED_STRIDE;
        const int node = e[0], nc = e[2], vs = e[3], ident = e[4];
        acc_t U[NB];
        int H[NB];
        if (vs < 0) {
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
#pragma unroll
                for (int r = 0; r < 4; ++r) U[nb][r] = (R)a.pi[16 * w + MF<R>::crow(lg, r)];
                H[nb] = 0;
            }
        } else {
            load_tile<R, T>(Xs, a.V + ((size_t)vs * a.n_tiles + tile) * (64 * T), tid);
            __syncthreads();
            if (ident) {
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                    for (int r = 0; r < 4; ++r) U[nb][r] = Xs[(16 * w + MF<R>::crow(lg, r)) * T + 16 * nb + lr];
            } else {
                const R *pa = a.PTA + ((size_t)node * 4 + w) * 1024;
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) U[nb] = splat<acc_t>(0.0);
#pragma unroll
                for (int kk = 0; kk < 16; ++kk) {
                    const R av = pa[kk * 64 + l];
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb) U[nb] = MF<R>::mma(av, Xs[(4 * kk + lg) * T + 16 * nb + lr], U[nb]);
                }
            }
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) H[nb] = a.G[(size_t)vs * a.P_pad + p0 + 16 * nb + lr];
            __syncthreads();
        }
        acc_t X[MAXC][NB];
        int Ex[MAXC][NB];
#pragma unroll
        for (int c = 0; c < MAXC; ++c)
            if (c < nc) child_message<R, T>(a, e[6 + 3 * c], e[7 + 3 * c], tile, p0, Xs, X[c], Ex[c], tid, w, l, lr, lg);
#pragma unroll
        for (int c = 0; c < MAXC; ++c) {
            if (c >= nc) continue;
            acc_t Vv[NB];
            int Gv[NB];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                Vv[nb] = U[nb];
                Gv[nb] = H[nb];
            }
#pragma unroll
            for (int c2 = 0; c2 < MAXC; ++c2) {
                if (c2 == c || c2 >= nc) continue;
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) {
                    Vv[nb] *= X[c2][nb];
                    Gv[nb] += Ex[c2][nb];
                }
            }
            int ex[NB];
            normalize_cols<R, T>(Vv, ex, cmax, w, lr, lg);
            const int vsc = e[8 + 3 * c];
            R *dst = a.V + ((size_t)vsc * a.n_tiles + tile) * (64 * T);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                Gv[nb] += ex[nb];
#pragma unroll
                for (int r = 0; r < 4; ++r) dst[(16 * w + MF<R>::crow(lg, r)) * T + 16 * nb + lr] = Vv[nb][r];
                if (w == 0 && lg == 0) a.G[(size_t)vsc * a.P_pad + p0 + 16 * nb + lr] = Gv[nb];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// k_adjoint: A_b partial = sum_{p in chunk} c_p V_b[:,p] D_b[:,p]^T with
// c_p = w_p / L_p * 2^(G_b + E_b - E_root).  MFMA in the working type within
// a tile, FP64 accumulation across tiles.
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ R tip_onehot(unsigned code, int j, const uint64_t *amb, int S) {
    if (code < 64u) return (R)(code == (unsigned)j ? 1 : 0);
    if (code == TIP_ONES) return (R)(j < S ? 1 : 0);
    return (R)((amb[code & 0x7FFFu] >> j) & 1ull);
}

template <typename R, int T>
__global__ __launch_bounds__(256) void k_adjoint(AdjArgs<R> a) {
    typedef typename MF<R>::acc acc_t;
    __shared__ R Vs[64 * T];
    __shared__ R Ds[64 * T];
    __shared__ R cs[T];
    __shared__ unsigned codes[T];
    const int chunk = blockIdx.x, bi = blockIdx.y;
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, lr = l & 15, lg = l >> 4;
    const int32_t *br = a.blist + bi * BR_STRIDE;
    const int vslot = br[1], ref = br[2];
    const int kind = ref >> 28, idx = ref & REF_MASK;
    f64x4 acc[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[cb] = splat<f64x4>(0.0);
    const int t0 = chunk * a.tiles_per_chunk;
    int t1 = t0 + a.tiles_per_chunk;
    if (t1 > a.n_tiles) t1 = a.n_tiles;
    for (int tile = t0; tile < t1; ++tile) {
        const int p0 = tile * T;
        load_tile<R, T>(Vs, a.V + ((size_t)vslot * a.n_tiles + tile) * (64 * T), tid);
        if (kind == KIND_TIP) {
            if (tid < T) codes[tid] = a.tips[(size_t)idx * a.P_pad + p0 + tid];
        } else {
            load_tile<R, T>(Ds, a.D + ((size_t)idx * a.n_tiles + tile) * (64 * T), tid);
        }
        if (tid < T) {
            const int p = p0 + tid;
            const double L = a.Lroot[p], wt = a.weights[p];
            double c = 0.0;
            if (wt > 0.0 && L > 0.0) {
                int ex = a.G[(size_t)vslot * a.P_pad + p] - a.Eroot[p];
                if (kind != KIND_TIP) ex += a.E[(size_t)idx * a.P_pad + p];
                c = ldexp(wt / L, ex);
            }
            cs[tid] = (R)c;
        }
        __syncthreads();
        acc_t loc[4];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) loc[cb] = splat<acc_t>(0.0);
#pragma unroll
        for (int kk = 0; kk < T / 4; ++kk) {
            const int pl = 4 * kk + lg;
            const R av = Vs[(16 * w + lr) * T + pl] * cs[pl];
            if (kind == KIND_TIP) {
                const unsigned code = codes[pl];
#pragma unroll
                for (int cb = 0; cb < 4; ++cb)
                    loc[cb] = MF<R>::mma(av, tip_onehot<R>(code, 16 * cb + lr, a.amb, a.S), loc[cb]);
            } else {
#pragma unroll
                for (int cb = 0; cb < 4; ++cb) loc[cb] = MF<R>::mma(av, Ds[(16 * cb + lr) * T + pl], loc[cb]);
            }
        }
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[cb][r] += (double)loc[cb][r];
        __syncthreads();
    }
    double *out = a.Apart + ((size_t)bi * a.n_chunks + chunk) * 4096;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int r = 0; r < 4; ++r) out[(16 * w + MF<R>::crow(lg, r)) * 64 + 16 * cb + lr] = acc[cb][r];
}

// ---------------------------------------------------------------------------
// k_proj: M_phi = V_pi dQ_phi U_pi and dmu_phi = -sum_i pi_i dQ_phi[i][i]
// (likelihood.hpp:1497-1562).  phi 0 = omega, 1..5 = theta AC, AT, CG, CT, GT.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_proj(ProjArgs a) {
    const int m = blockIdx.x, phi = blockIdx.y, tid = threadIdx.x;
    __shared__ double dQ[64 * 64];
    __shared__ double T1[64 * 64];
    __shared__ double Ms[64 * 64];
    const double *prm = a.params;
    const double alpha = prm[a.pl.off_alpha], beta = prm[a.pl.off_omega + m];
    const double *th = prm + a.pl.off_theta;
    const int S = a.S;
    const int pairs[5][2] = {{0, 1}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    for (int idx = tid; idx < 4096; idx += 256) {
        const int i = idx >> 6, j = idx & 63;
        double q = 0.0;
        if (i < S && j < S && i != j) {
            const int k = i * 64 + j;
            if (a.nd[k] == 1) {
                const int from = a.dfrom[3 * k], to = a.dto[3 * k], ps = a.dpos[3 * k];
                if (phi == 0) {
                    if (!a.syn[k]) q = nuc_rate(th, from, to) * a.pos[ps * 4 + to];
                } else {
                    const int u = pairs[phi - 1][0], v = pairs[phi - 1][1];
                    if ((from == u && to == v) || (from == v && to == u)) q = (a.syn[k] ? alpha : beta) * a.pos[ps * 4 + to];
                }
            }
        }
        dQ[idx] = q;
        Ms[idx] = a.U[m * 4096 + idx];
    }
    __syncthreads();
    if (tid < 64) {
        double s = 0.0;
        for (int j = 0; j < 64; ++j)
            if (j != tid) s += dQ[tid * 64 + j];
        dQ[tid * 64 + tid] = -s;
    }
    __syncthreads();
    if (tid == 0) {
        double d = 0.0;
        for (int i = 0; i < S; ++i) d -= a.pi[i] * dQ[i * 64 + i];
        a.dmu[m * 6 + phi] = d;
    }
    const int i0 = (tid >> 4) * 4, c0 = (tid & 15) * 4;
    double t[4][4];
    for (int ii = 0; ii < 4; ++ii)
        for (int jj = 0; jj < 4; ++jj) t[ii][jj] = 0.0;
    for (int k = 0; k < 64; ++k)
        for (int ii = 0; ii < 4; ++ii) {
            const double d = dQ[(i0 + ii) * 64 + k];
            for (int jj = 0; jj < 4; ++jj) t[ii][jj] = fma(d, Ms[k * 64 + c0 + jj], t[ii][jj]);
        }
    for (int ii = 0; ii < 4; ++ii)
        for (int jj = 0; jj < 4; ++jj) T1[(i0 + ii) * 64 + c0 + jj] = t[ii][jj];
    __syncthreads();
    for (int idx = tid; idx < 4096; idx += 256) Ms[idx] = a.V[m * 4096 + idx];
    __syncthreads();
    for (int ii = 0; ii < 4; ++ii)
        for (int jj = 0; jj < 4; ++jj) t[ii][jj] = 0.0;
    for (int k = 0; k < 64; ++k)
        for (int ii = 0; ii < 4; ++ii) {
            const double v = Ms[(i0 + ii) * 64 + k];
            for (int jj = 0; jj < 4; ++jj) t[ii][jj] = fma(v, T1[k * 64 + c0 + jj], t[ii][jj]);
        }
    double *M = a.M + ((size_t)m * 6 + phi) * 4096;
    for (int ii = 0; ii < 4; ++ii)
        for (int jj = 0; jj < 4; ++jj) M[(i0 + ii) * 64 + c0 + jj] = t[ii][jj];
}

// ---------------------------------------------------------------------------
// k_frechet: per (branch, component) spectral projections (FP64).
// out[0] = <A_b, P_m(tau)>, out[1] = w <A_b, Q_m P_m(tau)>,
// out[2+phi] = w <M_phi (.) E_b, U^T A_b V^T>.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_frechet(FrechetArgs a) {
    const int bi = blockIdx.x, jj = blockIdx.y, tid = threadIdx.x;
    const int node = a.blist[bi * BR_STRIDE], cls = a.blist[bi * BR_STRIDE + 3];
    const int J = a.comp_ptr[cls] + jj;
    if (J >= a.comp_ptr[cls + 1]) return;
    const int m = a.comp_mat[J];
    const double wj = a.params[a.pl.off_w + J];
    const double tau = a.tau[node];
    __shared__ double As[64 * 64];
    __shared__ double Bs[64 * 64];
    __shared__ double lam[64], ex[64];
    __shared__ double red[8][4];
    for (int idx = tid; idx < 4096; idx += 256) {
        double s = 0.0;
        for (int c = 0; c < a.n_chunks; ++c) s += a.Apart[((size_t)bi * a.n_chunks + c) * 4096 + idx];
        As[idx] = s;
        Bs[idx] = a.V[m * 4096 + idx];
    }
    if (tid < 64) {
        lam[tid] = a.lam[m * 64 + tid];
        ex[tid] = exp(lam[tid] * tau);
    }
    __syncthreads();
    const int i0 = (tid >> 4) * 4, c0 = (tid & 15) * 4;
    double t[4][4];
    for (int ii = 0; ii < 4; ++ii)
        for (int q = 0; q < 4; ++q) t[ii][q] = 0.0;
    // T1 = A V^T : T1[i][j] = sum_k A[i][k] V[j][k]
    for (int k = 0; k < 64; ++k)
        for (int ii = 0; ii < 4; ++ii) {
            const double av = As[(i0 + ii) * 64 + k];
            for (int q = 0; q < 4; ++q) t[ii][q] = fma(av, Bs[(c0 + q) * 64 + k], t[ii][q]);
        }
    __syncthreads();
    for (int ii = 0; ii < 4; ++ii)
        for (int q = 0; q < 4; ++q) As[(i0 + ii) * 64 + c0 + q] = t[ii][q];
    for (int idx = tid; idx < 4096; idx += 256) Bs[idx] = a.U[m * 4096 + idx];
    __syncthreads();
    // At = U^T T1 : At[i][j] = sum_k U[k][i] T1[k][j]
    for (int ii = 0; ii < 4; ++ii)
        for (int q = 0; q < 4; ++q) t[ii][q] = 0.0;
    for (int k = 0; k < 64; ++k)
        for (int ii = 0; ii < 4; ++ii) {
            const double u = Bs[k * 64 + i0 + ii];
            for (int q = 0; q < 4; ++q) t[ii][q] = fma(u, As[k * 64 + c0 + q], t[ii][q]);
        }
    double acc[8];
    for (int q = 0; q < 8; ++q) acc[q] = 0.0;
    const double *M = a.M + (size_t)m * 6 * 4096;
    for (int ii = 0; ii < 4; ++ii) {
        const int i = i0 + ii;
        for (int q = 0; q < 4; ++q) {
            const int j = c0 + q;
            const double at = t[ii][q];
            if (i == j) {
                acc[0] += ex[i] * at;
                acc[1] += lam[i] * ex[i] * at;
            }
            const double d = lam[i] - lam[j];
            const double E = (fabs(d) < 1e-9) ? tau * ex[i] : (ex[i] - ex[j]) / d;
            const double ea = E * at;
            for (int phi = 0; phi < 6; ++phi) acc[2 + phi] += M[(size_t)phi * 4096 + i * 64 + j] * ea;
        }
    }
    // Deterministic block reduction.
    const int lane = tid & 63, wv = tid >> 6;
    for (int q = 0; q < 8; ++q) {
        double v = acc[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
        if (lane == 0) red[q][wv] = v;
    }
    __syncthreads();
    if (tid < 8) {
        double v = ((red[tid][0] + red[tid][1]) + red[tid][2]) + red[tid][3];
        if (tid >= 1) v *= wj;
        a.out[((size_t)bi * a.max_comp + jj) * 8 + tid] = v;
    }
}

// ---------------------------------------------------------------------------
// k_final: deterministic reductions into the result vector.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void neum(double &s, double &c, double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x))
        c += (s - t) + x;
    else
        c += (x - t) + s;
    s = t;
}

__global__ __launch_bounds__(256) void k_final(FinalArgs a) {
    const int tid = threadIdx.x;
    __shared__ double ps[256], pc[256];
    __shared__ double sum_tau[64];
    const int P = a.n_patterns;
    const int per = (P + 255) / 256;
    double s = 0.0, c = 0.0;
    for (int p = tid * per; p < (tid + 1) * per && p < P; ++p) {
        const double w = a.weights[p];
        const double ll = a.site_ll[p];
        if (a.want_grad && !(a.Lroot[p] > 0.0)) continue;
        neum(s, c, w * ll);
    }
    ps[tid] = s;
    pc[tid] = c;
    __syncthreads();
    if (tid == 0) {
        double S0 = 0.0, C0 = 0.0;
        for (int i = 0; i < 256; ++i) {
            neum(S0, C0, ps[i]);
            C0 += pc[i];
        }
        a.res[0] = S0 + C0;
    }
    if (!a.want_grad) return;
    const double *prm = a.params;
    const int NR = a.n_real_nodes, NM = a.pl.n_mat, NCP = a.pl.n_comp;
    double *gbl = a.res + 1, *gom = gbl + NR, *gw = gom + NM, *gth = gw + NCP;
    // per-class conversion and sum_b tau_b dlnL/dtau_b
    if (tid < a.n_class) {
        double st = 0.0;
        for (int bi = 0; bi < a.n_branch; ++bi) {
            if (a.blist[bi * BR_STRIDE + 3] != tid) continue;
            const int node = a.blist[bi * BR_STRIDE];
            double dt = 0.0;
            for (int j = 0; j < a.max_comp; ++j) dt += a.fr[((size_t)bi * a.max_comp + j) * 8 + 1];
            st += a.tau[node] * dt;
        }
        sum_tau[tid] = st;
    }
    __syncthreads();
    for (int n = tid; n < NR; n += 256) gbl[n] = 0.0;
    __syncthreads();
    for (int bi = tid; bi < a.n_branch; bi += 256) {
        const int node = a.blist[bi * BR_STRIDE], cls = a.blist[bi * BR_STRIDE + 3];
        double conv = 1.0;
        if (a.convention == 1) {
            double mb = 0.0;
            for (int j = a.comp_ptr[cls]; j < a.comp_ptr[cls + 1]; ++j) mb += prm[a.pl.off_w + j] * a.mu[a.comp_mat[j]];
            if (mb < 1e-12) mb = 1e-12;
            conv = 3.0 / mb;
        }
        double dt = 0.0;
        for (int j = 0; j < a.max_comp; ++j) dt += a.fr[((size_t)bi * a.max_comp + j) * 8 + 1];
        if (node < NR) gbl[node] = conv * dt;
    }
    // omega (phi 0) and theta (phi 1..5): one output per thread, branch order.
    if (tid < NM + 5) {
        const bool is_om = tid < NM;
        const int phi = is_om ? 0 : (tid - NM + 1);
        double g = 0.0;
        for (int bi = 0; bi < a.n_branch; ++bi) {
            const int cls = a.blist[bi * BR_STRIDE + 3];
            for (int j = a.comp_ptr[cls]; j < a.comp_ptr[cls + 1]; ++j) {
                if (is_om && a.comp_mat[j] != tid) continue;
                g += a.fr[((size_t)bi * a.max_comp + (j - a.comp_ptr[cls])) * 8 + 2 + phi];
            }
        }
        if (a.convention == 1) {
            for (int cl = 0; cl < a.n_class; ++cl) {
                double mb = 0.0;
                for (int j = a.comp_ptr[cl]; j < a.comp_ptr[cl + 1]; ++j) mb += prm[a.pl.off_w + j] * a.mu[a.comp_mat[j]];
                if (mb < 1e-12) mb = 1e-12;
                const double f = sum_tau[cl] / mb;
                for (int j = a.comp_ptr[cl]; j < a.comp_ptr[cl + 1]; ++j) {
                    const int mm = a.comp_mat[j];
                    if (is_om && mm != tid) continue;
                    g -= f * prm[a.pl.off_w + j] * a.dmu[mm * 6 + phi];
                }
            }
        }
        if (is_om)
            gom[tid] = g;
        else
            gth[phi - 1] = g;
    }
    // mixture weights
    if (tid >= 64 && tid - 64 < NCP) {
        const int J = tid - 64;
        int cls = 0;
        while (cls + 1 <= a.n_class && a.comp_ptr[cls + 1] <= J) ++cls;
        const int jj = J - a.comp_ptr[cls];
        double g = 0.0;
        for (int bi = 0; bi < a.n_branch; ++bi) {
            if (a.blist[bi * BR_STRIDE + 3] != cls) continue;
            g += a.fr[((size_t)bi * a.max_comp + jj) * 8 + 0];
        }
        if (a.convention == 1) {
            double mb = 0.0;
            for (int j = a.comp_ptr[cls]; j < a.comp_ptr[cls + 1]; ++j) mb += prm[a.pl.off_w + j] * a.mu[a.comp_mat[j]];
            if (mb < 1e-12) mb = 1e-12;
            g -= sum_tau[cls] / mb * a.mu[a.comp_mat[J]];
        }
        gw[J] = g;
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
void launch_model(const ModelArgs &a, int n_mat, hipStream_t s) {
    hipLaunchKernelGGL(k_model, dim3(n_mat), dim3(256), 0, s, a);
}

template <typename R> void launch_pmat(const PmatArgs<R> &a, hipStream_t s) {
    hipLaunchKernelGGL(k_pmat<R>, dim3(a.n_branch), dim3(256), 0, s, a);
}

template <typename R> void launch_inside(const PruneArgs<R> &a, int T, hipStream_t s) {
    switch (T) {
        case 16: hipLaunchKernelGGL((k_inside<R, 16>), dim3(a.n_tiles), dim3(256), 0, s, a); break;
        case 32: hipLaunchKernelGGL((k_inside<R, 32>), dim3(a.n_tiles), dim3(256), 0, s, a); break;
        default: hipLaunchKernelGGL((k_inside<R, 64>), dim3(a.n_tiles), dim3(256), 0, s, a); break;
    }
}

template <typename R> void launch_outside(const PruneArgs<R> &a, int T, hipStream_t s) {
    switch (T) {
        case 16: hipLaunchKernelGGL((k_outside<R, 16>), dim3(a.n_tiles), dim3(256), 0, s, a); break;
        case 32: hipLaunchKernelGGL((k_outside<R, 32>), dim3(a.n_tiles), dim3(256), 0, s, a); break;
        default: hipLaunchKernelGGL((k_outside<R, 64>), dim3(a.n_tiles), dim3(256), 0, s, a); break;
    }
}

template <typename R> void launch_adjoint(const AdjArgs<R> &a, int T, hipStream_t s) {
    const dim3 grid(a.n_chunks, a.n_branch);
    switch (T) {
        case 16: hipLaunchKernelGGL((k_adjoint<R, 16>), grid, dim3(256), 0, s, a); break;
        case 32: hipLaunchKernelGGL((k_adjoint<R, 32>), grid, dim3(256), 0, s, a); break;
        default: hipLaunchKernelGGL((k_adjoint<R, 64>), grid, dim3(256), 0, s, a); break;
    }
}

template <typename R> void launch_pmat_host(const PmatArgs<R> &a, const double *Phost, hipStream_t s) {
    hipLaunchKernelGGL(k_pmat_host<R>, dim3(a.n_branch), dim3(256), 0, s, a, Phost);
}

void launch_proj(const ProjArgs &a, int n_mat, hipStream_t s) {
    hipLaunchKernelGGL(k_proj, dim3(n_mat, 6), dim3(256), 0, s, a);
}

void launch_frechet(const FrechetArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_frechet, dim3(a.n_branch, a.max_comp), dim3(256), 0, s, a);
}

void launch_final(const FinalArgs &a, hipStream_t s) { hipLaunchKernelGGL(k_final, dim3(1), dim3(256), 0, s, a); }

template void launch_pmat<float>(const PmatArgs<float> &, hipStream_t);
template void launch_pmat<double>(const PmatArgs<double> &, hipStream_t);
template void launch_pmat_host<float>(const PmatArgs<float> &, const double *, hipStream_t);
template void launch_pmat_host<double>(const PmatArgs<double> &, const double *, hipStream_t);
template void launch_inside<float>(const PruneArgs<float> &, int, hipStream_t);
template void launch_inside<double>(const PruneArgs<double> &, int, hipStream_t);
template void launch_outside<float>(const PruneArgs<float> &, int, hipStream_t);
template void launch_outside<double>(const PruneArgs<double> &, int, hipStream_t);
template void launch_adjoint<float>(const AdjArgs<float> &, int, hipStream_t);
template void launch_adjoint<double>(const AdjArgs<double> &, int, hipStream_t);

}  // namespace hy
