/*
 * hyoracle.c -- FP64 CPU restatement of HyPhy 3's codon likelihood core.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and CPU baseline for the MI355X
 * library.  Loaded only by tests/, the smoke() entry point and bench.py's
 * CPU legs (cpu_baseline, the reference-fit and E-step checks).  Nothing in
 * the product links, loads or calls this code.
 *
 * Reference files restated here (paths relative to the reference root):
 *   include/hyphy/core/genetic_code.hpp   (sense-codon numbering, diff table)
 *   include/hyphy/core/rate_matrix.hpp    (MG94 Q, symmetric eigen, P(t))
 *   include/hyphy/core/likelihood.hpp     (pruning, inside-outside, gradients)
 *   include/hyphy/analyses/busted.hpp     (branch-site mixed P, tile rule)
 * The reference's Eigen GEMMs and SelfAdjointEigenSolver are replaced by plain
 * loops and a Householder + implicit-QR symmetric eigensolver written here
 * (Golub & Van Loan, Matrix Computations, 4th ed., Alg. 8.3.1 and 8.3.3).
 */
#include "hyoracle.h"

#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Genetic codes.  Universal table in ACGT-lexicographic codon order          */
/* (genetic_code.hpp:248-253), other registry entries as per-codon overrides */
/* (genetic_code.hpp:274-370).                                                */
/* ------------------------------------------------------------------------- */
static const char UNIVERSAL_ACGT[65] =
    "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF";

typedef struct { int c64; char aa; } code_override;

static const code_override OV_1[] = {{8, '*'}, {10, '*'}, {12, 'M'}, {56, 'W'}, {-1, 0}};
static const code_override OV_2[] = {{12, 'M'}, {28, 'T'}, {29, 'T'}, {30, 'T'}, {31, 'T'}, {56, 'W'}, {-1, 0}};
static const code_override OV_3[] = {{56, 'W'}, {-1, 0}};
static const code_override OV_4[] = {{8, 'S'}, {10, 'S'}, {12, 'M'}, {56, 'W'}, {-1, 0}};
static const code_override OV_5[] = {{48, 'Q'}, {50, 'Q'}, {-1, 0}};
static const code_override OV_6[] = {{0, 'N'}, {8, 'S'}, {10, 'S'}, {56, 'W'}, {-1, 0}};
static const code_override OV_7[] = {{56, 'C'}, {-1, 0}};
static const code_override OV_8[] = {{30, 'S'}, {-1, 0}};
static const code_override OV_9[] = {{8, 'G'}, {10, 'G'}, {12, 'M'}, {56, 'W'}, {-1, 0}};
static const code_override OV_10[] = {{0, 'N'}, {8, 'S'}, {10, 'S'}, {48, 'Y'}, {56, 'W'}, {-1, 0}};
static const code_override OV_11[] = {{50, 'Q'}, {-1, 0}};
static const code_override OV_12[] = {{50, 'L'}, {-1, 0}};
static const code_override OV_13[] = {{0, 'N'}, {8, 'S'}, {10, 'S'}, {12, 'M'}, {56, 'W'}, {-1, 0}};
static const code_override OV_14[] = {{50, 'L'}, {52, '*'}, {-1, 0}};
static const code_override OV_15[] = {{60, '*'}, {-1, 0}};
static const code_override OV_16[] = {{8, 'S'}, {10, 'K'}, {56, 'W'}, {-1, 0}};
static const code_override OV_17[] = {{56, 'G'}, {-1, 0}};
static const code_override OV_18[] = {{30, 'A'}, {-1, 0}};
static const code_override OV_19[] = {{48, 'Y'}, {50, 'Y'}, {-1, 0}};
static const code_override OV_20[] = {{48, 'E'}, {50, 'E'}, {-1, 0}};
static const code_override OV_21[] = {{8, 'S'}, {10, 'K'}, {48, 'Y'}, {56, 'W'}, {-1, 0}};

static const code_override *const CODE_OVERRIDES[22] = {
    NULL, OV_1, OV_2, OV_3, OV_4, OV_5, OV_6, OV_7, OV_8, OV_9, OV_10,
    OV_11, OV_12, OV_13, OV_14, OV_15, OV_16, OV_17, OV_18, OV_19, OV_20, OV_21};

int or_code_build(int32_t code_id, or_code *out) {
    if (code_id < 0 || code_id >= 22 || out == NULL) return -1;
    memset(out, 0, sizeof(*out));
    out->code_id = code_id;
    memcpy(out->aa, UNIVERSAL_ACGT, 64);
    if (CODE_OVERRIDES[code_id]) {
        for (const code_override *o = CODE_OVERRIDES[code_id]; o->c64 >= 0; ++o) out->aa[o->c64] = o->aa;
    }
    int s = 0;
    for (int c = 0; c < 64; ++c) {
        out->c64_to_sense[c] = -1;
        out->sense_to_c64[c] = -1;
    }
    for (int c = 0; c < 64; ++c) {
        if (out->aa[c] == '*') continue;
        out->c64_to_sense[c] = (int8_t)s;
        out->sense_to_c64[s] = (int8_t)c;
        ++s;
    }
    out->S = s;
    /* Pairwise difference table (genetic_code.hpp:174-199). */
    for (int u = 0; u < s; ++u) {
        int cu = out->sense_to_c64[u];
        int nu[3] = {cu >> 4, (cu >> 2) & 3, cu & 3};
        for (int v = 0; v < s; ++v) {
            int idx = u * 64 + v;
            for (int k = 0; k < 3; ++k) {
                out->dpos[idx * 3 + k] = -1;
                out->dfrom[idx * 3 + k] = -1;
                out->dto[idx * 3 + k] = -1;
            }
            if (u == v) continue;
            int cv = out->sense_to_c64[v];
            int nv[3] = {cv >> 4, (cv >> 2) & 3, cv & 3};
            int d = 0;
            for (int p = 0; p < 3; ++p) {
                if (nu[p] != nv[p]) {
                    out->dpos[idx * 3 + d] = (int8_t)p;
                    out->dfrom[idx * 3 + d] = (int8_t)nu[p];
                    out->dto[idx * 3 + d] = (int8_t)nv[p];
                    ++d;
                }
            }
            out->nd[idx] = (int8_t)d;
            out->syn[idx] = (uint8_t)(out->aa[cu] == out->aa[cv]);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* MG94 rate matrix (rate_matrix.hpp:165-234).                               */
/* ------------------------------------------------------------------------- */
static void nuc_rates(const or_mg94_params *p, double r[4][4]) {
    /* theta_AG == 1 (rate_matrix.hpp:180-185) */
    double t_ac = p->theta[0], t_at = p->theta[1], t_cg = p->theta[2], t_ct = p->theta[3], t_gt = p->theta[4];
    double m[4][4] = {{0.0, t_ac, 1.0, t_at}, {t_ac, 0.0, t_cg, t_ct}, {1.0, t_cg, 0.0, t_gt}, {t_at, t_ct, t_gt, 0.0}};
    memcpy(r, m, sizeof(m));
}

void or_mg94_Q(const or_code *code, const or_mg94_params *p, const double *pos, const double *pi, double *Q,
               double *mu) {
    const int S = code->S;
    double r[4][4];
    nuc_rates(p, r);
    for (int i = 0; i < S; ++i) {
        double row = 0.0;
        for (int j = 0; j < S; ++j) {
            double q = 0.0;
            if (i != j) {
                const int idx = i * 64 + j;
                const double rate_mod = code->syn[idx] ? p->alpha : p->beta;
                const int nd = code->nd[idx];
                const int8_t *dp = code->dpos + idx * 3, *df = code->dfrom + idx * 3, *dt = code->dto + idx * 3;
                if (nd == 1) {
                    q = rate_mod * r[df[0]][dt[0]] * pos[dp[0] * 4 + dt[0]];
                } else if (nd == 2 && p->delta > 0.0) {
                    q = rate_mod * p->delta * (r[df[0]][dt[0]] * r[df[1]][dt[1]]) *
                        (pos[dp[0] * 4 + dt[0]] * pos[dp[1] * 4 + dt[1]]);
                } else if (nd == 3 && p->psi > 0.0) {
                    q = rate_mod * p->psi * (r[df[0]][dt[0]] * r[df[1]][dt[1]] * r[df[2]][dt[2]]) *
                        (pos[dp[0] * 4 + dt[0]] * pos[dp[1] * 4 + dt[1]] * pos[dp[2] * 4 + dt[2]]);
                }
            }
            Q[i * S + j] = q;
            row += q;
        }
        Q[i * S + i] = -row;
    }
    double avg = 0.0;
    for (int i = 0; i < S; ++i) avg -= pi[i] * Q[i * S + i];
    if (mu) *mu = avg;
}

/* ------------------------------------------------------------------------- */
/* Symmetric eigensolver: Householder tridiagonalisation + implicit QR with  */
/* Wilkinson shifts, accumulating the orthogonal factor.                     */
/* ------------------------------------------------------------------------- */
int or_sym_eigen(int32_t n, double *A, double *evals, double *Z) {
    /* Z <- I */
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Z[i * n + j] = (i == j) ? 1.0 : 0.0;
    double *v = (double *)malloc(sizeof(double) * (size_t)n * 3);
    double *pv = v + n, *qv = v + 2 * n;
    for (int k = 0; k + 2 < n; ++k) {
        const int m = n - k - 1; /* length of x = A[k+1.., k] */
        double xnorm = 0.0;
        for (int i = 0; i < m; ++i) xnorm = hypot(xnorm, A[(k + 1 + i) * n + k]);
        if (xnorm == 0.0) continue;
        const double x0 = A[(k + 1) * n + k];
        const double alpha = (x0 >= 0.0) ? -xnorm : xnorm;
        for (int i = 0; i < m; ++i) v[i] = A[(k + 1 + i) * n + k];
        v[0] -= alpha;
        double vnorm = 0.0;
        for (int i = 0; i < m; ++i) vnorm = hypot(vnorm, v[i]);
        if (vnorm == 0.0) continue;
        for (int i = 0; i < m; ++i) v[i] /= vnorm;
        /* p = A_sub v, K = v^T p, q = p - K v ; A_sub -= 2 v q^T + 2 q v^T */
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            const double *row = A + (size_t)(k + 1 + i) * n + (k + 1);
            for (int j = 0; j < m; ++j) s += row[j] * v[j];
            pv[i] = s;
        }
        double K = 0.0;
        for (int i = 0; i < m; ++i) K += v[i] * pv[i];
        for (int i = 0; i < m; ++i) qv[i] = pv[i] - K * v[i];
        for (int i = 0; i < m; ++i) {
            double *row = A + (size_t)(k + 1 + i) * n + (k + 1);
            for (int j = 0; j < m; ++j) row[j] -= 2.0 * (v[i] * qv[j] + qv[i] * v[j]);
        }
        A[(k + 1) * n + k] = alpha;
        A[k * n + (k + 1)] = alpha;
        for (int i = 1; i < m; ++i) {
            A[(k + 1 + i) * n + k] = 0.0;
            A[k * n + (k + 1 + i)] = 0.0;
        }
        /* Z[:, k+1..] <- Z[:, k+1..] (I - 2 v v^T) */
        for (int r = 0; r < n; ++r) {
            double *zr = Z + (size_t)r * n + (k + 1);
            double s = 0.0;
            for (int j = 0; j < m; ++j) s += zr[j] * v[j];
            s *= 2.0;
            for (int j = 0; j < m; ++j) zr[j] -= s * v[j];
        }
    }
    free(v);
    double *d = evals;
    double *e = (double *)calloc((size_t)(n > 1 ? n : 2), sizeof(double));
    for (int i = 0; i < n; ++i) d[i] = A[i * n + i];
    for (int i = 0; i + 1 < n; ++i) e[i] = A[(i + 1) * n + i];
    int hi = n - 1;
    int iter = 0;
    const int max_iter = 60 * n + 60;
    while (hi > 0) {
        if (fabs(e[hi - 1]) <= DBL_EPSILON * (fabs(d[hi - 1]) + fabs(d[hi])) || e[hi - 1] == 0.0) {
            e[hi - 1] = 0.0;
            --hi;
            continue;
        }
        int lo = hi - 1;
        while (lo > 0 && !(fabs(e[lo - 1]) <= DBL_EPSILON * (fabs(d[lo - 1]) + fabs(d[lo])))) --lo;
        if (lo > 0) e[lo - 1] = 0.0;
        if (++iter > max_iter) {
            free(e);
            return -1;
        }
        /* Wilkinson shift from the trailing 2x2 block. */
        const double a = d[hi - 1], b = e[hi - 1], c = d[hi];
        const double delta = 0.5 * (a - c);
        const double denom = delta + copysign(hypot(delta, b), delta == 0.0 ? 1.0 : delta);
        const double mu = c - b * b / denom;
        double x = d[lo] - mu;
        double z = e[lo];
        for (int k = lo; k < hi; ++k) {
            double r = hypot(x, z);
            double cs = 1.0, sn = 0.0;
            if (r != 0.0) {
                cs = x / r;
                sn = z / r;
            }
            if (k > lo) e[k - 1] = r;
            const double dk = d[k], dk1 = d[k + 1], ek = e[k];
            d[k] = cs * cs * dk + 2.0 * cs * sn * ek + sn * sn * dk1;
            d[k + 1] = sn * sn * dk - 2.0 * cs * sn * ek + cs * cs * dk1;
            e[k] = cs * sn * (dk1 - dk) + (cs * cs - sn * sn) * ek;
            if (k + 1 < hi) {
                z = sn * e[k + 1];
                e[k + 1] = cs * e[k + 1];
                x = e[k];
            }
            for (int rr = 0; rr < n; ++rr) {
                double *zr = Z + (size_t)rr * n;
                const double a0 = zr[k], a1 = zr[k + 1];
                zr[k] = cs * a0 + sn * a1;
                zr[k + 1] = -sn * a0 + cs * a1;
            }
        }
    }
    free(e);
    return 0;
}

/* rate_matrix.hpp:161-163,236-248: symmetrise D^1/2 Q D^-1/2, eigen, clamp. */
int or_mg94_decompose(int32_t S, const double *Q, const double *pi, double *evals, double *U_pi, double *V_pi) {
    double *sq = (double *)malloc(sizeof(double) * S * 2);
    double *isq = sq + S;
    for (int i = 0; i < S; ++i) {
        sq[i] = sqrt(pi[i] > 1e-12 ? pi[i] : 1e-12);
        isq[i] = 1.0 / sq[i];
    }
    double *Sm = (double *)malloc(sizeof(double) * S * S);
    double *Z = (double *)malloc(sizeof(double) * S * S);
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j) Sm[i * S + j] = sq[i] * Q[i * S + j] * isq[j];
    for (int i = 0; i < S; ++i)
        for (int j = i + 1; j < S; ++j) {
            double s = 0.5 * (Sm[i * S + j] + Sm[j * S + i]);
            Sm[i * S + j] = s;
            Sm[j * S + i] = s;
        }
    int rc = or_sym_eigen(S, Sm, evals, Z);
    for (int k = 0; k < S; ++k)
        if (evals[k] > 0.0) evals[k] = 0.0;
    /* U_pi = diag(1/sqrt pi) Z ; V_pi = Z^T diag(sqrt pi) */
    for (int i = 0; i < S; ++i)
        for (int k = 0; k < S; ++k) {
            U_pi[i * S + k] = isq[i] * Z[i * S + k];
            V_pi[k * S + i] = Z[i * S + k] * sq[i];
        }
    free(Z);
    free(Sm);
    free(sq);
    return rc;
}

/* rate_matrix.hpp:251-257 */
void or_transition(int32_t S, const double *evals, const double *U, const double *V, double t, double *P) {
    double ex[64];
    for (int k = 0; k < S; ++k) ex[k] = exp(evals[k] * t);
    for (int i = 0; i < S; ++i) {
        double *pr = P + i * S;
        for (int j = 0; j < S; ++j) pr[j] = 0.0;
        for (int k = 0; k < S; ++k) {
            const double uk = U[i * S + k] * ex[k];
            const double *vk = V + k * S;
            for (int j = 0; j < S; ++j) pr[j] += uk * vk[j];
        }
        for (int j = 0; j < S; ++j)
            if (pr[j] < 0.0) pr[j] = 0.0;
    }
}

/* Independent exp(t A) by scaling and squaring of a degree-30 Taylor series;
 * used only to pin the spectral P(t). */
int or_expm_taylor(int32_t n, const double *A, double t, double *out) {
    double norm = 0.0;
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += fabs(A[i * n + j] * t);
        if (s > norm) norm = s;
    }
    int sq = 0;
    while (norm > 0.25) {
        norm *= 0.5;
        ++sq;
    }
    const double scale = ldexp(t, -sq);
    double *X = (double *)malloc(sizeof(double) * n * n * 3);
    double *T = X + n * n, *tmp = X + 2 * n * n;
    for (int i = 0; i < n * n; ++i) X[i] = A[i] * scale;
    for (int i = 0; i < n * n; ++i) out[i] = 0.0;
    for (int i = 0; i < n; ++i) out[i * n + i] = 1.0;
    memcpy(T, out, sizeof(double) * n * n);
    for (int k = 1; k <= 30; ++k) {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double s = 0.0;
                for (int l = 0; l < n; ++l) s += T[i * n + l] * X[l * n + j];
                tmp[i * n + j] = s / k;
            }
        memcpy(T, tmp, sizeof(double) * n * n);
        for (int i = 0; i < n * n; ++i) out[i] += T[i];
    }
    for (int s = 0; s < sq; ++s) {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double acc = 0.0;
                for (int l = 0; l < n; ++l) acc += out[i * n + l] * out[l * n + j];
                tmp[i * n + j] = acc;
            }
        memcpy(out, tmp, sizeof(double) * n * n);
    }
    free(X);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Deterministic compensated summation (likelihood.hpp:34-51).               */
/* ------------------------------------------------------------------------- */
static inline void neumaier_acc(double *sum, double *c, double x) {
    double t = *sum + x;
    if (fabs(*sum) >= fabs(x))
        *c += (*sum - t) + x;
    else
        *c += (x - t) + *sum;
    *sum = t;
}

double or_neumaier_sum(const double *x, int64_t n) {
    double s = 0.0, c = 0.0;
    for (int64_t i = 0; i < n; ++i) neumaier_acc(&s, &c, x[i]);
    return s + c;
}

/* ------------------------------------------------------------------------- */
/* Tile workspaces and kernels.  Node matrices are [node][state][column].   */
/* ------------------------------------------------------------------------- */
#define SCALER_THRESH 1e-150
#define SCALER_UP 1e150
#define SCALER_DOWN 1e-150
#define LOG_SCALER 345.38776394910684

typedef struct {
    int S, B, N, maxc;
    double *D, *V, *PD, *U, *M;
    double *logsc;
    int *boost;
    double *pref, *suff;
    int *prefb, *suffb;
    double *L, *logL;
} ws_t;

static void ws_init(ws_t *w, const or_tree *tr, int B, int outside) {
    const int N = tr->num_nodes, S = tr->S;
    int maxc = 1;
    for (int n = 0; n < N; ++n) {
        int c = tr->child_ptr[n + 1] - tr->child_ptr[n];
        if (c > maxc) maxc = c;
    }
    w->S = S;
    w->B = B;
    w->N = N;
    w->maxc = maxc;
    const size_t nm = (size_t)N * S * B;
    w->D = (double *)calloc(nm, sizeof(double));
    w->M = (double *)calloc((size_t)S * B, sizeof(double));
    w->logsc = (double *)malloc(sizeof(double) * N * B);
    w->boost = (int *)malloc(sizeof(int) * N * B);
    w->L = (double *)malloc(sizeof(double) * B);
    w->logL = (double *)malloc(sizeof(double) * B);
    if (outside) {
        w->V = (double *)calloc(nm, sizeof(double));
        w->PD = (double *)calloc(nm, sizeof(double));
        w->U = (double *)calloc(nm, sizeof(double));
        w->pref = (double *)malloc(sizeof(double) * maxc * S * B);
        w->suff = (double *)malloc(sizeof(double) * maxc * S * B);
        w->prefb = (int *)malloc(sizeof(int) * maxc * B);
        w->suffb = (int *)malloc(sizeof(int) * maxc * B);
    } else {
        w->V = w->PD = w->U = w->pref = w->suff = NULL;
        w->prefb = w->suffb = NULL;
    }
}

static void ws_free(ws_t *w) {
    free(w->D);
    free(w->M);
    free(w->logsc);
    free(w->boost);
    free(w->L);
    free(w->logL);
    free(w->V);
    free(w->PD);
    free(w->U);
    free(w->pref);
    free(w->suff);
    free(w->prefb);
    free(w->suffb);
}

/* init_leaf_likelihood (likelihood.hpp:53-87) for one column. */
static void leaf_column(const or_tree *tr, int taxon, int pattern, double *D, int B, int b) {
    const int S = tr->S;
    if (taxon < 0) {
        for (int s = 0; s < S; ++s) D[s * B + b] = 1.0;
        return;
    }
    const int8_t st = tr->states[(size_t)pattern * tr->num_taxa + taxon];
    const uint64_t mask = tr->masks[(size_t)pattern * tr->num_taxa + taxon];
    if (st >= 0 && st < S) {
        for (int s = 0; s < S; ++s) D[s * B + b] = 0.0;
        D[st * B + b] = 1.0;
    } else if (mask != 0) {
        const uint64_t all = (S >= 64) ? ~(uint64_t)0 : (((uint64_t)1 << S) - 1);
        if ((mask & all) == all) {
            for (int s = 0; s < S; ++s) D[s * B + b] = 1.0;
        } else {
            for (int s = 0; s < S; ++s) D[s * B + b] = (mask & ((uint64_t)1 << s)) ? 1.0 : 0.0;
        }
    } else {
        for (int s = 0; s < S; ++s) D[s * B + b] = 1.0;
    }
}

/* C[S x B] = P[S x S] * X[S x B].  Four rows of C per pass share each row of
 * X (register blocking); every element still sums over k in ascending order,
 * so the result does not depend on the blocking. */
static void gemm_px(int S, int B, const double *P, const double *X, double *C) {
    int i = 0;
    for (; i + 4 <= S; i += 4) {
        double *c0 = C + (size_t)i * B, *c1 = c0 + B, *c2 = c1 + B, *c3 = c2 + B;
        for (int b = 0; b < B; ++b) c0[b] = c1[b] = c2[b] = c3[b] = 0.0;
        const double *p0 = P + (size_t)i * S, *p1 = p0 + S, *p2 = p1 + S, *p3 = p2 + S;
        for (int k = 0; k < S; ++k) {
            const double q0 = p0[k], q1 = p1[k], q2 = p2[k], q3 = p3[k];
            const double *x = X + (size_t)k * B;
            for (int b = 0; b < B; ++b) {
                const double xv = x[b];
                c0[b] += q0 * xv;
                c1[b] += q1 * xv;
                c2[b] += q2 * xv;
                c3[b] += q3 * xv;
            }
        }
    }
    for (; i < S; ++i) {
        double *c = C + (size_t)i * B;
        for (int b = 0; b < B; ++b) c[b] = 0.0;
        const double *pr = P + (size_t)i * S;
        for (int k = 0; k < S; ++k) {
            const double p = pr[k];
            const double *x = X + (size_t)k * B;
            for (int b = 0; b < B; ++b) c[b] += p * x[b];
        }
    }
}

/* C[S x B] = P^T X, four rows of C per pass (each element sums over i in
 * ascending order, as unblocked). */
static void gemm_ptx(int S, int B, const double *P, const double *X, double *C) {
    int j = 0;
    for (; j + 4 <= S; j += 4) {
        double *c0 = C + (size_t)j * B, *c1 = c0 + B, *c2 = c1 + B, *c3 = c2 + B;
        for (int b = 0; b < B; ++b) c0[b] = c1[b] = c2[b] = c3[b] = 0.0;
        for (int i = 0; i < S; ++i) {
            const double *pr = P + (size_t)i * S + j;
            const double q0 = pr[0], q1 = pr[1], q2 = pr[2], q3 = pr[3];
            const double *x = X + (size_t)i * B;
            for (int b = 0; b < B; ++b) {
                const double xv = x[b];
                c0[b] += q0 * xv;
                c1[b] += q1 * xv;
                c2[b] += q2 * xv;
                c3[b] += q3 * xv;
            }
        }
    }
    for (; j < S; ++j) {
        double *c = C + (size_t)j * B;
        for (int b = 0; b < B; ++b) c[b] = 0.0;
        for (int i = 0; i < S; ++i) {
            const double p = P[(size_t)i * S + j];
            const double *x = X + (size_t)i * B;
            for (int b = 0; b < B; ++b) c[b] += p * x[b];
        }
    }
}

static int boost_col(double *X, int S, int B, int b) {
    double mx = X[b];
    for (int s = 1; s < S; ++s)
        if (X[s * B + b] > mx) mx = X[s * B + b];
    int cnt = 0;
    while (mx < SCALER_THRESH && mx > 0.0) {
        for (int s = 0; s < S; ++s) X[s * B + b] *= SCALER_UP;
        mx *= SCALER_UP;
        ++cnt;
    }
    return cnt;
}

static inline double *nmat(double *base, const ws_t *w, int n) { return base + (size_t)n * w->S * w->B; }

/* compute_inside_batch_in_place (likelihood.hpp:343-476), bridge-free. */
static void inside_tile(const or_tree *tr, int p0, int p1, const double *P, const double *pi, ws_t *w) {
    const int S = tr->S, B = p1 - p0, WB = w->B;
    for (int oi = 0; oi < tr->num_nodes; ++oi) {
        const int nid = tr->post_order[oi];
        double *Dn = nmat(w->D, w, nid);
        if (tr->is_leaf[nid]) {
            for (int b = 0; b < B; ++b) {
                leaf_column(tr, tr->leaf_taxon[nid], p0 + b, Dn, WB, b);
                w->logsc[nid * WB + b] = 0.0;
            }
            continue;
        }
        for (int i = 0; i < S * WB; ++i) Dn[i] = 1.0;
        for (int b = 0; b < B; ++b) w->logsc[nid * WB + b] = 0.0;
        for (int ci = tr->child_ptr[nid]; ci < tr->child_ptr[nid + 1]; ++ci) {
            const int cid = tr->child_idx[ci];
            gemm_px(S, WB, P + (size_t)cid * S * S, nmat(w->D, w, cid), w->M);
            for (int i = 0; i < S * WB; ++i) Dn[i] *= w->M[i];
            for (int b = 0; b < B; ++b) {
                w->logsc[nid * WB + b] += w->logsc[cid * WB + b];
                int cnt = boost_col(Dn, S, WB, b);
                if (cnt > 0) w->logsc[nid * WB + b] -= cnt * LOG_SCALER;
            }
        }
        for (int b = 0; b < B; ++b) {
            int cnt = boost_col(Dn, S, WB, b);
            if (cnt > 0) w->logsc[nid * WB + b] -= cnt * LOG_SCALER;
        }
    }
    const double *Dr = nmat(w->D, w, tr->root);
    for (int b = 0; b < B; ++b) {
        double L = 0.0;
        for (int s = 0; s < S; ++s) L += Dr[s * WB + b] * pi[s];
        w->L[b] = L;
        w->logL[b] = (L > 0.0) ? log(L) + w->logsc[tr->root * WB + b] : -1e20;
    }
}

static void scale_col_pow(double *X, int S, int B, int b, int net) {
    while (net > 0) {
        for (int s = 0; s < S; ++s) X[s * B + b] *= SCALER_UP;
        --net;
    }
    while (net < 0) {
        for (int s = 0; s < S; ++s) X[s * B + b] *= SCALER_DOWN;
        ++net;
    }
}

/* compute_inside_outside_batch_in_place (likelihood.hpp:479-762), bridge-free. */
static void inside_outside_tile(const or_tree *tr, int p0, int p1, const double *P, const double *pi, ws_t *w) {
    const int S = tr->S, B = p1 - p0, WB = w->B;
    for (int oi = 0; oi < tr->num_nodes; ++oi) {
        const int nid = tr->post_order[oi];
        double *Dn = nmat(w->D, w, nid);
        if (tr->is_leaf[nid]) {
            for (int b = 0; b < B; ++b) {
                leaf_column(tr, tr->leaf_taxon[nid], p0 + b, Dn, WB, b);
                w->logsc[nid * WB + b] = 0.0;
                w->boost[nid * WB + b] = 0;
            }
            continue;
        }
        for (int ci = tr->child_ptr[nid]; ci < tr->child_ptr[nid + 1]; ++ci) {
            const int cid = tr->child_idx[ci];
            gemm_px(S, WB, P + (size_t)cid * S * S, nmat(w->D, w, cid), nmat(w->PD, w, cid));
        }
        for (int i = 0; i < S * WB; ++i) Dn[i] = 1.0;
        for (int b = 0; b < B; ++b) {
            w->logsc[nid * WB + b] = 0.0;
            w->boost[nid * WB + b] = 0;
        }
        for (int ci = tr->child_ptr[nid]; ci < tr->child_ptr[nid + 1]; ++ci) {
            const int cid = tr->child_idx[ci];
            const double *pd = nmat(w->PD, w, cid);
            for (int i = 0; i < S * WB; ++i) Dn[i] *= pd[i];
            for (int b = 0; b < B; ++b) {
                w->logsc[nid * WB + b] += w->logsc[cid * WB + b];
                w->boost[nid * WB + b] += boost_col(Dn, S, WB, b);
            }
        }
        for (int b = 0; b < B; ++b) {
            w->boost[nid * WB + b] += boost_col(Dn, S, WB, b);
            w->logsc[nid * WB + b] -= w->boost[nid * WB + b] * LOG_SCALER;
        }
    }
    const double *Dr = nmat(w->D, w, tr->root);
    for (int b = 0; b < B; ++b) {
        double L = 0.0;
        for (int s = 0; s < S; ++s) L += Dr[s * WB + b] * pi[s];
        w->L[b] = L;
        w->logL[b] = (L > 0.0) ? log(L) + w->logsc[tr->root * WB + b] : -1e20;
    }
    /* Pre-order pass */
    double *Ur = nmat(w->U, w, tr->root);
    for (int s = 0; s < S; ++s)
        for (int b = 0; b < WB; ++b) Ur[s * WB + b] = pi[s];
    for (int oi = tr->num_nodes - 1; oi >= 0; --oi) {
        const int uid = tr->post_order[oi];
        if (tr->is_leaf[uid]) continue;
        const int c0 = tr->child_ptr[uid], nc = tr->child_ptr[uid + 1] - c0;
        if (nc == 0) continue;
        double *pref = w->pref, *suff = w->suff;
        for (int i = 0; i < S * WB; ++i) pref[i] = 1.0;
        for (int b = 0; b < B; ++b) w->prefb[b] = 0;
        for (int i = 1; i < nc; ++i) {
            const double *pd = nmat(w->PD, w, tr->child_idx[c0 + i - 1]);
            double *cur = pref + (size_t)i * S * WB, *prev = pref + (size_t)(i - 1) * S * WB;
            for (int k = 0; k < S * WB; ++k) cur[k] = prev[k] * pd[k];
            for (int b = 0; b < B; ++b) w->prefb[i * WB + b] = w->prefb[(i - 1) * WB + b] + boost_col(cur, S, WB, b);
        }
        double *last = suff + (size_t)(nc - 1) * S * WB;
        for (int i = 0; i < S * WB; ++i) last[i] = 1.0;
        for (int b = 0; b < B; ++b) w->suffb[(nc - 1) * WB + b] = 0;
        for (int i = nc - 1; i > 0; --i) {
            const double *pd = nmat(w->PD, w, tr->child_idx[c0 + i]);
            double *cur = suff + (size_t)(i - 1) * S * WB, *nxt = suff + (size_t)i * S * WB;
            for (int k = 0; k < S * WB; ++k) cur[k] = nxt[k] * pd[k];
            for (int b = 0; b < B; ++b) w->suffb[(i - 1) * WB + b] = w->suffb[i * WB + b] + boost_col(cur, S, WB, b);
        }
        const double *Uu = nmat(w->U, w, uid);
        for (int i = 0; i < nc; ++i) {
            const int vid = tr->child_idx[c0 + i];
            double *Vv = nmat(w->V, w, vid);
            const double *pr = pref + (size_t)i * S * WB, *sf = suff + (size_t)i * S * WB;
            for (int k = 0; k < S * WB; ++k) Vv[k] = Uu[k] * pr[k] * sf[k];
            for (int b = 0; b < B; ++b) {
                int net = w->boost[uid * WB + b] - w->prefb[i * WB + b] - w->suffb[i * WB + b];
                scale_col_pow(Vv, S, WB, b, net);
            }
            gemm_ptx(S, WB, P + (size_t)vid * S * S, Vv, nmat(w->U, w, vid));
        }
    }
}

/* The workspace of compute_inside_batch_in_place (outside = 0) or
 * compute_inside_outside_batch_in_place (outside = 1) for patterns [p0, p1)
 * in one tile, copied out as [node][state][pattern] (D, V, PD, U; V/PD/U and
 * boost only when outside) and [node][pattern] (boost, logsc). */
int or_workspace(const or_tree *tr, const double *P, const double *pi, int32_t p0, int32_t p1, int32_t outside,
                 double *D, double *V, double *PD, double *U, int32_t *boost, double *logsc, double *L, double *logL) {
    const int B = p1 - p0, N = tr->num_nodes, S = tr->S;
    if (B <= 0 || p0 < 0 || p1 > tr->num_patterns) return -1;
    ws_t w;
    ws_init(&w, tr, B, outside);
    if (outside)
        inside_outside_tile(tr, p0, p1, P, pi, &w);
    else
        inside_tile(tr, p0, p1, P, pi, &w);
    const size_t nm = (size_t)N * S * B;
    memcpy(D, w.D, nm * sizeof(double));
    if (outside) {
        memcpy(V, w.V, nm * sizeof(double));
        memcpy(PD, w.PD, nm * sizeof(double));
        memcpy(U, w.U, nm * sizeof(double));
        memcpy(boost, w.boost, sizeof(int) * N * B);
    }
    memcpy(logsc, w.logsc, sizeof(double) * N * B);
    memcpy(L, w.L, sizeof(double) * B);
    memcpy(logL, w.logL, sizeof(double) * B);
    ws_free(&w);
    return 0;
}

static int clamp_threads(int threads, int work) {
    int t = threads > 0 ? threads : 1;
#ifdef _OPENMP
    int mx = omp_get_max_threads();
    if (threads <= 0) t = mx;
#endif
    if (work < 1) work = 1;
    if (t > work) t = work;
    return t < 1 ? 1 : t;
}

/* compute_tree_log_likelihood_from_P (likelihood.hpp:766-806). tile <= 0
 * selects the reference's 64-pattern tile. */
double or_loglik_from_P(const or_tree *tr, const double *P, const double *pi, int32_t tile, int32_t threads,
                        double *site_ll) {
    const int np = tr->num_patterns;
    if (np == 0) return 0.0;
    const int T = tile > 0 ? (tile < np ? tile : np) : (np < 64 ? np : 64);
    const int ntiles = (np + T - 1) / T;
    double *pll = (double *)calloc((size_t)np, sizeof(double));
    const int nt = clamp_threads(threads, ntiles);
#pragma omp parallel num_threads(nt)
    {
        ws_t w;
        ws_init(&w, tr, T, 0);
#pragma omp for schedule(static)
        for (int t = 0; t < ntiles; ++t) {
            const int p0 = t * T, p1 = (p0 + T < np) ? p0 + T : np;
            inside_tile(tr, p0, p1, P, pi, &w);
            for (int p = p0; p < p1; ++p) {
                pll[p] = tr->weights[p] * w.logL[p - p0];
                if (site_ll) site_ll[p] = w.logL[p - p0];
            }
        }
        ws_free(&w);
    }
    double s = or_neumaier_sum(pll, np);
    free(pll);
    return s;
}

/* compute_branch_length_gradients tile loop (likelihood.hpp:1150-1210) for
 * arbitrary per-branch derivative matrices dP. */
double or_branch_grad_from_P(const or_tree *tr, const double *P, const double *dP, const double *pi, int32_t tile,
                             int32_t threads, double *grad, double *site_ll) {
    const int np = tr->num_patterns, N = tr->num_nodes, S = tr->S;
    for (int n = 0; n < N; ++n) grad[n] = 0.0;
    if (np == 0) return 0.0;
    const int T = tile > 0 ? (tile < np ? tile : np) : (np < 64 ? np : 64);
    const int ntiles = (np + T - 1) / T;
    double *pll = (double *)calloc((size_t)np, sizeof(double));
    double *pg = (double *)calloc((size_t)N * np, sizeof(double));
    const int nt = clamp_threads(threads, ntiles);
#pragma omp parallel num_threads(nt)
    {
        ws_t w;
        ws_init(&w, tr, T, 1);
        double *dPD = (double *)malloc(sizeof(double) * S * T);
        double *invL = (double *)malloc(sizeof(double) * T);
#pragma omp for schedule(static)
        for (int t = 0; t < ntiles; ++t) {
            const int p0 = t * T, p1 = (p0 + T < np) ? p0 + T : np, B = p1 - p0;
            inside_outside_tile(tr, p0, p1, P, pi, &w);
            for (int b = 0; b < B; ++b) {
                const double L = w.L[b];
                if (L > 0.0) {
                    pll[p0 + b] = tr->weights[p0 + b] * w.logL[b];
                    invL[b] = tr->weights[p0 + b] / L;
                } else {
                    invL[b] = 0.0;
                }
                if (site_ll) site_ll[p0 + b] = w.logL[b];
            }
            for (int nid = 0; nid < N; ++nid) {
                if (nid == tr->root) continue;
                gemm_px(S, T, dP + (size_t)nid * S * S, nmat(w.D, &w, nid), dPD);
                const double *Vn = nmat(w.V, &w, nid);
                for (int b = 0; b < B; ++b) {
                    if (invL[b] > 0.0) {
                        double acc = 0.0;
                        for (int s = 0; s < S; ++s) acc += Vn[s * T + b] * dPD[s * T + b];
                        pg[(size_t)nid * np + p0 + b] = invL[b] * acc;
                    }
                }
            }
        }
        free(dPD);
        free(invL);
        ws_free(&w);
    }
    double total = or_neumaier_sum(pll, np);
    for (int n = 0; n < N; ++n) grad[n] = or_neumaier_sum(pg + (size_t)n * np, np);
    free(pg);
    free(pll);
    return total;
}

/* compute_mg94_log_likelihood (likelihood.hpp:257-292). */
double or_mg94_loglik(const or_tree *tr, const or_code *code, const or_mg94_params *p, const double *pos,
                      const double *pi, const double *bl, double branch_scale, int32_t threads, double *site_ll) {
    const int S = code->S, N = tr->num_nodes;
    double *Q = (double *)malloc(sizeof(double) * S * S * 3 + sizeof(double) * S);
    double *U = Q + S * S, *V = U + S * S, *ev = V + S * S;
    double mu;
    or_mg94_Q(code, p, pos, pi, Q, &mu);
    or_mg94_decompose(S, Q, pi, ev, U, V);
    double *P = (double *)malloc(sizeof(double) * (size_t)N * S * S);
    for (int n = 0; n < N; ++n) {
        double *Pn = P + (size_t)n * S * S;
        if (n == tr->root) {
            for (int i = 0; i < S * S; ++i) Pn[i] = 0.0;
            for (int i = 0; i < S; ++i) Pn[i * S + i] = 1.0;
        } else {
            or_transition(S, ev, U, V, bl[n] * branch_scale, Pn);
        }
    }
    double ll = or_loglik_from_P(tr, P, pi, 0, threads, site_ll);
    free(P);
    free(Q);
    return ll;
}

/* compute_branch_length_gradients (likelihood.hpp:1102-1211): raw times. */
double or_branch_length_gradients(const or_tree *tr, const or_code *code, const or_mg94_params *p,
                                  const double *pos, const double *pi, const double *bl, int32_t threads,
                                  double *grad) {
    const int S = code->S, N = tr->num_nodes;
    double *Q = (double *)malloc(sizeof(double) * S * S * 3 + sizeof(double) * S);
    double *U = Q + S * S, *V = U + S * S, *ev = V + S * S;
    double mu;
    or_mg94_Q(code, p, pos, pi, Q, &mu);
    or_mg94_decompose(S, Q, pi, ev, U, V);
    double *P = (double *)malloc(sizeof(double) * (size_t)N * S * S * 2);
    double *dP = P + (size_t)N * S * S;
    for (int n = 0; n < N; ++n) {
        double *Pn = P + (size_t)n * S * S, *dPn = dP + (size_t)n * S * S;
        if (n == tr->root) {
            for (int i = 0; i < S * S; ++i) {
                Pn[i] = 0.0;
                dPn[i] = 0.0;
            }
            for (int i = 0; i < S; ++i) Pn[i * S + i] = 1.0;
            continue;
        }
        or_transition(S, ev, U, V, bl[n], Pn);
        for (int i = 0; i < S; ++i)
            for (int j = 0; j < S; ++j) {
                double s = 0.0;
                for (int k = 0; k < S; ++k) s += Q[i * S + k] * Pn[k * S + j];
                dPn[i * S + j] = s;
            }
    }
    double ll = or_branch_grad_from_P(tr, P, dP, pi, 0, threads, grad, NULL);
    free(P);
    free(Q);
    return ll;
}

/* ------------------------------------------------------------------------- */
/* Branch-site mixture with full spectral-Frechet gradients.                 */
/* Forward/P_mix: busted.hpp:509-581, 1180-1216.  Branch gradients:          */
/* busted.hpp:1287-1327.  Parameter gradients: likelihood.hpp:1424-1591      */
/* generalised to per-component Frechet terms weighted by the mixture       */
/* weights.                                                                  */
/* ------------------------------------------------------------------------- */
static void build_dQ_omega(const or_code *code, const or_mg94_params *p, const double *pos, double *dQ) {
    /* likelihood.hpp:1529-1541 */
    const int S = code->S;
    double r[4][4];
    nuc_rates(p, r);
    for (int i = 0; i < S; ++i) {
        double row = 0.0;
        for (int j = 0; j < S; ++j) {
            double q = 0.0;
            const int idx = i * 64 + j;
            if (i != j && code->nd[idx] == 1 && !code->syn[idx])
                q = r[code->dfrom[idx * 3]][code->dto[idx * 3]] * pos[code->dpos[idx * 3] * 4 + code->dto[idx * 3]];
            dQ[i * S + j] = q;
            row += q;
        }
        dQ[i * S + i] = -row;
    }
}

static void build_dQ_theta(const or_code *code, const or_mg94_params *p, const double *pos, int u, int v,
                           double *dQ) {
    /* likelihood.hpp:1543-1562 */
    const int S = code->S;
    for (int i = 0; i < S; ++i) {
        double row = 0.0;
        for (int j = 0; j < S; ++j) {
            double q = 0.0;
            const int idx = i * 64 + j;
            if (i != j && code->nd[idx] == 1) {
                const int from = code->dfrom[idx * 3], to = code->dto[idx * 3];
                if ((from == u && to == v) || (from == v && to == u)) {
                    const double rm = code->syn[idx] ? p->alpha : p->beta;
                    q = rm * pos[code->dpos[idx * 3] * 4 + to];
                }
            }
            dQ[i * S + j] = q;
            row += q;
        }
        dQ[i * S + i] = -row;
    }
}

/* <V_pi dQ U_pi, H> (likelihood.hpp:1497-1506, part1) */
static double frechet_project(int S, const double *Vp, const double *dQ, const double *Up, const double *H) {
    double *T = (double *)malloc(sizeof(double) * S * S);
    /* T = dQ U */
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j) {
            double s = 0.0;
            for (int k = 0; k < S; ++k) s += dQ[i * S + k] * Up[k * S + j];
            T[i * S + j] = s;
        }
    double acc = 0.0;
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j) {
            double s = 0.0;
            for (int k = 0; k < S; ++k) s += Vp[i * S + k] * T[k * S + j];
            acc += s * H[i * S + j];
        }
    free(T);
    return acc;
}

static const int THETA_PAIRS[5][2] = {{0, 1}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

int or_mixture_eval(const or_tree *tr, const or_code *code, const or_mg94_params *base, const double *pos,
                    const double *pi, const or_mixture *mix, const double *bl, int32_t want_grad, int32_t tile,
                    int32_t threads, or_mixture_out *out) {
    const int S = code->S, N = tr->num_nodes, np = tr->num_patterns, NM = mix->n_mat, NC = mix->n_class;
    const int ncomp = mix->comp_ptr[NC];
    const size_t SS = (size_t)S * S;
    /* 1. Rate matrices and spectra (busted.hpp:530-548). */
    double *Q = (double *)malloc(sizeof(double) * SS * NM * 3);
    double *U = Q + SS * NM, *Vm = U + SS * NM;
    double *ev = (double *)malloc(sizeof(double) * S * NM);
    double *mu = (double *)malloc(sizeof(double) * NM);
    or_mg94_params *pm = (or_mg94_params *)malloc(sizeof(or_mg94_params) * NM);
    for (int m = 0; m < NM; ++m) {
        pm[m] = *base;
        pm[m].beta = mix->omegas[m];
        or_mg94_Q(code, &pm[m], pos, pi, Q + SS * m, &mu[m]);
        if (or_mg94_decompose(S, Q + SS * m, pi, ev + S * m, U + SS * m, Vm + SS * m) != 0) {
            free(Q); free(ev); free(mu); free(pm);
            return -2;
        }
    }
    /* Error sink (busted.hpp:193-206): only with raw time, whose branch
     * lengths do not depend on the components' rates. */
    double *Psink = NULL;
    for (int j = 0; j < ncomp; ++j)
        if (mix->comp_mat[j] < 0) {
            if (mix->comp_mat[j] != -1 || mix->convention != 0 || !mix->sink_freqs) {
                free(Q); free(ev); free(mu); free(pm);
                return -3;
            }
            if (!Psink) {
                Psink = (double *)calloc(SS, sizeof(double));
                for (int i = 0; i < S; ++i) {
                    const double d = 1.0 - mix->sink_freqs[i];
                    const double inv = 1.0 / (d > 1e-12 ? d : 1e-12);
                    for (int k = 0; k < S; ++k)
                        if (k != i) Psink[(size_t)i * S + k] = mix->sink_freqs[k] * inv;
                }
            }
        }
    /* 2. Per-class time conversion. */
    double *mubar = (double *)malloc(sizeof(double) * NC * 2);
    double *conv = mubar + NC;
    for (int c = 0; c < NC; ++c) {
        double mb = 0.0;
        for (int j = mix->comp_ptr[c]; j < mix->comp_ptr[c + 1]; ++j)
            if (mix->comp_mat[j] >= 0) mb += mix->comp_w[j] * mu[mix->comp_mat[j]];
        mubar[c] = mb > 1e-12 ? mb : 1e-12;
        conv[c] = mix->convention == 1 ? 3.0 / mubar[c] : 1.0;
    }
    /* 3. Mixed P and dP per branch (busted.hpp:1191-1216). */
    double *P = (double *)calloc(SS * N, sizeof(double));
    double *dP = (double *)calloc(SS * N, sizeof(double));
    double *tau = (double *)calloc((size_t)N, sizeof(double));
    double *Pk = (double *)malloc(sizeof(double) * SS);
    for (int n = 0; n < N; ++n) {
        double *Pn = P + SS * n, *dPn = dP + SS * n;
        if (n == tr->root) {
            for (int i = 0; i < S; ++i) Pn[i * S + i] = 1.0;
            continue;
        }
        const int c = mix->branch_class[n];
        tau[n] = bl[n] * conv[c];
        for (int j = mix->comp_ptr[c]; j < mix->comp_ptr[c + 1]; ++j) {
            const int m = mix->comp_mat[j];
            const double w = mix->comp_w[j];
            if (m < 0) { /* the sink: w * P_sink, no time derivative (busted.hpp:575-576) */
                for (size_t i = 0; i < SS; ++i) Pn[i] += w * Psink[i];
                continue;
            }
            or_transition(S, ev + S * m, U + SS * m, Vm + SS * m, tau[n], Pk);
            for (size_t i = 0; i < SS; ++i) Pn[i] += w * Pk[i];
            const double *Qm = Q + SS * m;
            for (int i = 0; i < S; ++i)
                for (int jj = 0; jj < S; ++jj) {
                    double s = 0.0;
                    for (int k = 0; k < S; ++k) s += Qm[i * S + k] * Pk[k * S + jj];
                    dPn[i * S + jj] += w * s;
                }
        }
    }
    free(Pk);
    const int T = tile > 0 ? (tile < np ? tile : np) : (np < 64 ? (np > 0 ? np : 1) : 64);
    const int ntiles = np > 0 ? (np + T - 1) / T : 0;
    double *pll = (double *)calloc((size_t)(np > 0 ? np : 1), sizeof(double));
    int rc = 0;
    if (!want_grad) {
        out->log_likelihood = or_loglik_from_P(tr, P, pi, T, threads, out->site_ll);
        goto done;
    }
    {
        double *pg = (double *)calloc((size_t)N * (np > 0 ? np : 1), sizeof(double));
        const int nt = clamp_threads(threads, ntiles);
        double *Athr = (double *)calloc((size_t)nt * N * SS, sizeof(double));
#pragma omp parallel num_threads(nt)
        {
            int tid = 0;
#ifdef _OPENMP
            tid = omp_get_thread_num();
#endif
            double *A = Athr + (size_t)tid * N * SS;
            ws_t w;
            ws_init(&w, tr, T, 1);
            double *dPD = (double *)malloc(sizeof(double) * S * T);
            double *invL = (double *)malloc(sizeof(double) * T);
            double *cV = (double *)malloc(sizeof(double) * S * T);
#pragma omp for schedule(static)
            for (int t = 0; t < ntiles; ++t) {
                const int p0 = t * T, p1 = (p0 + T < np) ? p0 + T : np, B = p1 - p0;
                inside_outside_tile(tr, p0, p1, P, pi, &w);
                for (int b = 0; b < T; ++b) invL[b] = 0.0;
                for (int b = 0; b < B; ++b) {
                    const double L = w.L[b];
                    if (L > 0.0) {
                        pll[p0 + b] = tr->weights[p0 + b] * w.logL[b];
                        invL[b] = tr->weights[p0 + b] / L;
                    }
                    if (out->site_ll) out->site_ll[p0 + b] = w.logL[b];
                }
                for (int nid = 0; nid < N; ++nid) {
                    if (nid == tr->root) continue;
                    const double *Dn = nmat(w.D, &w, nid), *Vn = nmat(w.V, &w, nid);
                    gemm_px(S, T, dP + SS * nid, Dn, dPD);
                    for (int b = 0; b < B; ++b) {
                        if (invL[b] > 0.0) {
                            double acc = 0.0;
                            for (int s = 0; s < S; ++s) acc += Vn[s * T + b] * dPD[s * T + b];
                            pg[(size_t)nid * np + p0 + b] = invL[b] * acc;
                        }
                    }
                    /* A_b += (w/L) V D^T  (likelihood.hpp:1362-1366,1435) */
                    for (int s = 0; s < S; ++s)
                        for (int b = 0; b < T; ++b) cV[s * T + b] = (b < B) ? Vn[s * T + b] * invL[b] : 0.0;
                    double *An = A + SS * nid;
                    for (int i = 0; i < S; ++i) {
                        const double *vi = cV + (size_t)i * T;
                        for (int j = 0; j < S; ++j) {
                            const double *dj = Dn + (size_t)j * T;
                            double acc = 0.0;
                            for (int b = 0; b < B; ++b) acc += vi[b] * dj[b];
                            An[i * S + j] += acc;
                        }
                    }
                }
            }
            free(dPD);
            free(invL);
            free(cV);
            ws_free(&w);
        }
        /* Reduce thread partials in thread order. */
        double *A = Athr;
        for (int th = 1; th < nt; ++th) {
            const double *Ap = Athr + (size_t)th * N * SS;
            for (size_t i = 0; i < (size_t)N * SS; ++i) A[i] += Ap[i];
        }
        double *dtau = (double *)calloc((size_t)N, sizeof(double));
        for (int n = 0; n < N; ++n)
            if (n != tr->root) dtau[n] = or_neumaier_sum(pg + (size_t)n * np, np);
        free(pg);
        /* Frechet accumulation (likelihood.hpp:1440-1471). */
        double *H = (double *)calloc(SS * NM, sizeof(double));
        double *sum_tau = (double *)calloc((size_t)NC, sizeof(double));
        double *At = (double *)malloc(sizeof(double) * SS);
        double *tmp = (double *)malloc(sizeof(double) * SS);
        double *Pm = (double *)malloc(sizeof(double) * SS);
        double *gw = (double *)calloc((size_t)(ncomp > 0 ? ncomp : 1), sizeof(double));
        for (int n = 0; n < N; ++n) {
            if (n == tr->root || tr->parent[n] < 0) continue;
            const int c = mix->branch_class[n];
            const double tb = tau[n];
            sum_tau[c] += tb * dtau[n];
            const double *An = A + SS * n;
            for (int j = mix->comp_ptr[c]; j < mix->comp_ptr[c + 1]; ++j) {
                const int m = mix->comp_mat[j];
                const double wj = mix->comp_w[j];
                if (m < 0) { /* d/dw_sink = <A_b, P_sink> */
                    double s_w = 0.0;
                    for (size_t i = 0; i < SS; ++i) s_w += An[i] * Psink[i];
                    gw[j] += s_w;
                    continue;
                }
                const double *Um = U + SS * m, *Vp = Vm + SS * m, *lam = ev + S * m;
                /* d/dw_j = <A_b, P_m(tau_b)> */
                or_transition(S, lam, Um, Vp, tb, Pm);
                double s_w = 0.0;
                for (size_t i = 0; i < SS; ++i) s_w += An[i] * Pm[i];
                gw[j] += s_w;
                /* A_tilde = U^T A V^T */
                for (int i = 0; i < S; ++i)
                    for (int jj = 0; jj < S; ++jj) {
                        double s = 0.0;
                        for (int k = 0; k < S; ++k) s += An[i * S + k] * Vp[jj * S + k];
                        tmp[i * S + jj] = s;
                    }
                for (int i = 0; i < S; ++i)
                    for (int jj = 0; jj < S; ++jj) {
                        double s = 0.0;
                        for (int k = 0; k < S; ++k) s += Um[k * S + i] * tmp[k * S + jj];
                        At[i * S + jj] = s;
                    }
                double *Hm = H + SS * m;
                for (int i = 0; i < S; ++i) {
                    const double ei = exp(lam[i] * tb);
                    for (int jj = 0; jj < S; ++jj) {
                        const double diff = lam[i] - lam[jj];
                        double E;
                        if (fabs(diff) < 1e-9)
                            E = tb * ei;
                        else
                            E = (ei - exp(lam[jj] * tb)) / diff;
                        Hm[i * S + jj] += wj * At[i * S + jj] * E;
                    }
                }
            }
        }
        /* Parameter projections (likelihood.hpp:1497-1591). */
        double *dQ = (double *)malloc(sizeof(double) * SS);
        double *gom = (double *)calloc((size_t)NM, sizeof(double));
        double gth[5] = {0, 0, 0, 0, 0};
        /* dmu/dphi per matrix */
        double *dmu_om = (double *)calloc((size_t)NM, sizeof(double));
        double *dmu_th = (double *)calloc((size_t)NM * 5, sizeof(double));
        for (int m = 0; m < NM; ++m) {
            const double *Um = U + SS * m, *Vp = Vm + SS * m, *Hm = H + SS * m;
            build_dQ_omega(code, &pm[m], pos, dQ);
            gom[m] = frechet_project(S, Vp, dQ, Um, Hm);
            double dm = 0.0;
            for (int i = 0; i < S; ++i) dm -= pi[i] * dQ[i * S + i];
            dmu_om[m] = dm;
            for (int k = 0; k < 5; ++k) {
                build_dQ_theta(code, &pm[m], pos, THETA_PAIRS[k][0], THETA_PAIRS[k][1], dQ);
                gth[k] += frechet_project(S, Vp, dQ, Um, Hm);
                double d2 = 0.0;
                for (int i = 0; i < S; ++i) d2 -= pi[i] * dQ[i * S + i];
                dmu_th[m * 5 + k] = d2;
            }
        }
        if (mix->convention == 1) {
            /* -(dmubar/mubar) * sum_b tau_b dlnL/dtau_b per class (likelihood.hpp:1505-1506) */
            for (int c = 0; c < NC; ++c) {
                const double f = sum_tau[c] / mubar[c];
                for (int j = mix->comp_ptr[c]; j < mix->comp_ptr[c + 1]; ++j) {
                    const int m = mix->comp_mat[j];
                    const double wj = mix->comp_w[j];
                    gom[m] -= f * wj * dmu_om[m];
                    for (int k = 0; k < 5; ++k) gth[k] -= f * wj * dmu_th[m * 5 + k];
                    gw[j] -= f * mu[m];
                }
            }
        }
        out->log_likelihood = or_neumaier_sum(pll, np);
        if (out->grad_bl)
            for (int n = 0; n < N; ++n) out->grad_bl[n] = (n == tr->root) ? 0.0 : conv[mix->branch_class[n]] * dtau[n];
        if (out->dtau)
            for (int n = 0; n < N; ++n) out->dtau[n] = dtau[n];
        if (out->grad_omega)
            for (int m = 0; m < NM; ++m) out->grad_omega[m] = gom[m];
        if (out->grad_w)
            for (int j = 0; j < ncomp; ++j) out->grad_w[j] = gw[j];
        if (out->grad_theta)
            for (int k = 0; k < 5; ++k) out->grad_theta[k] = gth[k];
        free(dQ);
        free(gom);
        free(dmu_om);
        free(dmu_th);
        free(gw);
        free(Pm);
        free(tmp);
        free(At);
        free(sum_tau);
        free(H);
        free(dtau);
        free(Athr);
    }
done:
    free(Psink);
    free(pll);
    free(tau);
    free(dP);
    free(P);
    free(mubar);
    free(pm);
    free(mu);
    free(ev);
    free(Q);
    return rc;
}

/* BUSTED synonymous-rate classes, BUSTEDAnalyzer::evaluate_log_l
 * (busted.hpp:554-581 P_mix[m] at syn_rates[m] * tau, 607-649 per-pattern
 * log-sum-exp with log syn_weights, classes with weight <= 1e-12 skipped,
 * 652-656 Neumaier over w_p * pll_p) and compute_busted_branch_gradients
 * (busted.hpp:1191-1216, 1335-1396: per-pattern class posteriors post_m(p),
 * gradient terms w_p * post_m(p) * dL_m/L_m, post <= 1e-12 dropped).  The
 * class-m terms form a branch-site mixture evaluation whose pattern weights are
 * w_p * post_m(p), so each class is one or_mixture_eval call at branch lengths
 * syn_rates[m] * bl.  Beyond the reference's branch gradients this also
 * returns omega / weight / theta gradients (summed over classes) and
 * d lnL / d syn_rates[m] = sum_b bl_b dlnL_m/d(bl_m,b),
 * d lnL / d syn_weights[m] = sum_p w_p post_m(p) / v_m. */
int or_srv_eval(const or_tree *tr, const or_code *code, const or_mg94_params *base, const double *pos,
                const double *pi, const or_mixture *mix, const double *bl, int32_t n_syn, const double *syn_rates,
                const double *syn_weights, int32_t want_grad, int32_t threads, or_mixture_out *out,
                double *grad_syn_rates, double *grad_syn_weights) {
    const int N = tr->num_nodes, np = tr->num_patterns, M = n_syn, NM = mix->n_mat;
    const int ncomp = mix->comp_ptr[mix->n_class];
    if (M < 1) return -1;
    double *blm = (double *)malloc(sizeof(double) * N);
    double *sll = (double *)malloc(sizeof(double) * (size_t)M * (np > 0 ? np : 1));
    double *post = (double *)calloc((size_t)M * (np > 0 ? np : 1), sizeof(double));
    double *pll = (double *)malloc(sizeof(double) * (np > 0 ? np : 1));
    double *wp = (double *)malloc(sizeof(double) * (np > 0 ? np : 1));
    int rc = 0;
    for (size_t i = 0; i < (size_t)M * np; ++i) sll[i] = -1e20;
    for (int m = 0; m < M && rc == 0; ++m) {
        if (!(syn_weights[m] > 1e-12)) continue;
        for (int n = 0; n < N; ++n) blm[n] = syn_rates[m] * bl[n];
        or_mixture_out o = {0};
        o.site_ll = sll + (size_t)m * np;
        rc = or_mixture_eval(tr, code, base, pos, pi, mix, blm, 0, 0, threads, &o);
    }
    if (rc == 0) {
        for (int p = 0; p < np; ++p) {
            double mx = -1e20;
            pll[p] = -1e20;
            for (int m = 0; m < M; ++m) {
                const double x = sll[(size_t)m * np + p];
                if (syn_weights[m] > 1e-12 && x > -1e19 && log(syn_weights[m]) + x > mx) mx = log(syn_weights[m]) + x;
            }
            if (mx <= -1e19) continue;
            double se = 0.0;
            for (int m = 0; m < M; ++m) {
                const double x = sll[(size_t)m * np + p];
                if (syn_weights[m] > 1e-12 && x > -1e19) {
                    post[(size_t)m * np + p] = exp(log(syn_weights[m]) + x - mx);
                    se += post[(size_t)m * np + p];
                }
            }
            if (se > 0.0) {
                pll[p] = mx + log(se);
                for (int m = 0; m < M; ++m) post[(size_t)m * np + p] /= se;
            } else {
                for (int m = 0; m < M; ++m) post[(size_t)m * np + p] = 0.0;
            }
        }
        for (int p = 0; p < np; ++p) wp[p] = tr->weights[p] * pll[p];
        out->log_likelihood = or_neumaier_sum(wp, np);
        if (out->site_ll)
            for (int p = 0; p < np; ++p) out->site_ll[p] = pll[p];
    }
    if (rc == 0 && want_grad) {
        double *gbl = (double *)calloc((size_t)N, sizeof(double)), *tb = (double *)calloc((size_t)N, sizeof(double));
        double *gom = (double *)calloc((size_t)NM, sizeof(double)), *to = (double *)calloc((size_t)NM, sizeof(double));
        double *gw = (double *)calloc((size_t)(ncomp > 0 ? ncomp : 1), sizeof(double));
        double *tw = (double *)calloc((size_t)(ncomp > 0 ? ncomp : 1), sizeof(double));
        double gth[5] = {0, 0, 0, 0, 0}, tt[5];
        for (int m = 0; m < M && rc == 0; ++m) {
            if (grad_syn_rates) grad_syn_rates[m] = 0.0;
            if (grad_syn_weights) grad_syn_weights[m] = 0.0;
            if (!(syn_weights[m] > 1e-12)) continue;
            double sw = 0.0;
            for (int p = 0; p < np; ++p) {
                const double q = post[(size_t)m * np + p];
                wp[p] = q > 1e-12 ? tr->weights[p] * q : 0.0;
                sw += tr->weights[p] * q;
            }
            if (grad_syn_weights) grad_syn_weights[m] = sw / syn_weights[m];
            or_tree trm = *tr;
            trm.weights = wp;
            for (int n = 0; n < N; ++n) blm[n] = syn_rates[m] * bl[n];
            or_mixture_out o = {0};
            o.grad_bl = tb;
            o.grad_omega = to;
            o.grad_w = tw;
            o.grad_theta = tt;
            rc = or_mixture_eval(&trm, code, base, pos, pi, mix, blm, 1, 0, threads, &o);
            if (rc != 0) break;
            double gr = 0.0;
            for (int n = 0; n < N; ++n) {
                gbl[n] += syn_rates[m] * tb[n];
                gr += bl[n] * tb[n];
            }
            if (grad_syn_rates) grad_syn_rates[m] = gr;
            for (int k = 0; k < NM; ++k) gom[k] += to[k];
            for (int j = 0; j < ncomp; ++j) gw[j] += tw[j];
            for (int k = 0; k < 5; ++k) gth[k] += tt[k];
        }
        if (rc == 0) {
            if (out->grad_bl) memcpy(out->grad_bl, gbl, sizeof(double) * N);
            if (out->grad_omega) memcpy(out->grad_omega, gom, sizeof(double) * NM);
            if (out->grad_w) memcpy(out->grad_w, gw, sizeof(double) * ncomp);
            if (out->grad_theta) memcpy(out->grad_theta, gth, sizeof(gth));
        }
        free(gbl);
        free(tb);
        free(gom);
        free(to);
        free(gw);
        free(tw);
    }
    free(blm);
    free(sll);
    free(post);
    free(pll);
    free(wp);
    return rc;
}

/* compute_joint_frechet_gradients (likelihood.hpp:1228-1594): one component
 * per class, 3/mu time convention, optional test/background partition. */
double or_joint_frechet(const or_tree *tr, const or_code *code, const or_mg94_params *p, const double *pos,
                        const double *pi, const double *bl, const uint8_t *is_test, double omega_bg, int32_t threads,
                        double *grad_bl, double *grad_omega, double *grad_omega_bg, double *grad_theta) {
    const int N = tr->num_nodes;
    const int has_part = (is_test != NULL && omega_bg >= 0.0);
    double omegas[2] = {p->beta, omega_bg};
    int32_t *bc = (int32_t *)calloc((size_t)N, sizeof(int32_t));
    for (int n = 0; n < N; ++n) bc[n] = (has_part && !is_test[n]) ? 1 : 0;
    int32_t comp_ptr[3] = {0, 1, 2};
    int32_t comp_mat[2] = {0, 1};
    double comp_w[2] = {1.0, 1.0};
    or_mixture mix;
    mix.n_mat = has_part ? 2 : 1;
    mix.n_class = has_part ? 2 : 1;
    mix.convention = 1;
    mix.pad0 = 0;
    mix.omegas = omegas;
    mix.branch_class = bc;
    mix.comp_ptr = comp_ptr;
    mix.comp_mat = comp_mat;
    mix.comp_w = comp_w;
    double gom[2] = {0, 0};
    or_mixture_out out;
    memset(&out, 0, sizeof(out));
    out.grad_bl = grad_bl;
    out.grad_omega = gom;
    out.grad_theta = grad_theta;
    or_mixture_eval(tr, code, p, pos, pi, &mix, bl, 1, 0, threads, &out);
    if (grad_omega) *grad_omega = gom[0];
    if (grad_omega_bg) *grad_omega_bg = has_part ? gom[1] : 0.0;
    free(bc);
    return out.log_likelihood;
}
