// Internal interface between the host engine (hy_engine.cpp) and the HIP
// kernels (hy_kernels.hip).  Not part of the public C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace hy {

// Child / branch reference encoding: (kind << 28) | index.
// Artificial nodes that resolve polytomies are internal nodes whose edge
// matrix is an exact identity (uploaded once per problem).
constexpr int KIND_INT = 0;    // internal node, index = D slot, edge matrix P(node)
constexpr int KIND_TIP = 1;    // leaf, index = tip row, edge matrix P(node)
constexpr int REF_MASK = 0x0FFFFFFF;
constexpr uint16_t TIP_ONES = 0xFFFFu;  // all-ones partial (gap, unmatched leaf, padding)
constexpr uint16_t TIP_AMB = 0x8000u;   // | index into the ambiguity-mask table

// Inside schedule entry (int32 x SCHED_STRIDE), one per internal node in a
// post-order that visits larger subtrees first: [0] node, [1] global D index,
// [2] #children (<= 3), [3] LDS slot receiving the node's partials (-1: HBM
// only), [4..5] unused, then per child c: [6+3c] node id of the child
// (selects its edge matrix), [7+3c] child reference (kind << 28 | tip row or
// global D index), [8+3c] LDS slot holding the child's partials (-1: reload
// from HBM).
constexpr int SCHED_STRIDE = 16;
constexpr int MAXC = 3;

// Outside schedule entry (int32 x OSCHED_STRIDE), reverse of the inside
// order: [0] node, [1] global D index, [2] #children, [3] V slot of the node
// (-1 at the root), [4] 1 if the edge above the node is an identity edge,
// [5] source of the node's U message (U_ROOT, U_SPILL or an LDS slot), then per
// child c: [6+4c] node id, [7+4c] child reference, [8+4c] V slot of the
// child, [9+4c] destination of the child's U (LDS slot, U_SPILL, or -1 for tips).
constexpr int OSCHED_STRIDE = 20;
constexpr int U_ROOT = -3;
constexpr int U_SPILL = -2;
constexpr int U_NEXT = -4;  // message handed to the next entry in registers

// LDS slot capacities of the pruning kernels for tile width T and element
// size esz (bytes of partials per slot = 64 * T * esz); each kernel stays
// within the 160 KiB of LDS a gfx950 workgroup may use.
// Inside partials in HBM are stored "B-operand-major": element (state s,
// pattern t) of a 64 x T tile sits at ((t/16*4 + s%4)*16 + t%16)*16 + s/4, so
// that the 16 MFMA B operands of one lane are contiguous.
constexpr int bmaj_index(int s, int t) { return (((t >> 4) * 4 + (s & 3)) * 16 + (t & 15)) * 16 + (s >> 2); }

constexpr int inside_slots(int T, int esz, bool lean = false) { return lean ? 6 : (T == 16 ? 12 : (T == 32 ? 6 : 3)); }
// "Lean" inside pass (FP32, T = 16, problems with many more tiles than CUs):
// six LDS slots, tip codes read from HBM and no register double buffering,
// so that four workgroups share a CU; throughput instead of latency.
constexpr bool lean_inside(int prec_is_fp32, int T, long n_tiles, int n_cu) {
    return prec_is_fp32 && T == 16 && n_tiles > 2L * n_cu;
}
// "Lean" outside pass (FP32, T = 16, subtree parts): four LDS slots for
// pending messages, tip codes read from HBM and the next node's operands
// requested after the current node's products (no register double
// buffering), so that four workgroups share a CU.
constexpr int outside_slots(int T, int esz, bool lean = false) {
    return lean ? 4 : (T == 16 ? 10 : (T == 32 ? 4 : (esz == 8 ? 1 : 2)));
}
// The top part of a split tree runs the default variants against schedules
// built for the lean ones: their LDS slot counts must cover the lean ones.
static_assert(inside_slots(16, 4, true) <= inside_slots(16, 4, false), "lean inside slots");
static_assert(outside_slots(16, 4, true) <= outside_slots(16, 4, false), "lean outside slots");
// Tip codes of a tile kept in LDS when (taxa + 1) * T fits.
constexpr int tip_lds_codes(int T, int esz) {
    return T == 16 ? 12288 : (esz == 4 ? 24576 : (T == 32 ? 16384 : 8192));
}

// Row-major transition matrices: per node a block of PRS elements, the 64x64
// matrix followed by its 64 row sums (tip gathers then address both from one
// base).
constexpr int PRS = 4096 + 64;

// Branch list entry: [0] node, [1] outside slot, [2] child reference (kind | index),
// [3] mixture class, [4] gene (batched evaluations; 0 otherwise), [5..7] unused.
constexpr int BR_STRIDE = 8;

// One problem (gene) of a fused batched evaluation: its slice of every
// batch-wide array.  Single-problem evaluations pass no table.
struct GeneInfo {
    int tile0, ntiles;            // pruning tiles [tile0, tile0 + ntiles)
    int pbase, P_pad, P;          // pattern offset, padded and real pattern count
    int sched_off, n_sched;       // inside schedule entries
    int osched_off, n_osched;     // outside schedule entries
    int tip_rows, tiles_per_chunk;
    int br0, n_br;                // branches [br0, br0 + n_br) of the branch list
    int node0, n_real_nodes;      // first extended node; nodes of the caller's tree
    int mat0, n_mat, comp0, n_comp, cls0, n_cls;
    int res_off, pad0;
    long long tip_base, d_base, e_base, v_base, g_base, u_base;  // element offsets
};

// Per-evaluation parameter block (device, FP64), offsets in doubles.
struct ParamLayout {
    int n_nodes;   // extended node count
    int n_mat, n_comp;
    int off_bl, off_alpha, off_omega, off_w, off_theta, off_delta, off_psi;
    int off_seq;  // call number of the evaluation (k_final publishes it when done)
    int total;
};

// Packed description of codon pair (i, j), one int32 per entry of the 64 x 64
// generator: bits 0-1 number of differing positions (0 = none or not a
// sense pair), bit 2 synonymous, then per difference k (k = 0..2) 7 bits at
// 3 + 7k: exchangeability index (0..4 = AC, AT, CG, CT, GT; 5 = AG) and
// 4 bits position-frequency index (position * 4 + target nucleotide).
constexpr int QDESC_SYN = 4;
constexpr int qdesc_theta_index(int a, int b) {
    return (a > b) ? qdesc_theta_index(b, a)
                   : (a == 0 && b == 1) ? 0 : (a == 0 && b == 3) ? 1 : (a == 1 && b == 2) ? 2
                   : (a == 1 && b == 3) ? 3 : (a == 2 && b == 3) ? 4 : 5;
}

struct ModelArgs {
    const int32_t *qdesc;                   // [64*64] packed pair descriptors
    const int8_t *nd, *dpos, *dfrom, *dto;  // [64*64], [64*64*3] x3
    const uint8_t *syn;                     // [64*64]
    const double *pos;                      // [12]
    const double *pi;                       // [64], zero beyond S
    int S;
    const double *params;
    ParamLayout pl;
    double *lam;  // [n_mat][64]
    double *U;    // [n_mat][64][64]
    double *V;    // [n_mat][64][64]
    double *mu;   // [n_mat]
    double *Z;    // [n_mat][64][64] eigenvectors of the symmetrised generator (warm-start state)
    int warm;     // 1: refine Z from the previous call
    int *status;  // [n_mat] 100+k: k refinement steps of a warm start (100: kept); 50+k: k steps after a
                  // tridiagonal solve; 1..40: Jacobi sweeps; negative = no convergence
    const int32_t *mat_gene;  // batched evaluation: gene of each matrix (pi, pos, alpha, theta, delta, psi)
    double tol;               // warm refinement: accepted predicted correction (relative)
    // When set: the parameter block is read from pinned host memory and
    // copied to params_dev (grid-strided over the workgroups) for the
    // kernels that follow, replacing a host-to-device copy.
    const double *params_host;
    double *params_dev;
    int n_params;
    // Slot addressing (spectrum pools): matrix m writes lam, U, V, mu, status
    // and its eigenvectors Z to slot oslot[m] (null: slot m) and, when warm,
    // refines the eigenvectors held in slot zsrc[m] (null: its own slot;
    // negative: a cold solve).  The source slot is never an output slot of
    // the same launch.
    const int32_t *oslot;
    const int32_t *zsrc;
};

// Per-site refits (FEL): candidate k is pattern pat[k] at (alpha[k], beta[k]);
// its spectrum (MG94 with omega = beta / max(alpha, 1e-12)) is slot spec[k]
// of the spectrum pool, branch times max(alpha, 1e-12) * t_b.  One workgroup
// per candidate walks the pattern-walk schedule (GTR_STRIDE entries).
struct FelArgs {
    const int32_t *sched;
    int n_sched;
    const int32_t *nref;     // [n_ext] child references (kind << 28 | tip row or D slot)
    const uint8_t *ident;    // [n_ext] identity edge
    const uint16_t *tips;    // [taxa + 1][tip_stride] codon codes
    int tip_stride;
    const uint64_t *amb;
    const int32_t *pat;      // [n_cand]
    const double *alpha;     // [n_cand]
    const int32_t *spec;     // [n_cand] spectrum slot
    const double *lam, *U, *V;  // spectrum pool ([slot][64], [slot][64][64])
    const double *mu;        // [slot] generator mean rate
    const int *status;       // [slot] eigensolver status
    const double *bl;        // [n_ext] branch lengths (model time at alpha = 1)
    const double *pi;        // [64]
    int S, n_int, n_cand;
    double *D;               // [n_cand][n_int][64] partials
    int32_t *E;              // [n_cand][n_int]
    double *out;             // [3][n_cand]: pattern lnL, mean rate, eigensolver status
};
void launch_fel(const FelArgs &a, int n_cand, hipStream_t s);

// MEME's per-site candidates (hy_meme_eval): test branches mix the spectra
// A (weight p1) and B (1 - p1), background branches use C.
struct MemeArgs {
    const int32_t *sched;
    int n_sched;
    const int32_t *nref, *pat;
    const uint8_t *ident, *is_test;  // [n_ext]
    const uint16_t *tips;
    int tip_stride;
    const uint64_t *amb;
    const double *alpha, *p1;        // [n_cand] (alpha already >= 1e-12)
    const int32_t *specA, *specB, *specC;
    const int32_t *cond;             // [n_cand] -1, or 2 node + class: that branch takes class 0 (A) or 1 (B) alone
    const double *lam, *Z;           // spectrum pool: [slot][64], [slot][64][64] eigenvectors of the symmetrised generator
    const double *bl, *pi;
    int S, n_int;
    double *D;
    int32_t *E;
    double *out;                     // [n_cand] pattern lnL
};
void launch_meme(const MemeArgs &a, int n_cand, hipStream_t s);

template <typename R>
struct PmatArgs {
    const int32_t *blist;  // [n_branch][BR_STRIDE]
    int n_branch;
    const int32_t *comp_ptr, *comp_mat;
    int convention;
    const double *params;
    ParamLayout pl;
    const double *lam, *U, *V, *mu;
    int S;
    double *tau;  // [n_nodes]
    R *Prow;      // [n_nodes][PRS]   row-major P, then its row sums
    R *PA;        // [n_nodes][4][16][64]   A-operand swizzle of P
    R *PTA;       // [n_nodes][4][16][64]   A-operand swizzle of P^T
    const double *sink;  // [64][64] row-major error-sink matrix, read for components with comp_mat < 0
};

template <typename R>
struct PruneArgs {
    const int32_t *sched;
    int n_sched;
    const int32_t *osched;
    int n_osched;
    int tip_rows;  // taxa + 1 (all-ones row)
    int S;
    R *Uspill;     // [n_int][P_pad][64] U messages that did not fit in LDS
    const R *PA, *PTA, *Prow;  // Prow: [n_nodes][PRS]
    const uint16_t *tips;  // [rows][P_pad]
    const uint64_t *amb;
    int P_pad, n_tiles, n_patterns;
    const double *pi;
    R *D;
    int32_t *E;    // [n_int][P_pad]
    R *V;
    int32_t *G;    // [n_vslot][P_pad]
    double *Lroot; // [P_pad]
    int32_t *Eroot;
    double *site_ll;
    const GeneInfo *genes;  // batched evaluation: gene table and tile -> gene map
    const int32_t *tile_gene;
    // Subtree parts: row part0 + blockIdx.y of the table {inside offset,
    // count, outside offset, count} selects the schedules (null: the whole
    // schedules).  root_out: the inside schedule ends at the root (site
    // log-likelihoods are written).
    const int4 *parts;
    int part0, root_out;
};

template <typename R>
struct AdjArgs {
    const int32_t *blist;
    int n_branch, n_chunks, tiles_per_chunk;
    const R *D, *V;
    const int32_t *E, *G, *Eroot;
    const double *Lroot, *weights;
    const uint16_t *tips;
    const uint64_t *amb;
    int P_pad, n_tiles, S;
    double *Apart;  // [n_branch][n_chunks][64][64]
    const GeneInfo *genes;  // batched evaluation (branch gene in the branch list)
    int n_amb;              // entries of the ambiguity-mask table
    // Branch selection (checkpointed evaluations): workgroup row y handles
    // branch list entry bsel[y] (null: entry y)
    const int32_t *bsel;
};

struct ProjArgs {
    const int32_t *qdesc;  // [64*64] packed pair descriptors
    const int8_t *nd, *dpos, *dfrom, *dto;
    const uint8_t *syn;
    const double *pos, *pi;
    int S;
    const double *params;
    ParamLayout pl;
    const double *U, *V;
    double *M;    // [n_mat][6][64][64]
    double *dmu;  // [n_mat][6]
    const int32_t *mat_gene;  // batched evaluation: gene of each matrix
};

struct FrechetArgs {
    const int32_t *blist;
    int n_branch, n_chunks, max_comp, n_mat;
    const int32_t *comp_ptr, *comp_mat;
    const double *params;
    ParamLayout pl;
    const double *Apart, *lam, *U, *V, *M, *tau;  // Apart: [n_branch][n_chunks][4096]
    int grp_size, n_grp;             // (branch, component) pairs per workgroup; groups per matrix
    const int32_t *wl_ptr, *wl_idx;  // per matrix: pair slots (branch * max_comp + component) using it
    double *out;                     // [n_branch][max_comp][8]: [0] <A, P_m>, [1] w_j d lnL / d tau share
    double *proj;                    // [n_mat][6][n_grp] <M_phi, H_m> of each group's partial H_m
    const int32_t *mat_gene;         // [n_mat] gene of each matrix in a gene batch, or null
};

struct FinalArgs {
    const double *site_ll, *weights;
    const double *Lroot;
    int n_patterns, want_grad;
    const int32_t *blist;
    int n_branch, max_comp, n_class, n_real_nodes;
    const int32_t *comp_ptr, *comp_mat;
    int convention;
    const double *params;
    ParamLayout pl;
    const double *fr, *tau, *mu, *dmu, *proj;
    int proj_stride;  // groups per matrix in proj ([n_mat][6][proj_stride])
    double *res;  // [0] lnL, [1..n_real] grad_bl, then omega[n_mat], w[n_comp], theta[5]
    const GeneInfo *genes;  // batched evaluation: one workgroup per gene, results at res_off
    // single evaluations: pinned host word that receives params[off_seq] once
    // every result is visible to the host (nullptr: none)
    unsigned long long *done;
};

// Launchers (hy_kernels.hip).
void launch_model(const ModelArgs &a, int n_mat, hipStream_t s);
// P builds (and, when pj is given, the dQ projections of n_mat matrices) in one launch
template <typename R> void launch_pmat(const PmatArgs<R> &a, const ProjArgs *pj, int n_mat, hipStream_t s);
template <typename R> void launch_pmat_host(const PmatArgs<R> &a, const double *Phost, hipStream_t s);
// ny: workgroups along y (subtree parts); the grid is n_tiles x ny
template <typename R> void launch_inside(const PruneArgs<R> &a, int T, bool lean, int ny, hipStream_t s);
// rev: every edge matrix is reversible with respect to a.pi (built from the model)
template <typename R> void launch_outside(const PruneArgs<R> &a, int T, bool rev, bool lean, int ny, hipStream_t s);
template <typename R> void launch_adjoint(const AdjArgs<R> &a, int T, hipStream_t s);
// <A_b, P_sink> of the error-sink components: one workgroup per (branch,
// component) slot, written to out[slot * 8] (and a zero time derivative at
// out[slot * 8 + 1])
struct SinkArgs {
    const double *Apart;  // [n_branch][n_chunks][4096]
    int n_chunks, max_comp;
    const int32_t *slots;  // branch * max_comp + component index
    const double *sink;    // [64][64] row-major
    double *out;           // [n_branch][max_comp][8]
};
void launch_sink(const SinkArgs &a, int n_slots, hipStream_t s);
void launch_frechet(const FrechetArgs &a, hipStream_t s);  // chunk sums, spectral transforms, projections
void launch_final(const FinalArgs &a, int n_genes, hipStream_t s);  // one workgroup per gene

// Nucleotide GTR stage (S = 4): one thread per pattern walks the tree.
// Schedule entry (int32 x GTR_STRIDE), internal nodes in post-order (the root
// last): [0] node, [1] D slot, [2] #children (<= 3), [3..5] child node ids.
constexpr int GTR_STRIDE = 8;
constexpr int GTR_BLOCK = 256;  // patterns per workgroup (P padded to a multiple)
template <typename R>
struct GtrArgs {
    const int32_t *sched;
    int n_sched;
    const int32_t *nref;  // [n_ext] reference of a node as a child: (kind << 28) | tip row or D slot
    const int32_t *nbr;   // [n_ext] branch index of the edge above a node (-1: root, identity edges)
    const R *Pm;          // [n_ext][16] row-major transition matrices
    const uint16_t *tips; // [taxa + 1][P_pad] codes: 0..3, TIP_ONES, TIP_AMB | k
    const uint64_t *amb;  // 4-bit masks of partially ambiguous codes
    const double *pi;     // [4]
    const double *weights;  // [P_pad], zero beyond P
    int P_pad, P, n_waves;
    R *D;            // [n_int][4][P_pad] inside partials (mantissas)
    int32_t *E;      // [n_int][P_pad] their power-of-two exponents
    R *U;            // [n_int][4][P_pad] outside messages at the node's end of its branch
    int32_t *H;      // [n_int][P_pad]
    double *Lroot;   // [P_pad]
    int32_t *Eroot;  // [P_pad]
    double *site_ll; // [P_pad]
    double *Apart;   // [n_branch][n_waves][16] per-wave sums of c_p V D^T
};
struct GtrReduceArgs {
    const double *Apart, *site_ll, *weights;
    int n_branch, n_waves, P;
    double *A;    // [n_branch][16]
    double *res;  // [0] lnL
};
template <typename R> void launch_gtr(const GtrArgs<R> &a, bool grad, hipStream_t s);
void launch_gtr_reduce(const GtrReduceArgs &a, bool grad, hipStream_t s);

// aBSREL's local projection (hy_branch_eval): candidate transition matrices
// of one branch scored against the resident partials of the last evaluation.
template <typename R>
struct BranchCtxArgs {
    const double *P;        // [n_cand][64][64] row-major FP64, zero beyond S
    const R *V, *D;         // the evaluation's outside / inside partial buffers
    const int32_t *E, *G;   // their exponents ([rows][P_pad])
    const uint16_t *tips;   // [taxa + 1][P_pad] codes
    const uint64_t *amb;    // ambiguity masks
    int v_slot;             // V row of the branch
    int d_slot, tip_row;    // internal child: D row (tip_row < 0); tip child: its code row
    int n_tiles, P_pad, n_patterns, S;
    double *out_L;          // [n_cand][P_pad] V_b^T P_c D_b in the partials' scaling
    int32_t *out_e;         // [P_pad] G_b + E_b: L = out_L 2^out_e
};
struct BranchSumArgs {
    const double *L;        // [n_cand][P_pad]
    const int32_t *e;       // [P_pad]
    const double *weights;  // pattern weights
    int P_pad, n_patterns;
    double *out;            // [n_cand] sum_p w_p ln L_p (w_p (-1e20) where L_p <= 0)
};
template <typename R> void launch_branch_ctx(const BranchCtxArgs<R> &a, int T, int n_cand, hipStream_t s);
void launch_branch_sum(const BranchSumArgs &a, int n_cand, hipStream_t s);

// Fixed device scratch of the allocation-free entry point
// (hy_loglik_from_P_noalloc): a file-scope __device__ array in
// hy_kernels.hip, NOALLOC_SCRATCH_BYTES per device.  noalloc_scratch returns
// its device address on `device` (looked up once per device with
// hipGetSymbolAddress, then cached).
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(128) << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;
hipError_t noalloc_scratch(int device, void **out);

}  // namespace hy
