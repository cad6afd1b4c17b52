// Host side of the MI355X codon likelihood engine: the C ABI declared in
// include/hyphy_mi355x.h.  Owns device memory and one HIP stream per handle,
// turns the caller's tree into a post-order pruning schedule, and launches the
// kernels of hy_kernels.hip.
#include "hyphy_mi355x.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "hy_internal.h"
#include "hy_lbfgs.h"
#include "hyphy/core/genetic_code.hpp"

using namespace hy;

namespace {

thread_local std::string g_create_error;

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// Spectrum pool of the per-site refits: slot -> omega (NaN: free) and the
// tick of its last use; omega -> slot by exact value and in order (nearest
// cached omega = warm start).  Valid for one model key (exchangeabilities,
// delta, psi and the problem/model generation).
struct FelPool {
    int cap = 0;
    std::vector<double> omega;
    std::vector<uint64_t> used;
    std::unordered_map<uint64_t, int> slot_of;
    std::map<double, int> by_omega;
    uint64_t tick = 0;
    double key[8] = {};
    bool keyed = false;
    int64_t n_cand = 0, n_hit = 0, n_warm = 0, n_cold = 0;  // counters (hy_fel_counters)
    void clear() {
        std::fill(omega.begin(), omega.end(), std::numeric_limits<double>::quiet_NaN());
        std::fill(used.begin(), used.end(), 0);
        slot_of.clear();
        by_omega.clear();
    }
    void drop(int slot) {
        if (std::isnan(omega[slot])) return;
        uint64_t bits;
        std::memcpy(&bits, &omega[slot], 8);
        slot_of.erase(bits);
        by_omega.erase(omega[slot]);
        omega[slot] = std::numeric_limits<double>::quiet_NaN();
    }
};

}  // namespace

// The stream an engine evaluates on.  An engine made by hy_create owns one;
// engines made together by hy_create_group share a few round-robin, so that
// thousands of per-gene engines do not hold thousands of streams.  Destroyed
// with its last engine.
struct StreamSet {
    int device = 0;
    hipStream_t s1 = nullptr;
    ~StreamSet() {
        (void)hipSetDevice(device);
        if (s1) (void)hipStreamDestroy(s1);
    }
};

struct GroupState;

struct hy_engine_s {
    int device = 0;
    std::shared_ptr<StreamSet> streams;
    std::shared_ptr<GroupState> group;  // engines made by hy_create_group
    // Fused batch context (an internal engine holding the merged problem of a
    // group): gene table, tile -> gene and matrix -> gene maps.
    int n_genes = 1;
    bool lean = false;  // schedules and inside launch of the lean variant
    DevBuf d_genes, d_tile_gene, d_mat_gene;
    hipStream_t stream = nullptr;
    // Captured evaluation graphs (parameter upload, every kernel, result
    // download), replayed while the configuration epoch is unchanged.
    struct GraphEntry {
        int prec, grad, warm;
        uint64_t epoch;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    uint64_t epoch = 1;
    bool use_graphs = true;
    std::string err;
    bool profile = false;
    hipEvent_t ev[9] = {};
    double timings[8] = {0, 0, 0, 0, 0, 0, 0, 0};

    // model frame
    bool have_model = false;
    int code_id = 0, S = 61;
    DevBuf d_nd, d_dpos, d_dfrom, d_dto, d_syn, d_qdesc, d_pos, d_pi, d_pi_custom, d_Phost;
    std::vector<double> pos_host, pi_host;  // [12], [64] as uploaded

    // problem
    bool have_problem = false;
    int num_nodes = 0, n_ext = 0, root = -1, num_taxa = 0, P = 0, n_int = 0, n_vslot = 0, n_branch = 0;
    std::vector<int32_t> sched, osched, blist;
    std::vector<int32_t> order, int_slot, vslot, leaf_taxon_ext;
    std::vector<std::vector<int32_t>> ext_children;
    std::vector<uint8_t> ext_ident;
    int ext_root = -1, sched_key = 0;
    int n_cu = 256, last_tile = 0;
    const void *ident_tiles[2] = {nullptr, nullptr};  // PA buffer whose identity tiles are current
    DevBuf d_osched, d_Uspill[2];
    int n_uspill = 0;
    bool lean_out = false;  // lean outside variant (FP32, T = 16, subtree parts)
    // Subtree parts of the pruning passes (single problems with few tiles):
    // per part {inside offset, count, outside offset, count} in schedule
    // entries; the last row is the top part, which holds the root.  n_parts
    // counts the subtree parts (0: one schedule for the whole tree).
    std::vector<int32_t> parts;
    DevBuf d_parts;
    int n_parts = 0;
    bool no_parts = false;  // HY_NO_PARTS=1: one schedule per pass
    // Checkpointed evaluations: when the partials of every node (D, V, their
    // exponents, spilled messages) exceed the budget, only the top part's
    // partials and the part roots' stay resident and each part's interior is
    // recomputed into a shared segment before its outside and adjoint passes.
    // mem_budget: bytes allowed for those buffers (0: the device's free
    // memory); ck_groups: offsets of the adjoint branch groups in d_bsel.
    bool ck = false;
    size_t mem_budget = 0, partial_bytes = 0;
    int ck_dslots = 0, ck_vslots = 0;
    std::string ck_error;  // non-empty: the checkpointed layout failed its host-side check
    std::vector<int32_t> ck_groups, ck_bsel, blist_ck;
    DevBuf d_blist_ck, d_bsel;
    std::vector<int32_t> real_parent;
    std::vector<uint16_t> codes;  // [num_taxa][P]
    std::vector<uint64_t> amb;
    std::vector<double> weights;
    DevBuf d_sched, d_blist, d_tips, d_amb, d_weights;
    int n_amb = 0;  // ambiguity masks in d_amb
    // Per-evaluation pattern weights replacing d_weights in the adjoint and the
    // final reductions (posterior-weighted passes of the synonymous-rate
    // mixture); nullptr = the problem's weights.  Evaluations with it set run
    // eagerly (graphs bake the weight pointer in).
    DevBuf d_wover;
    const double *wover = nullptr;
    // caller's pattern weights (hy_set_pattern_weights), in place of the
    // alignment's for every later evaluation until cleared
    DevBuf d_wuser;
    bool wuser = false;
    // hy_loglik_from_P_noalloc: device address of the fixed scratch on this
    // handle's device (looked up at creation; nullptr with the HIP error in
    // scratch_err when that failed) and the host layouts of its fixed
    // configuration (16-pattern tiles, one whole-tree inside schedule), built
    // by hy_set_problem: tip codes [taxa + 1][na_P_pad], padded pattern
    // weights, the inside schedule and the identity edges' nodes
    void *scratch = nullptr;
    hipError_t scratch_err = hipSuccess;
    // nucleotide GTR stage (problems set with S = 4): node tables, padded tips
    // and weights (built at the first hy_gtr_eval after hy_set_problem) and
    // per-evaluation buffers
    bool gtr_ready = false;
    int g_P_pad = 0, g_n_sched = 0;
    DevBuf d_gsched, d_gnref, d_gnbr, d_gtips, d_gw, d_gP, d_gpi, d_gD, d_gE, d_gU, d_gH, d_gL, d_gLE, d_gsite,
        d_gApart, d_gA, d_gres;
    std::vector<uint8_t> g_ident;  // extended nodes whose edge is an identity
    // per-site refits (hy_fel_eval): a pool of MG94 spectra keyed by omega
    // (eigenvectors kept for warm starts), per-call staging and scratch
    FelPool fel;
    DevBuf d_fZ, d_flam, d_fU, d_fV, d_fmu, d_fstatus, d_fstage, d_fbl, d_fD, d_fE, d_fout, d_fident;
    DevBuf d_bP, d_bL, d_be, d_blnl;  // hy_branch_eval: candidates, products, exponents, sums
    // single spectra (hy_mg94_spectrum): last eigenvectors kept for a warm start
    DevBuf d_sprm, d_slam, d_sU, d_sV, d_smu, d_sZ, d_sstatus;
    uint64_t s_model_gen = ~uint64_t(0);
    void *f_pin = nullptr;  // pinned staging: inputs, then outputs
    size_t f_pin_bytes = 0;
    void *p_pin = nullptr;  // pinned staging of the pool's spectrum solves
    size_t p_pin_bytes = 0;
    DevBuf d_pstage;
    std::vector<double> f_bl;  // branch lengths last uploaded to d_fbl
    bool f_ident_ready = false;
    uint64_t model_gen = 0;  // hy_set_problem / hy_set_model calls
    std::vector<uint16_t> na_tips;
    std::vector<double> na_weights;
    std::vector<int32_t> na_sched, na_ident;
    int na_P_pad = 0;
    double *h_site = nullptr;  // pinned staging of per-pattern lnL
    size_t h_site_cap = 0;
    // Partials left by the last evaluation (read back by hy_read_partials):
    // precision index (-1 none) and whether the outside pass ran.
    int part_ri = -1;
    bool part_outside = false;
    int T_req = 0, T = 0, n_tiles = 0, P_pad = 0;
    // hy_autotune's measured choice per precision ([0] FP32, [1] FP64): tile
    // width (0: cost model) and subtree parts (-1: rule, 0 off, 1 on)
    int tuned_T[2] = {0, 0}, tuned_parts[2] = {-1, -1};

    // mixture
    int n_mat = 1, n_class = 1, n_comp = 1, convention = HY_TIME_RAW, max_comp = 1;
    std::vector<int32_t> comp_ptr{0, 1}, comp_mat{0}, branch_class;
    DevBuf d_comp_ptr, d_comp_mat;

    // per-eval buffers
    ParamLayout pl{};
    double *h_params = nullptr;
    size_t h_params_cap = 0;
    double *h_res = nullptr;
    size_t h_res_cap = 0;
    int *h_status = nullptr;
    // completion word: k_final stores the call number of a single evaluation
    // here after its results (h_res) are visible to the host
    unsigned long long *h_done = nullptr, *h_done_dev = nullptr;
    uint64_t seq = 0;
    bool publish_done = false;  // run_eval: k_final publishes the call number
    bool wait_done = false;     // the pending evaluation publishes h_done
    // device addresses of the pinned h_params / h_res / h_status: k_model reads
    // the parameters from h_params under graph capture (zc_params) and writes
    // the eigensolver status, k_final writes the results, with no copies
    const double *h_params_dev = nullptr;
    double *h_res_dev = nullptr;
    int *h_status_dev = nullptr;
    bool zc_params = false;
    DevBuf d_params, d_lam, d_U, d_V, d_mu, d_tau, d_M, d_dmu, d_Z;
    bool warm_enable = true, warm_valid = false;
    DevBuf d_Prow[2], d_PA[2], d_PTA[2];  // Prow: [n_ext][PRS] (P row-major, then its row sums)
    DevBuf d_D[2], d_Vp[2], d_E, d_G, d_Lroot, d_Eroot, d_site, d_Apart, d_fr;
    DevBuf d_wl, d_proj;
    // BUSTED's error sink: components with comp_mat == HY_SINK, the 64 x 64
    // sink matrix (hy_set_sink) and the (branch, component) slots using it
    bool has_sink = false, sink_set = false;
    DevBuf d_sink, d_sinkwl;
    std::vector<int32_t> sink_wl;
    int grp_size = 1, n_grp = 1;  // Fréchet stage grouping of (branch, component) pairs per matrix
    std::vector<int32_t> wl;  // uploaded copy of the per-matrix W slot lists
    int n_chunks = 1, tiles_per_chunk = 1;

    size_t device_bytes = 0;
};

struct GroupState {
    hy_engine ctx = nullptr;
    std::vector<hy_engine> members;  // genes of the merged problem, in order
    std::vector<uint64_t> epochs;
    int prec = -1;
    int tuned_T[2] = {0, 0};  // hy_autotune_batch's tile per precision ([0] FP32, [1] FP64; 0: cost model)
    bool no_fuse = false;  // HY_NO_FUSE=1: evaluate group batches engine by engine
    struct Slice {
        int node0, n_ext, n_real, mat0, n_mat, comp0, n_comp, pbase, P;
        size_t res_off;
    };
    std::vector<Slice> slices;
    ~GroupState() {
        if (ctx) hy_destroy(ctx);
    }
};

namespace {

// Pinned buffers the kernels read or write directly (parameters, results,
// eigensolver status): coherent, so device reads see the host's latest
// writes and host reads see the kernels' writes without cache maintenance.
constexpr unsigned kZeroCopyFlags = hipHostMallocMapped | hipHostMallocCoherent;

int fail(hy_engine h, int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (h)
        h->err = buf;
    else
        g_create_error = buf;
    return code;
}

#define HIP_TRY(h, expr)                                                                          \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess) return fail(h, HY_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

void release(hy_engine h, DevBuf &b) {
    if (b.p) {
        (void)hipFree(b.p);
        h->device_bytes -= b.bytes;
    }
    b.p = nullptr;
    b.bytes = 0;
}

int ensure(hy_engine h, DevBuf &b, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (b.p && b.bytes >= bytes) return HY_OK;
    release(h, b);
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(h, HY_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    }
    b.bytes = bytes;
    h->device_bytes += bytes;
    ++h->epoch;  // device pointers changed: captured graphs are stale
    return HY_OK;
}

int upload(hy_engine h, DevBuf &b, const void *src, size_t bytes) {
    int rc = ensure(h, b, bytes);
    if (rc) return rc;
    HIP_TRY(h, hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return HY_OK;
}

#define TRY(expr)                \
    do {                         \
        int _rc = (expr);        \
        if (_rc != HY_OK) return _rc; \
    } while (0)

// Instantiated evaluation graphs are bounded process-wide: each holds runtime
// resources, and a run that instantiated about 3,000 of them (one per gene
// engine) ended in a host segmentation fault inside the runtime.  At the cap
// (HY_MAX_GRAPHS, default 512) evaluations launch their kernels directly, the
// same kernels in the same order.
std::atomic<int> g_live_graphs{0};

int graph_cap() {
    const char *e = std::getenv("HY_MAX_GRAPHS");
    if (!e || !*e) return 512;
    const int v = std::atoi(e);
    return v < 0 ? 0 : v;
}

bool acquire_graph_slot() {
    const int cap = graph_cap();
    int cur = g_live_graphs.load();
    while (cur < cap)
        if (g_live_graphs.compare_exchange_weak(cur, cur + 1)) return true;
    return false;
}

void destroy_graph_exec(hipGraphExec_t &exec) {
    if (!exec) return;
    (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    g_live_graphs.fetch_sub(1);
}

// Pattern-tile width: the pruning kernels are latency-bound per tree node, so
// the fewest "rounds" of resident workgroups win; a wider tile costs more per
// node (measured ~1.3x for 32, ~2x for 64).  FP32 tiles of 16 fit two
// workgroups per CU, every other variant one.
// FP64 runs tiles of at most 32 patterns: its 64-pattern outside pass needs
// more registers than a lane has (the compiler spilled 469-627 VGPRs to
// scratch), so those kernels are not built and a request for 64 runs at 32.
constexpr int max_tile(int prec) { return prec == HY_PREC_FP64 ? 32 : 64; }

int choose_tile(int P, int req, int prec, int n_cu) {
    if (req == 16 || req == 32 || req == 64) return std::min(req, max_tile(prec));
    const double node_cost[3] = {1.0, 1.3, 2.0};
    int best = 16;
    double best_cost = 1e300;
    for (int k = 0; (16 << k) <= max_tile(prec); ++k) {
        const int T = 16 << k;
        const int slots = n_cu * ((prec == HY_PREC_FP32 && T == 16) ? 2 : 1);
        const int tiles = (P + T - 1) / T;
        const double cost = ((tiles + slots - 1) / slots) * node_cost[k];
        if (cost < best_cost) {  // ties go to the narrower tile (measured faster)
            best_cost = cost;
            best = T;
        }
    }
    return best;
}

void build_layout(hy_engine h) {
    ParamLayout &pl = h->pl;
    pl.n_nodes = h->n_ext;
    pl.n_mat = h->n_mat;
    pl.n_comp = h->n_comp;
    int o = 0;
    pl.off_bl = o;
    o += h->n_ext;
    pl.off_alpha = o++;
    pl.off_omega = o;
    o += h->n_mat;
    pl.off_w = o;
    o += h->n_comp;
    pl.off_theta = o;
    o += 5;
    pl.off_delta = o++;
    pl.off_psi = o++;
    pl.off_seq = o++;
    pl.total = o;
}

// Inside / outside schedules and subtree parts of one tiling (see make_schedules).
struct Schedules {
    std::vector<int32_t> sched, osched, parts;
    int n_parts = 0, n_uspill = 0;
    // Checkpointed layout (ck): D index and V slot of every extended node.
    // Persistent slots first (D: the top part's nodes and the part roots;
    // V: branches below top-part nodes), then one segment the parts share in
    // turn (D: a part's other internal nodes; V: branches below them).
    std::vector<int32_t> islot, vsl;
    int n_dslot = 0, n_vslot = 0, n_dpersist = 0, n_vpersist = 0;
    // branch list indices grouped by the part of the branch's parent: the
    // top part first, then part 0, 1, ...; group g is bsel[bgroup[g] ..
    // bgroup[g + 1])
    std::vector<int32_t> bsel, bgroup;
};
void make_schedules(const hy_engine_s *h, int T, int esz, bool split, bool lean, bool lean_out, Schedules &o,
                    bool ck = false);
void build_schedules(hy_engine h, int T, int esz, bool split);

// Identity P of the artificial nodes in the Prow (with row sums), PA/PTA
// ([w][l][kk]) layouts of the working type (esz = 4 or 8 bytes).
int upload_identity_tiles(hy_engine h, int ri, int esz) {
    std::vector<double> prow(PRS, 0.0), pa(4096, 0.0);
    for (int i = 0; i < 64; ++i) prow[i * 64 + i] = 1.0;
    for (int i = 0; i < h->S; ++i) prow[4096 + i] = 1.0;  // row sums over the sense codons
    for (int idx = 0; idx < 4096; ++idx) {
        const int w = idx >> 10, l = (idx >> 4) & 63, kk = idx & 15;
        pa[idx] = (16 * w + (l & 15)) == (4 * kk + (l >> 4)) ? 1.0 : 0.0;
    }
    std::vector<float> prow32(prow.begin(), prow.end()), pa32(pa.begin(), pa.end());
    const void *srcs[3] = {esz == 8 ? (const void *)prow.data() : prow32.data(),
                           esz == 8 ? (const void *)pa.data() : pa32.data(),
                           esz == 8 ? (const void *)pa.data() : pa32.data()};
    DevBuf *dst[3] = {&h->d_Prow[ri], &h->d_PA[ri], &h->d_PTA[ri]};
    const size_t per[3] = {PRS, 4096, 4096};
    for (int n = 0; n < h->n_ext; ++n) {
        if (!h->ext_ident[n]) continue;
        for (int k = 0; k < 3; ++k)
            HIP_TRY(h, hipMemcpyAsync((char *)dst[k]->p + (size_t)n * per[k] * esz, srcs[k], per[k] * esz,
                                      hipMemcpyHostToDevice, h->stream));
    }
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return HY_OK;
}

// (Re)allocate everything that depends on the tile width and precision.
// Fréchet stage grouping: pairs (branch, component) per workgroup so that
// about two workgroups per CU run, each matrix's U and V loaded once per group.
template <typename Lists>
void set_frechet_groups(hy_engine h, const Lists &per) {
    size_t total = 0, longest = 1;
    for (const auto &l : per) {
        total += l.size();
        longest = std::max(longest, l.size());
    }
    const size_t target = (size_t)2 * h->n_cu;
    // about two groups per CU, and no group longer than the mean list (many
    // short lists of uneven length, e.g. gene batches, would otherwise leave
    // the longest lists to a few workgroups at the end)
    const size_t mean = per.empty() ? 1 : std::max<size_t>(1, (total + per.size() - 1) / per.size());
    h->grp_size = (int)std::max<size_t>(1, std::min((total + target - 1) / target, mean));
    h->n_grp = (int)((longest + h->grp_size - 1) / h->grp_size);
}

int prepare_buffers(hy_engine h, int prec) {
    const int ri = prec == HY_PREC_FP64 ? 1 : 0;
    const int T = choose_tile(h->P, h->T_req ? h->T_req : h->tuned_T[ri], prec, h->n_cu);
    const size_t rs = prec == HY_PREC_FP64 ? 8 : 4;
    if (T != h->T) {
        ++h->epoch;
        h->T = T;
        h->n_tiles = (h->P + T - 1) / T;
        h->P_pad = h->n_tiles * T;
        // tip rows: taxa + one all-ones row
        std::vector<uint16_t> tips((size_t)(h->num_taxa + 1) * h->P_pad, TIP_ONES);
        for (int t = 0; t < h->num_taxa; ++t)
            std::memcpy(&tips[(size_t)t * h->P_pad], &h->codes[(size_t)t * h->P], sizeof(uint16_t) * h->P);
        TRY(upload(h, h->d_tips, tips.data(), tips.size() * sizeof(uint16_t)));
        std::vector<double> w(h->P_pad, 0.0);
        std::copy(h->weights.begin(), h->weights.end(), w.begin());
        TRY(upload(h, h->d_weights, w.data(), w.size() * sizeof(double)));
        // the partial buffers of both precisions are kept (ensure() grows
        // them when the new tiling needs more): switching precision back and
        // forth between tile widths then allocates nothing
    }
    // Few tiles (fewer than two per CU): the pruning passes split the tree
    // into subtree parts walked by separate workgroups (a shorter dependent
    // chain per workgroup, more workgroups in flight); FP32 T = 16 then runs
    // the lean inside variant (four workgroups per CU).
    bool split = !h->no_parts && h->n_tiles < 2 * h->n_cu;
    if (!h->T_req && h->tuned_parts[ri] >= 0) split = !h->no_parts && h->tuned_parts[ri] == 1;
    const size_t Pp = (size_t)h->P_pad;
    // Checkpointing: the resident layout's partial buffers (D, V, U spills at
    // worst, exponents) against the budget, by default the device's free
    // memory plus what this handle's partial buffers already hold
    {
        const size_t full = (size_t)(2 * h->n_int + h->n_vslot) * Pp * 64 * rs + (size_t)(h->n_int + h->n_vslot) * Pp * 4;
        size_t held = h->d_E.bytes + h->d_G.bytes;
        for (int k = 0; k < 2; ++k) held += h->d_D[k].bytes + h->d_Vp[k].bytes + h->d_Uspill[k].bytes;
        bool ck;
        if (h->mem_budget) {
            ck = full > h->mem_budget;
        } else if (!h->ck && h->d_D[ri].bytes >= (size_t)h->n_int * Pp * 64 * rs &&
                   h->d_Vp[ri].bytes >= (size_t)h->n_vslot * Pp * 64 * rs) {
            ck = false;  // the resident layout is allocated already (no query per evaluation)
        } else {
            size_t fr = 0, tot = 0;
            size_t budget = ~size_t(0);
            if (hipMemGetInfo(&fr, &tot) == hipSuccess) {
                const size_t margin = size_t(1) << 30;
                budget = fr + held > margin ? fr + held - margin : 1;
            }
            ck = full > budget;
        }
        if (ck != h->ck) {
            // switching layouts: the partial buffers change size (release,
            // so the checkpointed layout really holds less memory)
            for (int k = 0; k < 2; ++k) {
                release(h, h->d_D[k]);
                release(h, h->d_Vp[k]);
                release(h, h->d_Uspill[k]);
            }
            release(h, h->d_E);
            release(h, h->d_G);
            h->ck = ck;
            ++h->epoch;
        }
    }
    if (h->ck) split = true;
    h->lean = !h->ck && (lean_inside(prec == HY_PREC_FP32, T, h->n_tiles, h->n_cu) ||
                         (split && prec == HY_PREC_FP32 && T == 16));
    h->lean_out = !h->ck && split && prec == HY_PREC_FP32 && T == 16;
    if (h->sched_key != T * 16 + (int)rs + (h->lean ? 4096 : 0) + (split ? 8192 : 0) + (h->lean_out ? 16384 : 0) +
                            (h->ck ? 32768 : 0)) {
        ++h->epoch;
        build_schedules(h, T, (int)rs, split);
        if (h->ck && !h->ck_error.empty()) {
            h->sched_key = 0;
            return fail(h, HY_ERR_STATE, "internal: checkpointed layout check failed (%s)", h->ck_error.c_str());
        }
        TRY(upload(h, h->d_sched, h->sched.data(), h->sched.size() * 4));
        TRY(upload(h, h->d_osched, h->osched.data(), h->osched.size() * 4));
        TRY(upload(h, h->d_parts, h->parts.data(), h->parts.size() * 4));
        if (h->ck) {
            // the adjoint's branch list with the checkpointed slots, and its groups
            TRY(upload(h, h->d_blist_ck, h->blist_ck.data(), h->blist_ck.size() * 4));
            TRY(upload(h, h->d_bsel, h->ck_bsel.data(), std::max<size_t>(h->ck_bsel.size(), 1) * 4));
        }
    }
    const size_t nd = h->ck ? (size_t)h->ck_dslots : (size_t)h->n_int;
    const size_t nv = h->ck ? (size_t)h->ck_vslots : (size_t)h->n_vslot;
    TRY(ensure(h, h->d_D[ri], nd * Pp * 64 * rs));
    TRY(ensure(h, h->d_Vp[ri], nv * Pp * 64 * rs));
    TRY(ensure(h, h->d_Uspill[ri], h->n_uspill ? nd * Pp * 64 * rs : 256));
    TRY(ensure(h, h->d_E, nd * Pp * 4));
    TRY(ensure(h, h->d_G, nv * Pp * 4));
    h->partial_bytes = h->d_D[ri].bytes + h->d_Vp[ri].bytes + h->d_Uspill[ri].bytes + h->d_E.bytes + h->d_G.bytes;
    TRY(ensure(h, h->d_Lroot, Pp * 8));
    TRY(ensure(h, h->d_Eroot, Pp * 4));
    TRY(ensure(h, h->d_site, Pp * 8));
    const size_t ne = (size_t)h->n_ext;
    TRY(ensure(h, h->d_Prow[ri], ne * PRS * rs));
    TRY(ensure(h, h->d_PA[ri], ne * 4096 * rs));
    TRY(ensure(h, h->d_PTA[ri], ne * 4096 * rs));
    if (h->ident_tiles[ri] != h->d_PA[ri].p) {
        // Edges added to resolve polytomies (identity transition): exact
        // identity tiles in every P layout; k_pmat never overwrites them.
        TRY(upload_identity_tiles(h, ri, (int)rs));
        h->ident_tiles[ri] = h->d_PA[ri].p;
    }
    // adjoint chunking: workgroups take whole 64-pattern groups (64 / T tiles);
    // as many chunks per branch as fill the resident workgroup slots (two per
    // CU in FP32, one in FP64) in one round, so no partial last round
    int nb = std::max(1, h->n_branch);
    const int G = 64 / T;
    const int n_groups = (h->n_tiles + G - 1) / G;
    const int slots = h->n_cu * (rs == 4 ? 2 : 1);
    // checkpointed: one adjoint launch per branch group, sized for the largest
    int nbl = nb;
    if (h->ck) {
        nbl = 1;
        for (size_t g = 0; g + 1 < h->ck_groups.size(); ++g) nbl = std::max(nbl, h->ck_groups[g + 1] - h->ck_groups[g]);
    }
    const int nchg = std::min(n_groups, std::max(1, slots / nbl));
    h->tiles_per_chunk = ((n_groups + nchg - 1) / nchg) * G;
    h->n_chunks = (h->n_tiles + h->tiles_per_chunk - 1) / h->tiles_per_chunk;
    TRY(ensure(h, h->d_Apart, (size_t)nb * h->n_chunks * 4096 * 8));
    TRY(ensure(h, h->d_fr, (size_t)nb * h->max_comp * 8 * 8));
    {
        // per matrix: the W slots (branch * max_comp + component index) whose
        // component uses it, branch-major (fixed summation order)
        std::vector<std::vector<int32_t>> per(h->n_mat);
        std::vector<int32_t> swl;
        for (int bi = 0; bi < h->n_branch; ++bi) {
            const int cls = h->blist[bi * BR_STRIDE + 3];
            for (int j = h->comp_ptr[cls]; j < h->comp_ptr[cls + 1]; ++j) {
                const int32_t slot = bi * h->max_comp + (j - h->comp_ptr[cls]);
                if (h->comp_mat[j] < 0)
                    swl.push_back(slot);
                else
                    per[h->comp_mat[j]].push_back(slot);
            }
        }
        if (swl != h->sink_wl || (!swl.empty() && !h->d_sinkwl.p)) {
            if (!swl.empty()) TRY(upload(h, h->d_sinkwl, swl.data(), swl.size() * 4));
            h->sink_wl = swl;
            ++h->epoch;  // the sink launch's grid changed
        }
        std::vector<int32_t> wl(h->n_mat + 1, 0);
        for (int m = 0; m < h->n_mat; ++m) wl[m + 1] = wl[m] + (int32_t)per[m].size();
        for (int m = 0; m < h->n_mat; ++m) wl.insert(wl.end(), per[m].begin(), per[m].end());
        if (wl != h->wl || !h->d_wl.p) {
            TRY(upload(h, h->d_wl, wl.data(), wl.size() * 4));
            h->wl = wl;
        }
        set_frechet_groups(h, per);
    }
    // per (matrix, parameter, group) projections: sized after the grouping
    TRY(ensure(h, h->d_proj, (size_t)h->n_mat * 6 * h->n_grp * 8));
    const int nm = h->n_mat;
    TRY(ensure(h, h->d_lam, (size_t)nm * 64 * 8));
    TRY(ensure(h, h->d_U, (size_t)nm * 4096 * 8));
    if (!h->d_Z.p || h->d_Z.bytes < (size_t)nm * 4096 * 8) h->warm_valid = false;
    TRY(ensure(h, h->d_Z, (size_t)nm * 4096 * 8));
    TRY(ensure(h, h->d_V, (size_t)nm * 4096 * 8));
    TRY(ensure(h, h->d_mu, (size_t)nm * 8));
    TRY(ensure(h, h->d_M, (size_t)nm * 6 * 4096 * 8));
    TRY(ensure(h, h->d_dmu, (size_t)nm * 6 * 8));
    TRY(ensure(h, h->d_tau, ne * 8));
    build_layout(h);
    TRY(ensure(h, h->d_params, (size_t)h->pl.total * 8));
    const size_t nres = 1 + (size_t)h->num_nodes + h->n_mat + h->n_comp + 5;
    if (h->h_params_cap < (size_t)h->pl.total) {
        if (h->h_params) (void)hipHostFree(h->h_params);
        HIP_TRY(h, hipHostMalloc((void **)&h->h_params, sizeof(double) * h->pl.total, kZeroCopyFlags));
        h->h_params_cap = h->pl.total;
        HIP_TRY(h, hipHostGetDevicePointer((void **)&h->h_params_dev, h->h_params, 0));
    }
    if (h->h_res_cap < nres) {
        if (h->h_res) (void)hipHostFree(h->h_res);
        HIP_TRY(h, hipHostMalloc((void **)&h->h_res, sizeof(double) * nres, kZeroCopyFlags));
        h->h_res_cap = nres;
        HIP_TRY(h, hipHostGetDevicePointer((void **)&h->h_res_dev, h->h_res, 0));
    }
    if (!h->h_status) {
        HIP_TRY(h, hipHostMalloc((void **)&h->h_status, sizeof(int) * 64, kZeroCopyFlags));
        HIP_TRY(h, hipHostGetDevicePointer((void **)&h->h_status_dev, h->h_status, 0));
    }
    if (!h->h_done) {
        HIP_TRY(h, hipHostMalloc((void **)&h->h_done, 64, kZeroCopyFlags));
        HIP_TRY(h, hipHostGetDevicePointer((void **)&h->h_done_dev, h->h_done, 0));
        *h->h_done = 0;
    }
    return HY_OK;
}

// Inside / outside schedules with LDS slot assignment for tile width T and
// element size esz.  With split, the tree is cut into a top part holding the
// root and subtree parts hanging off it: every part gets its own schedule
// (and its own LDS slots), subtree parts are walked by separate workgroups
// and exchange their results with the top part through HBM (inside: the
// part root's partials; outside: the message into the part root, spilled).
void make_schedules(const hy_engine_s *h, int T, int esz, bool split, bool lean, bool lean_out, Schedules &o,
                    bool ck) {
    const int NSI = inside_slots(T, esz, lean), NSO = outside_slots(T, esz, lean_out);
    const auto &ch = h->ext_children;
    const int n_ext = (int)ch.size();
    // D index / V slot maps: the problem's own, or the checkpointed layout
    // (filled below once the parts are known)
    const int32_t *isl = h->int_slot.data(), *vsl = h->vslot.data();
    auto ref_of = [&](int c) -> int32_t {
        if (ch[c].empty()) {
            const int row = (h->leaf_taxon_ext[c] >= 0) ? h->leaf_taxon_ext[c] : h->num_taxa;
            return (KIND_TIP << 28) | row;
        }
        return (KIND_INT << 28) | isl[c];  // identity edges carry an exact identity P
    };
    // ---- parts: grow the top from the root by cutting the largest subtree
    // while it holds more than a quarter of the internal nodes (at least 8);
    // checkpointed: more than about sqrt(internal nodes) (at least 4)
    std::vector<int> part_of(n_ext, -1);  // -1: top part
    std::vector<int> proots;
    if (ck && split && (int)h->order.size() >= 8) {
        std::vector<int> nint(n_ext, 0);
        for (int u : h->order) {
            nint[u] = 1;
            for (int c : ch[u]) nint[u] += nint[c];
        }
        const int target = std::max(4, (int)std::ceil(std::sqrt((double)h->order.size())));
        std::vector<int> cand;
        for (int c : ch[h->ext_root])
            if (!ch[c].empty()) cand.push_back(c);
        for (;;) {
            auto big = std::max_element(cand.begin(), cand.end(), [&](int x, int y) { return nint[x] < nint[y]; });
            if (big == cand.end() || nint[*big] <= target) break;
            const int u = *big;
            cand.erase(big);
            for (int c : ch[u])
                if (!ch[c].empty()) cand.push_back(c);
        }
        std::stable_sort(cand.begin(), cand.end(), [&](int x, int y) { return nint[x] > nint[y]; });
        proots = cand;
        for (size_t p = 0; p < proots.size(); ++p) {
            std::vector<int> st{proots[p]};
            while (!st.empty()) {
                const int u = st.back();
                st.pop_back();
                part_of[u] = (int)p;
                for (int c : ch[u]) st.push_back(c);
            }
        }
    } else if (split && (int)h->order.size() >= 16) {
        std::vector<int> nint(n_ext, 0);
        for (int u : h->order) {
            nint[u] = 1;
            for (int c : ch[u]) nint[u] += nint[c];
        }
        // (64 taxa x 5000 codons, 62 internal nodes: quarter 0.507 ms, third 0.501, sixth 0.505,
        // eighth 0.521 per evaluation; parts then no longer cover the latency)
        const int target = std::max(8, ((int)h->order.size() + 3) / 4);
        std::vector<int> cand;
        for (int c : ch[h->ext_root])
            if (!ch[c].empty()) cand.push_back(c);
        for (;;) {
            auto big = std::max_element(cand.begin(), cand.end(), [&](int x, int y) { return nint[x] < nint[y]; });
            if (big == cand.end() || nint[*big] <= target) break;
            const int u = *big;
            cand.erase(big);
            for (int c : ch[u])
                if (!ch[c].empty()) cand.push_back(c);
        }
        // largest parts first (their workgroups start first)
        std::stable_sort(cand.begin(), cand.end(), [&](int x, int y) { return nint[x] > nint[y]; });
        if (cand.size() >= 2) {
            proots = cand;
            for (size_t p = 0; p < proots.size(); ++p) {
                std::vector<int> st{proots[p]};
                while (!st.empty()) {
                    const int u = st.back();
                    st.pop_back();
                    part_of[u] = (int)p;
                    for (int c : ch[u]) st.push_back(c);
                }
            }
        }
    }
    const int np = (int)proots.size();
    std::vector<uint8_t> is_proot(n_ext, 0);
    for (int r : proots) is_proot[r] = 1;
    // post-order of every part (top last), filtered from the whole-tree order
    std::vector<std::vector<int>> porder(np + 1);
    for (int u : h->order) porder[part_of[u] < 0 ? np : part_of[u]].push_back(u);
    o.islot.clear();
    o.vsl.clear();
    o.bsel.clear();
    o.bgroup.clear();
    o.n_dslot = o.n_vslot = 0;
    if (ck) {
        std::vector<int> par(n_ext, -1);
        for (int u = 0; u < n_ext; ++u)
            for (int c : ch[u]) par[c] = u;
        // D: persistent top-part nodes and part roots, then the segment
        o.islot.assign(n_ext, -1);
        int pd = 0;
        for (int u : h->order)
            if (part_of[u] < 0 || is_proot[u]) o.islot[u] = pd++;
        int seg = 0;
        for (int p = 0; p < np; ++p) {
            int k = 0;
            for (int u : porder[p])
                if (!is_proot[u]) o.islot[u] = pd + k++;
            seg = std::max(seg, k);
        }
        o.n_dslot = std::max(1, pd + seg);
        o.n_dpersist = pd;
        // V: branches below top-part nodes persistent, the others in the
        // segment of their parent's part (the problem's V slot order within each)
        std::vector<int> byv(n_ext, -1);
        for (int c = 0; c < n_ext; ++c)
            if (c != h->ext_root) byv[h->vslot[c]] = c;
        o.vsl.assign(n_ext, -1);
        std::vector<int> vk(np, 0);
        int pv = 0;
        for (int v = 0; v < n_ext; ++v) {
            const int c = byv[v];
            if (c >= 0 && part_of[par[c]] < 0) o.vsl[c] = pv++;
        }
        int vseg = 0;
        for (int v = 0; v < n_ext; ++v) {
            const int c = byv[v];
            if (c < 0 || part_of[par[c]] < 0) continue;
            const int p = part_of[par[c]];
            o.vsl[c] = pv + vk[p]++;
            vseg = std::max(vseg, vk[p]);
        }
        o.n_vslot = std::max(1, pv + vseg);
        o.n_vpersist = pv;
        // adjoint groups of the branch list
        std::vector<std::vector<int32_t>> grp(np + 1);
        for (int bi = 0; bi < h->n_branch; ++bi) {
            const int c = h->blist[(size_t)bi * BR_STRIDE];
            grp[part_of[par[c]] < 0 ? 0 : part_of[par[c]] + 1].push_back(bi);
        }
        o.bgroup.push_back(0);
        for (auto &g : grp) {
            o.bsel.insert(o.bsel.end(), g.begin(), g.end());
            o.bgroup.push_back((int32_t)o.bsel.size());
        }
        isl = o.islot.data();
        vsl = o.vsl.data();
    }
    o.parts.assign((size_t)(np + 1) * 4, 0);
    // ---- inside: free-list LDS slots, spill to HBM when none is free ----
    std::vector<int32_t> where(n_ext, -1);  // LDS slot of a computed internal node (-1: HBM)
    o.sched.clear();
    for (int p = 0; p <= np; ++p) {
        std::vector<int> free_slots;
        for (int sl = NSI - 1; sl >= 0; --sl) free_slots.push_back(sl);
        o.parts[p * 4 + 0] = (int32_t)(o.sched.size() / SCHED_STRIDE);
        o.parts[p * 4 + 1] = (int32_t)porder[p].size();
        for (int u : porder[p]) {
            int32_t e[SCHED_STRIDE] = {0};
            e[0] = u;
            e[1] = isl[u];
            e[2] = (int32_t)ch[u].size();
            for (size_t c = 0; c < ch[u].size(); ++c) {
                const int cn = ch[u][c];
                e[6 + 3 * c] = cn;
                e[7 + 3 * c] = ref_of(cn);
                e[8 + 3 * c] = ch[cn].empty() ? -1 : where[cn];
            }
            if (ch[u].empty()) {
                // single-node tree: the root is a leaf whose partials are its tip
                // indicator (one tip child on an identity edge, node id -1)
                e[2] = 1;
                e[6] = -1;
                e[7] = ref_of(u);
                e[8] = -1;
            }
            for (int cn : ch[u])
                if (!ch[cn].empty() && where[cn] >= 0) free_slots.push_back(where[cn]);
            int out = -1;
            if (u != h->ext_root && !is_proot[u] && !free_slots.empty()) {
                out = free_slots.back();
                free_slots.pop_back();
            }
            where[u] = out;
            e[3] = out;
            o.sched.insert(o.sched.end(), e, e + SCHED_STRIDE);
        }
    }
    // ---- outside: top part first, then the subtree parts, each in reverse
    // order; pending U messages in LDS slots, into a part root through HBM
    std::vector<int32_t> usrc(n_ext, U_SPILL);
    usrc[h->ext_root] = U_ROOT;
    o.osched.clear();
    int spills = 0;
    for (int q = 0; q <= np; ++q) {
        const int p = (q == 0) ? np : q - 1;
        std::vector<int> free_slots;
        for (int sl = NSO - 1; sl >= 0; --sl) free_slots.push_back(sl);
        o.parts[p * 4 + 2] = (int32_t)(o.osched.size() / OSCHED_STRIDE);
        o.parts[p * 4 + 3] = (int32_t)porder[p].size();
        const std::vector<int> &ord = porder[p];
        const int n_order = (int)ord.size();
        for (int i = n_order - 1; i >= 0; --i) {
            const int u = ord[i];
            const int next = i > 0 ? ord[i - 1] : -1;  // entry processed after u
            int32_t e[OSCHED_STRIDE] = {0};
            e[0] = u;
            e[1] = isl[u];
            e[2] = (int32_t)ch[u].size();
            e[3] = (u == h->ext_root) ? -1 : vsl[u];
            e[4] = h->ext_ident[u];
            e[5] = usrc[u];
            if (usrc[u] >= 0) free_slots.push_back(usrc[u]);
            // internal children first, tips last (the kernel keeps register-prefetched
            // operands for the first two children only)
            std::vector<int32_t> och(ch[u]);
            std::stable_partition(och.begin(), och.end(), [&](int x) { return !ch[x].empty(); });
            for (size_t c = 0; c < och.size(); ++c) {
                const int cn = och[c];
                e[6 + 4 * c] = cn;
                e[7 + 4 * c] = ref_of(cn);
                e[8 + 4 * c] = vsl[cn];
                int dst = -1;
                if (!ch[cn].empty()) {
                    if (is_proot[cn]) {
                        dst = U_SPILL;
                        ++spills;
                    } else if (cn == next) {
                        dst = U_NEXT;
                    } else if (!free_slots.empty()) {
                        dst = free_slots.back();
                        free_slots.pop_back();
                    } else {
                        dst = U_SPILL;
                        ++spills;
                    }
                    usrc[cn] = dst;
                }
                e[9 + 4 * c] = dst;
            }
            o.osched.insert(o.osched.end(), e, e + OSCHED_STRIDE);
        }
    }
    o.n_parts = np;
    o.n_uspill = spills;
}

// Host-side check of a checkpointed layout before anything reaches the
// device: every D index, V slot and tip row within its buffer; the top part's
// entries use persistent slots only; each part's entries reference only the
// part's own nodes (the segment is shared by the parts in turn); every branch
// in exactly one adjoint group.  Returns an empty string when valid.
std::string check_ck_layout(const hy_engine_s *h, const Schedules &o) {
    char msg[256];
    const int np = o.n_parts;
    if ((int)o.parts.size() != 4 * (np + 1) || (int)o.bgroup.size() != np + 2) return "part tables";
    const auto &ch = h->ext_children;
    const int n_ext = (int)ch.size();
    auto ref_ok = [&](int ref, bool top) {
        const int kind = ref >> 28, idx = ref & REF_MASK;
        if (kind == KIND_TIP) return idx >= 0 && idx <= h->num_taxa;
        return idx >= 0 && idx < (top ? o.n_dpersist : o.n_dslot);
    };
    for (int p = 0; p <= np; ++p) {
        const bool top = p == np;
        std::vector<uint8_t> mine(n_ext, 0);
        const int i0 = o.parts[4 * p], ni = o.parts[4 * p + 1], q0 = o.parts[4 * p + 2], nq = o.parts[4 * p + 3];
        if (i0 < 0 || ni < 0 || (size_t)(i0 + ni) * SCHED_STRIDE > o.sched.size() || q0 < 0 || nq < 0 ||
            (size_t)(q0 + nq) * OSCHED_STRIDE > o.osched.size())
            return "part ranges";
        for (int k = i0; k < i0 + ni; ++k) {
            const int32_t *e = &o.sched[(size_t)k * SCHED_STRIDE];
            if (e[0] < 0 || e[0] >= n_ext) return "inside node";
            mine[e[0]] = 1;
        }
        for (int k = i0; k < i0 + ni; ++k) {
            const int32_t *e = &o.sched[(size_t)k * SCHED_STRIDE];
            if (e[1] < 0 || e[1] >= o.n_dslot || (top && e[1] >= o.n_dpersist) || e[2] < 1 || e[2] > MAXC) {
                snprintf(msg, sizeof(msg), "inside entry %d of part %d", k, p);
                return msg;
            }
            for (int c = 0; c < e[2]; ++c) {
                const int cn = e[6 + 3 * c], ref = e[7 + 3 * c];
                if (!ref_ok(ref, top)) return "inside child reference";
                if ((ref >> 28) == KIND_INT && !top && !mine[cn]) return "inside child outside its part";
            }
        }
        for (int k = q0; k < q0 + nq; ++k) {
            const int32_t *e = &o.osched[(size_t)k * OSCHED_STRIDE];
            if (e[0] < 0 || e[0] >= n_ext || !mine[e[0]]) return "outside node";
            if (e[1] < 0 || e[1] >= o.n_dslot || (top && e[1] >= o.n_dpersist)) return "outside D index";
            if (e[0] != h->ext_root && (e[3] < 0 || e[3] >= o.n_vslot)) return "outside V slot";
            for (int c = 0; c < e[2]; ++c) {
                const int cn = e[6 + 4 * c], ref = e[7 + 4 * c], vs = e[8 + 4 * c];
                if (!ref_ok(ref, false) || vs < 0 || vs >= (top ? o.n_vpersist : o.n_vslot))
                    return "outside child reference or V slot";
                if ((ref >> 28) == KIND_INT && !top && !mine[cn]) return "outside child outside its part";
            }
        }
    }
    std::vector<int> seen(h->n_branch, 0);
    for (int32_t bi : o.bsel) {
        if (bi < 0 || bi >= h->n_branch) return "branch selection";
        ++seen[bi];
    }
    for (int bi = 0; bi < h->n_branch; ++bi) {
        if (seen[bi] != 1) return "branch not in exactly one group";
        const int c = h->blist[(size_t)bi * BR_STRIDE];
        if (o.vsl[c] < 0 || o.vsl[c] >= o.n_vslot || (!ch[c].empty() && (o.islot[c] < 0 || o.islot[c] >= o.n_dslot)))
            return "branch slots";
    }
    return std::string();
}

void build_schedules(hy_engine h, int T, int esz, bool split) {
    Schedules o;
    make_schedules(h, T, esz, split, h->lean, h->lean_out, o, h->ck);
    if (h->ck) h->ck_error = check_ck_layout(h, o);
    h->sched = std::move(o.sched);
    h->osched = std::move(o.osched);
    h->parts = std::move(o.parts);
    h->n_parts = o.n_parts;
    h->n_uspill = o.n_uspill;
    if (h->ck) {
        h->ck_dslots = o.n_dslot;
        h->ck_vslots = o.n_vslot;
        h->ck_groups = std::move(o.bgroup);
        h->ck_bsel = std::move(o.bsel);
        // branch list entries [1] (V slot) and [2] (child reference) in the
        // checkpointed layout, same order as the branch list
        h->blist_ck = h->blist;
        for (int bi = 0; bi < h->n_branch; ++bi) {
            int32_t *e = &h->blist_ck[(size_t)bi * BR_STRIDE];
            const int c = e[0];
            e[1] = o.vsl[c];
            if ((e[2] >> 28) == KIND_INT) e[2] = (KIND_INT << 28) | o.islot[c];
        }
    }
    h->sched_key = T * 16 + esz + (h->lean ? 4096 : 0) + (split ? 8192 : 0) + (h->lean_out ? 16384 : 0) +
                   (h->ck ? 32768 : 0);
}


void refresh_blist_classes(hy_engine h) {
    for (int bi = 0; bi < h->n_branch; ++bi) {
        const int node = h->blist[bi * BR_STRIDE];
        h->blist[bi * BR_STRIDE + 3] = h->branch_class.empty() ? 0 : h->branch_class[node];
    }
}

// Internal flag of run_eval: the outside pass without the adjoint stages
// (workspace read-back of inside / outside partials).
constexpr unsigned RUN_OUTSIDE = 1u << 30;

// Outside pass: the top part first (it sends the messages into the subtree
// parts through HBM), then the subtree parts.
template <typename R>
void launch_outside_parts(hy_engine h, PruneArgs<R> pr, bool rev, hipStream_t s) {
    if (h->n_parts > 0) {
        pr.parts = (const int4 *)h->d_parts.p;
        pr.part0 = h->n_parts;
        // the top part is one round of workgroups: the default variant (its
        // LDS slots cover the lean schedule's), the parts the lean one
        launch_outside<R>(pr, h->T, rev, false, 1, s);
        pr.part0 = 0;
        launch_outside<R>(pr, h->T, rev, h->lean_out, h->n_parts, s);
    } else {
        pr.parts = nullptr;
        launch_outside<R>(pr, h->T, rev, h->lean_out, 1, s);
    }
}

template <typename R>
int run_eval(hy_engine h, unsigned flags, const double *custom_pi) {
    const int ri = sizeof(R) == 8 ? 1 : 0;
    if (h->T > max_tile(ri ? HY_PREC_FP64 : HY_PREC_FP32))
        return fail(h, HY_ERR_STATE, "internal: %d-pattern tiles are not built for this precision", h->T);
    const bool grad = (flags & HY_WANT_GRAD) != 0;
    h->part_ri = ri;
    h->part_outside = grad || (flags & RUN_OUTSIDE) != 0;
    hipStream_t s = h->stream;
    const bool prof = h->profile;
    if (prof) HIP_TRY(h, hipEventRecord(h->ev[0], s));

    const double *params = (const double *)h->d_params.p;
    if (!custom_pi) {
        ModelArgs ma{};
        ma.qdesc = (const int32_t *)h->d_qdesc.p;
        ma.nd = (const int8_t *)h->d_nd.p;
        ma.dpos = (const int8_t *)h->d_dpos.p;
        ma.dfrom = (const int8_t *)h->d_dfrom.p;
        ma.dto = (const int8_t *)h->d_dto.p;
        ma.syn = (const uint8_t *)h->d_syn.p;
        ma.pos = (const double *)h->d_pos.p;
        ma.pi = (const double *)h->d_pi.p;
        ma.S = h->S;
        ma.params = params;
        ma.pl = h->pl;
        ma.lam = (double *)h->d_lam.p;
        ma.U = (double *)h->d_U.p;
        ma.V = (double *)h->d_V.p;
        ma.mu = (double *)h->d_mu.p;
        ma.Z = (double *)h->d_Z.p;
        ma.warm = (h->warm_enable && h->warm_valid) ? 1 : 0;
        ma.status = h->h_status_dev;
        if (h->zc_params) {
            ma.params_host = h->h_params_dev;
            ma.params_dev = (double *)h->d_params.p;
            ma.n_params = h->pl.total;
        }
        ma.mat_gene = (const int32_t *)h->d_mat_gene.p;
        // accepted predicted eigenvector correction per precision
        ma.tol = sizeof(R) == 4 ? 1e-10 : 1e-14;
        launch_model(ma, h->n_mat, s);
        // M_phi = V dQ_phi U and d mu / d phi depend on the spectra only: they
        // are built by the workgroups of the P-build launch beyond the branches
        ProjArgs pj{};
        if (grad) {
            pj.qdesc = (const int32_t *)h->d_qdesc.p;
            pj.nd = (const int8_t *)h->d_nd.p;
            pj.dpos = (const int8_t *)h->d_dpos.p;
            pj.dfrom = (const int8_t *)h->d_dfrom.p;
            pj.dto = (const int8_t *)h->d_dto.p;
            pj.syn = (const uint8_t *)h->d_syn.p;
            pj.pos = (const double *)h->d_pos.p;
            pj.pi = (const double *)h->d_pi.p;
            pj.S = h->S;
            pj.params = params;
            pj.pl = h->pl;
            pj.U = (const double *)h->d_U.p;
            pj.V = (const double *)h->d_V.p;
            pj.M = (double *)h->d_M.p;
            pj.dmu = (double *)h->d_dmu.p;
            pj.mat_gene = (const int32_t *)h->d_mat_gene.p;
        }
        if (prof) HIP_TRY(h, hipEventRecord(h->ev[1], s));
        PmatArgs<R> pa{};
        pa.blist = (const int32_t *)h->d_blist.p;
        pa.n_branch = h->n_branch;
        pa.comp_ptr = (const int32_t *)h->d_comp_ptr.p;
        pa.comp_mat = (const int32_t *)h->d_comp_mat.p;
        pa.convention = h->convention;
        pa.params = params;
        pa.pl = h->pl;
        pa.lam = (const double *)h->d_lam.p;
        pa.U = (const double *)h->d_U.p;
        pa.V = (const double *)h->d_V.p;
        pa.mu = (const double *)h->d_mu.p;
        pa.S = h->S;
        pa.tau = (double *)h->d_tau.p;
        pa.Prow = (R *)h->d_Prow[ri].p;
        pa.PA = (R *)h->d_PA[ri].p;
        pa.PTA = (R *)h->d_PTA[ri].p;
        pa.sink = (const double *)h->d_sink.p;
        launch_pmat<R>(pa, grad ? &pj : nullptr, h->n_mat, s);
    } else {
        if (prof) HIP_TRY(h, hipEventRecord(h->ev[1], s));
        PmatArgs<R> pa{};
        pa.blist = (const int32_t *)h->d_blist.p;
        pa.n_branch = h->n_branch;
        pa.S = h->S;
        pa.Prow = (R *)h->d_Prow[ri].p;
        pa.PA = (R *)h->d_PA[ri].p;
        pa.PTA = (R *)h->d_PTA[ri].p;
        if (h->n_branch > 0) launch_pmat_host<R>(pa, (const double *)h->d_Phost.p, s);
    }
    if (prof) HIP_TRY(h, hipEventRecord(h->ev[2], s));

    if (!h->d_sched.p || !h->d_osched.p || h->sched.empty() || h->osched.empty())
        return fail(h, HY_ERR_STATE, "internal: pruning schedules not on the device");
    PruneArgs<R> pr{};
    pr.sched = (const int32_t *)h->d_sched.p;
    pr.n_sched = (int)(h->sched.size() / SCHED_STRIDE);
    pr.osched = (const int32_t *)h->d_osched.p;
    pr.n_osched = (int)(h->osched.size() / OSCHED_STRIDE);
    pr.tip_rows = h->num_taxa + 1;
    pr.S = h->S;
    pr.Uspill = (R *)h->d_Uspill[ri].p;
    pr.PA = (const R *)h->d_PA[ri].p;
    pr.PTA = (const R *)h->d_PTA[ri].p;
    pr.Prow = (const R *)h->d_Prow[ri].p;
    pr.tips = (const uint16_t *)h->d_tips.p;
    pr.amb = (const uint64_t *)h->d_amb.p;
    pr.P_pad = h->P_pad;
    pr.n_tiles = h->n_tiles;
    pr.n_patterns = h->P;
    pr.pi = custom_pi ? (const double *)h->d_pi_custom.p : (const double *)h->d_pi.p;
    pr.D = (R *)h->d_D[ri].p;
    pr.E = (int32_t *)h->d_E.p;
    pr.V = (R *)h->d_Vp[ri].p;
    pr.G = (int32_t *)h->d_G.p;
    pr.Lroot = (double *)h->d_Lroot.p;
    pr.Eroot = (int32_t *)h->d_Eroot.p;
    pr.site_ll = (double *)h->d_site.p;
    pr.genes = (const GeneInfo *)h->d_genes.p;
    pr.tile_gene = (const int32_t *)h->d_tile_gene.p;
    pr.root_out = 1;
    if (h->ck && h->n_parts > 0) {
        // checkpointed: the parts one after another (they share the segment),
        // each leaving its root's partials; then the top part
        pr.parts = (const int4 *)h->d_parts.p;
        pr.root_out = 0;
        for (int p = 0; p < h->n_parts; ++p) {
            pr.part0 = p;
            launch_inside<R>(pr, h->T, false, 1, s);
        }
        pr.part0 = h->n_parts;
        pr.root_out = 1;
    } else if (h->n_parts > 0) {
        // subtree parts, then the top part (root likelihood)
        pr.parts = (const int4 *)h->d_parts.p;
        pr.part0 = 0;
        pr.root_out = 0;
        launch_inside<R>(pr, h->T, h->lean, h->n_parts, s);
        pr.part0 = h->n_parts;
        pr.root_out = 1;
    }
    // with parts, the top part (one round of workgroups) runs the default
    // variant, whose LDS slots cover the lean schedule's
    launch_inside<R>(pr, h->T, h->lean && h->n_parts == 0, 1, s);
    if (prof) HIP_TRY(h, hipEventRecord(h->ev[3], s));
    if (grad) {
        const bool rev = custom_pi == nullptr && !h->has_sink;
        AdjArgs<R> aa{};
        aa.blist = (const int32_t *)(h->ck ? h->d_blist_ck.p : h->d_blist.p);
        aa.n_branch = h->n_branch;
        aa.n_chunks = h->n_chunks;
        aa.tiles_per_chunk = h->tiles_per_chunk;
        aa.D = pr.D;
        aa.V = pr.V;
        aa.E = pr.E;
        aa.G = pr.G;
        aa.Eroot = pr.Eroot;
        aa.Lroot = pr.Lroot;
        aa.weights = h->wover ? h->wover : (const double *)h->d_weights.p;
        aa.tips = pr.tips;
        aa.amb = pr.amb;
        aa.n_amb = h->n_amb;
        aa.P_pad = h->P_pad;
        aa.n_tiles = h->n_tiles;
        aa.S = h->S;
        aa.Apart = (double *)h->d_Apart.p;
        aa.genes = (const GeneInfo *)h->d_genes.p;
        if (h->ck && h->n_parts > 0) {
            // checkpointed: outside pass of the top part and the adjoint of the
            // branches below it; then per part: its inside partials again
            // (into the segment), its outside pass, the adjoint of its branches
            auto adjoint_group = [&](int g) {
                const int b0 = h->ck_groups[g], nbg = h->ck_groups[g + 1] - b0;
                if (nbg <= 0) return;
                AdjArgs<R> ag = aa;
                ag.bsel = (const int32_t *)h->d_bsel.p + b0;
                ag.n_branch = nbg;
                launch_adjoint<R>(ag, h->T, s);
            };
            PruneArgs<R> po = pr;
            po.parts = (const int4 *)h->d_parts.p;
            po.part0 = h->n_parts;
            launch_outside<R>(po, h->T, rev, false, 1, s);
            adjoint_group(0);
            for (int p = 0; p < h->n_parts; ++p) {
                PruneArgs<R> pi = pr;
                pi.parts = (const int4 *)h->d_parts.p;
                pi.part0 = p;
                pi.root_out = 0;
                launch_inside<R>(pi, h->T, false, 1, s);
                po.part0 = p;
                launch_outside<R>(po, h->T, rev, false, 1, s);
                adjoint_group(p + 1);
            }
            if (prof) HIP_TRY(h, hipEventRecord(h->ev[4], s));
        } else {
            launch_outside_parts<R>(h, pr, rev, s);
            if (prof) HIP_TRY(h, hipEventRecord(h->ev[4], s));
            if (h->n_branch > 0) launch_adjoint<R>(aa, h->T, s);
        }
        if (prof) HIP_TRY(h, hipEventRecord(h->ev[5], s));
        FrechetArgs fa{};
        fa.blist = (const int32_t *)h->d_blist.p;
        fa.n_branch = h->n_branch;
        fa.n_chunks = h->n_chunks;
        fa.max_comp = h->max_comp;
        fa.comp_ptr = (const int32_t *)h->d_comp_ptr.p;
        fa.comp_mat = (const int32_t *)h->d_comp_mat.p;
        fa.params = params;
        fa.pl = h->pl;
        fa.Apart = (const double *)h->d_Apart.p;
        fa.lam = (const double *)h->d_lam.p;
        fa.U = (const double *)h->d_U.p;
        fa.V = (const double *)h->d_V.p;
        fa.M = (const double *)h->d_M.p;
        fa.tau = (const double *)h->d_tau.p;
        fa.out = (double *)h->d_fr.p;
        fa.n_mat = h->n_mat;
        fa.grp_size = h->grp_size;
        fa.n_grp = h->n_grp;
        fa.wl_ptr = (const int32_t *)h->d_wl.p;
        fa.wl_idx = (const int32_t *)h->d_wl.p + (h->n_mat + 1);
        fa.proj = (double *)h->d_proj.p;
        fa.mat_gene = (const int32_t *)h->d_mat_gene.p;
        if (h->d_proj.bytes < (size_t)h->n_mat * 6 * h->n_grp * 8 ||
            h->d_Apart.bytes < (size_t)std::max(1, h->n_branch) * h->n_chunks * 4096 * 8)
            return fail(h, HY_ERR_STATE, "internal: Frechet stage buffers smaller than the grouping needs");
        if (h->n_branch > 0) launch_frechet(fa, s);
        if (!h->sink_wl.empty()) {
            if (!h->d_sink.p) return fail(h, HY_ERR_STATE, "internal: error-sink matrix not on the device");
            SinkArgs sa{};
            sa.Apart = fa.Apart;
            sa.n_chunks = h->n_chunks;
            sa.max_comp = h->max_comp;
            sa.slots = (const int32_t *)h->d_sinkwl.p;
            sa.sink = (const double *)h->d_sink.p;
            sa.out = fa.out;
            launch_sink(sa, (int)h->sink_wl.size(), s);
        }
        if (prof) HIP_TRY(h, hipEventRecord(h->ev[6], s));
    } else {
        if (flags & RUN_OUTSIDE) launch_outside_parts<R>(h, pr, custom_pi == nullptr && !h->has_sink, s);
        if (prof)
            for (int k = 4; k <= 6; ++k) HIP_TRY(h, hipEventRecord(h->ev[k], s));
    }
    FinalArgs fn{};
    fn.site_ll = (const double *)h->d_site.p;
    fn.weights = h->wover ? h->wover : (const double *)h->d_weights.p;
    fn.Lroot = (const double *)h->d_Lroot.p;
    fn.n_patterns = h->P;
    fn.want_grad = grad ? 1 : 0;
    fn.blist = (const int32_t *)h->d_blist.p;
    fn.n_branch = h->n_branch;
    fn.max_comp = h->max_comp;
    fn.n_class = h->n_class;
    fn.n_real_nodes = h->num_nodes;
    fn.comp_ptr = (const int32_t *)h->d_comp_ptr.p;
    fn.comp_mat = (const int32_t *)h->d_comp_mat.p;
    fn.convention = h->convention;
    fn.params = params;
    fn.pl = h->pl;
    fn.fr = (const double *)h->d_fr.p;
    fn.tau = (const double *)h->d_tau.p;
    fn.mu = (const double *)h->d_mu.p;
    fn.dmu = (const double *)h->d_dmu.p;
    fn.proj = (const double *)h->d_proj.p;
    fn.proj_stride = h->n_grp;
    fn.res = h->h_res_dev;
    fn.genes = (const GeneInfo *)h->d_genes.p;
    // a single evaluation publishes its call number for eval_complete's wait
    fn.done = (h->publish_done && h->n_genes <= 1 && !fn.genes && !prof) ? h->h_done_dev : nullptr;
    launch_final(fn, h->n_genes, s);
    HIP_TRY(h, hipGetLastError());
    if (prof) HIP_TRY(h, hipEventRecord(h->ev[7], s));
    return HY_OK;
}

}  // namespace

namespace {

hipError_t make_stream_set(int device, std::shared_ptr<StreamSet> &out) {
    auto ss = std::make_shared<StreamSet>();
    ss->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ss->s1, hipStreamNonBlocking);
    if (e == hipSuccess) out = ss;
    return e;
}

hy_engine new_engine(int device, int n_cu, const std::shared_ptr<StreamSet> &ss) {
    hy_engine h = new hy_engine_s();
    h->device = device;
    h->n_cu = n_cu > 0 ? n_cu : 256;
    h->streams = ss;
    h->stream = ss->s1;
    const char *pe = std::getenv("HY_PROFILE");
    h->profile = pe && pe[0] == '1';
    const char *ng = std::getenv("HY_NO_GRAPH");
    h->use_graphs = !(ng && ng[0] == '1');
    const char *np = std::getenv("HY_NO_PARTS");
    h->no_parts = np && np[0] == '1';
    if (h->profile)
        for (auto &ev : h->ev) (void)hipEventCreate(&ev);
    // fixed scratch of the allocation-free entry point on this device
    h->scratch_err = noalloc_scratch(device, &h->scratch);
    if (h->scratch_err != hipSuccess) h->scratch = nullptr;
    return h;
}

}  // namespace

extern "C" {

int hy_create(int device, hy_engine *out) {
    if (!out) return fail(nullptr, HY_ERR_ARG, "hy_create: out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return fail(nullptr, HY_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= n) return fail(nullptr, HY_ERR_ARG, "device %d out of range (%d devices)", device, n);
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return fail(nullptr, HY_ERR_HIP, "hipGetDeviceProperties: %s", hipGetErrorString(e));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(nullptr, HY_ERR_UNSUPPORTED, "device %d is %s; this library is built for gfx950 (MI355X)", device,
                    prop.gcnArchName);
    std::shared_ptr<StreamSet> ss;
    e = make_stream_set(device, ss);
    if (e != hipSuccess) return fail(nullptr, HY_ERR_HIP, "stream creation failed: %s", hipGetErrorString(e));
    *out = new_engine(device, prop.multiProcessorCount, ss);
    return HY_OK;
}

int hy_create_group(int device, int n, int n_stream_sets, hy_engine *out) {
    if (!out || n < 1) return fail(nullptr, HY_ERR_ARG, "hy_create_group: need n >= 1 and an output array");
    for (int i = 0; i < n; ++i) out[i] = nullptr;
    int nd = 0;
    hipError_t e = hipGetDeviceCount(&nd);
    if (e != hipSuccess || nd == 0)
        return fail(nullptr, HY_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= nd) return fail(nullptr, HY_ERR_ARG, "device %d out of range (%d devices)", device, nd);
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return fail(nullptr, HY_ERR_HIP, "hipGetDeviceProperties: %s", hipGetErrorString(e));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(nullptr, HY_ERR_UNSUPPORTED, "device %d is %s; this library is built for gfx950 (MI355X)", device,
                    prop.gcnArchName);
    const int ns = std::max(1, std::min(n, n_stream_sets > 0 ? n_stream_sets : 8));
    std::vector<std::shared_ptr<StreamSet>> sets(ns);
    for (auto &ss : sets) {
        e = make_stream_set(device, ss);
        if (e != hipSuccess) return fail(nullptr, HY_ERR_HIP, "stream creation failed: %s", hipGetErrorString(e));
    }
    // the context engine that evaluates the group's merged problem
    std::shared_ptr<StreamSet> cs;
    e = make_stream_set(device, cs);
    if (e != hipSuccess) return fail(nullptr, HY_ERR_HIP, "stream creation failed: %s", hipGetErrorString(e));
    auto gs = std::make_shared<GroupState>();
    gs->ctx = new_engine(device, prop.multiProcessorCount, cs);
    const char *nf = std::getenv("HY_NO_FUSE");
    gs->no_fuse = nf && nf[0] == '1';
    for (int i = 0; i < n; ++i) {
        out[i] = new_engine(device, prop.multiProcessorCount, sets[i % ns]);
        // Grouped engines evaluated one by one launch eagerly: thousands of
        // instantiated graphs (each holding runtime resources) are not kept
        // per gene.  Batches of a group run fused (one graph).
        out[i]->use_graphs = false;
        out[i]->group = gs;
    }
    return HY_OK;
}

void hy_destroy(hy_engine h) {
    if (!h) return;
    if (h->group) h->group->members.clear();  // the merged problem no longer matches
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    DevBuf *bufs[] = {&h->d_genes, &h->d_tile_gene, &h->d_mat_gene, &h->d_qdesc, &h->d_nd, &h->d_dpos, &h->d_dfrom, &h->d_dto, &h->d_syn, &h->d_pos, &h->d_pi,
                      &h->d_pi_custom, &h->d_Phost, &h->d_sched, &h->d_osched, &h->d_parts, &h->d_blist, &h->d_tips, &h->d_amb, &h->d_weights,
                      &h->d_comp_ptr, &h->d_comp_mat, &h->d_params, &h->d_lam, &h->d_U, &h->d_V, &h->d_mu,
                      &h->d_tau, &h->d_M, &h->d_dmu, &h->d_Z, &h->d_E, &h->d_G, &h->d_Lroot, &h->d_Eroot,
                      &h->d_site, &h->d_Apart, &h->d_fr, &h->d_wl, &h->d_proj, &h->d_wover, &h->d_wuser,
                      &h->d_sink, &h->d_sinkwl, &h->d_blist_ck, &h->d_bsel, &h->d_gsched, &h->d_gnref,
                      &h->d_gnbr, &h->d_gtips, &h->d_gw, &h->d_gP, &h->d_gpi, &h->d_gD, &h->d_gE, &h->d_gU,
                      &h->d_gH, &h->d_gL, &h->d_gLE, &h->d_gsite, &h->d_gApart, &h->d_gA, &h->d_gres,
                      &h->d_fZ, &h->d_flam, &h->d_fU, &h->d_fV, &h->d_fmu, &h->d_fstatus, &h->d_fstage, &h->d_fbl,
                      &h->d_fD, &h->d_fE, &h->d_fout, &h->d_fident, &h->d_sprm, &h->d_slam, &h->d_sU, &h->d_sV,
                      &h->d_smu, &h->d_sZ, &h->d_sstatus, &h->d_bP, &h->d_bL, &h->d_be, &h->d_blnl,
                      &h->d_pstage};
    for (DevBuf *b : bufs) release(h, *b);
    for (int k = 0; k < 2; ++k) {
        release(h, h->d_Prow[k]);
        release(h, h->d_PA[k]);
        release(h, h->d_PTA[k]);
        release(h, h->d_D[k]);
        release(h, h->d_Vp[k]);
        release(h, h->d_Uspill[k]);
    }
    if (h->h_params) (void)hipHostFree(h->h_params);
    if (h->h_res) (void)hipHostFree(h->h_res);
    if (h->h_status) (void)hipHostFree(h->h_status);
    if (h->h_done) (void)hipHostFree(h->h_done);
    if (h->h_site) (void)hipHostFree(h->h_site);
    if (h->f_pin) (void)hipHostFree(h->f_pin);
    if (h->p_pin) (void)hipHostFree(h->p_pin);
    if (h->profile)
        for (auto &ev : h->ev) (void)hipEventDestroy(ev);
    for (auto &g : h->graphs) destroy_graph_exec(g.exec);
    delete h;
}

const char *hy_last_error(hy_engine h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int hy_set_model(hy_engine h, int code_id, const double *pos_freqs, const double *pi) {
    if (h) ++h->epoch;  // captured graphs are stale
    if (!h) return HY_ERR_ARG;
    if (!pos_freqs || !pi) return fail(h, HY_ERR_ARG, "hy_set_model: NULL frequencies");
    HIP_TRY(h, hipSetDevice(h->device));
    std::shared_ptr<const hyphy::core::GeneticCode> gc;
    try {
        gc = hyphy::core::GeneticCode::from_id(code_id);
    } catch (const std::exception &ex) {
        return fail(h, HY_ERR_ARG, "%s", ex.what());
    }
    const int S = gc->num_sense_codons;
    if (h->have_problem && S != h->S)
        return fail(h, HY_ERR_ARG, "genetic code %d has %d sense codons but the problem was set with S=%d", code_id, S,
                    h->S);
    std::vector<int8_t> nd(4096, 0), dpos(4096 * 3, -1), dfrom(4096 * 3, -1), dto(4096 * 3, -1);
    std::vector<uint8_t> syn(4096, 0);
    std::vector<int32_t> qdesc(4096, 0);
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j) {
            if (i == j) continue;
            const auto &d = gc->diff_matrix[i][j];
            const int k = i * 64 + j;
            nd[k] = d.num_diffs;
            syn[k] = d.synonymous ? 1 : 0;
            int32_t qd = (d.num_diffs & 3) | (d.synonymous ? QDESC_SYN : 0);
            for (int q = 0; q < 3; ++q) {
                dpos[3 * k + q] = d.diff_pos[q];
                dfrom[3 * k + q] = d.diff_from[q];
                dto[3 * k + q] = d.diff_to[q];
                if (q < d.num_diffs)
                    qd |= (qdesc_theta_index(d.diff_from[q], d.diff_to[q]) | ((d.diff_pos[q] * 4 + d.diff_to[q]) << 3))
                          << (3 + 7 * q);
            }
            qdesc[k] = qd;
        }
    std::vector<double> pp(64, 0.0);
    for (int i = 0; i < S; ++i) {
        if (!(pi[i] >= 0.0)) return fail(h, HY_ERR_ARG, "hy_set_model: pi[%d] is negative or NaN", i);
        pp[i] = pi[i];
    }
    for (int k = 0; k < 12; ++k)
        if (!(pos_freqs[k] >= 0.0)) return fail(h, HY_ERR_ARG, "hy_set_model: pos_freqs[%d] is negative or NaN", k);
    TRY(upload(h, h->d_nd, nd.data(), nd.size()));
    TRY(upload(h, h->d_dpos, dpos.data(), dpos.size()));
    TRY(upload(h, h->d_dfrom, dfrom.data(), dfrom.size()));
    TRY(upload(h, h->d_dto, dto.data(), dto.size()));
    TRY(upload(h, h->d_syn, syn.data(), syn.size()));
    TRY(upload(h, h->d_qdesc, qdesc.data(), qdesc.size() * 4));
    TRY(upload(h, h->d_pos, pos_freqs, 12 * sizeof(double)));
    TRY(upload(h, h->d_pi, pp.data(), 64 * sizeof(double)));
    // a sink matrix is indexed by the sense codons of the code it was built for
    if (h->have_model && code_id != h->code_id) h->sink_set = false;
    h->code_id = code_id;
    h->S = S;
    h->pos_host.assign(pos_freqs, pos_freqs + 12);
    h->pi_host = pp;
    h->have_model = true;
    h->warm_valid = false;
    ++h->model_gen;
    return HY_OK;
}

int hy_set_problem(hy_engine h, int num_nodes, const int32_t *parent, const int32_t *child_ptr,
                   const int32_t *child_idx, const int32_t *leaf_taxon, int num_taxa, int num_patterns,
                   const int8_t *states, const uint64_t *masks, const double *weights, int S) {
    if (h) ++h->epoch;  // captured graphs are stale
    if (!h) return HY_ERR_ARG;
    HIP_TRY(h, hipSetDevice(h->device));
    if (num_nodes < 1 || !parent || !child_ptr || !leaf_taxon)
        return fail(h, HY_ERR_ARG, "hy_set_problem: empty tree or NULL topology arrays");
    if (num_patterns < 0 || num_taxa < 0 || S < 1 || S > 64)
        return fail(h, HY_ERR_ARG, "hy_set_problem: bad sizes (patterns=%d taxa=%d S=%d)", num_patterns, num_taxa, S);
    if (num_patterns > 0 && (!states || !masks || !weights))
        return fail(h, HY_ERR_ARG, "hy_set_problem: NULL pattern arrays");
    // --- topology checks ---
    int root = -1;
    for (int n = 0; n < num_nodes; ++n) {
        if (parent[n] < 0) {
            if (root >= 0) return fail(h, HY_ERR_ARG, "hy_set_problem: more than one root (%d and %d)", root, n);
            root = n;
        } else if (parent[n] >= num_nodes) {
            return fail(h, HY_ERR_ARG, "hy_set_problem: parent[%d]=%d out of range", n, parent[n]);
        }
    }
    if (root < 0) return fail(h, HY_ERR_ARG, "hy_set_problem: no root (parent == -1)");
    if (child_ptr[0] != 0) return fail(h, HY_ERR_ARG, "hy_set_problem: child_ptr[0] must be 0");
    std::vector<std::vector<int32_t>> ch(num_nodes);
    for (int n = 0; n < num_nodes; ++n) {
        if (child_ptr[n + 1] < child_ptr[n]) return fail(h, HY_ERR_ARG, "hy_set_problem: child_ptr not monotone");
        for (int k = child_ptr[n]; k < child_ptr[n + 1]; ++k) {
            const int c = child_idx[k];
            if (c < 0 || c >= num_nodes || parent[c] != n)
                return fail(h, HY_ERR_ARG, "hy_set_problem: child %d of node %d inconsistent with parent[]", c, n);
            ch[n].push_back(c);
        }
    }
    if (child_ptr[num_nodes] != num_nodes - 1)
        return fail(h, HY_ERR_ARG, "hy_set_problem: %d child links for %d nodes (a tree has n-1)", child_ptr[num_nodes],
                    num_nodes);
    for (int n = 0; n < num_nodes; ++n)
        if (ch[n].empty() && (leaf_taxon[n] < -1 || leaf_taxon[n] >= num_taxa))
            return fail(h, HY_ERR_ARG, "hy_set_problem: leaf_taxon[%d]=%d out of range", n, leaf_taxon[n]);
    // --- resolve polytomies wider than MAXC with identity edges ---
    std::vector<int32_t> ext_parent(parent, parent + num_nodes);
    std::vector<uint8_t> ident(num_nodes, 0);
    for (size_t u = 0; u < ch.size(); ++u) {
        if ((int)ch[u].size() <= MAXC) continue;
        const int a = (int)ch.size();
        std::vector<int32_t> rest(ch[u].begin() + (MAXC - 1), ch[u].end());
        ch[u].resize(MAXC - 1);
        ch[u].push_back(a);
        ch.push_back(rest);
        ext_parent.push_back((int32_t)u);
        ident.push_back(1);
        for (int c : ch[a]) ext_parent[c] = a;
    }
    const int n_ext = (int)ch.size();
    // --- post-order (children in order) ---
    std::vector<int32_t> post;
    post.reserve(n_ext);
    {
        std::vector<std::pair<int, int>> st;
        st.push_back({root, 0});
        std::vector<uint8_t> seen(n_ext, 0);
        while (!st.empty()) {
            auto &[node, k] = st.back();
            if (k == 0) {
                if (seen[node]) return fail(h, HY_ERR_ARG, "hy_set_problem: cycle at node %d", node);
                seen[node] = 1;
            }
            if (k < (int)ch[node].size()) {
                const int c = ch[node][k++];
                st.push_back({c, 0});
            } else {
                post.push_back(node);
                st.pop_back();
            }
        }
        if ((int)post.size() != n_ext) return fail(h, HY_ERR_ARG, "hy_set_problem: tree is not connected");
    }
    // internal nodes get a global D index in post-order; non-root nodes a V slot
    std::vector<int32_t> int_slot(n_ext, -1), vslot(n_ext, -1);
    int n_int = 0, n_v = 0;
    for (int u : post) {
        if (!ch[u].empty()) int_slot[u] = n_int++;
        if (u != root) vslot[u] = n_v++;
    }
    auto ref_of = [&](int c) -> int32_t {
        if (ch[c].empty()) {
            const int row = (c < num_nodes && leaf_taxon[c] >= 0) ? leaf_taxon[c] : num_taxa;
            return (KIND_TIP << 28) | row;
        }
        return (KIND_INT << 28) | int_slot[c];
    };
    // internal-node counts of subtrees (visit larger subtrees first)
    std::vector<int> sub(n_ext, 0);
    for (int u : post) {
        int s = ch[u].empty() ? 0 : 1;
        for (int c : ch[u]) s += sub[c];
        sub[u] = s;
    }
    std::vector<int32_t> order;  // internal nodes, post-order, larger subtrees first
    {
        std::vector<std::pair<int, int>> st{{root, 0}};
        std::vector<std::vector<int32_t>> vis(n_ext);
        while (!st.empty()) {
            auto &[node, k] = st.back();
            if (k == 0) {
                vis[node] = ch[node];
                std::stable_sort(vis[node].begin(), vis[node].end(), [&](int x, int y) { return sub[x] > sub[y]; });
            }
            if (k < (int)vis[node].size()) {
                const int c = vis[node][k++];
                if (!ch[c].empty()) st.push_back({c, 0});
            } else {
                if (!ch[node].empty()) order.push_back(node);
                st.pop_back();
            }
        }
    }
    if (order.empty()) {
        // single-node tree: the root is a leaf; prune it as an internal node with no children
        n_int = 1;
        int_slot[root] = 0;
        order.push_back(root);
    }
    std::vector<int32_t> blist;
    for (int b = 0; b < num_nodes; ++b) {
        if (b == root) continue;
        const int32_t e[BR_STRIDE] = {b, vslot[b], ref_of(b), 0, 0, 0, 0, 0};
        blist.insert(blist.end(), e, e + BR_STRIDE);
    }
    // --- tip codes ---
    std::vector<uint16_t> codes((size_t)num_taxa * std::max(num_patterns, 0));
    std::vector<uint64_t> amb;
    std::unordered_map<uint64_t, int> amb_idx;
    const uint64_t all = S >= 64 ? ~uint64_t(0) : ((uint64_t(1) << S) - 1);
    for (int p = 0; p < num_patterns; ++p)
        for (int t = 0; t < num_taxa; ++t) {
            const int8_t st = states[(size_t)p * num_taxa + t];
            const uint64_t m = masks[(size_t)p * num_taxa + t];
            uint16_t code;
            if (st >= 0 && st < S) {
                code = (uint16_t)st;
            } else if (m == 0 || (m & all) == all) {
                code = TIP_ONES;
            } else {
                auto it = amb_idx.find(m & all);
                int k;
                if (it == amb_idx.end()) {
                    k = (int)amb.size();
                    if (k >= 0x7FFF) return fail(h, HY_ERR_UNSUPPORTED, "more than 32766 distinct ambiguity masks");
                    amb_idx[m & all] = k;
                    amb.push_back(m & all);
                } else {
                    k = it->second;
                }
                code = (uint16_t)(TIP_AMB | k);
            }
            codes[(size_t)t * num_patterns + p] = code;
        }
    for (int p = 0; p < num_patterns; ++p)
        if (!(weights[p] >= 0.0)) return fail(h, HY_ERR_ARG, "hy_set_problem: weight[%d] negative or NaN", p);
    if (amb.empty()) amb.push_back(0);
    // --- commit ---
    if (h->have_model && S != h->S) {  // a model for another state count no longer applies
        h->have_model = false;
        h->sink_set = false;
    }
    h->num_nodes = num_nodes;
    h->n_ext = n_ext;
    h->root = root;
    h->num_taxa = num_taxa;
    h->P = num_patterns;
    h->S = S;
    h->n_int = n_int;
    h->n_vslot = std::max(n_v, 1);
    h->n_branch = (int)(blist.size() / BR_STRIDE);
    h->order = std::move(order);
    h->ext_children = ch;
    h->ext_ident = ident;
    h->int_slot = int_slot;
    h->vslot = vslot;
    h->ext_root = root;
    h->leaf_taxon_ext.assign(n_ext, -1);
    for (int n = 0; n < num_nodes; ++n) h->leaf_taxon_ext[n] = leaf_taxon[n];
    h->sched.clear();
    h->sched_key = 0;
    h->ident_tiles[0] = h->ident_tiles[1] = nullptr;
    h->blist = std::move(blist);
    h->real_parent.assign(parent, parent + num_nodes);
    h->codes = std::move(codes);
    h->amb = std::move(amb);
    h->weights.assign(weights, weights + num_patterns);
    h->wuser = false;  // caller pattern weights belong to the previous problem
    h->T = 0;
    h->branch_class.assign(num_nodes, 0);
    h->n_mat = 1;
    h->n_class = 1;
    h->n_comp = 1;
    h->max_comp = 1;
    h->convention = HY_TIME_RAW;
    h->comp_ptr = {0, 1};
    h->comp_mat = {0};
    // the default mixture has no error sink, and a sink matrix of the previous
    // problem (other S or code) must not be reused: hy_set_sink is needed again
    h->has_sink = false;
    h->sink_set = false;
    h->sink_wl.clear();
    ++h->epoch;
    refresh_blist_classes(h);
    h->warm_valid = false;
    TRY(upload(h, h->d_blist, h->blist.data(), std::max<size_t>(h->blist.size(), BR_STRIDE) * 4));
    TRY(upload(h, h->d_amb, h->amb.data(), h->amb.size() * 8));
    h->n_amb = (int)h->amb.size();
    TRY(upload(h, h->d_comp_ptr, h->comp_ptr.data(), h->comp_ptr.size() * 4));
    TRY(upload(h, h->d_comp_mat, h->comp_mat.data(), h->comp_mat.size() * 4));
    // host layouts of the allocation-free path: 16-pattern tiles, one
    // whole-tree inside schedule (its LDS slot count is the same in both
    // precisions)
    {
        constexpr int T = 16;
        h->na_P_pad = (h->P + T - 1) / T * T;
        h->na_tips.assign((size_t)(h->num_taxa + 1) * h->na_P_pad, TIP_ONES);
        for (int t = 0; t < h->num_taxa; ++t)
            std::memcpy(&h->na_tips[(size_t)t * h->na_P_pad], &h->codes[(size_t)t * h->P], sizeof(uint16_t) * h->P);
        h->na_weights.assign(h->na_P_pad, 0.0);
        std::copy(h->weights.begin(), h->weights.end(), h->na_weights.begin());
        Schedules o;
        make_schedules(h, T, 4, false, false, false, o);
        h->na_sched = std::move(o.sched);
        h->na_ident.clear();
        for (int n = 0; n < h->n_ext; ++n)
            if (h->ext_ident[n]) h->na_ident.push_back(n);
    }
    h->gtr_ready = false;
    h->f_ident_ready = false;
    ++h->model_gen;
    h->have_problem = true;
    return HY_OK;
}

int hy_set_mixture(hy_engine h, int n_mat, int n_class, const int32_t *branch_class, const int32_t *comp_ptr,
                   const int32_t *comp_mat, int time_convention) {
    if (h) ++h->epoch;  // captured graphs are stale
    if (!h) return HY_ERR_ARG;
    if (!h->have_problem) return fail(h, HY_ERR_STATE, "hy_set_mixture before hy_set_problem");
    if (n_mat < 1 || n_mat > 48 || n_class < 1 || n_class > 64 || !comp_ptr || !comp_mat)
        return fail(h, HY_ERR_ARG, "hy_set_mixture: bad sizes (n_mat=%d n_class=%d)", n_mat, n_class);
    if (time_convention != HY_TIME_RAW && time_convention != HY_TIME_CONV)
        return fail(h, HY_ERR_ARG, "hy_set_mixture: unknown time convention %d", time_convention);
    if (comp_ptr[0] != 0) return fail(h, HY_ERR_ARG, "hy_set_mixture: comp_ptr[0] must be 0");
    int max_comp = 0;
    for (int c = 0; c < n_class; ++c) {
        const int k = comp_ptr[c + 1] - comp_ptr[c];
        if (k < 1) return fail(h, HY_ERR_ARG, "hy_set_mixture: class %d has no component", c);
        max_comp = std::max(max_comp, k);
    }
    const int n_comp = comp_ptr[n_class];
    if (n_comp > 192) return fail(h, HY_ERR_ARG, "hy_set_mixture: at most 192 components");
    bool has_sink = false;
    for (int j = 0; j < n_comp; ++j) {
        if (comp_mat[j] == HY_SINK) {
            has_sink = true;
            continue;
        }
        if (comp_mat[j] < 0 || comp_mat[j] >= n_mat)
            return fail(h, HY_ERR_ARG, "hy_set_mixture: comp_mat[%d]=%d out of range", j, comp_mat[j]);
    }
    if (has_sink && time_convention != HY_TIME_RAW)
        return fail(h, HY_ERR_ARG, "hy_set_mixture: error-sink components need HY_TIME_RAW");
    std::vector<int32_t> bc(h->num_nodes, 0);
    if (branch_class) {
        for (int n = 0; n < h->num_nodes; ++n) {
            if (n == h->root) continue;
            if (branch_class[n] < 0 || branch_class[n] >= n_class)
                return fail(h, HY_ERR_ARG, "hy_set_mixture: branch_class[%d]=%d out of range", n, branch_class[n]);
            bc[n] = branch_class[n];
        }
    } else if (n_class != 1) {
        return fail(h, HY_ERR_ARG, "hy_set_mixture: branch_class required when n_class > 1");
    }
    h->n_mat = n_mat;
    h->n_class = n_class;
    h->n_comp = n_comp;
    h->max_comp = max_comp;
    h->convention = time_convention;
    h->comp_ptr.assign(comp_ptr, comp_ptr + n_class + 1);
    h->comp_mat.assign(comp_mat, comp_mat + n_comp);
    if (has_sink != h->has_sink) ++h->epoch;  // outside variant and sink launch change
    h->has_sink = has_sink;
    h->branch_class = bc;
    refresh_blist_classes(h);
    h->warm_valid = false;
    TRY(upload(h, h->d_blist, h->blist.data(), std::max<size_t>(h->blist.size(), BR_STRIDE) * 4));
    TRY(upload(h, h->d_comp_ptr, h->comp_ptr.data(), h->comp_ptr.size() * 4));
    TRY(upload(h, h->d_comp_mat, h->comp_mat.data(), h->comp_mat.size() * 4));
    return HY_OK;
}

int hy_set_sink(hy_engine h, const double *freqs) {
    if (!h) return HY_ERR_ARG;
    if (!h->have_problem) return fail(h, HY_ERR_STATE, "hy_set_sink before hy_set_problem");
    if (!freqs) return fail(h, HY_ERR_ARG, "hy_set_sink: NULL frequencies");
    const int S = h->S;
    for (int i = 0; i < S; ++i)
        if (!(freqs[i] >= 0.0 && freqs[i] <= 1.0))
            return fail(h, HY_ERR_ARG, "hy_set_sink: frequency %d = %g outside [0, 1]", i, freqs[i]);
    // P_sink(i, j) = f_j / max(1e-12, 1 - f_i), i != j (busted.hpp:193-206)
    std::vector<double> P(4096, 0.0);
    for (int i = 0; i < S; ++i) {
        const double inv = 1.0 / std::max(1e-12, 1.0 - freqs[i]);
        for (int j = 0; j < S; ++j)
            if (j != i) P[(size_t)i * 64 + j] = freqs[j] * inv;
    }
    const void *before = h->d_sink.p;
    TRY(upload(h, h->d_sink, P.data(), P.size() * 8));
    if (h->d_sink.p != before) ++h->epoch;
    h->sink_set = true;
    return HY_OK;
}

int hy_set_tile(hy_engine h, int patterns_per_tile) {
    if (h) ++h->epoch;  // captured graphs are stale
    if (!h) return HY_ERR_ARG;
    if (patterns_per_tile != 0 && patterns_per_tile != 16 && patterns_per_tile != 32 && patterns_per_tile != 64)
        return fail(h, HY_ERR_ARG, "hy_set_tile: tile must be 0, 16, 32 or 64 (got %d)", patterns_per_tile);
    h->T_req = patterns_per_tile;
    return HY_OK;
}

static int finish_host(hy_engine h, bool check_status, double *out_logL, double *out_grad_bl, double *out_grad_omega,
                       double *out_grad_w, double *out_grad_theta, bool grad);

static int finish_eval(hy_engine h, bool check_status, double *out_logL, double *out_site_ll, double *out_grad_bl,
                       double *out_grad_omega, double *out_grad_w, double *out_grad_theta, bool grad) {
    if (out_site_ll && h->P > 0)
        HIP_TRY(h, hipMemcpyAsync(out_site_ll, h->d_site.p, sizeof(double) * h->P, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return finish_host(h, check_status, out_logL, out_grad_bl, out_grad_omega, out_grad_w, out_grad_theta, grad);
}

// Evaluates through a captured graph: all kernels, with k_model reading the
// parameters from the pinned h_params (packed before the launch) and the
// results and eigensolver status written straight to pinned host memory.  Captures on first use per
// (precision, gradient, warm-start) and whenever the configuration epoch moved.
// Replays (capturing first if needed) the evaluation graph of this kind.
// Returns HY_OK when launched, an error code, or GRAPH_NO_SLOT when the
// process-wide graph cap is reached: the caller then launches eagerly.
constexpr int GRAPH_NO_SLOT = 1;

static int launch_graph(hy_engine h, int prec, bool grad) {
    const int warm = (h->warm_enable && h->warm_valid) ? 1 : 0;
    hy_engine_s::GraphEntry *ge = nullptr;
    for (auto &g : h->graphs)
        if (g.prec == prec && g.grad == (int)grad && g.warm == warm) ge = &g;
    if (ge && ge->exec && ge->epoch != h->epoch) {
        // an evaluation awaited on the completion word may still be retiring
        // on the stream: let it finish before its graph is destroyed
        HIP_TRY(h, hipStreamSynchronize(h->stream));
        destroy_graph_exec(ge->exec);
    }
    if (!ge || !ge->exec) {
        if (!acquire_graph_slot()) return GRAPH_NO_SLOT;
        hipGraph_t graph = nullptr;
        HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
        const unsigned flags = HY_WANT_LOGL | (grad ? HY_WANT_GRAD : 0u);
        h->zc_params = true;
        const int rc = prec == HY_PREC_FP64 ? run_eval<double>(h, flags, nullptr) : run_eval<float>(h, flags, nullptr);
        h->zc_params = false;
        const hipError_t e2 = hipStreamEndCapture(h->stream, &graph);
        if (rc != HY_OK || e2 != hipSuccess) {
            g_live_graphs.fetch_sub(1);  // the slot taken above
            if (graph) (void)hipGraphDestroy(graph);
            return rc != HY_OK ? rc : fail(h, HY_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e2));
        }
        hipGraphExec_t exec = nullptr;
        const hipError_t e3 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e3 != hipSuccess) {
            g_live_graphs.fetch_sub(1);
            return fail(h, HY_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e3));
        }
        if (!ge) {
            // the cold-start graph of this kind is not used again once spectra
            // are warm: release it rather than keep its runtime resources
            for (auto it = h->graphs.begin(); it != h->graphs.end(); ++it)
                if (warm && it->prec == prec && it->grad == (int)grad && it->warm == 0) {
                    HIP_TRY(h, hipStreamSynchronize(h->stream));
                    destroy_graph_exec(it->exec);
                    h->graphs.erase(it);
                    break;
                }
            h->graphs.push_back({prec, (int)grad, warm, h->epoch, exec});
            ge = &h->graphs.back();
        } else {
            ge->exec = exec;
            ge->epoch = h->epoch;
        }
    }
    HIP_TRY(h, hipGraphLaunch(ge->exec, h->stream));
    return HY_OK;
}

static int finish_host(hy_engine h, bool check_status, double *out_logL, double *out_grad_bl, double *out_grad_omega,
                       double *out_grad_w, double *out_grad_theta, bool grad) {
    if (h->profile) {
        float ms;
        for (int k = 0; k < 7; ++k) {
            (void)hipEventElapsedTime(&ms, h->ev[k], h->ev[k + 1]);
            h->timings[k] = ms;
        }
        (void)hipEventElapsedTime(&ms, h->ev[0], h->ev[7]);
        h->timings[7] = ms;
    }
    if (check_status) {
        h->warm_valid = true;
        for (int m = 0; m < h->n_mat; ++m)
            if (h->h_status[m] < 0) {
                h->warm_valid = false;
                return fail(h, HY_ERR_NUMERIC, "symmetric eigensolver did not converge for rate matrix %d", m);
            }
    }
    const double *r = h->h_res;
    if (out_logL) *out_logL = r[0];
    if (grad) {
        const double *gbl = r + 1, *gom = gbl + h->num_nodes, *gw = gom + h->n_mat, *gth = gw + h->n_comp;
        if (out_grad_bl) std::memcpy(out_grad_bl, gbl, sizeof(double) * h->num_nodes);
        if (out_grad_omega) std::memcpy(out_grad_omega, gom, sizeof(double) * h->n_mat);
        if (out_grad_w) std::memcpy(out_grad_w, gw, sizeof(double) * h->n_comp);
        if (out_grad_theta) std::memcpy(out_grad_theta, gth, sizeof(double) * 5);
    }
    return HY_OK;
}

}  // extern "C"

namespace {

// Zero outputs of an evaluation over an empty alignment.
void zero_outputs(hy_engine h, unsigned flags, double *out_logL, double *out_grad_bl, double *out_grad_omega,
                  double *out_grad_w, double *out_grad_theta) {
    if (out_logL) *out_logL = 0.0;
    if (flags & HY_WANT_GRAD) {
        if (out_grad_bl) std::fill(out_grad_bl, out_grad_bl + h->num_nodes, 0.0);
        if (out_grad_omega) std::fill(out_grad_omega, out_grad_omega + h->n_mat, 0.0);
        if (out_grad_w) std::fill(out_grad_w, out_grad_w + h->n_comp, 0.0);
        if (out_grad_theta) std::fill(out_grad_theta, out_grad_theta + 5, 0.0);
    }
}

int check_eval_args(hy_engine h, int precision, const double *bl, const double *omegas, const double *comp_w,
                    const double *theta, const char *who) {
    if (!h->have_problem || !h->have_model) return fail(h, HY_ERR_STATE, "%s before hy_set_problem/hy_set_model", who);
    if (precision != HY_PREC_FP64 && precision != HY_PREC_FP32)
        return fail(h, HY_ERR_ARG, "%s: unknown precision %d", who, precision);
    if (!bl || !omegas || !comp_w || !theta) return fail(h, HY_ERR_ARG, "%s: NULL parameter array", who);
    if (h->has_sink && !h->sink_set) return fail(h, HY_ERR_STATE, "%s: error-sink component without hy_set_sink", who);
    return HY_OK;
}

// Packs the parameters and queues one evaluation (captured graph, or eager
// launches) and its result downloads on the handle's stream; the per-pattern
// lnL goes to pinned staging when want_site.  eval_complete() waits for it.
int eval_submit(hy_engine h, int precision, unsigned flags, const double *bl, double alpha, const double *omegas,
                const double *comp_w, const double *theta, double delta, double psi, bool want_site) {
    HIP_TRY(h, hipSetDevice(h->device));
    TRY(prepare_buffers(h, precision));
    double *hp = h->h_params;
    const ParamLayout &pl = h->pl;
    for (int n = 0; n < h->n_ext; ++n) hp[pl.off_bl + n] = (n < h->num_nodes) ? bl[n] : 0.0;
    hp[pl.off_alpha] = alpha;
    for (int m = 0; m < h->n_mat; ++m) hp[pl.off_omega + m] = omegas[m];
    for (int j = 0; j < h->n_comp; ++j) hp[pl.off_w + j] = comp_w[j];
    for (int k = 0; k < 5; ++k) hp[pl.off_theta + k] = theta[k];
    hp[pl.off_delta] = delta;
    hp[pl.off_psi] = psi;
    hp[pl.off_seq] = (double)(++h->seq);
    const bool grad = (flags & HY_WANT_GRAD) != 0;
    if (want_site && h->h_site_cap < (size_t)h->P) {
        if (h->h_site) (void)hipHostFree(h->h_site);
        h->h_site = nullptr;
        h->h_site_cap = 0;
        HIP_TRY(h, hipHostMalloc((void **)&h->h_site, sizeof(double) * h->P, hipHostMallocDefault));
        h->h_site_cap = h->P;
    }
    int grc = GRAPH_NO_SLOT;
    h->wait_done = false;
    h->publish_done = true;
    // the caller's pattern weights stand in for the alignment's unless an
    // internal override (SRV posteriors) is active; both launch eagerly
    const bool user_w = h->wuser && !h->wover;
    if (user_w) h->wover = (const double *)h->d_wuser.p;
    if (h->use_graphs && !h->profile && !h->wover) grc = launch_graph(h, precision, grad);
    if (grc < 0) {
        h->publish_done = false;
        if (user_w) h->wover = nullptr;
        return grc;
    }
    if (grc == GRAPH_NO_SLOT) {
        const hipError_t e = hipMemcpyAsync(h->d_params.p, hp, sizeof(double) * pl.total, hipMemcpyHostToDevice,
                                            h->stream);
        const int rc = e != hipSuccess ? fail(h, HY_ERR_HIP, "hipMemcpyAsync: %s", hipGetErrorString(e))
                       : precision == HY_PREC_FP64 ? run_eval<double>(h, flags, nullptr)
                                                   : run_eval<float>(h, flags, nullptr);
        if (rc) {
            h->publish_done = false;
            if (user_w) h->wover = nullptr;
            return rc;
        }
    }
    h->publish_done = false;
    if (user_w) h->wover = nullptr;
    // partials left resident by this evaluation (a replayed graph does not
    // pass through run_eval, which records them when it launches eagerly)
    h->part_ri = precision == HY_PREC_FP64 ? 1 : 0;
    h->part_outside = grad;
    // the results can be awaited on the completion word unless a download
    // (site likelihoods) follows the evaluation on the stream
    h->wait_done = !want_site && !h->profile;
    if (want_site)
        HIP_TRY(h, hipMemcpyAsync(h->h_site, h->d_site.p, sizeof(double) * h->P, hipMemcpyDeviceToHost, h->stream));
    return HY_OK;
}

int eval_complete(hy_engine h, bool grad, double *out_logL, double *out_site_ll, double *out_grad_bl,
                  double *out_grad_omega, double *out_grad_w, double *out_grad_theta) {
    HIP_TRY(h, hipSetDevice(h->device));
    if (h->wait_done) {
        // k_final stores this call's number in the pinned completion word once
        // every result is visible to the host: spin on it (a stream wait adds
        // the end-of-dispatch signalling to every evaluation); after 0.2 s the
        // stream wait takes over, which also reports a failed launch
        h->wait_done = false;
        const unsigned long long want = h->seq;
        bool seen = false;
        const auto t0 = std::chrono::steady_clock::now();
        for (unsigned it = 1;; ++it) {
            if (__atomic_load_n(h->h_done, __ATOMIC_ACQUIRE) == want) {
                seen = true;
                break;
            }
            __builtin_ia32_pause();
            if ((it & 1023u) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(200)) break;
        }
        if (!seen) {
            HIP_TRY(h, hipStreamSynchronize(h->stream));
            if (__atomic_load_n(h->h_done, __ATOMIC_ACQUIRE) != want)
                return fail(h, HY_ERR_HIP, "evaluation finished without publishing its results");
        }
    } else {
        HIP_TRY(h, hipStreamSynchronize(h->stream));
    }
    if (out_site_ll) std::memcpy(out_site_ll, h->h_site, sizeof(double) * h->P);
    return finish_host(h, true, out_logL, out_grad_bl, out_grad_omega, out_grad_w, out_grad_theta, grad);
}

}  // namespace

extern "C" {

int hy_eval(hy_engine h, int precision, unsigned flags, const double *bl, double alpha, const double *omegas,
            const double *comp_w, const double *theta, double delta, double psi, double *out_logL,
            double *out_site_ll, double *out_grad_bl, double *out_grad_omega, double *out_grad_w,
            double *out_grad_theta) {
    if (!h) return HY_ERR_ARG;
    TRY(check_eval_args(h, precision, bl, omegas, comp_w, theta, "hy_eval"));
    if (h->P == 0) {
        zero_outputs(h, flags, out_logL, out_grad_bl, out_grad_omega, out_grad_w, out_grad_theta);
        return HY_OK;
    }
    TRY(eval_submit(h, precision, flags, bl, alpha, omegas, comp_w, theta, delta, psi, out_site_ll != nullptr));
    return eval_complete(h, (flags & HY_WANT_GRAD) != 0, out_logL, out_site_ll, out_grad_bl, out_grad_omega,
                         out_grad_w, out_grad_theta);
}

int hy_autotune(hy_engine h, int precision, unsigned flags, const double *bl, double alpha, const double *omegas,
                const double *comp_w, const double *theta, double delta, double psi, int reps, int *out_tile,
                int *out_parts, double *out_ms) {
    if (!h) return HY_ERR_ARG;
    TRY(check_eval_args(h, precision, bl, omegas, comp_w, theta, "hy_autotune"));
    if (reps < 1 || reps > 1000) return fail(h, HY_ERR_ARG, "hy_autotune: reps must be in 1..1000 (got %d)", reps);
    const int ri = precision == HY_PREC_FP64 ? 1 : 0;
    const unsigned fl = (flags & HY_WANT_GRAD) | HY_WANT_LOGL;
    const int saved_req = h->T_req;
    h->T_req = 0;
    int best_T = 0, best_parts = -1;
    double best = 1e300;
    if (h->P > 0) {
        for (int T = 16; T <= max_tile(precision); T *= 2)
            for (int parts = 0; parts <= 1; ++parts) {
                if (parts && h->no_parts) continue;
                h->tuned_T[ri] = T;
                h->tuned_parts[ri] = parts;
                ++h->epoch;
                // one untimed evaluation (re-tiling, schedules, graph capture)
                int rc = eval_submit(h, precision, fl, bl, alpha, omegas, comp_w, theta, delta, psi, false);
                if (rc == HY_OK) rc = eval_complete(h, false, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
                std::vector<double> t;
                for (int r = 0; rc == HY_OK && r < reps; ++r) {
                    const auto t0 = std::chrono::steady_clock::now();
                    rc = eval_submit(h, precision, fl, bl, alpha, omegas, comp_w, theta, delta, psi, false);
                    if (rc == HY_OK) rc = eval_complete(h, false, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
                    t.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
                }
                if (rc != HY_OK) {
                    h->T_req = saved_req;
                    h->tuned_T[ri] = 0;
                    h->tuned_parts[ri] = -1;
                    ++h->epoch;
                    return rc;
                }
                std::sort(t.begin(), t.end());
                const double med = t[t.size() / 2];
                if (med < best) {
                    best = med;
                    best_T = T;
                    best_parts = h->n_parts > 0 ? 1 : 0;
                }
            }
    }
    h->tuned_T[ri] = best_T;
    h->tuned_parts[ri] = best_parts;
    h->T_req = saved_req;
    ++h->epoch;
    if (out_tile) *out_tile = best_T;
    if (out_parts) *out_parts = best_parts;
    if (out_ms) *out_ms = h->P > 0 ? best : 0.0;
    return HY_OK;
}

int hy_set_tuning(hy_engine h, int precision, int patterns_per_tile, int parts) {
    if (!h) return HY_ERR_ARG;
    if (precision != HY_PREC_FP64 && precision != HY_PREC_FP32)
        return fail(h, HY_ERR_ARG, "hy_set_tuning: unknown precision %d", precision);
    if (patterns_per_tile != 0 && patterns_per_tile != 16 && patterns_per_tile != 32 && patterns_per_tile != 64)
        return fail(h, HY_ERR_ARG, "hy_set_tuning: tile must be 0, 16, 32 or 64 (got %d)", patterns_per_tile);
    if (parts < -1 || parts > 1) return fail(h, HY_ERR_ARG, "hy_set_tuning: parts must be -1, 0 or 1 (got %d)", parts);
    const int ri = precision == HY_PREC_FP64 ? 1 : 0;
    h->tuned_T[ri] = patterns_per_tile;
    h->tuned_parts[ri] = parts;
    ++h->epoch;
    return HY_OK;
}

int hy_eval_srv(hy_engine h, int precision, unsigned flags, const double *bl, double alpha, const double *omegas,
                const double *comp_w, const double *theta, double delta, double psi, int n_syn,
                const double *syn_rates, const double *syn_weights, double *out_logL, double *out_site_ll,
                double *out_grad_bl, double *out_grad_omega, double *out_grad_w, double *out_grad_theta,
                double *out_grad_syn_rates, double *out_grad_syn_weights) {
    // BUSTEDAnalyzer::evaluate_log_l / compute_busted_branch_gradients with M
    // synonymous-rate classes (busted.hpp:554-581, 607-649, 1191-1216,
    // 1335-1396): class m scales every branch time by syn_rates[m]; per
    // pattern lnL_p = log sum_m v_m L_m(p) (log-sum-exp, classes with
    // v_m <= 1e-12 skipped); gradients are posterior-weighted sums of the
    // class gradients, computed here as one evaluation per class whose
    // pattern weights are w_p * post_m(p).
    if (!h) return HY_ERR_ARG;
    TRY(check_eval_args(h, precision, bl, omegas, comp_w, theta, "hy_eval_srv"));
    if (h->wuser)
        return fail(h, HY_ERR_STATE, "hy_eval_srv: clear the caller's pattern weights (hy_set_pattern_weights NULL) first");
    if (n_syn < 1 || !syn_rates || !syn_weights) return fail(h, HY_ERR_ARG, "hy_eval_srv: need n_syn >= 1 rates and weights");
    for (int m = 0; m < n_syn; ++m)
        if (!(syn_rates[m] >= 0.0) || !std::isfinite(syn_rates[m]) || !(syn_weights[m] >= 0.0) || !std::isfinite(syn_weights[m]))
            return fail(h, HY_ERR_ARG, "hy_eval_srv: syn_rates[%d] / syn_weights[%d] must be finite and >= 0", m, m);
    const bool grad = (flags & HY_WANT_GRAD) != 0;
    if (grad) {
        if (out_grad_syn_rates) std::fill(out_grad_syn_rates, out_grad_syn_rates + n_syn, 0.0);
        if (out_grad_syn_weights) std::fill(out_grad_syn_weights, out_grad_syn_weights + n_syn, 0.0);
    }
    if (h->P == 0) {
        zero_outputs(h, flags, out_logL, out_grad_bl, out_grad_omega, out_grad_w, out_grad_theta);
        return HY_OK;
    }
    const int P = h->P, N = h->num_nodes, M = n_syn;
    std::vector<double> blm(N), sll((size_t)M * P, -1e20);
    std::vector<char> live(M);
    for (int m = 0; m < M; ++m) {
        live[m] = syn_weights[m] > 1e-12;
        if (!live[m]) continue;
        for (int n = 0; n < N; ++n) blm[n] = syn_rates[m] * bl[n];
        TRY(eval_submit(h, precision, HY_WANT_LOGL, blm.data(), alpha, omegas, comp_w, theta, delta, psi, true));
        TRY(eval_complete(h, false, nullptr, sll.data() + (size_t)m * P, nullptr, nullptr, nullptr, nullptr));
    }
    // Per-pattern log-sum-exp (busted.hpp:621-646) and class posteriors.
    std::vector<double> pll(P, -1e20), post((size_t)M * P, 0.0);
    for (int p = 0; p < P; ++p) {
        double mx = -1e20;
        for (int m = 0; m < M; ++m)
            if (live[m] && sll[(size_t)m * P + p] > -1e19) mx = std::max(mx, std::log(syn_weights[m]) + sll[(size_t)m * P + p]);
        if (mx <= -1e19) continue;
        double se = 0.0;
        for (int m = 0; m < M; ++m)
            if (live[m] && sll[(size_t)m * P + p] > -1e19) {
                const double e = std::exp(std::log(syn_weights[m]) + sll[(size_t)m * P + p] - mx);
                post[(size_t)m * P + p] = e;
                se += e;
            }
        if (se > 0.0) {
            pll[p] = mx + std::log(se);
            for (int m = 0; m < M; ++m) post[(size_t)m * P + p] /= se;
        } else {
            for (int m = 0; m < M; ++m) post[(size_t)m * P + p] = 0.0;
        }
    }
    {
        double s = 0.0, c = 0.0;  // Neumaier in pattern order (busted.hpp:652-656)
        for (int p = 0; p < P; ++p) {
            const double x = h->weights[p] * pll[p], t = s + x;
            c += (std::fabs(s) >= std::fabs(x)) ? (s - t) + x : (x - t) + s;
            s = t;
        }
        if (out_logL) *out_logL = s + c;
    }
    if (out_site_ll) std::memcpy(out_site_ll, pll.data(), sizeof(double) * P);
    if (!grad) return HY_OK;
    std::vector<double> gbl(N, 0.0), gom(h->n_mat, 0.0), gw(h->n_comp, 0.0), gth(5, 0.0);
    std::vector<double> tb(N), to(h->n_mat), tw(h->n_comp), tt(5), wp((size_t)h->P_pad, 0.0);
    int rc = HY_OK;
    for (int m = 0; m < M && rc == HY_OK; ++m) {
        if (!live[m]) continue;
        // posterior-weighted pattern weights; terms with post <= 1e-12 are
        // dropped as in busted.hpp:1383-1386
        double sw = 0.0;
        for (int p = 0; p < P; ++p) {
            const double q = post[(size_t)m * P + p];
            wp[p] = q > 1e-12 ? h->weights[p] * q : 0.0;
            sw += h->weights[p] * q;
        }
        if (out_grad_syn_weights) out_grad_syn_weights[m] = sw / syn_weights[m];
        rc = prepare_buffers(h, precision);
        if (rc == HY_OK) rc = ensure(h, h->d_wover, (size_t)h->P_pad * 8);
        if (rc == HY_OK) rc = upload(h, h->d_wover, wp.data(), (size_t)h->P_pad * 8);
        if (rc != HY_OK) break;
        for (int n = 0; n < N; ++n) blm[n] = syn_rates[m] * bl[n];
        h->wover = (const double *)h->d_wover.p;
        rc = eval_submit(h, precision, HY_WANT_LOGL | HY_WANT_GRAD, blm.data(), alpha, omegas, comp_w, theta, delta,
                         psi, false);
        if (rc == HY_OK) rc = eval_complete(h, true, nullptr, nullptr, tb.data(), to.data(), tw.data(), tt.data());
        h->wover = nullptr;
        if (rc != HY_OK) break;
        double gr = 0.0;
        for (int n = 0; n < N; ++n) {
            gbl[n] += syn_rates[m] * tb[n];
            gr += bl[n] * tb[n];
        }
        if (out_grad_syn_rates) out_grad_syn_rates[m] = gr;
        for (int k = 0; k < h->n_mat; ++k) gom[k] += to[k];
        for (int j = 0; j < h->n_comp; ++j) gw[j] += tw[j];
        for (int k = 0; k < 5; ++k) gth[k] += tt[k];
    }
    if (rc != HY_OK) return rc;
    if (out_grad_bl) std::memcpy(out_grad_bl, gbl.data(), sizeof(double) * N);
    if (out_grad_omega) std::memcpy(out_grad_omega, gom.data(), sizeof(double) * h->n_mat);
    if (out_grad_w) std::memcpy(out_grad_w, gw.data(), sizeof(double) * h->n_comp);
    if (out_grad_theta) std::memcpy(out_grad_theta, gth.data(), sizeof(double) * 5);
    return HY_OK;
}

int hy_busted_expected_weights(hy_engine h, int precision, const double *bl, double alpha, const double *omegas,
                               const double *comp_w, const double *theta, double delta, double psi, int n_syn,
                               const double *syn_rates, const double *syn_weights, int test_class,
                               double max_eps_sink, double *out_weights, double *out_syn_weights) {
    // BUSTEDAnalyzer::compute_expected_weights (busted.hpp:662-976) and
    // compute_expected_syn_weights (busted.hpp:978-1135) without bridges.
    // The reference accumulates, per test branch b, component k and class m,
    // w_p w_k post_m(p) V_b . (P_k D_b) / L_m(p); under the raw time
    // convention that sum over p and b is w_k dlnL/dw_k (the
    // posterior-weighted mixture-weight gradient), and sum_p w_p post_m(p) is
    // v_m dlnL/dv_m, so one gradient evaluation gives both updates.  The
    // error sink's share eps sum post_m V_b . P_sink D_b / L_m (:846-860) is
    // likewise eps dlnL/d eps, the sink component's weight gradient.
    if (!h) return HY_ERR_ARG;
    if (!h->have_problem || !h->have_model)
        return fail(h, HY_ERR_STATE, "hy_busted_expected_weights before hy_set_problem/hy_set_model");
    if (h->wuser)
        return fail(h, HY_ERR_STATE,
                    "hy_busted_expected_weights: clear the caller's pattern weights (hy_set_pattern_weights NULL) first");
    if (h->convention != HY_TIME_RAW)
        return fail(h, HY_ERR_STATE, "hy_busted_expected_weights: the mixture must use HY_TIME_RAW branch times");
    if (test_class < 0 || test_class >= h->n_class)
        return fail(h, HY_ERR_ARG, "hy_busted_expected_weights: test class %d out of range", test_class);
    if (n_syn < 1 || !syn_rates || !syn_weights || !out_weights || !out_syn_weights)
        return fail(h, HY_ERR_ARG, "hy_busted_expected_weights: bad synonymous classes or outputs");
    const int j0 = h->comp_ptr[test_class], j1 = h->comp_ptr[test_class + 1];
    // the test class's error-sink component (at most one), and its omega components
    int j_sink = -1;
    std::vector<int> jw;
    for (int j = j0; j < j1; ++j) {
        if (h->comp_mat[j] >= 0) {
            jw.push_back(j);
        } else if (j_sink >= 0) {
            return fail(h, HY_ERR_ARG, "hy_busted_expected_weights: test class %d has more than one error sink",
                        test_class);
        } else {
            j_sink = j;
        }
    }
    if (j_sink >= 0 && !(max_eps_sink >= 1e-7 && max_eps_sink <= 1.0))
        return fail(h, HY_ERR_ARG, "hy_busted_expected_weights: max_eps_sink %g outside [1e-7, 1]", max_eps_sink);
    const int K = (int)jw.size();
    std::vector<double> site(std::max(h->P, 1)), gbl(h->num_nodes), gom(h->n_mat), gw(h->n_comp), gth(5),
        gsr(n_syn), gsw(n_syn);
    double ll = 0.0;
    TRY(hy_eval_srv(h, precision, HY_WANT_LOGL | HY_WANT_SITE_LL | HY_WANT_GRAD, bl, alpha, omegas, comp_w, theta,
                    delta, psi, n_syn, syn_rates, syn_weights, &ll, site.data(), gbl.data(), gom.data(), gw.data(),
                    gth.data(), gsr.data(), gsw.data()));
    // pattern weight of the patterns with a positive likelihood in some class
    double total = 0.0;
    for (int p = 0; p < h->P; ++p)
        if (site[p] > -1e19) total += h->weights[p];
    int n_test = 0;
    for (int n = 0; n < h->num_nodes; ++n)
        if (h->real_parent[n] >= 0 && (h->branch_class.empty() ? 0 : h->branch_class[n]) == test_class) ++n_test;
    auto normalise = [](double *x, int n, const double *num, double den, const double *prev) {
        if (den > 0.0) {
            double sum = 0.0;
            for (int i = 0; i < n; ++i) {
                x[i] = std::clamp(num[i] / den, 1e-6, 1.0 - 1e-6);
                sum += x[i];
            }
            for (int i = 0; i < n; ++i) x[i] /= sum;
        } else {
            for (int i = 0; i < n; ++i) x[i] = prev[i];
        }
    };
    // out_weights follows the class's component order: omega components get
    // their normalised shares, the sink component its new eps
    const double den = (double)n_test * total;
    std::vector<double> num(std::max(K, 1)), prev(std::max(K, 1)), neww(std::max(K, 1), 1.0);
    // (with a sink the class's weights are (1 - eps) w_k; the unchanged
    // fallback is the reference's w_k)
    const double bio = (j_sink >= 0 && comp_w[j_sink] < 1.0) ? 1.0 - comp_w[j_sink] : 1.0;
    for (int k = 0; k < K; ++k) {
        num[k] = comp_w[jw[k]] * gw[jw[k]];
        prev[k] = comp_w[jw[k]] / bio;
    }
    if (K > 1) normalise(neww.data(), K, num.data(), den, prev.data());
    for (int k = 0; k < K; ++k) out_weights[jw[k] - j0] = neww[k];
    if (j_sink >= 0) {
        const double eps = comp_w[j_sink];
        out_weights[j_sink - j0] = den > 0.0 ? std::clamp(eps * gw[j_sink] / std::max(1.0, den), 1e-7, max_eps_sink)
                                             : eps;
    }
    if (n_syn <= 1) {
        out_syn_weights[0] = 1.0;
    } else {
        std::vector<double> num(n_syn);
        for (int m = 0; m < n_syn; ++m) num[m] = syn_weights[m] * gsw[m];
        normalise(out_syn_weights, n_syn, num.data(), total, syn_weights);
    }
    return HY_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Fused batches: the problems (genes) of a group's engines merged into one
// problem of an internal engine, so that every stage is one launch over all
// genes.  Kernels find each gene's slice through the gene table (tiles and
// branches), matrix -> gene and tile -> gene maps; node, matrix, component,
// class and branch indices are batch-global, partial and pattern indices
// gene-local.
// ---------------------------------------------------------------------------

namespace {

bool fusable(const std::vector<hy_engine> &hs) {
    if (hs.size() < 2) return false;
    const hy_engine h0 = hs[0];
    if (!h0->group || !h0->group->ctx) return false;
    for (hy_engine h : hs)
        if (!h || h->group != h0->group || !h->have_problem || !h->have_model || h->P == 0 || h->S != h0->S ||
            h->code_id != h0->code_id || h->convention != h0->convention || h->wuser)
            return false;  // (caller pattern weights: evaluated engine by engine)
    return true;
}

// Builds the merged problem of the engines hs in the group's context engine.
int build_batch(GroupState &gs, const std::vector<hy_engine> &hs, int prec) {
    hy_engine c = gs.ctx;
    const int esz = prec == HY_PREC_FP64 ? 8 : 4, ri = prec == HY_PREC_FP64 ? 1 : 0;
    const int G = (int)hs.size();
    // tile width: the batch autotuner's, else fewest rounds of resident
    // workgroups x per-node cost
    int T = gs.tuned_T[ri];
    if (T != 16 && T != 32 && T != 64) {
        T = 16;
        const double node_cost[3] = {1.0, 1.3, 2.0};
        double best = 1e300;
        for (int k = 0; (16 << k) <= max_tile(prec); ++k) {
            const int Tk = 16 << k;
            long tiles = 0;
            for (hy_engine h : hs) tiles += (h->P + Tk - 1) / Tk;
            const long slots = (long)c->n_cu * ((prec == HY_PREC_FP32 && Tk == 16) ? 2 : 1);
            const double cost = (double)((tiles + slots - 1) / slots) * node_cost[k];
            if (cost < best) {
                best = cost;
                T = Tk;
            }
        }
    }
    T = std::min(T, max_tile(prec));
    long total_tiles = 0;
    for (hy_engine h : hs) total_tiles += (h->P + T - 1) / T;
    const bool lean = lean_inside(prec == HY_PREC_FP32, T, total_tiles, c->n_cu);
    std::vector<GeneInfo> genes(G);
    std::vector<int32_t> sched, osched, blist, comp_ptr{0}, comp_mat, tile_gene, mat_gene, wl;
    std::vector<uint16_t> tips;
    std::vector<uint64_t> amb;
    std::vector<double> weights, pos, pi;
    std::vector<uint8_t> ident;
    gs.slices.assign(G, {});
    long long tip_base = 0, d_base = 0, e_base = 0, v_base = 0, g_base = 0, u_base = 0;
    int tile0 = 0, pbase = 0, node0 = 0, mat0 = 0, comp0 = 0, cls0 = 0, br0 = 0, max_taxa = 0, max_comp = 1;
    size_t res_off = 0;
    for (int g = 0; g < G; ++g) max_comp = std::max(max_comp, hs[g]->max_comp);
    for (int g = 0; g < G; ++g) {
        hy_engine h = hs[g];
        // the member's own schedule state (host copy and the key telling its
        // prepare_buffers that the device copy is current) is left untouched
        std::vector<int32_t> own_sched, own_osched;
        own_sched.swap(h->sched);
        own_osched.swap(h->osched);
        const int own_uspill = h->n_uspill, own_key = h->sched_key, own_nparts = h->n_parts;
        std::vector<int32_t> own_parts(h->parts);
        const bool own_lean = h->lean, own_lean_out = h->lean_out;
        h->lean = lean;
        h->lean_out = false;  // batch schedules: whole trees, the default outside variant
        build_schedules(h, T, esz, false);
        h->lean = own_lean;
        h->lean_out = own_lean_out;
        std::vector<int32_t> g_sched, g_osched;
        g_sched.swap(h->sched);
        g_osched.swap(h->osched);
        const int g_uspill = h->n_uspill;
        h->sched.swap(own_sched);
        h->osched.swap(own_osched);
        h->n_uspill = own_uspill;
        h->sched_key = own_key;
        h->n_parts = own_nparts;
        h->parts.swap(own_parts);
        const int ntiles = (h->P + T - 1) / T, P_pad = ntiles * T;
        GeneInfo &gi = genes[g];
        gi = GeneInfo{};
        gi.tile0 = tile0;
        gi.ntiles = ntiles;
        gi.pbase = pbase;
        gi.P_pad = P_pad;
        gi.P = h->P;
        gi.sched_off = (int)(sched.size() / SCHED_STRIDE);
        gi.n_sched = (int)(g_sched.size() / SCHED_STRIDE);
        gi.osched_off = (int)(osched.size() / OSCHED_STRIDE);
        gi.n_osched = (int)(g_osched.size() / OSCHED_STRIDE);
        gi.tip_rows = h->num_taxa + 1;
        gi.tiles_per_chunk = ntiles;
        gi.br0 = br0;
        gi.n_br = h->n_branch;
        gi.node0 = node0;
        gi.n_real_nodes = h->num_nodes;
        gi.mat0 = mat0;
        gi.n_mat = h->n_mat;
        gi.comp0 = comp0;
        gi.n_comp = h->n_comp;
        gi.cls0 = cls0;
        gi.n_cls = h->n_class;
        gi.res_off = (int)res_off;
        gi.tip_base = tip_base;
        gi.d_base = d_base;
        gi.e_base = e_base;
        gi.v_base = v_base;
        gi.g_base = g_base;
        gi.u_base = u_base;
        gs.slices[g] = {node0, h->n_ext, h->num_nodes, mat0, h->n_mat, comp0, h->n_comp, pbase, h->P, res_off};
        // schedules: edge (node) ids become batch-global
        for (size_t i = 0; i < g_sched.size(); i += SCHED_STRIDE) {
            int32_t e[SCHED_STRIDE];
            std::copy(g_sched.begin() + i, g_sched.begin() + i + SCHED_STRIDE, e);
            e[0] += node0;
            for (int cc = 0; cc < e[2] && cc < MAXC; ++cc)
                if (e[6 + 3 * cc] >= 0) e[6 + 3 * cc] += node0;
            sched.insert(sched.end(), e, e + SCHED_STRIDE);
        }
        for (size_t i = 0; i < g_osched.size(); i += OSCHED_STRIDE) {
            int32_t e[OSCHED_STRIDE];
            std::copy(g_osched.begin() + i, g_osched.begin() + i + OSCHED_STRIDE, e);
            e[0] += node0;
            for (int cc = 0; cc < e[2] && cc < MAXC; ++cc) e[6 + 4 * cc] += node0;
            osched.insert(osched.end(), e, e + OSCHED_STRIDE);
        }
        for (int b = 0; b < h->n_branch; ++b) {
            int32_t e[BR_STRIDE];
            std::copy(h->blist.begin() + (size_t)b * BR_STRIDE, h->blist.begin() + (size_t)(b + 1) * BR_STRIDE, e);
            e[0] += node0;
            e[3] += cls0;
            e[4] = g;
            blist.insert(blist.end(), e, e + BR_STRIDE);
        }
        for (int cc = 0; cc < h->n_class; ++cc) comp_ptr.push_back(comp0 + h->comp_ptr[cc + 1]);
        if (h->has_sink) return fail(c, HY_ERR_UNSUPPORTED, "batch: error-sink components are not batched");
        for (int j = 0; j < h->n_comp; ++j) comp_mat.push_back(mat0 + h->comp_mat[j]);
        // tip codes [taxa + 1][P_pad], ambiguity indices batch-global
        const int amb0 = (int)amb.size();
        if (amb0 + (int)h->amb.size() >= 0x7FFF) return fail(c, HY_ERR_UNSUPPORTED, "batch: more than 32766 ambiguity masks");
        amb.insert(amb.end(), h->amb.begin(), h->amb.end());
        const size_t t0 = tips.size();
        tips.resize(t0 + (size_t)(h->num_taxa + 1) * P_pad, TIP_ONES);
        for (int t = 0; t < h->num_taxa; ++t)
            for (int p = 0; p < h->P; ++p) {
                uint16_t code = h->codes[(size_t)t * h->P + p];
                if (code != TIP_ONES && (code & TIP_AMB)) code = (uint16_t)(TIP_AMB | ((code & 0x7FFF) + amb0));
                tips[t0 + (size_t)t * P_pad + p] = code;
            }
        weights.resize(pbase + P_pad, 0.0);
        std::copy(h->weights.begin(), h->weights.end(), weights.begin() + pbase);
        for (int m = 0; m < h->n_mat; ++m) mat_gene.push_back(g);
        for (int t = 0; t < ntiles; ++t) tile_gene.push_back(g);
        pos.insert(pos.end(), h->pos_host.begin(), h->pos_host.end());
        pi.insert(pi.end(), h->pi_host.begin(), h->pi_host.end());
        ident.insert(ident.end(), h->ext_ident.begin(), h->ext_ident.end());
        max_taxa = std::max(max_taxa, h->num_taxa);
        tile0 += ntiles;
        pbase += P_pad;
        node0 += h->n_ext;
        mat0 += h->n_mat;
        comp0 += h->n_comp;
        cls0 += h->n_class;
        br0 += h->n_branch;
        tip_base += (long long)(h->num_taxa + 1) * P_pad;
        d_base += (long long)h->n_int * ntiles * 64 * T;
        e_base += (long long)h->n_int * P_pad;
        v_base += (long long)h->n_vslot * ntiles * 64 * T;
        g_base += (long long)h->n_vslot * P_pad;
        if (g_uspill) u_base += (long long)h->n_int * ntiles * 64 * T;
        res_off += 1 + (size_t)h->num_nodes + h->n_mat + h->n_comp + 5;
    }
    // Host-side bounds check of every index the kernels will follow through
    // the merged tables (gene-local partial / tip / slot indices, batch-global
    // nodes, classes, components and matrices); nothing is launched otherwise.
    {
        const int NSI = inside_slots(T, esz, lean), NSO = outside_slots(T, esz);
        auto bad = [&](int g, const char *what) { return fail(c, HY_ERR_STATE, "batch tables: gene %d: %s", g, what); };
        for (int g = 0; g < G; ++g) {
            const hy_engine h = hs[g];
            const GeneInfo &gi = genes[g];
            auto ref_ok = [&](int32_t ref) {
                const int kind = ref >> 28, idx = ref & REF_MASK;
                return kind == KIND_TIP ? idx < gi.tip_rows : (kind == KIND_INT && idx < h->n_int);
            };
            for (int k = 0; k < gi.n_sched; ++k) {
                const int32_t *e = &sched[(size_t)(gi.sched_off + k) * SCHED_STRIDE];
                if (e[1] < 0 || e[1] >= h->n_int || e[2] < 1 || e[2] > MAXC || e[3] >= NSI) return bad(g, "inside entry");
                for (int cc = 0; cc < e[2]; ++cc) {
                    const int pn = e[6 + 3 * cc];
                    if (pn != -1 && (pn < gi.node0 || pn >= gi.node0 + h->n_ext)) return bad(g, "inside edge");
                    if (!ref_ok(e[7 + 3 * cc]) || e[8 + 3 * cc] >= NSI) return bad(g, "inside child");
                }
            }
            for (int k = 0; k < gi.n_osched; ++k) {
                const int32_t *e = &osched[(size_t)(gi.osched_off + k) * OSCHED_STRIDE];
                if (e[1] < 0 || e[1] >= h->n_int || e[2] < 1 || e[2] > MAXC || e[3] >= h->n_vslot || e[5] >= NSO)
                    return bad(g, "outside entry");
                for (int cc = 0; cc < e[2]; ++cc) {
                    const int pn = e[6 + 4 * cc];
                    if (pn < gi.node0 || pn >= gi.node0 + h->n_ext) return bad(g, "outside edge");
                    if (!ref_ok(e[7 + 4 * cc]) || e[8 + 4 * cc] < 0 || e[8 + 4 * cc] >= h->n_vslot || e[9 + 4 * cc] >= NSO)
                        return bad(g, "outside child");
                }
            }
            for (int b = 0; b < gi.n_br; ++b) {
                const int32_t *e = &blist[(size_t)(gi.br0 + b) * BR_STRIDE];
                if (e[0] < gi.node0 || e[0] >= gi.node0 + h->n_ext || e[1] < 0 || e[1] >= h->n_vslot || !ref_ok(e[2]) ||
                    e[3] < gi.cls0 || e[3] >= gi.cls0 + gi.n_cls || e[4] != g)
                    return bad(g, "branch entry");
            }
        }
        for (size_t k = 0; k + 1 < comp_ptr.size(); ++k)
            if (comp_ptr[k] < 0 || comp_ptr[k + 1] <= comp_ptr[k] || comp_ptr[k + 1] > comp0) return bad(-1, "class components");
        for (int32_t m : comp_mat)
            if (m < 0 || m >= mat0) return bad(-1, "component matrix");
        for (uint16_t code : tips)
            if (code != TIP_ONES && (code & TIP_AMB) && (size_t)(code & 0x7FFF) >= amb.size()) return bad(-1, "tip code");
        if ((int)tile_gene.size() != tile0 || (int)mat_gene.size() != mat0) return bad(-1, "maps");
    }
    // per matrix: the W slots (branch * max_comp + component index) using it
    {
        std::vector<std::vector<int32_t>> per(mat0);
        for (int bi = 0; bi < br0; ++bi) {
            const int cls = blist[(size_t)bi * BR_STRIDE + 3];
            for (int j = comp_ptr[cls]; j < comp_ptr[cls + 1]; ++j) per[comp_mat[j]].push_back(bi * max_comp + (j - comp_ptr[cls]));
        }
        wl.assign(mat0 + 1, 0);
        for (int m = 0; m < mat0; ++m) wl[m + 1] = wl[m] + (int32_t)per[m].size();
        for (int m = 0; m < mat0; ++m) wl.insert(wl.end(), per[m].begin(), per[m].end());
        set_frechet_groups(c, per);
    }
    // context engine state
    hy_engine h0 = hs[0];
    c->S = h0->S;
    c->code_id = h0->code_id;
    c->convention = h0->convention;
    c->n_genes = G;
    c->lean = lean;
    c->T = T;
    c->T_req = T;
    c->n_tiles = tile0;
    c->P_pad = pbase;
    c->P = pbase;
    c->num_taxa = max_taxa;
    c->n_ext = node0;
    c->num_nodes = node0;
    c->n_mat = mat0;
    c->n_comp = comp0;
    c->n_class = cls0;
    c->max_comp = max_comp;
    c->n_branch = br0;
    c->n_chunks = 1;
    c->tiles_per_chunk = 1;
    c->sched = sched;
    c->osched = osched;
    c->blist = blist;
    c->ext_ident = ident;
    c->warm_valid = false;
    c->have_problem = c->have_model = true;
    {
        ParamLayout &pl = c->pl;
        pl.n_nodes = node0;
        pl.n_mat = mat0;
        pl.n_comp = comp0;
        int o = 0;
        pl.off_bl = o;
        o += node0;
        pl.off_alpha = o;
        o += G;
        pl.off_omega = o;
        o += mat0;
        pl.off_w = o;
        o += comp0;
        pl.off_theta = o;
        o += 5 * G;
        pl.off_delta = o;
        o += G;
        pl.off_psi = o;
        o += G;
        pl.total = o;
    }
    const size_t tile_elems = (size_t)64 * T;
    TRY(upload(c, c->d_genes, genes.data(), genes.size() * sizeof(GeneInfo)));
    TRY(upload(c, c->d_tile_gene, tile_gene.data(), std::max<size_t>(tile_gene.size(), 1) * 4));
    TRY(upload(c, c->d_mat_gene, mat_gene.data(), mat_gene.size() * 4));
    TRY(upload(c, c->d_sched, sched.data(), sched.size() * 4));
    TRY(upload(c, c->d_osched, osched.data(), osched.size() * 4));
    TRY(upload(c, c->d_blist, blist.data(), std::max<size_t>(blist.size(), BR_STRIDE) * 4));
    TRY(upload(c, c->d_comp_ptr, comp_ptr.data(), comp_ptr.size() * 4));
    TRY(upload(c, c->d_comp_mat, comp_mat.data(), comp_mat.size() * 4));
    TRY(upload(c, c->d_tips, tips.data(), std::max<size_t>(tips.size(), 1) * 2));
    TRY(upload(c, c->d_amb, amb.data(), std::max<size_t>(amb.size(), 1) * 8));
    c->n_amb = (int)amb.size();
    TRY(upload(c, c->d_weights, weights.data(), std::max<size_t>(weights.size(), 1) * 8));
    TRY(upload(c, c->d_pos, pos.data(), pos.size() * 8));
    TRY(upload(c, c->d_pi, pi.data(), pi.size() * 8));
    TRY(upload(c, c->d_wl, wl.data(), wl.size() * 4));
    c->wl = wl;
    // genetic-code tables of the first gene (all genes share the code)
    DevBuf *src[6] = {&h0->d_qdesc, &h0->d_nd, &h0->d_dpos, &h0->d_dfrom, &h0->d_dto, &h0->d_syn};
    DevBuf *dst[6] = {&c->d_qdesc, &c->d_nd, &c->d_dpos, &c->d_dfrom, &c->d_dto, &c->d_syn};
    for (int k = 0; k < 6; ++k) {
        TRY(ensure(c, *dst[k], src[k]->bytes));
        HIP_TRY(c, hipMemcpy(dst[k]->p, src[k]->p, src[k]->bytes, hipMemcpyDeviceToDevice));
    }
    TRY(ensure(c, c->d_D[ri], (size_t)std::max<long long>(d_base, 1) * esz));
    TRY(ensure(c, c->d_Vp[ri], (size_t)std::max<long long>(v_base, 1) * esz));
    TRY(ensure(c, c->d_Uspill[ri], (size_t)std::max<long long>(u_base, 1) * esz));
    TRY(ensure(c, c->d_E, (size_t)std::max<long long>(e_base, 1) * 4));
    TRY(ensure(c, c->d_G, (size_t)std::max<long long>(g_base, 1) * 4));
    TRY(ensure(c, c->d_Lroot, (size_t)pbase * 8));
    TRY(ensure(c, c->d_Eroot, (size_t)pbase * 4));
    TRY(ensure(c, c->d_site, (size_t)pbase * 8));
    TRY(ensure(c, c->d_Prow[ri], (size_t)node0 * PRS * esz));
    TRY(ensure(c, c->d_PA[ri], (size_t)node0 * 4096 * esz));
    TRY(ensure(c, c->d_PTA[ri], (size_t)node0 * 4096 * esz));
    TRY(upload_identity_tiles(c, ri, esz));
    c->ident_tiles[ri] = c->d_PA[ri].p;
    TRY(ensure(c, c->d_tau, (size_t)node0 * 8));
    TRY(ensure(c, c->d_lam, (size_t)mat0 * 64 * 8));
    TRY(ensure(c, c->d_U, (size_t)mat0 * 4096 * 8));
    TRY(ensure(c, c->d_V, (size_t)mat0 * 4096 * 8));
    TRY(ensure(c, c->d_Z, (size_t)mat0 * 4096 * 8));
    TRY(ensure(c, c->d_mu, (size_t)mat0 * 8));
    TRY(ensure(c, c->d_M, (size_t)mat0 * 6 * 4096 * 8));
    TRY(ensure(c, c->d_dmu, (size_t)mat0 * 6 * 8));
    TRY(ensure(c, c->d_Apart, (size_t)br0 * 4096 * 8));
    TRY(ensure(c, c->d_fr, (size_t)br0 * max_comp * 8 * 8));
    TRY(ensure(c, c->d_proj, (size_t)mat0 * 6 * c->n_grp * 8));
    TRY(ensure(c, c->d_params, (size_t)c->pl.total * 8));
    (void)tile_elems;
    if (c->h_params_cap < (size_t)c->pl.total) {
        if (c->h_params) (void)hipHostFree(c->h_params);
        c->h_params = nullptr;
        c->h_params_cap = 0;
        HIP_TRY(c, hipHostMalloc((void **)&c->h_params, sizeof(double) * c->pl.total, kZeroCopyFlags));
        c->h_params_cap = c->pl.total;
        HIP_TRY(c, hipHostGetDevicePointer((void **)&c->h_params_dev, c->h_params, 0));
    }
    if (c->h_res_cap < res_off) {
        if (c->h_res) (void)hipHostFree(c->h_res);
        c->h_res = nullptr;
        c->h_res_cap = 0;
        HIP_TRY(c, hipHostMalloc((void **)&c->h_res, sizeof(double) * res_off, kZeroCopyFlags));
        c->h_res_cap = res_off;
        HIP_TRY(c, hipHostGetDevicePointer((void **)&c->h_res_dev, c->h_res, 0));
    }
    if (c->h_status) (void)hipHostFree(c->h_status);
    c->h_status = nullptr;
    c->h_status_dev = nullptr;
    HIP_TRY(c, hipHostMalloc((void **)&c->h_status, sizeof(int) * std::max(mat0, 64), kZeroCopyFlags));
    HIP_TRY(c, hipHostGetDevicePointer((void **)&c->h_status_dev, c->h_status, 0));
    if (c->h_site_cap < (size_t)pbase) {
        if (c->h_site) (void)hipHostFree(c->h_site);
        c->h_site = nullptr;
        c->h_site_cap = 0;
        HIP_TRY(c, hipHostMalloc((void **)&c->h_site, sizeof(double) * pbase, hipHostMallocDefault));
        c->h_site_cap = pbase;
    }
    ++c->epoch;
    gs.members = hs;
    gs.epochs.clear();
    for (hy_engine h : hs) gs.epochs.push_back(h->epoch);
    gs.prec = prec;
    return HY_OK;
}

int eval_fused(GroupState &gs, hy_batch_item *items, int n, int prec, unsigned flags) {
    hy_engine c = gs.ctx;
    const bool grad = (flags & HY_WANT_GRAD) != 0;
    bool want_site = false;
    for (int i = 0; i < n; ++i) want_site |= items[i].out_site_ll != nullptr;
    HIP_TRY(c, hipSetDevice(c->device));
    double *hp = c->h_params;
    const ParamLayout &pl = c->pl;
    for (int g = 0; g < n; ++g) {
        const hy_batch_item &it = items[g];
        const GroupState::Slice &sl = gs.slices[g];
        for (int k = 0; k < sl.n_ext; ++k) hp[pl.off_bl + sl.node0 + k] = k < sl.n_real ? it.bl[k] : 0.0;
        hp[pl.off_alpha + g] = it.alpha;
        for (int m = 0; m < sl.n_mat; ++m) hp[pl.off_omega + sl.mat0 + m] = it.omegas[m];
        for (int j = 0; j < sl.n_comp; ++j) hp[pl.off_w + sl.comp0 + j] = it.comp_w[j];
        for (int k = 0; k < 5; ++k) hp[pl.off_theta + 5 * g + k] = it.theta[k];
        hp[pl.off_delta + g] = it.delta;
        hp[pl.off_psi + g] = it.psi;
    }
    int grc = GRAPH_NO_SLOT;
    if (c->use_graphs && !c->profile) grc = launch_graph(c, prec, grad);
    if (grc < 0) return grc;
    if (grc == GRAPH_NO_SLOT) {
        HIP_TRY(c, hipMemcpyAsync(c->d_params.p, hp, sizeof(double) * pl.total, hipMemcpyHostToDevice, c->stream));
        const int rc = prec == HY_PREC_FP64 ? run_eval<double>(c, flags, nullptr) : run_eval<float>(c, flags, nullptr);
        if (rc) return rc;
    }
    if (want_site)
        HIP_TRY(c, hipMemcpyAsync(c->h_site, c->d_site.p, sizeof(double) * c->P_pad, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    TRY(finish_host(c, true, nullptr, nullptr, nullptr, nullptr, nullptr, false));
    for (int g = 0; g < n; ++g) {
        hy_batch_item &it = items[g];
        const GroupState::Slice &sl = gs.slices[g];
        const double *r = c->h_res + sl.res_off;
        if (it.out_logL) *it.out_logL = r[0];
        if (it.out_site_ll) std::memcpy(it.out_site_ll, c->h_site + sl.pbase, sizeof(double) * sl.P);
        if (grad) {
            const double *gbl = r + 1, *gom = gbl + sl.n_real, *gw = gom + sl.n_mat, *gth = gw + sl.n_comp;
            if (it.out_grad_bl) std::memcpy(it.out_grad_bl, gbl, sizeof(double) * sl.n_real);
            if (it.out_grad_omega) std::memcpy(it.out_grad_omega, gom, sizeof(double) * sl.n_mat);
            if (it.out_grad_w) std::memcpy(it.out_grad_w, gw, sizeof(double) * sl.n_comp);
            if (it.out_grad_theta) std::memcpy(it.out_grad_theta, gth, sizeof(double) * 5);
        }
        it.status = HY_OK;
    }
    return HY_OK;
}

}  // namespace

extern "C" {

int hy_eval_batch(int n_items, hy_batch_item *items, int precision, unsigned flags) {
    // Every item's evaluation is queued on its own handle's stream before any
    // is waited for, so independent genes overlap on the device.
    if (n_items < 0 || (n_items > 0 && !items)) return HY_ERR_ARG;
    {
        std::vector<hy_engine> hs;
        for (int i = 0; i < n_items; ++i)
            if (items[i].h) hs.push_back(items[i].h);
        std::sort(hs.begin(), hs.end());
        if (std::adjacent_find(hs.begin(), hs.end()) != hs.end()) {
            for (int i = 0; i < n_items; ++i) items[i].status = HY_ERR_ARG;
            return HY_ERR_ARG;  // a handle holds one pending evaluation at a time
        }
    }
    {
        // one fused evaluation when every item is a distinct engine of one group
        std::vector<hy_engine> hs;
        bool ok = true;
        for (int i = 0; i < n_items && ok; ++i) {
            hy_batch_item &it = items[i];
            ok = it.h && it.bl && it.omegas && it.comp_w && it.theta;
            if (ok) hs.push_back(it.h);
        }
        if (ok && (precision == HY_PREC_FP64 || precision == HY_PREC_FP32) && fusable(hs) && !hs[0]->group->no_fuse) {
            GroupState &gs = *hs[0]->group;
            bool same = gs.members == hs && gs.prec == precision;
            for (size_t i = 0; same && i < hs.size(); ++i) same = gs.epochs[i] == hs[i]->epoch;
            int rc = same ? HY_OK : build_batch(gs, hs, precision);
            if (rc == HY_OK) rc = eval_fused(gs, items, n_items, precision, flags);
            if (rc != HY_OK) {
                gs.members.clear();  // rebuild on the next call
                for (int i = 0; i < n_items; ++i) items[i].status = rc;
                hs[0]->err = gs.ctx->err;
            }
            return rc;
        }
    }
    int first = HY_OK;
    std::vector<char> queued(n_items, 0);
    for (int i = 0; i < n_items; ++i) {
        hy_batch_item &it = items[i];
        hy_engine h = it.h;
        it.status = HY_OK;
        if (!h) {
            it.status = HY_ERR_ARG;
        } else if ((it.status = check_eval_args(h, precision, it.bl, it.omegas, it.comp_w, it.theta, "hy_eval_batch")) ==
                   HY_OK) {
            if (h->P == 0) {
                zero_outputs(h, flags, it.out_logL, it.out_grad_bl, it.out_grad_omega, it.out_grad_w, it.out_grad_theta);
            } else {
                it.status = eval_submit(h, precision, flags, it.bl, it.alpha, it.omegas, it.comp_w, it.theta, it.delta,
                                        it.psi, it.out_site_ll != nullptr);
                queued[i] = it.status == HY_OK;
            }
        }
        if (it.status != HY_OK && first == HY_OK) first = it.status;
    }
    for (int i = 0; i < n_items; ++i) {
        if (!queued[i]) continue;
        hy_batch_item &it = items[i];
        it.status = eval_complete(it.h, (flags & HY_WANT_GRAD) != 0, it.out_logL, it.out_site_ll, it.out_grad_bl,
                                  it.out_grad_omega, it.out_grad_w, it.out_grad_theta);
        if (it.status != HY_OK && first == HY_OK) first = it.status;
    }
    return first;
}

int hy_autotune_batch(int n_items, hy_batch_item *items, int precision, unsigned flags, int reps, int *out_tile,
                      double *out_ms) {
    // Times the fused batch evaluation at every pattern-tile width and keeps
    // the fastest for this precision in the group (the batch is rebuilt).
    if (n_items < 2 || !items || reps < 1 || reps > 1000) return HY_ERR_ARG;
    if (precision != HY_PREC_FP64 && precision != HY_PREC_FP32) return HY_ERR_ARG;
    std::vector<hy_engine> hs;
    for (int i = 0; i < n_items; ++i) {
        if (!items[i].h) return HY_ERR_ARG;
        hs.push_back(items[i].h);
    }
    if (!fusable(hs) || hs[0]->group->no_fuse)
        return fail(hs[0], HY_ERR_STATE, "hy_autotune_batch: the items are not one fusable group batch");
    GroupState &gs = *hs[0]->group;
    const int ri = precision == HY_PREC_FP64 ? 1 : 0;
    int best_T = 0;
    double best = 1e300;
    for (int T = 16; T <= max_tile(precision); T *= 2) {
        gs.tuned_T[ri] = T;
        gs.members.clear();  // rebuild the merged problem at this tile width
        int rc = hy_eval_batch(n_items, items, precision, flags);
        std::vector<double> t;
        for (int r = 0; rc == HY_OK && r < reps; ++r) {
            const auto t0 = std::chrono::steady_clock::now();
            rc = hy_eval_batch(n_items, items, precision, flags);
            t.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        }
        if (rc != HY_OK) {
            gs.tuned_T[ri] = 0;
            gs.members.clear();
            return rc;
        }
        std::sort(t.begin(), t.end());
        if (t[t.size() / 2] < best) {
            best = t[t.size() / 2];
            best_T = T;
        }
    }
    gs.tuned_T[ri] = best_T;
    gs.members.clear();
    if (out_tile) *out_tile = best_T;
    if (out_ms) *out_ms = best;
    return HY_OK;
}

}  // extern "C"

namespace {

int eval_from_P(hy_engine h, int precision, const double *P, const double *pi, double *out_logL,
                double *out_site_ll, unsigned flags, const char *who) {
    if (!h) return HY_ERR_ARG;
    if (!h->have_problem) return fail(h, HY_ERR_STATE, "%s before hy_set_problem", who);
    if (precision != HY_PREC_FP64 && precision != HY_PREC_FP32)
        return fail(h, HY_ERR_ARG, "%s: unknown precision %d", who, precision);
    if (!P || !pi) return fail(h, HY_ERR_ARG, "%s: NULL P or pi", who);
    if (h->P == 0) {
        if (out_logL) *out_logL = 0.0;
        h->part_ri = -1;
        return HY_OK;
    }
    HIP_TRY(h, hipSetDevice(h->device));
    TRY(prepare_buffers(h, precision));
    if (h->ck && (flags & RUN_OUTSIDE))
        return fail(h, HY_ERR_UNSUPPORTED,
                    "%s: the partials of this problem exceed the memory budget (checkpointed layout); the "
                    "workspace read-back needs every node's partials resident (raise hy_set_memory_budget)",
                    who);
    std::vector<double> pp(64, 0.0);
    for (int i = 0; i < h->S; ++i) pp[i] = pi[i];
    TRY(upload(h, h->d_pi_custom, pp.data(), 64 * sizeof(double)));
    TRY(upload(h, h->d_Phost, P, sizeof(double) * (size_t)h->num_nodes * h->S * h->S));
    const int rc = precision == HY_PREC_FP64 ? run_eval<double>(h, flags, pp.data())
                                             : run_eval<float>(h, flags, pp.data());
    if (rc) return rc;
    return finish_eval(h, false, out_logL, out_site_ll, nullptr, nullptr, nullptr, nullptr, false);
}

}  // namespace

extern "C" {

int hy_loglik_from_P(hy_engine h, int precision, const double *P, const double *pi, double *out_logL,
                     double *out_site_ll) {
    return eval_from_P(h, precision, P, pi, out_logL, out_site_ll, HY_WANT_LOGL, "hy_loglik_from_P");
}

int hy_inside_outside_from_P(hy_engine h, int precision, const double *P, const double *pi, double *out_logL,
                             double *out_site_ll) {
    return eval_from_P(h, precision, P, pi, out_logL, out_site_ll, HY_WANT_LOGL | RUN_OUTSIDE,
                       "hy_inside_outside_from_P");
}

}  // extern "C"

namespace {

// Byte offsets of the allocation-free path's buffers in the fixed scratch.
struct NoallocLayout {
    size_t P, pi, tips, w, sched, Prow, PA, PTA, D, E, L, LE, site, res, total;
};

NoallocLayout noalloc_layout(const hy_engine_s *h, size_t rs) {
    NoallocLayout o{};
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t off = at;
        at += (bytes + 255) / 256 * 256;
        return off;
    };
    const size_t Pp = (size_t)h->na_P_pad, ne = (size_t)h->n_ext;
    o.P = take((size_t)h->num_nodes * h->S * h->S * 8);
    o.pi = take(64 * 8);
    o.tips = take(h->na_tips.size() * sizeof(uint16_t));
    o.w = take(Pp * 8);
    o.sched = take(h->na_sched.size() * 4);
    o.Prow = take(ne * PRS * rs);
    o.PA = take(ne * 4096 * rs);
    o.PTA = take(ne * 4096 * rs);
    o.D = take((size_t)h->n_int * Pp * 64 * rs);
    o.E = take((size_t)h->n_int * Pp * 4);
    o.L = take(Pp * 8);
    o.LE = take(Pp * 4);
    o.site = take(Pp * 8);
    o.res = take(8);
    o.total = at;
    return o;
}

// Identity edge tiles (polytomy nodes) of one element type in the Prow and
// PA / PTA layouts, as upload_identity_tiles builds them.
template <typename R>
void identity_tiles(int S, std::array<R, PRS> &prow, std::array<R, 4096> &pa) {
    prow.fill(R(0));
    pa.fill(R(0));
    for (int i = 0; i < 64; ++i) prow[i * 64 + i] = R(1);
    for (int i = 0; i < S; ++i) prow[4096 + i] = R(1);
    for (int idx = 0; idx < 4096; ++idx) {
        const int w = idx >> 10, l = (idx >> 4) & 63, kk = idx & 15;
        if ((16 * w + (l & 15)) == (4 * kk + (l >> 4))) pa[idx] = R(1);
    }
}

template <typename R>
int loglik_noalloc(hy_engine h, const double *P, const double *pi, double *out_logL, double *out_site_ll,
                   const NoallocLayout &lay) {
    hipStream_t s = h->stream;
    char *base = (char *)h->scratch;  // device address from hipGetSymbolAddress
    auto at = [&](size_t off) { return (void *)(base + off); };
    // inputs: caller's P and pi, the problem's host layouts
    HIP_TRY(h, hipMemcpyAsync(at(lay.P), P, (size_t)h->num_nodes * h->S * h->S * 8, hipMemcpyHostToDevice, s));
    std::array<double, 64> pp{};
    for (int i = 0; i < h->S; ++i) pp[i] = pi[i];
    HIP_TRY(h, hipMemcpyAsync(at(lay.pi), pp.data(), sizeof(pp), hipMemcpyHostToDevice, s));
    HIP_TRY(h, hipMemcpyAsync(at(lay.tips), h->na_tips.data(), h->na_tips.size() * sizeof(uint16_t),
                              hipMemcpyHostToDevice, s));
    HIP_TRY(h, hipMemcpyAsync(at(lay.w), h->na_weights.data(), h->na_weights.size() * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(h, hipMemcpyAsync(at(lay.sched), h->na_sched.data(), h->na_sched.size() * 4, hipMemcpyHostToDevice, s));
    if (!h->na_ident.empty()) {
        // the scratch keeps whatever the last call on this device left:
        // identity tiles are written on every call
        std::array<R, PRS> prow;
        std::array<R, 4096> pa;
        identity_tiles<R>(h->S, prow, pa);
        for (int n : h->na_ident) {
            HIP_TRY(h, hipMemcpyAsync((char *)at(lay.Prow) + (size_t)n * PRS * sizeof(R), prow.data(), sizeof(prow),
                                      hipMemcpyHostToDevice, s));
            HIP_TRY(h, hipMemcpyAsync((char *)at(lay.PA) + (size_t)n * 4096 * sizeof(R), pa.data(), sizeof(pa),
                                      hipMemcpyHostToDevice, s));
            HIP_TRY(h, hipMemcpyAsync((char *)at(lay.PTA) + (size_t)n * 4096 * sizeof(R), pa.data(), sizeof(pa),
                                      hipMemcpyHostToDevice, s));
        }
        // (host arrays are read by the copies before this function returns)
        HIP_TRY(h, hipStreamSynchronize(s));
    }
    // P tiles of every branch
    PmatArgs<R> pa{};
    pa.blist = (const int32_t *)h->d_blist.p;
    pa.n_branch = h->n_branch;
    pa.S = h->S;
    pa.Prow = (R *)at(lay.Prow);
    pa.PA = (R *)at(lay.PA);
    pa.PTA = (R *)at(lay.PTA);
    if (h->n_branch > 0) launch_pmat_host<R>(pa, (const double *)at(lay.P), s);
    // inside pass over the whole tree, one launch of 16-pattern tiles; every
    // D, E, L, LE and site entry read later is written here first
    PruneArgs<R> pr{};
    pr.sched = (const int32_t *)at(lay.sched);
    pr.n_sched = (int)(h->na_sched.size() / SCHED_STRIDE);
    pr.tip_rows = h->num_taxa + 1;
    pr.S = h->S;
    pr.PA = (const R *)at(lay.PA);
    pr.PTA = (const R *)at(lay.PTA);
    pr.Prow = (const R *)at(lay.Prow);
    pr.tips = (const uint16_t *)at(lay.tips);
    pr.amb = (const uint64_t *)h->d_amb.p;
    pr.P_pad = h->na_P_pad;
    pr.n_tiles = h->na_P_pad / 16;
    pr.n_patterns = h->P;
    pr.pi = (const double *)at(lay.pi);
    pr.D = (R *)at(lay.D);
    pr.E = (int32_t *)at(lay.E);
    pr.Lroot = (double *)at(lay.L);
    pr.Eroot = (int32_t *)at(lay.LE);
    pr.site_ll = (double *)at(lay.site);
    pr.root_out = 1;
    launch_inside<R>(pr, 16, false, 1, s);
    // deterministic compensated lnL (no gradient)
    FinalArgs fn{};
    fn.site_ll = pr.site_ll;
    fn.weights = (const double *)at(lay.w);
    fn.Lroot = pr.Lroot;
    fn.n_patterns = h->P;
    fn.want_grad = 0;
    fn.n_real_nodes = h->num_nodes;
    fn.res = (double *)at(lay.res);
    launch_final(fn, 1, s);
    HIP_TRY(h, hipGetLastError());
    if (out_logL) HIP_TRY(h, hipMemcpyAsync(out_logL, at(lay.res), 8, hipMemcpyDeviceToHost, s));
    if (out_site_ll)
        HIP_TRY(h, hipMemcpyAsync(out_site_ll, at(lay.site), (size_t)h->P * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(h, hipStreamSynchronize(s));
    return HY_OK;
}

}  // namespace

extern "C" {

int hy_loglik_from_P_noalloc(hy_engine h, int precision, const double *P, const double *pi, double *out_logL,
                             double *out_site_ll) {
    const char *who = "hy_loglik_from_P_noalloc";
    if (!h) return HY_ERR_ARG;
    if (!h->have_problem) return fail(h, HY_ERR_STATE, "%s before hy_set_problem", who);
    if (precision != HY_PREC_FP64 && precision != HY_PREC_FP32)
        return fail(h, HY_ERR_ARG, "%s: unknown precision %d", who, precision);
    if (!P || !pi) return fail(h, HY_ERR_ARG, "%s: NULL P or pi", who);
    if (!h->scratch)
        return fail(h, HY_ERR_HIP, "%s: the fixed scratch of device %d is not addressable: %s", who, h->device,
                    hipGetErrorString(h->scratch_err));
    if (h->P == 0) {
        if (out_logL) *out_logL = 0.0;
        return HY_OK;
    }
    const size_t rs = precision == HY_PREC_FP64 ? 8 : 4;
    const NoallocLayout lay = noalloc_layout(h, rs);
    if (lay.total > NOALLOC_SCRATCH_BYTES)
        return fail(h, HY_ERR_ARG,
                    "%s: this problem (%d internal nodes x %d patterns, %s) needs %zu bytes of scratch, more than "
                    "the fixed %zu; use hy_loglik_from_P",
                    who, h->n_int, h->P, rs == 8 ? "FP64" : "FP32", lay.total, (size_t)NOALLOC_SCRATCH_BYTES);
    HIP_TRY(h, hipSetDevice(h->device));
    return rs == 8 ? loglik_noalloc<double>(h, P, pi, out_logL, out_site_ll, lay)
                   : loglik_noalloc<float>(h, P, pi, out_logL, out_site_ll, lay);
}

}  // extern "C"

namespace {

// ---- nucleotide GTR stage ----
// GTR exchangeability pairs in theta order AC, AT, CG, CT, GT (AG == 1).
constexpr int GTR_A[5] = {0, 0, 1, 1, 2}, GTR_B[5] = {1, 3, 2, 3, 3};

// Cyclic Jacobi eigensolver of a symmetric 4 x 4 matrix (row-major S,
// destroyed): eigenvalues lam, eigenvectors in the columns of Z.
void sym_eigen4(double *S, double *lam, double *Z) {
    for (int i = 0; i < 16; ++i) Z[i] = (i % 5 == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) (i == j ? tot : off) += S[4 * i + j] * S[4 * i + j];
        if (off <= 1e-32 * (tot + off)) break;
        for (int p = 0; p < 3; ++p)
            for (int q = p + 1; q < 4; ++q) {
                const double apq = S[4 * p + q];
                if (apq == 0.0) continue;
                const double th = (S[4 * q + q] - S[4 * p + p]) / (2.0 * apq);
                const double t = (th >= 0.0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < 4; ++k) {  // columns p, q
                    const double a = S[4 * k + p], b = S[4 * k + q];
                    S[4 * k + p] = c * a - sn * b;
                    S[4 * k + q] = sn * a + c * b;
                }
                for (int k = 0; k < 4; ++k) {  // rows p, q
                    const double a = S[4 * p + k], b = S[4 * q + k];
                    S[4 * p + k] = c * a - sn * b;
                    S[4 * q + k] = sn * a + c * b;
                }
                for (int k = 0; k < 4; ++k) {
                    const double a = Z[4 * k + p], b = Z[4 * k + q];
                    Z[4 * k + p] = c * a - sn * b;
                    Z[4 * k + q] = sn * a + c * b;
                }
            }
    }
    for (int i = 0; i < 4; ++i) lam[i] = S[5 * i];
}

// GTRMatrix::update (rate_matrix.hpp:29-90): normalised Q, mu, and the
// spectrum of D^1/2 Q D^-1/2 (pi floored at 1e-12 in the square roots,
// eigenvalues clamped <= 0): P(t) = U diag(e^{lam t}) V.
struct Gtr4 {
    double Q[16], mu, lam[4], U[16], V[16];
};

Gtr4 gtr_model(const double *theta, const double *pi) {
    Gtr4 g{};
    double r[16] = {0};
    for (int k = 0; k < 5; ++k) r[4 * GTR_A[k] + GTR_B[k]] = r[4 * GTR_B[k] + GTR_A[k]] = theta[k];
    r[4 * 0 + 2] = r[4 * 2 + 0] = 1.0;  // A <-> G
    for (int i = 0; i < 4; ++i) {
        double rs = 0.0;
        for (int j = 0; j < 4; ++j)
            if (j != i) rs += (g.Q[4 * i + j] = r[4 * i + j] * pi[j]);
        g.Q[5 * i] = -rs;
    }
    g.mu = 0.0;
    for (int i = 0; i < 4; ++i) g.mu -= pi[i] * g.Q[5 * i];
    if (g.mu > 0.0)
        for (double &q : g.Q) q /= g.mu;
    double sq[4], isq[4], S[16], Z[16];
    for (int i = 0; i < 4; ++i) {
        sq[i] = std::sqrt(std::max(pi[i], 1e-12));
        isq[i] = 1.0 / sq[i];
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) S[4 * i + j] = sq[i] * g.Q[4 * i + j] * isq[j];
    for (int i = 0; i < 4; ++i)
        for (int j = i + 1; j < 4; ++j) S[4 * i + j] = S[4 * j + i] = 0.5 * (S[4 * i + j] + S[4 * j + i]);
    sym_eigen4(S, g.lam, Z);
    for (int i = 0; i < 4; ++i) g.lam[i] = std::min(g.lam[i], 0.0);
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k) {
            g.U[4 * i + k] = isq[i] * Z[4 * i + k];
            g.V[4 * k + i] = Z[4 * i + k] * sq[i];
        }
    return g;
}

// transition_matrix (rate_matrix.hpp:93-99): U diag(e^{lam t}) V, clamped >= 0
void gtr_transition(const Gtr4 &g, double t, double *P) {
    double e[4];
    for (int k = 0; k < 4; ++k) e[k] = std::exp(g.lam[k] * t);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double v = 0.0;
            for (int k = 0; k < 4; ++k) v += g.U[4 * i + k] * e[k] * g.V[4 * k + j];
            P[4 * i + j] = std::max(v, 0.0);
        }
}

// Node tables and padded patterns of the GTR stage (once per problem).
int gtr_prepare(hy_engine h) {
    if (h->gtr_ready) return HY_OK;
    const auto &ch = h->ext_children;
    const int n_ext = (int)ch.size();
    std::vector<int32_t> nref(n_ext), nbr(n_ext, -1), sched;
    for (int n = 0; n < n_ext; ++n) {
        if (ch[n].empty()) {
            const int row = h->leaf_taxon_ext[n] >= 0 ? h->leaf_taxon_ext[n] : h->num_taxa;
            nref[n] = (KIND_TIP << 28) | row;
        } else {
            nref[n] = (KIND_INT << 28) | h->int_slot[n];
        }
    }
    for (int bi = 0; bi < h->n_branch; ++bi) nbr[h->blist[(size_t)bi * BR_STRIDE]] = bi;
    h->g_ident.assign(n_ext, 0);
    for (int n = 0; n < n_ext; ++n) h->g_ident[n] = h->ext_ident[n];
    for (int u : h->order) {
        int32_t e[GTR_STRIDE] = {u, h->int_slot[u], (int32_t)ch[u].size(), 0, 0, 0, 0, 0};
        for (size_t c = 0; c < ch[u].size(); ++c) e[3 + c] = ch[u][c];
        if (ch[u].empty()) {  // a single-leaf tree: the root's own tip on an identity edge
            e[2] = 1;
            e[3] = u;
            nref[u] = (KIND_TIP << 28) | (h->leaf_taxon_ext[u] >= 0 ? h->leaf_taxon_ext[u] : h->num_taxa);
            h->g_ident[u] = 1;
        }
        sched.insert(sched.end(), e, e + GTR_STRIDE);
    }
    if (sched.empty() || sched[sched.size() - GTR_STRIDE] != h->ext_root)
        return fail(h, HY_ERR_STATE, "internal: GTR schedule does not end at the root");
    const int Pp = std::max(GTR_BLOCK, (h->P + GTR_BLOCK - 1) / GTR_BLOCK * GTR_BLOCK);
    std::vector<uint16_t> tips((size_t)(h->num_taxa + 1) * Pp, TIP_ONES);
    for (int t = 0; t < h->num_taxa; ++t)
        std::memcpy(&tips[(size_t)t * Pp], &h->codes[(size_t)t * h->P], sizeof(uint16_t) * h->P);
    std::vector<double> w(Pp, 0.0);
    std::copy(h->weights.begin(), h->weights.end(), w.begin());
    TRY(upload(h, h->d_gsched, sched.data(), sched.size() * 4));
    TRY(upload(h, h->d_gnref, nref.data(), nref.size() * 4));
    TRY(upload(h, h->d_gnbr, nbr.data(), nbr.size() * 4));
    TRY(upload(h, h->d_gtips, tips.data(), tips.size() * 2));
    TRY(upload(h, h->d_gw, w.data(), w.size() * 8));
    h->g_P_pad = Pp;
    h->g_n_sched = (int)(sched.size() / GTR_STRIDE);
    h->gtr_ready = true;
    return HY_OK;
}

template <typename R>
int gtr_run(hy_engine h, bool grad, const Gtr4 &g, const double *bl, const double *pi, double *out_logL,
            double *out_site_ll, std::vector<double> &A) {
    const size_t rs = sizeof(R), Pp = (size_t)h->g_P_pad, ni = (size_t)std::max(h->n_int, 1);
    const int n_ext = h->n_ext, n_waves = (int)(Pp / 64);
    TRY(ensure(h, h->d_gP, (size_t)n_ext * 16 * rs));
    TRY(ensure(h, h->d_gpi, 4 * 8));
    TRY(ensure(h, h->d_gD, ni * 4 * Pp * rs));
    TRY(ensure(h, h->d_gE, ni * Pp * 4));
    TRY(ensure(h, h->d_gL, Pp * 8));
    TRY(ensure(h, h->d_gLE, Pp * 4));
    TRY(ensure(h, h->d_gsite, Pp * 8));
    TRY(ensure(h, h->d_gres, 8));
    if (grad) {
        TRY(ensure(h, h->d_gU, ni * 4 * Pp * rs));
        TRY(ensure(h, h->d_gH, ni * Pp * 4));
        TRY(ensure(h, h->d_gApart, (size_t)std::max(h->n_branch, 1) * n_waves * 16 * 8));
        TRY(ensure(h, h->d_gA, (size_t)std::max(h->n_branch, 1) * 16 * 8));
    }
    // transition matrices of every extended node (identity edges and the root: I)
    std::vector<R> Pm((size_t)n_ext * 16, R(0));
    for (int n = 0; n < n_ext; ++n) {
        double P[16];
        if (h->g_ident[n] || n == h->ext_root) {
            for (int i = 0; i < 16; ++i) P[i] = (i % 5 == 0) ? 1.0 : 0.0;
        } else {
            gtr_transition(g, bl[n], P);
        }
        for (int i = 0; i < 16; ++i) Pm[(size_t)n * 16 + i] = (R)P[i];
    }
    hipStream_t s = h->stream;
    HIP_TRY(h, hipMemcpyAsync(h->d_gP.p, Pm.data(), Pm.size() * rs, hipMemcpyHostToDevice, s));
    HIP_TRY(h, hipMemcpyAsync(h->d_gpi.p, pi, 4 * 8, hipMemcpyHostToDevice, s));
    GtrArgs<R> a{};
    a.sched = (const int32_t *)h->d_gsched.p;
    a.n_sched = h->g_n_sched;
    a.nref = (const int32_t *)h->d_gnref.p;
    a.nbr = (const int32_t *)h->d_gnbr.p;
    a.Pm = (const R *)h->d_gP.p;
    a.tips = (const uint16_t *)h->d_gtips.p;
    a.amb = (const uint64_t *)h->d_amb.p;
    a.pi = (const double *)h->d_gpi.p;
    a.weights = (const double *)h->d_gw.p;
    a.P_pad = (int)Pp;
    a.P = h->P;
    a.n_waves = n_waves;
    a.D = (R *)h->d_gD.p;
    a.E = (int32_t *)h->d_gE.p;
    a.U = (R *)h->d_gU.p;
    a.H = (int32_t *)h->d_gH.p;
    a.Lroot = (double *)h->d_gL.p;
    a.Eroot = (int32_t *)h->d_gLE.p;
    a.site_ll = (double *)h->d_gsite.p;
    a.Apart = (double *)h->d_gApart.p;
    launch_gtr<R>(a, grad, s);
    GtrReduceArgs r{};
    r.Apart = a.Apart;
    r.site_ll = a.site_ll;
    r.weights = a.weights;
    r.n_branch = h->n_branch;
    r.n_waves = n_waves;
    r.P = h->P;
    r.A = (double *)h->d_gA.p;
    r.res = (double *)h->d_gres.p;
    launch_gtr_reduce(r, grad, s);
    HIP_TRY(h, hipGetLastError());
    double ll = 0.0;
    HIP_TRY(h, hipMemcpyAsync(&ll, h->d_gres.p, 8, hipMemcpyDeviceToHost, s));
    if (out_site_ll) HIP_TRY(h, hipMemcpyAsync(out_site_ll, h->d_gsite.p, (size_t)h->P * 8, hipMemcpyDeviceToHost, s));
    if (grad) {
        A.assign((size_t)h->n_branch * 16, 0.0);
        if (h->n_branch > 0)
            HIP_TRY(h, hipMemcpyAsync(A.data(), h->d_gA.p, A.size() * 8, hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(h, hipStreamSynchronize(s));
    if (out_logL) *out_logL = ll;
    return HY_OK;
}

}  // namespace

extern "C" {

int hy_gtr_eval(hy_engine h, int precision, unsigned flags, const double *bl, const double *theta, const double *freqs,
                double *out_logL, double *out_site_ll, double *out_grad_bl, double *out_grad_theta) {
    const char *who = "hy_gtr_eval";
    if (!h) return HY_ERR_ARG;
    if (!h->have_problem) return fail(h, HY_ERR_STATE, "%s before hy_set_problem", who);
    if (h->S != 4) return fail(h, HY_ERR_STATE, "%s: the problem has %d states; the GTR stage needs S = 4", who, h->S);
    if (precision != HY_PREC_FP64 && precision != HY_PREC_FP32)
        return fail(h, HY_ERR_ARG, "%s: unknown precision %d", who, precision);
    if (!bl || !theta || !freqs) return fail(h, HY_ERR_ARG, "%s: NULL branch lengths, rates or frequencies", who);
    for (int k = 0; k < 5; ++k)
        if (!(theta[k] >= 0.0 && std::isfinite(theta[k])))
            return fail(h, HY_ERR_ARG, "%s: theta[%d] = %g", who, k, theta[k]);
    for (int k = 0; k < 4; ++k)
        if (!(freqs[k] >= 0.0 && freqs[k] <= 1.0)) return fail(h, HY_ERR_ARG, "%s: freqs[%d] = %g", who, k, freqs[k]);
    for (int n = 0; n < h->num_nodes; ++n)
        if (n != h->root && !(bl[n] >= 0.0 && std::isfinite(bl[n])))
            return fail(h, HY_ERR_ARG, "%s: branch length %d = %g", who, n, bl[n]);
    const bool grad = (flags & HY_WANT_GRAD) != 0;
    if (h->P == 0) {
        if (out_logL) *out_logL = 0.0;
        if (grad && out_grad_bl) std::fill(out_grad_bl, out_grad_bl + h->num_nodes, 0.0);
        if (grad && out_grad_theta) std::fill(out_grad_theta, out_grad_theta + 5, 0.0);
        return HY_OK;
    }
    HIP_TRY(h, hipSetDevice(h->device));
    TRY(gtr_prepare(h));
    const Gtr4 g = gtr_model(theta, freqs);
    // branch lengths of the extended nodes (identity edges: unused)
    std::vector<double> blx(h->n_ext, 0.0);
    for (int n = 0; n < h->num_nodes; ++n) blx[n] = (n == h->root) ? 0.0 : bl[n];
    std::vector<double> A;
    TRY(precision == HY_PREC_FP64
            ? gtr_run<double>(h, grad, g, blx.data(), freqs, out_logL, (flags & HY_WANT_SITE_LL) ? out_site_ll : nullptr, A)
            : gtr_run<float>(h, grad, g, blx.data(), freqs, out_logL, (flags & HY_WANT_SITE_LL) ? out_site_ll : nullptr, A));
    if (!grad) return HY_OK;
    // d lnL / d t_b = <A_b, Q P_b>; d lnL / d theta_k = sum_b <U^T A_b V^T, (V dQ_k U) o E(t_b)>
    // with dQ_k the derivative of the normalised generator
    std::vector<double> gbl(h->num_nodes, 0.0);
    double gth[5] = {0, 0, 0, 0, 0}, M[5][16];
    for (int k = 0; k < 5; ++k) {
        const int a = GTR_A[k], b = GTR_B[k];
        double dQ[16] = {0};
        dQ[4 * a + b] = freqs[b];
        dQ[4 * b + a] = freqs[a];
        dQ[5 * a] = -freqs[b];
        dQ[5 * b] = -freqs[a];
        const double dmu = 2.0 * freqs[a] * freqs[b];
        double dn[16], T[16];
        for (int i = 0; i < 16; ++i) dn[i] = g.mu > 0.0 ? (dQ[i] - g.Q[i] * dmu) / g.mu : dQ[i];
        for (int i = 0; i < 4; ++i)  // T = dn U
            for (int j = 0; j < 4; ++j) {
                double v = 0.0;
                for (int q = 0; q < 4; ++q) v += dn[4 * i + q] * g.U[4 * q + j];
                T[4 * i + j] = v;
            }
        for (int i = 0; i < 4; ++i)  // M = V T
            for (int j = 0; j < 4; ++j) {
                double v = 0.0;
                for (int q = 0; q < 4; ++q) v += g.V[4 * i + q] * T[4 * q + j];
                M[k][4 * i + j] = v;
            }
    }
    for (int bi = 0; bi < h->n_branch; ++bi) {
        const int n = h->blist[(size_t)bi * BR_STRIDE];
        const double t = blx[n], *Ab = &A[(size_t)bi * 16];
        double P[16], QP[16], At[16], T[16], e[4];
        gtr_transition(g, t, P);
        double d = 0.0;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                double v = 0.0;
                for (int q = 0; q < 4; ++q) v += g.Q[4 * i + q] * P[4 * q + j];
                QP[4 * i + j] = v;
                d += Ab[4 * i + j] * v;
            }
        gbl[n] = d;
        // At = U^T A V^T
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                double v = 0.0;
                for (int q = 0; q < 4; ++q) v += Ab[4 * i + q] * g.V[4 * j + q];
                T[4 * i + j] = v;
            }
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                double v = 0.0;
                for (int q = 0; q < 4; ++q) v += g.U[4 * q + i] * T[4 * q + j];
                At[4 * i + j] = v;
            }
        for (int i = 0; i < 4; ++i) e[i] = std::exp(g.lam[i] * t);
        for (int k = 0; k < 5; ++k) {
            double v = 0.0;
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) {
                    const double dl = g.lam[i] - g.lam[j];
                    const double E = std::fabs(dl) < 1e-9 ? t * e[i] : (e[i] - e[j]) / dl;
                    v += At[4 * i + j] * M[k][4 * i + j] * E;
                }
            gth[k] += v;
        }
    }
    if (out_grad_bl) std::copy(gbl.begin(), gbl.end(), out_grad_bl);
    if (out_grad_theta) std::copy(gth, gth + 5, out_grad_theta);
    return HY_OK;
}

// Grows a device buffer to `bytes`, keeping its first `keep` bytes.
static int grow_keep(hy_engine h, DevBuf &b, size_t bytes, size_t keep) {
    if (b.p && b.bytes >= bytes) return HY_OK;
    DevBuf nb;
    HIP_TRY(h, hipMalloc(&nb.p, bytes));
    nb.bytes = bytes;
    h->device_bytes += bytes;
    if (b.p && keep) HIP_TRY(h, hipMemcpyAsync(nb.p, b.p, std::min(keep, b.bytes), hipMemcpyDeviceToDevice, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    release(h, b);
    b = nb;
    ++h->epoch;
    return HY_OK;
}

// The spectrum pool holds `cap` slots (lam, U, V, mu, status, Z per slot).
static int fel_pool_reserve(hy_engine h, int cap) {
    FelPool &f = h->fel;
    if (cap <= f.cap) return HY_OK;
    const size_t o = (size_t)f.cap, n = (size_t)cap;
    TRY(grow_keep(h, h->d_flam, n * 64 * 8, o * 64 * 8));
    TRY(grow_keep(h, h->d_fU, n * 4096 * 8, o * 4096 * 8));
    TRY(grow_keep(h, h->d_fV, n * 4096 * 8, o * 4096 * 8));
    TRY(grow_keep(h, h->d_fZ, n * 4096 * 8, o * 4096 * 8));
    TRY(grow_keep(h, h->d_fmu, n * 8, o * 8));
    TRY(grow_keep(h, h->d_fstatus, n * 4, o * 4));
    f.omega.resize(n, std::numeric_limits<double>::quiet_NaN());
    f.used.resize(n, 0);
    f.cap = cap;
    return HY_OK;
}

// Spectra of a list of omegas (MG94 at alpha = 1 with the given theta,
// delta, psi) from the handle's pool, shared by the per-site refit entry
// points: pooled omegas are reused, every other distinct omega is solved in
// one k_model launch on the stream, refining the eigenvectors of the pooled
// omega nearest in log(omega).  pool_commit() enters the new spectra once
// the stream has finished.
struct PoolSolve {
    std::vector<int32_t> slot;    // per requested omega
    std::vector<double> miss_om;  // omegas solved by this call
    std::vector<int32_t> oslot;   // and their slots
    uint64_t tick = 0;
};

static int pool_solve(hy_engine h, const char *who, const double *theta, double delta, double psi, const double *bl,
                      const std::vector<double> &om, PoolSolve &ps) {
    for (int n = 0; n < h->num_nodes; ++n)
        if (n != h->root && !(bl[n] >= 0.0 && std::isfinite(bl[n])))
            return fail(h, HY_ERR_ARG, "%s: branch length %d = %g", who, n, bl[n]);
    for (int i = 0; i < 5; ++i)
        if (!(theta[i] >= 0.0 && std::isfinite(theta[i])))
            return fail(h, HY_ERR_ARG, "%s: theta[%d] = %g", who, i, theta[i]);
    if (!(delta >= 0.0 && psi >= 0.0 && std::isfinite(delta) && std::isfinite(psi)))
        return fail(h, HY_ERR_ARG, "%s: delta %g, psi %g", who, delta, psi);
    HIP_TRY(h, hipSetDevice(h->device));
    TRY(gtr_prepare(h));  // pattern-walk schedule, node tables, padded tip codes
    hipStream_t s = h->stream;
    // per-problem tables: identity edges, branch lengths (uploaded when they change)
    if (!h->f_ident_ready) {
        std::vector<uint8_t> ident(h->g_ident.begin(), h->g_ident.end());
        TRY(upload(h, h->d_fident, ident.data(), ident.size()));
        h->f_ident_ready = true;
    }
    {
        std::vector<double> blx(h->n_ext, 0.0);
        for (int n = 0; n < h->num_nodes; ++n) blx[n] = (n == h->root) ? 0.0 : bl[n];
        if (blx != h->f_bl || !h->d_fbl.p) {
            TRY(upload(h, h->d_fbl, blx.data(), blx.size() * 8));
            h->f_bl = blx;
        }
    }
    // spectrum pool: a new model key empties it
    FelPool &f = h->fel;
    const double key[8] = {theta[0], theta[1], theta[2], theta[3], theta[4], delta, psi, (double)h->model_gen};
    if (!f.keyed || std::memcmp(key, f.key, sizeof key) != 0) {
        f.clear();
        std::memcpy(f.key, key, sizeof key);
        f.keyed = true;
    }
    const uint64_t tick = ps.tick = ++f.tick;
    // distinct omegas of the call: pooled ones are hits, the others are solved
    const size_t K = om.size();
    std::vector<int32_t> &spec = ps.slot;
    spec.assign(K, -1);
    std::vector<double> &miss_om = ps.miss_om;
    miss_om.clear();
    std::unordered_map<uint64_t, int> call;  // omega bits -> slot, or -(1 + miss index)
    int n_hit = 0;
    for (size_t k = 0; k < K; ++k) {
        uint64_t bits;
        std::memcpy(&bits, &om[k], 8);
        auto c = call.find(bits);
        if (c != call.end()) {
            spec[k] = c->second;
            continue;
        }
        auto it = f.slot_of.find(bits);
        if (it != f.slot_of.end()) {
            f.used[it->second] = tick;
            call[bits] = spec[k] = it->second;
            ++n_hit;
            ++f.n_hit;
        } else {
            call[bits] = spec[k] = -1 - (int)miss_om.size();
            miss_om.push_back(om[k]);
        }
    }
    const int n_miss = (int)miss_om.size();
    if (n_hit + n_miss > f.cap) TRY(fel_pool_reserve(h, std::max({n_hit + n_miss, 2 * f.cap, 64})));
    // output slots of the misses: free slots first, then the least recently
    // used; never a slot this call reads
    std::vector<int32_t> &oslot = ps.oslot;
    oslot.assign(n_miss, 0);
    std::vector<int32_t> zsrc(n_miss, -1);
    if (n_miss > 0) {
        std::vector<int> cand;
        cand.reserve(f.cap);
        for (int sl = 0; sl < f.cap; ++sl)
            if (f.used[sl] != tick) cand.push_back(sl);
        if ((int)cand.size() < n_miss) return fail(h, HY_ERR_STATE, "internal: spectrum pool too small");
        std::partial_sort(cand.begin(), cand.begin() + n_miss, cand.end(), [&](int x, int y) {
            const bool fx = std::isnan(f.omega[x]), fy = std::isnan(f.omega[y]);
            if (fx != fy) return fx;
            return f.used[x] != f.used[y] ? f.used[x] < f.used[y] : x < y;
        });
        for (int i = 0; i < n_miss; ++i) {
            oslot[i] = cand[i];
            f.drop(cand[i]);
        }
    }
    // warm start from the pooled spectrum nearest in log(omega)
    auto dist = [](double x, double y) {
        return std::fabs(std::log(std::max(x, 1e-300)) - std::log(std::max(y, 1e-300)));
    };
    for (int i = 0; i < n_miss && !f.by_omega.empty(); ++i) {
        auto hi = f.by_omega.lower_bound(miss_om[i]);
        int best = -1;
        double bd = 0.0;
        if (hi != f.by_omega.end()) {
            best = hi->second;
            bd = dist(hi->first, miss_om[i]);
        }
        if (hi != f.by_omega.begin()) {
            auto lo = std::prev(hi);
            if (best < 0 || dist(lo->first, miss_om[i]) < bd) best = lo->second;
        }
        zsrc[i] = best;
    }
    for (size_t k = 0; k < K; ++k)
        if (spec[k] < 0) spec[k] = oslot[-1 - spec[k]];
    if (n_miss == 0) return HY_OK;
    // staging of the solve: [params] doubles, then [oslot][zsrc] ints
    ParamLayout pl{};
    pl.n_nodes = 0;
    pl.n_mat = n_miss;
    pl.n_comp = 1;
    pl.off_bl = 0;
    pl.off_alpha = 0;
    pl.off_omega = 1;
    pl.off_w = 1 + n_miss;
    pl.off_theta = pl.off_w + 1;
    pl.off_delta = pl.off_theta + 5;
    pl.off_psi = pl.off_delta + 1;
    pl.off_seq = pl.off_psi + 1;
    pl.total = pl.off_seq + 1;
    const size_t bytes = (size_t)pl.total * 8 + 2 * (size_t)n_miss * 4;
    if (h->p_pin_bytes < bytes) {
        if (h->p_pin) (void)hipHostFree(h->p_pin);
        h->p_pin = nullptr;
        h->p_pin_bytes = 0;
        HIP_TRY(h, hipHostMalloc(&h->p_pin, bytes * 2, hipHostMallocDefault));
        h->p_pin_bytes = bytes * 2;
    }
    double *pd = (double *)h->p_pin;
    std::fill(pd, pd + pl.total, 0.0);
    pd[pl.off_alpha] = 1.0;
    std::copy(miss_om.begin(), miss_om.end(), pd + pl.off_omega);
    pd[pl.off_w] = 1.0;
    std::copy(theta, theta + 5, pd + pl.off_theta);
    pd[pl.off_delta] = delta;
    pd[pl.off_psi] = psi;
    int32_t *pi32 = (int32_t *)(pd + pl.total);
    std::copy(oslot.begin(), oslot.end(), pi32);
    std::copy(zsrc.begin(), zsrc.end(), pi32 + n_miss);
    TRY(ensure(h, h->d_pstage, bytes));
    HIP_TRY(h, hipMemcpyAsync(h->d_pstage.p, h->p_pin, bytes, hipMemcpyHostToDevice, s));
    const double *dd = (const double *)h->d_pstage.p;
    const int32_t *di = (const int32_t *)(dd + pl.total);
    ModelArgs ma{};
    ma.qdesc = (const int32_t *)h->d_qdesc.p;
    ma.nd = (const int8_t *)h->d_nd.p;
    ma.dpos = (const int8_t *)h->d_dpos.p;
    ma.dfrom = (const int8_t *)h->d_dfrom.p;
    ma.dto = (const int8_t *)h->d_dto.p;
    ma.syn = (const uint8_t *)h->d_syn.p;
    ma.pos = (const double *)h->d_pos.p;
    ma.pi = (const double *)h->d_pi.p;
    ma.S = h->S;
    ma.params = dd;
    ma.pl = pl;
    ma.lam = (double *)h->d_flam.p;
    ma.U = (double *)h->d_fU.p;
    ma.V = (double *)h->d_fV.p;
    ma.mu = (double *)h->d_fmu.p;
    ma.Z = (double *)h->d_fZ.p;
    ma.warm = 1;
    ma.oslot = di;
    ma.zsrc = di + n_miss;
    ma.status = (int *)h->d_fstatus.p;
    ma.tol = 1e-14;
    launch_model(ma, n_miss, s);
    HIP_TRY(h, hipGetLastError());
    return HY_OK;
}

// After the stream has finished: the spectra solved by pool_solve join the
// pool; one whose eigensolver failed is dropped and reported.
static int pool_commit(hy_engine h, const char *who, const PoolSolve &ps) {
    FelPool &f = h->fel;
    const int n_miss = (int)ps.miss_om.size();
    if (n_miss == 0) return HY_OK;
    std::vector<int> st(f.cap);
    HIP_TRY(h, hipMemcpy(st.data(), h->d_fstatus.p, (size_t)f.cap * 4, hipMemcpyDeviceToHost));
    int bad = -1;
    for (int i = 0; i < n_miss; ++i) {
        const int sl = ps.oslot[i];
        if (st[sl] < 0) {
            bad = i;
            continue;
        }
        f.omega[sl] = ps.miss_om[i];
        f.used[sl] = ps.tick;
        uint64_t bits;
        std::memcpy(&bits, &ps.miss_om[i], 8);
        f.slot_of[bits] = sl;
        f.by_omega[ps.miss_om[i]] = sl;
        if (st[sl] >= 100)
            ++f.n_warm;
        else
            ++f.n_cold;
    }
    if (bad >= 0) return fail(h, HY_ERR_NUMERIC, "%s: eigensolver did not converge for omega %g", who, ps.miss_om[bad]);
    return HY_OK;
}

int hy_fel_eval(hy_engine h, int n_cand, const int32_t *pattern, const double *alpha, const double *beta,
                const double *theta, double delta, double psi, const double *bl, double *out_ll, double *out_scale) {
    // FELAnalyzer::analyze_pattern's eval_pattern_lnl (fel.hpp:367-385):
    // omega = beta / max(alpha, 1e-12), MG94 with alpha = 1 and that beta,
    // P_b = transition_matrix(max(alpha, 1e-12) t_b), the pattern's lnL
    // (compute_pattern_log_likelihood, fel.hpp:276-308)
    const char *who = "hy_fel_eval";
    if (!h) return HY_ERR_ARG;
    if (!h->have_problem || !h->have_model) return fail(h, HY_ERR_STATE, "%s before hy_set_problem/hy_set_model", who);
    if (n_cand < 0 || (n_cand > 0 && (!pattern || !alpha || !beta || !out_ll)) || !theta || !bl)
        return fail(h, HY_ERR_ARG, "%s: bad candidate arrays", who);
    if (n_cand == 0) return HY_OK;
    if (n_cand > (1 << 24)) return fail(h, HY_ERR_ARG, "%s: %d candidates in one call", who, n_cand);
    const size_t K = (size_t)n_cand;
    std::vector<double> om(K);
    for (size_t k = 0; k < K; ++k) {
        const int p = pattern[k];
        if (p < 0 || p >= h->P) return fail(h, HY_ERR_ARG, "%s: pattern %d out of range", who, p);
        if (!(alpha[k] >= 0.0 && beta[k] >= 0.0 && std::isfinite(alpha[k]) && std::isfinite(beta[k])))
            return fail(h, HY_ERR_ARG, "%s: candidate %zu has alpha %g, beta %g", who, k, alpha[k], beta[k]);
        om[k] = beta[k] / std::max(alpha[k], 1e-12);
        if (!std::isfinite(om[k])) return fail(h, HY_ERR_ARG, "%s: candidate %zu has omega %g", who, k, om[k]);
    }
    PoolSolve ps;
    TRY(pool_solve(h, who, theta, delta, psi, bl, om, ps));
    FelPool &f = h->fel;
    hipStream_t s = h->stream;
    // staging: [alpha] doubles, then [spec][pattern] ints
    const size_t in_bytes = K * 8 + 2 * K * 4, out_bytes = 3 * K * 8;
    const size_t pin_need = std::max(in_bytes, out_bytes);
    if (h->f_pin_bytes < pin_need) {
        if (h->f_pin) (void)hipHostFree(h->f_pin);
        h->f_pin = nullptr;
        h->f_pin_bytes = 0;
        HIP_TRY(h, hipHostMalloc(&h->f_pin, pin_need * 2, hipHostMallocDefault));
        h->f_pin_bytes = pin_need * 2;
    }
    double *pd = (double *)h->f_pin;
    std::copy(alpha, alpha + K, pd);
    int32_t *pi32 = (int32_t *)(pd + K);
    std::copy(ps.slot.begin(), ps.slot.end(), pi32);
    std::copy(pattern, pattern + K, pi32 + K);
    TRY(ensure(h, h->d_fstage, in_bytes));
    TRY(ensure(h, h->d_fD, K * std::max(h->n_int, 1) * 64 * 8));
    TRY(ensure(h, h->d_fE, K * std::max(h->n_int, 1) * 4));
    TRY(ensure(h, h->d_fout, out_bytes));
    HIP_TRY(h, hipMemcpyAsync(h->d_fstage.p, h->f_pin, in_bytes, hipMemcpyHostToDevice, s));
    const double *dd = (const double *)h->d_fstage.p;
    const int32_t *di = (const int32_t *)(dd + K);
    FelArgs fa{};
    fa.sched = (const int32_t *)h->d_gsched.p;
    fa.n_sched = h->g_n_sched;
    fa.nref = (const int32_t *)h->d_gnref.p;
    fa.ident = (const uint8_t *)h->d_fident.p;
    fa.tips = (const uint16_t *)h->d_gtips.p;
    fa.tip_stride = h->g_P_pad;
    fa.amb = (const uint64_t *)h->d_amb.p;
    fa.pat = di + K;
    fa.alpha = dd;
    fa.spec = di;
    fa.lam = (const double *)h->d_flam.p;
    fa.U = (const double *)h->d_fU.p;
    fa.V = (const double *)h->d_fV.p;
    fa.mu = (const double *)h->d_fmu.p;
    fa.status = (const int *)h->d_fstatus.p;
    fa.bl = (const double *)h->d_fbl.p;
    fa.pi = (const double *)h->d_pi.p;
    fa.S = h->S;
    fa.n_int = std::max(h->n_int, 1);
    fa.n_cand = n_cand;
    fa.D = (double *)h->d_fD.p;
    fa.E = (int32_t *)h->d_fE.p;
    fa.out = (double *)h->d_fout.p;
    launch_fel(fa, n_cand, s);
    HIP_TRY(h, hipGetLastError());
    HIP_TRY(h, hipMemcpyAsync(h->f_pin, h->d_fout.p, out_bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(h, hipStreamSynchronize(s));
    TRY(pool_commit(h, who, ps));
    f.n_cand += (int64_t)K;
    const double *res = (const double *)h->f_pin;
    std::copy(res, res + K, out_ll);
    if (out_scale) std::copy(res + K, res + 2 * K, out_scale);
    return HY_OK;
}

int hy_meme_eval(hy_engine h, int n_cand, const int32_t *pattern, const double *alpha, const double *omega1,
                 const double *p1, const double *beta_plus, const double *beta_bg, const uint8_t *is_test,
                 const double *theta, double delta, double psi, const double *bl, const int32_t *cond_node,
                 const int32_t *cond_class, double *out_ll) {
    // MEMEAnalyzer::analyze_pattern's eval_pattern_meme_lnl (meme.hpp:219-277):
    // with a = max(alpha, 1e-12), MG94(alpha = a, beta = a omega1),
    // MG94(a, beta_plus), MG94(a, beta_bg) are a times the alpha = 1
    // generators at omega1, beta_plus / a, beta_bg / a, so their spectra come
    // from the pool and their transition matrices at a t_b; test branches
    // P = p1 P_1 + (1 - p1) P_+, background branches P_bg
    const char *who = "hy_meme_eval";
    if (!h) return HY_ERR_ARG;
    if (!h->have_problem || !h->have_model) return fail(h, HY_ERR_STATE, "%s before hy_set_problem/hy_set_model", who);
    if (n_cand < 0 || (n_cand > 0 && (!pattern || !alpha || !omega1 || !p1 || !beta_plus || !out_ll)) || !theta ||
        !bl)
        return fail(h, HY_ERR_ARG, "%s: bad candidate arrays", who);
    if (n_cand == 0) return HY_OK;
    if (n_cand > (1 << 22)) return fail(h, HY_ERR_ARG, "%s: %d candidates in one call", who, n_cand);
    bool has_bg = false;
    if (is_test)
        for (int n = 0; n < h->num_nodes; ++n)
            if (n != h->root && !is_test[n]) has_bg = true;
    if (has_bg && !beta_bg) return fail(h, HY_ERR_ARG, "%s: background branches need beta_bg", who);
    const size_t K = (size_t)n_cand;
    const int NS = has_bg ? 3 : 2;
    std::vector<double> om(K * NS);
    for (size_t k = 0; k < K; ++k) {
        const int p = pattern[k];
        if (p < 0 || p >= h->P) return fail(h, HY_ERR_ARG, "%s: pattern %d out of range", who, p);
        const double bb = has_bg ? beta_bg[k] : 0.0;
        if (!(alpha[k] >= 0.0 && omega1[k] >= 0.0 && beta_plus[k] >= 0.0 && bb >= 0.0 && p1[k] >= 0.0 &&
              p1[k] <= 1.0 && std::isfinite(alpha[k]) && std::isfinite(omega1[k]) && std::isfinite(beta_plus[k]) &&
              std::isfinite(bb)))
            return fail(h, HY_ERR_ARG, "%s: candidate %zu out of range", who, k);
        if (cond_node && cond_node[k] >= 0) {
            const int cn = cond_node[k], cc = cond_class ? cond_class[k] : -1;
            if (cn >= h->num_nodes || cn == h->root || (is_test && !is_test[cn]) || (cc != 0 && cc != 1))
                return fail(h, HY_ERR_ARG, "%s: candidate %zu conditions node %d on class %d", who, k, cn, cc);
        }
        const double a = std::max(alpha[k], 1e-12);
        om[k * NS] = omega1[k];
        om[k * NS + 1] = beta_plus[k] / a;
        if (has_bg) om[k * NS + 2] = bb / a;
        for (int j = 0; j < NS; ++j)
            if (!std::isfinite(om[k * NS + j])) return fail(h, HY_ERR_ARG, "%s: candidate %zu: omega overflow", who, k);
    }
    PoolSolve ps;
    TRY(pool_solve(h, who, theta, delta, psi, bl, om, ps));
    FelPool &f = h->fel;
    hipStream_t s = h->stream;
    // staging: [alpha][p1] doubles, then [spec A][spec B][spec C][pattern][condition] ints, then per-node
    // test flags
    const size_t n_ext = (size_t)h->n_ext;
    const size_t in_bytes = 2 * K * 8 + 5 * K * 4 + n_ext, out_bytes = K * 8;
    const size_t pin_need = std::max(in_bytes, out_bytes);
    if (h->f_pin_bytes < pin_need) {
        if (h->f_pin) (void)hipHostFree(h->f_pin);
        h->f_pin = nullptr;
        h->f_pin_bytes = 0;
        HIP_TRY(h, hipHostMalloc(&h->f_pin, pin_need * 2, hipHostMallocDefault));
        h->f_pin_bytes = pin_need * 2;
    }
    double *pd = (double *)h->f_pin;
    for (size_t k = 0; k < K; ++k) {
        pd[k] = std::max(alpha[k], 1e-12);
        pd[K + k] = p1[k];
    }
    int32_t *pi32 = (int32_t *)(pd + 2 * K);
    for (size_t k = 0; k < K; ++k) {
        pi32[k] = ps.slot[k * NS];
        pi32[K + k] = ps.slot[k * NS + 1];
        pi32[2 * K + k] = has_bg ? ps.slot[k * NS + 2] : ps.slot[k * NS];
        pi32[3 * K + k] = pattern[k];
        pi32[4 * K + k] = (cond_node && cond_node[k] >= 0) ? 2 * cond_node[k] + cond_class[k] : -1;
    }
    uint8_t *tf = (uint8_t *)(pi32 + 5 * K);
    for (size_t n = 0; n < n_ext; ++n)  // extended (identity) nodes are never read as test or background
        tf[n] = (n < (size_t)h->num_nodes && is_test) ? (is_test[n] ? 1 : 0) : 1;
    TRY(ensure(h, h->d_fstage, in_bytes));
    TRY(ensure(h, h->d_fD, K * std::max(h->n_int, 1) * 64 * 8));
    TRY(ensure(h, h->d_fE, K * std::max(h->n_int, 1) * 4));
    TRY(ensure(h, h->d_fout, out_bytes));
    HIP_TRY(h, hipMemcpyAsync(h->d_fstage.p, h->f_pin, in_bytes, hipMemcpyHostToDevice, s));
    const double *dd = (const double *)h->d_fstage.p;
    const int32_t *di = (const int32_t *)(dd + 2 * K);
    MemeArgs ma{};
    ma.sched = (const int32_t *)h->d_gsched.p;
    ma.n_sched = h->g_n_sched;
    ma.nref = (const int32_t *)h->d_gnref.p;
    ma.ident = (const uint8_t *)h->d_fident.p;
    ma.tips = (const uint16_t *)h->d_gtips.p;
    ma.tip_stride = h->g_P_pad;
    ma.amb = (const uint64_t *)h->d_amb.p;
    ma.pat = di + 3 * K;
    ma.alpha = dd;
    ma.p1 = dd + K;
    ma.specA = di;
    ma.specB = di + K;
    ma.specC = di + 2 * K;
    ma.cond = di + 4 * K;
    ma.is_test = (const uint8_t *)(di + 5 * K);
    ma.lam = (const double *)h->d_flam.p;
    ma.Z = (const double *)h->d_fZ.p;
    ma.bl = (const double *)h->d_fbl.p;
    ma.pi = (const double *)h->d_pi.p;
    ma.S = h->S;
    ma.n_int = std::max(h->n_int, 1);
    ma.D = (double *)h->d_fD.p;
    ma.E = (int32_t *)h->d_fE.p;
    ma.out = (double *)h->d_fout.p;
    launch_meme(ma, n_cand, s);
    HIP_TRY(h, hipGetLastError());
    HIP_TRY(h, hipMemcpyAsync(h->f_pin, h->d_fout.p, out_bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(h, hipStreamSynchronize(s));
    TRY(pool_commit(h, who, ps));
    f.n_cand += (int64_t)K;
    std::copy((const double *)h->f_pin, (const double *)h->f_pin + K, out_ll);
    return HY_OK;
}

namespace {
// invariant patterns: every resolved codon the same (fel.hpp:333-350,
// meme.hpp:186-197)
void invariant_patterns(hy_engine h, std::vector<uint8_t> &inv) {
    inv.assign(h->P, 1);
    for (int p = 0; p < h->P; ++p) {
        int first = -1;
        for (int t = 0; t < h->num_taxa && inv[p]; ++t) {
            const uint16_t c = h->codes[(size_t)t * h->P + p];
            if (c >= 64u) continue;
            if (first < 0)
                first = c;
            else if ((int)c != first)
                inv[p] = 0;
        }
    }
}

struct MemeRunCtx {
    hy_engine h;
    const double *theta;
    double delta, psi;
    const double *bl;
};
int meme_run_eval(void *user, int n, const int32_t *pattern, const double *alpha, const double *omega1,
                  const double *p1, const double *beta_plus, const int32_t *cond_node, const int32_t *cond_class,
                  double *out_ll) {
    const MemeRunCtx *c = (const MemeRunCtx *)user;
    return hy_meme_eval(c->h, n, pattern, alpha, omega1, p1, beta_plus, nullptr, nullptr, c->theta, c->delta, c->psi,
                        c->bl, cond_node, cond_class, out_ll);
}

struct FelRunCtx {
    hy_engine h;
    const double *theta;
    double delta, psi;
    const double *bl;
};
int fel_run_eval(void *user, int n, const int32_t *pattern, const double *alpha, const double *beta, double *out_ll) {
    const FelRunCtx *c = (const FelRunCtx *)user;
    return hy_fel_eval(c->h, n, pattern, alpha, beta, c->theta, c->delta, c->psi, c->bl, out_ll, nullptr);
}
}  // namespace

int hy_mg94_spectrum(hy_engine h, double alpha, double beta, const double *theta, double delta, double psi,
                     double *lam, double *U, double *V, double *mu) {
    // MG94Matrix::update's spectrum (rate_matrix.hpp:165-248) for the
    // handle's model: eigenvalues (clamped to <= 0), U_pi = diag(1/sqrt(pi)) Z,
    // V_pi = Z^T diag(sqrt(pi)), the mean rate
    const char *who = "hy_mg94_spectrum";
    if (!h) return HY_ERR_ARG;
    if (!h->have_model) return fail(h, HY_ERR_STATE, "%s before hy_set_model", who);
    if (!theta || !lam || !U || !V) return fail(h, HY_ERR_ARG, "%s: NULL argument", who);
    const double chk[4] = {alpha, beta, delta, psi};
    for (double x : chk)
        if (!(x >= 0.0 && std::isfinite(x))) return fail(h, HY_ERR_ARG, "%s: negative or non-finite rate %g", who, x);
    for (int i = 0; i < 5; ++i)
        if (!(theta[i] >= 0.0 && std::isfinite(theta[i])))
            return fail(h, HY_ERR_ARG, "%s: theta[%d] = %g", who, i, theta[i]);
    HIP_TRY(h, hipSetDevice(h->device));
    ParamLayout pl{};
    pl.n_mat = 1;
    pl.n_comp = 1;
    pl.off_alpha = 0;
    pl.off_omega = 1;
    pl.off_w = 2;
    pl.off_theta = 3;
    pl.off_delta = 8;
    pl.off_psi = 9;
    pl.total = 10;
    const double prm[10] = {alpha, beta, 1.0, theta[0], theta[1], theta[2], theta[3], theta[4], delta, psi};
    const bool warm = h->d_sZ.p && h->s_model_gen == h->model_gen;
    TRY(upload(h, h->d_sprm, prm, sizeof prm));
    TRY(ensure(h, h->d_slam, 64 * 8));
    TRY(ensure(h, h->d_sU, 4096 * 8));
    TRY(ensure(h, h->d_sV, 4096 * 8));
    TRY(ensure(h, h->d_sZ, 4096 * 8));
    TRY(ensure(h, h->d_smu, 8));
    TRY(ensure(h, h->d_sstatus, 4));
    ModelArgs ma{};
    ma.qdesc = (const int32_t *)h->d_qdesc.p;
    ma.nd = (const int8_t *)h->d_nd.p;
    ma.dpos = (const int8_t *)h->d_dpos.p;
    ma.dfrom = (const int8_t *)h->d_dfrom.p;
    ma.dto = (const int8_t *)h->d_dto.p;
    ma.syn = (const uint8_t *)h->d_syn.p;
    ma.pos = (const double *)h->d_pos.p;
    ma.pi = (const double *)h->d_pi.p;
    ma.S = h->S;
    ma.params = (const double *)h->d_sprm.p;
    ma.pl = pl;
    ma.lam = (double *)h->d_slam.p;
    ma.U = (double *)h->d_sU.p;
    ma.V = (double *)h->d_sV.p;
    ma.mu = (double *)h->d_smu.p;
    ma.Z = (double *)h->d_sZ.p;
    ma.warm = warm ? 1 : 0;  // refines the previous call's eigenvectors (same model)
    ma.status = (int *)h->d_sstatus.p;
    ma.tol = 1e-14;
    hipStream_t s = h->stream;
    launch_model(ma, 1, s);
    HIP_TRY(h, hipGetLastError());
    std::vector<double> hl(64), hU(4096), hV(4096);
    double hmu = 0.0;
    int st = 0;
    HIP_TRY(h, hipMemcpyAsync(hl.data(), h->d_slam.p, 64 * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(h, hipMemcpyAsync(hU.data(), h->d_sU.p, 4096 * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(h, hipMemcpyAsync(hV.data(), h->d_sV.p, 4096 * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(h, hipMemcpyAsync(&hmu, h->d_smu.p, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(h, hipMemcpyAsync(&st, h->d_sstatus.p, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(h, hipStreamSynchronize(s));
    if (st < 0) {
        h->s_model_gen = ~uint64_t(0);
        return fail(h, HY_ERR_NUMERIC, "%s: eigensolver did not converge", who);
    }
    h->s_model_gen = h->model_gen;
    const int S = h->S;
    // eigenpairs in increasing eigenvalue order (the order of Eigen's solver)
    std::vector<int> ord(S);
    for (int k = 0; k < S; ++k) ord[k] = k;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return hl[x] < hl[y]; });
    for (int c = 0; c < S; ++c) {
        const int k = ord[c];
        lam[c] = hl[k];
        for (int i = 0; i < S; ++i) {
            U[(size_t)i * S + c] = hU[(size_t)i * 64 + k];  // U_pi(i, c)
            V[(size_t)c * S + i] = hV[(size_t)k * 64 + i];  // V_pi(c, i)
        }
    }
    if (mu) *mu = hmu;
    return HY_OK;
}

int hy_fel_counters(hy_engine h, int64_t *out4, int reset) {
    if (!h || !out4) return HY_ERR_ARG;
    FelPool &f = h->fel;
    out4[0] = f.n_cand;
    out4[1] = f.n_hit;
    out4[2] = f.n_warm;
    out4[3] = f.n_cold;
    if (reset) f.n_cand = f.n_hit = f.n_warm = f.n_cold = 0;
    return HY_OK;
}

int hy_fel_run(hy_engine h, const double *theta, double alpha, double beta, double delta, double psi,
               const double *bl, double *out, int *out_calls, int64_t *out_candidates, double *out_q_scale) {
    const char *who = "hy_fel_run";
    if (!h) return HY_ERR_ARG;
    if (!h->have_problem || !h->have_model) return fail(h, HY_ERR_STATE, "%s before hy_set_problem/hy_set_model", who);
    if (!theta || !bl || !out) return fail(h, HY_ERR_ARG, "%s: NULL argument", who);
    if (!(alpha >= 0.0 && beta >= 0.0 && std::isfinite(alpha) && std::isfinite(beta)))
        return fail(h, HY_ERR_ARG, "%s: alpha %g, beta %g", who, alpha, beta);
    if (h->P == 0) {
        if (out_calls) *out_calls = 0;
        if (out_candidates) *out_candidates = 0;
        return HY_OK;
    }
    // prepare_branch_lengths (fel.hpp:240-262): Q_scale of the global model;
    // the generator is linear in (alpha, beta), so Q_scale = a mu(1, beta / a)
    std::vector<double> bl0(h->num_nodes, 0.0);
    double tested = 0.0;
    for (int n = 0; n < h->num_nodes; ++n) {
        if (n == h->root) continue;
        if (!(bl[n] >= 0.0 && std::isfinite(bl[n])))
            return fail(h, HY_ERR_ARG, "%s: branch length %d = %g", who, n, bl[n]);
        bl0[n] = bl[n];
        tested += bl[n];
    }
    const int32_t p0 = 0;
    const double a0 = std::max(alpha, 1e-12);
    double ll0 = 0.0, mu0 = 0.0;
    TRY(hy_fel_eval(h, 1, &p0, &a0, &beta, theta, delta, psi, bl0.data(), &ll0, &mu0));
    const double q_scale = a0 * mu0;
    const double conv = q_scale > 1e-12 ? 3.0 / q_scale : 1.0;
    std::vector<double> blc(h->num_nodes, 0.0);
    for (int n = 0; n < h->num_nodes; ++n) blc[n] = (n == h->root) ? 0.0 : bl[n] * conv;
    std::vector<uint8_t> inv;
    invariant_patterns(h, inv);
    FelRunCtx ctx{h, theta, delta, psi, blc.data()};
    const int rc = hy_fel_drive(h->P, inv.data(), tested, fel_run_eval, &ctx, out, out_calls, out_candidates);
    if (rc != HY_OK) return rc;
    if (out_q_scale) *out_q_scale = q_scale;
    return HY_OK;
}

int hy_meme_run(hy_engine h, const double *theta, double alpha, double beta, double delta, double psi,
                const double *bl, double *out, double *ebf, double *fel, int *out_calls, int64_t *out_candidates) {
    const char *who = "hy_meme_run";
    if (!h) return HY_ERR_ARG;
    if (!h->have_problem || !h->have_model) return fail(h, HY_ERR_STATE, "%s before hy_set_problem/hy_set_model", who);
    if (!theta || !bl || !out) return fail(h, HY_ERR_ARG, "%s: NULL argument", who);
    if (out_calls) *out_calls = 0;
    if (out_candidates) *out_candidates = 0;
    if (h->P == 0) return HY_OK;
    // the FEL baseline of every pattern, and the conversion of the lengths
    // (MEME's own prepare_branch_lengths is FEL's, meme.hpp:150-163)
    std::vector<double> fr((size_t)h->P * HY_FEL_FIELDS);
    double q_scale = 0.0;
    TRY(hy_fel_run(h, theta, alpha, beta, delta, psi, bl, fr.data(), nullptr, nullptr, &q_scale));
    const double conv = q_scale > 1e-12 ? 3.0 / q_scale : 1.0;
    std::vector<double> blc(h->num_nodes, 0.0);
    for (int n = 0; n < h->num_nodes; ++n) blc[n] = (n == h->root) ? 0.0 : bl[n] * conv;
    std::vector<uint8_t> inv;
    invariant_patterns(h, inv);
    MemeRunCtx ctx{h, theta, delta, psi, blc.data()};
    const int rc = hy_meme_drive(h->P, inv.data(), fr.data(), h->num_nodes, h->root, meme_run_eval, &ctx, out, ebf,
                                 out_calls, out_candidates);
    if (rc != HY_OK) return rc;
    // total branch length at the fitted rates (meme.hpp:614-638): the mean
    // rates of MG94(alpha, beta1) and MG94(alpha, beta_plus) are alpha times
    // those of the alpha = 1 generators
    std::vector<int32_t> pat;
    std::vector<double> ca, cb;
    for (int p = 0; p < h->P; ++p) {
        if (inv[p]) continue;
        const double *r = out + (size_t)p * HY_MEME_FIELDS;
        pat.insert(pat.end(), {p, p});
        ca.insert(ca.end(), {r[HY_MEME_ALPHA], r[HY_MEME_ALPHA]});
        cb.insert(cb.end(), {r[HY_MEME_BETA1], r[HY_MEME_BETA_PLUS]});
    }
    if (!pat.empty()) {
        std::vector<double> ll(pat.size()), mu(pat.size());
        TRY(hy_fel_eval(h, (int)pat.size(), pat.data(), ca.data(), cb.data(), theta, delta, psi, blc.data(), ll.data(),
                        mu.data()));
        size_t q = 0;
        for (int p = 0; p < h->P; ++p) {
            if (inv[p]) continue;
            double *r = out + (size_t)p * HY_MEME_FIELDS;
            const double a = std::max(r[HY_MEME_ALPHA], 1e-12);
            const double rate = (r[HY_MEME_P1] * (a * mu[q]) + r[HY_MEME_P_PLUS] * (a * mu[q + 1])) / 3.0;
            q += 2;
            double len = 0.0;
            for (int n = 0; n < h->num_nodes; ++n)
                if (n != h->root) len += blc[n] * rate;
            r[HY_MEME_LENGTH] = len;
        }
    }
    if (fel) std::copy(fr.begin(), fr.end(), fel);
    return HY_OK;
}

int hy_read_partials(hy_engine h, int p_start, int p_end, double *D, int32_t *E, double *V, int32_t *G, double *L,
                     int32_t *LE) {
    if (!h) return HY_ERR_ARG;
    if (h->part_ri < 0) return fail(h, HY_ERR_STATE, "hy_read_partials: no evaluation has left partials");
    if (h->ck)
        return fail(h, HY_ERR_UNSUPPORTED,
                    "hy_read_partials: the last evaluation kept only checkpointed partials (memory budget)");
    if (p_start < 0 || p_end > h->P || p_start >= p_end)
        return fail(h, HY_ERR_ARG, "hy_read_partials: pattern range [%d, %d) outside [0, %d)", p_start, p_end, h->P);
    if ((V || G) && !h->part_outside)
        return fail(h, HY_ERR_STATE, "hy_read_partials: the last evaluation ran no outside pass");
    HIP_TRY(h, hipSetDevice(h->device));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    const int ri = h->part_ri, T = h->T, S = h->S, B = p_end - p_start, N = h->num_nodes;
    const size_t esz = ri ? 8 : 4, tile_elems = (size_t)64 * T;
    const int t0 = p_start / T, nt = (p_end - 1) / T - t0 + 1;
    auto value = [&](const std::vector<char> &buf, size_t idx) {
        return ri ? reinterpret_cast<const double *>(buf.data())[idx] : (double)reinterpret_cast<const float *>(buf.data())[idx];
    };
    // rows x (tile range) of a [rows][n_tiles][64 T] buffer
    auto fetch_tiles = [&](const void *base, int rows, std::vector<char> &out) -> int {
        out.resize((size_t)rows * nt * tile_elems * esz);
        HIP_TRY(h, hipMemcpy2DAsync(out.data(), nt * tile_elems * esz, (const char *)base + t0 * tile_elems * esz,
                                    (size_t)h->n_tiles * tile_elems * esz, nt * tile_elems * esz, rows,
                                    hipMemcpyDeviceToHost, h->stream));
        return HY_OK;
    };
    // rows x [p_start, p_end) of a [rows][P_pad] int32 buffer
    auto fetch_cols = [&](const void *base, int rows, std::vector<int32_t> &out) -> int {
        out.resize((size_t)rows * B);
        HIP_TRY(h, hipMemcpy2DAsync(out.data(), (size_t)B * 4, (const char *)base + (size_t)p_start * 4,
                                    (size_t)h->P_pad * 4, (size_t)B * 4, rows, hipMemcpyDeviceToHost, h->stream));
        return HY_OK;
    };
    std::vector<char> dbuf, vbuf;
    std::vector<int32_t> ebuf, gbuf;
    if (D || E) {
        TRY(fetch_tiles(h->d_D[ri].p, h->n_int, dbuf));
        TRY(fetch_cols(h->d_E.p, h->n_int, ebuf));
    }
    if (V || G) {
        TRY(fetch_tiles(h->d_Vp[ri].p, h->n_vslot, vbuf));
        TRY(fetch_cols(h->d_G.p, h->n_vslot, gbuf));
    }
    std::vector<double> lr(B);
    std::vector<int32_t> ler(B);
    if (L) HIP_TRY(h, hipMemcpyAsync(lr.data(), (const double *)h->d_Lroot.p + p_start, B * 8, hipMemcpyDeviceToHost, h->stream));
    if (LE) HIP_TRY(h, hipMemcpyAsync(ler.data(), (const int32_t *)h->d_Eroot.p + p_start, B * 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    for (int u = 0; u < N; ++u) {
        const int gi = h->int_slot[u], vs = (u == h->ext_root) ? -1 : h->vslot[u];
        for (int b = 0; b < B; ++b) {
            const int p = p_start + b, t = p / T - t0, tl = p % T;
            if (gi >= 0) {
                const size_t base = ((size_t)gi * nt + t) * tile_elems;
                if (D)
                    for (int st = 0; st < S; ++st) D[((size_t)u * B + b) * S + st] = value(dbuf, base + bmaj_index(st, tl));
                if (E) E[(size_t)u * B + b] = ebuf[(size_t)gi * B + b];
            }
            if (vs >= 0) {
                const size_t base = ((size_t)vs * nt + t) * tile_elems;
                if (V)
                    for (int st = 0; st < S; ++st) V[((size_t)u * B + b) * S + st] = value(vbuf, base + (size_t)st * T + tl);
                if (G) G[(size_t)u * B + b] = gbuf[(size_t)vs * B + b];
            }
        }
    }
    if (L) std::memcpy(L, lr.data(), B * 8);
    if (LE) std::memcpy(LE, ler.data(), B * 4);
    return HY_OK;
}

int hy_set_pattern_weights(hy_engine h, const double *weights) {
    if (!h) return HY_ERR_ARG;
    if (!h->have_problem) return fail(h, HY_ERR_STATE, "hy_set_pattern_weights before hy_set_problem");
    if (!weights) {
        h->wuser = false;
        return HY_OK;
    }
    for (int p = 0; p < h->P; ++p)
        if (!(weights[p] >= 0.0) || !std::isfinite(weights[p]))
            return fail(h, HY_ERR_ARG, "hy_set_pattern_weights: weight %d must be finite and >= 0", p);
    HIP_TRY(h, hipSetDevice(h->device));
    // sized for the widest tiling (a later re-tile pads with the same zeros)
    const size_t cap = ((size_t)h->P + 63) / 64 * 64;
    std::vector<double> w(cap, 0.0);
    std::copy(weights, weights + h->P, w.begin());
    TRY(upload(h, h->d_wuser, w.data(), cap * 8));
    h->wuser = true;
    return HY_OK;
}

int hy_branch_eval(hy_engine h, int precision, int node, int n_cand, const double *P_cand, double *out_lnl,
                   double *out_site_L, int32_t *out_site_e) {
    if (!h) return HY_ERR_ARG;
    if (!P_cand || !out_lnl || n_cand < 1) return fail(h, HY_ERR_ARG, "hy_branch_eval: no candidates or NULL output");
    if (precision != HY_PREC_FP64 && precision != HY_PREC_FP32)
        return fail(h, HY_ERR_ARG, "hy_branch_eval: unknown precision %d", precision);
    const int ri = precision == HY_PREC_FP64 ? 1 : 0;
    if (h->part_ri != ri || !h->part_outside)
        return fail(h, HY_ERR_STATE,
                    "hy_branch_eval: the last evaluation on this handle was not a gradient evaluation in this precision");
    if (h->ck)
        return fail(h, HY_ERR_UNSUPPORTED, "hy_branch_eval: the last evaluation kept only checkpointed partials");
    if (h->n_genes > 1) return fail(h, HY_ERR_UNSUPPORTED, "hy_branch_eval: not on a fused gene batch");
    if (node < 0 || node >= h->num_nodes || node == h->root)
        return fail(h, HY_ERR_ARG, "hy_branch_eval: node %d has no branch above it", node);
    int bi = -1;
    for (int b = 0; b < h->n_branch; ++b)
        if (h->blist[(size_t)b * BR_STRIDE] == node) bi = b;
    if (bi < 0) return fail(h, HY_ERR_ARG, "hy_branch_eval: node %d has no branch above it", node);
    const int32_t *e = &h->blist[(size_t)bi * BR_STRIDE];
    const int ref = e[2], kind = ref >> 28, idx = ref & REF_MASK;
    const int S = h->S, T = h->T;
    HIP_TRY(h, hipSetDevice(h->device));
    // candidates padded to 64 x 64 (zero rows and columns beyond S)
    std::vector<double> Pp((size_t)n_cand * 4096, 0.0);
    for (int c = 0; c < n_cand; ++c)
        for (int i = 0; i < S; ++i)
            std::memcpy(&Pp[((size_t)c * 64 + i) * 64], P_cand + ((size_t)c * S + i) * S, sizeof(double) * S);
    TRY(ensure(h, h->d_bP, Pp.size() * 8));
    TRY(ensure(h, h->d_bL, (size_t)n_cand * h->P_pad * 8));
    TRY(ensure(h, h->d_be, (size_t)h->P_pad * 4));
    TRY(ensure(h, h->d_blnl, (size_t)n_cand * 8));
    hipStream_t s = h->stream;
    HIP_TRY(h, hipMemcpyAsync(h->d_bP.p, Pp.data(), Pp.size() * 8, hipMemcpyHostToDevice, s));
    auto fill = [&](auto &a) {
        a.P = (const double *)h->d_bP.p;
        a.E = (const int32_t *)h->d_E.p;
        a.G = (const int32_t *)h->d_G.p;
        a.tips = (const uint16_t *)h->d_tips.p;
        a.amb = (const uint64_t *)h->d_amb.p;
        a.v_slot = e[1];
        a.d_slot = kind == KIND_TIP ? 0 : idx;
        a.tip_row = kind == KIND_TIP ? idx : -1;
        a.n_tiles = h->n_tiles;
        a.P_pad = h->P_pad;
        a.n_patterns = h->P;
        a.S = S;
        a.out_L = (double *)h->d_bL.p;
        a.out_e = (int32_t *)h->d_be.p;
    };
    if (ri) {
        BranchCtxArgs<double> a{};
        fill(a);
        a.V = (const double *)h->d_Vp[1].p;
        a.D = (const double *)h->d_D[1].p;
        launch_branch_ctx<double>(a, T, n_cand, s);
    } else {
        BranchCtxArgs<float> a{};
        fill(a);
        a.V = (const float *)h->d_Vp[0].p;
        a.D = (const float *)h->d_D[0].p;
        launch_branch_ctx<float>(a, T, n_cand, s);
    }
    HIP_TRY(h, hipGetLastError());
    BranchSumArgs sa{};
    sa.L = (const double *)h->d_bL.p;
    sa.e = (const int32_t *)h->d_be.p;
    sa.weights = (const double *)(h->wuser ? h->d_wuser.p : h->d_weights.p);
    sa.P_pad = h->P_pad;
    sa.n_patterns = h->P;
    sa.out = (double *)h->d_blnl.p;
    launch_branch_sum(sa, n_cand, s);
    HIP_TRY(h, hipGetLastError());
    HIP_TRY(h, hipMemcpyAsync(out_lnl, h->d_blnl.p, (size_t)n_cand * 8, hipMemcpyDeviceToHost, s));
    if (out_site_L)
        HIP_TRY(h, hipMemcpy2DAsync(out_site_L, (size_t)h->P * 8, h->d_bL.p, (size_t)h->P_pad * 8, (size_t)h->P * 8,
                                    n_cand, hipMemcpyDeviceToHost, s));
    if (out_site_e) HIP_TRY(h, hipMemcpyAsync(out_site_e, h->d_be.p, (size_t)h->P * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(h, hipStreamSynchronize(s));
    return HY_OK;
}

int hy_last_tile(hy_engine h, int *patterns_per_tile) {
    if (!h || !patterns_per_tile) return fail(h, HY_ERR_ARG, "hy_last_tile: NULL argument");
    *patterns_per_tile = h->T;
    return HY_OK;
}

int hy_last_parts(hy_engine h, int *n_parts) {
    if (!h || !n_parts) return fail(h, HY_ERR_ARG, "hy_last_parts: NULL argument");
    *n_parts = h->n_parts;
    return HY_OK;
}

int hy_set_warm_start(hy_engine h, int enable) {
    if (h) ++h->epoch;  // captured graphs are stale
    if (!h) return HY_ERR_ARG;
    h->warm_enable = enable != 0;
    if (!h->warm_enable) h->warm_valid = false;
    return HY_OK;
}

int hy_last_eigen_status(hy_engine h, int *status, int n) {
    if (!h || !status) return HY_ERR_ARG;
    if (!h->h_status) return fail(h, HY_ERR_STATE, "no evaluation yet");
    for (int m = 0; m < n && m < h->n_mat; ++m) status[m] = h->h_status[m];
    return HY_OK;
}

int hy_last_timings(hy_engine h, double *ms8) {
    if (!h || !ms8) return HY_ERR_ARG;
    if (!h->profile) return fail(h, HY_ERR_STATE, "timings need HY_PROFILE=1 at hy_create");
    std::memcpy(ms8, h->timings, sizeof(h->timings));
    return HY_OK;
}

int hy_set_memory_budget(hy_engine h, size_t bytes) {
    if (!h) return HY_ERR_ARG;
    h->mem_budget = bytes;
    h->sched_key = 0;  // the layout is decided again at the next evaluation
    ++h->epoch;
    return HY_OK;
}

int hy_last_checkpoint(hy_engine h, int *n_parts, size_t *partial_bytes) {
    if (!h) return HY_ERR_ARG;
    if (n_parts) *n_parts = h->ck ? std::max(h->n_parts, 0) : 0;
    if (partial_bytes) *partial_bytes = h->partial_bytes;
    return HY_OK;
}

int hy_graph_count(int *live, int *cap) {
    if (live) *live = g_live_graphs.load();
    if (cap) *cap = graph_cap();
    return HY_OK;
}

int hy_device_bytes(hy_engine h, size_t *out) {
    if (!h || !out) return HY_ERR_ARG;
    *out = h->device_bytes;
    return HY_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Maximum-likelihood fit with captured optimizer iterations (hy_fit_lbfgs):
// the joint phase of MG94Fitter::fit (optimizer.hpp:934-1110) with the
// quasi-Newton step of hy_lbfgs.h.  The start point is evaluated through
// hy_eval's path and stepped on the host; from there every trial point is
// one replay of a graph holding the evaluation and k_lbfgs, which writes the
// next trial's parameters into the pinned parameter block the next replay's
// k_model reads.
// ---------------------------------------------------------------------------
namespace {

constexpr double LB_LOG_BL_LO = -18.420680744, LB_LOG_BL_HI = 6.91;  // optimizer.hpp:966-967
constexpr double LB_LOG_RATE_LO = -10.0, LB_LOG_RATE_HI = 6.91;      // optimizer.hpp:971-972, 986-987

struct FitScratch {
    hy_engine h = nullptr;
    // optimizer state; the captured graph's parameter block, results and
    // eigensolver status words (device memory: an iteration touches no host
    // memory)
    void *state = nullptr, *hd = nullptr, *maps = nullptr, *params = nullptr, *res = nullptr, *status = nullptr;
    hipGraphExec_t exec = nullptr;
    ~FitScratch() {
        if (exec) {
            (void)hipStreamSynchronize(h->stream);
            destroy_graph_exec(exec);
        }
        if (state) (void)hipFree(state);
        if (hd) (void)hipFree(hd);
        if (maps) (void)hipFree(maps);
        if (params) (void)hipFree(params);
        if (res) (void)hipFree(res);
        if (status) (void)hipFree(status);
    }
};

// Captures one fit iteration: the evaluation (as hy_eval's graph, without
// the completion word, with its parameter block, results and status words in
// the device buffers la names instead of pinned host memory) followed by
// k_lbfgs.
template <typename R>
int capture_fit_iteration(hy_engine h, const LbArgs &la, hipGraphExec_t *out) {
    hipGraph_t graph = nullptr;
    HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
    const double *keep_params = h->h_params_dev;
    double *keep_res = h->h_res_dev;
    int *keep_status = h->h_status_dev;
    h->h_params_dev = la.params;
    h->h_res_dev = const_cast<double *>(la.res);
    h->h_status_dev = const_cast<int *>(la.eig_status);
    h->zc_params = true;
    h->publish_done = false;
    const int rc = run_eval<R>(h, HY_WANT_LOGL | HY_WANT_GRAD, nullptr);
    h->zc_params = false;
    h->h_params_dev = keep_params;
    h->h_res_dev = keep_res;
    h->h_status_dev = keep_status;
    if (rc == HY_OK) launch_lbfgs(la, h->stream);
    const hipError_t e2 = hipStreamEndCapture(h->stream, &graph);
    if (rc != HY_OK || e2 != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc != HY_OK ? rc : fail(h, HY_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e2));
    }
    const hipError_t e3 = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e3 != hipSuccess) return fail(h, HY_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e3));
    return HY_OK;
}

}  // namespace

extern "C" {

double hy_lbfgs_exp(double x) { return lb_exp(x); }

size_t hy_lbfgs_state_doubles(int n, int m) { return (n < 1 || m < 1) ? 0 : lb_state_doubles(n, m); }

int hy_lbfgs_init(int n, int m, double *state, hy_lbfgs_head *hd, const double *x0, const double *lb,
                  const double *ub, int max_iter, double tol_grad, double tol_f) {
    if (n < 1 || m < 1 || m > LB_MMAX || !state || !hd || !x0 || !lb || !ub || max_iter < 0) return HY_ERR_ARG;
    for (int i = 0; i < n; ++i)
        if (!(lb[i] <= ub[i]) || x0[i] != x0[i]) return HY_ERR_ARG;
    std::memset(hd, 0, sizeof *hd);
    hd->n = n;
    hd->m = m;
    hd->max_iter = max_iter;
    hd->max_back = 20;
    hd->tol_grad = tol_grad;
    hd->tol_f = tol_f;
    hd->c1 = 1e-4;
    std::fill(state, state + lb_state_doubles(n, m), 0.0);
    const LbVec v = lb_vec(state, n, m);
    for (int i = 0; i < n; ++i) {
        v.lb[i] = lb[i];
        v.ub[i] = ub[i];
        v.x[i] = std::min(std::max(x0[i], lb[i]), ub[i]);
        v.xt[i] = v.x[i];
    }
    return HY_OK;
}

int hy_lbfgs_step(double *state, hy_lbfgs_head *hd, double f_trial, const double *g_trial, int ok, double *x_next) {
    if (!state || !hd || hd->n < 1 || hd->m < 1 || hd->m > LB_MMAX) return HY_ERR_ARG;
    const LbVec v = lb_vec(state, hd->n, hd->m);
    if (g_trial) std::memcpy(v.gt, g_trial, sizeof(double) * hd->n);
    LbHostPar p;
    lb_step(p, v, *hd, f_trial, ok);
    if (x_next) std::memcpy(x_next, v.xt, sizeof(double) * hd->n);
    return HY_OK;
}

int hy_fit_lbfgs(hy_engine h, int precision, int captured, double *bl, double alpha, double *omegas,
                 const double *comp_w, double *theta, double delta, double psi, double scale, int m, int max_iter,
                 double tol_grad, double tol_f, int check_every, double *out_logL, int32_t *out_info) {
    if (!h) return HY_ERR_ARG;
    TRY(check_eval_args(h, precision, bl, omegas, comp_w, theta, "hy_fit_lbfgs"));
    if (m < 1 || m > LB_MMAX) return fail(h, HY_ERR_ARG, "hy_fit_lbfgs: m = %d correction pairs (1..%d)", m, LB_MMAX);
    if (max_iter < 0 || check_every < 1 || !(scale > 0.0))
        return fail(h, HY_ERR_ARG, "hy_fit_lbfgs: max_iter >= 0, check_every >= 1 and scale > 0 required");
    if (h->P == 0) return fail(h, HY_ERR_STATE, "hy_fit_lbfgs: empty alignment");
    if (h->wuser) return fail(h, HY_ERR_UNSUPPORTED, "hy_fit_lbfgs: caller pattern weights are set");
    HIP_TRY(h, hipSetDevice(h->device));
    TRY(prepare_buffers(h, precision));
    // free variables: log bl of every node with a parent, log omega, log theta
    std::vector<int32_t> res_map, par_map;
    std::vector<double> x0, lo, hi;
    const ParamLayout &pl = h->pl;
    auto add = [&](double v, int res, int par, double a, double b) -> bool {
        if (!(v >= 0.0) || !std::isfinite(v)) return false;
        x0.push_back(v > 0.0 ? std::log(v) : a);
        lo.push_back(a);
        hi.push_back(b);
        res_map.push_back(res);
        par_map.push_back(par);
        return true;
    };
    for (int n = 0; n < h->num_nodes; ++n)
        if (h->real_parent[n] >= 0 && !add(bl[n], 1 + n, pl.off_bl + n, LB_LOG_BL_LO, LB_LOG_BL_HI))
            return fail(h, HY_ERR_ARG, "hy_fit_lbfgs: branch length %d is negative or not finite", n);
    for (int k = 0; k < h->n_mat; ++k)
        if (!add(omegas[k], 1 + h->num_nodes + k, pl.off_omega + k, LB_LOG_RATE_LO, LB_LOG_RATE_HI))
            return fail(h, HY_ERR_ARG, "hy_fit_lbfgs: omega %d is negative or not finite", k);
    for (int k = 0; k < 5; ++k)
        if (!add(theta[k], 1 + h->num_nodes + h->n_mat + h->n_comp + k, pl.off_theta + k, LB_LOG_RATE_LO,
                 LB_LOG_RATE_HI))
            return fail(h, HY_ERR_ARG, "hy_fit_lbfgs: exchangeability %d is negative or not finite", k);
    const int n = (int)x0.size();
    std::vector<double> state(lb_state_doubles(n, m));
    hy_lbfgs_head hd;
    TRY(hy_lbfgs_init(n, m, state.data(), &hd, x0.data(), lo.data(), hi.data(), max_iter, tol_grad, tol_f));
    const LbVec v = lb_vec(state.data(), n, m);
    std::vector<double> cbl(bl, bl + h->num_nodes), com(omegas, omegas + h->n_mat), cth(theta, theta + 5);
    // host evaluation of the trial point xt and one host step
    auto host_iteration = [&]() -> int {
        for (int i = 0; i < n; ++i) {
            const double pv = lb_exp(v.xt[i]);
            const int par = par_map[i];
            if (par >= pl.off_theta) cth[par - pl.off_theta] = pv;
            else if (par >= pl.off_omega) com[par - pl.off_omega] = pv;
            else cbl[par - pl.off_bl] = pv;
        }
        const int rc = hy_eval(h, precision, HY_WANT_LOGL | HY_WANT_GRAD, cbl.data(), alpha, com.data(), comp_w,
                               cth.data(), delta, psi, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
        if (rc != HY_OK && rc != HY_ERR_NUMERIC) return rc;
        LbHostPar p;
        const double ft = lb_gather(p, v, n, scale, h->h_res, h->h_params, res_map.data(), par_map.data());
        lb_step(p, v, hd, ft, rc == HY_OK ? 1 : 0);
        return HY_OK;
    };
    // the start point solves its spectra cold, so a fit does not depend on
    // what the engine evaluated before it (warm starts refine the previous
    // spectra, which can move the last bits of an evaluation)
    h->warm_valid = false;
    TRY(host_iteration());
    int replays = 0;
    // engines that launch eagerly (members of a group, per-stage profiling)
    // step on the host: same iterates, no captured graph
    if (!h->use_graphs || h->profile) captured = 0;
    if (captured && hd.status == HY_LBFGS_RUNNING) {
        FitScratch fs;
        fs.h = h;
        const size_t sb = state.size() * sizeof(double), mb = sizeof(int32_t) * 2 * n;
        HIP_TRY(h, hipMalloc(&fs.state, sb));
        HIP_TRY(h, hipMalloc(&fs.hd, sizeof(hy_lbfgs_head)));
        HIP_TRY(h, hipMalloc(&fs.maps, mb));
        const size_t nres = 1 + (size_t)h->num_nodes + h->n_mat + h->n_comp + 5;
        HIP_TRY(h, hipMalloc(&fs.params, sizeof(double) * pl.total));
        HIP_TRY(h, hipMalloc(&fs.res, sizeof(double) * nres));
        HIP_TRY(h, hipMalloc(&fs.status, sizeof(int) * 64));
        std::vector<int32_t> maps(res_map);
        maps.insert(maps.end(), par_map.begin(), par_map.end());
        HIP_TRY(h, hipMemcpy(fs.state, state.data(), sb, hipMemcpyHostToDevice));
        HIP_TRY(h, hipMemcpy(fs.hd, &hd, sizeof hd, hipMemcpyHostToDevice));
        HIP_TRY(h, hipMemcpy(fs.maps, maps.data(), mb, hipMemcpyHostToDevice));
        // the first captured evaluation reads the trial the host step chose
        std::vector<double> prm(h->h_params, h->h_params + pl.total);
        for (int i = 0; i < n; ++i) prm[par_map[i]] = lb_exp(v.xt[i]);
        HIP_TRY(h, hipMemcpy(fs.params, prm.data(), sizeof(double) * pl.total, hipMemcpyHostToDevice));
        LbArgs la{};
        la.state = (double *)fs.state;
        la.hd = (hy_lbfgs_head *)fs.hd;
        la.res = (const double *)fs.res;
        la.params = (double *)fs.params;
        la.eig_status = (const int *)fs.status;
        la.n_mat = h->n_mat;
        la.scale = scale;
        la.res_map = (const int32_t *)fs.maps;
        la.par_map = (const int32_t *)fs.maps + n;
        if (!acquire_graph_slot())
            return fail(h, HY_ERR_NOMEM, "hy_fit_lbfgs: the process-wide cap on captured graphs is reached");
        const int crc = precision == HY_PREC_FP64 ? capture_fit_iteration<double>(h, la, &fs.exec)
                                                  : capture_fit_iteration<float>(h, la, &fs.exec);
        if (crc != HY_OK) {
            g_live_graphs.fetch_sub(1);
            fs.exec = nullptr;
            return crc;
        }
        // every trial costs one replay; a line search takes at most max_back + 1
        const long cap = ((long)max_iter + 1) * (hd.max_back + 2) + 2;
        hy_lbfgs_head hh = hd;
        while (hh.status == HY_LBFGS_RUNNING && replays < cap) {
            for (int j = 0; j < check_every; ++j) HIP_TRY(h, hipGraphLaunch(fs.exec, h->stream));
            replays += check_every;
            HIP_TRY(h, hipMemcpyAsync(&hh, fs.hd, sizeof hh, hipMemcpyDeviceToHost, h->stream));
            HIP_TRY(h, hipStreamSynchronize(h->stream));
        }
        HIP_TRY(h, hipMemcpy(state.data(), fs.state, sb, hipMemcpyDeviceToHost));
        // replays past the end re-evaluate x and leave the state unchanged
        hd = hh;
        h->part_ri = precision == HY_PREC_FP64 ? 1 : 0;
        h->part_outside = true;
        if (hd.status == HY_LBFGS_RUNNING)
            return fail(h, HY_ERR_STATE, "hy_fit_lbfgs: %d replays without the optimizer finishing", replays);
    } else {
        while (hd.status == HY_LBFGS_RUNNING) TRY(host_iteration());
    }
    for (int i = 0; i < n; ++i) {
        const double pv = lb_exp(v.x[i]);
        const int par = par_map[i];
        if (par >= pl.off_theta) theta[par - pl.off_theta] = pv;
        else if (par >= pl.off_omega) omegas[par - pl.off_omega] = pv;
        else bl[par - pl.off_bl] = pv;
    }
    if (out_logL) *out_logL = -hd.f / scale;
    if (out_info) {
        out_info[0] = hd.status;
        out_info[1] = hd.iter;
        out_info[2] = hd.n_eval;
        out_info[3] = replays;
    }
    return HY_OK;
}

}  // extern "C"
