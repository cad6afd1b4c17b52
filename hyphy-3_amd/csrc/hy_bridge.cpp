// Wheatstone-bridged trees on the MI355X engine (C ABI hy_network_*).
//
// A bridge at hub u with children {A, B, C} replaces u's binary resolution
// by a weighted superposition of its three quartet resolutions
// (likelihood.hpp:376-425, 520-569, 623-698; wheatstone_likelihood.hpp):
//   D_u = sum_r w_r (P_br(r) (M_x * M_y)) * M_z,  M_c = P_c D_c,
//   (x, y | z) = (A, B | C), (A, C | B), (B, C | A).
// Everything above u depends linearly on D_u, so for each pattern
//   L_p = sum_c W_c L_p(T_c)
// over the configurations c that choose one resolution per bridge, with
// W_c the product of the chosen weights and T_c the binary tree in which a
// new node v_k (id num_nodes + k) joins the chosen pair x, y and hangs from
// u by the bridge branch of that resolution.  The network is evaluated as
// that per-pattern mixture over engines that each hold one resolved tree:
// per-pattern likelihoods of every configuration, a log-sum-exp, and
// gradients as the configurations' gradients with pattern weights
// w_p W_c L_p(T_c) / L_p (hy_set_pattern_weights), summed:
//   d lnL / d theta        = sum_c (gradient of T_c, posterior-weighted)
//   d lnL / d t_{k,r}      = sum over c with r_k = r of d lnL(T_c) / d t(v_k)
//   d lnL / d w_{k,r}      = sum over c with r_k = r of sum_p w_p post_c(p) / w_{k,r}
// The configurations are built once per problem (3^n_bridges engines, at most
// HY_MAX_BRIDGES bridges); configurations whose weight product is zero are
// not evaluated.

#include "hyphy_mi355x.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

struct hy_network_s {
    int device = 0;
    std::string err;
    int num_nodes = 0, num_patterns = 0, n_bridges = 0, root = -1;
    std::vector<int32_t> hub, kids;      // [n_bridges], [n_bridges][3]
    std::vector<double> weights;         // alignment pattern weights
    std::vector<hy_engine> cfg;          // one engine per configuration
    std::vector<std::vector<int>> res;   // resolution of every bridge per configuration
    std::vector<int32_t> branch_class;   // last mixture (extended per configuration)
    int n_mat = 1, n_comp = 1;
    bool have_model = false;
};

namespace {

int nfail(hy_network n, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (n) n->err = buf;
    return code;
}

// Forwards an engine error of configuration c.
int child_fail(hy_network n, size_t c, int rc) {
    return nfail(n, rc, "configuration %zu: %s", c, hy_last_error(n->cfg[c]));
}

void drop_configs(hy_network n) {
    for (hy_engine e : n->cfg) hy_destroy(e);
    n->cfg.clear();
    n->res.clear();
}

}  // namespace

extern "C" {

int hy_network_create(int device, hy_network *out) {
    if (!out) return HY_ERR_ARG;
    *out = nullptr;
    hy_engine probe = nullptr;
    const int rc = hy_create(device, &probe);  // checks the device like hy_create
    if (rc != HY_OK) return rc;
    hy_destroy(probe);
    hy_network n = new hy_network_s;
    n->device = device;
    *out = n;
    return HY_OK;
}

void hy_network_destroy(hy_network n) {
    if (!n) return;
    drop_configs(n);
    delete n;
}

const char *hy_network_last_error(hy_network n) { return n ? n->err.c_str() : "NULL network"; }

int hy_network_set_problem(hy_network n, int num_nodes, const int32_t *parent, const int32_t *child_ptr,
                           const int32_t *child_idx, const int32_t *leaf_taxon, int num_taxa, int num_patterns,
                           const int8_t *states, const uint64_t *masks, const double *weights, int S, int n_bridges,
                           const int32_t *hub, const int32_t *hub_children) {
    if (!n) return HY_ERR_ARG;
    if (num_nodes < 1 || !parent || !child_ptr || !child_idx || !leaf_taxon || num_patterns < 0 ||
        (num_patterns > 0 && (!weights || !states || !masks)))
        return nfail(n, HY_ERR_ARG, "hy_network_set_problem: bad problem arrays");
    if (n_bridges < 0 || n_bridges > HY_MAX_BRIDGES || (n_bridges > 0 && (!hub || !hub_children)))
        return nfail(n, HY_ERR_ARG, "hy_network_set_problem: 0..%d bridges", HY_MAX_BRIDGES);
    int root = -1;
    for (int v = 0; v < num_nodes; ++v)
        if (parent[v] < 0) root = v;
    std::vector<char> is_hub(num_nodes, 0);
    for (int k = 0; k < n_bridges; ++k) {
        const int u = hub[k];
        if (u < 0 || u >= num_nodes || u == root || is_hub[u])
            return nfail(n, HY_ERR_ARG, "hy_network_set_problem: bridge %d: hub %d must be a distinct non-root node", k,
                         u);
        is_hub[u] = 1;
        std::vector<int32_t> ch(child_idx + child_ptr[u], child_idx + child_ptr[u + 1]), want(hub_children + 3 * k,
                                                                                            hub_children + 3 * k + 3);
        std::sort(ch.begin(), ch.end());
        std::sort(want.begin(), want.end());
        if (ch != want)
            return nfail(n, HY_ERR_ARG, "hy_network_set_problem: bridge %d: the hub's children must be the three "
                                        "bridge children", k);
    }
    drop_configs(n);
    n->num_nodes = num_nodes;
    n->num_patterns = num_patterns;
    n->n_bridges = n_bridges;
    n->root = root;
    n->hub.assign(hub, hub + n_bridges);
    n->kids.assign(hub_children, hub_children + 3 * n_bridges);
    n->weights.assign(weights, weights + num_patterns);
    n->have_model = false;
    n->n_mat = n->n_comp = 1;
    n->branch_class.assign(num_nodes, 0);
    // every configuration's resolved tree
    int n_cfg = 1;
    for (int k = 0; k < n_bridges; ++k) n_cfg *= 3;
    const int N = num_nodes + n_bridges;
    for (int c = 0; c < n_cfg; ++c) {
        std::vector<int> r(n_bridges);
        for (int k = 0, x = c; k < n_bridges; ++k, x /= 3) r[k] = x % 3;
        std::vector<int32_t> par(parent, parent + num_nodes);
        std::vector<std::vector<int32_t>> ch(num_nodes);
        for (int v = 0; v < num_nodes; ++v) ch[v].assign(child_idx + child_ptr[v], child_idx + child_ptr[v + 1]);
        par.resize(N);
        ch.resize(N);
        for (int k = 0; k < n_bridges; ++k) {
            const int a = n->kids[3 * k], b = n->kids[3 * k + 1], cc = n->kids[3 * k + 2];
            const int xs[3][3] = {{a, b, cc}, {a, cc, b}, {b, cc, a}};
            const int x = xs[r[k]][0], y = xs[r[k]][1], z = xs[r[k]][2], v = num_nodes + k, u = hub[k];
            ch[u] = {v, z};
            ch[v] = {x, y};
            par[x] = par[y] = v;
            par[v] = u;
        }
        std::vector<int32_t> cp(N + 1, 0), ci, lt(leaf_taxon, leaf_taxon + num_nodes);
        lt.resize(N, -1);
        for (int v = 0; v < N; ++v) {
            ci.insert(ci.end(), ch[v].begin(), ch[v].end());
            cp[v + 1] = (int32_t)ci.size();
        }
        if (ci.empty()) ci.push_back(0);
        hy_engine e = nullptr;
        int rc = hy_create(n->device, &e);
        if (rc != HY_OK) {
            drop_configs(n);
            return nfail(n, rc, "hy_network_set_problem: hy_create failed");
        }
        n->cfg.push_back(e);
        n->res.push_back(r);
        rc = hy_set_problem(e, N, par.data(), cp.data(), ci.data(), lt.data(), num_taxa, num_patterns, states, masks,
                            weights, S);
        if (rc != HY_OK) {
            const int out = child_fail(n, n->cfg.size() - 1, rc);
            drop_configs(n);
            return out;
        }
    }
    return HY_OK;
}

int hy_network_set_model(hy_network n, int code_id, const double *pos_freqs, const double *pi) {
    if (!n) return HY_ERR_ARG;
    if (n->cfg.empty()) return nfail(n, HY_ERR_STATE, "hy_network_set_model before hy_network_set_problem");
    for (size_t c = 0; c < n->cfg.size(); ++c) {
        const int rc = hy_set_model(n->cfg[c], code_id, pos_freqs, pi);
        if (rc != HY_OK) return child_fail(n, c, rc);
    }
    n->have_model = true;
    return HY_OK;
}

int hy_network_set_mixture(hy_network n, int n_mat, int n_class, const int32_t *branch_class,
                           const int32_t *comp_ptr, const int32_t *comp_mat, int time_convention,
                           const int32_t *bridge_class) {
    if (!n) return HY_ERR_ARG;
    if (n->cfg.empty()) return nfail(n, HY_ERR_STATE, "hy_network_set_mixture before hy_network_set_problem");
    if (n_class < 1 || !comp_ptr) return nfail(n, HY_ERR_ARG, "hy_network_set_mixture: bad classes");
    const int N = n->num_nodes + n->n_bridges;
    std::vector<int32_t> bc(N, 0);
    for (int v = 0; v < n->num_nodes; ++v) bc[v] = branch_class ? branch_class[v] : 0;
    for (int k = 0; k < n->n_bridges; ++k) {
        // the bridge branches' class: given, else the hub's
        const int cls = bridge_class ? bridge_class[k] : bc[n->hub[k]];
        if (cls < 0 || cls >= n_class)
            return nfail(n, HY_ERR_ARG, "hy_network_set_mixture: bridge %d class %d out of range", k, cls);
        bc[n->num_nodes + k] = cls;
    }
    for (size_t c = 0; c < n->cfg.size(); ++c) {
        const int rc = hy_set_mixture(n->cfg[c], n_mat, n_class, bc.data(), comp_ptr, comp_mat, time_convention);
        if (rc != HY_OK) return child_fail(n, c, rc);
    }
    n->n_mat = n_mat;
    n->n_comp = comp_ptr[n_class];
    n->branch_class = bc;
    return HY_OK;
}

int hy_network_set_sink(hy_network n, const double *freqs) {
    if (!n) return HY_ERR_ARG;
    for (size_t c = 0; c < n->cfg.size(); ++c) {
        const int rc = hy_set_sink(n->cfg[c], freqs);
        if (rc != HY_OK) return child_fail(n, c, rc);
    }
    return HY_OK;
}

int hy_network_eval(hy_network n, int precision, unsigned flags, const double *bl, const double *bridge_weights,
                    const double *bridge_lengths, double alpha, const double *omegas, const double *comp_w,
                    const double *theta, double delta, double psi, double *out_logL, double *out_site_ll,
                    double *out_grad_bl, double *out_grad_omega, double *out_grad_w, double *out_grad_theta,
                    double *out_grad_bridge_lengths, double *out_grad_bridge_weights) {
    if (!n) return HY_ERR_ARG;
    if (n->cfg.empty() || !n->have_model)
        return nfail(n, HY_ERR_STATE, "hy_network_eval before hy_network_set_problem / hy_network_set_model");
    if (!bl || !omegas || !comp_w || !theta || (n->n_bridges > 0 && (!bridge_weights || !bridge_lengths)))
        return nfail(n, HY_ERR_ARG, "hy_network_eval: NULL parameter array");
    const int K = n->n_bridges, N0 = n->num_nodes, N = N0 + K, P = n->num_patterns;
    for (int i = 0; i < 3 * K; ++i)
        if (!(bridge_weights[i] >= 0.0) || !std::isfinite(bridge_weights[i]) || !(bridge_lengths[i] >= 0.0) ||
            !std::isfinite(bridge_lengths[i]))
            return nfail(n, HY_ERR_ARG, "hy_network_eval: bridge weight / length %d must be finite and >= 0", i);
    const bool grad = (flags & HY_WANT_GRAD) != 0;
    const size_t n_cfg = n->cfg.size();
    // configuration weights W_c = prod_k w_{k, r_k}; zero ones are not evaluated
    std::vector<double> W(n_cfg, 1.0);
    for (size_t c = 0; c < n_cfg; ++c)
        for (int k = 0; k < K; ++k) W[c] *= bridge_weights[3 * k + n->res[c][k]];
    std::vector<double> blc(N);
    auto lengths = [&](size_t c) {
        std::copy(bl, bl + N0, blc.begin());
        for (int k = 0; k < K; ++k) blc[N0 + k] = bridge_lengths[3 * k + n->res[c][k]];
    };
    std::vector<double> sll((size_t)n_cfg * P, -1e20);
    double ll_unused = 0.0;
    for (size_t c = 0; c < n_cfg; ++c) {
        if (!(W[c] > 0.0)) continue;
        lengths(c);
        const int rc = hy_eval(n->cfg[c], precision, HY_WANT_LOGL | HY_WANT_SITE_LL, blc.data(), alpha, omegas,
                               comp_w, theta, delta, psi, &ll_unused, sll.data() + (size_t)c * P, nullptr, nullptr,
                               nullptr, nullptr);
        if (rc != HY_OK) return child_fail(n, c, rc);
    }
    // per-pattern log-sum-exp over the configurations with a positive likelihood
    std::vector<double> pll(P, -1e20), post((size_t)n_cfg * P, 0.0);
    for (int p = 0; p < P; ++p) {
        double mx = -1e300;
        bool any = false;
        for (size_t c = 0; c < n_cfg; ++c)
            if (W[c] > 0.0 && sll[(size_t)c * P + p] > -1e19) {
                mx = std::max(mx, std::log(W[c]) + sll[(size_t)c * P + p]);
                any = true;
            }
        if (!any) continue;
        double se = 0.0;
        for (size_t c = 0; c < n_cfg; ++c)
            if (W[c] > 0.0 && sll[(size_t)c * P + p] > -1e19) {
                const double e = std::exp(std::log(W[c]) + sll[(size_t)c * P + p] - mx);
                post[(size_t)c * P + p] = e;
                se += e;
            }
        pll[p] = mx + std::log(se);
        for (size_t c = 0; c < n_cfg; ++c) post[(size_t)c * P + p] /= se;
    }
    {
        double s = 0.0, cmp = 0.0;  // Neumaier in pattern order
        for (int p = 0; p < P; ++p) {
            const double x = n->weights[p] * pll[p], t = s + x;
            cmp += (std::fabs(s) >= std::fabs(x)) ? (s - t) + x : (x - t) + s;
            s = t;
        }
        if (out_logL) *out_logL = s + cmp;
    }
    if (out_site_ll) std::memcpy(out_site_ll, pll.data(), sizeof(double) * P);
    if (!grad) return HY_OK;
    std::vector<double> gbl(N0, 0.0), gom(n->n_mat, 0.0), gw(n->n_comp, 0.0), gth(5, 0.0), glen(3 * K, 0.0),
        gwt(3 * K, 0.0);
    std::vector<double> tb(N), to(n->n_mat), tw(n->n_comp), tt(5), wp(P);
    for (size_t c = 0; c < n_cfg; ++c) {
        if (!(W[c] > 0.0)) continue;
        double sw = 0.0;
        for (int p = 0; p < P; ++p) {
            const double q = post[(size_t)c * P + p];
            wp[p] = q > 0.0 ? n->weights[p] * q : 0.0;
            sw += wp[p];
        }
        for (int k = 0; k < K; ++k) gwt[3 * k + n->res[c][k]] += sw / bridge_weights[3 * k + n->res[c][k]];
        lengths(c);
        hy_engine e = n->cfg[c];
        int rc = hy_set_pattern_weights(e, wp.data());
        if (rc == HY_OK)
            rc = hy_eval(e, precision, HY_WANT_LOGL | HY_WANT_GRAD, blc.data(), alpha, omegas, comp_w, theta, delta,
                         psi, &ll_unused, nullptr, tb.data(), to.data(), tw.data(), tt.data());
        const int rc2 = hy_set_pattern_weights(e, nullptr);
        if (rc != HY_OK) return child_fail(n, c, rc);
        if (rc2 != HY_OK) return child_fail(n, c, rc2);
        for (int v = 0; v < N0; ++v) gbl[v] += tb[v];
        for (int k = 0; k < K; ++k) glen[3 * k + n->res[c][k]] += tb[N0 + k];
        for (int m = 0; m < n->n_mat; ++m) gom[m] += to[m];
        for (int j = 0; j < n->n_comp; ++j) gw[j] += tw[j];
        for (int j = 0; j < 5; ++j) gth[j] += tt[j];
    }
    if (out_grad_bl) std::memcpy(out_grad_bl, gbl.data(), sizeof(double) * N0);
    if (out_grad_omega) std::memcpy(out_grad_omega, gom.data(), sizeof(double) * n->n_mat);
    if (out_grad_w) std::memcpy(out_grad_w, gw.data(), sizeof(double) * n->n_comp);
    if (out_grad_theta) std::memcpy(out_grad_theta, gth.data(), sizeof(double) * 5);
    if (out_grad_bridge_lengths) std::memcpy(out_grad_bridge_lengths, glen.data(), sizeof(double) * 3 * K);
    if (out_grad_bridge_weights) std::memcpy(out_grad_bridge_weights, gwt.data(), sizeof(double) * 3 * K);
    return HY_OK;
}

}  // extern "C"
