// Compile check of the facade against the way BUSTED's M = 1 likelihood
// (busted.hpp:555-620 of the reference) drives the engine: mixed transition
// matrices per branch built from MG94Matrix spectra with Matrix::Zero, +=,
// scalar products and noalias(), a tile loop over pattern ranges with a
// workspace set up by init(), and the per-pattern log-likelihoods read back
// through ws.root_log_L(b).  Written for this check; it is compiled, not run.
#include <algorithm>
#include <cmath>
#include <vector>

#include "hyphy/core/likelihood.hpp"

using namespace hyphy::core;

double busted_like_loglik(const Tree &tree, const Alignment &aln, const std::vector<const MG94Matrix *> &mg,
                          const MG94Matrix &mg_bg, const std::vector<Scalar> &weights, const std::vector<bool> &is_test,
                          const std::vector<Scalar> &tau, Scalar eps_sink, const Matrix &P_sink) {
    const int S = aln.code ? aln.code->num_sense_codons : 61;
    const size_t num_nodes = tree.num_nodes(), K = mg.size();
    const std::vector<size_t> leaf_to_taxon = LikelihoodEngine::build_leaf_to_taxon_map(tree, aln);
    std::vector<Matrix> P_mix(num_nodes);
    for (size_t n = 0; n < num_nodes; ++n) {
        const int nid = tree.nodes[n].id;
        if (nid == tree.root_id) continue;
        if (is_test[nid]) {
            P_mix[nid] = Matrix::Zero(S, S);
            for (size_t k = 0; k < K; ++k) P_mix[nid] += weights[k] * mg[k]->transition_matrix(tau[nid]);
            if (eps_sink > 0.0) P_mix[nid].noalias() += eps_sink * P_sink;
        } else {
            P_mix[nid] = mg_bg.transition_matrix(tau[nid]);
        }
    }
    const size_t num_patterns = aln.patterns.size(), tile = 64;
    std::vector<Scalar> site(num_patterns, -1e20);
    LikelihoodEngine::BatchedInsideWorkspace ws;
    ws.init(num_nodes, S, tile);
    for (size_t p0 = 0; p0 < num_patterns; p0 += tile) {
        const size_t p1 = std::min(num_patterns, p0 + tile);
        LikelihoodEngine::compute_inside_batch_in_place(tree, aln, p0, p1, leaf_to_taxon, P_mix,
                                                        aln.codon_frequencies_f3x4, S, ws, nullptr);
        for (size_t b = 0; b < p1 - p0; ++b) site[p0 + b] = ws.root_log_L(long(b));
    }
    double total = 0.0;
    for (size_t p = 0; p < num_patterns; ++p) total += aln.patterns[p].weight * site[p];
    // the spectral fields callers read directly
    const Scalar lam0 = mg_bg.eigenvalues(0), u00 = mg_bg.U_pi(0, 0), v00 = mg_bg.V_pi(0, 0);
    return total + 0.0 * (lam0 + u00 + v00 + mg_bg.scale_factor);
}
