/*
 * hyphy_mi355x.h -- C ABI of the MI355X MG94xREV codon likelihood engine.
 *
 * One handle = one GPU, one HIP stream, one host thread at a time.  All
 * arguments are plain pointers and sizes; every function returns HY_OK (0) or
 * a negative HY_ERR_* code and records a message readable with
 * hy_last_error().  Host arrays are copied at call time; the caller keeps
 * ownership of everything it passes in.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   hy_set_problem   Tree + Alignment + leaf_to_taxon arguments of every
 *                    LikelihoodEngine entry (include/hyphy/core/likelihood.hpp:
 *                    343-354, 766-774, 1102-1108, 1228-1236; leaf map :150-178)
 *   hy_set_model     aln.pos_nuc_frequencies / aln.codon_frequencies_f3x4 /
 *                    aln.code as read by MG94Matrix::update
 *                    (include/hyphy/core/rate_matrix.hpp:165-170)
 *   hy_set_mixture   branch-site mixture layout of BUSTED's P_mix and the
 *                    test/background partition of the joint gradient
 *                    (include/hyphy/analyses/busted.hpp:561-581,1191-1216;
 *                    likelihood.hpp:1240-1275)
 *   hy_eval          compute_mg94_log_likelihood (likelihood.hpp:257-292),
 *                    compute_branch_length_gradients (:1102-1211),
 *                    compute_joint_frechet_gradients (:1228-1594),
 *                    BUSTEDAnalyzer::evaluate_log_l (busted.hpp:509-659) and
 *                    compute_busted_branch_gradients (busted.hpp:1138-1429)
 *                    with M = 1 synonymous class
 *   hy_eval_srv      the same two BUSTED entries with M > 1 synonymous-rate
 *                    classes (busted.hpp:554-581, 607-649, 1335-1396)
 *   hy_busted_expected_weights  BUSTED's expectation step for the omega and
 *                    synonymous-rate weights (busted.hpp:662-1135)
 *   hy_eval_batch    a batch of independent problems (genes), one per handle
 *   hy_loglik_from_P compute_tree_log_likelihood_from_P (likelihood.hpp:766-806)
 *   hy_gtr_eval      compute_gtr_log_likelihood (likelihood.hpp:181-254), the
 *                    nucleotide stage of GTRFitter (optimizer.hpp:33-655)
 *   hy_fel_eval      FEL's per-site candidate likelihoods, eval_pattern_lnl
 *                    (fel.hpp:367-385) over many (pattern, alpha, beta)
 *   hy_fel_drive     FELAnalyzer::analyze_pattern's refits (fel.hpp:326-445)
 *                    for every pattern, over any candidate evaluator
 *   hy_fel_run       FELAnalyzer::prepare_branch_lengths + run's refits
 *                    (fel.hpp:227-272, 778-834) on the GPU evaluator
 */
#ifndef HYPHY_MI355X_H
#define HYPHY_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HY_OK 0
#define HY_ERR_ARG (-1)         /* invalid argument or shape                  */
#define HY_ERR_HIP (-2)         /* HIP runtime failure                        */
#define HY_ERR_STATE (-3)       /* call order (e.g. eval before set_problem)  */
#define HY_ERR_NOMEM (-4)       /* device allocation failed                   */
#define HY_ERR_UNSUPPORTED (-5) /* e.g. Wheatstone-bridged trees               */
#define HY_ERR_NUMERIC (-6)     /* eigensolver did not converge               */

/* Arithmetic of the pruning / adjoint contractions. */
#define HY_PREC_FP64 0 /* FP64 MFMA throughout (on-device oracle path)         */
#define HY_PREC_FP32 1 /* FP32 MFMA contractions, FP64 spectra and reductions  */

/* Branch time convention (see hy_set_mixture). */
/* comp_mat entry naming BUSTED's error sink (see hy_set_mixture, hy_set_sink). */
#define HY_SINK (-1)
#define HY_TIME_RAW 0  /* tau_b = bl_b                                          */
#define HY_TIME_CONV 1 /* tau_b = bl_b * 3 / mubar_class                        */

/* hy_eval flags */
#define HY_WANT_LOGL 1u
#define HY_WANT_SITE_LL 2u
#define HY_WANT_GRAD 4u /* branch, omega, mixture-weight and GTR gradients     */

typedef struct hy_engine_s *hy_engine;

/* Create an engine on HIP device `device`. */
int hy_create(int device, hy_engine *out);
void hy_destroy(hy_engine h);
/* n engines on one device that share min(n, n_stream_sets) stream sets
 * (n_stream_sets <= 0: 8) round-robin, for batches of many small problems
 * (one engine per gene, evaluated together by hy_eval_batch).  Engines of a
 * group must be driven from one host thread at a time and launch their
 * kernels directly (no captured graphs).  Each is released with hy_destroy.
 * No counterpart in the reference. */
int hy_create_group(int device, int n, int n_stream_sets, hy_engine *out);
/* Message of the last failing call on this handle (or of a failed hy_create
 * when h is NULL).  Valid until the next call on the handle. */
const char *hy_last_error(hy_engine h);

/* Tree and compressed alignment.
 *   num_nodes            nodes of the (unrooted or rooted) tree, ids 0..num_nodes-1
 *   parent[num_nodes]    parent id, -1 at the root (exactly one root)
 *   child_ptr[num_nodes+1], child_idx[]   children of node n in Newick order:
 *                        child_idx[child_ptr[n] .. child_ptr[n+1]-1]
 *   leaf_taxon[num_nodes] alignment row of each leaf, -1 when unmatched
 *                        (an unmatched leaf contributes all-ones partials,
 *                        likelihood.hpp:62-65); ignored for internal nodes
 *   states[num_patterns*num_taxa]  sense-codon index or -1 (pattern-major)
 *   masks[num_patterns*num_taxa]   compatible-codon bit masks
 *                        (alignment.hpp:26-31, genetic_code.hpp:82-109)
 *   weights[num_patterns] pattern multiplicities
 *   S                    number of sense codons of the genetic code (<= 64;
 *                        4 for a nucleotide view)
 * Re-calling replaces the problem and frees the previous device buffers; a
 * model set for another S is dropped (hy_set_model again before evaluating). */
int hy_set_problem(hy_engine h, int num_nodes, const int32_t *parent, const int32_t *child_ptr,
                   const int32_t *child_idx, const int32_t *leaf_taxon, int num_taxa, int num_patterns,
                   const int8_t *states, const uint64_t *masks, const double *weights, int S);

/* Genetic code (registry id 0..21 in the reference's order,
 * genetic_code.hpp:274-370), CF3x4-corrected position frequencies
 * pos_freqs[3*4] (row = codon position, column = A,C,G,T) and the codon
 * frequencies pi[S]. */
int hy_set_model(hy_engine h, int code_id, const double *pos_freqs, const double *pi);

/* Branch-site mixture layout.
 *   n_mat rate matrices; matrix m has nonsynonymous rate omegas[m] at eval time.
 *   branch_class[num_nodes]: class of the branch above each node (root ignored).
 *   Class c mixes components comp_ptr[c] .. comp_ptr[c+1]-1; component j uses
 *   matrix comp_mat[j] with weight comp_w[j] (weights given at eval time):
 *       P_b = sum_j comp_w[j] * exp(Q_{comp_mat[j]} tau_b)
 *   time_convention HY_TIME_RAW or HY_TIME_CONV (mubar_c = sum_j w_j mu_j).
 *   comp_mat[j] == HY_SINK makes component j BUSTED's error sink
 *   (busted.hpp:571-577): P_b gains comp_w[j] * P_sink, independent of tau_b
 *   (matrix from hy_set_sink; HY_TIME_RAW only; not in gene batches).  The
 *   reference's (1 - eps) sum_k w_k P_k + eps P_sink is comp_w = ((1 - eps)
 *   w_k, eps); out_grad_w of the sink component is sum_b <A_b, P_sink>.
 * Default after hy_set_problem: one matrix, one class, one component, RAW. */
int hy_set_mixture(hy_engine h, int n_mat, int n_class, const int32_t *branch_class, const int32_t *comp_ptr,
                   const int32_t *comp_mat, int time_convention);

/* Error-sink matrix P_sink(i, j) = f_j / max(1e-12, 1 - f_i) for i != j, zero
 * diagonal, from the codon frequencies f[S] (the reference builds it from the
 * alignment's F3x4 codon frequencies: BUSTEDAnalyzer::build_sink_kernel,
 * busted.hpp:193-206).  Needed before evaluating a mixture with HY_SINK
 * components. */
int hy_set_sink(hy_engine h, const double *freqs);

/* Evaluate lnL (and gradients) at the given parameters.
 *   bl[num_nodes]        branch lengths (root entry ignored)
 *   alpha                synonymous rate (MG94Parameters::alpha)
 *   omegas[n_mat]        nonsynonymous rates (MG94Parameters::beta per matrix)
 *   comp_w[n_comp]       mixture weights
 *   theta[5]             GTR exchangeabilities AC, AT, CG, CT, GT (AG == 1)
 *   delta, psi           multi-hit rates (forward only; gradients assume 0)
 * Outputs (NULL to skip): out_logL, out_site_ll[num_patterns] (per-pattern lnL,
 * unweighted), out_grad_bl[num_nodes] (d lnL/d bl, root 0), out_grad_omega
 * [n_mat], out_grad_w[n_comp], out_grad_theta[5]. */
int hy_eval(hy_engine h, int precision, unsigned flags, const double *bl, double alpha, const double *omegas,
            const double *comp_w, const double *theta, double delta, double psi, double *out_logL,
            double *out_site_ll, double *out_grad_bl, double *out_grad_omega, double *out_grad_w,
            double *out_grad_theta);

/* Synonymous-rate (SRV) mixture on top of the branch-site mixture, as
 * BUSTEDAnalyzer::evaluate_log_l (busted.hpp:509-659) and
 * compute_busted_branch_gradients (busted.hpp:1138-1429) with M = n_syn:
 * class m multiplies every branch time by syn_rates[m] and has weight
 * syn_weights[m]; lnL = sum_p w_p log sum_m v_m L_m(p) (log-sum-exp; classes
 * with v_m <= 1e-12 are skipped, a pattern with no positive term gets
 * -1e20).  Gradients are posterior-weighted over the classes and extend the
 * reference's branch gradients to omega, mixture weights, theta and
 * out_grad_syn_rates[n_syn] / out_grad_syn_weights[n_syn] (d lnL / d v_m,
 * 0 for skipped classes).  out_site_ll receives the combined per-pattern lnL.
 * Costs about 2 * n_syn evaluations (a forward pass per class for the
 * posteriors, then a gradient pass per class). */
int hy_eval_srv(hy_engine h, int precision, unsigned flags, const double *bl, double alpha, const double *omegas,
                const double *comp_w, const double *theta, double delta, double psi, int n_syn,
                const double *syn_rates, const double *syn_weights, double *out_logL, double *out_site_ll,
                double *out_grad_bl, double *out_grad_omega, double *out_grad_w, double *out_grad_theta,
                double *out_grad_syn_rates, double *out_grad_syn_weights);

/* BUSTED expectation step: BUSTEDAnalyzer::compute_expected_weights
 * (busted.hpp:662-976, including the error-sink update :846-860, 964-966)
 * and compute_expected_syn_weights (busted.hpp:978-1135).  Needs a mixture
 * with HY_TIME_RAW branch times (bl = the reference's s * branch_tau).
 * out_weights[k] (k over the components of test_class, in order): for an
 * omega component, clamp(sum over test branches and patterns of w_p w_k
 * post_m(p) V_b . (P_k D_b) / L_m(p), divided by (test branches x the weight
 * of the patterns with a positive likelihood), 1e-6, 1 - 1e-6), renormalised
 * over the omega components ({1} when the class has one); for the class's
 * error-sink component (comp_mat == HY_SINK, weight eps, the omega components
 * then weighted (1 - eps) w_k as in the reference), the new eps =
 * clamp(eps sum ... V_b . P_sink D_b / L_m(p) / max(1, that divisor), 1e-7,
 * max_eps_sink) (max_eps_sink is ignored without a sink).
 * out_syn_weights[m] = clamp(sum_p w_p post_m(p) / that pattern weight, 1e-6,
 * 1 - 1e-6), renormalised ({1} when n_syn = 1).  Computed from one
 * hy_eval_srv gradient evaluation (the accumulated sums are w_k dlnL/dw_k,
 * eps dlnL/d eps and v_m dlnL/dv_m). */
int hy_busted_expected_weights(hy_engine h, int precision, const double *bl, double alpha, const double *omegas,
                               const double *comp_w, const double *theta, double delta, double psi, int n_syn,
                               const double *syn_rates, const double *syn_weights, int test_class,
                               double max_eps_sink, double *out_weights, double *out_syn_weights);

/* Many independent problems (genes) in one call: item i evaluates handle
 * items[i].h exactly as hy_eval would with the item's parameters and
 * outputs.  All evaluations are queued (each on its handle's stream) before
 * any is waited for, so the genes run concurrently on the device.  Handles
 * must be distinct.  Returns HY_OK or the first failing item's code; every
 * item's own code is left in items[i].status.  Replaces the per-gene loops
 * of the reference's drivers over LikelihoodEngine calls (one problem per
 * call, likelihood.hpp:257-292, 1228-1236). */
typedef struct hy_batch_item {
    hy_engine h;
    const double *bl;
    double alpha;
    const double *omegas;
    const double *comp_w;
    const double *theta;
    double delta, psi;
    double *out_logL, *out_site_ll, *out_grad_bl, *out_grad_omega, *out_grad_w, *out_grad_theta;
    int status;
} hy_batch_item;
int hy_eval_batch(int n_items, hy_batch_item *items, int precision, unsigned flags);

/* Autotuner of fused gene batches: times `reps` hy_eval_batch calls (after
 * an untimed one) of a batch that fuses (every item a distinct engine of one
 * group) at every pattern-tile width (16, 32, 64) and keeps the fastest for
 * this precision in the group; later fused batches of the group use it.
 * Outputs (each may be NULL): the tile and its median time per batch in
 * milliseconds.  The items' outputs receive the last evaluation's results.
 * No counterpart in the reference. */
int hy_autotune_batch(int n_items, hy_batch_item *items, int precision, unsigned flags, int reps, int *out_tile,
                      double *out_ms);

/* Nucleotide GTR stage: compute_gtr_log_likelihood (likelihood.hpp:181-254)
 * on a problem set with S = 4, i.e. the nucleotide view of the alignment
 * (Alignment::to_nucleotide_alignment, alignment.hpp:832-890: states 0..3 =
 * A, C, G, T or -1, masks = 4-bit IUPAC masks).  The generator is
 * GTRMatrix::update's (rate_matrix.hpp:29-99): exchangeabilities theta[5]
 * (AC, AT, CG, CT, GT; AG = 1), frequencies freqs[4] (also the root
 * distribution: the alignment's nucleotide frequencies), normalised to one
 * expected substitution per unit of bl[num_nodes] (root entry ignored).
 * flags as hy_eval: out_logL (Neumaier sum over patterns), out_site_ll
 * [num_patterns] with HY_WANT_SITE_LL, and with HY_WANT_GRAD
 * out_grad_bl[num_nodes] and out_grad_theta[5] (d lnL / d theta of the
 * normalised model; the reference's GTRFitter is derivative-free,
 * optimizer.hpp:33-655, so these have no counterpart there). */
int hy_gtr_eval(hy_engine h, int precision, unsigned flags, const double *bl, const double *theta, const double *freqs,
                double *out_logL, double *out_site_ll, double *out_grad_bl, double *out_grad_theta);

/* FEL per-site refit candidates, FP64: candidate k is pattern pattern[k] at
 * synonymous rate alpha[k] and non-synonymous rate beta[k], evaluated as
 * FELAnalyzer::analyze_pattern's eval_pattern_lnl (fel.hpp:367-385): MG94
 * with alpha = 1, beta = omega = beta[k] / max(alpha[k], 1e-12), the
 * exchangeabilities theta[5] (AC, AT, CG, CT, GT) and delta, psi of the base
 * model (rate_matrix.hpp:170-247, unnormalised), P_b = max(U diag(e^{l a t_b})
 * V, 0) with a = max(alpha[k], 1e-12) and t_b = bl[b] (the tree's lengths
 * after the analyzer's conversion, fel.hpp:240-271), then the pattern's
 * log-likelihood (compute_pattern_log_likelihood, fel.hpp:276-308; -1e20
 * where the likelihood is not positive) in out_ll[k].  out_scale (may be
 * NULL) receives each candidate's generator mean rate sum_i pi_i (-Q_ii)
 * (MG94Matrix::scale_factor).  Patterns may repeat within a call.  The
 * handle keeps a pool of spectra keyed by omega (for one theta, delta, psi
 * and problem): a candidate whose omega is pooled reuses its spectrum,
 * every other distinct omega is solved in one launch, refining the
 * eigenvectors of the pooled omega nearest in log(omega).  Out-of-range
 * patterns, negative or non-finite rates are refused with HY_ERR_ARG, an
 * eigensolver failure with HY_ERR_NUMERIC.  Requires hy_set_problem and
 * hy_set_model. */
int hy_fel_eval(hy_engine h, int n_cand, const int32_t *pattern, const double *alpha, const double *beta,
                const double *theta, double delta, double psi, const double *bl, double *out_ll, double *out_scale);

/* MEME per-site refit candidates, FP64: candidate k is pattern pattern[k]
 * at MEMEAnalyzer::analyze_pattern's eval_pattern_meme_lnl(a, w1, p1,
 * b_plus, b_bg) (meme.hpp:219-277; no bridges): with a = max(alpha[k], 1e-12)
 * the generators MG94(alpha = a, beta = a omega1[k]), MG94(a, beta_plus[k])
 * and, when a branch is background, MG94(a, beta_bg[k]) (theta[5], delta,
 * psi of the base model); test branches (is_test[node] != 0; NULL = all)
 * take P = p1 max(P_1, 0) + (1 - p1) max(P_+, 0) at t_b = bl[b] (the tree's
 * lengths after the analyzer's conversion), background branches
 * max(P_bg, 0); out_ll[k] the pattern's lnL (compute_pattern_log_likelihood,
 * fel.hpp:276-308; -1e20 where not positive).  The spectra are those of the
 * alpha = 1 generators at omega1, beta_plus / a and beta_bg / a (the
 * generators are a times those), drawn from the handle's pool like
 * hy_fel_eval's.  beta_bg may be NULL when every branch is a test branch.
 * cond_node / cond_class (both NULL, or one entry per candidate): a candidate
 * with cond_node[k] >= 0 takes max(P_1, 0) (class 0) or max(P_+, 0)
 * (class 1) alone on that test branch and the mixture elsewhere, the
 * conditional likelihoods of the branch Empirical Bayes Factors
 * (meme.hpp:552-600). */
int hy_meme_eval(hy_engine h, int n_cand, const int32_t *pattern, const double *alpha, const double *omega1,
                 const double *p1, const double *beta_plus, const double *beta_bg, const uint8_t *is_test,
                 const double *theta, double delta, double psi, const double *bl, const int32_t *cond_node,
                 const int32_t *cond_class, double *out_ll);

/* The spectrum of one MG94 generator on the handle's model (hy_set_model's
 * code and frequencies; no problem needed): MG94Matrix::update
 * (rate_matrix.hpp:165-248) with rates alpha, beta, theta[5] (AC, AT, CG,
 * CT, GT), delta, psi.  Outputs, S = the code's sense codons: lam[S] the
 * eigenvalues of the symmetrised generator clamped to <= 0 in increasing
 * order, U[S*S] = U_pi = diag(1/sqrt(pi)) Z and V[S*S] = V_pi =
 * Z^T diag(sqrt(pi)) row-major (P(t) = U_pi diag(e^{lam t}) V_pi), mu (may be
 * NULL) the mean rate sum_i pi_i (-Q_ii).  Solved on the GPU; a call with
 * the same model refines the previous call's eigenvectors. */
int hy_mg94_spectrum(hy_engine h, double alpha, double beta, const double *theta, double delta, double psi,
                     double *lam, double *U, double *V, double *mu);

/* Counters of hy_fel_eval since the handle was made (or the last reset):
 * out4 = {candidates, candidates whose omega was already pooled, spectra
 * solved by warm refinement, spectra solved cold}.  No reference
 * counterpart (diagnostics of the spectrum pool). */
int hy_fel_counters(hy_engine h, int64_t *out4, int reset);

/* One row per pattern of the FEL refit results (SiteResult, fel.hpp:76-89). */
enum {
    HY_FEL_ALPHA = 0,        /* synonymous rate of the alternative fit */
    HY_FEL_BETA = 1,         /* non-synonymous rate of the alternative fit */
    HY_FEL_BETA_NUISANCE = 2,
    HY_FEL_ALPHA_NULL = 3,   /* alpha = beta of the null fit */
    HY_FEL_LRT = 4,
    HY_FEL_P_VALUE = 5,      /* erfc(sqrt(LRT / 2)) */
    HY_FEL_LENGTH = 6,       /* alpha_null x tested tree length */
    HY_FEL_LOGL_ALT = 7,
    HY_FEL_LOGL_NULL = 8,
    HY_FEL_EVALS = 9,        /* candidate evaluations of the pattern */
    HY_FEL_FIELDS = 10
};

/* Candidate evaluator of hy_fel_drive: n candidates (pattern[k], alpha[k],
 * beta[k]) -> out_ll[k]; a non-zero return aborts the drive with that code. */
typedef int (*hy_fel_eval_fn)(void *user, int n, const int32_t *pattern, const double *alpha, const double *beta,
                              double *out_ll);

/* FELAnalyzer::analyze_pattern with every branch tested (fel.hpp:326-445)
 * for patterns 0 .. n_patterns-1: invariant patterns (all resolved codons
 * equal, fel.hpp:333-361) get zeros and p = 1; every other one a 12-point
 * start grid, NelderMead2D in [1e-4, 1000]^2 (tol 1e-4, 100 iterations,
 * nelder_mead.hpp:121-204), the null alpha = beta from its start value and a
 * 7-point grid, Brent1D on [1e-4, 50] (tol 1e-4, 60 iterations,
 * nelder_mead.hpp:24-101), the LRT and its p-value.  The refits run as
 * coroutines in lockstep: each call of `evaluate` holds the pending
 * candidates of every unfinished pattern (the start grid and the null grid
 * as one batch each), so each pattern sees the reference's own sequence of
 * evaluations.  out[n_patterns * HY_FEL_FIELDS]; out_calls and
 * out_candidates (may be NULL) count the evaluator calls and candidates.
 * Host only (no GPU needed unless the evaluator uses one). */
int hy_fel_drive(int n_patterns, const uint8_t *invariant, double tested_length, hy_fel_eval_fn evaluate,
                 void *user, double *out, int *out_calls, int64_t *out_candidates);

/* FEL on the handle's problem and model, every branch tested: the fitted
 * tree's lengths bl[num_nodes] are converted by 3 / Q_scale of the global
 * MG94 model (alpha, beta, theta[5], delta, psi; FELAnalyzer::
 * prepare_branch_lengths, fel.hpp:227-272), then hy_fel_drive refits every
 * pattern with hy_fel_eval as the evaluator.  out[num_patterns *
 * HY_FEL_FIELDS] per pattern (sites map to patterns as the alignment's
 * site_to_pattern); out_q_scale (may be NULL) receives Q_scale. */
int hy_fel_run(hy_engine h, const double *theta, double alpha, double beta, double delta, double psi,
               const double *bl, double *out, int *out_calls, int64_t *out_candidates, double *out_q_scale);

/* One row per pattern of the MEME refit results (MEMESiteResult,
 * meme.hpp:31-50). */
enum {
    HY_MEME_ALPHA = 0,
    HY_MEME_BETA1 = 1,          /* alpha x omega1 */
    HY_MEME_P1 = 2,
    HY_MEME_BETA_PLUS = 3,
    HY_MEME_P_PLUS = 4,
    HY_MEME_LRT = 5,
    HY_MEME_P_VALUE = 6,        /* 2/3 (0.45 erfc(sqrt(LRT / 2)) + 0.55 e^{-LRT / 2}) */
    HY_MEME_BRANCHES = 7,       /* branches with EBF >= 100 */
    HY_MEME_LENGTH = 8,         /* total branch length at the fitted rates */
    HY_MEME_LOGL = 9,
    HY_MEME_LOGL_FEL = 10,
    HY_MEME_P_VS_FEL = 11,      /* e^{-(lnL_MEME - lnL_FEL)} when positive, else 1 */
    HY_MEME_FEL_ALPHA = 12,
    HY_MEME_FEL_BETA = 13,
    HY_MEME_BETA_BG = 14,       /* 0: every branch tested */
    HY_MEME_EVALS = 15,         /* candidate evaluations of the pattern */
    HY_MEME_FIELDS = 16
};

/* Candidate evaluator of hy_meme_drive: n candidates (pattern[k], alpha[k],
 * omega1[k], p1[k], beta_plus[k], cond_node[k], cond_class[k]) -> out_ll[k],
 * with hy_meme_eval's meaning (cond_node[k] = -1: no conditioning); a
 * non-zero return aborts the drive with that code. */
typedef int (*hy_meme_eval_fn)(void *user, int n, const int32_t *pattern, const double *alpha,
                               const double *omega1, const double *p1, const double *beta_plus,
                               const int32_t *cond_node, const int32_t *cond_class, double *out_ll);

/* MEMEAnalyzer::analyze_pattern with every branch tested (meme.hpp:166-648,
 * no bridges) for patterns 0 .. n_patterns-1, from each pattern's FEL row
 * fel[p * HY_FEL_FIELDS] (hy_fel_drive's layout): invariant patterns get
 * zeros, p1 = 1 and p = 1; every other one the regime's start grid, a 4-D
 * NelderMeadND over (alpha, omega1, p1, beta_plus) in [1e-4, 1e4] x [0, 1] x
 * [1e-6, 1 - 1e-6] x [1e-4, 1e5] (tol 1e-5, 800 iterations,
 * nelder_mead.hpp:208-360); when beta_plus > alpha and p_plus > 1e-6 the
 * null beta_plus = alpha as a 3-D NelderMeadND from its 20-point grid, a
 * warm alternative refit when the null is better, the LRT and its mixture
 * p-value, and the branch EBFs (ebf[p * num_nodes + node], may be NULL; 1 for
 * the root and for patterns without them).  HY_MEME_LENGTH is left 0 (it
 * needs the generators' mean rates; hy_meme_run fills it).  Lockstep
 * coroutines as hy_fel_drive: independent points (grids, a simplex's first
 * vertices, shrink steps, the EBF conditionals) are asked for as one batch.
 * Host only. */
int hy_meme_drive(int n_patterns, const uint8_t *invariant, const double *fel, int num_nodes, int root,
                  hy_meme_eval_fn evaluate, void *user, double *out, double *ebf, int *out_calls,
                  int64_t *out_candidates);

/* MEME on the handle's problem and model, every branch tested: hy_fel_run
 * (the same conversion of bl by 3 / Q_scale, MEMEAnalyzer::
 * prepare_branch_lengths, meme.hpp:150-163), then hy_meme_drive with
 * hy_meme_eval as the evaluator, then each pattern's total branch length
 * sum_b t_b (p1 mu(alpha, beta1) + p_plus mu(alpha, beta_plus)) / 3
 * (meme.hpp:614-638).  out[num_patterns * HY_MEME_FIELDS], ebf[num_patterns *
 * num_nodes] (may be NULL), fel[num_patterns * HY_FEL_FIELDS] (may be NULL)
 * the FEL rows it started from. */
int hy_meme_run(hy_engine h, const double *theta, double alpha, double beta, double delta, double psi,
                const double *bl, double *out, double *ebf, double *fel, int *out_calls, int64_t *out_candidates);

/* lnL from caller-supplied transition matrices P[num_nodes*S*S] (row-major,
 * entry of the root ignored) and root frequencies pi[S]. */
int hy_loglik_from_P(hy_engine h, int precision, const double *P, const double *pi, double *out_logL,
                     double *out_site_ll);

/* As hy_loglik_from_P, followed by the outside (pre-order) pass, so that
 * hy_read_partials can return outside messages as well:
 * compute_inside_outside_batch_in_place (likelihood.hpp:479-762) over all
 * patterns. */
int hy_inside_outside_from_P(hy_engine h, int precision, const double *P, const double *pi, double *out_logL,
                             double *out_site_ll);

/* As hy_loglik_from_P (same inputs, outputs and result to rounding), but
 * allocating nothing: no device, pinned or library allocation on any call.
 * No counterpart in the reference; for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid it).
 * The problem must be set (hy_set_problem, which allocates, is set-up).
 * Every per-call device buffer (the caller's P and pi, tip codes, transition
 * tiles, inside partials and exponents, root and site likelihoods, the
 * result) lives in one fixed scratch array of 128 MiB per device, part of the
 * library's code object (held until the process ends), in a fixed
 * configuration: 16-pattern tiles, one whole-tree inside schedule.  A problem
 * needs about num_nodes * S^2 * 8 + (internal nodes) * patterns * (256 FP32 |
 * 512 FP64) bytes; larger ones are refused with HY_ERR_ARG (use
 * hy_loglik_from_P).  All calls on one device share the array: calls that
 * overlap on one device (from several threads or handles) race and must be
 * serialised by the caller.  Every value read is written by the same call,
 * so nothing carries over from the previous one.  Blocks until the results
 * are in out_logL / out_site_ll[num_patterns]. */
int hy_loglik_from_P_noalloc(hy_engine h, int precision, const double *P, const double *pi, double *out_logL,
                             double *out_site_ll);

/* Partials of the last evaluation on h for patterns [p_start, p_end)
 * (B = p_end - p_start), the read-back behind the reference's workspaces
 * (BatchedInsideWorkspace / BatchedInsideOutsideWorkspace,
 * likelihood.hpp:295-340).  Values are stored as mantissa x 2^exponent:
 *   D[num_nodes][B][S], E[num_nodes][B]  conditional likelihood below each
 *       internal node u, D_u = D[u] * 2^E[u] (leaf entries untouched);
 *   V[num_nodes][B][S], G[num_nodes][B]  outside message of the branch above
 *       each non-root node v, V_v = V[v] * 2^G[v], with
 *       V_v . (P_v D_v) = L for every v (needs an outside pass: a gradient
 *       hy_eval or hy_inside_outside_from_P);
 *   L[B], LE[B]  root likelihood L * 2^LE.
 * Any output may be NULL.  Values are in the evaluation's precision. */
int hy_read_partials(hy_engine h, int p_start, int p_end, double *D, int32_t *E, double *V, int32_t *G, double *L,
                     int32_t *LE);

/* Pattern weights of every later evaluation on h (lnL and gradients),
 * weights[num_patterns] >= 0 in place of the alignment's multiplicities, as
 * the reference's site-level mixtures weight each pattern's gradient by its
 * class posterior (RELAXAnalyzer::compute_relax_branch_gradients,
 * relax.hpp:1080-1125; BUSTED's synonymous classes, busted.hpp:1383-1386).
 * NULL restores the alignment's weights; hy_set_problem clears them.  While
 * set, hy_eval_srv and hy_busted_expected_weights refuse (they weight
 * patterns themselves) and gene batches run engine by engine. */
int hy_set_pattern_weights(hy_engine h, const double *weights);

/* aBSREL's local projection of one branch: the branch context of
 * compute_branch_context_matrices (absrel.hpp:723-780; V_mat and D_mat of the
 * branch above `node` with their log-scale offset) scored by
 * eval_branch_ll_from_matrices (absrel.hpp:438-455) for n_cand candidate
 * transition matrices at once, on the GPU and against the partials the last
 * gradient evaluation on h left resident (hy_eval with HY_WANT_GRAD in
 * `precision`; not after a checkpointed evaluation).  P_cand[n_cand*S*S]
 * row-major; out_lnl[n_cand] = sum_p w_p ln(V_p^T P_c D_p) with the partials'
 * scaling folded in (the reference's offset included; w_p (-1e20) for a
 * pattern whose product is not positive).  Optional out_site_L[n_cand *
 * num_patterns] (the products in the partials' scaling) and out_site_e
 * [num_patterns]: L_c,p = out_site_L[c][p] * 2^out_site_e[p], one exponent
 * per pattern for every candidate, so candidates can be mixed per pattern
 * (absrel.hpp:1149-1167). */
int hy_branch_eval(hy_engine h, int precision, int node, int n_cand, const double *P_cand, double *out_lnl,
                   double *out_site_L, int32_t *out_site_e);

/* Pattern-tile width used by the pruning kernels (16, 32 or 64 patterns per
 * workgroup; FP64 evaluations run 64 as 32; 0 = automatic, chosen per
 * precision from the pattern count and the device's CU count).  Applies from the next hy_eval; a change of tile
 * width (e.g. when alternating precisions) re-tiles the resident problem. */
int hy_set_tile(hy_engine h, int patterns_per_tile);

/* Tile width the last hy_eval / hy_loglik_from_P used.  No counterpart in the
 * reference (tuning introspection). */
int hy_last_tile(hy_engine h, int *patterns_per_tile);

/* Autotuner: times `reps` evaluations (flags as hy_eval, at the given
 * parameters, after one untimed evaluation) for every pattern-tile width
 * (16, 32, 64) with and without subtree parts, and keeps the fastest
 * configuration for this precision: later evaluations use it unless
 * hy_set_tile requested a tile.  Outputs (each may be NULL): the chosen tile,
 * whether subtree parts are used (0/1) and its median time per evaluation in
 * milliseconds (host clock around each synchronous evaluation).  No
 * counterpart in the reference, which has a fixed tile rule
 * (busted.hpp:385-401). */
int hy_autotune(hy_engine h, int precision, unsigned flags, const double *bl, double alpha, const double *omegas,
                const double *comp_w, const double *theta, double delta, double psi, int reps, int *out_tile,
                int *out_parts, double *out_ms);

/* Applies a tuning choice (as hy_autotune returns it) for one precision: tile
 * 16, 32 or 64 (0: the cost model) and subtree parts 1/0 (-1: the rule), e.g.
 * one measured in an earlier run.  hy_set_tile still takes precedence.  No
 * counterpart in the reference. */
int hy_set_tuning(hy_engine h, int precision, int patterns_per_tile, int parts);

/* Subtree parts the last evaluation split each pruning pass into (0: one
 * schedule for the whole tree; n > 0: n subtree parts plus the top part, two
 * launches per pass).  No counterpart in the reference (tuning introspection). */
int hy_last_parts(hy_engine h, int *n_parts);

/* Warm starts (default on): each hy_eval refines the previous call's
 * eigenvectors of every rate matrix (Ogita-Aishima iteration, FP64) and only
 * falls back to a cold Jacobi solve when that does not converge; results agree
 * with cold solves to rounding.  Invalidated by hy_set_problem, hy_set_model
 * and hy_set_mixture.  No counterpart in the reference, which re-solves every
 * MG94Matrix::update (rate_matrix.hpp:236-248). */
int hy_set_warm_start(hy_engine h, int enable);

/* Eigensolver status of the last hy_eval per rate matrix: 100 + k after k
 * Ogita-Aishima refinement steps of a warm start (100: the previous
 * eigenvectors already diagonalised the matrix to 1e-12 and were kept
 * unchanged); 50 + k after k such steps from a cold tridiagonal solve
 * (Householder tridiagonalisation, Sturm multisection, twisted
 * factorisations); 1 .. 40 Jacobi sweeps when neither converged; negative:
 * no convergence. */
int hy_last_eigen_status(hy_engine h, int *status, int n);

/* Timing of the last hy_eval's device work, in milliseconds, per stage:
 * [0] spectra, [1] P(t), [2] inside, [3] outside, [4] adjoint accumulation,
 * [5] Frechet projection, [6] reductions, [7] total.  Only filled when the
 * engine was created with HY_PROFILE=1 in the environment. */
int hy_last_timings(hy_engine h, double *ms8);

/* Memory budget of the partials (inside partials D, outside messages V,
 * spilled messages and the exponents) in bytes; 0 (default) = the device's
 * free memory at the evaluation (plus what the engine's partial buffers hold
 * already, less 1 GiB).  When every node's partials would exceed it, the
 * engine checkpoints (inside partials kept at about sqrt(internal nodes)
 * nodes and recomputed): the tree is cut into parts of at most about
 * sqrt(internal nodes) internal nodes; only the partials of the top part
 * and of the part roots stay resident, and each part's interior is
 * recomputed into one segment shared by the parts right before that part's
 * outside pass and adjoint.  Same results to rounding; costs about one more
 * inside pass; the workspace read-back (hy_inside_outside_from_P,
 * hy_read_partials) is refused then.  Applies from the next evaluation.  No
 * counterpart in the reference, which materialises every node's partials
 * (likelihood.hpp:1310-1315). */
int hy_set_memory_budget(hy_engine h, size_t bytes);

/* Checkpointing of the last evaluation: *n_parts parts recomputed (0: all
 * partials resident) and *partial_bytes held by the partial buffers of its
 * precision.  No counterpart in the reference. */
int hy_last_checkpoint(hy_engine h, int *n_parts, size_t *partial_bytes);

/* Instantiated evaluation graphs alive in this process and their cap.
 * Evaluations capture one HIP graph per (precision, gradient, warm start)
 * and replay it; at most *cap graphs (HY_MAX_GRAPHS in the environment,
 * default 512) exist at once across all handles, and evaluations that would
 * need another launch their kernels directly (same kernels, same results).
 * No counterpart in the reference. */
int hy_graph_count(int *live, int *cap);

/* Device memory currently held by the engine, in bytes. */
int hy_device_bytes(hy_engine h, size_t *out);

/* ---------------------------------------------------------------------
 * Wheatstone-bridged trees (Tree::bridges with WheatstoneBridge hubs,
 * tree.hpp:32-69; evaluated by LikelihoodEngine's bridge branches,
 * likelihood.hpp:376-425, 520-569, 623-698, and WheatstoneLikelihood,
 * wheatstone_likelihood.hpp:21-148, 538-800).  A bridge at hub u with
 * children {A, B, C} gives u the partial sum_r w_r (P_br(r) (M_x * M_y)) * M_z
 * over the resolutions (x, y | z) = (A, B | C), (A, C | B), (B, C | A),
 * M_c = P_c D_c.  The network is evaluated exactly as the per-pattern
 * mixture of the binary trees that resolve every bridge one way (weights =
 * products of the chosen w_r; the bridge branch joins the chosen pair to
 * u), one engine per configuration (3^n_bridges engines).
 * --------------------------------------------------------------------- */
#define HY_MAX_BRIDGES 5
typedef struct hy_network_s *hy_network;

int hy_network_create(int device, hy_network *out);
void hy_network_destroy(hy_network n);
const char *hy_network_last_error(hy_network n);

/* The tree and patterns as hy_set_problem; hub[k] is the hub node of
 * bridge k (not the root) and hub_children[3k..3k+2] its three children
 * (A, B, C in the reference's order), which must be the hub's children in
 * the tree.  Resets the mixture to one component. */
int hy_network_set_problem(hy_network n, int num_nodes, const int32_t *parent, const int32_t *child_ptr,
                           const int32_t *child_idx, const int32_t *leaf_taxon, int num_taxa, int num_patterns,
                           const int8_t *states, const uint64_t *masks, const double *weights, int S, int n_bridges,
                           const int32_t *hub, const int32_t *hub_children);
int hy_network_set_model(hy_network n, int code_id, const double *pos_freqs, const double *pi);
/* As hy_set_mixture; bridge_class[n_bridges] (NULL: the hub's class) is the
 * class of the bridge branches (the reference takes the bridge node's
 * partition, likelihood.hpp:1283-1287). */
int hy_network_set_mixture(hy_network n, int n_mat, int n_class, const int32_t *branch_class,
                           const int32_t *comp_ptr, const int32_t *comp_mat, int time_convention,
                           const int32_t *bridge_class);
int hy_network_set_sink(hy_network n, const double *freqs);
/* As hy_eval, with bridge_weights[3k + r] = w_r and bridge_lengths[3k + r]
 * = t_r of bridge k (WheatstoneBridge::weights / branch_lengths, in the
 * time convention of the bridge's class).  Outputs as hy_eval plus
 * out_grad_bridge_lengths[3 n_bridges] and out_grad_bridge_weights
 * [3 n_bridges] (JointWheatstoneMG94Gradients::grad_bridge_lengths /
 * grad_bridge_weights, wheatstone_likelihood.hpp:527-529; each weight
 * varied on its own).  Any output may be NULL. */
int hy_network_eval(hy_network n, int precision, unsigned flags, const double *bl, const double *bridge_weights,
                    const double *bridge_lengths, double alpha, const double *omegas, const double *comp_w,
                    const double *theta, double delta, double psi, double *out_logL, double *out_site_ll,
                    double *out_grad_bl, double *out_grad_omega, double *out_grad_w, double *out_grad_theta,
                    double *out_grad_bridge_lengths, double *out_grad_bridge_weights);

/* ---- Maximum-likelihood fit with captured optimizer iterations ----------
 * Replaces the joint phase of MG94Fitter::fit (optimizer.hpp:934-1110):
 * branch lengths, the omegas and the five GTR exchangeabilities in log space
 * within the reference's box (optimizer.hpp:966-972, 986-987), minimised by
 * a limited-memory quasi-Newton method (the reference drives
 * LBFGSpp::LBFGSBSolver with m = 8, optimizer.hpp:1074-1087).  Here the
 * method is a projected two-loop L-BFGS with a backtracking line search
 * whose every decision runs on the GPU: one captured graph holds one
 * evaluation (logL + full gradient) and the kernel that takes the optimizer
 * step, so an iteration needs no host round trip.  The same step runs on the
 * host through hy_lbfgs_step, bit for bit. */
#define HY_LBFGS_RUNNING 0
#define HY_LBFGS_CONV_GRAD 1 /* largest projected gradient component <= tol_grad */
#define HY_LBFGS_CONV_F 2    /* |f_prev - f| <= tol_f * max(1, |f|)              */
#define HY_LBFGS_MAX_ITER 3
#define HY_LBFGS_LS_FAIL 4 /* no sufficient decrease after max_back halvings   */
#define HY_LBFGS_NUMERIC 5 /* the start point did not evaluate to finite values */

typedef struct hy_lbfgs_head {
    int32_t n, m;        /* variables; correction pairs (1..8)                    */
    int32_t status;      /* HY_LBFGS_*                                            */
    int32_t iter;        /* accepted steps                                        */
    int32_t n_eval;      /* evaluations consumed                                  */
    int32_t k, head;     /* pairs held; ring position of the next pair            */
    int32_t n_back;      /* backtracks of the current line search                 */
    int32_t max_iter, max_back;
    int32_t pad[2];
    double f, f_prev;    /* objective at x and at the previous accepted point     */
    double alpha, gd;    /* current step; g.d of the current direction            */
    double tol_grad, tol_f, c1;
    double pgnorm;       /* largest projected gradient component at x             */
} hy_lbfgs_head;

/* The map from a log parameter to the parameter the fits evaluate: exp(x)
 * computed in the same operations on the host and on the GPU (within 2 ulp
 * of the C library's exp). */
double hy_lbfgs_exp(double x);
/* Doubles of an optimizer state (x, g, d, xt, gt, lb, ub, m pairs). */
size_t hy_lbfgs_state_doubles(int n, int m);
/* Initialises a state: xt = x0 clamped to [lb, ub] is the first point to
 * evaluate.  c1 = 1e-4, max_back = 20. */
int hy_lbfgs_init(int n, int m, double *state, hy_lbfgs_head *hd, const double *x0, const double *lb,
                  const double *ub, int max_iter, double tol_grad, double tol_f);
/* One optimizer step on the host: consumes f_trial and the gradient
 * g_trial[n] at the trial point (ok = 0: the evaluation failed) and leaves
 * the next trial point in x_next[n].  Any of the output pointers may be
 * NULL. */
int hy_lbfgs_step(double *state, hy_lbfgs_head *hd, double f_trial, const double *g_trial, int ok, double *x_next);
/* Fit on the engine's problem, mixture and time convention: the free
 * variables are log bl of every node with a parent, log omega (n_mat), log
 * theta (5); alpha, comp_w, delta and psi stay fixed.  The objective is
 * -scale * lnL.  captured = 1: after the start point, every evaluation and
 * step is one replay of a captured graph, with the host reading the status
 * every check_every replays; captured = 0: hy_eval and hy_lbfgs_step
 * alternate on the host (same iterates; also what engines that launch
 * eagerly do: hy_create_group members, profiled engines).  bl[num_nodes], omegas[n_mat] and
 * theta[5] are read as the start and overwritten with the final point;
 * out_info[4] = status, iterations, evaluations, replays. */
int hy_fit_lbfgs(hy_engine h, int precision, int captured, double *bl, double alpha, double *omegas,
                 const double *comp_w, double *theta, double delta, double psi, double scale, int m, int max_iter,
                 double tol_grad, double tol_f, int check_every, double *out_logL, int32_t *out_info);

#ifdef __cplusplus
}
#endif
#endif
