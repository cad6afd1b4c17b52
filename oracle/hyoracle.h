/*
 * hyoracle.h -- FP64 CPU restatement of HyPhy 3's MG94xREV codon likelihood
 * core (Felsenstein pruning, inside-outside adjoint, spectral Frechet
 * gradients, branch-site omega mixtures).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * MI355X path and the CPU timing baseline.  Only tests/, the smoke() entry
 * and bench.py's cpu_baseline leg may load it.  The product library never
 * links, loads or calls it.
 *
 * Every function cites the reference file:line whose algorithm it restates
 * (paths relative to the reference repository root).
 */
#ifndef HYORACLE_H
#define HYORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Genetic code with the reference's sense-codon numbering
 * (include/hyphy/core/genetic_code.hpp:127-200): sense codons in
 * lexicographic ACGT order, stops skipped. */
typedef struct {
    int32_t S;                   /* number of sense codons                    */
    int32_t code_id;             /* registry id (genetic_code.hpp:248-380)    */
    int8_t c64_to_sense[64];     /* -1 for stops                              */
    int8_t sense_to_c64[64];
    char aa[64];                 /* amino-acid letter per c64, '*' for stop  */
    int8_t nd[64 * 64];          /* number of nucleotide differences          */
    int8_t dpos[64 * 64 * 3];    /* codon position of the k-th difference    */
    int8_t dfrom[64 * 64 * 3];   /* source nucleotide (A=0,C=1,G=2,T=3)      */
    int8_t dto[64 * 64 * 3];     /* target nucleotide                         */
    uint8_t syn[64 * 64];        /* synonymous pair                           */
} or_code;

/* Tree + compressed alignment in flat arrays. */
typedef struct {
    int32_t num_nodes;
    int32_t root;
    int32_t num_taxa;
    int32_t num_patterns;
    int32_t S;
    int32_t pad0;
    const int32_t *parent;      /* [num_nodes], -1 at the root              */
    const int32_t *child_ptr;   /* [num_nodes+1] CSR offsets                */
    const int32_t *child_idx;   /* children in Newick order                 */
    const int32_t *post_order;  /* [num_nodes]                              */
    const int32_t *leaf_taxon;  /* [num_nodes] taxon row, -1 = no match     */
    const uint8_t *is_leaf;     /* [num_nodes]                              */
    const int8_t *states;       /* [num_patterns * num_taxa]                */
    const uint64_t *masks;      /* [num_patterns * num_taxa]                */
    const double *weights;      /* [num_patterns]                           */
} or_tree;

/* MG94 parameters (rate_matrix.hpp:106-118). theta = AC, AT, CG, CT, GT. */
typedef struct {
    double alpha, beta;
    double theta[5];
    double delta, psi;
} or_mg94_params;

/* Branch-site mixture specification.  Rate matrix m has nonsynonymous rate
 * omegas[m] (alpha fixed by params).  Branch b belongs to class
 * branch_class[b]; class c mixes the components
 * comp_ptr[c] .. comp_ptr[c+1]-1, each a (matrix comp_mat[j], weight comp_w[j])
 * pair: P_b = sum_j comp_w[j] * P_{comp_mat[j]}(tau_b).
 * convention 0: tau_b = bl_b (reference "raw" time, likelihood.hpp:1122-1125,
 * busted.hpp:1191-1216); 1: tau_b = bl_b * 3 / mubar_c with
 * mubar_c = sum_j comp_w[j] * mu_{comp_mat[j]} (likelihood.hpp:1246-1256 for
 * one component). */
typedef struct {
    int32_t n_mat;
    int32_t n_class;
    int32_t convention;
    int32_t pad0;
    const double *omegas;        /* [n_mat]                 */
    const int32_t *branch_class; /* [num_nodes]             */
    const int32_t *comp_ptr;     /* [n_class+1]             */
    const int32_t *comp_mat;     /* [comp_ptr[n_class]]     */
    const double *comp_w;        /* [comp_ptr[n_class]]     */
    /* BUSTED's error sink (busted.hpp:193-206, 571-576): a component with
     * comp_mat[j] == -1 contributes comp_w[j] * P_sink, the time-independent
     * P_sink(i, j) = f_j / max(1e-12, 1 - f_i) (i != j, zero diagonal) of the
     * codon frequencies f = sink_freqs [S]; needs convention 0. */
    const double *sink_freqs;
} or_mixture;

/* Outputs of or_mixture_eval. Arrays may be NULL when not wanted. */
typedef struct {
    double log_likelihood;
    double *site_ll;      /* [num_patterns] per-pattern lnL (unweighted)     */
    double *grad_bl;      /* [num_nodes] d lnL / d bl_b (root entry 0)        */
    double *grad_omega;   /* [n_mat]                                          */
    double *grad_w;       /* [comp_ptr[n_class]]                              */
    double *grad_theta;   /* [5]                                              */
    double *dtau;         /* [num_nodes] d lnL / d tau_b                       */
} or_mixture_out;

int or_code_build(int32_t code_id, or_code *out);

void or_mg94_Q(const or_code *code, const or_mg94_params *p, const double *pos_freqs /*3x4*/,
               const double *pi, double *Q /*SxS*/, double *mu);

int or_sym_eigen(int32_t n, double *A /*nxn, destroyed*/, double *evals, double *evecs /*nxn, column k = vector k*/);

int or_mg94_decompose(int32_t S, const double *Q, const double *pi, double *evals, double *U_pi, double *V_pi);

void or_transition(int32_t S, const double *evals, const double *U_pi, const double *V_pi, double t, double *P);

int or_expm_taylor(int32_t n, const double *A, double t, double *out);

double or_loglik_from_P(const or_tree *tr, const double *P /*num_nodes*S*S*/, const double *pi,
                        int32_t tile, int32_t threads, double *site_ll);

double or_branch_grad_from_P(const or_tree *tr, const double *P, const double *dP, const double *pi,
                             int32_t tile, int32_t threads, double *grad, double *site_ll);

double or_mg94_loglik(const or_tree *tr, const or_code *code, const or_mg94_params *p,
                      const double *pos_freqs, const double *pi, const double *bl, double branch_scale,
                      int32_t threads, double *site_ll);

double or_branch_length_gradients(const or_tree *tr, const or_code *code, const or_mg94_params *p,
                                  const double *pos_freqs, const double *pi, const double *bl,
                                  int32_t threads, double *grad);

double or_joint_frechet(const or_tree *tr, const or_code *code, const or_mg94_params *p,
                        const double *pos_freqs, const double *pi, const double *bl,
                        const uint8_t *is_test, double omega_bg, int32_t threads,
                        double *grad_bl, double *grad_omega, double *grad_omega_bg, double *grad_theta);

int or_mixture_eval(const or_tree *tr, const or_code *code, const or_mg94_params *base,
                    const double *pos_freqs, const double *pi, const or_mixture *mix,
                    const double *bl, int32_t want_grad, int32_t tile, int32_t threads,
                    or_mixture_out *out);

/* BUSTED synonymous-rate classes over the mixture (busted.hpp:509-659,
 * 1138-1429); grad_syn_rates / grad_syn_weights [n_syn] may be NULL. */
int or_srv_eval(const or_tree *tr, const or_code *code, const or_mg94_params *base,
                const double *pos_freqs, const double *pi, const or_mixture *mix, const double *bl,
                int32_t n_syn, const double *syn_rates, const double *syn_weights, int32_t want_grad,
                int32_t threads, or_mixture_out *out, double *grad_syn_rates, double *grad_syn_weights);
/* Reference workspaces (likelihood.hpp:295-340) of one pattern tile. */
int or_workspace(const or_tree *tr, const double *P, const double *pi, int32_t p0, int32_t p1, int32_t outside,
                 double *D, double *V, double *PD, double *U, int32_t *boost, double *logsc, double *L,
                 double *logL);
double or_neumaier_sum(const double *x, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
