/*
 * vmc_oracle.h -- fp64 CPU restatement of ANNonGPU's variational Monte Carlo path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product in
 * annongpu_amd/ and the timed CPU baseline of bench.py.  Only tests/, the smoke()
 * entry point and bench.py's cpu_baseline leg may load it; the product never does.
 *
 * Every routine cites the reference file:line whose behaviour it restates.
 * Complex numbers are C99 `double complex` (layout = two doubles, re then im),
 * so numpy complex128 arrays can be passed straight in.
 */
#ifndef VMC_ORACLE_H
#define VMC_ORACLE_H

#include <stdint.h>
#include <complex.h>

#define ORC_MAX_WORDS 4     /* spins: up to 256 sites  */
#define ORC_MAX_LAYERS 4    /* PsiDeep hidden layers     */
#define ORC_MAX_CNN_LAYERS 3

enum { ORC_RBM = 0, ORC_DEEP = 1, ORC_CNN = 2 };
enum { ORC_MOVE_FLIP = 0, ORC_MOVE_EXCHANGE = 1 };

typedef struct {
    int32_t kind;
    int32_t N;                 /* input units (= sites for the spin basis) */
    int32_t num_params;
    int32_t words;             /* 64-bit words per configuration           */
    double complex log_prefactor;
    /* PsiRBM (include/quantum_state/PsiRBM.hpp:193-218) */
    int32_t M;
    int32_t pad0;
    const double complex* W;   /* [N*M], row = visible unit                */
    double complex final_weight;
    /* PsiDeep (include/quantum_state/PsiDeep.hpp:70-135) */
    int32_t num_layers;        /* hidden layers                            */
    int32_t size[ORC_MAX_LAYERS];
    int32_t conn[ORC_MAX_LAYERS];              /* lhs connectivity         */
    int32_t pad1;
    const uint32_t* connections[ORC_MAX_LAYERS];   /* [conn*size]          */
    const double complex* weights[ORC_MAX_LAYERS]; /* [conn*size]          */
    const double complex* biases[ORC_MAX_LAYERS];  /* [size]               */
    const double complex* input_weights;           /* [N]                  */
    const double complex* final_weights;           /* [size[last]]         */
    /* PsiCNN, dim = 3 (include/quantum_state/PsiCNN.hpp:32-87) */
    int32_t extent[3];
    int32_t cnn_layers;
    int32_t channels[ORC_MAX_CNN_LAYERS];
    int32_t connectivity[ORC_MAX_CNN_LAYERS][3];
    int32_t num_sym;
    const uint32_t* symmetry_classes;          /* [N]                      */
    const double complex* params;              /* [num_params]             */
    double final_factor;
} orc_psi;

typedef struct {
    int32_t num_strings;
    int32_t words;
    const uint64_t* a;          /* [num_strings*words] */
    const uint64_t* b;          /* [num_strings*words] */
    const double complex* coefficients;
} orc_operator;

#ifdef __cplusplus
extern "C" {
#endif

/* RNG: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* activations: include/quantum_state/psi_functions.hpp:40-48, 102-116 */
double complex orc_logcosh(double complex z, uint32_t layer);
double complex orc_tanh(double complex z, uint32_t layer);

/* basis + Pauli strings: include/basis/PauliString.hpp:193-208,242-254 */
void orc_pauli_apply(int32_t words, const uint64_t* a, const uint64_t* b,
                     const uint64_t* conf, uint64_t* conf_out, double complex* factor);

/* wave function */
double complex orc_log_psi_s(const orc_psi* psi, const uint64_t* conf);
void orc_psi_O_k(const orc_psi* psi, const uint64_t* conf, double complex* O_k);

/* Operator::local_energy: include/operator/Operator.hpp:36-121 */
double complex orc_local_energy(const orc_operator* op, const orc_psi* psi,
                                const uint64_t* conf, double complex log_psi);

/* ExactSummation enumeration (include/ensembles/ExactSummation.hpp:33-72) */
void orc_exact_log_psi_vector(const orc_psi* psi, int32_t num_sites, double complex* out);

/* accumulate E, E2, <O>, F, S (source/network_functions/TDVP.cu.template:77-126,181-302)
 * over a list of configurations with weights.  S may be NULL. */
void orc_tdvp_accumulate(const orc_operator* op, const orc_psi* psi, int64_t num_samples,
                         const uint64_t* confs, const double* weights,
                         double complex* E, double* E2, double complex* O_mean,
                         double complex* F, double complex* S,
                         double complex* E_loc_out, double complex* O_samples_out,
                         int32_t num_threads);

/* Monte Carlo sampler: include/ensembles/MonteCarlo.hpp:55-177 with a Philox stream. */
typedef struct {
    int32_t num_chains;
    int32_t steps_per_chain;      /* samples per chain                      */
    int32_t num_sweeps;
    int32_t num_thermalization_sweeps;
    int32_t move;                 /* ORC_MOVE_FLIP or ORC_MOVE_EXCHANGE     */
    int32_t total_z;              /* sector for exchange moves (2*S_z)      */
    uint64_t seed;
    uint64_t counter;             /* proposals already drawn by each chain  */
    uint32_t call_index;          /* re-initialisation stream index         */
    int32_t init_random;          /* 1 = draw a fresh configuration         */
} orc_mc_params;

void orc_mc_run(const orc_psi* psi, const orc_mc_params* p, uint64_t* chain_confs,
                uint64_t* sample_confs, double complex* sample_log_psi,
                uint64_t* acceptances, uint64_t* rejections, int32_t num_threads);

/* dense SR solve (S + shift*I) x = F in fp64 */
int orc_cholesky_solve(int32_t n, const double complex* S, double shift,
                       const double complex* F, double complex* x);
int orc_cg_solve(int32_t n, const double complex* S, double shift, const double complex* F,
                 double complex* x, double rtol, int32_t max_iter, int32_t* iterations);
/* (S + shift) x = F by Jacobi CG on the matrix-free S of stored O_k samples [ns][P] */
int orc_sr_cg_matfree(int64_t ns, int32_t P, const double complex* O_samples, const double* w,
                      const double complex* O_mean, const double complex* F, double shift, double rtol,
                      int32_t max_iter, double complex* x, double* res, int32_t num_threads);

/* learned block proposal (proposal_oracle.c; no counterpart in the reference) */
void orc_mlp_logits(int32_t N, int32_t H, const double* params, const uint64_t* conf,
                    const int32_t* block, int32_t bs, double* logits);
double orc_mlp_loss_grad(int32_t N, int32_t H, const double* params, int64_t count,
                         const uint64_t* confs, int32_t words, const int32_t* blocks, int32_t bs,
                         double* grad);
void orc_adam_step(int64_t n, double* params, double* m, double* v, const double* grad, double lr,
                   double beta1, double beta2, double eps, int64_t t);
double orc_block_log_q(const double* block_logits, int32_t bs, uint32_t p_new, int32_t k);

#ifdef __cplusplus
}
#endif
#endif
