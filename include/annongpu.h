/*
 * annongpu.h -- C ABI of the MI355X (gfx950) variational Monte Carlo library.
 *
 * Plain pointers and sizes only.  Complex arrays are interleaved (re, im) fp64
 * pairs, i.e. numpy complex128 / C99 double complex / std::complex<double>.
 * Spin configurations are little-endian uint64 words, bit i set <=> s_i = +1
 * (reference: include/basis/Spins.h:104-118).  Every function returns 0 on
 * success and a non-zero ANN_E_* code on failure; ann_last_error() then holds a
 * thread-local message.  Results are on the host when the call that returns
 * them returns; ann_tdvp_eval and the sampling calls may return before their
 * device work ends (a later call that reads a result waits for it, and reports
 * a device error of that work).  Objects are not thread-safe; use one per thread.
 *
 * Each entry point names the reference interface it replaces.  The reference
 * binds those through pybind11 in pyANNonGPU/main.cpp.template (cited as MAIN).
 */
#ifndef ANNONGPU_H
#define ANNONGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ANN_OK 0
#define ANN_E_INVALID 1      /* bad argument / capacity exceeded            */
#define ANN_E_DEVICE 2       /* HIP runtime or kernel launch failure        */
#define ANN_E_UNSUPPORTED 3  /* combination not implemented                 */
#define ANN_E_NUMERIC 4      /* e.g. S + shift not positive definite        */

typedef struct ann_psi_s* ann_psi;
typedef struct ann_operator_s* ann_operator;
typedef struct ann_ensemble_s* ann_ensemble;
typedef struct ann_tdvp_s* ann_tdvp;
typedef struct ann_proposal_s* ann_proposal;
typedef struct ann_kl_s* ann_kl;

/* precision of the device path: FP64 everywhere, or FP32 sweeps / E_loc / O_k
 * with the angle GEMM on MFMA (bf16 x3 split of W, fp32 accumulation).  fp64 is
 * kept for every reduction, the SR solve and the returned results. */
#define ANN_FP64 0
#define ANN_FP32 1

#define ANN_MOVE_FLIP 0       /* single spin flip  (policies/Update_Policy.hpp:19-33) */
#define ANN_MOVE_EXCHANGE 1   /* swap of an anti-parallel pair, conserves S_z          */
#define ANN_MOVE_PAULI 2      /* a site set to a random Pauli type (Pauli-string ensembles) */

/* ---------------------------------------------------------------- runtime */
const char* ann_last_error(void);
/* setDevice (source/ANNonGPU.cu:7-9, MAIN:538) */
int ann_set_device(int device);
/* start_profiling / stop_profiling (source/ANNonGPU.cu:11-17, MAIN:539-540):
 * switch the library's per-kernel hipEvent timers on and off. */
int ann_profiler_start(void);
int ann_profiler_stop(void);
/* accumulated device time of one timed kernel class since ann_profiler_start, and
 * its number of launches; classes inside the CG loop are timed on every 16th launch
 * and the time is scaled to all launches */
int ann_profiler_query(const char* name, double* milliseconds, int64_t* launches);
int ann_profiler_reset(void);

/* -------------------------------------------------------- wave functions */
/* PsiRBM(W[N][M], final_weight, log_prefactor, gpu) -- PsiRBM.hpp:225-243, MAIN:120-154 */
int ann_psi_create_rbm(ann_psi* out, int32_t N, int32_t M, const double* W, const double* final_weight,
                       const double* log_prefactor);
/* PsiDeep(num_sites, a[N], b_list, connections_list, W_list, final_weights,
 * log_prefactor, gpu) -- PsiDeep.hpp:504-578, MAIN:71-114.  Layer l has
 * sizes[l] units and lhs connectivity conns[l]; connections[l] and weights[l]
 * are [conns[l]][sizes[l]] row-major, biases[l] is [sizes[l]]. */
int ann_psi_create_deep(ann_psi* out, int32_t num_sites, int32_t N, const double* a, int32_t num_layers,
                        const int32_t* sizes, const int32_t* conns, const uint32_t* const* connections,
                        const double* const* weights, const double* const* biases, const double* final_weights,
                        const double* log_prefactor);
/* PsiCNN(extent[3], num_channels_list, connectivity_list[L][3], symmetry_classes[N],
 * params, final_factor, log_prefactor, gpu) -- PsiCNN.hpp:308-340, MAIN:160-204 */
int ann_psi_create_cnn(ann_psi* out, const int32_t* extent, int32_t num_layers, const int32_t* channels,
                       const int32_t* connectivity, const uint32_t* symmetry_classes, int32_t num_params,
                       const double* params, double final_factor, const double* log_prefactor);
/* PsiClassicalFP_1/_2, PsiClassicalANN_1/_2 (num_sites, H_local, params, psi_ref, log_prefactor,
 * gpu) -- quantum_state/PsiClassical.hpp:48-160,204-225, source/quantum_state/PsiClassical.cu:9-61,
 * MAIN:208-270.  H_local[n] = ops[n] (copied), params[num_ops] interleaved complex;
 * log psi(s) = log_prefactor + sum_n params[n] L_n(s), L_n = the matrix-element sum of
 * H_local[n] (Operator.hpp:123-135).  ref (copied): a PsiCNN (the ANN form,
 * PsiClassical.hpp:245-256; spin configurations only), a PsiDeep (also on Pauli strings), or
 * null for PsiFullyPolarized (log psi_ref = 0, no parameters, PsiFullyPolarized.hpp:31-40).  With
 * order 2 and a network reference log psi_ref(s) enters log psi and the reference's
 * parameters follow the classical ones in params and O_k. */
int ann_psi_create_classical(ann_psi* out, int32_t num_sites, int32_t num_ops, const ann_operator* ops,
                             const double* params, ann_psi ref, int32_t order, const double* log_prefactor,
                             const double* ref_log_prefactor);
/* order (MAIN:242), len(H_local) (MAIN:221), and whether a network reference is held */
int ann_psi_classical_info(ann_psi psi, int32_t* order, int32_t* num_ops, int32_t* has_network_ref);
/* psi_ref (MAIN:236-239): a copy of the network reference (ANN_E_INVALID for the fully
 * polarised one) */
int ann_psi_classical_ref(ann_psi psi, ann_psi* out);
/* ref_log_prefactor property (MAIN:226-234) */
int ann_psi_get_ref_log_prefactor(ann_psi psi, double* value2);
int ann_psi_set_ref_log_prefactor(ann_psi psi, const double* value2);
/* copy() (MAIN:82,127,171) */
int ann_psi_copy(ann_psi src, ann_psi* out);
int ann_psi_destroy(ann_psi psi);
/* num_params / N / num_sites (MAIN:92-93,135-137,173-174) */
int ann_psi_info(ann_psi psi, int32_t* num_params, int32_t* N, int32_t* num_sites);
/* params property (MAIN:94-100,138-144,191-197); order of PsiDeep.cu:245-270 */
int ann_psi_get_params(ann_psi psi, double* params);
int ann_psi_set_params(ann_psi psi, const double* params);
/* log_prefactor property (MAIN:84-88) */
int ann_psi_get_log_prefactor(ann_psi psi, double* value2);
int ann_psi_set_log_prefactor(ann_psi psi, const double* value2);
/* additive: device precision (ANN_FP64 default) */
int ann_psi_set_precision(ann_psi psi, int32_t precision);

/* --------------------------------------------------------------- operator */
/* Operator(PauliExpression, gpu) -- operator/Operator.cpp:18-39, MAIN:321-333.
 * String n is (a[n][words], b[n][words]) in the PauliString encoding
 * (basis/PauliString.hpp:37-56: X=(1,0), Y=(0,1), Z=(1,1)). */
int ann_operator_create(ann_operator* out, int32_t num_strings, int32_t words, const uint64_t* a, const uint64_t* b,
                        const double* coefficients);
int ann_operator_destroy(ann_operator op);

/* -------------------------------------------------------------- ensembles */
/* MonteCarloSpins(num_samples, num_sweeps, num_thermalization_sweeps,
 * num_markov_chains, gpu) -- ensembles/MonteCarlo.hpp:249-259, MAIN:358-366.
 * Additive: move kind, S_z sector (2 S_z) for exchange moves, Philox seed. */
int ann_monte_carlo_create(ann_ensemble* out, uint32_t num_samples, uint32_t num_sweeps,
                           uint32_t num_thermalization_sweeps, uint32_t num_markov_chains, int32_t move,
                           int32_t total_z, uint64_t seed);
/* additive: keep chains between calls instead of re-initialising them
 * (the reference re-initialises and re-thermalises on every call). */
int ann_monte_carlo_set_persistent(ann_ensemble ens, int32_t persistent);
/* ExactSummationSpins(num_sites, gpu) -- source/ensembles/ExactSummation.cu:16-25, MAIN:395-399 */
int ann_exact_summation_create(ann_ensemble* out, uint32_t num_sites);
int ann_ensemble_destroy(ann_ensemble ens);
/* num_steps (MAIN:363,399) */
int ann_ensemble_num_steps(ann_ensemble ens, uint64_t* num_steps);
/* acceptance_rate (MAIN:364-366): raw counters of the last call */
int ann_ensemble_acceptance(ann_ensemble ens, uint64_t* accepted, uint64_t* rejected);
/* additive: block moves of the last sampling call with a learned proposal: proposals
 * made, redraws of the current pattern (ann_ensemble_acceptance counts them with the
 * rejected proposals: they do not move the chain) and flipped sites whose rows
 * entered a psi ratio */
int ann_ensemble_block_stats(ann_ensemble ens, uint64_t* proposals, uint64_t* null_moves, uint64_t* rows_evaluated);
/* additive: configurations of the last sampling call, [num_steps][words] */
int ann_ensemble_last_samples(ann_ensemble ens, uint64_t* confs, int64_t capacity, int64_t* count, int32_t* words);
/* additive (no counterpart in the reference): an ensemble of caller-given configurations
 * confs[count][words] (bit i of a configuration set <=> s_i = +1, words <= 4) with weights
 * (null: 1 / count each), evaluated as they are; it lets two precisions or two builds
 * see the same samples, the role the reference's ExactSummation plays at N <= 20 */
int ann_sample_list_create(ann_ensemble* out, uint32_t num_sites, const uint64_t* confs, int64_t count,
                           int32_t words, const double* weights);

/* ExactSummationPaulis(num_sites, gpu) -- ensembles/ExactSummation.hpp:125, MAIN:402-406:
 * the 4^N Pauli strings in PauliString::enumerate order (basis/PauliString.hpp:30-35) */
int ann_exact_summation_paulis_create(ann_ensemble* out, uint32_t num_sites);
/* MonteCarloPaulis(num_samples, num_sweeps, num_thermalization_sweeps, num_markov_chains,
 * gpu) -- ensembles/MonteCarlo.hpp:265-283, MAIN:369-380: Metropolis moves that set a
 * random site to a random Pauli type (policies/Update_Policy.hpp:41-56).  Additive: seed. */
int ann_monte_carlo_paulis_create(ann_ensemble* out, uint32_t num_samples, uint32_t num_sweeps,
                                  uint32_t num_thermalization_sweeps, uint32_t num_markov_chains, uint64_t seed);
/* additive: caller-given Pauli strings ab[count][2] = (a, b) over num_sites <= 64 sites,
 * with weights (null: 1 / count each) */
int ann_pauli_sample_list_create(ann_ensemble* out, uint32_t num_sites, const uint64_t* ab, int64_t count,
                                 const double* weights);

/* ------------------------------------------------------ network functions */
/* log_psi_s(psi, spins) -- PsiVector.cu.template:101-133, MAIN:495-497 */
int ann_log_psi_s(ann_psi psi, const uint64_t* conf, double* out2);
/* psi_O_k(psi, spins) -- PsiOkVector.cu.template:40-74, MAIN:498-500 */
int ann_psi_O_k(ann_psi psi, const uint64_t* conf, double* out);
/* log_psi_s / psi_O_k of a Pauli string ab = (a, b) (MAIN:495-500 on the PauliString basis) */
int ann_log_psi_s_pauli(ann_psi psi, const uint64_t* ab, double* out2);
int ann_psi_O_k_pauli(ann_psi psi, const uint64_t* ab, double* out);
/* psi_O_k_vector(psi, exact) = sum_s O_k(s), unweighted -- PsiOkVector.cu.template:8-37 */
int ann_psi_O_k_vector(ann_psi psi, ann_ensemble ens, double* out);
/* log_psi(psi, ens) = sum_s w_s log psi(s) -- PsiVector.cu.template:73-99 */
int ann_log_psi(ann_psi psi, ann_ensemble ens, double* out2);
/* psi_vector / log_psi_vector -- PsiVector.cu.template:14-71 */
int ann_psi_vector(ann_psi psi, ann_ensemble ens, double* out);
int ann_log_psi_vector(ann_psi psi, ann_ensemble ens, double* out);
/* psi.norm(exact) = sqrt(sum_s |psi(s)|^2) -- PsiNorm.cu.template:13-42 */
int ann_psi_norm(ann_psi psi, ann_ensemble ens, double* out);
/* apply_operator = psi(s) E_loc(s) -- ApplyOperator.cu.template:9-44 */
int ann_apply_operator(ann_psi psi, ann_operator op, ann_ensemble ens, double* out);
/* activation_function(x, layer) -- MAIN:534-536 */
int ann_activation_function(const double* x2, int32_t layer, double* out2);

/* ExpectationValue(gpu): __call__, fluctuation, gradient --
 * ExpectationValue.cu.template:19-50,175-275, MAIN:420-433 */
int ann_expectation_value(ann_operator op, ann_psi psi, ann_ensemble ens, double* out2);
int ann_expectation_fluctuation(ann_operator op, ann_psi psi, ann_ensemble ens, double* sigma, double* value2);
int ann_expectation_gradient(ann_operator op, ann_psi psi, ann_ensemble ens, double* gradient, double* value2);
/* ExpectationValue.exp_sigma_z(op, psi, ensemble) -- ExpectationValue.hpp:27,
 * ExpectationValue.cu.template:52-77, MAIN:428: the weighted sum over the samples of psi of
 * exp(fast_local_energy(s)) (Operator.hpp:123-135: the matrix-element sum of op's strings
 * at s), spin or Pauli-string samples; value2 = complex result */
int ann_expectation_exp_sigma_z(ann_operator op, ann_psi psi, ann_ensemble ens, double* value2);
/* ExpectationValue(op, psi, psi_sampling, ensemble) -- ExpectationValue.hpp:32-33,
 * ExpectationValue.cu.template:130-172, MAIN:431: <op> in psi from the samples of
 * psi_sampling, each weighted by |psi(s) / psi_sampling(s)|^2, divided by the sum of those
 * weights.  (The reference divides by an accumulator its kernel never adds to,
 * ExpectationValue.cu.template:135,155-171; this is the self-normalised estimate it names.) */
int ann_expectation_value_reweighted(ann_operator op, ann_psi psi, ann_psi psi_sampling, ann_ensemble ens,
                                     double* value2);
/* additive: per-sample local energies of the last sampling, [num_steps] */
int ann_local_energies(ann_operator op, ann_psi psi, ann_ensemble ens, double* e_loc, int64_t capacity);
/* No counterpart in the reference: for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  E_loc of `count` configurations given
 * in DEVICE memory, confs[count][words], into DEVICE memory e_loc[count] (interleaved
 * complex), for a PsiRBM / one-layer PsiDeep.  The angles live in a static 16 MiB device
 * array: count * padded hidden units * 8 B (fp32 mode; 16 B in fp64 mode) must fit it, else
 * ANN_E_INVALID.  The net's parameters and the operator's device tables must be on the
 * device already (one ordinary call with each puts them there), else ANN_E_UNSUPPORTED.
 * Synchronous on the library stream.  Calls that overlap on one GPU share that array
 * and race. */
int ann_local_energies_noalloc(ann_operator op, ann_psi psi, const uint64_t* confs, int64_t count, double* e_loc);

/* ------------------------------------------------------------------- TDVP */
/* HilbertSpaceDistance(num_params, gpu) __call__ / gradient (MAIN:436-441;
 * HilbertSpaceDistance.cu.template:113-165): the distance between psi_prime and U psi
 * (U = op, unitary or, with is_unitary = 0, the generator exp(E_loc)) on psi's samples;
 * gradient (psi_prime's num_params, interleaved complex; may be null) with respect to
 * psi_prime's parameters, divided by distance^nu */
int ann_hilbert_space_distance(ann_psi psi, ann_psi psi_prime, ann_operator op, int32_t is_unitary, ann_ensemble ens,
                               double nu, double* gradient, double* distance);

/* KullbackLeibler(num_params, gpu) -- network_functions/KullbackLeibler.hpp:14-117,
 * KullbackLeibler.cu.template:126-162, MAIN:445-462 */
int ann_kl_create(ann_kl* out, uint32_t num_params);
int ann_kl_destroy(ann_kl kl);
/* __call__ (mode 0), gradient (mode 1), gradient_with_noise (mode 2) --
 * KullbackLeibler.cu.template:171-323: psi the classical network, psi_prime the network
 * whose ensemble is sampled; the deviation log psi_prime - log_psi_scale log psi is compared
 * with the mean deviation of the previous call and counts where its square exceeds
 * threshold^2.  gradient (num_params interleaved complex, may be null) is with respect to
 * psi_prime's parameters, divided by value^nu; noise (num_params, mode 2) may be null. */
int ann_kl_eval(ann_kl kl, ann_psi psi, ann_psi psi_prime, ann_ensemble ens, double threshold, double nu,
                int32_t mode, double* gradient, double* noise, double* value);
/* total_weight and mean_deviation of the last call (MAIN:459-460) */
int ann_kl_info(ann_kl kl, double* total_weight, double* mean_deviation2);
/* log_psi_scale property (MAIN:461) */
int ann_kl_get_log_psi_scale(ann_kl kl, double* scale);
int ann_kl_set_log_psi_scale(ann_kl kl, double scale);

/* TDVP(num_params, gpu) -- network_functions/TDVP.hpp:16-101, MAIN:465-489 */
int ann_tdvp_create(ann_tdvp* out, uint32_t num_params);
int ann_tdvp_destroy(ann_tdvp tdvp);
/* eval (with_S = 1) and eval_F (with_S = 0) -- TDVP.cu.template:181-334 */
int ann_tdvp_eval(ann_tdvp tdvp, ann_operator op, ann_psi psi, ann_ensemble ens, int32_t with_S);
/* S_matrix [P][P], F_vector, O_k_vector, E_local, E2_local, var_H (MAIN:467-474) */
int ann_tdvp_get_S(ann_tdvp tdvp, double* S);
int ann_tdvp_get_F(ann_tdvp tdvp, double* F);
int ann_tdvp_get_O(ann_tdvp tdvp, double* O);
int ann_tdvp_get_energy(ann_tdvp tdvp, double* E2_complex, double* E2_local, double* var_H);
/* E_local_samples, total_weight (MAIN:470-474; TDVP.hpp:29-34) and the samples of the last
 * evaluation: confs[count][words], weights[count], e_local[count] (interleaved complex);
 * any output may be null, capacity is the number of samples the buffers hold */
int ann_tdvp_get_samples(ann_tdvp tdvp, uint64_t* confs, double* weights, double* e_local, int64_t capacity,
                         int64_t* count, int32_t* words, double* total_weight);
/* O_k_samples (MAIN:472; TDVP.cu.template:69,121: uncentred O_k of every sample),
 * out[count][P] interleaved complex; materialised on demand from the evaluation's state */
int ann_tdvp_get_O_samples(ann_tdvp tdvp, double* out, int64_t capacity);
/* S_dot_vector(vec, ensemble) -- TDVP.cu.template:336-443, MAIN:483-487 */
int ann_tdvp_S_dot_vector(ann_tdvp tdvp, const double* v, double* out);
/* additive (absent from the reference): solve (S + shift I) x = F.
 * method 0 = fp64 Cholesky of the dense S (needs eval with_S = 1),
 * method 1 = conjugate gradients on the matrix-free S (fp64 vectors and dots,
 * Jacobi preconditioner) started from zero, method 2 = the same started from
 * the x passed in (warm start from the previous step's solution). */
#define ANN_SOLVE_CHOLESKY 0
#define ANN_SOLVE_CG 1
#define ANN_SOLVE_CG_WARM 2
/* CG preconditioned by Kronecker factors of S (spin covariance x covariance of the
 * hidden-unit factors g, blocks of 64 units) on the W parameters and Jacobi on the
 * biases; PsiRBM and one-layer PsiDeep of up to 128 sites.  No counterpart in the
 * reference, whose solve is a dense numpy call: an alternative solver for the same
 * (S + shift) x = F.  The _WARM variant starts from the x passed in. */
#define ANN_SOLVE_CG_KFAC 3
#define ANN_SOLVE_CG_KFAC_WARM 4
int ann_tdvp_solve(ann_tdvp tdvp, double shift, int32_t method, double rtol, int32_t max_iter, double* x,
                   int32_t* iterations, double* residual);

/* ------------------------------------------------------- learned proposal */
/* Additive (absent from the reference): a two-layer tanh network
 * x -> tanh(W1 x + b1) -> W2 h + b2 (N -> hidden -> N) that proposes
 * S_z-preserving moves of whole blocks of block_size <= 8 sites.  block_sites is
 * [num_blocks][block_size].  A move picks a block, masks its spins in the input
 * and redraws them among the patterns with the same number of up spins with
 * probability proportional to exp(sum of the logits of the up sites); the
 * Metropolis-Hastings test uses |psi'/psi|^2 times the proposal ratio, so the
 * sampled distribution stays |psi|^2.  Parameters are fp32, packed
 * [W1 (hidden x N) | b1 | W2 (N x hidden) | b2]; hidden <= 256. */
int ann_proposal_create(ann_proposal* out, int32_t num_sites, int32_t hidden, int32_t num_blocks,
                        int32_t block_size, const int32_t* block_sites, uint64_t seed);
int ann_proposal_destroy(ann_proposal p);
int ann_proposal_num_params(ann_proposal p, int64_t* n);
int ann_proposal_get_params(ann_proposal p, float* params);
int ann_proposal_set_params(ann_proposal p, const float* params);
/* loss (mean over rows of sum over the row's block of softplus(l) - y l) and its
 * gradient for explicit rows: confs [count][words], sites [count][block_size] */
int ann_proposal_loss_grad(ann_proposal p, const uint64_t* confs, int64_t count, int32_t words,
                           const int32_t* sites, double* loss, float* grad);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the given gradient */
int ann_proposal_adam_step(ann_proposal p, const float* grad, double lr);
/* Adam training on random (sample, block) rows of the ensemble's last samples
 * (device resident); loss = loss of the last batch */
int ann_proposal_train(ann_proposal p, ann_ensemble ens, int32_t steps, int32_t batch, double lr, double* loss);
/* attach (p != NULL) or detach a proposal.  With one attached, the interval
 * between samples is num_sweeps block sweeps (ceil(N / block_size) block moves
 * each) plus plain_sweeps single-flip/exchange sweeps.  The ensemble does not
 * own p: keep it alive while attached. */
int ann_monte_carlo_set_proposal(ann_ensemble ens, ann_proposal p, uint32_t plain_sweeps);
/* additive: with exclude_current != 0 a block move draws its pattern from the network's
 * conditional law without the current pattern, and the Metropolis-Hastings ratio carries
 * the matching factor (Z - q(old)) / (Z - q(new)) (a Metropolized Gibbs move); no move
 * redraws the current configuration */
int ann_monte_carlo_set_proposal_mode(ann_ensemble ens, int32_t exclude_current);

#ifdef __cplusplus
}
#endif
#endif
