// cnn.hpp -- device views and launchers of the wave functions without a structured
// O: PsiCNN and multi-layer PsiDeep (cnn.hip), and the materialised-O kernels they
// share (generic_ok.hip).
#pragma once

#include "kernels.hpp"

namespace ann {

constexpr int kMaxCnnLayers = 3;

// PsiCNN in dim 3 (include/quantum_state/PsiCNN.hpp:32-87).  Parameter k of
// link (layer l, input channel ci, output channel cj), symmetry class c and
// kernel offset o sits at link_off[l] + ((ci * ch[l] + cj) * nsym + c) * vol[l] + o
// (pyANNonGPU/PsiCNN.py:67-83); offsets o run over the kernel box with the
// last extent fastest.  nbr[nbr_off[l] + x * vol[l] + o] is the periodic
// neighbour of site x at offset o (include/quantum_state/detail/Convolve.hpp:122-148).
template <typename T>
struct CnnView {
    using real_t = T;
    int N;
    int L;
    int nsym;
    int num_params;
    int max_ch;
    int ch[kMaxCnnLayers];
    int vol[kMaxCnnLayers];
    int link_off[kMaxCnnLayers];
    int nbr_off[kMaxCnnLayers];
    int zoff[kMaxCnnLayers + 1];   // offsets of each layer's pre-activations, zoff[L] = total
    T final_factor;
    const cplx<T>* params;
    const int* nbr;
    const uint32_t* sym;
    // cons[(l * N + y) * kMaxWords + k]: word k of the mask of the sites x whose layer-l
    // pre-activations read site y (the consumers of y, for the local updates)
    const uint64_t* cons;
};

constexpr int kMaxMlpLayers = 4;

// PsiDeep with more than one hidden layer (include/quantum_state/PsiDeep.hpp:171-215,
// 345-445).  Hidden layer l (0-based) has size[l] units, each fed by conn[l] units
// of the layer below (the spins for l = 0): pre-activation
//   z_l[j] = b_l[j] + sum_i W_l[i][j] in_l[cn_l[i][j]],  in_0 = s,  in_l = act_{l-1}(z_{l-1}),
// log psi = log_prefactor + sum_j f_j act_{L-1}(z_{L-1}[j]), act_0 = log-cosh series,
// act_{l >= 1} = the layer >= 1 series.  Parameters: [a (N) | per layer b_l, W_l
// ([i][j])]; b_l of layer l starts at poff[l].  rb/ri (layers l >= 1) list, for each
// unit i of layer l - 1, the weights i' size[l] + j of layer l that read it (the
// back-propagation gathers them in a fixed order).
template <typename T>
struct MlpView {
    using real_t = T;
    int N, L, num_params;
    int size[kMaxMlpLayers], conn[kMaxMlpLayers];
    int zoff[kMaxMlpLayers + 1];
    int poff[kMaxMlpLayers];
    const uint32_t* cn[kMaxMlpLayers];
    const cplx<T>* W[kMaxMlpLayers];
    const cplx<T>* b[kMaxMlpLayers];
    const cplx<T>* f;
    const int* rb[kMaxMlpLayers];
    const int* ri[kMaxMlpLayers];
};

// A PsiDeep reading Pauli-string configurations (w[0] = a, w[1] = b over N sites)
// through three input units per site: unit 3i + t - 1 is +1 when site i holds Pauli
// type t (1 = X, 2 = Y, 3 = Z) and -1 otherwise (PauliString::network_unit_at,
// include/basis/PauliString.hpp:84-90; PsiDeep.hpp:283-305).
template <typename V>
struct PauliInput {
    using real_t = typename V::real_t;
    V net;             // the network over 3 N input units
    int N;             // Pauli sites
    int num_params;
};

// PsiClassical (include/quantum_state/PsiClassical.hpp:48-160):
//   log psi(s) = sum_n params[n] L_n(s) [+ log psi_ref(s) when the reference enters],
//   L_n(s) = sum over the strings m of H_local[n] of c_m <s|P_m|s'> (Operator::fast_local_energy,
//   operator/Operator.hpp:123-135), the matrix-element coefficient of PauliString::apply on
//   spins (PauliString.hpp:242-254) or on Pauli strings (:257-276).
// O_k = L_k(s) for k < num_ops, then the reference's O_k.  The fully polarised reference
// (PsiFullyPolarized.hpp:31-40) has log psi_ref = 0 and no parameters: has_ref = 0.
// R: the reference network's view (MlpView<T> for a PsiDeep, CnnView<T> for a PsiCNN)
template <typename T, typename R = MlpView<T>>
struct ClassicalView {
    using real_t = T;
    int N;               // sites
    int basis;           // Basis
    int num_ops;
    int num_params;      // num_ops + the reference's parameters when has_ref
    int has_ref;         // a network reference enters log psi (order 2)
    const int* op_begin;     // [num_ops + 1] string ranges
    const uint64_t* sa;      // [strings][kMaxWords]
    const uint64_t* sb;      // [strings][kMaxWords]
    const double* coef;      // [strings][2]: c_m, times (-i)^{#Y of P_m} for the spin basis
    const cplx<T>* params;   // [num_ops]
    R ref;                   // over N spins (a PsiDeep also over the 3 N units of a Pauli string)
    cplx<T> ref_lp;          // the reference's log_prefactor
};

// Kernels of the wave functions without a structured O (V = CnnView<T> or MlpView<T>):
// one 64-lane wavefront per configuration, a full forward pass per configuration,
// the layer state in LDS.
template <typename V>
void launch_generic_log_psi(const V& net, const uint64_t* confs, int words, int64_t count,
                            cplx<double> log_prefactor, cplx<double>* log_psi, double* weight, hipStream_t stream);
template <typename V>
void launch_generic_mc(const V& net, uint64_t* chain_confs, int words, int num_chains, int move,
                       int64_t num_proposals, uint64_t counter, uint64_t seed, uint32_t* accepted,
                       hipStream_t stream);
template <typename V>
void launch_generic_local_energy(const V& net, const OpView& op, const uint64_t* confs, int64_t count,
                                 cplx<double>* e_loc, hipStream_t stream);
// O[s][k] = d log psi(s) / d param_k - center[k] (center may be null)
template <typename V>
void launch_generic_ok(const V& net, const uint64_t* confs, int words, int64_t count, const cplx<double>* center,
                       cplx<typename V::real_t>* O, hipStream_t stream);

// E_loc on Pauli-string configurations (confs [count][2]) for the views that read them
// (PauliInput<MlpView<T>>, ClassicalView<T> with basis BASIS_PAULIS):
// E_loc(x) = sum_n c_n phase_n(x) psi(x_n') / psi(x), the identity string adding c_n
// (operator/Operator.hpp:36-84 with PauliString::apply on Pauli strings)
template <typename V>
void launch_generic_local_energy_pauli(const V& net, const PauliOpView& op, const uint64_t* confs, int64_t count,
                                       cplx<double>* e_loc, hipStream_t stream);

// out[s] = exp(L(s)), L = the matrix-element sum of the operator's strings at
// configuration s (Operator::fast_local_energy, operator/Operator.hpp:123-135): on spins
// from the grouped strings, on Pauli strings (confs [count][2]) from the string phases
void launch_exp_fast_energy(const OpView& op, const uint64_t* confs, int64_t count, cplx<double>* out,
                            hipStream_t stream);
void launch_exp_fast_energy_pauli(const PauliOpView& op, const uint64_t* confs, int64_t count, cplx<double>* out,
                                  hipStream_t stream);
// out[s] = w[s] exp(2 (Re lp[s] - Re lp_sampling[s]))
void launch_reweight(const double* w, const cplx<double>* lp, const cplx<double>* lp_sampling, int64_t count,
                     double* out, hipStream_t stream);

// ------------------------------------------------------- materialised O [n][P]
// u[s] = sum_k O[s][k] v[k]   (fp64 accumulation)
template <typename T>
void launch_ok_rowdot(const cplx<T>* O, int64_t n, int P, const cplx<double>* v, cplx<double>* u,
                      hipStream_t stream);
// out[k] = sum_s coef[s] * op(O[s][k]) + shift * v[k]; op = identity (mode 0),
// conjugate (mode 1) or |.|^2 (mode 2).  partial holds
// col_sum_partials(n) * P complex values.
int64_t col_sum_partials(int64_t n);
template <typename T>
void launch_ok_colsum(const cplx<T>* O, int64_t n, int P, const cplx<double>* coef, int mode, double shift,
                      const cplx<double>* v, cplx<double>* partial, cplx<double>* out, hipStream_t stream);
// diag[k] = Re(out2[k]) - |obar[k]|^2  (Jacobi preconditioner of S)
void launch_diag_from_sums(const cplx<double>* sq, const cplx<double>* obar, int P, double* diag,
                           hipStream_t stream);

}  // namespace ann
