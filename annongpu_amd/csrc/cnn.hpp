// cnn.hpp -- device view and launchers of PsiCNN (cnn.hip) and of the
// materialised-O kernels used by wave functions without a structured O
// (generic_ok.hip).
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
};

// LDS bytes per wavefront
template <typename T>
size_t cnn_forward_lds(const CnnView<T>& net);
template <typename T>
size_t cnn_ok_lds(const CnnView<T>& net);

template <typename T>
void launch_cnn_log_psi(const CnnView<T>& net, const uint64_t* confs, int words, int64_t count,
                        cplx<double> log_prefactor, cplx<double>* log_psi, double* weight, hipStream_t stream);
template <typename T>
void launch_cnn_mc(const CnnView<T>& net, uint64_t* chain_confs, int words, int num_chains, int move,
                   int64_t num_proposals, uint64_t counter, uint64_t seed, uint32_t* accepted, hipStream_t stream);
template <typename T>
void launch_cnn_local_energy(const CnnView<T>& net, const OpView& op, const uint64_t* confs, int64_t count,
                             cplx<double>* e_loc, hipStream_t stream);
// O[s][k] = d log psi(s) / d param_k - center[k] (center may be null)
template <typename T>
void launch_cnn_ok(const CnnView<T>& net, const uint64_t* confs, int words, int64_t count,
                   const cplx<double>* center, cplx<T>* O, hipStream_t stream);

// ------------------------------------------------------- materialised O [n][P]
// u[s] = sum_k O[s][k] v[k]   (fp64 accumulation)
template <typename T>
void launch_ok_rowdot(const cplx<T>* O, int64_t n, int P, const cplx<double>* v, cplx<double>* u,
                      hipStream_t stream);
// out[k] = sum_s coef[s] * op(O[s][k]) + shift * v[k]; op = identity (mode 0),
// conjugate (mode 1) or |.|^2 (mode 2, real part only).  partial holds
// col_sum_partials(n) * P complex values.
int64_t col_sum_partials(int64_t n);
template <typename T>
void launch_ok_colsum(const cplx<T>* O, int64_t n, int P, const cplx<double>* coef, int mode, double shift,
                      const cplx<double>* v, cplx<double>* partial, cplx<double>* out, hipStream_t stream);
// diag[k] = Re(out2[k]) - |obar[k]|^2  (Jacobi preconditioner of S)
void launch_diag_from_sums(const cplx<double>* sq, const cplx<double>* obar, int P, double* diag,
                           hipStream_t stream);

}  // namespace ann
