// kernels.hpp -- device views and launch wrappers of the gfx950 VMC kernels.
//
// A "dense net" is PsiRBM (include/quantum_state/PsiRBM.hpp:193-218) or a PsiDeep
// with one hidden layer (include/quantum_state/PsiDeep.hpp:70-135).  On the device
// both are stored as a dense visible x hidden matrix Wd[v][j] (the gather table of
// PsiDeep folded in), hidden units padded to Mpad = 64 * U so that one wavefront
// owns one chain and every lane owns U hidden units.
#pragma once

#include "common.hpp"

#include <vector>

namespace ann {

// MOVE_PAULI: a random site of a Pauli-string configuration set to a uniformly random
// Pauli type, the current one included (include/ensembles/policies/Update_Policy.hpp:41-56)
enum MoveKind : int { MOVE_FLIP = 0, MOVE_EXCHANGE = 1, MOVE_PAULI = 2 };
// Configurations: spins (bit i = site i up, words = ceil(N / 64)) or Pauli strings of
// at most 64 sites, stored as the two words (a, b): site i holds I (0, 0), X (1, 0),
// Y (0, 1) or Z (1, 1) (include/basis/PauliString.hpp:18-72)
enum Basis : int { BASIS_SPINS = 0, BASIS_PAULIS = 1 };
enum ParamLayout : int { LAYOUT_RBM = 0, LAYOUT_DEEP1 = 1 };

template <typename T>
struct DenseView {
    int N;             // visible units (sites)
    int M;             // hidden units
    int Mpad;          // 64 * U
    int C;             // connectivity (PsiDeep lhs_connectivity; N for PsiRBM)
    int layout;        // ParamLayout
    int num_params;
    int real_f;        // all final weights real -> cheaper real-part kernels
    const cplx<T>* Wd;     // [N][Mpad]
    const cplx<T>* b;      // [Mpad] (zero for PsiRBM)
    const cplx<T>* f;      // [Mpad] final weights (zero in the padding)
    const uint32_t* conn;  // [C*M] PsiDeep connections (nullptr for PsiRBM)
};

// The W parameter of (site v, hidden unit j) without a table: PsiRBM v M + j; a
// PsiDeep whose connection i of unit j is site (j shift + i) mod N (the circulant
// one-dimensional layout, shift >= 0) N + M + i M + j with i = (v - j shift) mod N
// when i < C, else none (-1).  shift < 0: a PsiDeep with other connections (the
// kernels then read tables).
// sm = 1 (a circulant PsiDeep with C = N, where (site, unit) -> parameter is a
// bijection): vectors in site-major order instead, W of (v, j) at N + M + v M + j
// (the order the CG solve keeps its vectors in; to_site_major converts).
struct WParamMap {
    int rbm, N, M, C, shift, sm = 0;
    __host__ __device__ bool analytic() const { return rbm || shift >= 0; }
    __host__ __device__ int param(int v, int j) const { return param_at(v, j, rbm || sm ? 0 : (j * shift) % N); }
    // the same with jm = (j shift) mod N given (no division per site)
    __host__ __device__ int param_at(int v, int j, int jm) const {
        if (rbm) return v * M + j;
        if (sm) return N + M + v * M + j;
        int i = v - jm;
        if (i < 0) i += N;
        return i < C ? N + M + i * M + j : -1;
    }
    // site of connection i of unit j (circulant PsiDeep)
    __host__ __device__ int site(int i, int j) const { return (j * shift + i) % N; }
};

// Pauli-sum operator compiled for the device: strings grouped by flip mask so
// that X and Y strings that reach the same s' share one wave-function ratio.
struct OpView {
    int num_groups;
    int words;
    const uint64_t* group_flip;      // [num_groups][words]
    const int* group_begin;          // [num_groups + 1] ranges into the string arrays
    const uint64_t* string_zy;       // [num_strings][words]  Z|Y mask (b bits)
    const double* string_coef;       // [num_strings][2] coefficient * (-i)^{#Y}
};

// Pauli-sum operator acting on Pauli-string configurations (string n maps x to
// (a_n ^ x.a, b_n ^ x.b) with the phase of PauliString::apply, PauliString.hpp:257-276)
struct PauliOpView {
    int num_strings;
    const uint64_t* ab;              // [num_strings][2]: a, b of the string
    const double* coef;              // [num_strings][2]
};

// ---------------------------------------------------------------- launches
// Every wrapper launches on `stream` and throws HipError on a launch failure.

void launch_enumerate_confs(uint64_t* confs, int64_t first_index, int64_t count, int words,
                            hipStream_t stream);
// Pauli strings by index (PauliString::enumerate, PauliString.hpp:30-35): site i takes
// bit 2i of the index as its a bit and bit 2i + 1 as its b bit; confs [count][2]
void launch_enumerate_paulis(uint64_t* confs, int64_t first_index, int64_t count, hipStream_t stream);

template <typename T>
void launch_theta_valu(const DenseView<T>& net, const uint64_t* confs, int words, int64_t count,
                       cplx<T>* theta, hipStream_t stream);

// theta = S * Wd + b on MFMA: spins are exact in bf16, Wd is split into three bf16
// terms (hi, mid, lo) and accumulated in fp32.  wfrag is the fragment-ordered
// split matrix built by pack_theta_fragments().
void launch_theta_mfma(const DenseView<float>& net, const uint16_t* wfrag, int npad,
                       const uint64_t* confs, int words, int64_t count, cplx<float>* theta,
                       hipStream_t stream);
void pack_theta_fragments(const double* Wd_interleaved, int N, int Mpad, int nsplit,
                          std::vector<uint16_t>& out, int& npad);
// Wd64/Wd32, the PsiDeep bias and the 3-term fragments (npad rows) rebuilt on the device
// from the parameter vector (PsiRBM W[v][j]; one-layer PsiDeep a, b, W[i][j] with
// conn[i][j], or with conn[i][j] = (j shift + i) mod N when shift >= 0); buffers laid
// out and padded as by the host packing, padding untouched
void launch_net_repack(int layout, int N, int M, int Mpad, int C, int shift, const cplx<double>* params,
                       const uint32_t* conn, int npad, cplx<double>* Wd64, cplx<float>* Wd32, cplx<double>* b64,
                       cplx<float>* b32, uint16_t* wfrag, hipStream_t stream);

template <typename T>
void launch_net_log_psi(const DenseView<T>& net, const cplx<T>* theta, int64_t count,
                        cplx<double> log_prefactor, cplx<double>* log_psi, double* weight_exact,
                        hipStream_t stream);

template <typename T>
void launch_mc_sweep(const DenseView<T>& net, uint64_t* chain_confs, cplx<T>* theta, int num_chains,
                     int words, int move, int64_t num_proposals, uint64_t counter, uint64_t seed,
                     uint32_t* accepted, hipStream_t stream);

template <typename T>
void launch_local_energy(const DenseView<T>& net, const OpView& op, const uint64_t* confs,
                         const cplx<T>* theta, int64_t count, cplx<double>* e_loc,
                         hipStream_t stream);

template <typename T>
void launch_g_factor(const DenseView<T>& net, const cplx<T>* theta, int64_t count, cplx<T>* g,
                     hipStream_t stream);

template <typename T>
void launch_materialize_ok(const DenseView<T>& net, const uint64_t* confs, int words,
                           const cplx<T>* g, int64_t count, const cplx<double>* center,
                           cplx<T>* O, hipStream_t stream);

// Yw[v][j] = sum_s coef_s spin_sv X_sj, Yb[j] = sum_s coef_s X_sj,
// Ya[v] = sum_s coef_s spin_sv, fp64 accumulation (fp64 MFMA).
// X_sj = g_sj (conj_g = 0), conj(g_sj) (1) or |g_sj|^2 (2).
template <typename T>
void launch_spin_weighted_sums(const DenseView<T>& net, const uint64_t* confs, int words,
                               const cplx<T>* g, int conj_g, const cplx<double>* coef,
                               int64_t count, cplx<double>* Yw, cplx<double>* Yb,
                               cplx<double>* Ya, hipStream_t stream);

// u_s = coef_s * (O_s . v - shift_u) for the structured O of a dense net; V is the
// reindexed W part of v ([N][Mpad]), vb the bias part, va the visible part.
template <typename T>
void launch_structured_Ov(const DenseView<T>& net, const uint64_t* confs, int words,
                          const cplx<T>* g, const cplx<T>* SV, const cplx<double>* vb,
                          const cplx<double>* va, int64_t count, cplx<double>* u,
                          hipStream_t stream);

// weighted fp64 reductions
// the spin GEMM of launch_spin_weighted_sums on bf16 matrix cores (fp32 accumulation
// per sample chunk, chunks added in fp64)
// ws (optional): the caller's transposed spin bits of the samples (launch_spin_bits with
// the plan's vpad) and workspace of spin_sums_work sizes; without it the call builds the
// bits and allocates its workspace on the stream
struct SpinSumsWork {
    const uint64_t* bits;
    float* partial;          // plan.chunks * plan.n
    double* full;            // plan.n
    cplx<float>* c32;        // sr_count_pad(count)
};
void launch_spin_sums_bf16(const DenseView<float>& net, const uint64_t* confs, int words, const cplx<float>* g,
                           int mode, const cplx<double>* coef, int64_t count, cplx<double>* Yw, cplx<double>* Yb,
                           cplx<double>* Ya, hipStream_t stream, const SpinSumsWork* ws = nullptr);

// the spin GEMM's partition of the samples: npairs column pairs (16 hidden units
// each, plus the coefficient columns when with_coef) x chunks of `chunk` samples x
// zgroups row groups; fp32 partials [chunks][n] floats, n = vpad * ncols (columns of
// the hidden-unit padding are not written)
struct SpinPlan {
    int vpad = 0, ncols = 0, npairs = 0, zgroups = 0, chunks = 0;
    int64_t chunk = 0, n = 0, bit_blocks = 0;
};
SpinPlan spin_plan(int N, int M, int Mpad, int64_t count, bool with_coef);
// transposed spin bits of the configurations: bits[(count+63)/64][vpad], bit b of
// word [blk][v] = spin v of sample 64 blk + b (row N all ones)
void launch_spin_bits(const uint64_t* confs, int words, int64_t count, int N, int vpad, uint64_t* bits,
                      hipStream_t stream);
// rows of g and of the fp32 coefficients the spin GEMM and O.v kernels read: count
// rounded up to whole 256-sample rounds, the rows beyond count zero
int64_t sr_count_pad(int64_t count);
// Y = S^T X partials per sample chunk (X from g and coef as selected by mode, see
// sr_mfma.hip); with_coef = 0 leaves the coefficient columns zero; g and coef
// have sr_count_pad(count) rows
void launch_spin_gemm32(const DenseView<float>& net, const uint64_t* bits, const cplx<float>* g, int mode,
                        const cplx<float>* coef, int with_coef, int64_t count, const SpinPlan& p, float* partial,
                        hipStream_t stream);

// workspace of the fused fp32 matvec: frag sr_frag_elems(N, Mpad), va N, vb Mpad,
// upart sr_ov_pairs(M) * count, u sr_count_pad(count), spin_part chunks * n, opart 128,
// ypart 2 * sr_y_parts(net, wm) (at most 512)
struct SrWorkspace {
    uint16_t* frag;
    cplx<double>* va;
    cplx<double>* vb;
    cplx<float>* upart;
    cplx<float>* u;
    float* spin_part;
    double* opart;
    double* ypart;
};
// y = (S + shift) v with S = sum_s w_s conj(O_s) O_s^T - conj(<O>) <O>^T for the
// structured O: launch_sr_pack, launch_sr_ov32, launch_sr_finish (u_s = w_s (O_s.v -
// <O>.v) into ws.u), launch_spin_gemm32 (mode 1, coefficients ws.u, partials into
// ws.spin_part), launch_sr_y_bf16 (y, and the partials of v^H y in ws.ypart,
// [sr_y_parts(net, wm)][2])
// the CG direction update p_new = z + (rz_new / rz_old) p_old, fused into the first
// kernel of the matvec (z null: plain matvec of v).  rz partials as in cg.hip
// ([parts][4], element 0); nothing changes once done[0] is set (p_new = p_old).
struct CgDir {
    const cplx<double>* z = nullptr;
    const cplx<double>* p_old = nullptr;
    const double* newpart = nullptr;
    const double* oldpart = nullptr;
    int parts = 0;
    const int* done = nullptr;
    cplx<double>* p_new = nullptr;
};
int sr_y_parts(const DenseView<float>& net, const WParamMap& wm);
int sr_kt32(int N);
int sr_ov_pairs(int M);
size_t sr_frag_elems(int N, int Mpad);
// the first half in three launches: V fragments (and the fused direction update),
// O.v per column pair, u_s = w_s (O_s.v - <O>.v)
void launch_sr_pack(const DenseView<float>& net, const WParamMap& wm, const int* inv, const cplx<double>* v,
                    const CgDir& dir, const cplx<double>* obar, SrWorkspace ws, hipStream_t stream);
void launch_sr_ov32(const DenseView<float>& net, const uint64_t* confs, int words, int64_t count,
                    const cplx<float>* g, SrWorkspace ws, hipStream_t stream);
void launch_sr_finish(const DenseView<float>& net, const double* w, const uint64_t* confs, int words, int64_t count,
                      SrWorkspace ws, hipStream_t stream);
void launch_sr_y_bf16(const DenseView<float>& net, const WParamMap& wm, const cplx<double>* v, double shift,
                      const SpinPlan& plan, SrWorkspace ws, cplx<double>* y, hipStream_t stream);

// site-major order of a circulant PsiDeep's parameter vector (WParamMap sm):
// forward out[N + M + v M + j] = in[N + M + i M + j] with v = (j shift + i) mod N
// (C = N), the a and b parts copied; backward the inverse
void launch_site_major(const cplx<double>* in, int N, int M, int shift, bool forward, cplx<double>* out,
                       hipStream_t stream);

void launch_weighted_sums(const cplx<double>* e_loc, const double* w, int64_t count,
                          double* out4 /* E.re, E.im, E2, sum w */, hipStream_t stream);

// HilbertSpaceDistance per-sample terms: c_omega = w omega, c_ratio = w ratio and
// sums4 = [sum w omega (re, im), sum w ratio, sum w norm]
void launch_hsd_terms(const cplx<double>* e_loc, const cplx<double>* lp, const cplx<double>* lpp, const double* w,
                      int64_t n, bool unitary, cplx<double>* c_omega, cplx<double>* c_ratio, double* sums4,
                      hipStream_t stream);

// KullbackLeibler per-sample terms on psi' samples: c_w = weight, c_d = weight dev mask,
// c_d2 = weight |dev|^2 mask and sums6 = [sum weight, sum weight (l' - l) (re, im),
// sum weight dev mask (re, im), sum weight |dev|^2 mask]
void launch_kl_terms(const cplx<double>* lp, const cplx<double>* lpp, const double* wp, int64_t n, double scale,
                     cplx<double> last_mean_deviation, double threshold, cplx<double>* c_w, cplx<double>* c_d,
                     cplx<double>* c_d2, double* sums6, hipStream_t stream);

// dense Hermitian Gram S = O^H diag(w) O  (O row-major [count][P], complex T)
template <typename T>
void launch_gram(const cplx<T>* O, const double* w, int64_t count, int P, cplx<double>* S,
                 hipStream_t stream);

// y = S x for dense S in fp64 (complex)
void launch_dense_matvec(const cplx<double>* S, const cplx<double>* x, int P, double shift,
                         cplx<double>* y, hipStream_t stream);

// in-place blocked Cholesky of the lower triangle (fp64 complex, matrix cores); info != 0
// on failure; work holds cholesky_work_elems() complex values
size_t cholesky_work_elems();
void cholesky_lower(cplx<double>* A, int P, int* d_info, cplx<double>* work, hipStream_t stream);
void cholesky_solve_inplace(const cplx<double>* L, int P, cplx<double>* x, hipStream_t stream);

}  // namespace ann
