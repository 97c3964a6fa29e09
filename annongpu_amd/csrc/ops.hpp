// ops.hpp -- launch wrappers of sampler_ops.hip (chain initialisation, SR vector
// plumbing, fp64 vector kernels).
#pragma once

#include "kernels.hpp"

namespace ann {

void launch_init_chains(uint64_t* confs, int num_chains, int words, int N, int move, int total_z,
                        uint32_t call_index, uint64_t seed, hipStream_t stream);

// inv[vis][j] ([N][Mpad]): the unique i with conn(i, j) = vis, -1 if none, -2 if several
template <typename T>
void launch_unpack_v(const DenseView<T>& net, const int* inv, const cplx<double>* v, cplx<double>* va,
                     cplx<double>* vb, cplx<T>* V, hipStream_t stream);

void launch_pack_y(int layout, int N, int M, int Mpad, int C, const uint32_t* conn, const cplx<double>* Ya,
                   const cplx<double>* Yb, const cplx<double>* Yw, const cplx<double>* v, double shift, int P,
                   cplx<double>* y, hipStream_t stream);

void launch_pack_fragments(const cplx<float>* V, int N, int Mpad, int npad, uint16_t* out, hipStream_t stream);

void launch_cdot(const cplx<double>* a, const cplx<double>* b, int64_t n, int conj_a, double* out2,
                 hipStream_t stream);
// out4 = (<a,b>.re, <a,b>.im, <c,d>.re, <c,d>.im), conjugating a and c when conj_a;
// scratch holds 256 doubles
void launch_dot2(const cplx<double>* a, const cplx<double>* b, const cplx<double>* c, const cplx<double>* d,
                 int64_t n, int conj_a, double* scratch, double* out4, hipStream_t stream);
void launch_cg_update_dev(cplx<double>* x, cplx<double>* r, const cplx<double>* p, const cplx<double>* Ap,
                          const double* rz, const double* pAp, int64_t n, hipStream_t stream);
void launch_cg_dir_dev(cplx<double>* p, const cplx<double>* z, const double* rz_new, const double* rz_old, int64_t n,
                       hipStream_t stream);
// preconditioned CG (cg.hip): dot products as kCgParts block partials [b][4] = (r^H z, -, r^H r, -);
// a matvec's partials of p^H Ap are [b][2]
constexpr int kCgParts = 256;
void launch_pcg_init(const cplx<double>* r, cplx<double>* z, cplx<double>* p, const double* diag, double shift,
                     int64_t n, int64_t jac_end, double* part, hipStream_t stream);

// K-FAC preconditioner of the W parameters of a dense net (kfac.hip): factors from the
// first kKfacSamples samples (spin covariance by popcount for equally weighted
// samples, else moments per group of 32 samples; g moments per group of 128 samples;
// fixed-order reductions), inverted once per solve; the step kernel takes the CG update and the
// preconditioning of every parameter (Jacobi below jac_end, K-FAC on the W part)
constexpr int64_t kKfacSamples = 16384;
constexpr int kKfacMaxSites = 128;
constexpr int kKfacMaxUnits = 2048;
struct KfacWork {
    double* spart;        // kfac_spart_elems: spin moments per sample group
    cplx<double>* gpart;  // kfac_gpart_elems: g moments per sample group and unit block
    double* C;            // [N][N]
    double* Ci;           // [N][N] (C + gc)^-1
    cplx<double>* Dblk;   // kfac_dblock_elems: D blocks [64][64], inverted in place
    double* gam;          // [2] gc, gd
};
struct KfacView {
    int N, M;
    const int* idx;              // [N][M] parameter index of (site, unit), -1 where none
    WParamMap wm;                // the same without the table when wm.analytic()
    const double* Ci;
    const cplx<double>* Dblk;
};
// one CG iteration after the matvec with the K-FAC preconditioner, or (Ap null) the
// preconditioning of the initial residual; the partials are [kCgParts][4]
struct KfacStep {
    cplx<double>* x;             // x += alpha p (step)
    const cplx<double>* r_in;    // r of the previous iteration
    cplx<double>* r_out;         // r - alpha Ap (step; a buffer other than r_in)
    cplx<double>* z;
    cplx<double>* p_init;        // initial preconditioning: p = z (else null)
    const cplx<double>* p;
    const cplx<double>* Ap;
    const double* diag;          // Jacobi diagonal of the parameters k < jac_end (null: identity)
    double shift;
    int64_t jac_end;
    const double* rzpart;        // partials of the previous iteration
    const double* apart;         // p^H Ap partials [a_parts][2]
    int a_parts;
    double* newpart;
    const int* done;
};
int kfac_dblocks(int M);
size_t kfac_dblock_elems(int M);
size_t kfac_spart_elems(int N, int64_t count);
size_t kfac_gpart_elems(int M, int64_t count);
// bits: the transposed spin bits (launch_spin_bits, vpad rows), read when w_uniform
// (every sample weighs the same: the spin covariance is counted by popcount)
void launch_kfac_setup(const uint64_t* confs, int words, const uint64_t* bits, int vpad, bool w_uniform,
                       const cplx<float>* g32, const cplx<double>* g64, int Mpad, const double* w, int64_t count, int N,
                       int M, double shift, KfacWork& kw, hipStream_t stream);
// the set-up in two parts: the factors C and D from the samples (kw.C, kw.Dblk), then
// the damping for `shift` and the inverses (Ci, Dblk inverted in place)
void launch_kfac_moments(const uint64_t* confs, int words, const uint64_t* bits, int vpad, bool w_uniform,
                         const cplx<float>* g32, const cplx<double>* g64, int Mpad, const double* w, int64_t count,
                         int N, int M, KfacWork& kw, hipStream_t stream);
void launch_kfac_finish(int N, int M, double shift, KfacWork& kw, hipStream_t stream);
// idx of KfacView: RBM v M + j, PsiDeep N + M + inv[v][j] M + j
void launch_kfac_index(int layout, int N, int M, int Mpad, const int* inv, int* idx, hipStream_t stream);
void launch_kfac_step(const KfacView& kv, const KfacStep& a, hipStream_t stream);
void launch_pcg_pap(const cplx<double>* p, const cplx<double>* Ap, int64_t n, double* part, hipStream_t stream);
// done = {converged, iteration}: step and dir leave x, r, z, p untouched once it is set,
// launch_pcg_check sets it when r^H r of the iteration's partials is <= (rtol ||F||)^2,
// ||F||^2 = fnorm2[0] on the device.  The
// Jacobi preconditioner covers the parameters k < jac_end (all of them without K-FAC);
// r^H z is the sum of slots 0 (Jacobi) and 1 (K-FAC) of the partials.
void launch_pcg_step(cplx<double>* x, cplx<double>* r, cplx<double>* z, const cplx<double>* p,
                     const cplx<double>* Ap, const double* diag, double shift, int64_t n, int64_t jac_end,
                     const double* rzpart, const double* apart, int a_parts, double* newpart, const int* done,
                     hipStream_t stream);
void launch_pcg_dir(cplx<double>* p, const cplx<double>* z, int64_t n, const double* newpart, const double* oldpart,
                    const int* done, hipStream_t stream);
// mirror (optional): device address of host-visible memory that receives done
void launch_pcg_check(const double* part, const double* fnorm2, double rtol, int iter, int* done, hipStream_t stream,
                      int* mirror = nullptr);
// out[4] = (r^H r, iteration, ||F||^2, converged) of the solve's result (part: the
// partials of every iteration, [iteration][kCgParts][4]; it: iterations run)
void launch_pcg_final(const double* part, int it, const int* done, const double* fnorm2, double* out,
                      hipStream_t stream);
void launch_center_scale(cplx<double>* u, const double* w, const double* c2, int64_t n, hipStream_t stream);
void launch_cg_step(cplx<double>* x, cplx<double>* r, const cplx<double>* p, const cplx<double>* Ap, double alpha,
                    int64_t n, hipStream_t stream);
void launch_cg_dir(cplx<double>* p, const cplx<double>* r, double beta, int64_t n, hipStream_t stream);
void launch_precond(cplx<double>* z, const cplx<double>* r, const double* diag, double shift, int64_t n,
                    hipStream_t stream);
void launch_sr_diag(int layout, int N, int M, int Mpad, const double* wsum, const cplx<double>* Dj,
                    const cplx<double>* Obar, int P, double* diag, hipStream_t stream);
void launch_rank1(cplx<double>* S, const cplx<double>* o, int P, double alpha, hipStream_t stream);
void launch_add_diag(cplx<double>* S, int P, double shift, hipStream_t stream);
// A <- D^-1/2 A D^-1/2 with D = diag(A) (entries <= 0 left unscaled), dinv[i] = D_ii^-1/2:
// the Cholesky factor of the scaled matrix solves the same system (x = D^-1/2 y for
// (D^-1/2 A D^-1/2) y = D^-1/2 b) and is far better conditioned when the parameters'
// scales differ by orders of magnitude
void launch_equilibrate(cplx<double>* A, int P, double* dinv, hipStream_t stream);
// x[i] *= dinv[i]
void launch_scale_vec(cplx<double>* x, int P, const double* dinv, hipStream_t stream);
template <typename Tin, typename Tout>
void launch_convert(const cplx<Tin>* a, cplx<Tout>* b, int64_t n, hipStream_t stream);
void launch_fill(double* a, double v, int64_t n, hipStream_t stream);
void launch_weight_coef(const double* w, const cplx<double>* e, int64_t n, cplx<double>* coef, hipStream_t stream);
// coef = w (e - E): F = sum_s coef_s conj(O_s) without the cancellation of E conj(<O>)
void launch_weight_coef_centered(const double* w, const cplx<double>* e, cplx<double> E, int64_t n,
                                 cplx<double>* coef, hipStream_t stream);
// the same with E = (sums[0], sums[1]) on the device (launch_weighted_sums' output)
void launch_weight_coef_centered_dev(const double* w, const cplx<double>* e, const double* sums, int64_t n,
                                     cplx<double>* coef, hipStream_t stream);

}  // namespace ann
