// distance.hip -- per-sample terms of the distance between two wave functions on
// the samples of the first (HilbertSpaceDistance), reduced in fp64.
//
// For each sample s of psi's ensemble (weight w_s), with E_s the local energy of the
// operator U on psi and d_s = conj(log psi'(s) - log psi(s)):
//   unitary U:  omega_s = exp(d_s) E_s,      norm_s = |E_s|^2
//   otherwise:  omega_s = exp(E_s + d_s),    norm_s = exp(2 Re E_s)
//   ratio_s = exp(2 (Re log psi'(s) - Re log psi(s)))
// (source/network_functions/HilbertSpaceDistance.cu.template:31-84).  The kernel
// writes w omega and w ratio per sample (the coefficients of the O_k sums of the
// gradient) and the sums of w omega, w ratio and w norm.
//
// KullbackLeibler (source/network_functions/KullbackLeibler.cu.template:16-112): the
// samples come from psi' (weights w'_s); with l = scale log psi(s), l' = log psi'(s):
//   weight_s = exp(2 (Re l - Re l')) w'_s,  dev_s = l' - l - last mean deviation,
//   mask_s = |dev_s|^2 > threshold^2,
// and the kernel writes the coefficients weight_s, weight_s dev_s mask_s and
// weight_s |dev_s|^2 mask_s per sample and the sums of weight, weight (l' - l),
// weight dev mask and weight |dev|^2 mask.
#include "kernels.hpp"

namespace ann {

namespace {

constexpr int kHsdBlock = 1024;

__global__ __launch_bounds__(kHsdBlock) void hsd_terms_kernel(const cplx<double>* e_loc, const cplx<double>* lp,
                                                              const cplx<double>* lpp, const double* w, int64_t n,
                                                              int unitary, cplx<double>* c_omega,
                                                              cplx<double>* c_ratio, double* sums) {
    __shared__ double s0[kHsdBlock], s1[kHsdBlock], s2[kHsdBlock], s3[kHsdBlock];
    double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += kHsdBlock) {
        const double ws = w[i];
        const cplx<double> E = e_loc[i];
        const double dr = lpp[i].re - lp[i].re, di = -(lpp[i].im - lp[i].im);
        cplx<double> om;
        double norm;
        if (unitary) {
            const double m = exp(dr);
            const double cr = m * cos(di), ci = m * sin(di);
            om = {cr * E.re - ci * E.im, cr * E.im + ci * E.re};
            norm = E.re * E.re + E.im * E.im;
        } else {
            const double m = exp(dr + E.re);
            om = {m * cos(di + E.im), m * sin(di + E.im)};
            norm = exp(2.0 * E.re);
        }
        const double ratio = exp(2.0 * dr);
        const cplx<double> wo = {ws * om.re, ws * om.im};
        c_omega[i] = wo;
        c_ratio[i] = {ws * ratio, 0.0};
        a += wo.re;
        b += wo.im;
        c += ws * ratio;
        d += ws * norm;
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = b;
    s2[threadIdx.x] = c;
    s3[threadIdx.x] = d;
    __syncthreads();
    for (int o = kHsdBlock / 2; o > 0; o >>= 1) {
        if (int(threadIdx.x) < o) {
            s0[threadIdx.x] += s0[threadIdx.x + o];
            s1[threadIdx.x] += s1[threadIdx.x + o];
            s2[threadIdx.x] += s2[threadIdx.x + o];
            s3[threadIdx.x] += s3[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sums[0] = s0[0];
        sums[1] = s1[0];
        sums[2] = s2[0];
        sums[3] = s3[0];
    }
}

__global__ __launch_bounds__(kHsdBlock) void kl_terms_kernel(const cplx<double>* lp, const cplx<double>* lpp,
                                                             const double* wp, int64_t n, double scale,
                                                             cplx<double> last, double thr2, cplx<double>* c_w,
                                                             cplx<double>* c_d, cplx<double>* c_d2, double* sums) {
    constexpr int kSums = 6;
    __shared__ double red[kSums][kHsdBlock];
    double acc[kSums] = {0, 0, 0, 0, 0, 0};
    for (int64_t i = threadIdx.x; i < n; i += kHsdBlock) {
        const double lr = scale * lp[i].re, li = scale * lp[i].im;
        const double wt = exp(2.0 * (lr - lpp[i].re)) * wp[i];
        const double gr = lpp[i].re - lr, gi = lpp[i].im - li;
        const double dr = gr - last.re, di = gi - last.im;
        const double d2 = dr * dr + di * di;
        const double m = d2 > thr2 ? wt : 0.0;
        c_w[i] = {wt, 0.0};
        c_d[i] = {m * dr, m * di};
        c_d2[i] = {m * d2, 0.0};
        acc[0] += wt;
        acc[1] += wt * gr;
        acc[2] += wt * gi;
        acc[3] += m * dr;
        acc[4] += m * di;
        acc[5] += m * d2;
    }
#pragma unroll
    for (int k = 0; k < kSums; k++) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int o = kHsdBlock / 2; o > 0; o >>= 1) {
        if (int(threadIdx.x) < o) {
#pragma unroll
            for (int k = 0; k < kSums; k++) red[k][threadIdx.x] += red[k][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x < kSums) sums[threadIdx.x] = red[threadIdx.x][0];
}

}  // namespace

void launch_kl_terms(const cplx<double>* lp, const cplx<double>* lpp, const double* wp, int64_t n, double scale,
                     cplx<double> last_mean_deviation, double threshold, cplx<double>* c_w, cplx<double>* c_d,
                     cplx<double>* c_d2, double* sums6, hipStream_t stream) {
    if (n <= 0) throw std::invalid_argument("no samples");
    kl_terms_kernel<<<1, kHsdBlock, 0, stream>>>(lp, lpp, wp, n, scale, last_mean_deviation, threshold * threshold,
                                                 c_w, c_d, c_d2, sums6);
    ANN_LAUNCH_CHECK();
}

void launch_hsd_terms(const cplx<double>* e_loc, const cplx<double>* lp, const cplx<double>* lpp, const double* w,
                      int64_t n, bool unitary, cplx<double>* c_omega, cplx<double>* c_ratio, double* sums4,
                      hipStream_t stream) {
    if (n <= 0) throw std::invalid_argument("no samples");
    hsd_terms_kernel<<<1, kHsdBlock, 0, stream>>>(e_loc, lp, lpp, w, n, unitary ? 1 : 0, c_omega, c_ratio, sums4);
    ANN_LAUNCH_CHECK();
}

}  // namespace ann
