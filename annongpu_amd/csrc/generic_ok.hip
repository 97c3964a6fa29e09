// generic_ok.hip -- kernels over a materialised O matrix [n][P] (row = sample,
// parameter fastest), used for wave functions whose O_k has no low-rank
// structure (PsiCNN).  They compute the reference's sample sums
//   <O_k>, <conj(O_k) E_loc>, S.v = sum_s w_s conj(O_s) (O_s.v) - conj(<O>) (<O>.v)
// (source/network_functions/TDVP.cu.template:77-126, 336-443) as two
// HBM-streaming passes: a row pass (one wavefront per sample, coalesced over k)
// and a column pass (one thread per parameter, coalesced over k, rows split
// over blocks, partial sums reduced in a second kernel).
#include "cnn.hpp"

namespace ann {

namespace {

template <typename T>
__global__ __launch_bounds__(256) void ok_rowdot_kernel(const cplx<T>* __restrict__ O, int64_t n, int P,
                                                        const cplx<double>* __restrict__ v,
                                                        cplx<double>* __restrict__ u) {
    const int lane = threadIdx.x & 63;
    const int64_t s = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (s >= n) return;
    const cplx<T>* row = O + s * P;
    double re = 0.0, im = 0.0;
    for (int k = lane; k < P; k += 64) {
        const cplx<T> o = row[k];
        const cplx<double> x = v[k];
        re += double(o.re) * x.re - double(o.im) * x.im;
        im += double(o.re) * x.im + double(o.im) * x.re;
    }
    re = wave_sum(re);
    im = wave_sum(im);
    if (lane == 0) u[s] = {re, im};
}

constexpr int kColRows = 64;   // rows per block of the column pass

template <typename T>
__global__ __launch_bounds__(256) void ok_colsum_kernel(const cplx<T>* __restrict__ O, int64_t n, int P,
                                                        const cplx<double>* __restrict__ coef, int mode,
                                                        cplx<double>* __restrict__ partial) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    const int64_t r0 = int64_t(blockIdx.y) * kColRows;
    if (k >= P) return;
    const int64_t r1 = r0 + kColRows < n ? r0 + kColRows : n;
    double re = 0.0, im = 0.0;
    for (int64_t s = r0; s < r1; s++) {
        const cplx<T> o = O[s * P + k];
        const cplx<double> c = coef[s];
        const double ore = double(o.re), oim = mode == 1 ? -double(o.im) : double(o.im);
        if (mode == 2) {
            const double m2 = ore * ore + oim * oim;
            re += c.re * m2;
            im += c.im * m2;
        } else {
            re += c.re * ore - c.im * oim;
            im += c.re * oim + c.im * ore;
        }
    }
    partial[int64_t(blockIdx.y) * P + k] = {re, im};
}

__global__ __launch_bounds__(256) void colsum_finish_kernel(const cplx<double>* __restrict__ partial, int64_t parts,
                                                            int P, double shift, const cplx<double>* __restrict__ v,
                                                            cplx<double>* __restrict__ out) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= P) return;
    double re = 0.0, im = 0.0;
    for (int64_t q = 0; q < parts; q++) {
        re += partial[q * P + k].re;
        im += partial[q * P + k].im;
    }
    if (v) {
        re += shift * v[k].re;
        im += shift * v[k].im;
    }
    out[k] = {re, im};
}

__global__ void diag_from_sums_kernel(const cplx<double>* sq, const cplx<double>* obar, int P, double* diag) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P) return;
    diag[k] = sq[k].re - (obar[k].re * obar[k].re + obar[k].im * obar[k].im);
}

}  // namespace

int64_t col_sum_partials(int64_t n) { return (n + kColRows - 1) / kColRows; }

template <typename T>
void launch_ok_rowdot(const cplx<T>* O, int64_t n, int P, const cplx<double>* v, cplx<double>* u,
                      hipStream_t stream) {
    if (n <= 0) return;
    ok_rowdot_kernel<T><<<unsigned((n + 3) / 4), 256, 0, stream>>>(O, n, P, v, u);
    ANN_LAUNCH_CHECK();
}

template <typename T>
void launch_ok_colsum(const cplx<T>* O, int64_t n, int P, const cplx<double>* coef, int mode, double shift,
                      const cplx<double>* v, cplx<double>* partial, cplx<double>* out, hipStream_t stream) {
    if (P <= 0) return;
    const int64_t parts = col_sum_partials(n);
    if (parts > 65535) throw std::invalid_argument("too many samples for the column pass");
    if (parts > 0) {
        ok_colsum_kernel<T><<<dim3(unsigned((P + 255) / 256), unsigned(parts)), 256, 0, stream>>>(O, n, P, coef,
                                                                                                   mode, partial);
        ANN_LAUNCH_CHECK();
    }
    colsum_finish_kernel<<<unsigned((P + 255) / 256), 256, 0, stream>>>(partial, parts, P, shift, v, out);
    ANN_LAUNCH_CHECK();
}

void launch_diag_from_sums(const cplx<double>* sq, const cplx<double>* obar, int P, double* diag,
                           hipStream_t stream) {
    if (P <= 0) return;
    diag_from_sums_kernel<<<unsigned((P + 255) / 256), 256, 0, stream>>>(sq, obar, P, diag);
    ANN_LAUNCH_CHECK();
}

template void launch_ok_rowdot<float>(const cplx<float>*, int64_t, int, const cplx<double>*, cplx<double>*,
                                      hipStream_t);
template void launch_ok_rowdot<double>(const cplx<double>*, int64_t, int, const cplx<double>*, cplx<double>*,
                                       hipStream_t);
template void launch_ok_colsum<float>(const cplx<float>*, int64_t, int, const cplx<double>*, int, double,
                                      const cplx<double>*, cplx<double>*, cplx<double>*, hipStream_t);
template void launch_ok_colsum<double>(const cplx<double>*, int64_t, int, const cplx<double>*, int, double,
                                       const cplx<double>*, cplx<double>*, cplx<double>*, hipStream_t);

}  // namespace ann
