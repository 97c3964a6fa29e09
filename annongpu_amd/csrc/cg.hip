// cg.hip -- vector kernels of the preconditioned conjugate-gradient solve of
// (S + shift) x = F, fp64.  Scalars never leave the device: every dot product
// is written as per-block partial sums, and the kernel that needs the scalar
// adds the partials itself (all blocks in the same fixed order), so one CG
// iteration costs the matvec plus three launches.
#include "ops.hpp"

namespace ann {

namespace {

__device__ inline double block_sum(double v, double* sh) {
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    return sh[0] + sh[1] + sh[2] + sh[3];
}

__device__ inline double total(const double* part, int nparts, int stride, int idx, double* sh) {
    const double x = int(threadIdx.x) < nparts ? part[int64_t(threadIdx.x) * stride + idx] : 0.0;
    return block_sum(x, sh);
}

__device__ inline cplx<double> precond(const cplx<double> r, const double* diag, double shift, int64_t i) {
    if (!diag) return r;
    const double d = diag[i] + shift;
    return {r.re / d, r.im / d};
}

// r^H z of one iteration's partials: slot 0 (Jacobi part) + slot 1 (K-FAC part)
__device__ inline double total_rz(const double* part, int nparts, double* sh) {
    return total(part, nparts, 4, 0, sh) + total(part, nparts, 4, 1, sh);
}

// z = M^-1 r and p = z for the parameters k < jac_end (the rest are preconditioned
// by the K-FAC kernels afterwards), part[b] = (r^H z over k < jac_end, 0, r^H r)
__global__ __launch_bounds__(256) void cg_init_kernel(const cplx<double>* r, cplx<double>* z, cplx<double>* p,
                                                      const double* diag, double shift, int64_t n, int64_t jac_end,
                                                      double* part) {
    __shared__ double sh[4];
    double rz = 0.0, rr = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const cplx<double> ri = r[i];
        if (i < jac_end) {
            const cplx<double> zi = precond(ri, diag, shift, i);
            z[i] = zi;
            p[i] = zi;
            rz += ri.re * zi.re + ri.im * zi.im;
        }
        rr += ri.re * ri.re + ri.im * ri.im;
    }
    rz = block_sum(rz, sh);
    rr = block_sum(rr, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x * 4 + 0] = rz;
        part[blockIdx.x * 4 + 1] = 0.0;
        part[blockIdx.x * 4 + 2] = rr;
    }
}

// part[b] = Re(p^H Ap) partials (matvecs that do not produce them themselves)
__global__ __launch_bounds__(256) void cg_pap_kernel(const cplx<double>* p, const cplx<double>* Ap, int64_t n,
                                                     double* part) {
    __shared__ double sh[4];
    double d = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        d += p[i].re * Ap[i].re + p[i].im * Ap[i].im;
    d = block_sum(d, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x * 2] = d;
        part[blockIdx.x * 2 + 1] = 0.0;
    }
}

// alpha = rz / pAp; x += alpha p; r -= alpha Ap; z = M^-1 r for k < jac_end; new
// partials (r^H z over k < jac_end, 0, r^H r).  Nothing changes once done[0] is set
// (the solve has converged).
__global__ __launch_bounds__(256) void cg_step_kernel(cplx<double>* x, cplx<double>* r, cplx<double>* z,
                                                      const cplx<double>* p, const cplx<double>* Ap,
                                                      const double* diag, double shift, int64_t n,
                                                      int64_t jac_end, const double* rzpart, int rz_parts,
                                                      const double* apart, int a_parts, double* newpart,
                                                      const int* done) {
    __shared__ double sh[4];
    if (done[0]) return;
    const double rz = total_rz(rzpart, rz_parts, sh);
    const double pap = total(apart, a_parts, 2, 0, sh);
    const double alpha = pap > 0.0 ? rz / pap : 0.0;
    double rz2 = 0.0, rr = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        x[i].re += alpha * p[i].re;
        x[i].im += alpha * p[i].im;
        cplx<double> ri = r[i];
        ri.re -= alpha * Ap[i].re;
        ri.im -= alpha * Ap[i].im;
        r[i] = ri;
        if (i < jac_end) {
            const cplx<double> zi = precond(ri, diag, shift, i);
            z[i] = zi;
            rz2 += ri.re * zi.re + ri.im * zi.im;
        }
        rr += ri.re * ri.re + ri.im * ri.im;
    }
    rz2 = block_sum(rz2, sh);
    rr = block_sum(rr, sh);
    if (threadIdx.x == 0) {
        newpart[blockIdx.x * 4 + 0] = rz2;
        newpart[blockIdx.x * 4 + 1] = 0.0;
        newpart[blockIdx.x * 4 + 2] = rr;
    }
}

// p = z + (rz_new / rz_old) p (nothing once done[0] is set)
__global__ __launch_bounds__(256) void cg_dir_kernel(cplx<double>* p, const cplx<double>* z, int64_t n,
                                                     const double* newpart, const double* oldpart, int parts,
                                                     const int* done) {
    __shared__ double sh[4];
    if (done[0]) return;
    const double rn = total_rz(newpart, parts, sh);
    const double ro = total_rz(oldpart, parts, sh);
    const double beta = ro != 0.0 ? rn / ro : 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        p[i].re = z[i].re + beta * p[i].re;
        p[i].im = z[i].im + beta * p[i].im;
    }
}

// convergence test of iteration `iter` on the device: r^H r from its partials;
// the first iteration with r^H r <= tol2 sets done = {1, iter}
// tolerance (rtol ||F||)^2 with ||F||^2 = fnorm2[0]
__device__ inline double cg_tol2(const double* fnorm2, double rtol) {
    const double tol = rtol * sqrt(fnorm2[0]);
    return tol * tol;
}

// (mirror: a host-visible copy of done, written every check)
__global__ __launch_bounds__(256) void cg_check_kernel(const double* part, int parts, const double* fnorm2,
                                                       double rtol, int iter, int* done, int* mirror) {
    __shared__ double sh[4];
    int d0 = done[0], d1 = done[1];
    if (!d0) {
        const double rr = total(part, parts, 4, 2, sh);
        if (rr <= cg_tol2(fnorm2, rtol)) {
            d0 = 1;
            d1 = iter;
            if (threadIdx.x == 0) {
                done[0] = 1;
                done[1] = iter;
            }
        }
    }
    if (mirror && threadIdx.x == 0) {
        mirror[0] = d0;
        mirror[1] = d1;
    }
}

// out = (r^H r, iteration, ||F||^2, converged) of the solve's result: the iterate of
// the first converged check, else iteration `it`
__global__ __launch_bounds__(256) void cg_final_kernel(const double* part, int parts, int64_t stride, int it,
                                                       const int* done, const double* fnorm2, double* out) {
    __shared__ double sh[4];
    const int conv = done[0];
    const int iter = conv ? done[1] : it;
    const double rr = total(part + int64_t(iter) * stride, parts, 4, 2, sh);
    if (threadIdx.x == 0) {
        out[0] = rr;
        out[1] = double(iter);
        out[2] = fnorm2[0];
        out[3] = double(conv);
    }
}

}  // namespace

void launch_pcg_check(const double* part, const double* fnorm2, double rtol, int iter, int* done, hipStream_t stream,
                      int* mirror) {
    cg_check_kernel<<<1, 256, 0, stream>>>(part, kCgParts, fnorm2, rtol, iter, done, mirror);
    ANN_LAUNCH_CHECK();
}

void launch_pcg_final(const double* part, int it, const int* done, const double* fnorm2, double* out,
                      hipStream_t stream) {
    cg_final_kernel<<<1, 256, 0, stream>>>(part, kCgParts, int64_t(kCgParts) * 4, it, done, fnorm2, out);
    ANN_LAUNCH_CHECK();
}

void launch_pcg_init(const cplx<double>* r, cplx<double>* z, cplx<double>* p, const double* diag, double shift,
                    int64_t n, int64_t jac_end, double* part, hipStream_t stream) {
    cg_init_kernel<<<kCgParts, 256, 0, stream>>>(r, z, p, diag, shift, n, jac_end, part);
    ANN_LAUNCH_CHECK();
}

void launch_pcg_pap(const cplx<double>* p, const cplx<double>* Ap, int64_t n, double* part, hipStream_t stream) {
    cg_pap_kernel<<<kCgParts, 256, 0, stream>>>(p, Ap, n, part);
    ANN_LAUNCH_CHECK();
}

void launch_pcg_step(cplx<double>* x, cplx<double>* r, cplx<double>* z, const cplx<double>* p,
                    const cplx<double>* Ap, const double* diag, double shift, int64_t n, int64_t jac_end,
                    const double* rzpart, const double* apart, int a_parts, double* newpart, const int* done,
                    hipStream_t stream) {
    if (a_parts > 256) throw std::invalid_argument("too many partial sums");
    cg_step_kernel<<<kCgParts, 256, 0, stream>>>(x, r, z, p, Ap, diag, shift, n, jac_end, rzpart, kCgParts, apart,
                                                 a_parts, newpart, done);
    ANN_LAUNCH_CHECK();
}

void launch_pcg_dir(cplx<double>* p, const cplx<double>* z, int64_t n, const double* newpart, const double* oldpart,
                   const int* done, hipStream_t stream) {
    cg_dir_kernel<<<kCgParts, 256, 0, stream>>>(p, z, n, newpart, oldpart, kCgParts, done);
    ANN_LAUNCH_CHECK();
}

}  // namespace ann
