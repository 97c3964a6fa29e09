// linalg.hip -- fp64 reductions, the dense Hermitian Gram S = O^H diag(w) O, and
// the fp64 Cholesky / triangular solves of the shifted SR system (S + lambda I) x = F.
//
// The reference builds S with one complex atomic per (sample, k, k') and leaves
// the solve to the caller (source/network_functions/TDVP.cu.template:181-302);
// here S is a tiled GEMM with fp64 accumulation and the solve stays on the GPU.
#include "kernels.hpp"

namespace ann {

namespace {

constexpr int kRed = 1024;

__global__ __launch_bounds__(kRed) void weighted_sums_kernel(const cplx<double>* e, const double* w,
                                                             int64_t n, double* out) {
    __shared__ double s0[kRed], s1[kRed], s2[kRed], s3[kRed];
    double a = 0, b = 0, c = 0, d = 0;
    for (int64_t i = threadIdx.x; i < n; i += kRed) {
        const double wi = w[i];
        const cplx<double> ei = e[i];
        a += wi * ei.re;
        b += wi * ei.im;
        c += wi * (ei.re * ei.re + ei.im * ei.im);
        d += wi;
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = b;
    s2[threadIdx.x] = c;
    s3[threadIdx.x] = d;
    __syncthreads();
    for (int o = kRed / 2; o > 0; o >>= 1) {
        if (int(threadIdx.x) < o) {
            s0[threadIdx.x] += s0[threadIdx.x + o];
            s1[threadIdx.x] += s1[threadIdx.x + o];
            s2[threadIdx.x] += s2[threadIdx.x + o];
            s3[threadIdx.x] += s3[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s0[0];
        out[1] = s1[0];
        out[2] = s2[0];
        out[3] = s3[0];
    }
}

// ------------------------------------------------------------------ Gram
constexpr int kGT = 64;   // output tile
constexpr int kGK = 16;   // samples per LDS stage

template <typename T>
__global__ __launch_bounds__(256) void gram_kernel(const cplx<T>* O, const double* w, int64_t count, int P,
                                                   cplx<double>* S) {
    // upper-triangular tile pair (ti <= tj) from the linear block index
    const int64_t t = blockIdx.x;
    int tj = int((sqrt(8.0 * double(t) + 1.0) - 1.0) * 0.5);
    while (int64_t(tj) * (tj + 1) / 2 > t) tj--;
    while (int64_t(tj + 1) * (tj + 2) / 2 <= t) tj++;
    const int ti = int(t - int64_t(tj) * (tj + 1) / 2);
    __shared__ double Ar[kGK][kGT], Ai[kGK][kGT], Br[kGK][kGT], Bi[kGK][kGT];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double cr[4][4], ci[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) cr[a][b] = ci[a][b] = 0.0;
    const int row0 = ti * kGT, col0 = tj * kGT;
    for (int64_t s0 = 0; s0 < count; s0 += kGK) {
        for (int e = threadIdx.x; e < kGK * kGT; e += 256) {
            const int ks = e / kGT, c = e % kGT;
            const int64_t s = s0 + ks;
            double ar = 0, ai = 0, br = 0, bi = 0;
            if (s < count) {
                const double ws = w[s];
                if (row0 + c < P) {
                    const cplx<T> o = O[s * P + row0 + c];
                    ar = ws * double(o.re);
                    ai = -ws * double(o.im);   // w * conj(O_k)
                }
                if (col0 + c < P) {
                    const cplx<T> o = O[s * P + col0 + c];
                    br = double(o.re);
                    bi = double(o.im);
                }
            }
            Ar[ks][c] = ar;
            Ai[ks][c] = ai;
            Br[ks][c] = br;
            Bi[ks][c] = bi;
        }
        __syncthreads();
#pragma unroll 4
        for (int ks = 0; ks < kGK; ks++) {
            double xr[4], xi[4], yr[4], yi[4];
#pragma unroll
            for (int a = 0; a < 4; a++) {
                xr[a] = Ar[ks][ty * 4 + a];
                xi[a] = Ai[ks][ty * 4 + a];
                yr[a] = Br[ks][tx * 4 + a];
                yi[a] = Bi[ks][tx * 4 + a];
            }
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    cr[a][b] += xr[a] * yr[b] - xi[a] * yi[b];
                    ci[a][b] += xr[a] * yi[b] + xi[a] * yr[b];
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int r = row0 + ty * 4 + a, c = col0 + tx * 4 + b;
            if (r < P && c < P) {
                S[int64_t(r) * P + c] = {cr[a][b], ci[a][b]};
                if (ti != tj) S[int64_t(c) * P + r] = {cr[a][b], -ci[a][b]};
            }
        }
}

__global__ __launch_bounds__(256) void dense_matvec_kernel(const cplx<double>* S, const cplx<double>* x,
                                                           int P, double shift, cplx<double>* y) {
    const int r = blockIdx.x;
    __shared__ double sr[256], si[256];
    double re = 0, im = 0;
    for (int k = threadIdx.x; k < P; k += 256) {
        const cplx<double> a = S[int64_t(r) * P + k], b = x[k];
        re += a.re * b.re - a.im * b.im;
        im += a.re * b.im + a.im * b.re;
    }
    sr[threadIdx.x] = re;
    si[threadIdx.x] = im;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (int(threadIdx.x) < o) {
            sr[threadIdx.x] += sr[threadIdx.x + o];
            si[threadIdx.x] += si[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) y[r] = {sr[0] + shift * x[r].re, si[0] + shift * x[r].im};
}

// -------------------------------------------------------------- Cholesky
constexpr int kNB = 32;

// factor the diagonal block A[k:k+nb, k:k+nb] (lower) in LDS
__global__ __launch_bounds__(256) void potrf_diag_kernel(cplx<double>* A, int P, int k, int nb, int* info) {
    __shared__ double Lr[kNB][kNB + 1], Li[kNB][kNB + 1];
    for (int e = threadIdx.x; e < nb * nb; e += 256) {
        const int i = e / nb, j = e % nb;
        const cplx<double> a = A[int64_t(k + i) * P + k + j];
        Lr[i][j] = a.re;
        Li[i][j] = a.im;
    }
    __syncthreads();
    for (int j = 0; j < nb; j++) {
        if (threadIdx.x == 0) {
            const double d = Lr[j][j];
            if (!(d > 0.0)) {
                *info = k + j + 1;
                Lr[j][j] = 1.0;
            } else {
                Lr[j][j] = sqrt(d);
            }
            Li[j][j] = 0.0;
        }
        __syncthreads();
        const double inv = 1.0 / Lr[j][j];
        for (int i = j + 1 + threadIdx.x; i < nb; i += 256) {
            Lr[i][j] *= inv;
            Li[i][j] *= inv;
        }
        __syncthreads();
        // trailing update A[i][l] -= L[i][j] conj(L[l][j]), j < l <= i
        const int rem = nb - j - 1;
        for (int e = threadIdx.x; e < rem * rem; e += 256) {
            const int i = j + 1 + e / rem, l = j + 1 + e % rem;
            if (l <= i) {
                const double ar = Lr[i][j], ai = Li[i][j], br = Lr[l][j], bi = -Li[l][j];
                Lr[i][l] -= ar * br - ai * bi;
                Li[i][l] -= ar * bi + ai * br;
            }
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < nb * nb; e += 256) {
        const int i = e / nb, j = e % nb;
        if (j <= i) A[int64_t(k + i) * P + k + j] = {Lr[i][j], Li[i][j]};
    }
}

// rows i >= k+nb: solve X L_kk^H = A[i, k:k+nb]
__global__ __launch_bounds__(256) void trsm_panel_kernel(cplx<double>* A, int P, int k, int nb) {
    __shared__ double Lr[kNB][kNB + 1], Li[kNB][kNB + 1];
    for (int e = threadIdx.x; e < nb * nb; e += 256) {
        const int i = e / nb, j = e % nb;
        const cplx<double> a = A[int64_t(k + i) * P + k + j];
        Lr[i][j] = a.re;
        Li[i][j] = a.im;
    }
    __syncthreads();
    const int row = k + nb + blockIdx.x * 256 + threadIdx.x;
    if (row >= P) return;
    double xr[kNB], xi[kNB];
#pragma unroll
    for (int j = 0; j < kNB; j++) {
        if (j < nb) {
            const cplx<double> a = A[int64_t(row) * P + k + j];
            double r = a.re, im = a.im;
            for (int m = 0; m < j; m++) {
                // x_m * conj(L[j][m])
                r -= xr[m] * Lr[j][m] + xi[m] * Li[j][m];
                im -= xi[m] * Lr[j][m] - xr[m] * Li[j][m];
            }
            const double inv = 1.0 / Lr[j][j];
            xr[j] = r * inv;
            xi[j] = im * inv;
        }
    }
#pragma unroll
    for (int j = 0; j < kNB; j++)
        if (j < nb) A[int64_t(row) * P + k + j] = {xr[j], xi[j]};
}

// trailing update A[i][j] -= sum_m A[i][k+m] conj(A[j][k+m]) on lower tiles
__global__ __launch_bounds__(256) void herk_update_kernel(cplx<double>* A, int P, int k, int nb, int start) {
    const int64_t t = blockIdx.x;
    int bi = int((sqrt(8.0 * double(t) + 1.0) - 1.0) * 0.5);
    while (int64_t(bi) * (bi + 1) / 2 > t) bi--;
    while (int64_t(bi + 1) * (bi + 2) / 2 <= t) bi++;
    const int bj = int(t - int64_t(bi) * (bi + 1) / 2);   // bj <= bi
    const int r0 = start + bi * 32, c0 = start + bj * 32;
    __shared__ double Xr[32][kNB + 1], Xi[32][kNB + 1], Yr[32][kNB + 1], Yi[32][kNB + 1];
    for (int e = threadIdx.x; e < 32 * nb; e += 256) {
        const int r = e / nb, m = e % nb;
        cplx<double> x = {0, 0}, y = {0, 0};
        if (r0 + r < P) x = A[int64_t(r0 + r) * P + k + m];
        if (c0 + r < P) y = A[int64_t(c0 + r) * P + k + m];
        Xr[r][m] = x.re;
        Xi[r][m] = x.im;
        Yr[r][m] = y.re;
        Yi[r][m] = y.im;
    }
    __syncthreads();
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8 threads, 4 rows each
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const int r = ty * 4 + a;
        const int gi = r0 + r, gj = c0 + tx;
        if (gi < P && gj < P && gj <= gi) {
            double sr = 0, si = 0;
            for (int m = 0; m < nb; m++) {
                // x * conj(y)
                sr += Xr[r][m] * Yr[tx][m] + Xi[r][m] * Yi[tx][m];
                si += Xi[r][m] * Yr[tx][m] - Xr[r][m] * Yi[tx][m];
            }
            cplx<double>& dst = A[int64_t(gi) * P + gj];
            dst.re -= sr;
            dst.im -= si;
        }
    }
}

// forward: solve L_kk y_k = b_k in place (one block), then b_i -= L_ik y_k for i > block
__global__ __launch_bounds__(256) void trsv_lower_block_kernel(const cplx<double>* L, int P, int k, int nb,
                                                               cplx<double>* x) {
    __shared__ double yr[kNB], yi[kNB];
    if (threadIdx.x == 0) {
        for (int j = 0; j < nb; j++) {
            double r = x[k + j].re, im = x[k + j].im;
            for (int m = 0; m < j; m++) {
                const cplx<double> l = L[int64_t(k + j) * P + k + m];
                r -= l.re * yr[m] - l.im * yi[m];
                im -= l.re * yi[m] + l.im * yr[m];
            }
            const double d = L[int64_t(k + j) * P + k + j].re;
            yr[j] = r / d;
            yi[j] = im / d;
            x[k + j] = {yr[j], yi[j]};
        }
    }
}
__global__ __launch_bounds__(256) void gemv_lower_update_kernel(const cplx<double>* L, int P, int k, int nb,
                                                                cplx<double>* x) {
    const int i = k + nb + blockIdx.x * 256 + threadIdx.x;
    if (i >= P) return;
    double r = 0, im = 0;
    for (int m = 0; m < nb; m++) {
        const cplx<double> l = L[int64_t(i) * P + k + m];
        const cplx<double> y = x[k + m];
        r += l.re * y.re - l.im * y.im;
        im += l.re * y.im + l.im * y.re;
    }
    x[i].re -= r;
    x[i].im -= im;
}
// backward: solve L_kk^H x_k = y_k, then y_i -= sum_m conj(L[k+m][i]) x_{k+m} for i < k
__global__ __launch_bounds__(256) void trsv_upper_block_kernel(const cplx<double>* L, int P, int k, int nb,
                                                               cplx<double>* x) {
    __shared__ double zr[kNB], zi[kNB];
    if (threadIdx.x == 0) {
        for (int j = nb - 1; j >= 0; j--) {
            double r = x[k + j].re, im = x[k + j].im;
            for (int m = j + 1; m < nb; m++) {
                const cplx<double> l = L[int64_t(k + m) * P + k + j];   // conj(L[m][j]) * z_m
                r -= l.re * zr[m] + l.im * zi[m];
                im -= l.re * zi[m] - l.im * zr[m];
            }
            const double d = L[int64_t(k + j) * P + k + j].re;
            zr[j] = r / d;
            zi[j] = im / d;
            x[k + j] = {zr[j], zi[j]};
        }
    }
}
__global__ __launch_bounds__(256) void gemv_upper_update_kernel(const cplx<double>* L, int P, int k, int nb,
                                                                cplx<double>* x) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= k) return;
    double r = 0, im = 0;
    for (int m = 0; m < nb; m++) {
        const cplx<double> l = L[int64_t(k + m) * P + i];
        const cplx<double> z = x[k + m];
        r += l.re * z.re + l.im * z.im;
        im += l.re * z.im - l.im * z.re;
    }
    x[i].re -= r;
    x[i].im -= im;
}

}  // namespace

void launch_weighted_sums(const cplx<double>* e_loc, const double* w, int64_t count, double* out4,
                          hipStream_t stream) {
    weighted_sums_kernel<<<1, kRed, 0, stream>>>(e_loc, w, count, out4);
    ANN_LAUNCH_CHECK();
}

template <typename T>
void launch_gram(const cplx<T>* O, const double* w, int64_t count, int P, cplx<double>* S, hipStream_t stream) {
    const int64_t nt = (P + kGT - 1) / kGT;
    const int64_t blocks = nt * (nt + 1) / 2;
    gram_kernel<T><<<unsigned(blocks), 256, 0, stream>>>(O, w, count, P, S);
    ANN_LAUNCH_CHECK();
}
template void launch_gram<float>(const cplx<float>*, const double*, int64_t, int, cplx<double>*, hipStream_t);
template void launch_gram<double>(const cplx<double>*, const double*, int64_t, int, cplx<double>*, hipStream_t);

void launch_dense_matvec(const cplx<double>* S, const cplx<double>* x, int P, double shift, cplx<double>* y,
                         hipStream_t stream) {
    dense_matvec_kernel<<<unsigned(P), 256, 0, stream>>>(S, x, P, shift, y);
    ANN_LAUNCH_CHECK();
}

void cholesky_lower(cplx<double>* A, int P, int* d_info, hipStream_t stream) {
    for (int k = 0; k < P; k += kNB) {
        const int nb = P - k < kNB ? P - k : kNB;
        potrf_diag_kernel<<<1, 256, 0, stream>>>(A, P, k, nb, d_info);
        ANN_LAUNCH_CHECK();
        const int rest = P - k - nb;
        if (rest <= 0) break;
        trsm_panel_kernel<<<unsigned((rest + 255) / 256), 256, 0, stream>>>(A, P, k, nb);
        ANN_LAUNCH_CHECK();
        const int64_t nt = (rest + 31) / 32;
        herk_update_kernel<<<unsigned(nt * (nt + 1) / 2), 256, 0, stream>>>(A, P, k, nb, k + nb);
        ANN_LAUNCH_CHECK();
    }
}

void cholesky_solve_inplace(const cplx<double>* L, int P, cplx<double>* x, hipStream_t stream) {
    for (int k = 0; k < P; k += kNB) {
        const int nb = P - k < kNB ? P - k : kNB;
        trsv_lower_block_kernel<<<1, 64, 0, stream>>>(L, P, k, nb, x);
        ANN_LAUNCH_CHECK();
        const int rest = P - k - nb;
        if (rest > 0) {
            gemv_lower_update_kernel<<<unsigned((rest + 255) / 256), 256, 0, stream>>>(L, P, k, nb, x);
            ANN_LAUNCH_CHECK();
        }
    }
    const int last = ((P - 1) / kNB) * kNB;
    for (int k = last; k >= 0; k -= kNB) {
        const int nb = P - k < kNB ? P - k : kNB;
        trsv_upper_block_kernel<<<1, 64, 0, stream>>>(L, P, k, nb, x);
        ANN_LAUNCH_CHECK();
        if (k > 0) {
            gemv_upper_update_kernel<<<unsigned((k + 255) / 256), 256, 0, stream>>>(L, P, k, nb, x);
            ANN_LAUNCH_CHECK();
        }
    }
}

}  // namespace ann
