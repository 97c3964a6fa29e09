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

// ---------------------------------------- complex fp64 tiles on matrix cores
// A 64 x 64 complex output tile per workgroup of 4 waves; wave (wr, wc) owns the
// 32 x 32 quarter at rows 32 wr, columns 32 wc as 2 x 2 tiles of
// v_mfma_f64_16x16x4f64 with real and imaginary accumulators apart.  Operands come
// from LDS stages of 16 k values, each four [16][64] fp64 arrays: the real and
// imaginary parts of A[i][k] (stored [k][i]) and of B[k][j] (stored [k][j]), so that
// the tile accumulates the plain complex product D += A B (a loader stores conjugates
// and weights as its product needs them).  f64 16x16x4 operand layout: lane l gives
// A[row l%16][k l/16] and B[k l/16][column l%16]; result item r of lane l is row
// l/16 + 4 r, column l%16.
typedef double double4_t __attribute__((ext_vector_type(4)));
constexpr int kCT = 64;        // output tile
constexpr int kCS = 16;        // k values per LDS stage
constexpr int kCStage = 4 * kCS * kCT;   // doubles per stage (Ar, Ai, Br, Bi)

struct CTileAcc {
    double4_t re[2][2], im[2][2];
};

__device__ inline void ctile_zero(CTileAcc& c) {
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int u = 0; u < 2; u++) c.re[t][u] = c.im[t][u] = {0.0, 0.0, 0.0, 0.0};
}

// the 16 k values of one stage into the wave's quarter
__device__ inline void ctile_stage(const double* st, int lane, int wr, int wc, CTileAcc& c) {
    const double* Ar = st;
    const double* Ai = st + kCS * kCT;
    const double* Br = st + 2 * kCS * kCT;
    const double* Bi = st + 3 * kCS * kCT;
#pragma unroll
    for (int q = 0; q < kCS / 4; q++) {
        const int kk = 4 * q + (lane >> 4);
        double ar[2], ai[2], br[2], bi[2];
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int i = wr * 32 + 16 * t + (lane & 15);
            ar[t] = Ar[kk * kCT + i];
            ai[t] = Ai[kk * kCT + i];
            const int j = wc * 32 + 16 * t + (lane & 15);
            br[t] = Br[kk * kCT + j];
            bi[t] = Bi[kk * kCT + j];
        }
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
            for (int u = 0; u < 2; u++) {
                c.re[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[t], br[u], c.re[t][u], 0, 0, 0);
                c.re[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[t], bi[u], c.re[t][u], 0, 0, 0);
                c.im[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[t], bi[u], c.im[t][u], 0, 0, 0);
                c.im[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[t], br[u], c.im[t][u], 0, 0, 0);
            }
    }
}

// tile-local (row, column) of result item r of sub-tile (t, u) of the wave's quarter
__device__ inline int ctile_row(int lane, int wr, int t, int r) { return wr * 32 + 16 * t + (lane >> 4) + 4 * r; }
__device__ inline int ctile_col(int lane, int wc, int u) { return wc * 32 + 16 * u + (lane & 15); }

// (ti, tj), ti <= tj, of upper-triangular tile pair t (tj(tj+1)/2 <= t < (tj+1)(tj+2)/2)
__device__ inline void upper_pair(int64_t t, int& ti, int& tj) {
    int j = int((sqrt(8.0 * double(t) + 1.0) - 1.0) * 0.5);
    while (int64_t(j) * (j + 1) / 2 > t) j--;
    while (int64_t(j + 1) * (j + 2) / 2 <= t) j++;
    tj = j;
    ti = int(t - int64_t(j) * (j + 1) / 2);
}

// XCD-aware order: consecutive workgroups go to the 8 XCDs in turn, so logical tile t
// is given to workgroups of one XCD in runs (tiles that share operands share an L2)
__device__ inline int64_t xcd_tile(int64_t b, int64_t n) {
    const int64_t per = (n + 7) / 8;
    return (b % 8) * per + b / 8;
}

// ------------------------------------------------------------------ Gram
// S[k][k'] = sum_s w_s conj(O_sk) O_sk' for the upper-triangular tile pairs, mirrored
// into the lower triangle; samples are the k of the product, 16 per stage, staged in
// LDS double-buffered with the next stage's global loads in flight during the MFMAs.
template <typename T>
__global__ __launch_bounds__(256) void gram_mfma_kernel(const cplx<T>* O, const double* w, int64_t count, int P,
                                                        int64_t pairs, cplx<double>* S) {
    extern __shared__ double lds[];   // [2][kCStage]
    const int64_t t = xcd_tile(blockIdx.x, pairs);
    if (t >= pairs) return;
    int ti, tj;
    upper_pair(t, ti, tj);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int row0 = ti * kCT, col0 = tj * kCT;
    // loader: thread -> sample (tid / 16) of the stage, 4 consecutive parameters (tid % 16) * 4
    const int ls = threadIdx.x >> 4, lc = (threadIdx.x & 15) * 4;
    cplx<T> ra[4], ca[4];
    double ws = 0.0;
    auto fetch = [&](int64_t s0) {
        const int64_t s = s0 + ls;
        const bool ok = s < count;
        ws = ok ? w[s] : 0.0;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            ra[e] = ok && row0 + lc + e < P ? O[s * P + row0 + lc + e] : cplx<T>{T(0), T(0)};
            ca[e] = ok && col0 + lc + e < P ? O[s * P + col0 + lc + e] : cplx<T>{T(0), T(0)};
        }
    };
    auto stash = [&](double* st) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int o = ls * kCT + lc + e;
            st[o] = ws * double(ra[e].re);                      // A = w conj(O) of the rows
            st[kCS * kCT + o] = -ws * double(ra[e].im);
            st[2 * kCS * kCT + o] = double(ca[e].re);           // B = O of the columns
            st[3 * kCS * kCT + o] = double(ca[e].im);
        }
    };
    CTileAcc acc;
    ctile_zero(acc);
    const int64_t stages = (count + kCS - 1) / kCS;
    fetch(0);
    stash(lds);
    __syncthreads();
    for (int64_t g = 0; g < stages; g++) {
        if (g + 1 < stages) fetch((g + 1) * kCS);
        ctile_stage(lds + (g & 1) * kCStage, lane, wr, wc, acc);
        if (g + 1 < stages) stash(lds + ((g + 1) & 1) * kCStage);
        __syncthreads();
    }
#pragma unroll
    for (int tt = 0; tt < 2; tt++)
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gr = row0 + ctile_row(lane, wr, tt, r), gc = col0 + ctile_col(lane, wc, u);
                if (gr < P && gc < P) {
                    S[int64_t(gr) * P + gc] = {acc.re[tt][u][r], acc.im[tt][u][r]};
                    if (ti != tj) S[int64_t(gc) * P + gr] = {acc.re[tt][u][r], -acc.im[tt][u][r]};
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
// Right-looking blocked Cholesky of the lower triangle with 64-column panels: the
// diagonal block is factored (and its triangular inverse formed) in LDS by one
// workgroup; the panel below it is A L^-H, a matrix-core product with that inverse;
// the trailing lower tiles take A -= L L^H on matrix cores (K = 64).
constexpr int kNB = 64;           // panel width (the factor's block size)
constexpr int kTS = 32;           // block size of the triangular solves
constexpr int kDL = kNB + 1;      // padded LDS row of the diagonal kernel

// factor the diagonal block A[k:k+nb, k:k+nb] (lower) and write L^-1 to Linv [kNB][kNB]
__global__ __launch_bounds__(256) void potrf_inv_kernel(cplx<double>* A, int P, int k, int nb, int* info,
                                                        cplx<double>* Linv) {
    extern __shared__ double dl[];
    double* Lr = dl;                  // [kNB][kDL]
    double* Li = dl + kNB * kDL;
    double* Xr = dl + 2 * kNB * kDL;
    double* Xi = dl + 3 * kNB * kDL;
    for (int e = threadIdx.x; e < nb * nb; e += 256) {
        const int i = e / nb, j = e % nb;
        const cplx<double> a = A[int64_t(k + i) * P + k + j];
        Lr[i * kDL + j] = a.re;
        Li[i * kDL + j] = a.im;
    }
    __syncthreads();
    for (int j = 0; j < nb; j++) {
        if (threadIdx.x == 0) {
            const double d = Lr[j * kDL + j];
            if (!(d > 0.0)) {
                atomicCAS(info, 0, k + j + 1);
                Lr[j * kDL + j] = 1.0;
            } else {
                Lr[j * kDL + j] = sqrt(d);
            }
            Li[j * kDL + j] = 0.0;
        }
        __syncthreads();
        const double inv = 1.0 / Lr[j * kDL + j];
        for (int i = j + 1 + threadIdx.x; i < nb; i += 256) {
            Lr[i * kDL + j] *= inv;
            Li[i * kDL + j] *= inv;
        }
        __syncthreads();
        // trailing update A[i][l] -= L[i][j] conj(L[l][j]), j < l <= i
        const int rem = nb - j - 1;
        for (int e = threadIdx.x; e < rem * rem; e += 256) {
            const int i = j + 1 + e / rem, l = j + 1 + e % rem;
            if (l <= i) {
                const double ar = Lr[i * kDL + j], ai = Li[i * kDL + j], br = Lr[l * kDL + j], bi = -Li[l * kDL + j];
                Lr[i * kDL + l] -= ar * br - ai * bi;
                Li[i * kDL + l] -= ar * bi + ai * br;
            }
        }
        __syncthreads();
    }
    // X = L^-1 by forward substitution, four threads per column (partial sums over m
    // interleaved by 4, combined by shuffles within the quad)
    const int col = threadIdx.x >> 2, part = threadIdx.x & 3;
    for (int i = 0; i < nb; i++) {
        double sr = 0.0, si = 0.0;
        if (col < nb && i >= col)
            for (int m = col + part; m < i; m += 4) {
                const double lr = Lr[i * kDL + m], li = Li[i * kDL + m], xr = Xr[m * kDL + col], xi = Xi[m * kDL + col];
                sr += lr * xr - li * xi;
                si += lr * xi + li * xr;
            }
        sr += __shfl_xor(sr, 1);
        si += __shfl_xor(si, 1);
        sr += __shfl_xor(sr, 2);
        si += __shfl_xor(si, 2);
        if (col < nb && part == 0) {
            const double d = Lr[i * kDL + i];   // real, positive
            if (i < col) {
                Xr[i * kDL + col] = 0.0;
                Xi[i * kDL + col] = 0.0;
            } else {
                Xr[i * kDL + col] = ((i == col ? 1.0 : 0.0) - sr) / d;
                Xi[i * kDL + col] = -si / d;
            }
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < nb * nb; e += 256) {
        const int i = e / nb, j = e % nb;
        if (j <= i) A[int64_t(k + i) * P + k + j] = {Lr[i * kDL + j], Li[i * kDL + j]};
    }
    for (int e = threadIdx.x; e < kNB * kNB; e += 256) {
        const int i = e / kNB, j = e % kNB;
        Linv[e] = i < nb && j < nb ? cplx<double>{Xr[i * kDL + j], Xi[i * kDL + j]} : cplx<double>{0.0, 0.0};
    }
}

// panel rows r0 + 64 blockIdx.x ..: A[i, k:k+nb] <- A[i, k:k+nb] L^-H (in place: every
// stage is read before the tile is written)
__global__ __launch_bounds__(256) void panel_mfma_kernel(cplx<double>* A, int P, int k, int nb, int r0,
                                                         const cplx<double>* Linv) {
    extern __shared__ double lds[];   // [kNB / kCS][kCStage]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int rb = r0 + blockIdx.x * kCT;
    for (int e = threadIdx.x; e < kNB * kCT; e += 256) {
        const int i = e / kNB, m = e % kNB;   // A operand [m][i] = A[rb + i][k + m] (rows read contiguously)
        double* st = lds + (m / kCS) * kCStage;
        const int o = (m % kCS) * kCT + i;
        cplx<double> a = {0.0, 0.0};
        if (m < nb && rb + i < P) a = A[int64_t(rb + i) * P + k + m];
        st[o] = a.re;
        st[kCS * kCT + o] = a.im;
        // B operand [m][j] = conj(Linv[j][m])
        const cplx<double> b = Linv[i * kNB + m];
        st[2 * kCS * kCT + o] = b.re;
        st[3 * kCS * kCT + o] = -b.im;
    }
    __syncthreads();
    CTileAcc acc;
    ctile_zero(acc);
    for (int g = 0; g < kNB / kCS; g++) ctile_stage(lds + g * kCStage, lane, wr, wc, acc);
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gr = rb + ctile_row(lane, wr, t, r), j = ctile_col(lane, wc, u);
                if (gr < P && j < nb) A[int64_t(gr) * P + k + j] = {acc.re[t][u][r], acc.im[t][u][r]};
            }
}

// trailing update of the lower 64 x 64 tiles (bi >= bj) from s0:
// A[i][j] -= sum_m A[i][k+m] conj(A[j][k+m])
__global__ __launch_bounds__(256) void herk_mfma_kernel(cplx<double>* A, int P, int k, int nb, int s0,
                                                        int64_t pairs) {
    extern __shared__ double lds[];   // [kNB / kCS][kCStage]
    const int64_t t = xcd_tile(blockIdx.x, pairs);
    if (t >= pairs) return;
    int bj, bi;
    upper_pair(t, bj, bi);   // bj <= bi
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int r0 = s0 + bi * kCT, c0 = s0 + bj * kCT;
    for (int e = threadIdx.x; e < kNB * kCT; e += 256) {
        const int i = e / kNB, m = e % kNB;   // rows of the panel read contiguously
        double* st = lds + (m / kCS) * kCStage;
        const int o = (m % kCS) * kCT + i;
        cplx<double> x = {0.0, 0.0}, y = {0.0, 0.0};
        if (m < nb && r0 + i < P) x = A[int64_t(r0 + i) * P + k + m];
        if (m < nb && c0 + i < P) y = A[int64_t(c0 + i) * P + k + m];
        st[o] = x.re;
        st[kCS * kCT + o] = x.im;
        st[2 * kCS * kCT + o] = y.re;       // B [m][j] = conj(A[c0 + j][k + m])
        st[3 * kCS * kCT + o] = -y.im;
    }
    __syncthreads();
    CTileAcc acc;
    ctile_zero(acc);
    for (int g = 0; g < kNB / kCS; g++) ctile_stage(lds + g * kCStage, lane, wr, wc, acc);
#pragma unroll
    for (int tt = 0; tt < 2; tt++)
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gi = r0 + ctile_row(lane, wr, tt, r), gj = c0 + ctile_col(lane, wc, u);
                if (gi < P && gj <= gi) {
                    cplx<double>& dst = A[int64_t(gi) * P + gj];
                    dst.re -= acc.re[tt][u][r];
                    dst.im -= acc.im[tt][u][r];
                }
            }
}

// forward: solve L_kk y_k = b_k in place (one block), then b_i -= L_ik y_k for i > block
__global__ __launch_bounds__(256) void trsv_lower_block_kernel(const cplx<double>* L, int P, int k, int nb,
                                                               cplx<double>* x) {
    __shared__ double yr[kTS], yi[kTS];
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
    __shared__ double zr[kTS], zi[kTS];
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
    const int64_t nt = (P + kCT - 1) / kCT;
    const int64_t pairs = nt * (nt + 1) / 2;
    const size_t lds_bytes = sizeof(double) * 2 * kCStage;   // 64 KiB
    ANN_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(gram_mfma_kernel<T>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_bytes)));
    const int64_t grid = (pairs + 7) / 8 * 8;
    gram_mfma_kernel<T><<<unsigned(grid), 256, lds_bytes, stream>>>(O, w, count, P, pairs, S);
    ANN_LAUNCH_CHECK();
}
template void launch_gram<float>(const cplx<float>*, const double*, int64_t, int, cplx<double>*, hipStream_t);
template void launch_gram<double>(const cplx<double>*, const double*, int64_t, int, cplx<double>*, hipStream_t);

void launch_dense_matvec(const cplx<double>* S, const cplx<double>* x, int P, double shift, cplx<double>* y,
                         hipStream_t stream) {
    dense_matvec_kernel<<<unsigned(P), 256, 0, stream>>>(S, x, P, shift, y);
    ANN_LAUNCH_CHECK();
}

size_t cholesky_work_elems() { return size_t(kNB) * kNB; }

void cholesky_lower(cplx<double>* A, int P, int* d_info, cplx<double>* work, hipStream_t stream) {
    const size_t diag_lds = sizeof(double) * 4 * kNB * kDL;
    const size_t tile_lds = sizeof(double) * (kNB / kCS) * kCStage;   // 128 KiB
    ANN_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(potrf_inv_kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(diag_lds)));
    ANN_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(panel_mfma_kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(tile_lds)));
    ANN_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(herk_mfma_kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(tile_lds)));
    for (int k = 0; k < P; k += kNB) {
        const int nb = P - k < kNB ? P - k : kNB;
        potrf_inv_kernel<<<1, 256, diag_lds, stream>>>(A, P, k, nb, d_info, work);
        ANN_LAUNCH_CHECK();
        const int s0 = k + nb, rest = P - s0;
        if (rest <= 0) break;
        panel_mfma_kernel<<<unsigned((rest + kCT - 1) / kCT), 256, tile_lds, stream>>>(A, P, k, nb, s0, work);
        ANN_LAUNCH_CHECK();
        const int64_t nt = (rest + kCT - 1) / kCT;
        const int64_t pairs = nt * (nt + 1) / 2;
        herk_mfma_kernel<<<unsigned((pairs + 7) / 8 * 8), 256, tile_lds, stream>>>(A, P, k, nb, s0, pairs);
        ANN_LAUNCH_CHECK();
    }
}

void cholesky_solve_inplace(const cplx<double>* L, int P, cplx<double>* x, hipStream_t stream) {
    for (int k = 0; k < P; k += kTS) {
        const int nb = P - k < kTS ? P - k : kTS;
        trsv_lower_block_kernel<<<1, 64, 0, stream>>>(L, P, k, nb, x);
        ANN_LAUNCH_CHECK();
        const int rest = P - k - nb;
        if (rest > 0) {
            gemv_lower_update_kernel<<<unsigned((rest + 255) / 256), 256, 0, stream>>>(L, P, k, nb, x);
            ANN_LAUNCH_CHECK();
        }
    }
    const int last = ((P - 1) / kTS) * kTS;
    for (int k = last; k >= 0; k -= kTS) {
        const int nb = P - k < kTS ? P - k : kTS;
        trsv_upper_block_kernel<<<1, 64, 0, stream>>>(L, P, k, nb, x);
        ANN_LAUNCH_CHECK();
        if (k > 0) {
            gemv_upper_update_kernel<<<unsigned((k + 255) / 256), 256, 0, stream>>>(L, P, k, nb, x);
            ANN_LAUNCH_CHECK();
        }
    }
}

}  // namespace ann
