// kfac.hip -- Kronecker-factored preconditioner of (S + shift) for the dense nets.
//
// The W part of O is O_{(v,j)}(s) = s_v g_j(s) (v = the site of a connection,
// j = hidden unit), so S restricted to W is approximately C (x) D with the
// centred covariances C_vv' = Cov(s_v, s_v') (real, N x N) and
// D_jj' = Cov(g_j^*, g_j') (Hermitian, M x M).  The preconditioner applies
//   Z = (C + gc I)^-1 X ((D + gd I)^-1)^T,   X[v][j] = r at parameter (v, j),
// with the damping split gc = sqrt(shift) pi, gd = sqrt(shift) / pi,
// pi^2 = (tr C / N) / (tr D / M); D is kept block-diagonal in blocks of 64
// hidden units so that every block (and C) is inverted by Gauss-Jordan inside
// one workgroup's LDS.  The biases a, b keep the Jacobi (diagonal) preconditioner.
// It only has to be Hermitian positive definite and fixed during a solve: the
// CG solution is that of (S + shift) x = F whatever it is.
#include "ops.hpp"

#include <algorithm>
#include <stdexcept>

namespace ann {

namespace {

constexpr int kDBlock = 64;   // hidden units per diagonal block of D
constexpr int kGroup = 32;    // samples per workgroup of the moment kernels

// raw spin moments of one group of 32 samples: part[gi] = (sum w s_v s_u [N*N],
// sum w s_v [N], sum w); the group's bits and weights are staged in LDS
__global__ __launch_bounds__(256) void kfac_moments_spin_kernel(const uint64_t* confs, int words, const double* w,
                                                                int64_t nsub, int N, double* part) {
    __shared__ uint64_t sc[kGroup][kMaxWords];
    __shared__ double sw[kGroup];
    const int gi = blockIdx.x;
    const int64_t s0 = int64_t(gi) * kGroup;
    for (int e = threadIdx.x; e < kGroup * kMaxWords; e += blockDim.x) {
        const int s = e / kMaxWords, k = e % kMaxWords;
        sc[s][k] = s0 + s < nsub && k < words ? confs[(s0 + s) * words + k] : 0;
    }
    for (int s = threadIdx.x; s < kGroup; s += blockDim.x) sw[s] = s0 + s < nsub ? w[s0 + s] : 0.0;
    __syncthreads();
    double* out = part + int64_t(gi) * (N * N + N + 1);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int v = wave; v < N; v += 4)
        for (int u = lane; u < N; u += 64) {
            double acc = 0.0;
            for (int s = 0; s < kGroup; s++) {
                const bool x = (sc[s][v >> 6] >> (v & 63)) & 1, y = (sc[s][u >> 6] >> (u & 63)) & 1;
                acc += x == y ? sw[s] : -sw[s];
            }
            out[v * N + u] = acc;
        }
    for (int e = threadIdx.x; e < N + 1; e += blockDim.x) {
        double acc = 0.0;
        if (e < N)
            for (int s = 0; s < kGroup; s++) acc += (sc[s][e >> 6] >> (e & 63)) & 1 ? sw[s] : -sw[s];
        else
            for (int s = 0; s < kGroup; s++) acc += sw[s];
        out[N * N + e] = acc;
    }
}

// raw g moments of one group of 32 samples and one block of 64 units:
// part[gi][b] = (sum w conj(g_j) g_k [64][64], sum w g_j [64])
__global__ __launch_bounds__(256) void kfac_moments_g_kernel(const cplx<float>* g32, const cplx<double>* g64,
                                                             int Mpad, const double* w, int64_t nsub, int M,
                                                             int nblk, cplx<double>* part) {
    __shared__ cplx<double> sg[kGroup][kDBlock];
    __shared__ double sw[kGroup];
    const int gi = blockIdx.x, b = blockIdx.y;
    const int64_t s0 = int64_t(gi) * kGroup;
    for (int e = threadIdx.x; e < kGroup * kDBlock; e += blockDim.x) {
        const int s = e / kDBlock, jj = e % kDBlock;
        const int j = b * kDBlock + jj;
        cplx<double> x = {0.0, 0.0};
        if (s0 + s < nsub && j < M) {
            const int64_t o = (s0 + s) * Mpad + j;
            x = g32 ? cplx<double>{double(g32[o].re), double(g32[o].im)} : g64[o];
        }
        sg[s][jj] = x;
    }
    for (int s = threadIdx.x; s < kGroup; s += blockDim.x) sw[s] = s0 + s < nsub ? w[s0 + s] : 0.0;
    __syncthreads();
    cplx<double>* out = part + (int64_t(gi) * nblk + b) * (kDBlock * kDBlock + kDBlock);
    for (int e = threadIdx.x; e < kDBlock * kDBlock + kDBlock; e += blockDim.x) {
        double re = 0.0, im = 0.0;
        if (e < kDBlock * kDBlock) {
            const int jj = e / kDBlock, kk = e % kDBlock;
            for (int s = 0; s < kGroup; s++) {
                const cplx<double> x = sg[s][jj], y = sg[s][kk];
                re += sw[s] * (x.re * y.re + x.im * y.im);   // conj(x) y
                im += sw[s] * (x.re * y.im - x.im * y.re);
            }
        } else {
            const int jj = e - kDBlock * kDBlock;
            for (int s = 0; s < kGroup; s++) {
                re += sw[s] * sg[s][jj].re;
                im += sw[s] * sg[s][jj].im;
            }
        }
        out[e] = {re, im};
    }
}

// sums of the group moments (group order) and centring:
// C = sum w s s^T / W - m m^T, D block b = sum w conj(g) g^T / W - conj(mg) mg^T
__global__ void kfac_reduce_kernel(const double* spart, const cplx<double>* gpart, int groups, int N, int M,
                                   int nblk, double* C, cplx<double>* Dblk) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t ns = int64_t(N) * N + N + 1, ng = int64_t(kDBlock) * kDBlock + kDBlock;
    double W = 0.0;
#pragma unroll 8
    for (int gi = 0; gi < groups; gi++) W += spart[gi * ns + ns - 1];
    const double iw = W > 0.0 ? 1.0 / W : 0.0;
    if (e < int64_t(N) * N) {
        const int v = int(e / N), u = int(e % N);
        double s2 = 0.0, mv = 0.0, mu = 0.0;
#pragma unroll 8
        for (int gi = 0; gi < groups; gi++) {
            s2 += spart[gi * ns + e];
            mv += spart[gi * ns + int64_t(N) * N + v];
            mu += spart[gi * ns + int64_t(N) * N + u];
        }
        C[e] = s2 * iw - (mv * iw) * (mu * iw);
    } else if (e < int64_t(N) * N + int64_t(nblk) * kDBlock * kDBlock) {
        const int64_t f = e - int64_t(N) * N;
        const int b = int(f / (kDBlock * kDBlock));
        const int jj = int(f % (kDBlock * kDBlock)) / kDBlock, kk = int(f % kDBlock);
        cplx<double> out = {0.0, 0.0};
        if (b * kDBlock + jj < M && b * kDBlock + kk < M) {
            double re = 0.0, im = 0.0, ajr = 0.0, aji = 0.0, akr = 0.0, aki = 0.0;
#pragma unroll 8
            for (int gi = 0; gi < groups; gi++) {
                const cplx<double>* q = gpart + (int64_t(gi) * nblk + b) * ng;
                re += q[jj * kDBlock + kk].re;
                im += q[jj * kDBlock + kk].im;
                ajr += q[kDBlock * kDBlock + jj].re;
                aji += q[kDBlock * kDBlock + jj].im;
                akr += q[kDBlock * kDBlock + kk].re;
                aki += q[kDBlock * kDBlock + kk].im;
            }
            ajr *= iw; aji *= iw; akr *= iw; aki *= iw;
            out = {re * iw - (ajr * akr + aji * aki), im * iw - (ajr * aki - aji * akr)};
        }
        Dblk[f] = out;
    }
}

// damping split from the traces: gam[0] = gc, gam[1] = gd (one block)
__global__ void kfac_damping_kernel(const double* C, int N, const cplx<double>* Dblk, int M, double shift,
                                    double* gam) {
    __shared__ double sh[2][256];
    double tc = 0.0, td = 0.0;
    for (int v = threadIdx.x; v < N; v += blockDim.x) tc += C[int64_t(v) * N + v];
    for (int j = threadIdx.x; j < M; j += blockDim.x) {
        const int b = j / kDBlock, jj = j % kDBlock;
        td += Dblk[(int64_t(b) * kDBlock + jj) * kDBlock + jj].re;
    }
    sh[0][threadIdx.x] = tc;
    sh[1][threadIdx.x] = td;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if (int(threadIdx.x) < st) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + st];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double ac = sh[0][0] / N, ad = sh[1][0] / M;
        const double pi = ac > 0.0 && ad > 0.0 ? sqrt(ac / ad) : 1.0;
        gam[0] = sqrt(shift) * pi;
        gam[1] = sqrt(shift) / pi;
    }
}

// in-place Gauss-Jordan inverse of the Hermitian positive definite A + gamma I
// (n <= 128 real / 64 complex, whole matrix in LDS; no pivoting is needed for a
// positive definite matrix)
template <typename E>
__device__ inline E gj_mul(E a, E b);
template <>
__device__ inline double gj_mul<double>(double a, double b) { return a * b; }
template <>
__device__ inline cplx<double> gj_mul<cplx<double>>(cplx<double> a, cplx<double> b) { return cmul(a, b); }
__device__ inline double gj_inv(double a) { return 1.0 / a; }
__device__ inline cplx<double> gj_inv(cplx<double> a) {
    const double d = a.re * a.re + a.im * a.im;
    return {a.re / d, -a.im / d};
}
__device__ inline double gj_sub(double a, double b) { return a - b; }
__device__ inline cplx<double> gj_sub(cplx<double> a, cplx<double> b) { return {a.re - b.re, a.im - b.im}; }
__device__ inline double gj_addr(double a, double g) { return a + g; }
__device__ inline cplx<double> gj_addr(cplx<double> a, double g) { return {a.re + g, a.im}; }

template <typename E>
__device__ void gj_invert(E* A, int n) {   // A: n x n in LDS, row-major; blockDim a multiple of 64
    __shared__ E piv_row[128];
    __shared__ E piv_col[128];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    for (int k = 0; k < n; k++) {
        __syncthreads();
        const E p = gj_inv(A[k * n + k]);
        // row k scaled, column k kept for the update
        for (int c = threadIdx.x; c < n; c += blockDim.x) {
            piv_row[c] = c == k ? p : gj_mul(A[k * n + c], p);
            piv_col[c] = A[c * n + k];
        }
        __syncthreads();
        // waves over rows, lanes over columns (no index division)
        for (int r = wave; r < n; r += waves) {
            const E f = piv_col[r];
            for (int c = lane; c < n; c += 64) {
                E& a = A[r * n + c];
                if (r == k)
                    a = piv_row[c];
                else
                    a = c == k ? gj_sub(E{}, gj_mul(f, p)) : gj_sub(a, gj_mul(f, piv_row[c]));
            }
        }
    }
    __syncthreads();
}

// Ci = (C + gc I)^-1 (one block, C in LDS)
constexpr int kInvThreads = 1024;

__global__ __launch_bounds__(kInvThreads) void kfac_invert_c_kernel(const double* C, int N, const double* gam,
                                                                    double* Ci) {
    extern __shared__ double lds_c[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    for (int r = wave; r < N; r += waves)
        for (int c = lane; c < N; c += 64) lds_c[r * N + c] = r == c ? gj_addr(C[r * N + c], gam[0]) : C[r * N + c];
    gj_invert<double>(lds_c, N);
    for (int r = wave; r < N; r += waves)
        for (int c = lane; c < N; c += 64) Ci[r * N + c] = lds_c[r * N + c];
}

// Di block b = (D_b + gd I)^-1 (one block per diagonal block of D), stored transposed
// in place: Dblk[b][k][j] = Di_b[j][k]
__global__ __launch_bounds__(kInvThreads) void kfac_invert_d_kernel(cplx<double>* Dblk, int M, const double* gam) {
    extern __shared__ cplx<double> lds_d[];
    const int b = blockIdx.x;
    const int n = std::min(kDBlock, M - b * kDBlock);
    cplx<double>* src = Dblk + int64_t(b) * kDBlock * kDBlock;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    for (int r = wave; r < n; r += waves)
        for (int c = lane; c < n; c += 64) {
            const cplx<double> x = src[r * kDBlock + c];
            lds_d[r * n + c] = r == c ? gj_addr(x, gam[1]) : x;
        }
    gj_invert<cplx<double>>(lds_d, n);
    for (int r = wave; r < n; r += waves)
        for (int c = lane; c < n; c += 64) src[c * kDBlock + r] = lds_d[r * n + c];
}

// parameter index of W at (site v, unit j): RBM v M + j; PsiDeep N + M + i M + j with
// i = inv[v][j]; -1 where (v, j) carries no parameter
__global__ void kfac_index_kernel(int layout, int N, int M, int Mpad, const int* inv, int* idx) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= int64_t(N) * M) return;
    const int v = int(e / M), j = int(e % M);
    if (layout == LAYOUT_RBM) {
        idx[e] = int(e);
    } else {
        const int i = inv[int64_t(v) * Mpad + j];
        idx[e] = i >= 0 ? N + M + i * M + j : -1;
    }
}

__device__ inline double readlane_f64(double x, int l) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane(int(b & 0xffffffff), l);
    const int hi = __builtin_amdgcn_readlane(int(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// T[v][j] = sum_{k in block(j)} X[v][k] Di[j][k], X[v][k] = r at (v, k): one wave per
// (site v, unit block b); lane l gathers X[v][64 b + l] and owns j = 64 b + l; the
// X values reach every lane by readlane, Di (stored transposed) is read coalesced
__global__ __launch_bounds__(256) void kfac_apply_d_kernel(int N, int M, const int* idx, const cplx<double>* DiT,
                                                           const cplx<double>* r, cplx<double>* T) {
    const int lane = threadIdx.x & 63;
    const int nblk = (M + kDBlock - 1) / kDBlock;
    const int task = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (task >= N * nblk) return;   // whole waves; no barrier in this kernel
    const int v = task / nblk, b = task % nblk;
    const int n = min(kDBlock, M - b * kDBlock);
    const int j = b * kDBlock + lane;
    const int q = lane < n ? idx[int64_t(v) * M + j] : -1;
    const cplx<double> x = q >= 0 ? r[q] : cplx<double>{0.0, 0.0};
    const cplx<double>* di = DiT + int64_t(b) * kDBlock * kDBlock + lane;
    // all 64 k: beyond n both X (lanes >= n) and the stored block are zero
    double re = 0.0, im = 0.0;
#pragma unroll 8
    for (int k = 0; k < kDBlock; k++) {
        const double xr = readlane_f64(x.re, k), xi = readlane_f64(x.im, k);
        const cplx<double> d = di[k * kDBlock];
        re += xr * d.re - xi * d.im;
        im += xr * d.im + xi * d.re;
    }
    if (lane < n) T[int64_t(v) * M + j] = {re, im};
}

__device__ inline double block_sum_kf(double v, double* sh) {
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    return sh[0] + sh[1] + sh[2] + sh[3];
}

// Z[v][j] = sum_u Ci[v][u] T[u][j]: one wave per (site v, 64 units); the row of Ci is
// wave-uniform, T is read coalesced.  z (and p when p_init is set) at every W
// parameter; part[b][1] = Re sum conj(r) z of the block's waves.  Launched with
// kCgParts blocks of 4 waves (grid-stride over the tasks).
__global__ __launch_bounds__(256) void kfac_apply_c_kernel(int N, int M, const int* idx, const double* Ci,
                                                           const cplx<double>* T, const cplx<double>* r,
                                                           cplx<double>* z, cplx<double>* p_init, double* part) {
    __shared__ double sh[4];
    const int lane = threadIdx.x & 63;
    const int jblk = (M + 63) / 64;
    double rz = 0.0;
    for (int task = blockIdx.x * 4 + (threadIdx.x >> 6); task < N * jblk; task += gridDim.x * 4) {
        const int v = task / jblk, j = (task % jblk) * 64 + lane;
        const double* ci = Ci + int64_t(v) * N;
        const cplx<double>* tc = T + min(j, M - 1);
        double re = 0.0, im = 0.0;
#pragma unroll 8
        for (int u = 0; u < N; u++) {
            const double c = ci[u];
            const cplx<double> t = tc[int64_t(u) * M];
            re += c * t.re;
            im += c * t.im;
        }
        const int q = j < M ? idx[int64_t(v) * M + j] : -1;
        if (q >= 0) {
            z[q] = {re, im};
            if (p_init) p_init[q] = {re, im};
            const cplx<double> rk = r[q];
            rz += rk.re * re + rk.im * im;
        }
    }
    rz = block_sum_kf(rz, sh);
    if (threadIdx.x == 0) part[blockIdx.x * 4 + 1] = rz;
}

}  // namespace

void launch_kfac_setup(const uint64_t* confs, int words, const cplx<float>* g32, const cplx<double>* g64, int Mpad,
                       const double* w, int64_t count, int N, int M, double shift, KfacWork& kw,
                       hipStream_t stream) {
    if (N < 1 || N > kKfacMaxSites) throw std::invalid_argument("the K-FAC preconditioner supports 1..128 sites");
    if (M < 1 || M > kKfacMaxUnits) throw std::invalid_argument("the K-FAC preconditioner supports 1..2048 hidden units");
    const int64_t nsub = std::min<int64_t>(count, kKfacSamples);
    const int groups = int((nsub + kGroup - 1) / kGroup);
    const int nblk = kfac_dblocks(M);
    kfac_moments_spin_kernel<<<unsigned(groups), 256, 0, stream>>>(confs, words, w, nsub, N, kw.spart);
    ANN_LAUNCH_CHECK();
    kfac_moments_g_kernel<<<dim3(unsigned(groups), unsigned(nblk)), 256, 0, stream>>>(g32, g64, Mpad, w, nsub, M, nblk,
                                                                                      kw.gpart);
    ANN_LAUNCH_CHECK();
    const int64_t nout = int64_t(N) * N + int64_t(nblk) * kDBlock * kDBlock;
    kfac_reduce_kernel<<<unsigned((nout + 255) / 256), 256, 0, stream>>>(kw.spart, kw.gpart, groups, N, M, nblk,
                                                                         kw.C, kw.Dblk);
    ANN_LAUNCH_CHECK();
    kfac_damping_kernel<<<1, 256, 0, stream>>>(kw.C, N, kw.Dblk, M, shift, kw.gam);
    ANN_LAUNCH_CHECK();
    // dynamic LDS beyond the default limit is opted in (up to 128 KiB here, of 160 KiB per CU)
    const size_t bytes_c = sizeof(double) * N * N;
    ANN_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kfac_invert_c_kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes_c)));
    kfac_invert_c_kernel<<<1, kInvThreads, bytes_c, stream>>>(kw.C, N, kw.gam, kw.Ci);
    ANN_LAUNCH_CHECK();
    const size_t bytes_d = sizeof(cplx<double>) * kDBlock * kDBlock;
    ANN_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kfac_invert_d_kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes_d)));
    kfac_invert_d_kernel<<<unsigned(nblk), kInvThreads, bytes_d, stream>>>(kw.Dblk, M, kw.gam);
    ANN_LAUNCH_CHECK();
}

void launch_kfac_index(int layout, int N, int M, int Mpad, const int* inv, int* idx, hipStream_t stream) {
    const int64_t n = int64_t(N) * M;
    kfac_index_kernel<<<unsigned((n + 255) / 256), 256, 0, stream>>>(layout, N, M, Mpad, inv, idx);
    ANN_LAUNCH_CHECK();
}

void launch_kfac_apply(const KfacView& kv, const cplx<double>* r, cplx<double>* z, cplx<double>* p_init,
                       double* part, cplx<double>* T, hipStream_t stream) {
    const int64_t dtasks = int64_t(kv.N) * kfac_dblocks(kv.M);
    kfac_apply_d_kernel<<<unsigned((dtasks + 3) / 4), 256, 0, stream>>>(kv.N, kv.M, kv.idx, kv.Dblk, r, T);
    ANN_LAUNCH_CHECK();
    kfac_apply_c_kernel<<<kCgParts, 256, 0, stream>>>(kv.N, kv.M, kv.idx, kv.Ci, T, r, z, p_init, part);
    ANN_LAUNCH_CHECK();
}

int kfac_dblocks(int M) { return (M + kDBlock - 1) / kDBlock; }
size_t kfac_dblock_elems(int M) { return size_t(kfac_dblocks(M)) * kDBlock * kDBlock; }
int kfac_groups(int64_t count) { return int((std::min<int64_t>(count, kKfacSamples) + kGroup - 1) / kGroup); }
size_t kfac_spart_elems(int N, int64_t count) { return size_t(kfac_groups(count)) * (size_t(N) * N + N + 1); }
size_t kfac_gpart_elems(int M, int64_t count) {
    return size_t(kfac_groups(count)) * kfac_dblocks(M) * (kDBlock * kDBlock + kDBlock);
}

}  // namespace ann
