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

// spin covariance from the transposed spin bits when every sample has the same weight
// (Monte Carlo): over the first nsub samples, sum_s s_v s_u = nsub - 2 #(s_v != s_u)
// and sum_s s_v = 2 #up_v - nsub, counted 64 samples at a time by popcount.  Block v,
// lane u; bits[blk][vpad] (bit b = spin of sample 64 blk + b up, zero beyond count).
__global__ __launch_bounds__(kKfacMaxSites) void kfac_c_popcount_kernel(const uint64_t* bits, int vpad,
                                                                      int64_t nsub, int N, double* C) {
    const int v = blockIdx.x, u = threadIdx.x;
    const int uc = u < N ? u : N - 1;
    const int64_t blocks = (nsub + 63) / 64;
    const uint64_t tail = (nsub & 63) ? (uint64_t(1) << (nsub & 63)) - 1 : ~uint64_t(0);
    int64_t diff = 0, up_u = 0, up_v = 0;
#pragma unroll 16
    for (int64_t blk = 0; blk < blocks; blk++) {
        const uint64_t m = blk == blocks - 1 ? tail : ~uint64_t(0);
        const uint64_t bv = bits[blk * vpad + v] & m, bu = bits[blk * vpad + uc] & m;
        diff += __popcll(bv ^ bu);
        up_u += __popcll(bu);
        up_v += __popcll(bv);
    }
    if (u < N) {
        const double n = double(nsub);
        const double mv = (2.0 * double(up_v) - n) / n, mu = (2.0 * double(up_u) - n) / n;
        C[int64_t(v) * N + u] = (n - 2.0 * double(diff)) / n - mv * mu;
    }
}

// raw g moments of one group of kGroupG samples and one block of 64 units, staged
// through LDS kChunkG samples at a time:
// part[gi][b] = (sum w conj(g_j) g_k [64][64], sum w g_j [64], sum w)
constexpr int kGroupG = 128;
constexpr int kChunkG = 32;
constexpr int kGOut = kDBlock * kDBlock + kDBlock + 1;
__global__ __launch_bounds__(256) void kfac_moments_g_kernel(const cplx<float>* g32, const cplx<double>* g64,
                                                             int Mpad, const double* w, int64_t nsub, int M,
                                                             int nblk, cplx<double>* part) {
    __shared__ cplx<double> sg[kChunkG][kDBlock];
    __shared__ double sw[kChunkG];
    constexpr int kPer = kDBlock * kDBlock / 256 + 1;   // 16 products + 1 (mean or weight)
    static_assert(kDBlock == 64, "256 threads cover a 64 x 64 block in 16 rounds");
    const int gi = blockIdx.x, b = blockIdx.y;
    double ar[kPer], ai[kPer];
#pragma unroll
    for (int i = 0; i < kPer; i++) ar[i] = ai[i] = 0.0;
    for (int c0 = 0; c0 < kGroupG; c0 += kChunkG) {
        const int64_t s0 = int64_t(gi) * kGroupG + c0;
        if (s0 >= nsub) break;
        __syncthreads();   // the previous chunk is consumed
        for (int e = threadIdx.x; e < kChunkG * kDBlock; e += blockDim.x) {
            const int s = e / kDBlock, jj = e % kDBlock;
            const int j = b * kDBlock + jj;
            cplx<double> x = {0.0, 0.0};
            if (s0 + s < nsub && j < M) {
                const int64_t o = (s0 + s) * Mpad + j;
                x = g32 ? cplx<double>{double(g32[o].re), double(g32[o].im)} : g64[o];
            }
            sg[s][jj] = x;
        }
        for (int s = threadIdx.x; s < kChunkG; s += blockDim.x) sw[s] = s0 + s < nsub ? w[s0 + s] : 0.0;
        __syncthreads();
        // samples outer (not unrolled), the thread's outputs inner: output i of thread t
        // is (jj, kk) = (t / 64 + 4 i, t % 64) for i < 16, then the mean of unit t
        // (t < 64) and the weight sum (t = 64)
        const int t = threadIdx.x;
#pragma unroll 1
        for (int s = 0; s < kChunkG; s++) {
            const double ws = sw[s];
            const cplx<double> y = sg[s][t & 63];
#pragma unroll
            for (int i = 0; i < kPer - 1; i++) {
                const cplx<double> x = sg[s][(t >> 6) + 4 * i];
                ar[i] += ws * (x.re * y.re + x.im * y.im);   // conj(x) y
                ai[i] += ws * (x.re * y.im - x.im * y.re);
            }
            if (t < kDBlock) {
                ar[kPer - 1] += ws * y.re;
                ai[kPer - 1] += ws * y.im;
            } else if (t == kDBlock) {
                ar[kPer - 1] += ws;
            }
        }
    }
    cplx<double>* out = part + (int64_t(gi) * nblk + b) * kGOut;
    const int t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < kPer - 1; i++) out[((t >> 6) + 4 * i) * kDBlock + (t & 63)] = {ar[i], ai[i]};
    if (t <= kDBlock) out[kDBlock * kDBlock + t] = {ar[kPer - 1], ai[kPer - 1]};
}

// sums of the group moments and centring, C = sum w s s^T / W - m m^T (when spart is
// given; else C is already set) by one thread per entry, D block b = sum w conj(g) g^T
// / W - conj(mg) mg^T by a quad of lanes per entry (lane q of the quad adds the groups
// q, q + 4, ..., the quad sum is taken in a fixed order)
__device__ inline double quad_sum_kf(double v) {
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    return v;
}

__global__ void kfac_reduce_kernel(const double* spart, int sgroups, const cplx<double>* gpart, int ggroups, int N,
                                   int M, int nblk, double* C, cplx<double>* Dblk) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t nc = spart ? (int64_t(N) * N + 3) / 4 * 4 : 0;   // C entries, rounded to whole quads
    if (e < nc) {
        if (e >= int64_t(N) * N) return;
        const int64_t ns = int64_t(N) * N + N + 1;
        double W = 0.0;
#pragma unroll 8
        for (int gi = 0; gi < sgroups; gi++) W += spart[gi * ns + ns - 1];
        const double iw = W > 0.0 ? 1.0 / W : 0.0;
        const int v = int(e / N), u = int(e % N);
        double s2 = 0.0, mv = 0.0, mu = 0.0;
#pragma unroll 8
        for (int gi = 0; gi < sgroups; gi++) {
            s2 += spart[gi * ns + e];
            mv += spart[gi * ns + int64_t(N) * N + v];
            mu += spart[gi * ns + int64_t(N) * N + u];
        }
        C[e] = s2 * iw - (mv * iw) * (mu * iw);
        return;
    }
    // D entries: 4 lanes each, whole quads (the block size is a multiple of 4)
    const int64_t qd = (e - nc) >> 2;
    const int q = int((e - nc) & 3);
    const int64_t nd = int64_t(nblk) * kDBlock * kDBlock;
    if (qd >= nd) return;   // (a whole quad leaves together)
    const int b = int(qd / (kDBlock * kDBlock));
    const int jj = int(qd % (kDBlock * kDBlock)) / kDBlock, kk = int(qd % kDBlock);
    const bool ok = b * kDBlock + jj < M && b * kDBlock + kk < M;
    double re = 0.0, im = 0.0, ajr = 0.0, aji = 0.0, akr = 0.0, aki = 0.0, W = 0.0;
#pragma unroll 8
    for (int gi = q; gi < ggroups; gi += 4) {
        const cplx<double>* g = gpart + (int64_t(gi) * nblk + b) * kGOut;
        const cplx<double> x = g[jj * kDBlock + kk], mj = g[kDBlock * kDBlock + jj], mk = g[kDBlock * kDBlock + kk];
        re += x.re;
        im += x.im;
        ajr += mj.re;
        aji += mj.im;
        akr += mk.re;
        aki += mk.im;
        W += g[kGOut - 1].re;
    }
    re = quad_sum_kf(re);
    im = quad_sum_kf(im);
    ajr = quad_sum_kf(ajr);
    aji = quad_sum_kf(aji);
    akr = quad_sum_kf(akr);
    aki = quad_sum_kf(aki);
    W = quad_sum_kf(W);
    if (q == 0) {
        cplx<double> out = {0.0, 0.0};
        if (ok) {
            const double iw = W > 0.0 ? 1.0 / W : 0.0;
            ajr *= iw; aji *= iw; akr *= iw; aki *= iw;
            out = {re * iw - (ajr * akr + aji * aki), im * iw - (ajr * aki - aji * akr)};
        }
        Dblk[qd] = out;
    }
}

// damping split from the traces: gam[0] = gc, gam[1] = gd (one block).  The product
// gc gd = kKfacDamping^2 shift is below the shift of the system: on an equilibrated
// 16,384-sample Heisenberg 10x10 problem (M = 400, shift 1e-3) the iterations to
// 1e-2 were 128 / 107 / 78 / 73 / 62 / 63 for multipliers 10 / 3 / 1 / 0.5 / 0.3 / 0.1
// with factors from all samples (96 with 2,048 samples and multiplier 1)
constexpr double kKfacDamping = 0.3;
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
        gam[0] = kKfacDamping * sqrt(shift) * pi;
        gam[1] = kKfacDamping * sqrt(shift) / pi;
    }
}

// Gauss-Jordan inverse of the Hermitian positive definite A + gamma I (n <= 128
// real / 64 complex; no pivoting is needed for a positive definite matrix)
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

// The matrix lives in registers: thread (wave w, lane l) of 16 waves holds rows
// w + 16 i (i < RI) and columns l + 64 h (h < CH).  Per pivot k the pivot row and
// column go through LDS (double-buffered, so one barrier per pivot); with p = 1/A_kk
// and the pivot row scaled, R_c = A_kc p (R_k = p), row k becomes R and every other
// row r becomes A_rc - A_rk R_c (column k: -A_rk p).
// load(r, c) gives the entries, store(r, c, x) takes the inverse (after every load).
template <typename E, int RI, int CH, typename Load, typename Store>
__device__ void gj_invert_regs(int n, Load load, Store store) {
    static_assert(CH == 1 || CH == 2, "columns lane and lane + 64");
    __shared__ E prow[2][64 * CH];
    __shared__ E pcol[2][16 * RI];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    E a[RI][CH];
#pragma unroll
    for (int i = 0; i < RI; i++)
#pragma unroll
        for (int h = 0; h < CH; h++) {
            const int r = wave + 16 * i, c = lane + 64 * h;
            a[i][h] = r < n && c < n ? load(r, c) : E{};
        }
    // entries outside n x n are never stored: they take part in the updates without
    // guards (no branch per entry) and never feed an entry that is stored
    for (int k = 0; k < n; k++) {
        const int b = k & 1, ik = k >> 4, hk = k >> 6;
        // the owner wave of row k writes its register slot ik (wave-uniform branches,
        // static register indices)
        if (wave == (k & 15)) {
#pragma unroll
            for (int i = 0; i < RI; i++)
                if (i == ik) {
#pragma unroll
                    for (int h = 0; h < CH; h++) prow[b][lane + 64 * h] = a[i][h];
                }
        }
        if (lane == (k & 63)) {   // the owner lane of column k (slot hk)
#pragma unroll
            for (int i = 0; i < RI; i++) {
                E x = a[i][0];
                if constexpr (CH > 1) x = hk == 1 ? a[i][1] : x;
                pcol[b][wave + 16 * i] = x;
            }
        }
        __syncthreads();
        const E p = gj_inv(prow[b][k]);
        E pr[CH];
#pragma unroll
        for (int h = 0; h < CH; h++) pr[h] = lane + 64 * h == k ? p : gj_mul(prow[b][lane + 64 * h], p);
#pragma unroll
        for (int i = 0; i < RI; i++) {
            const E f = pcol[b][wave + 16 * i];
            const bool rk = wave + 16 * i == k;
#pragma unroll
            for (int h = 0; h < CH; h++) {
                const E fp = gj_mul(f, pr[h]);
                const E upd = gj_sub(lane + 64 * h == k ? E{} : a[i][h], fp);
                a[i][h] = rk ? pr[h] : upd;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < RI; i++)
#pragma unroll
        for (int h = 0; h < CH; h++) {
            const int r = wave + 16 * i, c = lane + 64 * h;
            if (r < n && c < n) store(r, c, a[i][h]);
        }
}

// Ci = (C + gc I)^-1 (block 0) and Di_b = (D_b + gd I)^-1 (block 1 + b, stored
// transposed in place: Dblk[b][k][j] = Di_b[j][k]), each in the registers of its
// block; the two inversions run side by side in one launch.  Entries beyond a D
// block's n stay as the reduction left them (zero).
constexpr int kInvThreads = 1024;
static_assert(kKfacMaxSites <= 128 && kDBlock <= 64, "register tiles of the inversion");

__global__ __launch_bounds__(kInvThreads) void kfac_invert_kernel(const double* C, int N, const double* gam,
                                                                  double* Ci, cplx<double>* Dblk, int M) {
    if (blockIdx.x == 0) {
        const double gc = gam[0];
        gj_invert_regs<double, 8, 2>(
            N,
            [&](int r, int c) {
                const double x = C[r * N + c];
                return r == c ? gj_addr(x, gc) : x;
            },
            [&](int r, int c, double x) { Ci[r * N + c] = x; });
        return;
    }
    const int b = blockIdx.x - 1;
    const int n = min(kDBlock, M - b * kDBlock);
    cplx<double>* src = Dblk + int64_t(b) * kDBlock * kDBlock;
    const double gd = gam[1];
    gj_invert_regs<cplx<double>, 4, 1>(
        n,
        [&](int r, int c) {
            const cplx<double> x = src[r * kDBlock + c];
            return r == c ? gj_addr(x, gd) : x;
        },
        [&](int r, int c, cplx<double> x) { src[c * kDBlock + r] = x; });
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

constexpr int kStepRows = 4;                      // sites per task of the fused step
constexpr int kStepWaves = 8;                     // the site sum of a task split over 8 waves
constexpr int kStepThreads = 64 * kStepWaves;
constexpr int kStepSiteIters = (kKfacMaxSites + kStepWaves - 1) / kStepWaves;
constexpr int kStepK = kDBlock / kStepWaves;      // k of Di per wave

__device__ inline double block_sum_step(double v, double* sh) {
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kStepWaves; w++) t += sh[w];
    return t;
}

// One CG iteration after the matvec, with the K-FAC preconditioner (kCgParts blocks
// of 8 waves):
//   alpha = r^H z / p^H Ap (from the partials), x' = x + alpha p, r' = r - alpha Ap,
//   z = (diag + shift)^-1 r' for the parameters k < jac_end (Jacobi),
//   Z = Ci R' Di on the W parameters, R'[v][j] = r' at parameter (v, j),
//   newpart[b] = (r'^H z over k < jac_end, r'^H z over the W parameters, r'^H r', 0).
// Without Ap (the initial preconditioning) r' = r, x is not touched and p_init = z.
// Z is formed as (Ci R') Di so that a task (sites v0 .. v0+3, unit block b of 64)
// needs only its own rows of Ci R' (Y: the site sum split over the 8 waves) and the
// block's Di (stored transposed, DiT[k][j]: wave w takes k = 8w .. 8w+7); the wave
// partials are added in a fixed order through LDS.  A task issues all its loads
// (alpha's partials, the done flag, its rows of Ci into LDS, R' at its 64 units, the
// wave's Di rows, the owned parameters) before the first use of any.  Every task
// reads r at all sites of its unit block, so r' goes to r_out, a buffer other than
// r_in; each parameter is written by exactly one thread (its task, or the Jacobi
// loop of the last blocks, which have no task at these sizes).  Nothing is written
// once done[0] is set.
template <bool STEP, bool ANALYTIC>
__global__ __launch_bounds__(kStepThreads) void kfac_step_kernel(KfacStep a, int N, int M, const int* idx,
                                                                 WParamMap wm, const double* Ci,
                                                                 const cplx<double>* DiT) {
    // loads are unconditional (indices clamped, values selected afterwards), so that
    // none of them waits for another
    __shared__ double sh[kStepWaves];
    __shared__ cplx<double> red[kStepWaves][kStepRows][64];
    __shared__ cplx<double> ys[kStepRows][64];
    __shared__ double cis[kStepRows][kKfacMaxSites];
    const int t = threadIdx.x, lane = t & 63, wave = uniform_int(t >> 6);
    const cplx<double> zero{0.0, 0.0};
    // alpha from its partials; the loads are issued with a task's other loads (inside
    // each loop body, so they are not hoisted ahead of them), the sums taken at first use
    struct Partials {
        double2 rz[4];
        double ap[4];
    };
    auto load_partials = [&](Partials& pt) {
        if constexpr (STEP) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int b = lane + 64 * i;
                pt.rz[i] = *reinterpret_cast<const double2*>(a.rzpart + b * 4);
                pt.ap[i] = a.apart[min(b, a.a_parts - 1) * 2];
            }
        }
    };
    auto get_alpha = [&](const Partials& pt) {
        if constexpr (!STEP) return 0.0;
        double x = 0.0, y = 0.0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            x += pt.rz[i].x + pt.rz[i].y;
            y += lane + 64 * i < a.a_parts ? pt.ap[i] : 0.0;
        }
        x = wave_sum(x);
        y = wave_sum(y);
        return y > 0.0 ? x / y : 0.0;
    };
    auto rnew = [&](cplx<double> r, cplx<double> ap, double alpha) -> cplx<double> {
        if constexpr (STEP) return {r.re - alpha * ap.re, r.im - alpha * ap.im};
        return r;
    };
    double rz0 = 0.0, rz1 = 0.0, rr = 0.0;
    // the owner's update of parameter q: r' = rq, z = zq, p and x as loaded
    auto own = [&](int q, cplx<double> rq, cplx<double> zq, cplx<double> pq, cplx<double> xq, double alpha,
                   double& rz) {
        if constexpr (STEP) {
            a.x[q] = {xq.re + alpha * pq.re, xq.im + alpha * pq.im};
            a.r_out[q] = rq;
        }
        a.z[q] = zq;
        if (a.p_init) a.p_init[q] = zq;
        rz += rq.re * zq.re + rq.im * zq.im;
        rr += rq.re * rq.re + rq.im * rq.im;
    };
    bool frozen = false;
    const int nblk = (M + kDBlock - 1) / kDBlock;
    const int nsg = (N + kStepRows - 1) / kStepRows;
    const int tasks = nsg * nblk;
    // tasks in unit-block-major order, handed out in contiguous runs per XCD (workgroup
    // i runs on XCD i mod 8): the tasks of one unit block share their columns of r and
    // Ap, which an XCD's L2 then fetches for about tasks / 8 / nsg blocks instead of all
    const int xslot = gridDim.x % kXcds == 0
                          ? int(blockIdx.x % kXcds) * int(gridDim.x / kXcds) + int(blockIdx.x / kXcds)
                          : int(blockIdx.x);
    for (int task = xslot; task < tasks; task += gridDim.x) {
        const int b = task / nsg, v0 = task % nsg * kStepRows;
        const int j = b * kDBlock + lane;
        const bool jok = j < M;
        const int jc = jok ? j : M - 1;
        // parameter indices (-1: none) and their clamped form for the loads
        const int vo = v0 + (t >> 6);
        const bool owner = t < kStepRows * 64 && vo < N && jok;
        const int voc = vo < N ? vo : N - 1;
        const int jm = ANALYTIC && !wm.rbm ? (jc * wm.shift) % N : 0;
        int qo = ANALYTIC ? wm.param_at(voc, jc, jm) : idx[voc * M + jc];
        qo = owner ? qo : -1;
        int qs[kStepSiteIters];
#pragma unroll
        for (int i = 0; i < kStepSiteIters; i++) {
            const int u = wave + kStepWaves * i;
            const int uc = u < N ? u : N - 1;
            const int q = ANALYTIC ? wm.param_at(uc, jc, jm) : idx[uc * M + jc];
            qs[i] = u < N && jok ? q : -1;
        }
        // Ci rows of the task
        const int cr = t / kKfacMaxSites, cu = t % kKfacMaxSites;
        const bool cok = cr < kStepRows && cu < N && v0 + cr < N;
        const double cload = Ci[cok ? (v0 + cr) * N + cu : 0];
        const int qoc = qo >= 0 ? qo : 0;
        cplx<double> ro = a.r_in[qoc], ao = zero, po = zero, xo = zero;
        if constexpr (STEP) {
            ao = a.Ap[qoc];
            po = a.p[qoc];
            xo = a.x[qoc];
        }
        Partials pt;
        load_partials(pt);
        // R' at this wave's sites in two batches of 8 (the registers of 16 would spill)
        constexpr int kBatch = kStepSiteIters / 2;
        cplx<double> rv[kBatch], av[kBatch];
        auto load_sites = [&](int i0) {
#pragma unroll
            for (int i = 0; i < kBatch; i++) {
                const int qc = qs[i0 + i] >= 0 ? qs[i0 + i] : 0;
                rv[i] = a.r_in[qc];
                av[i] = zero;
                if constexpr (STEP) av[i] = a.Ap[qc];
            }
        };
        // Y[r][k] = sum_u Ci[v0 + r][u] R'[u][64 b + k] over this wave's sites u = wave + 8 i
        cplx<double> acc[kStepRows];
#pragma unroll
        for (int r = 0; r < kStepRows; r++) acc[r] = zero;
        auto add_sites = [&](int i0, double alpha) {
#pragma unroll
            for (int i = 0; i < kBatch; i++) {
                const int u = wave + kStepWaves * (i0 + i);   // < 128; cis and x are zero for u >= N
                cplx<double> x = rnew(rv[i], av[i], alpha);
                if (qs[i0 + i] < 0) x = zero;
#pragma unroll
                for (int r = 0; r < kStepRows; r++) {
                    const double c = cis[r][u];
                    acc[r].re += c * x.re;
                    acc[r].im += c * x.im;
                }
            }
        };
        load_sites(0);
        frozen = a.done && a.done[0] != 0;
        if (cr < kStepRows) cis[cr][cu] = cok ? cload : 0.0;
        const double alpha = get_alpha(pt);
        __syncthreads();
        add_sites(0, alpha);
        load_sites(kBatch);
        cplx<double> d[kStepK];
        const cplx<double>* di = DiT + (b * kDBlock + wave * kStepK) * kDBlock + lane;
#pragma unroll
        for (int i = 0; i < kStepK; i++) d[i] = di[i * kDBlock];
        add_sites(kBatch, alpha);
#pragma unroll
        for (int r = 0; r < kStepRows; r++) red[wave][r][lane] = acc[r];
        __syncthreads();
        if (t < kStepRows * 64) {
            const int r = t >> 6;
            cplx<double> y = red[0][r][lane];
#pragma unroll
            for (int w = 1; w < kStepWaves; w++) {
                y.re += red[w][r][lane].re;
                y.im += red[w][r][lane].im;
            }
            ys[r][lane] = y;
        }
        __syncthreads();
        // Z[r][j] = sum_k Y[r][k] Di[j][k], this wave's k = 8 wave + i
#pragma unroll
        for (int r = 0; r < kStepRows; r++) {
            double zr = 0.0, zi = 0.0;
#pragma unroll
            for (int i = 0; i < kStepK; i++) {
                const cplx<double> y = ys[r][wave * kStepK + i];
                zr += y.re * d[i].re - y.im * d[i].im;
                zi += y.re * d[i].im + y.im * d[i].re;
            }
            red[wave][r][lane] = {zr, zi};
        }
        __syncthreads();
        if (qo >= 0 && !frozen) {
            const int r = t >> 6;
            cplx<double> zq = red[0][r][lane];
#pragma unroll
            for (int w = 1; w < kStepWaves; w++) {
                zq.re += red[w][r][lane].re;
                zq.im += red[w][r][lane].im;
            }
            own(qo, rnew(ro, ao, alpha), zq, po, xo, alpha, rz1);
        }
        __syncthreads();
    }
    // Jacobi part, from the last block down (no task at these sizes); alpha is taken
    // before the loop, where every lane of the wave is active for its reduction
    const int jb = int(gridDim.x) - 1 - int(blockIdx.x);
    if (int64_t(jb) * kStepThreads < a.jac_end) {
        Partials pt;
        load_partials(pt);
        const double alpha = get_alpha(pt);
        const bool fz = a.done && a.done[0] != 0;
        for (int k = jb * kStepThreads + t; k < a.jac_end; k += int(gridDim.x) * kStepThreads) {
            const cplx<double> r0 = a.r_in[k];
            cplx<double> ap = zero, pq = zero, xq = zero;
            if constexpr (STEP) {
                ap = a.Ap[k];
                pq = a.p[k];
                xq = a.x[k];
            }
            const double d = a.diag ? a.diag[k] + a.shift : 1.0;
            const cplx<double> rq = rnew(r0, ap, alpha);
            if (!fz) own(k, rq, {rq.re / d, rq.im / d}, pq, xq, alpha, rz0);
        }
    }
    if (a.done) frozen = a.done[0] != 0;
    if (frozen) return;   // uniform: no barrier is skipped by part of the block
    rz0 = block_sum_step(rz0, sh);
    rz1 = block_sum_step(rz1, sh);
    rr = block_sum_step(rr, sh);
    if (t == 0) {
        a.newpart[blockIdx.x * 4 + 0] = rz0;
        a.newpart[blockIdx.x * 4 + 1] = rz1;
        a.newpart[blockIdx.x * 4 + 2] = rr;
        a.newpart[blockIdx.x * 4 + 3] = 0.0;
    }
}

}  // namespace

void launch_kfac_moments(const uint64_t* confs, int words, const uint64_t* bits, int vpad, bool w_uniform,
                         const cplx<float>* g32, const cplx<double>* g64, int Mpad, const double* w, int64_t count,
                         int N, int M, KfacWork& kw, hipStream_t stream) {
    if (N < 1 || N > kKfacMaxSites) throw std::invalid_argument("the K-FAC preconditioner supports 1..128 sites");
    if (M < 1 || M > kKfacMaxUnits) throw std::invalid_argument("the K-FAC preconditioner supports 1..2048 hidden units");
    const int64_t nsub = std::min<int64_t>(count, kKfacSamples);
    const int nblk = kfac_dblocks(M);
    int sgroups = 0;
    if (w_uniform) {
        if (!bits || vpad < N) throw std::invalid_argument("the K-FAC spin moments need the transposed spin bits");
        kfac_c_popcount_kernel<<<unsigned(N), kKfacMaxSites, 0, stream>>>(bits, vpad, nsub, N, kw.C);
    } else {
        sgroups = int((nsub + kGroup - 1) / kGroup);
        kfac_moments_spin_kernel<<<unsigned(sgroups), 256, 0, stream>>>(confs, words, w, nsub, N, kw.spart);
    }
    ANN_LAUNCH_CHECK();
    const int ggroups = int((nsub + kGroupG - 1) / kGroupG);
    kfac_moments_g_kernel<<<dim3(unsigned(ggroups), unsigned(nblk)), 256, 0, stream>>>(g32, g64, Mpad, w, nsub, M,
                                                                                       nblk, kw.gpart);
    ANN_LAUNCH_CHECK();
    // C entries (general weights only) then 4 lanes per D entry; N * N is rounded up to
    // a multiple of 4 so that every quad of D lanes is whole
    const int64_t nc = w_uniform ? 0 : (int64_t(N) * N + 3) / 4 * 4;
    const int64_t nout = nc + 4 * int64_t(nblk) * kDBlock * kDBlock;
    kfac_reduce_kernel<<<unsigned((nout + 255) / 256), 256, 0, stream>>>(w_uniform ? nullptr : kw.spart, sgroups,
                                                                         kw.gpart, ggroups, N, M, nblk, kw.C,
                                                                         kw.Dblk);
    ANN_LAUNCH_CHECK();
}

void launch_kfac_finish(int N, int M, double shift, KfacWork& kw, hipStream_t stream) {
    if (N < 1 || N > kKfacMaxSites || M < 1 || M > kKfacMaxUnits) throw std::invalid_argument("K-FAC factor sizes");
    const int nblk = kfac_dblocks(M);
    kfac_damping_kernel<<<1, 256, 0, stream>>>(kw.C, N, kw.Dblk, M, shift, kw.gam);
    ANN_LAUNCH_CHECK();
    kfac_invert_kernel<<<unsigned(1 + nblk), kInvThreads, 0, stream>>>(kw.C, N, kw.gam, kw.Ci, kw.Dblk, M);
    ANN_LAUNCH_CHECK();
}

void launch_kfac_setup(const uint64_t* confs, int words, const uint64_t* bits, int vpad, bool w_uniform,
                       const cplx<float>* g32, const cplx<double>* g64, int Mpad, const double* w, int64_t count, int N,
                       int M, double shift, KfacWork& kw, hipStream_t stream) {
    launch_kfac_moments(confs, words, bits, vpad, w_uniform, g32, g64, Mpad, w, count, N, M, kw, stream);
    launch_kfac_finish(N, M, shift, kw, stream);
}

void launch_kfac_index(int layout, int N, int M, int Mpad, const int* inv, int* idx, hipStream_t stream) {
    const int64_t n = int64_t(N) * M;
    kfac_index_kernel<<<unsigned((n + 255) / 256), 256, 0, stream>>>(layout, N, M, Mpad, inv, idx);
    ANN_LAUNCH_CHECK();
}

void launch_kfac_step(const KfacView& kv, const KfacStep& a, hipStream_t stream) {
    if (a.a_parts > 256) throw std::invalid_argument("too many partial sums");
    if (a.Ap && (!a.x || !a.p || !a.r_out || a.r_out == a.r_in || !a.apart || !a.rzpart))
        throw std::invalid_argument("incomplete K-FAC CG step");
    if (int64_t(kv.N) * kv.M + a.jac_end > INT32_MAX) throw std::invalid_argument("too many parameters");
    const bool affine = kv.wm.analytic();
    auto run = [&](auto kern) {
        kern<<<kCgParts, kStepThreads, 0, stream>>>(a, kv.N, kv.M, kv.idx, kv.wm, kv.Ci, kv.Dblk);
    };
    if (a.Ap)
        affine ? run(kfac_step_kernel<true, true>) : run(kfac_step_kernel<true, false>);
    else
        affine ? run(kfac_step_kernel<false, true>) : run(kfac_step_kernel<false, false>);
    ANN_LAUNCH_CHECK();
}

int kfac_dblocks(int M) { return (M + kDBlock - 1) / kDBlock; }
size_t kfac_dblock_elems(int M) { return size_t(kfac_dblocks(M)) * kDBlock * kDBlock; }
size_t kfac_spart_elems(int N, int64_t count) {
    const int64_t groups = (std::min<int64_t>(count, kKfacSamples) + kGroup - 1) / kGroup;
    return size_t(groups) * (size_t(N) * N + N + 1);
}
size_t kfac_gpart_elems(int M, int64_t count) {
    const int64_t groups = (std::min<int64_t>(count, kKfacSamples) + kGroupG - 1) / kGroupG;
    return size_t(groups) * kfac_dblocks(M) * kGOut;
}

}  // namespace ann
