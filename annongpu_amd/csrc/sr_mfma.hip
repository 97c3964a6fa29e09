// sr_mfma.hip -- the two GEMMs of the structured stochastic-reconfiguration
// matvec on bf16 matrix cores (fp32 mode of the dense nets).
//
// With O_k(s) = s_v g_j(s) (W part), g_j(s) (b part) and s_i (a part), the
// reference's S.v (source/network_functions/TDVP.cu.template:336-443) is
//   u_s = sum_j g_sj (theta_v(s)_j + vb_j) + sum_i s_i va_i,
//         theta_v(s) = sum_v s_v V[v][:]                    (first GEMM)
//   y   = S^T (c o g) with c_s = w_s (u_s - <O>.v)          (second GEMM)
// and every P-vector statistic of TDVP.eval is the second GEMM with other c.
//
// Both GEMMs have the spins as one operand (+-1: exact in bf16) and fp32 data as
// the other, which is split into three bf16 terms (hi + mid + lo carry 24
// significant bits), so the products are exact and the sums are fp32 matrix-core
// accumulations.  v_mfma_f32_16x16x16_bf16: A lane l = row l%16, k = 4(l/16)+e;
// B lane l = k = 4(l/16)+e, column l%16; D lane l = row 4(l/16)+r, column l%16.
//
// ov_mfma_kernel fuses the first GEMM with the contraction against g, so the
// [count][Mpad] angles of v are never written to memory; spin_gemm_bf16_kernel
// streams g once per call and writes fp32 partial sums per sample chunk, which a
// second kernel adds in fp64 in a fixed order.
#include "kernels.hpp"
#include "wave_net.hpp"

#include <algorithm>
#include <type_traits>

namespace ann {

namespace {

using namespace wave;

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ inline short bf16_bits(float x) {
    const uint32_t b = __float_as_uint(x);
    return short(uint16_t((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16));
}
__device__ inline float bf16_value(short h) { return __uint_as_float(uint32_t(uint16_t(h)) << 16); }

// x = hi + mid + lo exactly (truncating split: each step keeps the top 8
// significant bits, the remainder is exact in fp32 and has 8 bits fewer)
__device__ inline void split3(float x, short& hi, short& mid, short& lo) {
    const uint32_t b0 = __float_as_uint(x) & 0xFFFF0000u;
    const float r1 = x - __uint_as_float(b0);
    const uint32_t b1 = __float_as_uint(r1) & 0xFFFF0000u;
    const float r2 = r1 - __uint_as_float(b1);
    hi = short(b0 >> 16);
    mid = short(b1 >> 16);
    lo = short(__float_as_uint(r2) >> 16);
}

// +-1 / 0 of spin v of a configuration, as bf16 bits (v == N is the constant-1 row)
__device__ inline short spin_bf16(const uint64_t* w, int v, int N) {
    if (v < N) {
        const uint64_t word = v < 64 ? w[0] : v < 128 ? w[1] : v < 192 ? w[2] : w[3];
        return ((word >> (v & 63)) & 1) ? short(0x3F80) : short(0xBF80);
    }
    return v == N ? short(0x3F80) : short(0);
}

// 16-lane row sum (DPP within each row of 16 lanes)
__device__ inline double row16_sum(double v) {
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    v += dpp_f64<0x141>(v);
    v += dpp_f64<0x140>(v);
    return v;
}

constexpr int kMaxKT = 16;   // site tiles of 16: configurations of up to 256 sites
constexpr int kXcds = 8;     // workgroups are dispatched round-robin over the 8 XCDs

// XCD-aware tile order: the workgroups of one "row" (a sample chunk or slice,
// shared data: its spin fragments and coefficients) all land on the same XCD,
// so that data is fetched into one XCD's L2 only.  Linear workgroup id L runs on
// XCD L % 8; rows xcd, xcd + 8, ... are given to that XCD, cols fastest.
__device__ inline bool xcd_tile(int L, int rows, int cols, int& row, int& col) {
    const int xcd = L % kXcds, slot = L / kXcds;
    row = xcd + kXcds * (slot / cols);
    col = slot % cols;
    return row < rows;
}
inline int xcd_grid(int rows, int cols) { return kXcds * ((rows + kXcds - 1) / kXcds) * cols; }

template <int K>
struct KtTag {
    static constexpr int value = K;
};

template <typename F>
void dispatch_kt(int kt, F&& fn) {
    switch (kt) {
        case 1: fn(KtTag<1>{}); break;
        case 2: fn(KtTag<2>{}); break;
        case 3: fn(KtTag<3>{}); break;
        case 4: fn(KtTag<4>{}); break;
        case 5: fn(KtTag<5>{}); break;
        case 6: fn(KtTag<6>{}); break;
        case 7: fn(KtTag<7>{}); break;
        case 8: fn(KtTag<8>{}); break;
        case 9: fn(KtTag<9>{}); break;
        case 10: fn(KtTag<10>{}); break;
        case 11: fn(KtTag<11>{}); break;
        case 12: fn(KtTag<12>{}); break;
        case 13: fn(KtTag<13>{}); break;
        case 14: fn(KtTag<14>{}); break;
        case 15: fn(KtTag<15>{}); break;
        case 16: fn(KtTag<16>{}); break;
        default: throw std::invalid_argument("configurations of 1..256 sites are supported");
    }
}

// ------------------------------------------------------------ u = O.v (part)
// A workgroup owns one column group (4 tiles of the split layout: tile 2p = real
// parts of units 16p..16p+15, tile 2p+1 = imaginary parts) and keeps its bf16x3
// fragments of V in LDS; each wave then walks 16-sample groups, computes their
// angles of v on matrix cores and contracts them with g.  upart[col group][s].
template <int KT>
__global__ __launch_bounds__(256) void ov_mfma_kernel(DenseView<float> net, const short4_t* vfrag, int NT,
                                                      int slices, const short4_t* afrag, int64_t count,
                                                      const cplx<float>* g, const cplx<double>* vb,
                                                      cplx<double>* upart) {
    extern __shared__ __align__(16) unsigned char ov_smem[];
    short4_t* sv = reinterpret_cast<short4_t*>(ov_smem);   // [3][KT][4][64]
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    int slice, y;
    if (!xcd_tile(blockIdx.x, slices, NT / 4, slice, y)) return;
    const int nt0 = 4 * y;
    const int nfrag = 3 * KT * 4 * 64;
    // the whole slice is loaded before any of it is stored (the loads overlap)
    constexpr int kFill = 8;
    for (int i0 = threadIdx.x; i0 < nfrag; i0 += 256 * kFill) {
        short4_t tmp[kFill];
#pragma unroll
        for (int f = 0; f < kFill; f++) {
            const int i = i0 + 256 * f;
            const int l = i & 63, t = (i >> 6) & 3, kt = (i >> 8) % KT, sp = (i >> 8) / KT;
            tmp[f] = i < nfrag ? vfrag[((int64_t(sp) * KT + kt) * NT + nt0 + t) * 64 + l] : short4_t{0, 0, 0, 0};
        }
#pragma unroll
        for (int f = 0; f < kFill; f++)
            if (i0 + 256 * f < nfrag) sv[i0 + 256 * f] = tmp[f];
    }
    __syncthreads();
    cplx<float> bj[2];
    int jq[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        jq[q] = (2 * y + q) * 16 + (lane & 15);
        const cplx<double> b = vb && jq[q] < net.M ? vb[jq[q]] : cplx<double>{0.0, 0.0};
        bj[q] = {float(b.re), float(b.im)};
    }
    const int kq = 4 * (lane >> 4);
    const int64_t groups = (count + 15) / 16;
    for (int64_t grp = int64_t(slice) * 4 + wave; grp < groups; grp += int64_t(slices) * 4) {
        const int64_t row0 = grp * 16;
        // g of this lane's output elements, loaded ahead of the matrix-core work
        cplx<float> gv[4][2];
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const int64_t s = row0 + kq + r;
                gv[r][q] = jq[q] < net.M && s < count ? g[s * net.Mpad + jq[q]] : cplx<float>{0.f, 0.f};
            }
        float4_t acc[4];
#pragma unroll
        for (int t = 0; t < 4; t++) acc[t] = {0.f, 0.f, 0.f, 0.f};
        // all spin fragments of the group are loaded before the first MFMA
        const short4_t* ag = afrag + grp * KT * 64 + lane;
        short4_t a[KT];
#pragma unroll
        for (int kt = 0; kt < KT; kt++) a[kt] = ag[kt * 64];
#pragma unroll
        for (int kt = 0; kt < KT; kt++) {
#pragma unroll
            for (int sp = 0; sp < 3; sp++)
#pragma unroll
                for (int t = 0; t < 4; t++)
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[kt], sv[((sp * KT + kt) * 4 + t) * 64 + lane],
                                                                       acc[t], 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int64_t s = row0 + kq + r;
            float ur = 0.f, ui = 0.f;
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const float tr = acc[2 * q][r] + bj[q].re, ti = acc[2 * q + 1][r] + bj[q].im;
                ur += gv[r][q].re * tr - gv[r][q].im * ti;
                ui += gv[r][q].re * ti + gv[r][q].im * tr;
            }
            // row_mirror leaves the full 16-lane row sum in every lane of the row
            const double sr = row16_sum(double(ur)), si = row16_sum(double(ui));
            if ((lane & 15) == 15 && s < count) upart[int64_t(y) * count + s] = {sr, si};
        }
    }
}

// spin fragments of the sample configurations, computed once per TDVP
// evaluation and reused by every matvec:
//   ov : [group][kt][lane]   rows = samples 16 group + l%16, k = sites 16 kt + 4 (l/16) + e
//   sg : [group][vt][lane]   rows = sites 16 vt + l%16 (site N = the constant-1 row),
//                            k = samples 16 group + 4 (l/16) + e
__global__ void spin_frag_kernel(const uint64_t* confs, int words, int64_t count, int N, int KT, int VT,
                                 short4_t* ov, short4_t* sg) {
    const int64_t e_all = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t groups = (count + 15) / 16;
    const int per = (KT + VT) * 64;
    if (e_all >= groups * per) return;
    const int64_t grp = e_all / per;
    const int rem = int(e_all % per);
    const int l = rem & 63;
    const int kq = 4 * (l >> 4);
    if (rem < KT * 64) {
        const int kt = rem >> 6;
        const int64_t s = grp * 16 + (l & 15);
        uint64_t w[kMaxWords];
#pragma unroll
        for (int k = 0; k < kMaxWords; k++) w[k] = s < count && k < words ? confs[s * words + k] : 0;
        short4_t a;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int v = kt * 16 + kq + e;
            a[e] = s < count && v < N ? spin_bf16(w, v, N) : short(0);
        }
        ov[(grp * KT + kt) * 64 + l] = a;
    } else {
        const int vt = (rem >> 6) - KT;
        const int v = vt * 16 + (l & 15);
        short4_t a;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int64_t s = grp * 16 + kq + e;
            uint64_t w[kMaxWords];
#pragma unroll
            for (int k = 0; k < kMaxWords; k++) w[k] = s < count && k < words ? confs[s * words + k] : 0;
            a[e] = s < count ? spin_bf16(w, v, N) : short(0);
        }
        sg[(grp * VT + vt) * 64 + l] = a;
    }
}

// --------------------------------------------------------------- Y = S^T X
// X[s][c]: c < 2 Mpad from c_s g_sj (mode 0), c_s conj(g_sj) (mode 1) or
// Re(c_s) |g_sj|^2 in the real column (mode 2); columns 2 Mpad, 2 Mpad + 1 hold c_s.
constexpr int kSpinRowTiles = 7;     // rows 112 per wave (N + 1 <= 112 in one pass)

// A workgroup (4 waves) owns 64 real columns and a chunk of samples.  Per stage
// of 64 samples each thread loads 4 consecutive samples x 2 complex columns (one
// 16-byte load per sample, 16 threads cover a 256-byte row segment), forms
// c_s g_sj, splits it into bf16x3 and writes whole B-fragment quads (4 samples)
// to LDS; the stages are double-buffered (the next stage's loads are in flight
// while the current one multiplies) and each wave runs 7 row tiles x 4 k-steps
// x 3 splits of MFMA on its own 16-column tile.
constexpr int kSgStep = 64;

__global__ __launch_bounds__(256) void spin_gemm_bf16_kernel(DenseView<float> net, const short4_t* afrag,
                                                             const cplx<float>* g, int mode,
                                                             const cplx<double>* coef, int64_t count, int64_t chunk,
                                                             int chunks, int cblocks, int vpad, int ncols,
                                                             float* partial) {
    __shared__ short4_t sbf[2][3][4][4][64];        // [buf][split][k-step][col tile][lane]
    __shared__ short4_t saf[2][4][kSpinRowTiles][64];  // [buf][k-step][row tile][lane] spin fragments
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    int chunk_id, cb;
    if (!xcd_tile(blockIdx.x, chunks, cblocks, chunk_id, cb)) return;
    const int c0 = cb * 64;
    const int vt0 = blockIdx.y * kSpinRowTiles;
    const int vtiles = vpad / 16;
    const int64_t s0 = int64_t(chunk_id) * chunk;
    const int64_t s1 = s0 + chunk < count ? s0 + chunk : count;
    const int two_m = 2 * net.Mpad;
    // loader role: samples 4 sg .. 4 sg + 3 of the stage, complex columns jb, jb + 1
    const int sg = threadIdx.x >> 4;
    const int jp = (threadIdx.x & 15) * 2;
    const int jb = c0 / 2 + jp;
    float4 rg[4];
    cplx<float> rc[4];
    constexpr int kAfragPerThread = (4 * kSpinRowTiles * 64 + 255) / 256;
    short4_t ra[kAfragPerThread];
    const int vtl = vtiles - vt0 < kSpinRowTiles ? vtiles - vt0 : kSpinRowTiles;   // row tiles of this z group
    auto load = [&](int64_t st) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int64_t s = st + 4 * sg + e;
            const bool ok = s < s1;
            rg[e] = ok && jb < net.Mpad ? *reinterpret_cast<const float4*>(g + s * net.Mpad + jb)
                                        : float4{0.f, 0.f, 0.f, 0.f};
            const cplx<double> c = ok ? coef[s] : cplx<double>{0.0, 0.0};
            rc[e] = {float(c.re), float(c.im)};
        }
        // spin fragments of the stage: 4 k-steps x kSpinRowTiles tiles x 64 lanes, 4 per thread
#pragma unroll
        for (int q = 0; q < kAfragPerThread; q++) {
            const int i = threadIdx.x + 256 * q;           // [ks][t][lane]
            const int ks = i / (kSpinRowTiles * 64), t = (i / 64) % kSpinRowTiles, l = i & 63;
            const int64_t grp = (st >> 4) + ks;
            ra[q] = i < 4 * kSpinRowTiles * 64 && t < vtl && grp * 16 < s1
                        ? afrag[(grp * (vtiles) + vt0 + t) * 64 + l] : short4_t{0, 0, 0, 0};
        }
    };
    auto store = [&](int buf) {
        const int ks = sg >> 2, kl = (sg & 3) * 16;   // k-step of the quad, lane group of its 4 samples
#pragma unroll
        for (int i = 0; i < 4; i++) {                 // real columns 2 jp + i of the block
            const int col = 2 * jp + i;
            const int c = c0 + col;
            short4_t h[3];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const float gr = (i >> 1) ? rg[e].z : rg[e].x;
                const float gi0 = (i >> 1) ? rg[e].w : rg[e].y;
                const float gi = mode == 1 ? -gi0 : gi0;
                const float cr = rc[e].re, ci = rc[e].im;
                float x;
                if (c < two_m)
                    x = mode == 2 ? ((i & 1) ? 0.f : cr * (gr * gr + gi * gi))
                                  : ((i & 1) ? cr * gi + ci * gr : cr * gr - ci * gi);
                else
                    x = c == two_m ? cr : c == two_m + 1 ? ci : 0.f;
                short h0, h1, h2;
                split3(x, h0, h1, h2);
                h[0][e] = h0;
                h[1][e] = h1;
                h[2][e] = h2;
            }
            const int ct = col >> 4, cc = col & 15;
#pragma unroll
            for (int sp = 0; sp < 3; sp++) sbf[buf][sp][ks][ct][kl + cc] = h[sp];
        }
#pragma unroll
        for (int q = 0; q < kAfragPerThread; q++) {
            const int i = threadIdx.x + 256 * q;
            if (i < 4 * kSpinRowTiles * 64)
                saf[buf][i / (kSpinRowTiles * 64)][(i / 64) % kSpinRowTiles][i & 63] = ra[q];
        }
    };
    float4_t acc[kSpinRowTiles];
#pragma unroll
    for (int t = 0; t < kSpinRowTiles; t++) acc[t] = {0.f, 0.f, 0.f, 0.f};
    const int kq = 4 * (lane >> 4);
    (void)net;
    int buf = 0;
    load(s0);
    store(0);
    __syncthreads();
    for (int64_t st = s0; st < s1; st += kSgStep) {
        const bool more = st + kSgStep < s1;
        if (more) load(st + kSgStep);
#pragma unroll 1
        for (int ks = 0; ks < 4; ks++) {
            short4_t b[3];
#pragma unroll
            for (int sp = 0; sp < 3; sp++) b[sp] = sbf[buf][sp][ks][wave][lane];
#pragma unroll
            for (int t = 0; t < kSpinRowTiles; t++) {
                if (t >= vtl) continue;
                const short4_t a = saf[buf][ks][t][lane];
#pragma unroll
                for (int sp = 0; sp < 3; sp++)
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[sp], acc[t], 0, 0, 0);
            }
        }
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    const int c = c0 + 16 * wave + (lane & 15);
#pragma unroll
    for (int t = 0; t < kSpinRowTiles; t++) {
        if (vt0 + t >= vtiles || c >= ncols) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int v = (vt0 + t) * 16 + kq + r;
            partial[(int64_t(chunk_id) * vpad + v) * ncols + c] = acc[t][r];
        }
    }
}

__global__ void sum_partials_f32_kernel(const float* partial, int chunks, int64_t n, double* out) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double acc = 0.0;
#pragma unroll 8
    for (int c = 0; c < chunks; c++) acc += double(partial[c * n + e]);
    out[e] = acc;
}

__global__ void split_y_kernel(const double* Y, int N, int Mpad, int ncols, cplx<double>* Yw, cplx<double>* Yb,
                               cplx<double>* Ya) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t nw = int64_t(N + 1) * Mpad;
    if (e < nw) {
        const int v = int(e / Mpad), j = int(e % Mpad);
        const cplx<double> y = {Y[int64_t(v) * ncols + 2 * j], Y[int64_t(v) * ncols + 2 * j + 1]};
        if (v < N) {
            if (Yw) Yw[e] = y;
        } else if (Yb) {
            Yb[j] = y;
        }
    } else if (e < nw + N && Ya) {
        const int v = int(e - nw);
        Ya[v] = {Y[int64_t(v) * ncols + 2 * Mpad], Y[int64_t(v) * ncols + 2 * Mpad + 1]};
    }
}

// 256-thread block sum, the same fixed order in every block
__device__ inline double block_sum256(double v, double* sh) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    return sh[0] + sh[1] + sh[2] + sh[3];
}

// sum over blocks of partial[b * stride + idx], b < nparts <= 256 (all threads get it)
__device__ inline double parts_total(const double* partial, int nparts, int stride, int idx, double* sh) {
    const double x = int(threadIdx.x) < nparts ? partial[int64_t(threadIdx.x) * stride + idx] : 0.0;
    return block_sum256(x, sh);
}

constexpr int kDotParts = 64;   // blocks of the fused dot products

// v (parameter order) -> bf16x3 fragments of V (split layout), va, vb, and the
// partials of <O>.v.  Blocks [0, pack) pack, then [pack, pack + vab) write va /
// vb, then kDotParts blocks reduce <O>.v.
__global__ __launch_bounds__(256) void unpack_pack_kernel(DenseView<float> net, const int* inv,
                                                          const cplx<double>* v, int KT, int NT, int pack_blocks,
                                                          int vab_blocks, uint16_t* frag, cplx<double>* va,
                                                          cplx<double>* vb, const cplx<double>* obar,
                                                          double* opart) {
    __shared__ double sh[4];
    const int b = blockIdx.x;
    if (b < pack_blocks) {
        const int64_t e = int64_t(b) * 256 + threadIdx.x;
        const int ncol = 2 * net.Mpad;
        if (e >= int64_t(KT) * 16 * ncol) return;
        const int vis = int(e / ncol), col = int(e % ncol);
        const int nt = col / 16, cc = col % 16;
        const int j = (nt / 2) * 16 + cc;
        double re = 0.0, im = 0.0;
        if (vis < net.N && j < net.M) {
            if (net.layout == LAYOUT_RBM) {
                re = v[int64_t(vis) * net.M + j].re;
                im = v[int64_t(vis) * net.M + j].im;
            } else {
                // inv[vis][j]: the unique i with conn(i, j) = vis, -1 if none, -2 if several
                const int i0 = inv[int64_t(vis) * net.Mpad + j];
                if (i0 >= 0) {
                    re = v[net.N + net.M + int64_t(i0) * net.M + j].re;
                    im = v[net.N + net.M + int64_t(i0) * net.M + j].im;
                } else if (i0 == -2) {
                    for (int i = 0; i < net.C; i++) {
                        if (int(net.conn[int64_t(i) * net.M + j]) == vis) {
                            re += v[net.N + net.M + int64_t(i) * net.M + j].re;
                            im += v[net.N + net.M + int64_t(i) * net.M + j].im;
                        }
                    }
                }
            }
        }
        float r = float((nt & 1) ? im : re);
        const int kt = vis / 16, kk = vis % 16;
        const int lane = (kk / 4) * 16 + cc, el = kk % 4;
#pragma unroll
        for (int sp = 0; sp < 3; sp++) {
            const short h = bf16_bits(r);
            r -= bf16_value(h);
            frag[((((int64_t(sp) * KT + kt) * NT + nt) * 64 + lane) * 4) + el] = uint16_t(h);
        }
        return;
    }
    if (b < pack_blocks + vab_blocks) {
        const int e2 = (b - pack_blocks) * 256 + threadIdx.x;
        const bool deep = net.layout == LAYOUT_DEEP1;
        if (e2 < net.Mpad) {
            vb[e2] = deep && e2 < net.M ? v[net.N + e2] : cplx<double>{0.0, 0.0};
        } else if (e2 < net.Mpad + net.N) {
            const int i = e2 - net.Mpad;
            va[i] = deep ? v[i] : cplx<double>{0.0, 0.0};
        }
        return;
    }
    // <O>.v (no conjugation), grid-stride over the parameters
    const int db = b - pack_blocks - vab_blocks;
    double re = 0.0, im = 0.0;
    for (int64_t k = int64_t(db) * 256 + threadIdx.x; k < net.num_params; k += int64_t(kDotParts) * 256) {
        const cplx<double> o = obar[k], x = v[k];
        re += o.re * x.re - o.im * x.im;
        im += o.re * x.im + o.im * x.re;
    }
    re = block_sum256(re, sh);
    im = block_sum256(im, sh);
    if (threadIdx.x == 0) {
        opart[db * 2] = re;
        opart[db * 2 + 1] = im;
    }
}

// u_s = w_s (sum_b upart[b][s] + sum_i s_i va_i - <O>.v); four lanes per sample
// (a quad), each taking every fourth partial and site, summed by quad DPP
__device__ inline double quad_sum(double v) {
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    return v;
}

__global__ __launch_bounds__(256) void ov_finish_center_kernel(const cplx<double>* upart, int blocks, int64_t count,
                                                               const uint64_t* confs, int words, int N,
                                                               const cplx<double>* va, const double* opart,
                                                               const double* w, cplx<double>* u) {
    __shared__ double sh[4];
    __shared__ cplx<double> sva[64 * kMaxWords];   // va of up to 256 sites
    const double c_re = parts_total(opart, kDotParts, 2, 0, sh);
    const double c_im = parts_total(opart, kDotParts, 2, 1, sh);
    if (va)
        for (int i = threadIdx.x; i < N; i += blockDim.x) sva[i] = va[i];
    __syncthreads();
    const int64_t s = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 2;
    const int q = threadIdx.x & 3;
    const bool ok = s < count;
    double re = 0.0, im = 0.0;
    if (ok) {
        for (int b = q; b < blocks; b += 4) {
            re += upart[int64_t(b) * count + s].re;
            im += upart[int64_t(b) * count + s].im;
        }
        if (va) {
            uint64_t cw[kMaxWords];
#pragma unroll
            for (int k = 0; k < kMaxWords; k++) cw[k] = k < words ? confs[s * words + k] : 0;
            for (int i = q; i < N; i += 4) {
                const uint64_t word = i < 64 ? cw[0] : i < 128 ? cw[1] : i < 192 ? cw[2] : cw[3];
                const double si = ((word >> (i & 63)) & 1) ? 1.0 : -1.0;
                re += si * sva[i].re;
                im += si * sva[i].im;
            }
        }
    }
    re = quad_sum(re);
    im = quad_sum(im);
    if (ok && q == 0) u[s] = {w[s] * (re - c_re), w[s] * (im - c_im)};
}

// y_k (parameter order) = sum over chunks of the spin-GEMM partials + shift v_k,
// plus the block partials of v^H y
__global__ __launch_bounds__(256) void spin_y_kernel(const double* full, int ncols,
                                                     int layout, int N, int M, int Mpad, const uint32_t* conn,
                                                     const cplx<double>* v, double shift, int P,
                                                     cplx<double>* y, double* ypart) {
    __shared__ double sh[4];
    double dr = 0.0, di = 0.0;
    for (int k = blockIdx.x * 256 + threadIdx.x; k < P; k += gridDim.x * 256) {
        int row, col;
        if (layout == LAYOUT_RBM) {
            row = k / M;
            col = 2 * (k % M);
        } else if (k < N) {
            row = k;
            col = 2 * Mpad;
        } else if (k < N + M) {
            row = N;
            col = 2 * (k - N);
        } else {
            const int q = k - N - M;
            row = int(conn[q]);
            col = 2 * (q % M);
        }
        const int64_t off = int64_t(row) * ncols + col;
        double yr = full[off], yi = full[off + 1];
        const cplx<double> vk = v[k];
        yr += shift * vk.re;
        yi += shift * vk.im;
        y[k] = {yr, yi};
        dr += vk.re * yr + vk.im * yi;
        di += vk.re * yi - vk.im * yr;
    }
    dr = block_sum256(dr, sh);
    di = block_sum256(di, sh);
    if (threadIdx.x == 0) {
        ypart[blockIdx.x * 2] = dr;
        ypart[blockIdx.x * 2 + 1] = di;
    }
}

}  // namespace

int sr_y_parts(int P) { return std::min(256, std::max(1, (P + 255) / 256)); }

SpinPlan spin_plan(int N, int Mpad, int64_t count) {
    SpinPlan p;
    p.vpad = (N + 1 + 15) / 16 * 16;
    p.ncols = 2 * Mpad + 16;
    p.cblocks = (p.ncols + 63) / 64;
    p.zgroups = (p.vpad / 16 + kSpinRowTiles - 1) / kSpinRowTiles;
    // about 512 workgroups of 4 waves, sample chunks of whole 32-sample stages
    int64_t chunks = std::max<int64_t>(1, 512 / (int64_t(p.cblocks) * p.zgroups));
    chunks = std::min<int64_t>(chunks, std::max<int64_t>(1, (count + 255) / 256));
    int64_t chunk = (count + chunks - 1) / chunks;
    chunk = (chunk + kSgStep - 1) / kSgStep * kSgStep;
    p.chunk = chunk > 0 ? chunk : kSgStep;
    p.chunks = int((count + p.chunk - 1) / p.chunk);
    if (p.chunks < 1) p.chunks = 1;
    p.n = int64_t(p.vpad) * p.ncols;
    return p;
}

void launch_spin_frags(const uint64_t* confs, int words, int64_t count, int N, int npad, int vpad, uint16_t* ov,
                       uint16_t* sg, hipStream_t stream) {
    if (count <= 0) return;
    const int KT = npad / 16, VT = vpad / 16;
    const int64_t n = (count + 15) / 16 * (KT + VT) * 64;
    spin_frag_kernel<<<unsigned((n + 255) / 256), 256, 0, stream>>>(confs, words, count, N, KT, VT,
                                                                    reinterpret_cast<short4_t*>(ov),
                                                                    reinterpret_cast<short4_t*>(sg));
    ANN_LAUNCH_CHECK();
}

void launch_spin_gemm_bf16(const DenseView<float>& net, const uint16_t* sgfrag, const cplx<float>* g, int mode,
                           const cplx<double>* coef, int64_t count, const SpinPlan& p, float* partial,
                           hipStream_t stream) {
    if (count <= 0) return;
    if (net.N + 1 > 16 * kSpinRowTiles * p.zgroups) throw std::invalid_argument("spin GEMM plan too small");
    const dim3 grid{unsigned(xcd_grid(p.chunks, p.cblocks)), unsigned(p.zgroups)};
    spin_gemm_bf16_kernel<<<grid, 256, 0, stream>>>(net, reinterpret_cast<const short4_t*>(sgfrag), g, mode, coef,
                                                    count, p.chunk, p.chunks, p.cblocks, p.vpad, p.ncols, partial);
    ANN_LAUNCH_CHECK();
}

void launch_sr_ov_bf16(const DenseView<float>& net, const int* inv, const cplx<double>* v, const cplx<double>* obar,
                       const double* w, const uint64_t* confs, int words, int64_t count, const cplx<float>* g,
                       const uint16_t* ovfrag, int npad, SrWorkspace ws, hipStream_t stream) {
    if (count <= 0) return;
    const int KT = npad / 16, NT = 2 * net.Mpad / 16;
    const int64_t npack = int64_t(KT) * 16 * 2 * net.Mpad;
    const int pack_blocks = int((npack + 255) / 256);
    const int vab_blocks = (net.Mpad + net.N + 255) / 256;
    unpack_pack_kernel<<<unsigned(pack_blocks + vab_blocks + kDotParts), 256, 0, stream>>>(
        net, inv, v, KT, NT, pack_blocks, vab_blocks, ws.frag, ws.va, ws.vb, obar, ws.opart);
    ANN_LAUNCH_CHECK();
    const bool deep = net.layout == LAYOUT_DEEP1;
    // about 1024 workgroups in all, each column group's V fragments in LDS
    const int64_t groups = (count + 15) / 16;
    int slices = int(std::max<int64_t>(1, std::min<int64_t>((groups + 3) / 4, 1024 / (NT / 4))));
    slices = (slices + kXcds - 1) / kXcds * kXcds;
    const size_t lds = sizeof(short4_t) * 3 * KT * 4 * 64;
    if (lds > 160 * 1024) throw std::invalid_argument("too many sites for the LDS-resident V fragments");
    const dim3 grid{unsigned(xcd_grid(slices, NT / 4))};
    dispatch_kt(KT, [&](auto tag) {
        constexpr int kt = std::decay_t<decltype(tag)>::value;
        if (lds > 64 * 1024)
            ANN_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(ov_mfma_kernel<kt>),
                                              hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
        ov_mfma_kernel<kt><<<grid, 256, lds, stream>>>(net, reinterpret_cast<const short4_t*>(ws.frag), NT,
                                                       slices, reinterpret_cast<const short4_t*>(ovfrag), count, g,
                                                       deep ? ws.vb : nullptr, ws.upart);
    });
    ANN_LAUNCH_CHECK();
    ov_finish_center_kernel<<<unsigned((4 * count + 255) / 256), 256, 0, stream>>>(
        ws.upart, NT / 4, count, confs, words, net.N, deep ? ws.va : nullptr, ws.opart, w, ws.u);
    ANN_LAUNCH_CHECK();
}

void launch_sr_y_bf16(const DenseView<float>& net, const cplx<double>* v, double shift, const SpinPlan& plan,
                      SrWorkspace ws, cplx<double>* y, hipStream_t stream) {
    // chunk partials added in fp64 (coalesced), then gathered in parameter order
    sum_partials_f32_kernel<<<unsigned((plan.n + 255) / 256), 256, 0, stream>>>(ws.spin_part, plan.chunks, plan.n,
                                                                                ws.full);
    ANN_LAUNCH_CHECK();
    spin_y_kernel<<<unsigned(sr_y_parts(net.num_params)), 256, 0, stream>>>(
        ws.full, plan.ncols, net.layout, net.N, net.M, net.Mpad, net.conn, v, shift, net.num_params, y, ws.ypart);
    ANN_LAUNCH_CHECK();
}

void launch_spin_sums_bf16(const DenseView<float>& net, const uint64_t* confs, int words, const cplx<float>* g,
                           int mode, const cplx<double>* coef, int64_t count, cplx<double>* Yw, cplx<double>* Yb,
                           cplx<double>* Ya, hipStream_t stream) {
    if (count <= 0) return;
    const SpinPlan p = spin_plan(net.N, net.Mpad, count);
    float* partial = nullptr;
    double* full = nullptr;
    uint16_t* sg = nullptr;
    const size_t nsg = size_t((count + 15) / 16) * (p.vpad / 16) * 64 * 4;
    ANN_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&partial), sizeof(float) * p.n * p.chunks, stream));
    ANN_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&full), sizeof(double) * p.n, stream));
    ANN_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&sg), sizeof(uint16_t) * nsg, stream));
    launch_spin_frags(confs, words, count, net.N, 0, p.vpad, nullptr, sg, stream);
    launch_spin_gemm_bf16(net, sg, g, mode, coef, count, p, partial, stream);
    sum_partials_f32_kernel<<<unsigned((p.n + 255) / 256), 256, 0, stream>>>(partial, p.chunks, p.n, full);
    ANN_LAUNCH_CHECK();
    const int64_t ns = int64_t(net.N + 1) * net.Mpad + net.N;
    split_y_kernel<<<unsigned((ns + 255) / 256), 256, 0, stream>>>(full, net.N, net.Mpad, p.ncols, Yw, Yb, Ya);
    ANN_LAUNCH_CHECK();
    ANN_HIP_CHECK(hipFreeAsync(partial, stream));
    ANN_HIP_CHECK(hipFreeAsync(full, stream));
    ANN_HIP_CHECK(hipFreeAsync(sg, stream));
}

}  // namespace ann
