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

// ------------------------------------------------------------ u = O.v (part)
// A workgroup owns one column group (4 tiles of the split layout: tile 2p = real
// parts of units 16p..16p+15, tile 2p+1 = imaginary parts) and keeps its bf16x3
// fragments of V in LDS; each wave then walks 16-sample groups, computes their
// angles of v on matrix cores and contracts them with g.  upart[col group][s].
__global__ __launch_bounds__(256) void ov_mfma_kernel(DenseView<float> net, const short4_t* vfrag, int KT, int NT,
                                                      const uint64_t* confs, int words, int64_t count,
                                                      const cplx<float>* g, const cplx<double>* vb,
                                                      cplx<double>* upart) {
    extern __shared__ __align__(16) unsigned char ov_smem[];
    short4_t* sv = reinterpret_cast<short4_t*>(ov_smem);   // [3][KT][4][64]
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int y = blockIdx.y;
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
    cplx<double> bj[2];
    int jq[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        jq[q] = (2 * y + q) * 16 + (lane & 15);
        bj[q] = vb && jq[q] < net.M ? vb[jq[q]] : cplx<double>{0.0, 0.0};
    }
    const int kq = 4 * (lane >> 4);
    const int64_t groups = (count + 15) / 16;
    for (int64_t grp = int64_t(blockIdx.x) * 4 + wave; grp < groups; grp += int64_t(gridDim.x) * 4) {
        const int64_t row0 = grp * 16;
        const int64_t my_row = row0 + (lane & 15);
        uint64_t w[kMaxWords];
#pragma unroll
        for (int k = 0; k < kMaxWords; k++) w[k] = my_row < count && k < words ? confs[my_row * words + k] : 0;
        float4_t acc[4];
#pragma unroll
        for (int t = 0; t < 4; t++) acc[t] = {0.f, 0.f, 0.f, 0.f};
        for (int kt = 0; kt < KT; kt++) {
            short4_t a;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int v = kt * 16 + kq + e;
                a[e] = v < net.N ? spin_bf16(w, v, net.N) : short(0);   // V has no bias row
            }
#pragma unroll
            for (int sp = 0; sp < 3; sp++)
#pragma unroll
                for (int t = 0; t < 4; t++)
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, sv[((sp * KT + kt) * 4 + t) * 64 + lane],
                                                                       acc[t], 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int64_t s = row0 + kq + r;
            double ur = 0.0, ui = 0.0;
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const bool ok = jq[q] < net.M && s < count;
                const cplx<float> gv = ok ? g[s * net.Mpad + jq[q]] : cplx<float>{0.f, 0.f};
                const double tr = double(acc[2 * q][r]) + bj[q].re, ti = double(acc[2 * q + 1][r]) + bj[q].im;
                ur += double(gv.re) * tr - double(gv.im) * ti;
                ui += double(gv.re) * ti + double(gv.im) * tr;
            }
            // row_mirror leaves the full 16-lane row sum in every lane of the row
            ur = row16_sum(ur);
            ui = row16_sum(ui);
            if ((lane & 15) == 15 && s < count) upart[int64_t(y) * count + s] = {ur, ui};
        }
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

template <int WORDS>
__global__ __launch_bounds__(256) void spin_gemm_bf16_kernel(DenseView<float> net, const uint64_t* confs,
                                                             const cplx<float>* g, int mode,
                                                             const cplx<double>* coef, int64_t count, int64_t chunk,
                                                             int vpad, int ncols, float* partial) {
    __shared__ short4_t sbf[2][3][4][4][64];        // [buf][split][k-step][col tile][lane]
    __shared__ uint64_t sw[2][kSgStep][WORDS];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int c0 = blockIdx.x * 64;
    const int vt0 = blockIdx.z * kSpinRowTiles;
    const int vtiles = vpad / 16;
    const int64_t s0 = int64_t(blockIdx.y) * chunk;
    const int64_t s1 = s0 + chunk < count ? s0 + chunk : count;
    const int two_m = 2 * net.Mpad;
    // loader role: samples 4 sg .. 4 sg + 3 of the stage, complex columns jb, jb + 1
    const int sg = threadIdx.x >> 4;
    const int jp = (threadIdx.x & 15) * 2;
    const int jb = c0 / 2 + jp;
    float4 rg[4];
    cplx<float> rc[4];
    uint64_t rw[WORDS];
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
        if (threadIdx.x < kSgStep) {
            const int64_t ws_ = st + threadIdx.x;
#pragma unroll
            for (int k = 0; k < WORDS; k++) rw[k] = ws_ < s1 ? confs[ws_ * WORDS + k] : 0;
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
#pragma unroll
                for (int sp = 0; sp < 3; sp++) {
                    const short b16 = bf16_bits(x);
                    x -= bf16_value(b16);
                    h[sp][e] = b16;
                }
            }
            const int ct = col >> 4, cc = col & 15;
#pragma unroll
            for (int sp = 0; sp < 3; sp++) sbf[buf][sp][ks][ct][kl + cc] = h[sp];
        }
        if (threadIdx.x < kSgStep)
#pragma unroll
            for (int k = 0; k < WORDS; k++) sw[buf][threadIdx.x][k] = rw[k];
    };
    float4_t acc[kSpinRowTiles];
#pragma unroll
    for (int t = 0; t < kSpinRowTiles; t++) acc[t] = {0.f, 0.f, 0.f, 0.f};
    const int kq = 4 * (lane >> 4);
    int buf = 0;
    load(s0);
    store(0);
    __syncthreads();
    for (int64_t st = s0; st < s1; st += kSgStep) {
        const bool more = st + kSgStep < s1;
        if (more) load(st + kSgStep);
#pragma unroll 1
        for (int ks = 0; ks < 4; ks++) {
            uint64_t w4[4][kMaxWords];
#pragma unroll
            for (int e = 0; e < 4; e++)
#pragma unroll
                for (int k = 0; k < kMaxWords; k++) w4[e][k] = k < WORDS ? sw[buf][ks * 16 + kq + e][k < WORDS ? k : 0] : 0;
#pragma unroll
            for (int t = 0; t < kSpinRowTiles; t++) {
                if (vt0 + t >= vtiles) continue;
                const int v = (vt0 + t) * 16 + (lane & 15);
                short4_t a;
#pragma unroll
                for (int e = 0; e < 4; e++) a[e] = spin_bf16(w4[e], v, net.N);
#pragma unroll
                for (int sp = 0; sp < 3; sp++)
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, sbf[buf][sp][ks][wave][lane], acc[t], 0, 0, 0);
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
            partial[(int64_t(blockIdx.y) * vpad + v) * ncols + c] = acc[t][r];
        }
    }
}

__global__ void sum_partials_f32_kernel(const float* partial, int chunks, int64_t n, double* out) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double acc = 0.0;
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

// u_s = w_s (sum_b upart[b][s] + sum_i s_i va_i - <O>.v)
__global__ __launch_bounds__(256) void ov_finish_center_kernel(const cplx<double>* upart, int blocks, int64_t count,
                                                               const uint64_t* confs, int words, int N,
                                                               const cplx<double>* va, const double* opart,
                                                               const double* w, cplx<double>* u) {
    __shared__ double sh[4];
    __shared__ cplx<double> sva[4 * 64 * kMaxWords / 4];   // va of up to 256 sites
    const double c_re = parts_total(opart, kDotParts, 2, 0, sh);
    const double c_im = parts_total(opart, kDotParts, 2, 1, sh);
    if (va)
        for (int i = threadIdx.x; i < N; i += blockDim.x) sva[i] = va[i];
    __syncthreads();
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= count) return;
    double re = 0.0, im = 0.0;
#pragma unroll 8
    for (int b = 0; b < blocks; b++) {
        re += upart[int64_t(b) * count + s].re;
        im += upart[int64_t(b) * count + s].im;
    }
    if (va) {
        uint64_t cw[kMaxWords];
#pragma unroll
        for (int k = 0; k < kMaxWords; k++) cw[k] = k < words ? confs[s * words + k] : 0;
        for (int i = 0; i < N; i++) {
            const uint64_t word = i < 64 ? cw[0] : i < 128 ? cw[1] : i < 192 ? cw[2] : cw[3];
            const double si = ((word >> (i & 63)) & 1) ? 1.0 : -1.0;
            re += si * sva[i].re;
            im += si * sva[i].im;
        }
    }
    u[s] = {w[s] * (re - c_re), w[s] * (im - c_im)};
}

// y_k (parameter order) = sum over chunks of the spin-GEMM partials + shift v_k,
// plus the block partials of v^H y
__global__ __launch_bounds__(256) void spin_y_kernel(const float* partial, int chunks, int64_t n, int ncols,
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
        double yr = 0.0, yi = 0.0;
#pragma unroll 8
        for (int c = 0; c < chunks; c++) {
            yr += double(partial[c * n + off]);
            yi += double(partial[c * n + off + 1]);
        }
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

void launch_spin_gemm_bf16(const DenseView<float>& net, const uint64_t* confs, int words, const cplx<float>* g,
                           int mode, const cplx<double>* coef, int64_t count, const SpinPlan& p, float* partial,
                           hipStream_t stream) {
    if (count <= 0) return;
    if (net.N + 1 > 16 * kSpinRowTiles * p.zgroups) throw std::invalid_argument("spin GEMM plan too small");
    const dim3 grid{unsigned(p.cblocks), unsigned(p.chunks), unsigned(p.zgroups)};
    switch (words) {
        case 1: spin_gemm_bf16_kernel<1><<<grid, 256, 0, stream>>>(net, confs, g, mode, coef, count, p.chunk, p.vpad, p.ncols, partial); break;
        case 2: spin_gemm_bf16_kernel<2><<<grid, 256, 0, stream>>>(net, confs, g, mode, coef, count, p.chunk, p.vpad, p.ncols, partial); break;
        case 3: spin_gemm_bf16_kernel<3><<<grid, 256, 0, stream>>>(net, confs, g, mode, coef, count, p.chunk, p.vpad, p.ncols, partial); break;
        case 4: spin_gemm_bf16_kernel<4><<<grid, 256, 0, stream>>>(net, confs, g, mode, coef, count, p.chunk, p.vpad, p.ncols, partial); break;
        default: throw std::invalid_argument("configurations of 1..4 words are supported");
    }
    ANN_LAUNCH_CHECK();
}

void launch_sr_matvec_bf16(const DenseView<float>& net, const int* inv, const cplx<double>* v, const cplx<double>* obar,
                           const double* w, const uint64_t* confs, int words, int64_t count, const cplx<float>* g,
                           int npad, double shift, const SpinPlan& plan, SrWorkspace ws, cplx<double>* y,
                           hipStream_t stream) {
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
    const int per_col = int(std::max<int64_t>(1, std::min<int64_t>((groups + 3) / 4, 1024 / (NT / 4))));
    const size_t lds = sizeof(short4_t) * 3 * KT * 4 * 64;
    if (lds > 160 * 1024) throw std::invalid_argument("too many sites for the LDS-resident V fragments");
    if (lds > 64 * 1024)
        ANN_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(ov_mfma_kernel),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    const dim3 grid{unsigned(per_col), unsigned(NT / 4)};
    ov_mfma_kernel<<<grid, 256, lds, stream>>>(net, reinterpret_cast<const short4_t*>(ws.frag), KT, NT, confs, words,
                                               count, g, deep ? ws.vb : nullptr, ws.upart);
    ANN_LAUNCH_CHECK();
    ov_finish_center_kernel<<<unsigned((count + 255) / 256), 256, 0, stream>>>(
        ws.upart, NT / 4, count, confs, words, net.N, deep ? ws.va : nullptr, ws.opart, w, ws.u);
    ANN_LAUNCH_CHECK();
    launch_spin_gemm_bf16(net, confs, words, g, 1, ws.u, count, plan, ws.spin_part, stream);
    spin_y_kernel<<<unsigned(sr_y_parts(net.num_params)), 256, 0, stream>>>(
        ws.spin_part, plan.chunks, plan.n, plan.ncols, net.layout, net.N, net.M, net.Mpad, net.conn, v, shift,
        net.num_params, y, ws.ypart);
    ANN_LAUNCH_CHECK();
}

void launch_spin_sums_bf16(const DenseView<float>& net, const uint64_t* confs, int words, const cplx<float>* g,
                           int mode, const cplx<double>* coef, int64_t count, cplx<double>* Yw, cplx<double>* Yb,
                           cplx<double>* Ya, hipStream_t stream) {
    if (count <= 0) return;
    const SpinPlan p = spin_plan(net.N, net.Mpad, count);
    float* partial = nullptr;
    double* full = nullptr;
    ANN_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&partial), sizeof(float) * p.n * p.chunks, stream));
    ANN_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&full), sizeof(double) * p.n, stream));
    launch_spin_gemm_bf16(net, confs, words, g, mode, coef, count, p, partial, stream);
    sum_partials_f32_kernel<<<unsigned((p.n + 255) / 256), 256, 0, stream>>>(partial, p.chunks, p.n, full);
    ANN_LAUNCH_CHECK();
    const int64_t ns = int64_t(net.N + 1) * net.Mpad + net.N;
    split_y_kernel<<<unsigned((ns + 255) / 256), 256, 0, stream>>>(full, net.N, net.Mpad, p.ncols, Yw, Yb, Ya);
    ANN_LAUNCH_CHECK();
    ANN_HIP_CHECK(hipFreeAsync(partial, stream));
    ANN_HIP_CHECK(hipFreeAsync(full, stream));
}

}  // namespace ann
