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
// accumulations.  v_mfma_f32_16x16x32_bf16: A lane l = row l%16, k = 8(l/16)+e;
// B lane l = k = 8(l/16)+e, column l%16; D lane l = row 4(l/16)+r, column l%16
// (any k order shared by A and B gives the same product).
//
// Columns are handled in "pairs": pair p = hidden units 16p..16p+15, as two
// 16-column tiles (real parts, imaginary parts).  One wavefront owns one pair,
// keeps what it multiplies in registers and streams its column of g, so neither
// kernel touches LDS in its main loop:
//   ov32_kernel        V fragments of the pair in VGPRs (A operand), spins built
//                      from the configuration words (B operand), angles of v
//                      contracted with g on the fly; per-pair partial sums of O.v.
//   spin_gemm32_kernel spins built from a transposed bit array (bit b of
//                      bits[blk][v] = spin v of sample 64 blk + b), c o g split on
//                      the fly; the 4 waves of a workgroup share a pair, reduce
//                      through LDS once at the end and write one fp32 partial per
//                      sample chunk, added in fp64 in a fixed order by the consumer.
#include "kernels.hpp"
#include "wave_net.hpp"

#include <algorithm>
#include <type_traits>

namespace ann {

namespace {

using namespace wave;

typedef float float4_t __attribute__((ext_vector_type(4)));
typedef uint32_t dw4_t __attribute__((ext_vector_type(4)));   // 8 bf16
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

__device__ inline float4_t mfma32(dw4_t a, dw4_t b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b),
                                                   c, 0, 0, 0);
}

__device__ inline short bf16_bits(float x) {
    const uint32_t b = __float_as_uint(x);
    return short(uint16_t((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16));
}
__device__ inline float bf16_value(short h) { return __uint_as_float(uint32_t(uint16_t(h)) << 16); }

// x = hi + mid + lo exactly (truncating split: each step keeps the top 8
// significant bits, the remainder is exact in fp32 and has 8 bits fewer); the
// three terms are returned as fp32 bit patterns whose low halves are zero
__device__ inline void split3_bits(float x, uint32_t& hi, uint32_t& mid, uint32_t& lo) {
    hi = __float_as_uint(x) & 0xFFFF0000u;
    const float r1 = x - __uint_as_float(hi);
    mid = __float_as_uint(r1) & 0xFFFF0000u;
    const float r2 = r1 - __uint_as_float(mid);
    lo = __float_as_uint(r2) & 0xFFFF0000u;
}
// two truncated fp32 patterns -> one dword of two bf16 (element 2k low, 2k+1 high)
__device__ inline uint32_t pack_bf16x2(uint32_t even, uint32_t odd) { return (even >> 16) | odd; }

// fragment of 8 spins from 8 bits (bit e = element e, 1 = up = +1):
// dword k = bf16(element 2k) | bf16(element 2k+1) << 16, +1 = 0x3F80, -1 = 0xBF80
__device__ inline dw4_t spin_frag8(uint32_t x) {
    dw4_t a;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t t = x << (15 - 2 * k);   // bit 2k -> 15, bit 2k+1 -> 16 (x < 256)
        a[k] = ((t | (t << 15)) & 0x80008000u) ^ 0xBF80BF80u;
    }
    return a;
}

constexpr int kMaxVpad = 64 * kMaxWords + 16;            // spin rows incl. the bias row
constexpr int kBitWords = (kMaxVpad + 63) / 64;

template <int K>
struct KtTag {
    static constexpr int value = K;
};

// site tiles of 32: configurations of up to 256 sites
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
        default: throw std::invalid_argument("configurations of 1..256 sites are supported");
    }
}

// spin fragments of all 256 bytes in LDS: fragment of byte x = lut[x] (one
// ds_read_b128 instead of a dozen VALU operations per fragment)
__device__ inline void fill_spin_lut(dw4_t* lut) {
    for (int x = threadIdx.x; x < 256; x += blockDim.x) lut[x] = spin_frag8(uint32_t(x));
}

// ------------------------------------------------------------ u = O.v (part)
// Wave (pair p, slice): the V fragments of the pair (vfrag[((sp KT + kt) NT + 2p + h)
// 64 + lane], 8 bf16 each: sites 32 kt + 8 (lane/16) + e, unit 16p + lane%16,
// h = 0 real / 1 imaginary parts) stay in registers as the A operand (row = unit), the
// spins of 16 samples are the B operand (column = sample), so the product tile holds
// theta_v(s)_j with lane l owning sample l%16 at the units 16p + 4(l/16) + r, r < 4.
// Contracted with g there (one 32-byte load of the 4 units per lane) and summed over
// the 4 lane rows by two permlane swaps, lanes 0..15 store the 16 samples' partials in
// one coalesced 128-byte store.  The wave walks the 16-sample groups slice, slice +
// slices, ...; the next group's configuration words and g are loaded while the current
// group multiplies (configuration indices clamped to the last sample, g padded to
// sr_count_pad(count) rows: rows beyond count are computed and never stored).
// upart[p][s] = sum over the pair's units of g_sj (theta_v(s)_j + vb_j).
template <int KT>
__global__ __launch_bounds__(256) void ov32_kernel(const dw4_t* vfrag, int NT, int npairs, int slices,
                                                   const uint64_t* confs, int64_t count, int Mpad, int M,
                                                   const float2* g, const cplx<double>* vb, float2* upart) {
    constexpr int WORDS = (KT + 1) / 2;
    __shared__ dw4_t lut[256];
    fill_spin_lut(lut);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wid = blockIdx.x * kWavesPerBlock + uniform_int(threadIdx.x >> 6);
    if (wid >= npairs * slices) return;   // whole waves leave after the only barrier
    const int p = wid % npairs, slice = wid / npairs;
    const int q = lane >> 4, col = lane & 15;
    const int j0 = 16 * p + 4 * q;   // this lane's first unit
    // prologue loads in the order biases (vb: Mpad values, zero for PsiRBM), V fragments,
    // first group: each waits only for what was issued before it
    cplx<double> vbx[4];
#pragma unroll
    for (int r = 0; r < 4; r++) vbx[r] = vb[j0 + r];
    dw4_t bv[3][KT][2];
#pragma unroll
    for (int sp = 0; sp < 3; sp++)
#pragma unroll
        for (int kt = 0; kt < KT; kt++)
#pragma unroll
            for (int h = 0; h < 2; h++) bv[sp][kt][h] = vfrag[((sp * KT + kt) * NT + 2 * p + h) * 64 + lane];
    const int64_t groups = (count + 15) / 16;
    const int64_t last = count - 1;
    uint64_t wn[WORDS];
    float4 gn[2];
    auto load = [&](int64_t grp) {
        const int64_t sa = grp * 16 + col < last ? grp * 16 + col : last;   // sample of this lane's column
#pragma unroll
        for (int k = 0; k < WORDS; k++) wn[k] = confs[sa * WORDS + k];
        const float4* gp = reinterpret_cast<const float4*>(g + (grp * 16 + col) * Mpad + j0);
        gn[0] = gp[0];
        gn[1] = gp[1];
    };
    int64_t grp = slice;
    if (grp < groups) load(grp);
    float bre[4], bim[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        bre[r] = j0 + r < M ? float(vbx[r].re) : 0.f;
        bim[r] = j0 + r < M ? float(vbx[r].im) : 0.f;
    }
    for (; grp < groups; grp += slices) {
        uint64_t w[WORDS];
#pragma unroll
        for (int k = 0; k < WORDS; k++) w[k] = wn[k];
        const float4 g01 = gn[0], g23 = gn[1];
        if (grp + slices < groups) load(grp + slices);
        float4_t are = {0.f, 0.f, 0.f, 0.f}, aim = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < KT; kt++) {
            const int o = 32 * kt + 8 * q;
            const dw4_t b = lut[uint32_t(w[kt >> 1] >> (o & 63)) & 0xFFu];
#pragma unroll
            for (int sp = 0; sp < 3; sp++) {
                are = mfma32(bv[sp][kt][0], b, are);
                aim = mfma32(bv[sp][kt][1], b, aim);
            }
        }
        const float gx[4] = {g01.x, g01.z, g23.x, g23.z}, gy[4] = {g01.y, g01.w, g23.y, g23.w};
        float ur = 0.f, ui = 0.f;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float tr = are[r] + bre[r], ti = aim[r] + bim[r];
            ur += gx[r] * tr - gy[r] * ti;
            ui += gx[r] * ti + gy[r] * tr;
        }
        // sum over the 4 lane rows: lanes l, l ^ 16, l ^ 32, l ^ 48
        const auto a16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(ur), __float_as_uint(ur), false, false);
        const auto b16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(ui), __float_as_uint(ui), false, false);
        ur = __uint_as_float(a16[0]) + __uint_as_float(a16[1]);
        ui = __uint_as_float(b16[0]) + __uint_as_float(b16[1]);
        const auto a32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(ur), __float_as_uint(ur), false, false);
        const auto b32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(ui), __float_as_uint(ui), false, false);
        ur = __uint_as_float(a32[0]) + __uint_as_float(a32[1]);
        ui = __uint_as_float(b32[0]) + __uint_as_float(b32[1]);
        const int64_t s = grp * 16 + col;
        if (q == 0 && s < count) upart[int64_t(p) * count + s] = float2{ur, ui};
    }
}

// transposed spin bits: bit b of bits[blk][v] = spin v of sample 64 blk + b is up
// (v < N); row N is all ones (the bias row), rows above it zero; samples beyond
// count have zero bits.  One wave per block of 64 samples, one ballot per row.
__global__ __launch_bounds__(256) void spin_bits_kernel(const uint64_t* confs, int words, int64_t count, int N,
                                                        int vpad, uint64_t* bits) {
    const int lane = threadIdx.x & 63;
    const int64_t blk = int64_t(blockIdx.x) * kWavesPerBlock + uniform_int(threadIdx.x >> 6);
    if (blk >= (count + 63) / 64) return;
    const int64_t s = blk * 64 + lane;
    const bool ok = s < count;
    uint64_t w[kMaxWords];
#pragma unroll
    for (int k = 0; k < kMaxWords; k++) w[k] = ok && k < words ? confs[s * words + k] : 0;
#pragma unroll
    for (int k = 0; k < kBitWords; k++) {
        const uint64_t word = k < kMaxWords ? w[k < kMaxWords ? k : 0] : 0;
        uint64_t keep = 0;
        for (int vv = 0; vv < 64; vv++) {
            const int v = 64 * k + vv;
            if (v >= vpad) break;
            const bool up = v < N ? ((word >> vv) & 1) != 0 : v == N;
            const uint64_t m = __ballot(ok && up);
            if (vv == lane) keep = m;
        }
        if (64 * k + lane < vpad) bits[blk * vpad + 64 * k + lane] = keep;
    }
}

// --------------------------------------------------------------- Y = S^T X
// X[s][c]: c < 2 Mpad from c_s g_sj (MODE 0), c_s conj(g_sj) (MODE 1) or
// Re(c_s) |g_sj|^2 in the real column (MODE 2); columns 2 Mpad, 2 Mpad + 1 hold c_s.
// Output rows v < vpad (spins; row N = constant 1), columns in the interleaved
// [v][2 j + re/im] layout.
constexpr int kSpinRowTiles = 7;                // rows 112 per wave (N + 1 <= 112 in one pass)
constexpr int kSgStep = 32;                     // samples per matrix-core k step
constexpr int kSgAlign = kWavesPerBlock * kSgStep;   // chunk granularity: one step per wave
constexpr int kSgPad = 256;                          // sample padding of g and the coefficients

// Workgroup (pair p of npairs = Mpad/16 + 1, sample chunk) for the VTL row tiles
// starting at vt0: the 4 waves take the chunk's 32-sample steps round robin; per
// step a lane loads g of its 8 samples (k = 8(l/16)+e) at unit 16p + l%16 and
// their coefficients (g and coef have sr_count_pad(count) rows, those beyond
// count zero, so no load needs a guard), forms the two B fragments (real /
// imaginary column tile) as
// bf16x3, and takes the spin A fragment of each row tile from one byte of the
// transposed bits.  The next step's loads are issued before the current step
// multiplies.  Pair Mpad/16 holds the coefficient columns (launched only when
// with_coef); pairs beyond the hidden units are not launched, their partial
// columns are left as they are.
template <int MODE, int VTL>
__global__ __launch_bounds__(256, 2) void spin_gemm32_kernel(const uint64_t* bits, int vpad, int vt0,
                                                             const float2* g, int Mpad, int M, const float2* coef,
                                                             int64_t count, int64_t chunk, int npairs,
                                                             int with_coef, int ncols, float* partial) {
    // launched pairs: the unit pairs 0 .. ceil(M/16)-1, then the coefficient pair
    // Mpad/16 when with_coef (npairs counts the launched ones)
    __shared__ float4_t red[kWavesPerBlock][VTL][2][64];
    __shared__ dw4_t lut[256];
    fill_spin_lut(lut);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = uniform_int(threadIdx.x >> 6);
    const int q = lane >> 4, col = lane & 15;
    const int unit_pairs = (M + 15) / 16;
    const int pl = blockIdx.x % npairs;
    const int p = pl < unit_pairs ? pl : Mpad / 16;
    const int chunk_id = blockIdx.x / npairs;
    const bool coef_pair = pl >= unit_pairs;
    const bool active = !coef_pair || with_coef != 0;
    const int64_t s0 = int64_t(chunk_id) * chunk;
    const int64_t s1 = s0 + chunk < count ? s0 + chunk : count;
    const int j = 16 * p + col;
    float4_t acc[VTL][2];
#pragma unroll
    for (int t = 0; t < VTL; t++) acc[t][0] = acc[t][1] = float4_t{0.f, 0.f, 0.f, 0.f};
    const int nsteps = active && s1 > s0 ? int((s1 - s0 + kSgStep - 1) / kSgStep) : 0;
    const uint64_t* brow = bits + (vt0 * 16 + col);
    auto run = [&](auto coef_tag) {
        constexpr bool COEF = decltype(coef_tag)::value;
        float2 gn[8], cn[8];
        uint64_t bn[VTL];
        auto load = [&](int step) {
            const int64_t st = s0 + int64_t(kSgStep) * step;
            const float2* gp = g + (st + 8 * q) * Mpad + j;
            const float2* cp = coef + st + 8 * q;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                if constexpr (!COEF) gn[e] = gp[int64_t(e) * Mpad];
                cn[e] = cp[e];
            }
            const uint64_t* bb = brow + (st >> 6) * vpad;
#pragma unroll
            for (int t = 0; t < VTL; t++) bn[t] = bb[t * 16];
        };
        int step = wave;
        if (step < nsteps) load(step);
        for (; step < nsteps; step += kWavesPerBlock) {
            float2 gv[8], cv[8];
            uint64_t bw[VTL];
#pragma unroll
            for (int e = 0; e < 8; e++) {
                if constexpr (!COEF) gv[e] = gn[e];
                cv[e] = cn[e];
            }
#pragma unroll
            for (int t = 0; t < VTL; t++) bw[t] = bn[t];
            const int shift = int(((s0 + int64_t(kSgStep) * step) >> 5) & 1) * 32 + 8 * q;
            if (step + kWavesPerBlock < nsteps) load(step + kWavesPerBlock);
            // B fragments [split], real and imaginary column tiles
            dw4_t bre[3], bim[3];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                uint32_t r[2][3], i[2][3];
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int e = 2 * k + h;
                    float xr, xi;
                    if constexpr (COEF) {
                        xr = col == 0 ? cv[e].x : 0.f;
                        xi = col == 0 ? cv[e].y : 0.f;
                    } else {
                        const float gr = gv[e].x, gi = MODE == 1 ? -gv[e].y : gv[e].y;
                        const float cr = cv[e].x, ci = cv[e].y;
                        if constexpr (MODE == 2) {
                            xr = cr * (gr * gr + gi * gi);
                            xi = 0.f;
                        } else {
                            xr = cr * gr - ci * gi;
                            xi = cr * gi + ci * gr;
                        }
                    }
                    split3_bits(xr, r[h][0], r[h][1], r[h][2]);
                    split3_bits(xi, i[h][0], i[h][1], i[h][2]);
                }
#pragma unroll
                for (int sp = 0; sp < 3; sp++) {
                    bre[sp][k] = pack_bf16x2(r[0][sp], r[1][sp]);
                    bim[sp][k] = pack_bf16x2(i[0][sp], i[1][sp]);
                }
            }
#pragma unroll
            for (int t = 0; t < VTL; t++) {
                const dw4_t a = lut[uint32_t(bw[t] >> shift) & 0xFFu];
#pragma unroll
                for (int sp = 0; sp < 3; sp++) {
                    acc[t][0] = mfma32(a, bre[sp], acc[t][0]);
                    acc[t][1] = mfma32(a, bim[sp], acc[t][1]);
                }
            }
        }
    };
    if (coef_pair)
        run(std::true_type{});
    else
        run(std::false_type{});
    // the 4 waves' sums of each tile, added in a fixed order; wave w finishes row tiles w, w + 4
#pragma unroll
    for (int t = 0; t < VTL; t++) {
        red[wave][t][0][lane] = acc[t][0];
        red[wave][t][1][lane] = acc[t][1];
    }
    __syncthreads();
    const int c = 2 * j;
    if (c >= ncols) return;
#pragma unroll
    for (int t = 0; t < VTL; t++) {
        if ((t & (kWavesPerBlock - 1)) != wave) continue;
        float4_t sr = red[0][t][0][lane], si = red[0][t][1][lane];
#pragma unroll
        for (int w = 1; w < kWavesPerBlock; w++) {
            sr += red[w][t][0][lane];
            si += red[w][t][1][lane];
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int v = (vt0 + t) * 16 + 4 * q + r;
            *reinterpret_cast<float2*>(partial + (int64_t(chunk_id) * vpad + v) * ncols + c) = float2{sr[r], si[r]};
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

// fp32 coefficients, zero for count <= s < pad
__global__ void coef_f32_kernel(const cplx<double>* coef, int64_t count, int64_t pad, float2* out) {
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s < count)
        out[s] = float2{float(coef[s].re), float(coef[s].im)};
    else if (s < pad)
        out[s] = float2{0.f, 0.f};
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

constexpr int kDotParts = 64;   // blocks of the fused dot products

// beta = r^H z (this iteration) / r^H z (previous) from the CG partials (slots 0 + 1
// of [parts][4], parts <= 256): loads issued by load(), the sum taken by every wave
// alone at value() (same fixed order in every wave, no barrier; all lanes active)
struct CgBeta {
    double2 pn[4], po[4];
    __device__ void load(const CgDir& dir) {
        const int lane = threadIdx.x & 63;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int b = min(lane + 64 * i, dir.parts - 1);
            pn[i] = *reinterpret_cast<const double2*>(dir.newpart + b * 4);
            po[i] = *reinterpret_cast<const double2*>(dir.oldpart + b * 4);
        }
    }
    __device__ double value(const CgDir& dir) const {
        const int lane = threadIdx.x & 63;
        double rn = 0.0, ro = 0.0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const bool ok = lane + 64 * i < dir.parts;
            rn += ok ? pn[i].x + pn[i].y : 0.0;
            ro += ok ? po[i].x + po[i].y : 0.0;
        }
        rn = wave_sum(rn);
        ro = wave_sum(ro);
        return ro != 0.0 ? rn / ro : 0.0;
    }
};

// v (parameter order) -> bf16x3 fragments of V (pair layout, 32-site k tiles:
// frag[(((sp KT + kt) NT + nt) 64 + lane) 8 + el], site 32 kt + 8 (lane/16) + el,
// unit 16 (nt/2) + lane%16, nt even = real parts), va, vb, and the partials of
// <O>.v.  Blocks [0, pack) pack, then [pack, pack + vab) write va / vb, then
// kDotParts blocks reduce <O>.v.  With dir.z set, v is the new CG direction
// z + beta p_old, formed on the fly by every block (beta from the rz partials)
// and written to dir.p_new by the <O>.v blocks; it stays p_old once dir.done[0] is
// set.  The PsiDeep parameter of (site, unit) comes from wm when it is analytic,
// else from the table inv.
// Every role issues its loads (z and p_old, or v, at clamped indices) before it
// waits for the beta partials, so a block takes one memory round trip.
__global__ __launch_bounds__(256) void unpack_pack_kernel(DenseView<float> net, WParamMap wm, const int* inv,
                                                          const cplx<double>* v, CgDir dir, int KT, int NT,
                                                          int pack_blocks, int vab_blocks, uint16_t* frag,
                                                          cplx<double>* va, cplx<double>* vb,
                                                          const cplx<double>* obar, double* opart) {
    __shared__ double sh[4];
    const int b = blockIdx.x;
    const bool cg = dir.z != nullptr;
    CgBeta cb;
    int frozen = 0;
    if (cg) {
        cb.load(dir);
        frozen = dir.done[0];
    }
    // the value at parameter k (clamped to 0 for loads; the caller discards k < 0)
    struct Pending {
        cplx<double> a, b;
    };
    auto fetch = [&](int64_t k) -> Pending {
        const int64_t kc = k >= 0 ? k : 0;
        if (!cg) return {v[kc], {0.0, 0.0}};
        return {dir.z[kc], dir.p_old[kc]};
    };
    auto value = [&](const Pending& x, double beta) -> cplx<double> {
        if (!cg) return x.a;
        if (frozen) return x.b;
        return {x.a.re + beta * x.b.re, x.a.im + beta * x.b.im};
    };
    auto get_beta = [&]() { return cg ? cb.value(dir) : 0.0; };
    if (b < pack_blocks) {
        // one thread per (site, unit): it reads the parameter once and writes both the
        // real-part (column tile 2 (j / 16)) and the imaginary-part (2 (j / 16) + 1) entry
        const int64_t e = int64_t(b) * 256 + threadIdx.x;
        const bool in = e < int64_t(KT) * 32 * net.Mpad;
        const int vis = in ? int(e / net.Mpad) : 0, j = in ? int(e % net.Mpad) : 0;
        const int nt0 = 2 * (j / 16), cc = j % 16;
        // parameter of (site vis, unit j): -1 none, -2 several connections (summed below)
        int64_t k = -1;
        if (in && vis < net.N && j < net.M) {
            if (net.layout == LAYOUT_RBM) {
                k = int64_t(vis) * net.M + j;
            } else if (wm.shift >= 0) {
                k = wm.param(vis, j);
            } else {
                const int i0 = inv[int64_t(vis) * net.Mpad + j];
                k = i0 >= 0 ? net.N + net.M + int64_t(i0) * net.M + j : (i0 == -2 ? -2 : -1);
            }
        }
        const Pending pd = fetch(k);
        const double beta = get_beta();
        double re = 0.0, im = 0.0;
        if (k >= 0) {
            const cplx<double> x = value(pd, beta);
            re = x.re;
            im = x.im;
        } else if (k == -2) {
            for (int i = 0; i < net.C; i++) {
                if (int(net.conn[int64_t(i) * net.M + j]) == vis) {
                    const cplx<double> x = value(fetch(net.N + net.M + int64_t(i) * net.M + j), beta);
                    re += x.re;
                    im += x.im;
                }
            }
        }
        if (!in) return;
        const int kt = vis / 32, kk = vis % 32;
        const int lane = (kk / 8) * 16 + cc, el = kk % 8;
#pragma unroll
        for (int h2 = 0; h2 < 2; h2++) {
            float r = float(h2 ? im : re);
#pragma unroll
            for (int sp = 0; sp < 3; sp++) {
                const short h = bf16_bits(r);
                r -= bf16_value(h);
                frag[((((int64_t(sp) * KT + kt) * NT + nt0 + h2) * 64 + lane) * 8) + el] = uint16_t(h);
            }
        }
        return;
    }
    if (b < pack_blocks + vab_blocks) {
        const int e2 = (b - pack_blocks) * 256 + threadIdx.x;
        const bool deep = net.layout == LAYOUT_DEEP1;
        const int64_t k = !deep ? -1 : e2 < net.M ? net.N + e2 : (e2 >= net.Mpad && e2 < net.Mpad + net.N) ? e2 - net.Mpad : -1;
        const Pending pd = fetch(k);
        const double beta = get_beta();
        const cplx<double> x = k >= 0 ? value(pd, beta) : cplx<double>{0.0, 0.0};
        if (e2 < net.Mpad)
            vb[e2] = x;
        else if (e2 < net.Mpad + net.N)
            va[e2 - net.Mpad] = x;
        return;
    }
    // <O>.v (no conjugation) in rounds of 64 blocks x 256 parameters, four rounds per
    // batch of loads; these blocks also store the new direction
    const int db = b - pack_blocks - vab_blocks;
    const int64_t stride = int64_t(kDotParts) * 256;
    const int64_t rounds = (net.num_params + stride - 1) / stride;
    double re = 0.0, im = 0.0;
    double beta = 0.0;
    for (int64_t r0 = 0; r0 < rounds; r0 += 4) {
        Pending pd[4];
        cplx<double> o[4];
        int64_t ks[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int64_t k = (r0 + i) * stride + int64_t(db) * 256 + threadIdx.x;
            ks[i] = k < net.num_params ? k : -1;
            pd[i] = fetch(ks[i]);
            o[i] = obar[ks[i] >= 0 ? ks[i] : 0];
        }
        if (r0 == 0) beta = get_beta();
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (ks[i] < 0) continue;
            const cplx<double> x = value(pd[i], beta);
            if (cg) dir.p_new[ks[i]] = x;
            re += o[i].re * x.re - o[i].im * x.im;
            im += o[i].re * x.im + o[i].im * x.re;
        }
    }
    re = block_sum256(re, sh);
    im = block_sum256(im, sh);
    if (threadIdx.x == 0) {
        opart[db * 2] = re;
        opart[db * 2 + 1] = im;
    }
}

// u_s = w_s (sum_p upart[p][s] + sum_i s_i va_i - <O>.v), as fp32 for the spin GEMM
// (zero for count <= s < pad); four lanes per sample (a quad), each taking every
// fourth partial and site, summed by quad DPP
__device__ inline double quad_sum(double v) {
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    return v;
}

// (all loads at clamped indices first: the <O>.v partials, va, the pair partials and
// the configuration; <O>.v is then summed by every wave alone, kDotParts = 64 lanes)
static_assert(kDotParts == 64, "one <O>.v partial per lane");
constexpr int kFinishBatch = 8;   // pair partials per lane and load batch
__global__ __launch_bounds__(256) void ov_finish_center_kernel(const float2* upart, int npairs, int64_t count,
                                                               int64_t pad,
                                                               const uint64_t* confs, int words, int N,
                                                               const cplx<double>* va, const double* opart,
                                                               const double* w, float2* u) {
    __shared__ cplx<double> sva[64 * kMaxWords];   // va of up to 256 sites
    const int lane = threadIdx.x & 63;
    const double2 oc = *reinterpret_cast<const double2*>(opart + lane * 2);
    const cplx<double> vai = va && int(threadIdx.x) < N ? va[threadIdx.x] : cplx<double>{0.0, 0.0};
    const int64_t s = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 2;
    const int q = threadIdx.x & 3;
    const bool ok = s < count;
    const int64_t sc = ok ? s : count - 1;
    uint64_t cw[kMaxWords];
#pragma unroll
    for (int k = 0; k < kMaxWords; k++) cw[k] = confs[sc * words + (k < words ? k : words - 1)];
    double re = 0.0, im = 0.0;
    for (int b0 = q; b0 < npairs; b0 += 4 * kFinishBatch) {
        float2 x[kFinishBatch];
#pragma unroll
        for (int i = 0; i < kFinishBatch; i++) {
            const int b = b0 + 4 * i;
            x[i] = upart[int64_t(b < npairs ? b : npairs - 1) * count + sc];
        }
#pragma unroll
        for (int i = 0; i < kFinishBatch; i++) {
            if (b0 + 4 * i < npairs) {
                re += double(x[i].x);
                im += double(x[i].y);
            }
        }
    }
    const double c_re = wave_sum(oc.x), c_im = wave_sum(oc.y);
    if (va && int(threadIdx.x) < N) sva[threadIdx.x] = vai;
    __syncthreads();
    if (ok && va) {
        for (int i = q; i < N; i += 4) {
            const uint64_t word = i < 64 ? cw[0] : i < 128 ? cw[1] : i < 192 ? cw[2] : cw[3];
            const double si = ((word >> (i & 63)) & 1) ? 1.0 : -1.0;
            re += si * sva[i].re;
            im += si * sva[i].im;
        }
    }
    re = quad_sum(re);
    im = quad_sum(im);
    if (ok && q == 0) u[s] = float2{float(w[s] * (re - c_re)), float(w[s] * (im - c_im))};
    if (!ok && q == 0 && s < pad) u[s] = float2{0.f, 0.f};
}

// y_k (parameter order) = sum over sample chunks of the spin-GEMM partials (fp64,
// chunk order) + shift v_k, plus the block partials of v^H y.  The chunk partials of
// a parameter are loaded in batches of kYBatch at clamped indices (one memory round
// trip for up to kYBatch chunks); the site of a PsiDeep connection comes from wm
// when it is analytic, else from conn.
// Threads walk slots: the parameters in order, except the W block of a circulant
// PsiDeep, which is walked in tiles of 16 connections x 16 units (connections padded
// to a multiple of 16, padding slots idle): connection i of unit j sits on row
// (j shift + i) mod N, so a tile reads runs of up to 16 units on 16 + 15 shift rows
// instead of a different row in every lane.
constexpr int kYBatch = 32;
constexpr int kYTile = 16;
__host__ __device__ inline bool spin_y_tiled(int layout, const WParamMap& wm) {
    return layout == LAYOUT_DEEP1 && wm.shift >= 0 && !wm.sm;
}
__host__ __device__ inline int64_t spin_y_slots(int layout, int N, int M, int C, int P, const WParamMap& wm) {
    if (!spin_y_tiled(layout, wm)) return P;
    const int64_t cp = (C + kYTile - 1) / kYTile * kYTile, mp = (M + kYTile - 1) / kYTile * kYTile;
    return int64_t(N) + M + cp * mp;
}
__global__ __launch_bounds__(256) void spin_y_kernel(const float* partial, int chunks, int64_t n, int ncols,
                                                     int layout, int N, int M, int Mpad, int C, WParamMap wm,
                                                     const uint32_t* conn,
                                                     const cplx<double>* v, double shift, int P, int64_t slots,
                                                     cplx<double>* y, double* ypart) {
    __shared__ double sh[4];
    double dr = 0.0, di = 0.0;
    const bool tiled = spin_y_tiled(layout, wm);
    const int jtiles = (M + kYTile - 1) / kYTile;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < slots; e += int64_t(gridDim.x) * 256) {
        int row, col;
        int64_t k = e;
        if (layout == LAYOUT_RBM) {
            row = int(k / M);
            col = 2 * int(k % M);
        } else if (k < N) {
            row = int(k);
            col = 2 * Mpad;
        } else if (k < N + M) {
            row = N;
            col = 2 * int(k - N);
        } else if (wm.sm) {
            const int q = int(k) - N - M;
            row = q / M;
            col = 2 * (q % M);
        } else if (tiled) {
            const int64_t q = e - N - M;
            const int64_t tile = q / (kYTile * kYTile);
            const int w = int(q % (kYTile * kYTile));
            const int i = int(tile / jtiles) * kYTile + w / kYTile, j = int(tile % jtiles) * kYTile + w % kYTile;
            if (i >= C || j >= M) continue;
            k = int64_t(N) + M + int64_t(i) * M + j;
            row = wm.site(i, j);
            col = 2 * j;
        } else {
            const int q = int(k) - N - M;
            row = int(conn[q]);
            col = 2 * (q % M);
        }
        const int64_t off = int64_t(row) * ncols + col;
        const cplx<double> vk = v[k];
        double yr = 0.0, yi = 0.0;
        for (int c0 = 0; c0 < chunks; c0 += kYBatch) {
            float2 x[kYBatch];
#pragma unroll
            for (int i = 0; i < kYBatch; i++) {
                const int c = c0 + i < chunks ? c0 + i : chunks - 1;
                x[i] = *reinterpret_cast<const float2*>(partial + c * n + off);
            }
#pragma unroll
            for (int i = 0; i < kYBatch; i++) {
                if (c0 + i < chunks) {
                    yr += double(x[i].x);
                    yi += double(x[i].y);
                }
            }
        }
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

// site-major order (one thread per site-major slot; the reads of the backward
// direction are contiguous, the writes of the forward one)
__global__ void site_major_kernel(const cplx<double>* in, int N, int M, int shift, int forward, cplx<double>* out) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t P = int64_t(N) + M + int64_t(N) * M;
    if (e >= P) return;
    int64_t k = e;
    if (e >= N + M) {
        const int q = int(e - N - M), v = q / M, j = q % M;
        int i = v - int((int64_t(j) * shift) % N);
        if (i < 0) i += N;
        k = int64_t(N) + M + int64_t(i) * M + j;
    }
    if (forward)
        out[e] = in[k];
    else
        out[k] = in[e];
}

int sr_y_parts(const DenseView<float>& net, const WParamMap& wm) {
    const int64_t slots = spin_y_slots(net.layout, net.N, net.M, net.C, net.num_params, wm);
    return int(std::min<int64_t>(256, std::max<int64_t>(1, (slots + 255) / 256)));
}

int sr_kt32(int N) { return (N + 31) / 32; }

int64_t sr_count_pad(int64_t count) { return (count + kSgPad - 1) / kSgPad * kSgPad; }

SpinPlan spin_plan(int N, int M, int Mpad, int64_t count, bool with_coef) {
    SpinPlan p;
    p.vpad = (N + 1 + 15) / 16 * 16;
    p.ncols = 2 * Mpad + 16;
    p.npairs = (M + 15) / 16 + (with_coef ? 1 : 0);
    p.zgroups = (p.vpad / 16 + kSpinRowTiles - 1) / kSpinRowTiles;
    // about 512 workgroups (two per CU), chunks of whole 128-sample rounds
    int64_t chunks = std::max<int64_t>(1, 512 / (int64_t(p.npairs) * p.zgroups));
    chunks = std::min<int64_t>(chunks, std::max<int64_t>(1, (count + kSgAlign - 1) / kSgAlign));
    int64_t chunk = (count + chunks - 1) / chunks;
    chunk = (chunk + kSgAlign - 1) / kSgAlign * kSgAlign;
    p.chunk = chunk > 0 ? chunk : kSgAlign;
    p.chunks = int((count + p.chunk - 1) / p.chunk);
    if (p.chunks < 1) p.chunks = 1;
    p.n = int64_t(p.vpad) * p.ncols;
    p.bit_blocks = (count + 63) / 64;
    return p;
}

void launch_spin_bits(const uint64_t* confs, int words, int64_t count, int N, int vpad, uint64_t* bits,
                      hipStream_t stream) {
    if (count <= 0) return;
    if (vpad > kMaxVpad || words > kMaxWords) throw std::invalid_argument("too many sites for the spin bit array");
    spin_bits_kernel<<<grid_waves((count + 63) / 64), kBlock, 0, stream>>>(confs, words, count, N, vpad, bits);
    ANN_LAUNCH_CHECK();
}

template <int MODE>
void spin_gemm32_rows(int vtl, dim3 grid, hipStream_t stream, const uint64_t* bits, int vpad, int vt0,
                      const float2* g, int Mpad, int M, const float2* coef, int64_t count, int64_t chunk, int npairs,
                      int with_coef, int ncols, float* partial) {
    auto run = [&](auto kern) {
        kern<<<grid, 256, 0, stream>>>(bits, vpad, vt0, g, Mpad, M, coef, count, chunk, npairs, with_coef, ncols,
                                       partial);
    };
    switch (vtl) {
        case 1: run(spin_gemm32_kernel<MODE, 1>); break;
        case 2: run(spin_gemm32_kernel<MODE, 2>); break;
        case 3: run(spin_gemm32_kernel<MODE, 3>); break;
        case 4: run(spin_gemm32_kernel<MODE, 4>); break;
        case 5: run(spin_gemm32_kernel<MODE, 5>); break;
        case 6: run(spin_gemm32_kernel<MODE, 6>); break;
        case 7: run(spin_gemm32_kernel<MODE, 7>); break;
        default: throw std::invalid_argument("spin GEMM row group of 1..7 tiles");
    }
}

void launch_spin_gemm32(const DenseView<float>& net, const uint64_t* bits, const cplx<float>* g, int mode,
                        const cplx<float>* coef, int with_coef, int64_t count, const SpinPlan& p, float* partial,
                        hipStream_t stream) {
    if (count <= 0) return;
    if (net.N + 1 > 16 * kSpinRowTiles * p.zgroups || p.vpad > kMaxVpad)
        throw std::invalid_argument("spin GEMM plan too small");
    if (p.chunk % kSgAlign != 0) throw std::invalid_argument("spin GEMM chunks must be whole 128-sample rounds");
    if (p.npairs != (net.M + 15) / 16 + (with_coef ? 1 : 0)) throw std::invalid_argument("spin GEMM plan mismatch");
    if (int64_t(p.chunks) * p.chunk < count) throw std::invalid_argument("spin GEMM chunks do not cover the samples");
    const dim3 grid{unsigned(p.npairs * p.chunks)};
    const float2* g2 = reinterpret_cast<const float2*>(g);
    const float2* c2 = reinterpret_cast<const float2*>(coef);
    const int vtiles = p.vpad / 16;
    for (int z = 0; z < p.zgroups; z++) {
        const int vt0 = z * kSpinRowTiles;
        const int vtl = std::min(kSpinRowTiles, vtiles - vt0);
        if (mode == 0)
            spin_gemm32_rows<0>(vtl, grid, stream, bits, p.vpad, vt0, g2, net.Mpad, net.M, c2, count, p.chunk,
                                p.npairs, with_coef, p.ncols, partial);
        else if (mode == 1)
            spin_gemm32_rows<1>(vtl, grid, stream, bits, p.vpad, vt0, g2, net.Mpad, net.M, c2, count, p.chunk,
                                p.npairs, with_coef, p.ncols, partial);
        else
            spin_gemm32_rows<2>(vtl, grid, stream, bits, p.vpad, vt0, g2, net.Mpad, net.M, c2, count, p.chunk,
                                p.npairs, with_coef, p.ncols, partial);
        ANN_LAUNCH_CHECK();
    }
}

void launch_sr_pack(const DenseView<float>& net, const WParamMap& wm, const int* inv, const cplx<double>* v,
                    const CgDir& dir, const cplx<double>* obar, SrWorkspace ws, hipStream_t stream) {
    const int KT = sr_kt32(net.N), NT = 2 * net.Mpad / 16;
    if (KT > 8) throw std::invalid_argument("configurations of 1..256 sites are supported");
    if (dir.z && (!dir.p_old || !dir.p_new || !dir.newpart || !dir.oldpart || !dir.done || dir.parts < 1 ||
                  dir.parts > 256))
        throw std::invalid_argument("incomplete CG direction update");
    const int64_t npack = int64_t(KT) * 32 * net.Mpad;   // (site, unit) pairs
    const int pack_blocks = int((npack + 255) / 256);
    const int vab_blocks = (net.Mpad + net.N + 255) / 256;
    unpack_pack_kernel<<<unsigned(pack_blocks + vab_blocks + kDotParts), 256, 0, stream>>>(
        net, wm, inv, v, dir, KT, NT, pack_blocks, vab_blocks, ws.frag, ws.va, ws.vb, obar, ws.opart);
    ANN_LAUNCH_CHECK();
}

void launch_sr_ov32(const DenseView<float>& net, const uint64_t* confs, int words, int64_t count,
                    const cplx<float>* g, SrWorkspace ws, hipStream_t stream) {
    if (count <= 0) return;
    const int KT = sr_kt32(net.N), NT = 2 * net.Mpad / 16;
    if (KT > 8 || words != (KT + 1) / 2) throw std::invalid_argument("configurations of 1..256 sites are supported");
    const int npairs = sr_ov_pairs(net.M);
    const int64_t groups = (count + 15) / 16;
    // about 2048 waves (two per SIMD)
    const int slices = int(std::max<int64_t>(1, std::min<int64_t>(groups, (2048 + npairs - 1) / npairs)));
    const int64_t waves = int64_t(npairs) * slices;
    dispatch_kt(KT, [&](auto tag) {
        constexpr int kt = std::decay_t<decltype(tag)>::value;
        ov32_kernel<kt><<<unsigned((waves + kWavesPerBlock - 1) / kWavesPerBlock), kBlock, 0, stream>>>(
            reinterpret_cast<const dw4_t*>(ws.frag), NT, npairs, slices, confs, count, net.Mpad, net.M,
            reinterpret_cast<const float2*>(g), ws.vb, reinterpret_cast<float2*>(ws.upart));
    });
    ANN_LAUNCH_CHECK();
}

void launch_sr_finish(const DenseView<float>& net, const double* w, const uint64_t* confs, int words, int64_t count,
                      SrWorkspace ws, hipStream_t stream) {
    if (count <= 0) return;
    const bool deep = net.layout == LAYOUT_DEEP1;
    const int64_t pad = sr_count_pad(count);
    ov_finish_center_kernel<<<unsigned((4 * pad + 255) / 256), 256, 0, stream>>>(
        reinterpret_cast<const float2*>(ws.upart), sr_ov_pairs(net.M), count, pad, confs, words, net.N,
        deep ? ws.va : nullptr, ws.opart, w, reinterpret_cast<float2*>(ws.u));
    ANN_LAUNCH_CHECK();
}

int sr_ov_pairs(int M) { return (M + 15) / 16; }

size_t sr_frag_elems(int N, int Mpad) { return size_t(3) * sr_kt32(N) * (2 * Mpad / 16) * 64 * 8; }

void launch_site_major(const cplx<double>* in, int N, int M, int shift, bool forward, cplx<double>* out,
                       hipStream_t stream) {
    if (shift < 0) throw std::invalid_argument("site-major order needs a circulant PsiDeep");
    const int64_t P = int64_t(N) + M + int64_t(N) * M;
    site_major_kernel<<<unsigned((P + 255) / 256), 256, 0, stream>>>(in, N, M, shift, forward ? 1 : 0, out);
    ANN_LAUNCH_CHECK();
}

void launch_sr_y_bf16(const DenseView<float>& net, const WParamMap& wm, const cplx<double>* v, double shift,
                      const SpinPlan& plan, SrWorkspace ws, cplx<double>* y, hipStream_t stream) {
    spin_y_kernel<<<unsigned(sr_y_parts(net, wm)), 256, 0, stream>>>(
        ws.spin_part, plan.chunks, plan.n, plan.ncols, net.layout, net.N, net.M, net.Mpad, net.C, wm, net.conn, v,
        shift, net.num_params, spin_y_slots(net.layout, net.N, net.M, net.C, net.num_params, wm), y, ws.ypart);
    ANN_LAUNCH_CHECK();
}

void launch_spin_sums_bf16(const DenseView<float>& net, const uint64_t* confs, int words, const cplx<float>* g,
                           int mode, const cplx<double>* coef, int64_t count, cplx<double>* Yw, cplx<double>* Yb,
                           cplx<double>* Ya, hipStream_t stream, const SpinSumsWork* ws) {
    if (count <= 0) return;
    const SpinPlan p = spin_plan(net.N, net.M, net.Mpad, count, true);
    const int64_t pad = sr_count_pad(count);
    float* partial = nullptr;
    double* full = nullptr;
    uint64_t* own_bits = nullptr;
    cplx<float>* c32 = nullptr;
    const uint64_t* bits = nullptr;
    if (ws) {
        partial = ws->partial;
        full = ws->full;
        c32 = ws->c32;
        bits = ws->bits;
    } else {
        ANN_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&partial), sizeof(float) * p.n * p.chunks, stream));
        ANN_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&full), sizeof(double) * p.n, stream));
        ANN_HIP_CHECK(
            hipMallocAsync(reinterpret_cast<void**>(&own_bits), sizeof(uint64_t) * p.bit_blocks * p.vpad, stream));
        ANN_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&c32), sizeof(cplx<float>) * pad, stream));
        launch_spin_bits(confs, words, count, net.N, p.vpad, own_bits, stream);
        bits = own_bits;
    }
    coef_f32_kernel<<<unsigned((pad + 255) / 256), 256, 0, stream>>>(coef, count, pad, reinterpret_cast<float2*>(c32));
    ANN_LAUNCH_CHECK();
    // columns of the hidden-unit padding are not computed: zero
    ANN_HIP_CHECK(hipMemsetAsync(partial, 0, sizeof(float) * p.n * p.chunks, stream));
    launch_spin_gemm32(net, bits, g, mode, c32, 1, count, p, partial, stream);
    sum_partials_f32_kernel<<<unsigned((p.n + 255) / 256), 256, 0, stream>>>(partial, p.chunks, p.n, full);
    ANN_LAUNCH_CHECK();
    const int64_t ns = int64_t(net.N + 1) * net.Mpad + net.N;
    split_y_kernel<<<unsigned((ns + 255) / 256), 256, 0, stream>>>(full, net.N, net.Mpad, p.ncols, Yw, Yb, Ya);
    ANN_LAUNCH_CHECK();
    if (!ws) {
        ANN_HIP_CHECK(hipFreeAsync(partial, stream));
        ANN_HIP_CHECK(hipFreeAsync(full, stream));
        ANN_HIP_CHECK(hipFreeAsync(own_bits, stream));
        ANN_HIP_CHECK(hipFreeAsync(c32, stream));
    }
}

}  // namespace ann
