// proposal.hip -- learned block proposal: fp32 training kernels (batch build,
// LDS-tiled SGEMM forward/backward, loss, Adam) and the Metropolis-Hastings
// block sweep on gfx950.
//
// The block sweep keeps, per chain (one wavefront), the RBM angles in VGPRs as
// the single-spin sweep does, plus the first-layer pre-activations of the
// proposal network for the whole configuration (4 hidden units per lane) in
// exact 32-bit fixed point.  Masking a block subtracts its sites' columns; the
// integer sums do not depend on the order of updates, so the logits are an exact
// function of the spins outside the block and the proposal ratio of a move and
// its reverse is exactly exp(e_old - e_new).
#include "proposal.hpp"
#include "wave_net.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace ann {

namespace {

using namespace wave;

constexpr int kHPad = 256;   // hidden units of the inference form (4 per lane)

__device__ inline float bf16_to_f32(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

__device__ inline uint16_t f32_to_bf16(float x) {
    const uint32_t b = __float_as_uint(x);
    return uint16_t((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}

template <int WORDS>
__device__ inline uint32_t conf_bit(const uint64_t (&w)[WORDS], int i) {
    uint64_t x = w[0];
#pragma unroll
    for (int k = 1; k < WORDS; k++) x = (i >> 6) == k ? w[k] : x;
    return uint32_t((x >> (i & 63)) & 1u);
}

template <int WORDS>
__device__ inline void conf_flip(uint64_t (&w)[WORDS], int i) {
#pragma unroll
    for (int k = 0; k < WORDS; k++) w[k] ^= (i >> 6) == k ? uint64_t(1) << (i & 63) : 0;
}

// tanh(x) = 1 - 2 / (1 + e^{2x}) from the hardware exp and reciprocal (about 1e-6
// relative): the proposal only has to be a deterministic function of the spins
// outside the block, the Metropolis-Hastings ratio uses the same logits both ways
__device__ inline float tanh_hw(float x) {
    const float e = __expf(2.f * x);
    return 1.f - 2.f * __builtin_amdgcn_rcpf(1.f + e);
}

// the 8 logit partials of every lane summed over the wave (wave_sum8): lane l ends
// with the total of logit (l >> 3) & 7
__device__ inline float logits8_reduce(const float (&part)[kMaxBlock]) {
    static_assert(kMaxBlock == 8, "eight logits per block");
    return wave_sum8(part);
}

// ------------------------------------------------------------- block sweep
// Hidden unit 4 lane + qq of the proposal network sits in lane `lane`, slot qq, so a
// lane's four first-layer weights of a site are one 16-byte load and its four
// second-layer weights one 8-byte load.

// the first-layer columns W1q[site][4 lane .. 4 lane + 3] of a block's sites
template <int BS>
__device__ inline void load_block_cols(const BlockNet& bn, const int (&st)[kMaxBlock], int bs, int lane,
                                       int4 (&wq)[kMaxBlock]) {
    const int4* W1 = reinterpret_cast<const int4*>(bn.W1q);
#pragma unroll
    for (int j = 0; j < kMaxBlock; j++)
        wq[j] = (BS == kMaxBlock || j < bs) ? W1[int64_t(st[j]) * (kHPad / 4) + lane] : make_int4(0, 0, 0, 0);
}

// The block of proposal q is drawn from the Philox batch of its 64-proposal group, so
// the sites and the first-layer columns of proposal q + 1 are loaded while proposal q
// is evaluated (its second-layer rows are loaded at its start and arrive while the
// hidden units are formed; prefetching them too raised the registers past the point
// where the sweep got slower: 1.59 against 1.55 ms for 65 moves of 16,384 chains).
// BS = 8 (the benchmark's plaquettes) is compiled with the block size fixed; BS = 0
// reads it at run time.
template <typename T, int U, int WORDS, int BS>
__global__ __launch_bounds__(kBlock) void block_sweep_kernel(DenseView<T> net, BlockNet bn, uint64_t* chain_confs,
                                                             cplx<T>* theta, int num_chains, int64_t num_proposals,
                                                             uint64_t counter, uint32_t key0, uint32_t key1,
                                                             uint32_t* accepted, uint32_t* bstat) {
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * kWavesPerBlock + uniform_int(threadIdx.x >> 6);
    if (c >= num_chains) return;
    const int N = net.N;
    const int bs = BS == kMaxBlock ? kMaxBlock : bn.bs;
    uint64_t w[WORDS];
#pragma unroll
    for (int k = 0; k < WORDS; k++) w[k] = uniform_u64(chain_confs[int64_t(c) * WORDS + k]);
    cplx<T> th[U], f[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        th[u] = theta[int64_t(c) * net.Mpad + lane + 64 * u];
        f[u] = net.f[lane + 64 * u];
    }
    const bool realf = net.real_f != 0;
    auto log_amp = [&](const cplx<T>(&z)[U]) {
        return wave_sum(realf ? lane_re_sum<T, U, true>(z, f) : lane_re_sum<T, U, false>(z, f));
    };
    T cur = log_amp(th);
    // first-layer pre-activations of the whole configuration (4 hidden units per lane)
    const int4* W1 = reinterpret_cast<const int4*>(bn.W1q);
    int32_t a1[4];
    {
        const int4 b = reinterpret_cast<const int4*>(bn.b1q)[lane];
        a1[0] = b.x;
        a1[1] = b.y;
        a1[2] = b.z;
        a1[3] = b.w;
    }
    for (int i = 0; i < N; i++) {
        const int4 col = W1[int64_t(i) * (kHPad / 4) + lane];
        const bool up = conf_bit(w, i);
        a1[0] += up ? col.x : -col.x;
        a1[1] += up ? col.y : -col.y;
        a1[2] += up ? col.z : -col.z;
        a1[3] += up ? col.w : -col.w;
    }
    uint32_t acc = 0, nul = 0, rows = 0;
    // Philox batch of proposals q0 .. q0 + 63 (x: block, y: pattern, z: acceptance)
    auto batch = [&](int64_t q0, uint32_t& rx, uint32_t& ry, uint32_t& rz) {
        const uint64_t t = counter + uint64_t(q0) + uint64_t(lane);
        const u32x4 rr = philox(uint32_t(c), uint32_t(t), uint32_t(t >> 32), 3u, key0, key1);
        rx = rr.x;
        ry = rr.y;
        rz = rr.z;
    };
    auto block_sites = [&](uint32_t r_x, int (&st)[kMaxBlock]) {
        const int b = int(range32(r_x, uint32_t(bn.num_blocks)));
        const int* bsites = bn.blocks + int64_t(b) * bs;
#pragma unroll
        for (int j = 0; j < kMaxBlock; j++) st[j] = (BS == kMaxBlock || j < bs) ? bsites[j] : bsites[0];
    };
    uint32_t rx = 0, ry = 0, rz = 0;
    int st[kMaxBlock];
    int4 wq[kMaxBlock];
    if (num_proposals > 0) {
        batch(0, rx, ry, rz);
        block_sites(readlane_u32(rx, 0), st);
        load_block_cols<BS>(bn, st, bs, lane, wq);
    }
    for (int64_t q = 0; q < num_proposals; q++) {
        const int ql = int(q & 63);
        const uint32_t r_y = readlane_u32(ry, ql), r_z = readlane_u32(rz, ql);
        // the next proposal's block and columns, in flight while this one is evaluated
        int nst[kMaxBlock];
        int4 nwq[kMaxBlock];
        const bool more = q + 1 < num_proposals;
        if (more) {
            if (ql == 63) batch(q + 1, rx, ry, rz);
            block_sites(readlane_u32(rx, int((q + 1) & 63)), nst);
            load_block_cols<BS>(bn, nst, bs, lane, nwq);
        }
        uint2 w2[kMaxBlock];   // W2h[site][4 lane .. 4 lane + 3] (bf16 pairs)
        {
            const uint2* W2 = reinterpret_cast<const uint2*>(bn.W2h);
#pragma unroll
            for (int j = 0; j < kMaxBlock; j++)
                w2[j] = (BS == kMaxBlock || j < bs) ? W2[int64_t(st[j]) * (kHPad / 4) + lane] : make_uint2(0u, 0u);
        }
        float b2v[kMaxBlock];
        uint32_t p_old = 0;
#pragma unroll
        for (int j = 0; j < kMaxBlock; j++) {
            const bool in = BS == kMaxBlock || j < bs;
            b2v[j] = in ? bn.b2[st[j]] : 0.f;
            p_old |= in ? conf_bit(w, st[j]) << j : 0u;
        }
        const int k = __popc(p_old);
        // hidden units of the masked input
        float hd[4];
        {
            int32_t a[4] = {a1[0], a1[1], a1[2], a1[3]};
#pragma unroll
            for (int j = 0; j < kMaxBlock; j++) {
                const bool up = (p_old >> j) & 1u;
                a[0] = up ? a[0] - wq[j].x : a[0] + wq[j].x;
                a[1] = up ? a[1] - wq[j].y : a[1] + wq[j].y;
                a[2] = up ? a[2] - wq[j].z : a[2] + wq[j].z;
                a[3] = up ? a[3] - wq[j].w : a[3] + wq[j].w;
            }
#pragma unroll
            for (int qq = 0; qq < 4; qq++) hd[qq] = tanh_hw(float(a[qq]) * (1.0f / kFixScale));
        }
        // logits of the block's sites: lane partials, one transposed wave reduction
        float part[kMaxBlock];
#pragma unroll
        for (int j = 0; j < kMaxBlock; j++) {
            const uint2 v = w2[j];
            part[j] = __uint_as_float(v.x << 16) * hd[0] + __uint_as_float(v.x & 0xFFFF0000u) * hd[1] +
                      __uint_as_float(v.y << 16) * hd[2] + __uint_as_float(v.y & 0xFFFF0000u) * hd[3];
        }
        const float lsum8 = logits8_reduce(part);
        float lg[kMaxBlock];
#pragma unroll
        for (int j = 0; j < kMaxBlock; j++)
            lg[j] = (BS == kMaxBlock || j < bs)
                        ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(lsum8), 8 * j)) + b2v[j]
                        : 0.f;
        // patterns 4 lane + r with k up spins; weights exp(e_p - e_max).  Bits 2..7 of
        // the pattern are the lane's bits 0..5: their energy is shared by the 4 patterns.
        float e[4], wr[4];
        bool ok[4];
        float emax_l = -INFINITY;
        float e_hi = 0.f;
#pragma unroll
        for (int j = 2; j < kMaxBlock; j++) e_hi += (lane >> (j - 2)) & 1 ? lg[j] : 0.f;
        const int pc_hi = __popc(uint32_t(lane));
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const uint32_t pt = uint32_t(4 * lane + r);
            ok[r] = (BS == kMaxBlock || pt < (1u << bs)) && pc_hi + __popc(uint32_t(r)) == k;
            const float sv = e_hi + ((r & 1) ? lg[0] : 0.f) + ((r & 2) ? lg[1] : 0.f);
            e[r] = sv;
            if (ok[r]) emax_l = fmaxf(emax_l, sv);
        }
        const float emax = wave_max(emax_l);
        // with bn.exclusive the current pattern is not proposed (Metropolized Gibbs: the
        // proposal is the conditional law without its current value, corrected below)
        float lsum = 0.f;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const bool cur_p = bn.exclusive && uint32_t(4 * lane + r) == p_old;
            wr[r] = ok[r] && !cur_p ? __expf(e[r] - emax) : 0.f;
            lsum += wr[r];
        }
        const float incl = wave_incl_scan(lsum, lane);
        const float Z = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(incl), 63));
        const float target = (float(r_y) + 0.5f) * 2.3283064365386963e-10f * Z;
        const uint64_t hit = __ballot(incl > target);
        const uint64_t nz = __ballot(lsum > 0.f);
        const int L = hit ? int(__builtin_ctzll(hit)) : nz ? 63 - int(__builtin_clzll(nz)) : 0;
        float run = incl - lsum;
        int pick = -1, last = -1;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            if (wr[r] > 0.f) {
                run += wr[r];
                last = 4 * lane + r;
                if (pick < 0 && run > target) pick = 4 * lane + r;
            }
        }
        if (pick < 0) pick = last;
        if (pick < 0) pick = int(p_old);
        const uint32_t p_new =
            uint32_t(__builtin_amdgcn_readlane(pick, L)) & (BS == kMaxBlock ? 0xFFu : ((1u << bs) - 1u));
        const uint32_t flips = Z > 0.f ? p_old ^ p_new : 0u;
        if (flips != 0) {
            float e_old = 0.f, e_new = 0.f;
#pragma unroll
            for (int j = 0; j < kMaxBlock; j++) {
                e_old += (p_old >> j) & 1u ? lg[j] : 0.f;
                e_new += (p_new >> j) & 1u ? lg[j] : 0.f;
            }
            // log q(old | new) - log q(new | old); exclusive: q(new | old) = w_new / (Z_all - w_old)
            // with Z (the scan total) = Z_all - w_old, and q(old | new) = w_old / (Z_all - w_new)
            float lq = e_old - e_new;
            if (bn.exclusive) {
                const float w_old = __expf(e_old - emax), w_new = __expf(e_new - emax);
                lq += __logf(Z) - __logf(Z + w_old - w_new);
            }
            rows += uint32_t(__popc(flips));
            // psi ratio: theta' = theta - 2 s_j W_j over the flipped sites, whose rows are
            // loaded two at a time (a move flips an even number of sites: S_z is conserved)
            cplx<T> z[U];
#pragma unroll
            for (int u = 0; u < U; u++) z[u] = th[u];
            uint32_t fl = flips;
            while (fl) {
                const int j0 = int(__builtin_ctz(fl));
                fl &= fl - 1u;
                const int j1 = fl ? int(__builtin_ctz(fl)) : j0;
                const T c1 = fl ? ((p_old >> j1) & 1u ? T(-2) : T(2)) : T(0);
                fl &= fl ? fl - 1u : 0u;
                int s0 = st[0], s1 = st[0];
#pragma unroll
                for (int j = 0; j < kMaxBlock; j++) {
                    s0 = j == j0 ? st[j] : s0;
                    s1 = j == j1 ? st[j] : s1;
                }
                const T c0 = (p_old >> j0) & 1u ? T(-2) : T(2);
                cplx<T> row0[U], row1[U];
                load_row<T, U>(row0, net.Wd, s0, net.Mpad, lane);
                load_row<T, U>(row1, net.Wd, s1, net.Mpad, lane);
#pragma unroll
                for (int u = 0; u < U; u++) {
                    z[u].re += c0 * row0[u].re + c1 * row1[u].re;
                    z[u].im += c0 * row0[u].im + c1 * row1[u].im;
                }
            }
            const T nxt = log_amp(z);
            const T d = T(2) * (nxt - cur) + T(lq);
            const T uu = (T(r_z) + T(0.5)) * T(2.3283064365386963e-10);
            if (d > T(0) || uu <= dexp(d)) {
#pragma unroll
                for (int u = 0; u < U; u++) th[u] = z[u];
                cur = nxt;
#pragma unroll
                for (int j = 0; j < kMaxBlock; j++) {
                    if ((flips >> j) & 1u) {
                        conf_flip(w, st[j]);
                        // the site leaves the block's masked sum with its new spin: +-2 W1
                        const bool was_up = (p_old >> j) & 1u;
                        a1[0] += was_up ? -2 * wq[j].x : 2 * wq[j].x;
                        a1[1] += was_up ? -2 * wq[j].y : 2 * wq[j].y;
                        a1[2] += was_up ? -2 * wq[j].z : 2 * wq[j].z;
                        a1[3] += was_up ? -2 * wq[j].w : 2 * wq[j].w;
                    }
                }
                acc++;
            }
        } else {   // the current pattern was redrawn: nothing to evaluate, nothing changes
            nul++;
        }
        if (more) {
#pragma unroll
            for (int j = 0; j < kMaxBlock; j++) {
                st[j] = nst[j];
                wq[j] = nwq[j];
            }
        }
    }
#pragma unroll
    for (int u = 0; u < U; u++) theta[int64_t(c) * net.Mpad + lane + 64 * u] = th[u];
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < WORDS; k++) chain_confs[int64_t(c) * WORDS + k] = w[k];
        accepted[c] += acc;
        bstat[2 * c] += nul;
        bstat[2 * c + 1] += rows;
    }
}

__global__ void block_net_pack_kernel(const float* params, int N, int H, int32_t* W1q, int32_t* b1q, uint16_t* W2h,
                                      float* b2) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;   // e = i * 256 + h
    if (e >= N * kHPad) return;
    const int i = e / kHPad, h = e % kHPad;
    const float* W1 = params;
    const float* bb1 = W1 + int64_t(H) * N;
    const float* W2 = bb1 + H;
    const float* bb2 = W2 + int64_t(N) * H;
    // bf16-rounded weights, clamped to |w| <= 4 so that 257 terms stay inside int32
    auto fix = [](float x) {
        const float r = bf16_to_f32(f32_to_bf16(fminf(fmaxf(x, -4.f), 4.f)));
        return int32_t(rintf(r * kFixScale));
    };
    W1q[e] = h < H ? fix(W1[int64_t(h) * N + i]) : 0;
    W2h[e] = h < H ? f32_to_bf16(W2[int64_t(i) * H + h]) : uint16_t(0);
    if (i == 0) b1q[h] = h < H ? fix(bb1[h]) : 0;
    if (h == 0) b2[i] = bb2[i];
}

// ------------------------------------------------------------------ training
__global__ void draw_batch_kernel(int Bt, int64_t* sample, int* sites, int64_t count, const int* blocks,
                                  int num_blocks, int bs, uint32_t step_lo, uint32_t step_hi, uint32_t key0,
                                  uint32_t key1) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= Bt) return;
    const u32x4 x = philox(uint32_t(r), step_lo, step_hi, 4u, key0, key1);
    // 64-bit range reduction of the sample index for counts beyond 2^32
    const uint64_t wide = (uint64_t(x.x) << 32) | x.z;
    sample[r] = int64_t(__umul64hi(wide, uint64_t(count)));
    const int b = int(range32(x.y, uint32_t(num_blocks)));
    for (int j = 0; j < bs; j++) sites[r * bs + j] = blocks[b * bs + j];
}

__global__ void build_inputs_kernel(TrainBatch b, const uint64_t* confs, int words, int N, int bs) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= int64_t(b.Bt) * N) return;
    const int r = int(e / N), i = int(e % N);
    const uint64_t word = confs[b.sample[r] * words + (i >> 6)];
    const bool up = (word >> (i & 63)) & 1u;
    bool in_block = false;
    for (int j = 0; j < bs; j++) in_block = in_block || b.sites[r * bs + j] == i;
    b.X[e] = in_block ? 0.f : (up ? 1.f : -1.f);
    b.Y[e] = up ? 1.f : 0.f;
    b.Mask[e] = in_block ? 1.f : 0.f;
}

// C[m][n] = sum_k A(m, k) B(k, n) (+ bias[n], then tanh), fp32 on the matrix cores
// (v_mfma_f32_16x16x4_f32: fp32 operands, fp32 accumulation).  64x64 tiles staged
// through LDS 16 k at a time; wave w of the 4 takes the 32x32 quarter (w / 2, w % 2) as
// 2 x 2 MFMA tiles.  Operand layout of one 16x16x4 product: lane l supplies A[l % 16][l / 16]
// and B[l / 16][l % 16] and receives D[4 (l / 16) + r][l % 16], r < 4.
typedef float sg_float4_t __attribute__((ext_vector_type(4)));
constexpr int kTK = 16;
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void sgemm_kernel(int M, int Nn, int K, const float* __restrict__ A, int lda,
                                                    const float* __restrict__ B, int ldb,
                                                    const float* __restrict__ bias, float* __restrict__ C, int ldc,
                                                    int ksplit) {
    // split-K: block z covers k in [z ksplit, (z + 1) ksplit) and writes its own C slice
    const int kbeg = blockIdx.z * ksplit;
    const int kend = kbeg + ksplit < K ? kbeg + ksplit : K;
    C += int64_t(blockIdx.z) * M * ldc;
    __shared__ float As[kTK][64 + 4];
    __shared__ float Bs[kTK][64 + 4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int kq = lane >> 4, c = lane & 15;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    sg_float4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = sg_float4_t{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kbeg; k0 < kend; k0 += kTK) {
        for (int e = threadIdx.x; e < 64 * kTK; e += 256) {
            int m, k;
            if constexpr (TA) {
                k = e / 64;
                m = e % 64;
            } else {
                m = e / kTK;
                k = e % kTK;
            }
            const int gm = m0 + m, gk = k0 + k;
            float v = 0.f;
            if (gm < M && gk < kend) v = TA ? A[int64_t(gk) * lda + gm] : A[int64_t(gm) * lda + gk];
            As[k][m] = v;
            int n;
            if constexpr (TB) {
                n = e / kTK;
                k = e % kTK;
            } else {
                k = e / 64;
                n = e % 64;
            }
            const int gn = n0 + n, gk2 = k0 + k;
            float u = 0.f;
            if (gn < Nn && gk2 < kend) u = TB ? B[int64_t(gn) * ldb + gk2] : B[int64_t(gk2) * ldb + gn];
            Bs[k][n] = u;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kTK; kk += 4) {
            float a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; i++) a[i] = As[kk + kq][wm + 16 * i + c];
#pragma unroll
            for (int j = 0; j < 2; j++) b[j] = Bs[kk + kq][wn + 16 * j + c];
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * i + 4 * kq + r, n = n0 + wn + 16 * j + c;
                if (m >= M || n >= Nn) continue;
                float v = acc[i][j][r];
                if constexpr (EPI >= 1) v += bias[n];
                if constexpr (EPI == 2) v = tanhf(v);
                C[int64_t(m) * ldc + n] = v;
            }
}

template <bool TA, bool TB, int EPI>
void sgemm(int M, int Nn, int K, const float* A, int lda, const float* B, int ldb, const float* bias, float* C,
           int ldc, hipStream_t stream, int splits = 1) {
    const int ksplit = (K + splits - 1) / splits;
    const dim3 grid(unsigned((Nn + 63) / 64), unsigned((M + 63) / 64), unsigned(splits));
    sgemm_kernel<TA, TB, EPI><<<grid, 256, 0, stream>>>(M, Nn, K, A, lda, B, ldb, bias, C, ldc, ksplit);
    ANN_LAUNCH_CHECK();
}

// out[i] = sum_z part[z n + i] (fixed order)
__global__ void reduce_splits_kernel(const float* part, int splits, int64_t n, float* out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int z = 0; z < splits; z++) s += part[z * n + i];
    out[i] = s;
}

// one wavefront per row: d loss / d logits in place, per-row loss in fp64
__global__ __launch_bounds__(256) void loss_kernel(TrainBatch b, int N, double norm) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= b.Bt) return;
    double loss = 0.0;
    for (int o = lane; o < N; o += 64) {
        const int64_t e = int64_t(r) * N + o;
        float d = 0.f;
        if (b.Mask[e] != 0.f) {
            const float l = b.A2[e], y = b.Y[e];
            const float sp = l > 0.f ? l + log1pf(__expf(-l)) : log1pf(__expf(l));
            loss += double(sp - y * l);
            d = (1.f / (1.f + __expf(-l)) - y) * float(1.0 / norm);
        }
        b.A2[e] = d;
    }
    loss = wave_sum(loss);
    if (lane == 0) b.loss_part[r] = loss / norm;
}

__global__ void dtanh_kernel(float* dH, const float* Hd, int64_t n) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < n) dH[e] *= 1.f - Hd[e] * Hd[e];
}

// out[z][n] = sum of D[r][n] over the rows of split z: 64 columns x 4 row phases, LDS fold
__global__ __launch_bounds__(256) void colsum_kernel(const float* D, int rows, int cols, int rsplit, float* out) {
    __shared__ float part[4][64];
    const int n = blockIdx.x * 64 + (threadIdx.x & 63);
    const int ph = threadIdx.x >> 6;
    const int r0 = blockIdx.y * rsplit, r1 = r0 + rsplit < rows ? r0 + rsplit : rows;
    float s = 0.f;
    if (n < cols)
        for (int r = r0 + ph; r < r1; r += 4) s += D[int64_t(r) * cols + n];
    part[ph][threadIdx.x & 63] = s;
    __syncthreads();
    if (ph == 0 && n < cols)
        out[int64_t(blockIdx.y) * cols + n] =
            part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
}

__global__ void adam_kernel(float* p, float* m, float* v, const float* g, int64_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

inline unsigned blocks_of(int64_t n, int per = 256) { return unsigned((n + per - 1) / per); }

}  // namespace

void launch_block_net_pack(const float* params, int N, int H, int32_t* W1q, int32_t* b1q, uint16_t* W2h, float* b2,
                           hipStream_t stream) {
    if (H > kHPad) throw std::invalid_argument("the proposal network has at most 256 hidden units");
    block_net_pack_kernel<<<blocks_of(int64_t(N) * kHPad), 256, 0, stream>>>(params, N, H, W1q, b1q, W2h, b2);
    ANN_LAUNCH_CHECK();
}

template <typename T>
void launch_block_sweep(const DenseView<T>& net, const BlockNet& bn, uint64_t* chain_confs, cplx<T>* theta,
                        int num_chains, int words, int64_t num_proposals, uint64_t counter, uint64_t seed,
                        uint32_t* accepted, uint32_t* bstat, hipStream_t stream) {
    if (num_chains <= 0 || num_proposals <= 0) return;
    if (bn.bs < 1 || bn.bs > kMaxBlock) throw std::invalid_argument("block size must be 1..8");
    if (bn.N != net.N) throw std::invalid_argument("proposal and wave function have different sizes");
    const uint32_t k0 = uint32_t(seed), k1 = uint32_t(seed >> 32);
    dispatch_u(net.Mpad / 64, [&](auto tag) {
        constexpr int U = std::decay_t<decltype(tag)>::value;
        const unsigned grid = grid_waves(num_chains);
        auto go = [&](auto bs_tag) {
            constexpr int BS = decltype(bs_tag)::value;
            switch (words) {
                case 1: block_sweep_kernel<T, U, 1, BS><<<grid, kBlock, 0, stream>>>(net, bn, chain_confs, theta, num_chains, num_proposals, counter, k0, k1, accepted, bstat); break;
                case 2: block_sweep_kernel<T, U, 2, BS><<<grid, kBlock, 0, stream>>>(net, bn, chain_confs, theta, num_chains, num_proposals, counter, k0, k1, accepted, bstat); break;
                case 3: block_sweep_kernel<T, U, 3, BS><<<grid, kBlock, 0, stream>>>(net, bn, chain_confs, theta, num_chains, num_proposals, counter, k0, k1, accepted, bstat); break;
                case 4: block_sweep_kernel<T, U, 4, BS><<<grid, kBlock, 0, stream>>>(net, bn, chain_confs, theta, num_chains, num_proposals, counter, k0, k1, accepted, bstat); break;
                default: throw HipError("configurations of 1..4 words are supported");
            }
        };
        // 8-site blocks of the fp32 path with the block size fixed at compile time
        if (std::is_same<T, float>::value && bn.bs == kMaxBlock)
            go(std::integral_constant<int, kMaxBlock>{});
        else
            go(std::integral_constant<int, 0>{});
    });
    ANN_LAUNCH_CHECK();
}

template void launch_block_sweep<float>(const DenseView<float>&, const BlockNet&, uint64_t*, cplx<float>*, int, int,
                                        int64_t, uint64_t, uint64_t, uint32_t*, uint32_t*, hipStream_t);
template void launch_block_sweep<double>(const DenseView<double>&, const BlockNet&, uint64_t*, cplx<double>*, int,
                                         int, int64_t, uint64_t, uint64_t, uint32_t*, uint32_t*, hipStream_t);

void launch_draw_batch(const TrainBatch& b, int64_t count, const int* blocks, int num_blocks, int bs, uint64_t step,
                       uint64_t seed, hipStream_t stream) {
    draw_batch_kernel<<<blocks_of(b.Bt), 256, 0, stream>>>(b.Bt, b.sample, b.sites, count, blocks, num_blocks, bs,
                                                           uint32_t(step), uint32_t(step >> 32), uint32_t(seed),
                                                           uint32_t(seed >> 32));
    ANN_LAUNCH_CHECK();
}

void launch_build_inputs(const TrainBatch& b, const uint64_t* confs, int words, int N, int bs, hipStream_t stream) {
    build_inputs_kernel<<<blocks_of(int64_t(b.Bt) * N), 256, 0, stream>>>(b, confs, words, N, bs);
    ANN_LAUNCH_CHECK();
}

void launch_loss_grad(const TrainBatch& b, const float* params, int N, int H, double count_norm, float* grad,
                      hipStream_t stream) {
    const float* W1 = params;
    const float* b1 = W1 + int64_t(H) * N;
    const float* W2 = b1 + H;
    const float* b2 = W2 + int64_t(N) * H;
    float* gW1 = grad;
    float* gb1 = gW1 + int64_t(H) * N;
    float* gW2 = gb1 + H;
    float* gb2 = gW2 + int64_t(N) * H;
    const int Bt = b.Bt;
    // forward: Hd = tanh(X W1^T + b1), A2 = Hd W2^T + b2
    sgemm<false, true, 2>(Bt, H, N, b.X, N, W1, N, b1, b.Hd, H, stream);
    sgemm<false, true, 1>(Bt, N, H, b.Hd, H, W2, H, b2, b.A2, N, stream);
    loss_kernel<<<blocks_of(Bt, 4), 256, 0, stream>>>(b, N, count_norm);
    ANN_LAUNCH_CHECK();
    // backward; the reductions over the batch run split over kGradSplits row ranges
    // into b.gpart and are then added in a fixed order
    const int splits = std::max(1, std::min(kGradSplits, Bt / 64));
    const int rsplit = (Bt + splits - 1) / splits;
    const int64_t nW = int64_t(N) * H;
    float* pW = b.gpart;
    float* pb = b.gpart + splits * nW;
    sgemm<true, false, 0>(N, H, Bt, b.A2, N, b.Hd, H, nullptr, pW, H, stream, splits);
    reduce_splits_kernel<<<blocks_of(nW), 256, 0, stream>>>(pW, splits, nW, gW2);
    ANN_LAUNCH_CHECK();
    colsum_kernel<<<dim3(unsigned(blocks_of(N, 64)), unsigned(splits)), 256, 0, stream>>>(b.A2, Bt, N, rsplit, pb);
    reduce_splits_kernel<<<blocks_of(N), 256, 0, stream>>>(pb, splits, N, gb2);
    ANN_LAUNCH_CHECK();
    sgemm<false, false, 0>(Bt, H, N, b.A2, N, W2, H, nullptr, b.dH, H, stream);
    dtanh_kernel<<<blocks_of(int64_t(Bt) * H), 256, 0, stream>>>(b.dH, b.Hd, int64_t(Bt) * H);
    ANN_LAUNCH_CHECK();
    sgemm<true, false, 0>(H, N, Bt, b.dH, H, b.X, N, nullptr, pW, N, stream, splits);
    reduce_splits_kernel<<<blocks_of(nW), 256, 0, stream>>>(pW, splits, nW, gW1);
    ANN_LAUNCH_CHECK();
    colsum_kernel<<<dim3(unsigned(blocks_of(H, 64)), unsigned(splits)), 256, 0, stream>>>(b.dH, Bt, H, rsplit, pb);
    reduce_splits_kernel<<<blocks_of(H), 256, 0, stream>>>(pb, splits, H, gb1);
    ANN_LAUNCH_CHECK();
}

void launch_adam(float* params, float* m, float* v, const float* grad, int64_t n, float lr, float beta1, float beta2,
                 float eps, int64_t t, hipStream_t stream) {
    const float c1 = float(1.0 - std::pow(double(beta1), double(t)));
    const float c2 = float(1.0 - std::pow(double(beta2), double(t)));
    adam_kernel<<<blocks_of(n), 256, 0, stream>>>(params, m, v, grad, n, lr, beta1, beta2, eps, c1, c2);
    ANN_LAUNCH_CHECK();
}

}  // namespace ann
