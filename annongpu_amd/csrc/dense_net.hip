// dense_net.hip -- gfx950 kernels for PsiRBM / one-hidden-layer PsiDeep.
//
// Layout: one 64-lane wavefront per Markov chain (or per sample); lane l owns the
// hidden units j = l + 64u, u < U.  The angles theta_j = b_j + sum_v Wd[v][j] s_v
// of a chain live in VGPRs for the whole sweep; a proposal flipping sites i (and
// k) updates them with one (two) rows of Wd, the log-cosh sum is a wavefront
// reduction, and the Metropolis decision is made wave-uniform.
//
// Reference behaviour restated here:
//   log psi      include/quantum_state/PsiDeep.hpp:138-266, PsiRBM.hpp:69-119
//   update       include/quantum_state/PsiDeep.hpp:268-343, PsiRBM.hpp:121-157
//   Metropolis   include/ensembles/MonteCarlo.hpp:133-177
//   local energy include/operator/Operator.hpp:36-121
//   O_k          include/quantum_state/PsiDeep.hpp:345-445, PsiRBM.hpp:159-176
#include "wave_net.hpp"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <type_traits>

namespace ann {

namespace {

using namespace wave;

// ----------------------------------------------------------------- kernels

__global__ void enumerate_kernel(uint64_t* confs, int64_t first, int64_t count, int words) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    confs[i * words] = uint64_t(first + i);
    for (int w = 1; w < words; w++) confs[i * words + w] = 0;
}

__global__ void enumerate_paulis_kernel(uint64_t* confs, int64_t first, int64_t count) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint64_t x = uint64_t(first + i);
    uint64_t a = 0, b = 0;
    for (int k = 0; k < 32; k++) {
        a |= ((x >> (2 * k)) & 1u) << k;
        b |= ((x >> (2 * k + 1)) & 1u) << k;
    }
    confs[2 * i] = a;
    confs[2 * i + 1] = b;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void theta_valu_kernel(DenseView<T> net, const uint64_t* confs,
                                                            int words, int64_t count, cplx<T>* theta) {
    const int64_t s = blockIdx.x;
    if (s >= count) return;
    const Conf c = load_conf(confs, s, words);
    for (int j = threadIdx.x; j < net.Mpad; j += blockDim.x) {
        cplx<T> t = net.b[j];
        for (int v = 0; v < net.N; v++) {
            const cplx<T> w = net.Wd[int64_t(v) * net.Mpad + j];
            const T sv = spin_bit(c, v) ? T(1) : T(-1);
            t.re += sv * w.re;
            t.im += sv * w.im;
        }
        theta[s * net.Mpad + j] = t;
    }
}

// theta on MFMA.  One wavefront computes 16 samples x 64 real columns (4 tiles
// of v_mfma_f32_16x16x16_bf16).  A fragment: lane l holds row l%16, k = 4(l/16)+0..3
// (spins, exact in bf16); B fragment: k = 4(l/16)+0..3, column l%16 (pre-packed).
typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// Workgroup (column group of 4 tiles = 64 interleaved columns, 256 samples): the
// group's B fragments (nsplit x KT x 4 tiles, 6 KB per k tile) are staged in LDS once
// and shared by the 16 waves, each of which takes one tile of 16 samples (the
// fragments were read from L2 by every 16-sample wave before: 43 KB per wave at
// N = 100).  Per accumulator the matrix-core steps run in the order k tile, then split.
constexpr int kThetaWaves = 16;
__global__ __launch_bounds__(64 * kThetaWaves) void theta_mfma_kernel(DenseView<float> net, const short4_t* wfrag,
                                                                       int KT, int NT, int nsplit,
                                                                       const uint64_t* confs, int words,
                                                                       int64_t count, float* theta) {
    extern __shared__ short4_t sb[];   // [nsplit * KT][4][64]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nt0 = blockIdx.y * 4;
    const int nb = nsplit * KT * 4 * 64;
    for (int e = threadIdx.x; e < nb; e += 64 * kThetaWaves)
        sb[e] = wfrag[(int64_t(e >> 8) * NT + nt0 + ((e >> 6) & 3)) * 64 + (e & 63)];
    __syncthreads();
    const int64_t row0 = (int64_t(blockIdx.x) * kThetaWaves + wave) * 16;
    if (row0 >= count) return;   // no barrier follows
    const int64_t my_row = row0 + (lane & 15);
    // the configuration words as scalars (as an array they end up in scratch memory)
    const bool rok = my_row < count;
    const uint64_t w0 = rok ? confs[my_row * words] : 0;
    const uint64_t w1 = rok && words > 1 ? confs[my_row * words + 1] : 0;
    const uint64_t w2 = rok && words > 2 ? confs[my_row * words + 2] : 0;
    const uint64_t w3 = rok && words > 3 ? confs[my_row * words + 3] : 0;
    float4_t acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = {0.f, 0.f, 0.f, 0.f};
    const int kq = 4 * (lane >> 4);
    for (int kt = 0; kt < KT; kt++) {
        short4_t a;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int v = kt * 16 + kq + e;
            const int k = v >> 6;
            const uint64_t x = k == 0 ? w0 : k == 1 ? w1 : k == 2 ? w2 : w3;
            const short one = v < net.N ? short(0x3F80) : short(0);
            a[e] = (v < net.N && ((x >> (v & 63)) & 1u)) ? one : short(v < net.N ? 0xBF80 : 0);
        }
        for (int sp = 0; sp < nsplit; sp++) {
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const short4_t b = sb[((sp * KT + kt) * 4 + t) * 64 + lane];
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
            }
        }
    }
    const int ncol = 2 * net.Mpad;
    const float* bflat = reinterpret_cast<const float*>(net.b);
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int col = (nt0 + t) * 16 + (lane & 15);
        const float bias = bflat[col];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int64_t row = row0 + 4 * (lane >> 4) + r;
            if (row < count) theta[row * ncol + col] = acc[t][r] + bias;
        }
    }
}

template <typename T, int U>
__global__ __launch_bounds__(kBlock) void net_log_psi_kernel(DenseView<T> net, const cplx<T>* theta,
                                                             int64_t count, cplx<double> lp,
                                                             cplx<double>* log_psi, double* weight) {
    const int lane = threadIdx.x & 63;
    const int64_t s = int64_t(blockIdx.x) * kWavesPerBlock + uniform_int(threadIdx.x >> 6);
    if (s >= count) return;
    cplx<T> th[U], f[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        th[u] = theta[s * net.Mpad + lane + 64 * u];
        f[u] = net.f[lane + 64 * u];
    }
    const cplx<T> L = wave_csum(lane_sum<T, U>(th, f));
    if (lane == 0) {
        const cplx<double> r = {lp.re + double(L.re), lp.im + double(L.im)};
        log_psi[s] = r;
        if (weight) weight[s] = exp(2.0 * r.re);
    }
}

// Metropolis sweeps (MonteCarlo.hpp:133-177) with single flips or S_z-conserving
// exchanges.  Proposal t of chain c uses Philox((c, t_lo, t_hi, 0), seed):
// x -> site i = range32(x, N); y -> rank of the partner among the sites of
// opposite spin (exchange); z -> acceptance uniform (z + 0.5) 2^-32.
//
// Scalar-unit budget: the scalar ALU is shared by the four SIMDs of a CU, so the
// per-proposal scalar work is kept small: the Philox draws of 64 consecutive
// proposals are computed at once on the vector ALU (lane l draws proposal
// t0 + l) and read back with v_readlane; the partner is selected with a
// lane-parallel prefix count (mbcnt) and a ballot.  The configuration words live
// in scalar registers; the row of Wd for the next proposal's site is loaded while
// the current one is evaluated.
template <typename T, int U, int MOVE, bool REALF, int WORDS>
__global__ __launch_bounds__(kBlock) void mc_sweep_kernel(DenseView<T> net, uint64_t* chain_confs,
                                                          cplx<T>* theta, int num_chains,
                                                          int64_t num_proposals, uint64_t counter,
                                                          uint32_t key0, uint32_t key1,
                                                          uint32_t* accepted) {
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * kWavesPerBlock + uniform_int(threadIdx.x >> 6);
    if (c >= num_chains) return;
    const int N = net.N;
    uint64_t w[WORDS], msk[WORDS];
#pragma unroll
    for (int k = 0; k < WORDS; k++) {
        w[k] = uniform_u64(chain_confs[int64_t(c) * WORDS + k]);
        const int bits = N - 64 * k;
        msk[k] = bits >= 64 ? ~uint64_t(0) : bits > 0 ? (uint64_t(1) << bits) - 1 : 0;
    }
    cplx<T> th[U], f[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        th[u] = theta[int64_t(c) * net.Mpad + lane + 64 * u];
        f[u] = net.f[lane + 64 * u];
    }
    T cur = wave_sum(lane_re_sum<T, U, REALF>(th, f));
    uint32_t acc = 0;
    // random words of proposals [base, base + 64), one proposal per lane
    uint32_t rx, ry, rz;
    auto draw_batch = [&](uint64_t base) {
        const uint64_t t = base + uint64_t(lane);
        const u32x4 rr = philox(uint32_t(c), uint32_t(t), uint32_t(t >> 32), 0u, key0, key1);
        rx = rr.x;
        ry = rr.y;
        rz = rr.z;
    };
    draw_batch(counter);
    int i = int(range32(readlane_u32(rx, 0), uint32_t(N)));
    cplx<T> wa[U], wb[U];
    load_row<T, U>(wa, net.Wd, i, net.Mpad, lane);
    // one proposal: wi holds the row of site i, the row of the next site is loaded into wn
    auto proposal = [&](int64_t q, cplx<T> (&wi)[U], cplx<T> (&wn)[U]) {
        const int ql = int(q & 63);
        if (ql == 0 && q > 0) draw_batch(counter + uint64_t(q));
        const uint32_t r_y = readlane_u32(ry, ql);
        const uint32_t r_z = readlane_u32(rz, ql);
        uint64_t wsel = w[0];
#pragma unroll
        for (int k = 1; k < WORDS; k++) wsel = (i >> 6) == k ? w[k] : wsel;
        const int si_bit = int((wsel >> (i & 63)) & 1u);
        int kp = -1;
        cplx<T> wk[U];
        if constexpr (MOVE == MOVE_EXCHANGE) {
            uint64_t o[WORDS];
            uint32_t cnt[WORDS];
            uint32_t n_opp = 0;
#pragma unroll
            for (int k = 0; k < WORDS; k++) {
                o[k] = (si_bit ? ~w[k] : w[k]) & msk[k];
                cnt[k] = uint32_t(__popcll(o[k]));
                n_opp += cnt[k];
            }
            if (n_opp > 0) {
                uint32_t rank = range32(r_y, n_opp);
                uint64_t osel = o[0];
                int word = 0;
#pragma unroll
                for (int k = 0; k < WORDS - 1; k++) {
                    if (word == k && rank >= cnt[k]) {
                        rank -= cnt[k];
                        word = k + 1;
                        osel = o[k + 1];
                    }
                }
                kp = 64 * word + select64_wave(osel, rank, lane);
                load_row<T, U>(wk, net.Wd, kp, net.Mpad, lane);
            }
        }
        // next proposal's site and the prefetch of its row (always a valid row)
        const int qn = int((q + 1) & 63);
        uint32_t xn;
        if (qn != 0) {
            xn = readlane_u32(rx, qn);
        } else {
            const uint64_t tn = counter + uint64_t(q + 1);
            xn = philox(uint32_t(c), uint32_t(tn), uint32_t(tn >> 32), 0u, key0, key1).x;
        }
        const int i_next = int(range32(xn, uint32_t(N)));
        load_row<T, U>(wn, net.Wd, i_next, net.Mpad, lane);
        if (MOVE != MOVE_EXCHANGE || kp >= 0) {
            // theta' = theta - 2 s_i W_i (+ 2 s_i W_k for the exchange partner, s_k = -s_i)
            const T ci = si_bit ? T(-2) : T(2);
            cplx<T> z[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                if constexpr (MOVE == MOVE_EXCHANGE) {
                    z[u].re = th[u].re + ci * (wi[u].re - wk[u].re);
                    z[u].im = th[u].im + ci * (wi[u].im - wk[u].im);
                } else {
                    z[u].re = th[u].re + ci * wi[u].re;
                    z[u].im = th[u].im + ci * wi[u].im;
                }
            }
            T part = 0;
#pragma unroll
            for (int u = 0; u < U; u++) {
                if constexpr (REALF) {
                    part += f[u].re * logcosh0_re(z[u]);
                } else {
                    const cplx<T> l = logcosh0(z[u]);
                    part += f[u].re * l.re - f[u].im * l.im;
                }
            }
            const T nxt = wave_sum(part);
            const T d2 = T(2) * (nxt - cur);
            const T uu = (T(r_z) + T(0.5)) * T(2.3283064365386963e-10);
            // branch-free update of the angles (the decision is wave-uniform)
            const bool accept = d2 > T(0) || uu <= dexp(d2);
#pragma unroll
            for (int u = 0; u < U; u++) th[u] = accept ? z[u] : th[u];
            cur = accept ? nxt : cur;
#pragma unroll
            for (int k = 0; k < WORDS; k++) {
                uint64_t flip = (i >> 6) == k ? uint64_t(1) << (i & 63) : 0;
                if (kp >= 0 && (kp >> 6) == k) flip |= uint64_t(1) << (kp & 63);
                w[k] ^= accept ? flip : 0;
            }
            acc += accept ? 1u : 0u;
        }
        i = i_next;
    };
    for (int64_t q = 0; q < num_proposals; q += 2) {
        proposal(q, wa, wb);
        if (q + 1 < num_proposals) proposal(q + 1, wb, wa);
    }
#pragma unroll
    for (int u = 0; u < U; u++) theta[int64_t(c) * net.Mpad + lane + 64 * u] = th[u];
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < WORDS; k++) chain_confs[int64_t(c) * WORDS + k] = w[k];
        accepted[c] += acc;
    }
}

// E_loc(s) = sum_groups coef_g(s) * psi(s')/psi(s)  (Operator.hpp:36-121)
template <typename T, int U>
__global__ __launch_bounds__(kBlock, sizeof(T) == 4 ? 4 : 1) void local_energy_kernel(DenseView<T> net, OpView op,
                                                              const uint64_t* confs,
                                                              const cplx<T>* theta, int64_t count,
                                                              cplx<double>* e_loc) {
    const int lane = threadIdx.x & 63;
    const int64_t s = int64_t(blockIdx.x) * kWavesPerBlock + uniform_int(threadIdx.x >> 6);
    if (s >= count) return;
    const int words = op.words;
    const Conf conf = load_conf(confs, s, words);
    cplx<T> th[U], f[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        th[u] = theta[s * net.Mpad + lane + 64 * u];
        f[u] = net.f[lane + 64 * u];
    }
    const bool realf = net.real_f != 0;
    // log-cosh of every unit at s: a ratio is summed as per-unit differences
    // lc(theta'_j) - lc(theta_j), which keeps it accurate to the rounding of those
    // differences (a difference of the two totals loses the digits the totals share)
    // (kept in LDS: as registers they would push the kernel past 4 waves per SIMD)
    __shared__ cplx<T> lc_lds[kWavesPerBlock][U][64];
    cplx<T>(&lc0)[U][64] = lc_lds[threadIdx.x >> 6];
#pragma unroll
    for (int u = 0; u < U; u++) lc0[u][lane] = logcosh0(th[u]);
    double e_re = 0.0, e_im = 0.0;
    // off-diagonal terms wait in lane `pend` (log-ratio and coefficient) until 64 are
    // collected; then every lane evaluates its own exp / sincos once
    // (kept in T: the ratio is only as accurate as the T-precision angles anyway)
    T q_dre = 0, q_dim = 0, q_cre = 0, q_cim = 0;
    double a_re = 0.0, a_im = 0.0;
    int pend = 0;
    auto flush = [&]() {
        if (lane < pend) {
            const double m = exp(double(q_dre));
            double sn, cs;
            sincos(double(q_dim), &sn, &cs);
            const double r_re = m * cs, r_im = m * sn;
            a_re += q_cre * r_re - q_cim * r_im;
            a_im += q_cre * r_im + q_cim * r_re;
        }
        pend = 0;
    };
    // coefficient of group g at this configuration (the operator tables are read as
    // wave-uniform constants: scalar loads, which do not wait behind the row loads)
    auto group_coef = [&](int g, double& c_re, double& c_im) {   // (g, n: wave-uniform)
        c_re = 0.0;
        c_im = 0.0;
        g = uniform_int(g);
        const int b0 = ld_uniform(op.group_begin + g), b1 = ld_uniform(op.group_begin + g + 1);
        for (int n = uniform_int(b0); n < b1; n++) {
            int par = 0;
            // unrolled over the largest word count: the configuration stays in registers
#pragma unroll
            for (int w = 0; w < kMaxWords; w++)
                if (w < words) par += __popcll(~conf.w[w] & ld_uniform(op.string_zy + n * words + w));
            const double sg = (par & 1) ? -1.0 : 1.0;
            c_re += sg * ld_uniform(op.string_coef + 2 * n);
            c_im += sg * ld_uniform(op.string_coef + 2 * n + 1);
        }
    };
    for (int g = 0; g < op.num_groups; g++) {
        double c_re, c_im;
        group_coef(g, c_re, c_im);
        if (c_re == 0.0 && c_im == 0.0) continue;
        bool diag = true;
        for (int w = 0; w < words; w++) diag = diag && ld_uniform(op.group_flip + g * words + w) == 0;
        if (diag) {
            e_re += c_re;
            e_im += c_im;
            continue;
        }
        cplx<T> tn[U];
#pragma unroll
        for (int u = 0; u < U; u++) tn[u] = th[u];
        // unrolled over the largest word count: the configuration stays in registers
#pragma unroll
        for (int w = 0; w < kMaxWords; w++) {
            uint64_t m = w < words ? ld_uniform(op.group_flip + g * words + w) : 0;
            const uint64_t cw = conf.w[w];
            while (m) {
                const int b = __ffsll((unsigned long long)m) - 1;
                const int i = w * 64 + b;
                m &= m - 1;
                const T ci = ((cw >> b) & 1) ? T(-2) : T(2);
                const cplx<T>* wi = net.Wd + int64_t(i) * net.Mpad + lane;
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const cplx<T> wv = wi[64 * u];
                    tn[u].re += ci * wv.re;
                    tn[u].im += ci * wv.im;
                }
            }
        }
        const cplx<T> dL = wave_csum(realf ? lane_dsum<T, U, true>(tn, lc0, f, lane)
                                           : lane_dsum<T, U, false>(tn, lc0, f, lane));
        if (lane == pend) {
            q_dre = dL.re;
            q_dim = dL.im;
            q_cre = T(c_re);
            q_cim = T(c_im);
        }
        if (++pend == 64) flush();
    }
    flush();
    a_re = wave_sum(a_re);
    a_im = wave_sum(a_im);
    if (lane == 0) e_loc[s] = {e_re + a_re, e_im + a_im};
}

template <typename T>
__global__ void g_factor_kernel(DenseView<T> net, const cplx<T>* theta, int64_t total, cplx<T>* g) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const int j = int(e % net.Mpad);
    g[e] = cmul(net.f[j], tanh0(theta[e]));
}

// O_k rows in the reference parameter order (PsiDeep.cu:245-270 / PsiRBM W order)
template <typename T>
__global__ void materialize_ok_kernel(DenseView<T> net, const uint64_t* confs, int words,
                                      const cplx<T>* g, int64_t count, const cplx<double>* center,
                                      cplx<T>* O) {
    const int kblocks = (net.num_params + int(blockDim.x) - 1) / int(blockDim.x);
    const int64_t s = int64_t(blockIdx.x) / kblocks;
    const int k = int(int64_t(blockIdx.x) % kblocks) * blockDim.x + threadIdx.x;
    if (k >= net.num_params || s >= count) return;
    const Conf c = load_conf(confs, s, words);
    const cplx<T>* gs = g + s * net.Mpad;
    cplx<T> o;
    if (net.layout == LAYOUT_RBM) {
        const int i = k / net.M, j = k % net.M;
        o = cscale(gs[j], spin_bit(c, i) ? T(1) : T(-1));
    } else {
        if (k < net.N) {
            o = {spin_bit(c, k) ? T(1) : T(-1), T(0)};
        } else if (k < net.N + net.M) {
            o = gs[k - net.N];
        } else {
            const int q = k - net.N - net.M;
            const int j = q % net.M;
            o = cscale(gs[j], spin_bit(c, int(net.conn[q])) ? T(1) : T(-1));
        }
    }
    if (center) {
        o.re -= T(center[k].re);
        o.im -= T(center[k].im);
    }
    O[s * net.num_params + k] = o;
}

// Y[v][c] = sum_s spin_sv X_sc on fp64 MFMA (v_mfma_f64_16x16x4_f64), the
// transpose-GEMM behind <O>, F, the gradient and the O^H u half of S.v for a
// dense net.  Rows v < N are spins (+-1, exact), row N is all ones (bias
// parameters).  Columns: X_s,2j / 2j+1 = re / im of coef_s * g_sj (or
// conj g_sj), columns 2 Mpad / 2 Mpad+1 = re / im of coef_s (visible
// parameters).  One wavefront owns one 16-column tile for up to 8 row tiles and a
// chunk of samples; lane l feeds sample 4t + l/16 of step t (A: row l%16, B:
// column l%16, both at k = l/16).  Partial tiles per
// chunk are summed by sum_partials_kernel in a fixed order (deterministic).
typedef double double4_t __attribute__((ext_vector_type(4)));
constexpr int kRowTilesPerWave = 8;

template <typename T>
__global__ __launch_bounds__(64) void spin_gemm_kernel(DenseView<T> net, const uint64_t* confs, int words,
                                                       const cplx<T>* g, int conj_g, const cplx<double>* coef,
                                                       int64_t count, int64_t chunk, int vpad, int ncols,
                                                       double* partial) {
    const int lane = threadIdx.x;
    const int col = blockIdx.x * 16 + (lane & 15);
    const int kk = lane >> 4;
    const int vt0 = blockIdx.z * kRowTilesPerWave;
    const int vtiles = vpad / 16;
    const int64_t s0 = int64_t(blockIdx.y) * chunk;
    const int64_t s1 = s0 + chunk < count ? s0 + chunk : count;
    const int two_m = 2 * net.Mpad;
    const bool is_g = col < two_m;
    const bool is_coef = col >= two_m && col < two_m + 2;
    const int j = col >> 1;
    const bool imag_part = col & 1;
    double4_t acc[kRowTilesPerWave];
#pragma unroll
    for (int t = 0; t < kRowTilesPerWave; t++) acc[t] = {0.0, 0.0, 0.0, 0.0};
    constexpr int kStage = 4;   // MFMA steps whose operands are loaded together
    for (int64_t sq = s0; sq < s1; sq += 4 * kStage) {
        double x[kStage];
        uint64_t cw[kStage][kMaxWords];
#pragma unroll
        for (int st = 0; st < kStage; st++) {
            const int64_t s = sq + 4 * st + kk;
            x[st] = 0.0;
#pragma unroll
            for (int w = 0; w < kMaxWords; w++) cw[st][w] = 0;
            if (s < s1) {
                const cplx<double> cf = coef[s];
                if (is_g) {
                    const cplx<T> gv = g[s * net.Mpad + j];
                    const double gr = double(gv.re), gi = conj_g ? -double(gv.im) : double(gv.im);
                    if (conj_g == 2)   // coef_s |g_sj|^2 (diagonal of S)
                        x[st] = imag_part ? 0.0 : cf.re * (gr * gr + gi * gi);
                    else
                        x[st] = imag_part ? cf.re * gi + cf.im * gr : cf.re * gr - cf.im * gi;
                } else if (is_coef) {
                    x[st] = imag_part ? cf.im : cf.re;
                }
#pragma unroll
                for (int w = 0; w < kMaxWords; w++) cw[st][w] = w < words ? confs[s * words + w] : 0;
            }
        }
#pragma unroll
        for (int st = 0; st < kStage; st++) {
#pragma unroll
            for (int t = 0; t < kRowTilesPerWave; t++) {
                if (vt0 + t < vtiles) {
                    const int v = (vt0 + t) * 16 + (lane & 15);
                    double a = 0.0;
                    if (v < net.N) {
                        const uint64_t word = v < 64 ? cw[st][0] : v < 128 ? cw[st][1] : v < 192 ? cw[st][2] : cw[st][3];
                        a = ((word >> (v & 63)) & 1) ? 1.0 : -1.0;
                    } else if (v == net.N) {
                        a = 1.0;
                    }
                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, x[st], acc[t], 0, 0, 0);
                }
            }
        }
    }
#pragma unroll
    for (int t = 0; t < kRowTilesPerWave; t++) {
        if (vt0 + t < vtiles) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                // f64 16x16x4 result layout: item r of lane l is row l/16 + 4 r, column l%16
                const int v = (vt0 + t) * 16 + (lane >> 4) + 4 * r;
                partial[(int64_t(blockIdx.y) * vpad + v) * ncols + col] = acc[t][r];
            }
        }
    }
}

__global__ void sum_partials_kernel(const double* partial, int chunks, int64_t n, double* out) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double acc = 0.0;
    for (int c = 0; c < chunks; c++) acc += partial[c * n + e];
    out[e] = acc;
}

// Yw[v][j], Yb[j], Ya[v] from the GEMM result Y[vpad][ncols]
__global__ void split_spin_gemm_kernel(const double* Y, int N, int Mpad, int ncols, cplx<double>* Yw,
                                       cplx<double>* Yb, cplx<double>* Ya) {
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

// u_s = sum_i s_i va_i + sum_j g_sj (vb_j + SV_sj)
template <typename T, int U>
__global__ __launch_bounds__(kBlock) void structured_Ov_kernel(DenseView<T> net, const uint64_t* confs,
                                                               int words, const cplx<T>* g,
                                                               const cplx<T>* SV, const cplx<double>* vb,
                                                               const cplx<double>* va, int64_t count,
                                                               cplx<double>* u_out) {
    const int lane = threadIdx.x & 63;
    const int64_t s = int64_t(blockIdx.x) * kWavesPerBlock + uniform_int(threadIdx.x >> 6);
    if (s >= count) return;
    const Conf c = load_conf(confs, s, words);
    double re = 0, im = 0;
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int j = lane + 64 * u;
        if (j < net.M) {
            const cplx<T> gv = g[s * net.Mpad + j];
            const cplx<T> sv = SV[s * net.Mpad + j];
            double xr = double(sv.re), xi = double(sv.im);
            if (vb) {
                xr += vb[j].re;
                xi += vb[j].im;
            }
            re += double(gv.re) * xr - double(gv.im) * xi;
            im += double(gv.re) * xi + double(gv.im) * xr;
        }
    }
    if (va) {
        for (int i = lane; i < net.N; i += 64) {
            const double si = spin_bit(c, i) ? 1.0 : -1.0;
            re += si * va[i].re;
            im += si * va[i].im;
        }
    }
    re = wave_sum(re);
    im = wave_sum(im);
    if (lane == 0) u_out[s] = {re, im};
}

// bf16 of x rounded to nearest even (the host's to_bf16 bit for bit)
__device__ inline uint16_t bf16_rne(float x) {
    const uint32_t b = __float_as_uint(x);
    return uint16_t((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}

// Device copy of a dense net rebuilt from its parameter vector: Wd (fp64, fp32), the
// bias (PsiDeep) and the three-term bf16 theta fragments of pack_theta_fragments.
// Thread (v, j) owns site v, hidden unit j < M; the padding (j >= M, rows >= N)
// keeps the zeros of the full upload.  PsiDeep sums the connections of (v, j) in
// ascending i from zero, as the host packing does; with shift >= 0 (connection i of
// unit j is site (j shift + i) mod N, C <= N) the one connection i = (v - j shift) mod N
// is found directly.
__global__ void net_repack_kernel(int layout, int N, int M, int Mpad, int C, int shift, const cplx<double>* p,
                                  const uint32_t* conn, int KT, int NT, cplx<double>* Wd64, cplx<float>* Wd32,
                                  cplx<double>* b64, cplx<float>* b32, uint16_t* frag) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int v = blockIdx.y;
    if (j >= M) return;
    double wr = 0.0, wi = 0.0;
    if (layout == LAYOUT_RBM) {
        wr = p[size_t(v) * M + j].re;
        wi = p[size_t(v) * M + j].im;
    } else if (shift >= 0) {
        const cplx<double>* W = p + N + M;
        int i = v - int((int64_t(j) * shift) % N);
        if (i < 0) i += N;
        if (i < C) {
            wr += W[size_t(i) * M + j].re;
            wi += W[size_t(i) * M + j].im;
        }
    } else {
        const cplx<double>* W = p + N + M;
        for (int i = 0; i < C; i++)
            if (conn[size_t(i) * M + j] == uint32_t(v)) {
                wr += W[size_t(i) * M + j].re;
                wi += W[size_t(i) * M + j].im;
            }
    }
    if (layout != LAYOUT_RBM && v == 0) {
        b64[j] = p[N + j];
        b32[j] = {float(p[N + j].re), float(p[N + j].im)};
    }
    Wd64[size_t(v) * Mpad + j] = {wr, wi};
    Wd32[size_t(v) * Mpad + j] = {float(wr), float(wi)};
    const int kt = v / 16, kk = v % 16, lane_base = (kk / 4) * 16, e = kk % 4;
#pragma unroll
    for (int comp = 0; comp < 2; comp++) {
        const int col = 2 * j + comp, nt = col / 16, cc = col % 16;
        double r = comp ? wi : wr;
#pragma unroll
        for (int sp = 0; sp < 3; sp++) {
            const uint16_t h = bf16_rne(float(r));
            r -= double(__uint_as_float(uint32_t(h) << 16));
            frag[((((size_t(sp) * KT + kt) * NT + nt) * 64 + lane_base + cc) * 4) + e] = h;
        }
    }
}

uint16_t to_bf16(float x) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    const uint32_t round = 0x7FFFu + ((b >> 16) & 1u);
    return uint16_t((b + round) >> 16);
}
double from_bf16(uint16_t h) {
    const uint32_t b = uint32_t(h) << 16;
    float x;
    std::memcpy(&x, &b, 4);
    return double(x);
}

}  // namespace

// ---------------------------------------------------------------- wrappers

void launch_enumerate_confs(uint64_t* confs, int64_t first_index, int64_t count, int words,
                            hipStream_t stream) {
    if (count <= 0) return;
    enumerate_kernel<<<unsigned((count + 255) / 256), 256, 0, stream>>>(confs, first_index, count, words);
    ANN_LAUNCH_CHECK();
}

void launch_enumerate_paulis(uint64_t* confs, int64_t first_index, int64_t count, hipStream_t stream) {
    if (count <= 0) return;
    enumerate_paulis_kernel<<<unsigned((count + 255) / 256), 256, 0, stream>>>(confs, first_index, count);
    ANN_LAUNCH_CHECK();
}

template <typename T>
void launch_theta_valu(const DenseView<T>& net, const uint64_t* confs, int words, int64_t count,
                       cplx<T>* theta, hipStream_t stream) {
    if (count <= 0) return;
    theta_valu_kernel<T><<<unsigned(count), kBlock, 0, stream>>>(net, confs, words, count, theta);
    ANN_LAUNCH_CHECK();
}

void pack_theta_fragments(const double* Wd, int N, int Mpad, int nsplit, std::vector<uint16_t>& out,
                          int& npad) {
    npad = (N + 15) / 16 * 16;
    const int KT = npad / 16, ncol = 2 * Mpad, NT = ncol / 16;
    out.assign(size_t(nsplit) * KT * NT * 64 * 4, 0);
    for (int v = 0; v < N; v++) {
        for (int col = 0; col < ncol; col++) {
            double r = Wd[size_t(v) * ncol + col];
            const int kt = v / 16, kk = v % 16, nt = col / 16, cc = col % 16;
            const int lane = (kk / 4) * 16 + cc, e = kk % 4;
            for (int sp = 0; sp < nsplit; sp++) {
                const uint16_t h = to_bf16(float(r));
                r -= from_bf16(h);
                out[((((size_t(sp) * KT + kt) * NT + nt) * 64 + lane) * 4) + e] = h;
            }
        }
    }
}

void launch_net_repack(int layout, int N, int M, int Mpad, int C, int shift, const cplx<double>* params,
                       const uint32_t* conn, int npad, cplx<double>* Wd64, cplx<float>* Wd32, cplx<double>* b64,
                       cplx<float>* b32, uint16_t* wfrag, hipStream_t stream) {
    if (N <= 0 || M <= 0) return;
    if (npad < N || M > Mpad) throw std::invalid_argument("net_repack: inconsistent padding");
    if (layout != LAYOUT_RBM && shift < 0 && conn == nullptr)
        throw std::invalid_argument("net_repack: PsiDeep needs connections");
    if (shift >= 0 && C > N) throw std::invalid_argument("net_repack: a circulant layout has C <= N");
    dim3 grid(unsigned((M + 63) / 64), unsigned(N));
    net_repack_kernel<<<grid, 64, 0, stream>>>(layout, N, M, Mpad, C, shift, params, conn, npad / 16, 2 * Mpad / 16,
                                               Wd64, Wd32, b64, b32, wfrag);
    ANN_LAUNCH_CHECK();
}

void launch_theta_mfma(const DenseView<float>& net, const uint16_t* wfrag, int npad, const uint64_t* confs,
                       int words, int64_t count, cplx<float>* theta, hipStream_t stream) {
    if (count <= 0) return;
    const int KT = npad / 16, NT = 2 * net.Mpad / 16;
    if (NT % 4 != 0) throw std::invalid_argument("theta_mfma: column tiles in groups of 4");
    const int64_t per_block = int64_t(kThetaWaves) * 16;
    dim3 grid(unsigned((count + per_block - 1) / per_block), unsigned(NT / 4));
    // the fragments are 3 bf16 terms (pack_theta_fragments with nsplit 3); LDS for them
    // beyond the default 64 KiB (N > 160) is opted in, up to 96 KiB at 256 sites
    const size_t lds = size_t(3) * KT * 4 * 64 * sizeof(short4_t);
    if (lds > 64 * 1024)
        ANN_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(theta_mfma_kernel),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    theta_mfma_kernel<<<grid, 64 * kThetaWaves, lds, stream>>>(net, reinterpret_cast<const short4_t*>(wfrag), KT, NT, 3, confs,
                                                  words, count, reinterpret_cast<float*>(theta));
    ANN_LAUNCH_CHECK();
}

template <typename T>
void launch_net_log_psi(const DenseView<T>& net, const cplx<T>* theta, int64_t count, cplx<double> lp,
                        cplx<double>* log_psi, double* weight, hipStream_t stream) {
    if (count <= 0) return;
    dispatch_u(net.Mpad / 64, [&](auto tag) {
        constexpr int U = std::decay_t<decltype(tag)>::value;
        net_log_psi_kernel<T, U><<<grid_waves(count), kBlock, 0, stream>>>(net, theta, count, lp, log_psi,
                                                                            weight);
    });
    ANN_LAUNCH_CHECK();
}

template <typename T, int U, int WORDS>
void launch_sweep_words(const DenseView<T>& net, uint64_t* chain_confs, cplx<T>* theta, int num_chains, int move,
                        int64_t num_proposals, uint64_t counter, uint32_t k0, uint32_t k1, uint32_t* accepted,
                        hipStream_t stream) {
    const unsigned grid = grid_waves(num_chains);
    if (move == MOVE_EXCHANGE) {
        if (net.real_f)
            mc_sweep_kernel<T, U, MOVE_EXCHANGE, true, WORDS><<<grid, kBlock, 0, stream>>>(
                net, chain_confs, theta, num_chains, num_proposals, counter, k0, k1, accepted);
        else
            mc_sweep_kernel<T, U, MOVE_EXCHANGE, false, WORDS><<<grid, kBlock, 0, stream>>>(
                net, chain_confs, theta, num_chains, num_proposals, counter, k0, k1, accepted);
    } else {
        if (net.real_f)
            mc_sweep_kernel<T, U, MOVE_FLIP, true, WORDS><<<grid, kBlock, 0, stream>>>(
                net, chain_confs, theta, num_chains, num_proposals, counter, k0, k1, accepted);
        else
            mc_sweep_kernel<T, U, MOVE_FLIP, false, WORDS><<<grid, kBlock, 0, stream>>>(
                net, chain_confs, theta, num_chains, num_proposals, counter, k0, k1, accepted);
    }
}

template <typename T>
void launch_mc_sweep(const DenseView<T>& net, uint64_t* chain_confs, cplx<T>* theta, int num_chains,
                     int words, int move, int64_t num_proposals, uint64_t counter, uint64_t seed,
                     uint32_t* accepted, hipStream_t stream) {
    if (num_chains <= 0 || num_proposals <= 0) return;
    const uint32_t k0 = uint32_t(seed), k1 = uint32_t(seed >> 32);
    dispatch_u(net.Mpad / 64, [&](auto tag) {
        constexpr int U = std::decay_t<decltype(tag)>::value;
        switch (words) {
            case 1: launch_sweep_words<T, U, 1>(net, chain_confs, theta, num_chains, move, num_proposals, counter, k0, k1, accepted, stream); break;
            case 2: launch_sweep_words<T, U, 2>(net, chain_confs, theta, num_chains, move, num_proposals, counter, k0, k1, accepted, stream); break;
            case 3: launch_sweep_words<T, U, 3>(net, chain_confs, theta, num_chains, move, num_proposals, counter, k0, k1, accepted, stream); break;
            case 4: launch_sweep_words<T, U, 4>(net, chain_confs, theta, num_chains, move, num_proposals, counter, k0, k1, accepted, stream); break;
            default: throw HipError("configurations of 1..4 words are supported");
        }
    });
    ANN_LAUNCH_CHECK();
}

template <typename T>
void launch_local_energy(const DenseView<T>& net, const OpView& op, const uint64_t* confs,
                         const cplx<T>* theta, int64_t count, cplx<double>* e_loc, hipStream_t stream) {
    if (count <= 0) return;
    dispatch_u(net.Mpad / 64, [&](auto tag) {
        constexpr int U = std::decay_t<decltype(tag)>::value;
        local_energy_kernel<T, U><<<grid_waves(count), kBlock, 0, stream>>>(net, op, confs, theta, count,
                                                                             e_loc);
    });
    ANN_LAUNCH_CHECK();
}

template <typename T>
void launch_g_factor(const DenseView<T>& net, const cplx<T>* theta, int64_t count, cplx<T>* g,
                     hipStream_t stream) {
    const int64_t total = count * net.Mpad;
    if (total <= 0) return;
    g_factor_kernel<T><<<unsigned((total + 255) / 256), 256, 0, stream>>>(net, theta, total, g);
    ANN_LAUNCH_CHECK();
}

template <typename T>
void launch_materialize_ok(const DenseView<T>& net, const uint64_t* confs, int words, const cplx<T>* g,
                           int64_t count, const cplx<double>* center, cplx<T>* O, hipStream_t stream) {
    if (count <= 0) return;
    const int64_t blocks = count * ((net.num_params + 255) / 256);
    if (blocks > 2147483647) throw HipError("too many samples for O_k materialisation");
    materialize_ok_kernel<T><<<unsigned(blocks), 256, 0, stream>>>(net, confs, words, g, count, center, O);
    ANN_LAUNCH_CHECK();
}

template <typename T>
void launch_spin_weighted_sums(const DenseView<T>& net, const uint64_t* confs, int words, const cplx<T>* g,
                               int conj_g, const cplx<double>* coef, int64_t count, cplx<double>* Yw,
                               cplx<double>* Yb, cplx<double>* Ya, hipStream_t stream) {
    if (count <= 0) return;
    const int vpad = (net.N + 1 + 15) / 16 * 16;
    const int ncols = 2 * net.Mpad + 16;
    const int ctiles = ncols / 16;
    const int zgroups = (vpad / 16 + kRowTilesPerWave - 1) / kRowTilesPerWave;
    int64_t chunks = std::max<int64_t>(1, 4096 / (int64_t(ctiles) * zgroups));
    chunks = std::min<int64_t>(chunks, (count + 63) / 64);
    int64_t chunk = (count + chunks - 1) / chunks;
    chunk = (chunk + 3) / 4 * 4;
    chunks = (count + chunk - 1) / chunk;
    const int64_t n = int64_t(vpad) * ncols;
    double* partial = nullptr;
    ANN_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&partial), sizeof(double) * n * (chunks + 1), stream));
    double* full = partial + n * chunks;
    const dim3 grid{unsigned(ctiles), unsigned(chunks), unsigned(zgroups)};
    spin_gemm_kernel<T><<<grid, 64, 0, stream>>>(net, confs, words, g, conj_g, coef, count, chunk, vpad, ncols,
                                                 partial);
    ANN_LAUNCH_CHECK();
    sum_partials_kernel<<<unsigned((n + 255) / 256), 256, 0, stream>>>(partial, int(chunks), n, full);
    ANN_LAUNCH_CHECK();
    const int64_t ns = int64_t(net.N + 1) * net.Mpad + net.N;
    split_spin_gemm_kernel<<<unsigned((ns + 255) / 256), 256, 0, stream>>>(full, net.N, net.Mpad, ncols, Yw, Yb, Ya);
    ANN_LAUNCH_CHECK();
    ANN_HIP_CHECK(hipFreeAsync(partial, stream));
}

template <typename T>
void launch_structured_Ov(const DenseView<T>& net, const uint64_t* confs, int words, const cplx<T>* g,
                          const cplx<T>* SV, const cplx<double>* vb, const cplx<double>* va, int64_t count,
                          cplx<double>* u, hipStream_t stream) {
    if (count <= 0) return;
    dispatch_u(net.Mpad / 64, [&](auto tag) {
        constexpr int U = std::decay_t<decltype(tag)>::value;
        structured_Ov_kernel<T, U><<<grid_waves(count), kBlock, 0, stream>>>(net, confs, words, g, SV, vb, va,
                                                                              count, u);
    });
    ANN_LAUNCH_CHECK();
}

#define ANN_INSTANTIATE(T)                                                                                   \
    template void launch_theta_valu<T>(const DenseView<T>&, const uint64_t*, int, int64_t, cplx<T>*,        \
                                       hipStream_t);                                                         \
    template void launch_net_log_psi<T>(const DenseView<T>&, const cplx<T>*, int64_t, cplx<double>,        \
                                        cplx<double>*, double*, hipStream_t);                                \
    template void launch_mc_sweep<T>(const DenseView<T>&, uint64_t*, cplx<T>*, int, int, int, int64_t,      \
                                     uint64_t, uint64_t, uint32_t*, hipStream_t);                            \
    template void launch_local_energy<T>(const DenseView<T>&, const OpView&, const uint64_t*, const cplx<T>*, \
                                         int64_t, cplx<double>*, hipStream_t);                               \
    template void launch_g_factor<T>(const DenseView<T>&, const cplx<T>*, int64_t, cplx<T>*, hipStream_t);  \
    template void launch_materialize_ok<T>(const DenseView<T>&, const uint64_t*, int, const cplx<T>*,       \
                                           int64_t, const cplx<double>*, cplx<T>*, hipStream_t);             \
    template void launch_spin_weighted_sums<T>(const DenseView<T>&, const uint64_t*, int, const cplx<T>*,   \
                                               int, const cplx<double>*, int64_t, cplx<double>*,             \
                                               cplx<double>*, cplx<double>*, hipStream_t);                   \
    template void launch_structured_Ov<T>(const DenseView<T>&, const uint64_t*, int, const cplx<T>*,        \
                                          const cplx<T>*, const cplx<double>*, const cplx<double>*, int64_t, \
                                          cplx<double>*, hipStream_t);

ANN_INSTANTIATE(float)
ANN_INSTANTIATE(double)

}  // namespace ann
