// wave_net.hpp -- device helpers shared by the wavefront-per-chain kernels of
// the dense networks (dense_net.hip, proposal.hip): one 64-lane wavefront per
// chain or sample, lane l owning the hidden units j = l + 64u, u < U.
#pragma once

#include "kernels.hpp"

#include <string>

namespace ann {
namespace wave {

constexpr int kBlock = 256;            // 4 wavefronts
constexpr int kWavesPerBlock = kBlock / kWave;

template <typename T>
__device__ inline T uniform(T v) {
    if constexpr (sizeof(T) == 4) {
        return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
    } else {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_readfirstlane(int(b & 0xffffffff));
        const int hi = __builtin_amdgcn_readfirstlane(int(b >> 32));
        return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
    }
}

__device__ inline uint64_t uniform_u64(uint64_t v) {
    const int lo = __builtin_amdgcn_readfirstlane(int(v & 0xffffffffu));
    const int hi = __builtin_amdgcn_readfirstlane(int(v >> 32));
    return (uint64_t(uint32_t(hi)) << 32) | uint32_t(lo);
}

__device__ inline Conf load_conf(const uint64_t* confs, int64_t s, int words) {
    Conf c;
#pragma unroll
    for (int w = 0; w < kMaxWords; w++) c.w[w] = w < words ? uniform_u64(confs[s * words + w]) : 0;
    return c;
}

// Re(lc(z)) for the layer-0 polynomial, expanded in u = z^2 = a + i b:
// Re(u/2 - u^2/12 + u^3/45) = a(1/2 - a/12 + a^2/45) + b^2 (1/12 - a/15)
// with q = x y (b = 2q):  = a(1/2 - a/12 + a^2/45) + q^2 (1/3 - 4a/15)
template <typename T>
__device__ inline T logcosh0_re(cplx<T> z) {
    const T a = z.re * z.re - z.im * z.im;
    const T q = z.re * z.im;
    const T q2 = q * q;
    T p = a * T(1.0 / 45.0) - T(1.0 / 12.0);
    p = p * a + T(0.5);
    return p * a + q2 * (T(1.0 / 3.0) - a * T(4.0 / 15.0));
}

// Sum_j Re(f_j lc(theta_j)) over this lane's units
template <typename T, int U, bool REALF>
__device__ inline T lane_re_sum(const cplx<T> (&th)[U], const cplx<T> (&f)[U]) {
    T s = 0;
#pragma unroll
    for (int u = 0; u < U; u++) {
        if constexpr (REALF) {
            s += f[u].re * logcosh0_re(th[u]);
        } else {
            const cplx<T> l = logcosh0(th[u]);
            s += f[u].re * l.re - f[u].im * l.im;
        }
    }
    return s;
}

template <typename T, int U>
__device__ inline cplx<T> lane_sum(const cplx<T> (&th)[U], const cplx<T> (&f)[U]) {
    cplx<T> s = {0, 0};
#pragma unroll
    for (int u = 0; u < U; u++) s = cadd(s, cmul(f[u], logcosh0(th[u])));
    return s;
}

// the same for real final weights (f[u].im = 0)
template <typename T, int U>
__device__ inline cplx<T> lane_sum_realf(const cplx<T> (&th)[U], const cplx<T> (&f)[U]) {
    cplx<T> s = {0, 0};
#pragma unroll
    for (int u = 0; u < U; u++) {
        const cplx<T> l = logcosh0(th[u]);
        s.re += f[u].re * l.re;
        s.im += f[u].re * l.im;
    }
    return s;
}

// Sum_j f_j (lc(theta'_j) - lc0_j) over this lane's units (lc0[u][lane] = lc(theta_j), in LDS)
template <typename T, int U, bool REALF>
__device__ inline cplx<T> lane_dsum(const cplx<T> (&tn)[U], const cplx<T> (&lc0)[U][64], const cplx<T> (&f)[U],
                                    int lane) {
    cplx<T> s = {0, 0};
#pragma unroll
    for (int u = 0; u < U; u++) {
        const cplx<T> l = logcosh0(tn[u]);
        const cplx<T> l0 = lc0[u][lane];
        const cplx<T> d = {l.re - l0.re, l.im - l0.im};
        if constexpr (REALF) {
            s.re += f[u].re * d.re;
            s.im += f[u].re * d.im;
        } else {
            s = cadd(s, cmul(f[u], d));
        }
    }
    return s;
}

// eight lane partials summed over the wave in one transposed reduction: lane l ends
// with the total of part[(l >> 3) & 7].  Lanes l and l + 32 exchange halves
// (permlane32 swap), then rows 2r and 2r + 1 (permlane16 swap), then lanes l and l ^ 8
// (DPP row_ror:8); the last three levels are DPP sums inside groups of 8 lanes.
__device__ inline float wave_sum8(const float (&part)[8]) {
    const int lane = threadIdx.x & 63;
    float s4[4], s2[2];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(part[i]), __float_as_uint(part[4 + i]), false,
                                                        false);
        s4[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);   // part 4 (l >> 5) + i
    }
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(s4[i]), __float_as_uint(s4[2 + i]), false,
                                                        false);
        s2[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);   // part 4 (l >> 5) + 2 ((l >> 4) & 1) + i
    }
    const bool b3 = (lane >> 3) & 1;
    float s = (b3 ? s2[1] : s2[0]) + dpp_f32<0x128>(b3 ? s2[0] : s2[1]);
    s += dpp_f32<0x141>(s);   // row_half_mirror: lanes i and 7 - i of 8
    s += dpp_f32<0xB1>(s);    // quad_perm [1, 0, 3, 2]
    s += dpp_f32<0x4E>(s);    // quad_perm [2, 3, 0, 1]
    return s;
}

template <typename T>
__device__ inline cplx<T> wave_csum(cplx<T> v) {
    return {wave_sum(v.re), wave_sum(v.im)};
}

template <typename T, int U>
__device__ inline void load_row(cplx<T> (&w)[U], const cplx<T>* Wd, int row, int Mpad, int lane) {
    const cplx<T>* p = Wd + int64_t(row) * Mpad + lane;
#pragma unroll
    for (int u = 0; u < U; u++) w[u] = p[64 * u];
}

// position of the rank-th set bit of the wave-uniform mask o (rank < popcount(o))
__device__ inline int select64_wave(uint64_t o, uint32_t rank, int lane) {
    const uint32_t below = __builtin_amdgcn_mbcnt_hi(uint32_t(o >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(o), 0u));
    const bool hit = ((o >> lane) & 1u) && below == rank;
    const uint64_t m = __ballot(hit);
    return __builtin_ctzll(m);
}

__device__ inline uint32_t readlane_u32(uint32_t v, int l) {
    return uint32_t(__builtin_amdgcn_readlane(int(v), l));
}


template <int U>
struct UTag {
    static constexpr int value = U;
};

// dispatch on the number of hidden units per lane
template <typename F>
inline void dispatch_u(int U, F&& fn) {
    switch (U) {
        case 1: fn(UTag<1>{}); break;
        case 2: fn(UTag<2>{}); break;
        case 4: fn(UTag<4>{}); break;
        case 7: fn(UTag<7>{}); break;
        case 8: fn(UTag<8>{}); break;
        case 16: fn(UTag<16>{}); break;
        default: throw HipError("unsupported hidden width (units per lane " + std::to_string(U) + ")");
    }
}

inline unsigned grid_waves(int64_t n) { return unsigned((n + kWavesPerBlock - 1) / kWavesPerBlock); }
}  // namespace wave
}  // namespace ann
