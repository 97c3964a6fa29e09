// common.hpp -- shared device/host helpers for the gfx950 VMC kernels.
//
// Complex values are stored interleaved (re, im) in the real type T (float or
// double).  All kernels are written for 64-lane CDNA wavefronts.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace ann {

constexpr int kWave = 64;
constexpr int kMaxWords = 4;   // configurations of up to 256 sites
// accelerator complex dies (XCDs) of an MI355X, each with its own L2; the dispatcher
// hands workgroup i of a launch to XCD i mod 8
constexpr int kXcds = 8;

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define ANN_HIP_CHECK(expr)                                                               \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess)                                                             \
            throw ::ann::HipError(std::string(#expr) + ": " + hipGetErrorString(_e) +     \
                                  " (" __FILE__ ":" + std::to_string(__LINE__) + ")");   \
    } while (0)

#define ANN_LAUNCH_CHECK() ANN_HIP_CHECK(hipGetLastError())

template <typename T>
struct cplx {
    T re, im;
};

template <typename T>
__host__ __device__ inline cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__host__ __device__ inline cplx<T> cadd(cplx<T> a, cplx<T> b) {
    return {a.re + b.re, a.im + b.im};
}
template <typename T>
__host__ __device__ inline cplx<T> csub(cplx<T> a, cplx<T> b) {
    return {a.re - b.re, a.im - b.im};
}
template <typename T>
__host__ __device__ inline cplx<T> cscale(cplx<T> a, T s) {
    return {a.re * s, a.im * s};
}
template <typename T>
__host__ __device__ inline cplx<T> cconj(cplx<T> a) {
    return {a.re, -a.im};
}

__device__ inline float dexp(float x) { return __expf(x); }
__device__ inline double dexp(double x) { return exp(x); }
__device__ inline float dsin(float x) { return __sinf(x); }
__device__ inline double dsin(double x) { return sin(x); }
__device__ inline float dcos(float x) { return __cosf(x); }
__device__ inline double dcos(double x) { return cos(x); }

template <typename T>
__device__ inline cplx<T> cexp(cplx<T> z) {
    const T m = dexp(z.re);
    T s, c;
    if constexpr (sizeof(T) == 4) {
        sincosf(z.im, &s, &c);
    } else {
        sincos(z.im, &s, &c);
    }
    return {m * c, m * s};
}

// Activations of the reference (include/quantum_state/psi_functions.hpp:40-48,
// 102-116): truncated Taylor polynomials of log cosh and tanh.
//   layer 0 : lc(z) = z^2/2 - z^4/12 + z^6/45,   lc'(z) = z - z^3/3 + 2 z^5/15
//   layer>=1: lc(z) = z - z^3/3 + 2 z^5/15,      lc'(z) = 1 - z^2 + 2 z^4/3
// Evaluated in Horner form on u = z^2.
template <typename T>
__host__ __device__ inline cplx<T> logcosh0(cplx<T> z) {
    const cplx<T> u = cmul(z, z);
    // u * (1/2 + u * (-1/12 + u / 45))
    cplx<T> p = {T(-1.0 / 12.0) + u.re * T(1.0 / 45.0), u.im * T(1.0 / 45.0)};
    p = cmul(u, p);
    p.re += T(0.5);
    return cmul(u, p);
}
template <typename T>
__host__ __device__ inline cplx<T> tanh0(cplx<T> z) {
    const cplx<T> u = cmul(z, z);
    // z * (1 - u / 3 + 2 u^2 / 15) = z * (1 + u * (-1/3 + 2u/15))
    cplx<T> p = {T(-1.0 / 3.0) + u.re * T(2.0 / 15.0), u.im * T(2.0 / 15.0)};
    p = cmul(u, p);
    p.re += T(1.0);
    return cmul(z, p);
}
template <typename T>
__host__ __device__ inline cplx<T> dtanh1(cplx<T> z) {
    const cplx<T> u = cmul(z, z);
    // 1 - u + 2 u^2 / 3
    cplx<T> p = {T(-1.0) + u.re * T(2.0 / 3.0), u.im * T(2.0 / 3.0)};
    p = cmul(u, p);
    p.re += T(1.0);
    return p;
}
template <typename T>
__host__ __device__ inline cplx<T> act(cplx<T> z, int layer) {
    return layer == 0 ? logcosh0(z) : tanh0(z);
}
template <typename T>
__host__ __device__ inline cplx<T> dact(cplx<T> z, int layer) {
    return layer == 0 ? tanh0(z) : dtanh1(z);
}

// ---------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy
// as 1, 2, 3", SC'11).  Counter = (chain, proposal lo, proposal hi, stream).
struct u32x4 {
    uint32_t x, y, z, w;
};

__host__ __device__ inline u32x4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                        uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = uint64_t(0xD2511F53u) * c0;
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
        const uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = uint32_t(p0 >> 32) ^ c3 ^ k1;
        c1 = uint32_t(p1);
        c3 = uint32_t(p0);
        c0 = n0;
        c2 = n2;
    }
    return {c0, c1, c2, c3};
}

// -------------------------------------------------------- configurations
struct Conf {
    uint64_t w[kMaxWords];
};

// word k of a configuration in registers, picked by masks: a select between loads of
// the words can be folded into one load at a run-time index, which would move the
// whole array to scratch memory
__device__ inline uint64_t word_sel(const Conf& c, int k) {
    uint64_t r = 0;
#pragma unroll
    for (int j = 0; j < kMaxWords; j++) r |= c.w[j] & (uint64_t(0) - uint64_t(k == j));
    return r;
}
__device__ inline int spin_bit(const Conf& c, int i) { return int((word_sel(c, i >> 6) >> (i & 63)) & 1u); }
__device__ inline void flip_bit(Conf& c, int i) {
    const uint64_t m = uint64_t(1) << (i & 63);
    const int k = i >> 6;
    c.w[0] ^= k == 0 ? m : 0;
    c.w[1] ^= k == 1 ? m : 0;
    c.w[2] ^= k == 2 ? m : 0;
    c.w[3] ^= k == 3 ? m : 0;
}

// index of the rank-th set bit of mask (words of the configuration); defined below
__device__ inline int select64(uint64_t x, uint32_t rank);
__device__ inline int select_bit(const uint64_t* m, int words, uint32_t rank) {
    // unrolled over the largest word count: a register array is never indexed at run time
    int pos = -1;
#pragma unroll
    for (int w = 0; w < kMaxWords; w++) {
        const uint32_t c = w < words ? uint32_t(__popcll(m[w])) : 0u;
        if (pos < 0 && rank < c) pos = w * 64 + select64(m[w], rank);
        rank = pos < 0 ? rank - c : rank;
    }
    return pos;
}

// word idx of a configuration in registers without a run-time array index
// (m has kMaxWords words; picked by masks, as in word_sel)
__device__ inline uint64_t word_at(const uint64_t* m, int idx) {
    uint64_t r = 0;
#pragma unroll
    for (int k = 0; k < kMaxWords; k++) r |= m[k] & (uint64_t(0) - uint64_t(idx == k));
    return r;
}

// m ^= bit i, unrolled over the words
__device__ inline void flip_bit(uint64_t* m, int i) {
    const uint64_t b = uint64_t(1) << (i & 63);
#pragma unroll
    for (int k = 0; k < kMaxWords; k++) m[k] ^= b & (uint64_t(0) - uint64_t(k == (i >> 6)));
}

// ------------------------------------------------------ wave reductions
// GFX9 DPP all-reduce: quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror and
// row_mirror give 16-lane row sums, row_bcast:15 / row_bcast:31 fold the four
// rows into lane 63, which is then read into a scalar register.  The result is
// bit-identical in every lane (one summation order).
template <int CTRL, int ROW_MASK = 0xf>
__device__ inline int dpp_i32(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xf, false);
}
template <int CTRL, int ROW_MASK = 0xf>
__device__ inline float dpp_f32(float v) {
    return __int_as_float(dpp_i32<CTRL, ROW_MASK>(__float_as_int(v)));
}
template <int CTRL, int ROW_MASK = 0xf>
__device__ inline double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_i32<CTRL, ROW_MASK>(int(b & 0xffffffff));
    const int hi = dpp_i32<CTRL, ROW_MASK>(int(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ inline float wave_sum(float v) {
    v += dpp_f32<0xB1>(v);
    v += dpp_f32<0x4E>(v);
    v += dpp_f32<0x141>(v);
    v += dpp_f32<0x140>(v);
    v += dpp_f32<0x142, 0xa>(v);
    v += dpp_f32<0x143, 0xc>(v);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ inline double wave_sum(double v) {
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    v += dpp_f64<0x141>(v);
    v += dpp_f64<0x140>(v);
    v += dpp_f64<0x142, 0xa>(v);
    v += dpp_f64<0x143, 0xc>(v);
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(int(b & 0xffffffff), 63);
    const int hi = __builtin_amdgcn_readlane(int(b >> 32), 63);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ inline float wave_max(float v) {
    v = fmaxf(v, dpp_f32<0xB1>(v));
    v = fmaxf(v, dpp_f32<0x4E>(v));
    v = fmaxf(v, dpp_f32<0x141>(v));
    v = fmaxf(v, dpp_f32<0x140>(v));
    v = fmaxf(v, __shfl_xor(v, 16));
    v = fmaxf(v, __shfl_xor(v, 32));
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// inclusive prefix sum over the 64 lanes (Hillis-Steele on ds_bpermute)
__device__ inline float wave_incl_scan(float v, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const float y = __shfl_up(v, d);
        v += lane >= d ? y : 0.0f;
    }
    return v;
}

// (x * n) >> 32: integer in [0, n) from a 32-bit random word (Lemire's
// multiply-shift range reduction, without the rejection step)
__host__ __device__ inline uint32_t range32(uint32_t x, uint32_t n) {
    return uint32_t((uint64_t(x) * n) >> 32);
}

// position of the rank-th set bit of a 64-bit word (rank < popcount(x)),
// by binary search on popcounts
__device__ inline int select64(uint64_t x, uint32_t rank) {
    int pos = 0;
#pragma unroll
    for (int width = 32; width >= 1; width >>= 1) {
        const uint64_t low = x & ((uint64_t(1) << width) - 1);
        const uint32_t c = uint32_t(__popcll(low));
        if (rank >= c) {
            rank -= c;
            x >>= width;
            pos += width;
        }
    }
    return pos;
}

__device__ inline int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }
// a load from a wave-uniform address of read-only data through the constant address
// space (the backend emits a scalar load)
template <typename T>
__device__ inline T ld_uniform(const T* p) {
    return *(const __attribute__((address_space(4))) T*)(p);
}

}  // namespace ann
