// cnn.hip -- gfx950 kernels for the wave functions without a structured O:
// PsiCNN (periodic convolutions, dim 3) and PsiDeep with several hidden layers.
//
// One 64-lane wavefront per configuration; lane l owns the sites (units)
// x = l + 64 u.  The activations of every layer live in LDS (dynamic shared memory,
// one slice per wavefront).  Reference behaviour restated here:
//   PsiCNN forward  include/quantum_state/PsiCNN.hpp:97-175
//   PsiCNN O_k      include/quantum_state/PsiCNN.hpp:183-266
//   convolution     include/quantum_state/detail/Convolve.hpp:112-173
//   PsiDeep forward include/quantum_state/PsiDeep.hpp:171-266
//   PsiDeep O_k     include/quantum_state/PsiDeep.hpp:345-445
//   Metropolis      include/ensembles/MonteCarlo.hpp:133-177 (the reference keeps no
//                   incremental state for these networks, PsiCNN.hpp:177-181: every
//                   proposal re-runs the forward pass, as here)
#include "cnn.hpp"

#include <algorithm>
#include <type_traits>

namespace ann {

namespace {

// input of layer l at site y for channel ci: spins for layer 0, the previous
// layer's activation otherwise (stored in A)
template <typename T>
__device__ __forceinline__ cplx<T> layer_input(int l, const uint64_t* w, const cplx<T>* A, int N, int ci, int y) {
    if (l == 0) {
        return {((word_at(w, y >> 6) >> (y & 63)) & 1) ? T(1) : T(-1), T(0)};
    }
    return A[ci * N + y];
}

// forward pass; A (max_ch * N) receives the activations of each layer in turn
// (A2 is the ping-pong partner), Z (optional) receives the pre-activations of
// every layer at Z + zoff[l].  Returns the network part of log psi.
template <typename T>
__device__ __forceinline__ cplx<T> cnn_forward(const CnnView<T>& net, const uint64_t* w, cplx<T>* A, cplx<T>* A2, cplx<T>* Z,
                               int lane) {
    const int N = net.N;
    cplx<T> total = {0, 0};
    cplx<T>* in = A;
    cplx<T>* out = A2;
    for (int l = 0; l < net.L; l++) {
        const int prev = l > 0 ? net.ch[l - 1] : 1;
        const int vol = net.vol[l];
        const int C = net.ch[l];
        const bool last = l == net.L - 1;
        for (int cj = 0; cj < C; cj++) {
            for (int x = lane; x < N; x += 64) {
                cplx<T> z = {0, 0};
                const int cls = int(net.sym[x]);
                for (int ci = 0; ci < prev; ci++) {
                    const cplx<T>* wk = net.params + net.link_off[l] + (int64_t(ci) * C + cj) * net.nsym * vol +
                                        int64_t(cls) * vol;
                    const int* nb = net.nbr + net.nbr_off[l] + int64_t(x) * vol;
                    for (int k = 0; k < vol; k++) {
                        const cplx<T> a = layer_input<T>(l, w, in, N, ci, nb[k]);
                        const cplx<T> q = wk[k];
                        z.re += q.re * a.re - q.im * a.im;
                        z.im += q.re * a.im + q.im * a.re;
                    }
                }
                if (Z) Z[net.zoff[l] + cj * N + x] = z;
                const cplx<T> act_v = act(z, l);
                if (last) {
                    total = cadd(total, act_v);
                } else {
                    out[cj * N + x] = act_v;
                }
            }
        }
        __syncthreads();
        cplx<T>* t = in;
        in = out;
        out = t;
    }
    const cplx<T> s = {wave_sum(total.re), wave_sum(total.im)};
    return cscale(s, net.final_factor);
}

template <typename T>
__device__ __forceinline__ cplx<T> view_forward(const CnnView<T>& net, const uint64_t* w, unsigned char* smem, int lane) {
    cplx<T>* A = reinterpret_cast<cplx<T>*>(smem);
    return cnn_forward(net, w, A, A + net.max_ch * net.N, static_cast<cplx<T>*>(nullptr), lane);
}

// PsiDeep forward: Z[zoff[l] + j] = z_l[j], Act[zoff[l] + j] = act_l(z_l[j])
template <typename T>
__device__ __forceinline__ cplx<T> mlp_forward(const MlpView<T>& net, const uint64_t* w, cplx<T>* Z, cplx<T>* Act, int lane) {
    for (int l = 0; l < net.L; l++) {
        const int M = net.size[l], C = net.conn[l];
        for (int j = lane; j < M; j += 64) {
            cplx<T> z = net.b[l][j];
            for (int i = 0; i < C; i++) {
                const int src = int(net.cn[l][i * M + j]);
                const cplx<T> q = net.W[l][i * M + j];
                if (l == 0) {
                    const T sv = ((word_at(w, src >> 6) >> (src & 63)) & 1) ? T(1) : T(-1);
                    z.re += q.re * sv;
                    z.im += q.im * sv;
                } else {
                    const cplx<T> a = Act[net.zoff[l - 1] + src];
                    z.re += q.re * a.re - q.im * a.im;
                    z.im += q.re * a.im + q.im * a.re;
                }
            }
            Z[net.zoff[l] + j] = z;
            Act[net.zoff[l] + j] = act(z, l);
        }
        __syncthreads();
    }
    const int last = net.L - 1;
    cplx<T> total = {0, 0};
    for (int j = lane; j < net.size[last]; j += 64) total = cadd(total, cmul(net.f[j], Act[net.zoff[last] + j]));
    return {wave_sum(total.re), wave_sum(total.im)};
}

template <typename T>
__device__ __forceinline__ cplx<T> view_forward(const MlpView<T>& net, const uint64_t* w, unsigned char* smem, int lane) {
    cplx<T>* Z = reinterpret_cast<cplx<T>*>(smem);
    return mlp_forward(net, w, Z, Z + net.zoff[net.L], lane);
}

// the 3 N input units of a Pauli-string configuration (a = w[0], b = w[1], N <= 64) as
// spins: bit 3i + t - 1 set when site i holds Pauli type t = a_i + 2 b_i > 0
__device__ __forceinline__ void pauli_units(const uint64_t* w, int N, uint64_t* e) {
#pragma unroll
    for (int k = 0; k < kMaxWords; k++) e[k] = 0;
    for (int i = 0; i < N; i++) {
        const int t = int((w[0] >> i) & 1u) | (int((w[1] >> i) & 1u) << 1);
        if (t) flip_bit(e, 3 * i + t - 1);   // the bit was clear
    }
}

template <typename V>
__device__ __forceinline__ cplx<typename V::real_t> view_forward(const PauliInput<V>& net, const uint64_t* w,
                                                        unsigned char* smem, int lane) {
    uint64_t e[kMaxWords];
    pauli_units(w, net.N, e);
    return view_forward(net.net, e, smem, lane);
}

// matrix-element coefficient of string (pa, pb) applied to the Pauli string (xa, xb)
// (PauliString.hpp:257-276): (-1)^{#sites with a negative Levi-Civita sign}
// (-i)^{#sites where both are non-trivial and differ}
__device__ __forceinline__ cplx<double> pauli_phase(uint64_t pa, uint64_t pb, uint64_t xa, uint64_t xb) {
    const uint64_t px = pa & ~pb, py = ~pa & pb, pz = pa & pb;
    const uint64_t xx = xa & ~xb, xy = ~xa & xb, xz = xa & xb;
    const int neg = __popcll((px & xz) | (py & xx) | (pz & xy)) & 1;
    const int ne = __popcll((pa | pb) & (xa | xb) & ((pa ^ xa) | (pb ^ xb))) & 3;
    double re = ne == 0 ? 1.0 : ne == 2 ? -1.0 : 0.0;
    double im = ne == 1 ? -1.0 : ne == 3 ? 1.0 : 0.0;
    if (neg) {
        re = -re;
        im = -im;
    }
    return {re, im};
}

// L_n(s) of PsiClassical (Operator::fast_local_energy over the strings of H_local[n])
template <typename T, typename R>
__device__ __forceinline__ cplx<double> classical_term(const ClassicalView<T, R>& v, int n, const uint64_t* w) {
    double re = 0.0, im = 0.0;
    for (int m = v.op_begin[n]; m < v.op_begin[n + 1]; m++) {
        cplx<double> f;
        if (v.basis == BASIS_PAULIS) {
            f = pauli_phase(v.sa[int64_t(m) * kMaxWords], v.sb[int64_t(m) * kMaxWords], w[0], w[1]);
        } else {
            // spin basis: (-1)^{#(Z or Y) on a down spin}; the (-i)^{#Y} is in coef
            int par = 0;
#pragma unroll
            for (int k = 0; k < kMaxWords; k++) par += __popcll(~w[k] & v.sb[int64_t(m) * kMaxWords + k]);
            f = {(par & 1) ? -1.0 : 1.0, 0.0};
        }
        const double cr = v.coef[2 * m], ci = v.coef[2 * m + 1];
        re += cr * f.re - ci * f.im;
        im += cr * f.im + ci * f.re;
    }
    return {re, im};
}

template <typename T, typename R>
__device__ __forceinline__ cplx<T> view_forward(const ClassicalView<T, R>& v, const uint64_t* w, unsigned char* smem, int lane) {
    double re = 0.0, im = 0.0;
    for (int n = lane; n < v.num_ops; n += 64) {
        const cplx<double> L = classical_term(v, n, w);
        const double pr = double(v.params[n].re), pi = double(v.params[n].im);
        re += pr * L.re - pi * L.im;
        im += pr * L.im + pi * L.re;
    }
    cplx<T> r = {T(wave_sum(re)), T(wave_sum(im))};
    if (v.has_ref) {
        r = cadd(r, v.ref_lp);
        if (v.basis == BASIS_PAULIS) {
            uint64_t e[kMaxWords];
            pauli_units(w, v.N, e);
            r = cadd(r, view_forward(v.ref, e, smem, lane));
        } else {
            r = cadd(r, view_forward(v.ref, w, smem, lane));
        }
    }
    return r;
}

// the number of sites a Metropolis move picks from
template <typename T>
__device__ __forceinline__ int view_sites(const CnnView<T>& v) { return v.N; }
template <typename T>
__device__ __forceinline__ int view_sites(const MlpView<T>& v) { return v.N; }
template <typename V>
__device__ __forceinline__ int view_sites(const PauliInput<V>& v) { return v.N; }
template <typename T, typename R>
__device__ __forceinline__ int view_sites(const ClassicalView<T, R>& v) { return v.N; }

// s is uniform over the wavefront, so these become scalar loads
__device__ __forceinline__ void load_words(const uint64_t* confs, int64_t s, int words, uint64_t* w) {
    for (int k = 0; k < kMaxWords; k++) w[k] = k < words ? confs[s * words + k] : 0;
}

template <typename V>
__global__ __launch_bounds__(64) void generic_log_psi_kernel(V net, const uint64_t* confs, int words,
                                                             int64_t count, cplx<double> lp, cplx<double>* log_psi,
                                                             double* weight) {
    using T = typename V::real_t;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    for (int64_t s = blockIdx.x; s < count; s += gridDim.x) {
        uint64_t w[kMaxWords];
        load_words(confs, s, words, w);
        const cplx<T> L = view_forward(net, w, smem, lane);
        if (lane == 0) {
            const cplx<double> r = {lp.re + double(L.re), lp.im + double(L.im)};
            log_psi[s] = r;
            if (weight) weight[s] = exp(2.0 * r.re);
        }
        __syncthreads();
    }
}

// Metropolis chains, proposals as in the dense sampler (Philox stream per chain)
template <typename V>
__global__ __launch_bounds__(64) void generic_mc_kernel(V net, uint64_t* chain_confs, int words, int num_chains,
                                                        int move, int64_t num_proposals, uint64_t counter,
                                                        uint32_t key0, uint32_t key1, uint32_t* accepted) {
    using T = typename V::real_t;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    const int c = blockIdx.x;
    if (c >= num_chains) return;
    const int N = view_sites(net);
    uint64_t w[kMaxWords];
    load_words(chain_confs, c, words, w);
    uint64_t msk[kMaxWords];
    for (int k = 0; k < kMaxWords; k++) {
        const int bits = N - 64 * k;
        msk[k] = k >= words ? 0 : bits >= 64 ? ~uint64_t(0) : bits > 0 ? (uint64_t(1) << bits) - 1 : 0;
    }
    T cur = view_forward(net, w, smem, lane).re;
    uint32_t acc = 0;
    for (int64_t q = 0; q < num_proposals; q++) {
        const uint64_t t = counter + uint64_t(q);
        const u32x4 r = philox(uint32_t(c), uint32_t(t), uint32_t(t >> 32), 0u, key0, key1);
        const int i = int(range32(r.x, uint32_t(N)));
        const int si_bit = int((word_at(w, i >> 6) >> (i & 63)) & 1u);
        int kp = -1;
        if (move == MOVE_PAULI) {
            // site i takes Pauli type r.y >> 30 (possibly its current one)
            const uint32_t ty = r.y >> 30;
            const uint64_t bit = uint64_t(1) << i;
            uint64_t wn[kMaxWords];
            for (int k = 0; k < kMaxWords; k++) wn[k] = w[k];
            wn[0] = (w[0] & ~bit) | ((ty & 1u) ? bit : 0);
            wn[1] = (w[1] & ~bit) | ((ty & 2u) ? bit : 0);
            const T nxt = view_forward(net, wn, smem, lane).re;
            const T d2 = T(2) * (nxt - cur);
            const T uu = (T(r.z) + T(0.5)) * T(2.3283064365386963e-10);
            if (d2 > T(0) || uu <= dexp(d2)) {
                for (int k = 0; k < kMaxWords; k++) w[k] = wn[k];
                cur = nxt;
                acc++;
            }
            continue;
        }
        if (move == MOVE_EXCHANGE) {
            uint64_t o[kMaxWords];
            uint32_t n_opp = 0;
            for (int k = 0; k < kMaxWords; k++) {
                o[k] = (si_bit ? ~w[k] : w[k]) & msk[k];
                n_opp += uint32_t(__popcll(o[k]));
            }
            if (n_opp == 0) continue;
            kp = select_bit(o, words, range32(r.y, n_opp));
        }
        uint64_t wn[kMaxWords];
        for (int k = 0; k < kMaxWords; k++) wn[k] = w[k];
        flip_bit(wn, i);
        if (kp >= 0) flip_bit(wn, kp);
        const T nxt = view_forward(net, wn, smem, lane).re;
        const T d2 = T(2) * (nxt - cur);
        const T uu = (T(r.z) + T(0.5)) * T(2.3283064365386963e-10);
        if (d2 > T(0) || uu <= dexp(d2)) {
            for (int k = 0; k < kMaxWords; k++) w[k] = wn[k];
            cur = nxt;
            acc++;
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < kMaxWords; k++)
            if (k < words) chain_confs[int64_t(c) * words + k] = w[k];
        accepted[c] += acc;
    }
}

template <typename V>
__global__ __launch_bounds__(64) void generic_local_energy_kernel(V net, OpView op, const uint64_t* confs,
                                                                  int64_t count, cplx<double>* e_loc) {
    using T = typename V::real_t;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    const int words = op.words;
    for (int64_t s = blockIdx.x; s < count; s += gridDim.x) {
        uint64_t w[kMaxWords];
        load_words(confs, s, words, w);
        const cplx<T> L0 = view_forward(net, w, smem, lane);
        double e_re = 0.0, e_im = 0.0;
        for (int g = 0; g < op.num_groups; g++) {
            double c_re = 0.0, c_im = 0.0;
            for (int n = op.group_begin[g]; n < op.group_begin[g + 1]; n++) {
                int par = 0;
#pragma unroll
                for (int k = 0; k < kMaxWords; k++)
                    if (k < words) par += __popcll(~w[k] & op.string_zy[int64_t(n) * words + k]);
                const double sg = (par & 1) ? -1.0 : 1.0;
                c_re += sg * op.string_coef[2 * n];
                c_im += sg * op.string_coef[2 * n + 1];
            }
            if (c_re == 0.0 && c_im == 0.0) continue;
            bool diag = true;
            uint64_t wn[kMaxWords];
            for (int k = 0; k < kMaxWords; k++) {
                const uint64_t m = k < words ? op.group_flip[int64_t(g) * words + k] : 0;
                diag = diag && m == 0;
                wn[k] = w[k] ^ m;
            }
            if (diag) {
                e_re += c_re;
                e_im += c_im;
                continue;
            }
            const cplx<T> L1 = view_forward(net, wn, smem, lane);
            const double d_re = double(L1.re) - double(L0.re), d_im = double(L1.im) - double(L0.im);
            const double m = exp(d_re);
            double sn, cs;
            sincos(d_im, &sn, &cs);
            e_re += c_re * m * cs - c_im * m * sn;
            e_im += c_re * m * sn + c_im * m * cs;
        }
        if (lane == 0) e_loc[s] = {e_re, e_im};
        __syncthreads();
    }
}

// E_loc on Pauli-string configurations: string n maps x to x' = (a_n ^ x.a, b_n ^ x.b)
// with the coefficient c_n pauli_phase(string n, x); the identity string is diagonal
template <typename V>
__global__ __launch_bounds__(64) void generic_local_energy_pauli_kernel(V net, PauliOpView op, const uint64_t* confs,
                                                                        int64_t count, cplx<double>* e_loc) {
    using T = typename V::real_t;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    for (int64_t s = blockIdx.x; s < count; s += gridDim.x) {
        uint64_t w[kMaxWords];
        load_words(confs, s, 2, w);
        const cplx<T> L0 = view_forward(net, w, smem, lane);
        double e_re = 0.0, e_im = 0.0;
        for (int n = 0; n < op.num_strings; n++) {
            const uint64_t pa = op.ab[2 * n], pb = op.ab[2 * n + 1];
            const double cr = op.coef[2 * n], ci = op.coef[2 * n + 1];
            if ((pa | pb) == 0) {
                e_re += cr;
                e_im += ci;
                continue;
            }
            const cplx<double> f = pauli_phase(pa, pb, w[0], w[1]);
            const double fr = cr * f.re - ci * f.im, fi = cr * f.im + ci * f.re;
            uint64_t wn[kMaxWords] = {w[0] ^ pa, w[1] ^ pb, 0, 0};
            const cplx<T> L1 = view_forward(net, wn, smem, lane);
            const double d_re = double(L1.re) - double(L0.re), d_im = double(L1.im) - double(L0.im);
            const double m = exp(d_re);
            double sn, cs;
            sincos(d_im, &sn, &cs);
            e_re += fr * m * cs - fi * m * sn;
            e_im += fr * m * sn + fi * m * cs;
        }
        if (lane == 0) e_loc[s] = {e_re, e_im};
        __syncthreads();
    }
}

// O_k of PsiCNN by back-propagation; row s of O is written in parameter order
// (minus center if given).  LDS: Z (all pre-activations), A, A2, G (gradients).
template <typename T>
__device__ __forceinline__ void cnn_ok_row(const CnnView<T>& net, const uint64_t* w, unsigned char* smem, cplx<T>* row,
                                           const cplx<double>* center, int lane) {
    const int N = net.N;
    cplx<T>* Z = reinterpret_cast<cplx<T>*>(smem);
    cplx<T>* A = Z + net.zoff[net.L];
    cplx<T>* A2 = A + net.max_ch * N;
    // d logpsi / d output activation of the current layer: in the forward pass's second
    // activation buffer, which is dead once the forward pass is done
    cplx<T>* G = A2;
    cplx<T>* D = A2 + net.max_ch * N;    // d logpsi / d pre-activation
    {
        cnn_forward(net, w, A, A2, Z, lane);
        const int last = net.L - 1;
        for (int e = lane; e < net.ch[last] * N; e += 64) G[e] = {net.final_factor, T(0)};
        __syncthreads();
        for (int l = last; l >= 0; l--) {
            const int prev = l > 0 ? net.ch[l - 1] : 1;
            const int C = net.ch[l];
            const int vol = net.vol[l];
            for (int e = lane; e < C * N; e += 64) D[e] = cmul(G[e], dact(Z[net.zoff[l] + e], l));
            // activations of the previous layer (inputs of layer l)
            if (l > 0)
                for (int e = lane; e < prev * N; e += 64) A[e] = act(Z[net.zoff[l - 1] + e], l - 1);
            __syncthreads();
            for (int e = lane; e < prev * N; e += 64) G[e] = {0, 0};
            __syncthreads();
            for (int ci = 0; ci < prev; ci++) {
                for (int cj = 0; cj < C; cj++) {
                    const int64_t link = net.link_off[l] + (int64_t(ci) * C + cj) * net.nsym * vol;
                    for (int k = 0; k < vol; k++) {
                        for (int cls = 0; cls < net.nsym; cls++) {
                            cplx<T> part = {0, 0};
                            for (int x = lane; x < N; x += 64) {
                                if (int(net.sym[x]) != cls) continue;
                                const int y = net.nbr[net.nbr_off[l] + int64_t(x) * vol + k];
                                part = cadd(part, cmul(D[cj * N + x], layer_input<T>(l, w, A, N, ci, y)));
                            }
                            const cplx<T> tot = {wave_sum(part.re), wave_sum(part.im)};
                            if (lane == 0) {
                                const int64_t pidx = link + int64_t(cls) * vol + k;
                                cplx<T> v = tot;
                                if (center) {
                                    v.re -= T(center[pidx].re);
                                    v.im -= T(center[pidx].im);
                                }
                                row[pidx] = v;
                            }
                        }
                        if (l > 0) {
                            // scatter into the previous layer: x -> y is a translation for fixed k
                            for (int x = lane; x < N; x += 64) {
                                const int y = net.nbr[net.nbr_off[l] + int64_t(x) * vol + k];
                                const cplx<T> wq = net.params[link + int64_t(net.sym[x]) * vol + k];
                                G[ci * N + y] = cadd(G[ci * N + y], cmul(wq, D[cj * N + x]));
                            }
                            __syncthreads();
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
}

template <typename T>
__global__ __launch_bounds__(64) void cnn_ok_kernel(CnnView<T> net, const uint64_t* confs, int words, int64_t count,
                                                    const cplx<double>* center, cplx<T>* O) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    for (int64_t s = blockIdx.x; s < count; s += gridDim.x) {
        uint64_t w[kMaxWords];
        load_words(confs, s, words, w);
        cnn_ok_row(net, w, smem, O + s * net.num_params, center, lane);
        __syncthreads();
    }
}

// ------------------------------------------------ PsiCNN local updates
// A flip changes the layer-0 pre-activations only at the consumers of the flipped
// sites (their receptive field), the layer-1 ones only at the consumers of those, and
// so on; a proposal recomputes exactly those entries from the new inputs (no running
// sums, so nothing drifts) and the log psi difference is the sum over the last layer's
// affected sites of the new minus the old per-site output.  State per wavefront (LDS):
//   A    activations of layers 0 .. L-2 at the current configuration ([C_l][N] each)
//   Tn   the same for the proposal (valid on the affected sites)
//   V    last-layer activations of the proposal ([C_{L-1}][N], affected sites)
//   out  per-site output sum_c act(z_{L-1}[c][x]) at the current configuration, nout
//        the same for the proposal, lst the affected sites of the layer at hand.
template <typename T>
struct CnnInc {
    cplx<T>* A;
    cplx<T>* Tn;
    cplx<T>* V;
    cplx<T>* out;
    cplx<T>* nout;
    int* lst;
    int aoff[kMaxCnnLayers];
};

template <typename T>
__device__ __forceinline__ CnnInc<T> cnn_inc_layout(const CnnView<T>& net, unsigned char* smem) {
    CnnInc<T> st;
    int a = 0;
    for (int l = 0; l + 1 < net.L; l++) {
        st.aoff[l] = a;
        a += net.ch[l] * net.N;
    }
    st.A = reinterpret_cast<cplx<T>*>(smem);
    st.Tn = st.A + a;
    st.V = st.Tn + a;
    st.out = st.V + net.ch[net.L - 1] * net.N;
    st.nout = st.out + net.N;
    st.lst = reinterpret_cast<int*>(st.nout + net.N);
    return st;
}

template <typename T>
size_t cnn_inc_lds(const CnnView<T>& net) {
    size_t a = 0;
    for (int l = 0; l + 1 < net.L; l++) a += size_t(net.ch[l]) * net.N;
    return sizeof(cplx<T>) * (2 * a + size_t(net.ch[net.L - 1]) * net.N + 2 * size_t(net.N)) + sizeof(int) * net.N;
}

// the read-only tables of a PsiCNN (parameters, neighbour tables, symmetry classes) in
// LDS: every multiply-add of a local update reads a weight and a neighbour index, which
// from global memory would be a dependent cache round trip each
__host__ __device__ __forceinline__ int cnn_nbr_elems(int L, const int* nbr_off, const int* vol, int N) {
    return nbr_off[L - 1] + N * vol[L - 1];
}

template <typename T>
size_t cnn_table_lds(const CnnView<T>& net) {
    // neighbour indices and symmetry classes as bytes (at most 256 sites and classes)
    const size_t p = (sizeof(cplx<T>) * size_t(net.num_params) + 15) / 16 * 16;
    const size_t nb = (size_t(cnn_nbr_elems(net.L, net.nbr_off, net.vol, net.N)) + 15) / 16 * 16;
    return p + nb + (size_t(net.N) + 15) / 16 * 16;
}

// where the local updates read the tables: LDS copies or the view's global arrays (kept
// apart from the view so that the view stays in the kernel arguments, where its small
// per-layer arrays are indexed by scalar loads)
template <typename T, typename NI, typename SI>
struct CnnTables {
    const cplx<T>* params;
    const NI* nbr;
    const SI* sym;
};

// copies the tables to smem and returns where they are
template <typename T>
__device__ __forceinline__ CnnTables<T, uint8_t, uint8_t> stage_cnn_tables(const CnnView<T>& net,
                                                                           unsigned char* smem, int lane) {
    cplx<T>* p = reinterpret_cast<cplx<T>*>(smem);
    for (int k = lane; k < net.num_params; k += 64) p[k] = net.params[k];
    uint8_t* nb = smem + (sizeof(cplx<T>) * size_t(net.num_params) + 15) / 16 * 16;
    const int nnb = cnn_nbr_elems(net.L, net.nbr_off, net.vol, net.N);
    for (int k = lane; k < nnb; k += 64) nb[k] = uint8_t(net.nbr[k]);
    uint8_t* sy = nb + (size_t(nnb) + 15) / 16 * 16;
    for (int k = lane; k < net.N; k += 64) sy[k] = uint8_t(net.sym[k]);
    __syncthreads();
    return {p, nb, sy};
}

__device__ __forceinline__ bool mask_bit(const uint64_t (&m)[kMaxWords], int y) {
    const uint64_t word = y < 64 ? m[0] : y < 128 ? m[1] : y < 192 ? m[2] : m[3];
    return (word >> (y & 63)) & 1u;
}

// the sites of mask m as a list in LDS (in increasing order); returns their number
__device__ __forceinline__ int mask_list(const uint64_t (&m)[kMaxWords], int N, int* lst, int lane) {
    int n = 0;
#pragma unroll
    for (int u = 0; u < kMaxWords; u++) {
        const int x = lane + 64 * u;
        const bool in = x < N && mask_bit(m, x);
        const uint64_t b = __ballot(in);
        const int idx = n + int(__builtin_amdgcn_mbcnt_hi(uint32_t(b >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(b), 0u)));
        if (in) lst[idx] = x;
        n += __popcll(b);
    }
    __syncthreads();
    return n;
}

// pre-activation of output channel c at site x in layer l: sum over the input channels
// ci and the kernel offsets k of params[link(l, ci, c, class(x)) + k] times the input at
// neighbour k of x (the spins w for l = 0, else in[ci N + y]).  With VMAX > 0 (an upper
// bound of the kernel volume) the neighbour indices sit in registers and the loads of one
// input channel are issued together; VMAX = 0 is the plain loop for any volume.
template <typename T, int VMAX, typename TB>
__device__ __forceinline__ cplx<T> conv_point(const CnnView<T>& net, const TB& tb, int l, int c, int x,
                                              const uint64_t* w, const cplx<T>* in) {
    const int N = net.N;
    const int prev = l > 0 ? net.ch[l - 1] : 1;
    const int vol = net.vol[l];
    const int C = net.ch[l];
    const int cls = int(tb.sym[x]);
    const auto* nb = tb.nbr + net.nbr_off[l] + int64_t(x) * vol;
    cplx<T> z = {0, 0};
    auto input = [&](int ci, int y) -> cplx<T> {
        return l == 0 ? cplx<T>{((word_at(w, y >> 6) >> (y & 63)) & 1) ? T(1) : T(-1), T(0)} : in[ci * N + y];
    };
    if constexpr (VMAX > 0) {
        int nbk[VMAX];
#pragma unroll
        for (int k = 0; k < VMAX; k++) nbk[k] = k < vol ? nb[k] : 0;
        for (int ci = 0; ci < prev; ci++) {
            const cplx<T>* wk = tb.params + net.link_off[l] + ((int64_t(ci) * C + c) * net.nsym + cls) * vol;
#pragma unroll
            for (int k = 0; k < VMAX; k++) {
                if (k < vol) {
                    const cplx<T> a = input(ci, nbk[k]);
                    const cplx<T> q = wk[k];
                    z.re += q.re * a.re - q.im * a.im;
                    z.im += q.re * a.im + q.im * a.re;
                }
            }
        }
    } else {
        for (int ci = 0; ci < prev; ci++) {
            const cplx<T>* wk = tb.params + net.link_off[l] + ((int64_t(ci) * C + c) * net.nsym + cls) * vol;
            for (int k = 0; k < vol; k++) {
                const cplx<T> a = input(ci, nb[k]);
                const cplx<T> q = wk[k];
                z.re += q.re * a.re - q.im * a.im;
                z.im += q.re * a.im + q.im * a.re;
            }
        }
    }
    return z;
}

// full forward at configuration w into the state; returns sum_x out[x] (the network
// part of log psi before the final factor)
template <typename T, int VMAX, typename TB>
__device__ __forceinline__ cplx<T> cnn_inc_init(const CnnView<T>& net, const TB& tb, const uint64_t* w,
                                                CnnInc<T>& st, int lane) {
    const int N = net.N;
    for (int l = 0; l < net.L; l++) {
        const int C = net.ch[l];
        const bool last = l == net.L - 1;
        for (int x = lane; x < N; x += 64) {
            cplx<T> osum = {0, 0};
            for (int c = 0; c < C; c++) {
                const cplx<T> z = conv_point<T, VMAX, TB>(net, tb, l, c, x, w, l > 0 ? st.A + st.aoff[l - 1] : st.A);
                const cplx<T> av = act(z, l);
                if (last) {
                    osum = cadd(osum, av);
                } else {
                    st.A[st.aoff[l] + c * N + x] = av;
                    st.Tn[st.aoff[l] + c * N + x] = av;
                }
            }
            if (last) st.out[x] = osum;
        }
        __syncthreads();
    }
    cplx<T> part = {0, 0};
    for (int x = lane; x < N; x += 64) part = cadd(part, st.out[x]);
    return {wave_sum(part.re), wave_sum(part.im)};
}

// the change of sum_x out[x] when the spins of mask F flip; R[l + 1] receives the
// sites whose layer-l pre-activations change (for a commit or a reject)
template <typename T, int VMAX, typename TB>
__device__ __forceinline__ cplx<T> cnn_inc_delta(const CnnView<T>& net, const TB& tb, const uint64_t* w,
                                                 const uint64_t (&F)[kMaxWords],
                                 CnnInc<T>& st, uint64_t (&R)[kMaxCnnLayers + 1][kMaxWords], int lane) {
    const int N = net.N;
    uint64_t wn[kMaxWords];
#pragma unroll
    for (int k = 0; k < kMaxWords; k++) {
        wn[k] = w[k] ^ F[k];
        R[0][k] = F[k];
    }
    int n = 0;
    // layers unrolled: R stays in registers (no run-time index into it)
#pragma unroll
    for (int l = 0; l < kMaxCnnLayers; l++) {
        if (l >= net.L) break;
        // dilate: the consumers of the sites changed in the layer below
#pragma unroll
        for (int k = 0; k < kMaxWords; k++) R[l + 1][k] = 0;
#pragma unroll
        for (int k = 0; k < kMaxWords; k++) {
            uint64_t m = R[l][k];
            while (m) {
                const int y = 64 * k + int(__builtin_ctzll(m));
                m &= m - 1;
                const uint64_t* cy = net.cons + (int64_t(l) * N + y) * kMaxWords;
#pragma unroll
                for (int q = 0; q < kMaxWords; q++) R[l + 1][q] |= cy[q];
            }
        }
        n = mask_list(R[l + 1], N, st.lst, lane);
        const int C = net.ch[l];
        const bool last = l == net.L - 1;
        for (int e = lane; e < C * n; e += 64) {
            const int c = e / n, x = st.lst[e % n];
            // Tn mirrors A except at the sites of this proposal (R[l]), where it holds the
            // proposal's activations
            const cplx<T> z = conv_point<T, VMAX, TB>(net, tb, l, c, x, wn, l > 0 ? st.Tn + st.aoff[l - 1] : st.Tn);
            const cplx<T> av = act(z, l);
            if (last)
                st.V[c * N + x] = av;
            else
                st.Tn[st.aoff[l] + c * N + x] = av;
        }
        __syncthreads();
    }
    const int C = net.ch[net.L - 1];
    cplx<T> part = {0, 0};
    for (int t = lane; t < n; t += 64) {
        const int x = st.lst[t];
        cplx<T> sx = {0, 0};
        for (int c = 0; c < C; c++) sx = cadd(sx, st.V[c * N + x]);
        st.nout[x] = sx;
        part.re += sx.re - st.out[x].re;
        part.im += sx.im - st.out[x].im;
    }
    __syncthreads();
    return {wave_sum(part.re), wave_sum(part.im)};
}

// the proposal of the last cnn_inc_delta is dropped: Tn mirrors A again
template <typename T>
__device__ __forceinline__ void cnn_inc_reject(const CnnView<T>& net, CnnInc<T>& st, const uint64_t (&R)[kMaxCnnLayers + 1][kMaxWords],
                               int lane) {
    const int N = net.N;
#pragma unroll
    for (int l = 0; l + 1 < kMaxCnnLayers; l++) {
        if (l + 1 >= net.L) break;
        const int n = mask_list(R[l + 1], N, st.lst, lane);
        const int C = net.ch[l];
        for (int e = lane; e < C * n; e += 64) {
            const int o = st.aoff[l] + (e / n) * N + st.lst[e % n];
            st.Tn[o] = st.A[o];
        }
        __syncthreads();
    }
}

// the proposal of the last cnn_inc_delta becomes the current state
template <typename T>
__device__ __forceinline__ void cnn_inc_commit(const CnnView<T>& net, CnnInc<T>& st, const uint64_t (&R)[kMaxCnnLayers + 1][kMaxWords],
                               int lane) {
    const int N = net.N;
#pragma unroll
    for (int l = 0; l + 1 < kMaxCnnLayers; l++) {
        if (l + 1 >= net.L) break;
        const int n = mask_list(R[l + 1], N, st.lst, lane);
        const int C = net.ch[l];
        for (int e = lane; e < C * n; e += 64) {
            const int o = st.aoff[l] + (e / n) * N + st.lst[e % n];
            st.A[o] = st.Tn[o];
        }
        __syncthreads();
    }
    uint64_t RL[kMaxWords];   // R[L] without a run-time index
#pragma unroll
    for (int k = 0; k < kMaxWords; k++) RL[k] = net.L == 1 ? R[1][k] : net.L == 2 ? R[2][k] : R[3][k];
    for (int x = lane; x < N; x += 64)
        if (mask_bit(RL, x)) st.out[x] = st.nout[x];
    __syncthreads();
}

// Metropolis chains of a PsiCNN with local updates (same proposals and Philox stream as
// the full-forward sampler)
template <typename T, int VMAX, bool STAGE>
__global__ __launch_bounds__(64) void cnn_mc_inc_kernel(CnnView<T> gnet, uint64_t* chain_confs, int words,
                                                        int num_chains, int move, int64_t num_proposals,
                                                        uint64_t counter, uint32_t key0, uint32_t key1,
                                                        uint32_t* accepted, int stage) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    const int c = blockIdx.x;
    if (c >= num_chains) return;   // the whole wave: the barriers below see all of it
    // tables in LDS when they fit (stage = their size in bytes), the state after them
    const CnnView<T>& net = gnet;
    using TB = std::conditional_t<STAGE, CnnTables<T, uint8_t, uint8_t>, CnnTables<T, int, uint32_t>>;
    TB tb;
    if constexpr (STAGE)
        tb = stage_cnn_tables(net, smem, lane);
    else
        tb = TB{net.params, net.nbr, net.sym};
    const int N = net.N;
    CnnInc<T> st = cnn_inc_layout(net, smem + stage);
    uint64_t w[kMaxWords];
    load_words(chain_confs, c, words, w);
    uint64_t msk[kMaxWords];
    for (int k = 0; k < kMaxWords; k++) {
        const int bits = N - 64 * k;
        msk[k] = k >= words ? 0 : bits >= 64 ? ~uint64_t(0) : bits > 0 ? (uint64_t(1) << bits) - 1 : 0;
    }
    T cur = cnn_inc_init<T, VMAX, TB>(net, tb, w, st, lane).re;
    uint64_t R[kMaxCnnLayers + 1][kMaxWords];
    uint32_t acc = 0;
    for (int64_t q = 0; q < num_proposals; q++) {
        const uint64_t t = counter + uint64_t(q);
        const u32x4 r = philox(uint32_t(c), uint32_t(t), uint32_t(t >> 32), 0u, key0, key1);
        const int i = int(range32(r.x, uint32_t(N)));
        const int si_bit = int((word_at(w, i >> 6) >> (i & 63)) & 1u);
        int kp = -1;
        if (move == MOVE_EXCHANGE) {
            uint64_t o[kMaxWords];
            uint32_t n_opp = 0;
            for (int k = 0; k < kMaxWords; k++) {
                o[k] = (si_bit ? ~w[k] : w[k]) & msk[k];
                n_opp += uint32_t(__popcll(o[k]));
            }
            if (n_opp == 0) continue;
            kp = select_bit(o, words, range32(r.y, n_opp));
        }
        uint64_t F[kMaxWords] = {0, 0, 0, 0};
        flip_bit(F, i);
        if (kp >= 0) flip_bit(F, kp);
        const T d = cnn_inc_delta<T, VMAX, TB>(net, tb, w, F, st, R, lane).re;
        const T d2 = T(2) * net.final_factor * d;
        const T uu = (T(r.z) + T(0.5)) * T(2.3283064365386963e-10);
        if (d2 > T(0) || uu <= dexp(d2)) {
            for (int k = 0; k < kMaxWords; k++) w[k] ^= F[k];
            cnn_inc_commit(net, st, R, lane);
            acc++;
        } else {
            cnn_inc_reject(net, st, R, lane);
        }
    }
    (void)cur;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < kMaxWords; k++)
            if (k < words) chain_confs[int64_t(c) * words + k] = w[k];
        accepted[c] += acc;
    }
}

// E_loc of PsiCNN configurations with the psi ratios from local updates
template <typename T, int VMAX, bool STAGE>
__global__ __launch_bounds__(64) void cnn_local_energy_inc_kernel(CnnView<T> gnet, OpView op, const uint64_t* confs,
                                                                  int64_t count, cplx<double>* e_loc, int stage) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    const int words = op.words;
    const CnnView<T>& net = gnet;
    using TB = std::conditional_t<STAGE, CnnTables<T, uint8_t, uint8_t>, CnnTables<T, int, uint32_t>>;
    TB tb;
    if constexpr (STAGE)
        tb = stage_cnn_tables(net, smem, lane);
    else
        tb = TB{net.params, net.nbr, net.sym};
    CnnInc<T> st = cnn_inc_layout(net, smem + stage);
    uint64_t R[kMaxCnnLayers + 1][kMaxWords];
    for (int64_t s = blockIdx.x; s < count; s += gridDim.x) {
        uint64_t w[kMaxWords];
        load_words(confs, s, words, w);
        cnn_inc_init<T, VMAX, TB>(net, tb, w, st, lane);
        double e_re = 0.0, e_im = 0.0;
        for (int g = 0; g < op.num_groups; g++) {
            double c_re = 0.0, c_im = 0.0;
            for (int n = op.group_begin[g]; n < op.group_begin[g + 1]; n++) {
                int par = 0;
#pragma unroll
                for (int k = 0; k < kMaxWords; k++)
                    if (k < words) par += __popcll(~w[k] & op.string_zy[int64_t(n) * words + k]);
                const double sg = (par & 1) ? -1.0 : 1.0;
                c_re += sg * op.string_coef[2 * n];
                c_im += sg * op.string_coef[2 * n + 1];
            }
            if (c_re == 0.0 && c_im == 0.0) continue;
            bool diag = true;
            uint64_t F[kMaxWords];
            for (int k = 0; k < kMaxWords; k++) {
                F[k] = k < words ? op.group_flip[int64_t(g) * words + k] : 0;
                diag = diag && F[k] == 0;
            }
            if (diag) {
                e_re += c_re;
                e_im += c_im;
                continue;
            }
            const cplx<T> dl = cnn_inc_delta<T, VMAX, TB>(net, tb, w, F, st, R, lane);
            cnn_inc_reject(net, st, R, lane);
            const double d_re = double(net.final_factor) * double(dl.re), d_im = double(net.final_factor) * double(dl.im);
            const double m = exp(d_re);
            double sn, cs;
            sincos(d_im, &sn, &cs);
            e_re += c_re * m * cs - c_im * m * sn;
            e_im += c_re * m * sn + c_im * m * cs;
        }
        if (lane == 0) e_loc[s] = {e_re, e_im};
        __syncthreads();
    }
}

// O_k of a PsiDeep (PsiDeep.hpp:345-445) for the spins w into row[0 .. num_params) in
// parameter order (minus center if given), by back-propagation; LDS: Z, Act (forward)
// and D (the derivative of log psi by the pre-activations of every layer), zoff[L]
// values each
template <typename T>
__device__ __forceinline__ void mlp_ok_row(const MlpView<T>& net, const uint64_t* w, unsigned char* smem, cplx<T>* row,
                           const cplx<double>* center, int lane) {
    cplx<T>* Z = reinterpret_cast<cplx<T>*>(smem);
    cplx<T>* Act = Z + net.zoff[net.L];
    cplx<T>* D = Act + net.zoff[net.L];
    auto put = [&](int k, cplx<T> v) {
        if (center) {
            v.re -= T(center[k].re);
            v.im -= T(center[k].im);
        }
        row[k] = v;
    };
    mlp_forward(net, w, Z, Act, lane);
    auto spin = [&](int i) {
        const uint64_t word = i < 64 ? w[0] : i < 128 ? w[1] : i < 192 ? w[2] : w[3];
        return ((word >> (i & 63)) & 1) ? T(1) : T(-1);
    };
    for (int i = lane; i < net.N; i += 64) put(i, {spin(i), T(0)});   // a: O = s_i
    const int last = net.L - 1;
    for (int j = lane; j < net.size[last]; j += 64)
        D[net.zoff[last] + j] = cmul(net.f[j], dact(Z[net.zoff[last] + j], last));
    __syncthreads();
    for (int l = last; l >= 0; l--) {
        const int M = net.size[l], C = net.conn[l];
        for (int j = lane; j < M; j += 64) {
            const cplx<T> dj = D[net.zoff[l] + j];
            put(net.poff[l] + j, dj);
            for (int i = 0; i < C; i++) {
                const int src = int(net.cn[l][i * M + j]);
                const cplx<T> in = l == 0 ? cplx<T>{spin(src), T(0)} : Act[net.zoff[l - 1] + src];
                put(net.poff[l] + M + i * M + j, cmul(dj, in));
            }
        }
        if (l > 0) {
            // D of layer l - 1: the weights reading each unit, gathered in a fixed order
            for (int i = lane; i < net.size[l - 1]; i += 64) {
                cplx<T> acc = {0, 0};
                for (int e = net.rb[l][i]; e < net.rb[l][i + 1]; e++) {
                    const int k = net.ri[l][e];
                    acc = cadd(acc, cmul(net.W[l][k], D[net.zoff[l] + k % M]));
                }
                D[net.zoff[l - 1] + i] = cmul(acc, dact(Z[net.zoff[l - 1] + i], l - 1));
            }
        }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(64) void mlp_ok_kernel(MlpView<T> net, const uint64_t* confs, int words, int64_t count,
                                                    const cplx<double>* center, cplx<T>* O) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    for (int64_t s = blockIdx.x; s < count; s += gridDim.x) {
        uint64_t w[kMaxWords];
        load_words(confs, s, words, w);
        mlp_ok_row(net, w, smem, O + s * net.num_params, center, lane);
    }
}

// the same for Pauli-string configurations read through 3 units per site
template <typename T>
__global__ __launch_bounds__(64) void mlp_ok_pauli_kernel(PauliInput<MlpView<T>> pn, const uint64_t* confs,
                                                          int64_t count, const cplx<double>* center, cplx<T>* O) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    for (int64_t s = blockIdx.x; s < count; s += gridDim.x) {
        uint64_t w[kMaxWords], e[kMaxWords];
        load_words(confs, s, 2, w);
        pauli_units(w, pn.N, e);
        mlp_ok_row(pn.net, e, smem, O + s * pn.num_params, center, lane);
    }
}

// O_k of PsiClassical: L_n(s) for n < num_ops, then the reference network's O_k
// O_k rows of the reference network (same LDS layout as its own O_k kernel)
template <typename T>
__device__ __forceinline__ void ref_ok_row(const MlpView<T>& net, const uint64_t* w, unsigned char* smem, cplx<T>* row,
                                           const cplx<double>* center, int lane) {
    mlp_ok_row(net, w, smem, row, center, lane);
}
template <typename T>
__device__ __forceinline__ void ref_ok_row(const CnnView<T>& net, const uint64_t* w, unsigned char* smem, cplx<T>* row,
                                           const cplx<double>* center, int lane) {
    cnn_ok_row(net, w, smem, row, center, lane);
}

template <typename T, typename R>
__global__ __launch_bounds__(64) void classical_ok_kernel(ClassicalView<T, R> v, const uint64_t* confs, int words,
                                                          int64_t count, const cplx<double>* center, cplx<T>* O) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    for (int64_t s = blockIdx.x; s < count; s += gridDim.x) {
        uint64_t w[kMaxWords];
        load_words(confs, s, words, w);
        cplx<T>* row = O + s * v.num_params;
        for (int n = lane; n < v.num_ops; n += 64) {
            cplx<double> L = classical_term(v, n, w);
            if (center) {
                L.re -= center[n].re;
                L.im -= center[n].im;
            }
            row[n] = {T(L.re), T(L.im)};
        }
        if (v.has_ref) {
            const cplx<double>* c2 = center ? center + v.num_ops : nullptr;
            if (v.basis == BASIS_PAULIS) {
                uint64_t e[kMaxWords];
                pauli_units(w, v.N, e);
                ref_ok_row(v.ref, e, smem, row + v.num_ops, c2, lane);
            } else {
                ref_ok_row(v.ref, w, smem, row + v.num_ops, c2, lane);
            }
        }
        __syncthreads();
    }
}

// kernels with more than 64 KiB of LDS per block must opt in (at most the 160 KiB of a gfx950 CU)
template <typename K>
void allow_lds(K kernel, size_t bytes) {
    if (bytes > 160 * 1024) throw std::invalid_argument("the network state of one configuration exceeds the 160 KiB LDS of a CU");
    if (bytes > 64 * 1024)
        ANN_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
}

inline unsigned grid_for(int64_t count) {
    const int64_t g = count < 65536 ? count : 65536;
    return unsigned(g > 0 ? g : 1);
}

// exp(L(s)) of Operator::fast_local_energy (Operator.hpp:123-135): one thread per
// configuration; the string tables are read as wave-uniform loads
__global__ __launch_bounds__(256) void exp_fast_energy_kernel(OpView op, const uint64_t* confs, int64_t count,
                                                              cplx<double>* out) {
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= count) return;
    const int words = op.words;
    uint64_t w[kMaxWords];
#pragma unroll
    for (int k = 0; k < kMaxWords; k++) w[k] = k < words ? confs[s * words + k] : 0;
    double re = 0.0, im = 0.0;
    const int n_end = op.group_begin[op.num_groups];
    for (int n = 0; n < n_end; n++) {
        int par = 0;
#pragma unroll
        for (int k = 0; k < kMaxWords; k++)
            if (k < words) par += __popcll(~w[k] & op.string_zy[int64_t(n) * words + k]);
        const double sg = (par & 1) ? -1.0 : 1.0;
        re += sg * op.string_coef[2 * n];
        im += sg * op.string_coef[2 * n + 1];
    }
    const double m = exp(re);
    double sn, cs;
    sincos(im, &sn, &cs);
    out[s] = {m * cs, m * sn};
}

__global__ __launch_bounds__(256) void exp_fast_energy_pauli_kernel(PauliOpView op, const uint64_t* confs,
                                                                    int64_t count, cplx<double>* out) {
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= count) return;
    const uint64_t xa = confs[2 * s], xb = confs[2 * s + 1];
    double re = 0.0, im = 0.0;
    for (int n = 0; n < op.num_strings; n++) {
        const cplx<double> f = pauli_phase(op.ab[2 * n], op.ab[2 * n + 1], xa, xb);
        const double cr = op.coef[2 * n], ci = op.coef[2 * n + 1];
        re += cr * f.re - ci * f.im;
        im += cr * f.im + ci * f.re;
    }
    const double m = exp(re);
    double sn, cs;
    sincos(im, &sn, &cs);
    out[s] = {m * cs, m * sn};
}

__global__ void reweight_kernel(const double* w, const cplx<double>* lp, const cplx<double>* lps, int64_t count,
                                double* out) {
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s < count) out[s] = w[s] * exp(2.0 * (lp[s].re - lps[s].re));
}

}  // namespace

void launch_exp_fast_energy(const OpView& op, const uint64_t* confs, int64_t count, cplx<double>* out,
                            hipStream_t stream) {
    if (count <= 0) return;
    exp_fast_energy_kernel<<<unsigned((count + 255) / 256), 256, 0, stream>>>(op, confs, count, out);
    ANN_LAUNCH_CHECK();
}

void launch_exp_fast_energy_pauli(const PauliOpView& op, const uint64_t* confs, int64_t count, cplx<double>* out,
                                  hipStream_t stream) {
    if (count <= 0) return;
    exp_fast_energy_pauli_kernel<<<unsigned((count + 255) / 256), 256, 0, stream>>>(op, confs, count, out);
    ANN_LAUNCH_CHECK();
}

void launch_reweight(const double* w, const cplx<double>* lp, const cplx<double>* lp_sampling, int64_t count,
                     double* out, hipStream_t stream) {
    if (count <= 0) return;
    reweight_kernel<<<unsigned((count + 255) / 256), 256, 0, stream>>>(w, lp, lp_sampling, count, out);
    ANN_LAUNCH_CHECK();
}

template <typename T>
size_t forward_lds(const CnnView<T>& net) {
    return sizeof(cplx<T>) * 2 * size_t(net.max_ch) * net.N;
}
template <typename T>
size_t forward_lds(const MlpView<T>& net) {
    return sizeof(cplx<T>) * 2 * size_t(net.zoff[net.L]);
}
template <typename T>
size_t ok_lds(const CnnView<T>& net) {
    return sizeof(cplx<T>) * (size_t(net.zoff[net.L]) + 3 * size_t(net.max_ch) * net.N);
}
template <typename T>
size_t ok_lds(const MlpView<T>& net) {
    return sizeof(cplx<T>) * 3 * size_t(net.zoff[net.L]);
}
template <int V>
struct VmaxTag {
    static constexpr int value = V;
};

// the local-update kernels with the neighbour indices of one site in registers when
// every layer's kernel volume is at most 9 (3 x 3) or 27 (3 x 3 x 3), else the plain loop
template <typename T, typename F>
void dispatch_vmax(const CnnView<T>& net, F&& fn) {
    int vmax = 0;
    for (int l = 0; l < net.L; l++) vmax = std::max(vmax, net.vol[l]);
    if (vmax <= 9)
        fn(VmaxTag<9>{});
    else if (vmax <= 27)
        fn(VmaxTag<27>{});
    else
        fn(VmaxTag<0>{});
}

// bytes of the LDS copy of a PsiCNN's tables ahead of the local-update state (one chain
// per workgroup), 0 when they would leave fewer than four chains per CU.  Above that the
// copy is kept even where leaving the tables in global memory would fit one chain more:
// on the 12x12, 2 x 8-channel net, 4 chains per CU with the tables in LDS sample faster
// than 5 without (each multiply-add reads a weight and a neighbour index)
template <typename T>
int cnn_stage_bytes(const CnnView<T>& net) {
    const size_t t = cnn_table_lds(net);
    return 4 * (t + cnn_inc_lds(net)) <= 160 * 1024 ? int(t) : 0;
}

template <typename V>
size_t forward_lds(const PauliInput<V>& net) {
    return forward_lds(net.net);
}
template <typename V>
size_t ok_lds(const PauliInput<V>& net) {
    return ok_lds(net.net);
}
template <typename T, typename R>
size_t forward_lds(const ClassicalView<T, R>& v) {
    return v.has_ref ? forward_lds(v.ref) : 0;
}
template <typename T, typename R>
size_t ok_lds(const ClassicalView<T, R>& v) {
    return v.has_ref ? ok_lds(v.ref) : 0;
}

template <typename V>
void launch_generic_log_psi(const V& net, const uint64_t* confs, int words, int64_t count,
                            cplx<double> log_prefactor, cplx<double>* log_psi, double* weight, hipStream_t stream) {
    if (count <= 0) return;
    allow_lds(generic_log_psi_kernel<V>, forward_lds(net));
    generic_log_psi_kernel<V><<<grid_for(count), 64, forward_lds(net), stream>>>(net, confs, words, count,
                                                                                 log_prefactor, log_psi, weight);
    ANN_LAUNCH_CHECK();
}

template <typename V>
void launch_generic_mc(const V& net, uint64_t* chain_confs, int words, int num_chains, int move,
                       int64_t num_proposals, uint64_t counter, uint64_t seed, uint32_t* accepted,
                       hipStream_t stream) {
    if (num_chains <= 0 || num_proposals <= 0) return;
    using T = typename V::real_t;
    if constexpr (std::is_same<V, CnnView<T>>::value) {
        const int stage = cnn_stage_bytes(net);
        dispatch_vmax(net, [&](auto vt) {
            constexpr int VM = decltype(vt)::value;
            auto run = [&](auto kern) {
                allow_lds(kern, cnn_inc_lds(net) + stage);
                kern<<<unsigned(num_chains), 64, cnn_inc_lds(net) + stage, stream>>>(
                    net, chain_confs, words, num_chains, move, num_proposals, counter, uint32_t(seed),
                    uint32_t(seed >> 32), accepted, stage);
            };
            if (stage)
                run(cnn_mc_inc_kernel<T, VM, true>);
            else
                run(cnn_mc_inc_kernel<T, VM, false>);
        });
    } else {
        allow_lds(generic_mc_kernel<V>, forward_lds(net));
        generic_mc_kernel<V><<<unsigned(num_chains), 64, forward_lds(net), stream>>>(
            net, chain_confs, words, num_chains, move, num_proposals, counter, uint32_t(seed), uint32_t(seed >> 32),
            accepted);
    }
    ANN_LAUNCH_CHECK();
}

template <typename V>
void launch_generic_local_energy(const V& net, const OpView& op, const uint64_t* confs, int64_t count,
                                 cplx<double>* e_loc, hipStream_t stream) {
    if (count <= 0) return;
    using T = typename V::real_t;
    if constexpr (std::is_same<V, CnnView<T>>::value) {
        const int stage = cnn_stage_bytes(net);
        dispatch_vmax(net, [&](auto vt) {
            constexpr int VM = decltype(vt)::value;
            auto run = [&](auto kern) {
                allow_lds(kern, cnn_inc_lds(net) + stage);
                kern<<<grid_for(count), 64, cnn_inc_lds(net) + stage, stream>>>(net, op, confs, count, e_loc, stage);
            };
            if (stage)
                run(cnn_local_energy_inc_kernel<T, VM, true>);
            else
                run(cnn_local_energy_inc_kernel<T, VM, false>);
        });
    } else {
        allow_lds(generic_local_energy_kernel<V>, forward_lds(net));
        generic_local_energy_kernel<V><<<grid_for(count), 64, forward_lds(net), stream>>>(net, op, confs, count,
                                                                                         e_loc);
    }
    ANN_LAUNCH_CHECK();
}

template <typename V>
void launch_generic_ok(const V& net, const uint64_t* confs, int words, int64_t count, const cplx<double>* center,
                       cplx<typename V::real_t>* O, hipStream_t stream) {
    if (count <= 0) return;
    using T = typename V::real_t;
    if constexpr (std::is_same<V, CnnView<T>>::value) {
        allow_lds(cnn_ok_kernel<T>, ok_lds(net));
        cnn_ok_kernel<T><<<grid_for(count), 64, ok_lds(net), stream>>>(net, confs, words, count, center, O);
    } else if constexpr (std::is_same<V, MlpView<T>>::value) {
        allow_lds(mlp_ok_kernel<T>, ok_lds(net));
        mlp_ok_kernel<T><<<grid_for(count), 64, ok_lds(net), stream>>>(net, confs, words, count, center, O);
    } else if constexpr (std::is_same<V, PauliInput<MlpView<T>>>::value) {
        allow_lds(mlp_ok_pauli_kernel<T>, ok_lds(net));
        mlp_ok_pauli_kernel<T><<<grid_for(count), 64, ok_lds(net), stream>>>(net, confs, count, center, O);
    } else {
        using R = decltype(net.ref);
        allow_lds(classical_ok_kernel<T, R>, ok_lds(net));
        classical_ok_kernel<T, R><<<grid_for(count), 64, ok_lds(net), stream>>>(net, confs, words, count, center, O);
    }
    ANN_LAUNCH_CHECK();
}

template <typename V>
void launch_generic_local_energy_pauli(const V& net, const PauliOpView& op, const uint64_t* confs, int64_t count,
                                       cplx<double>* e_loc, hipStream_t stream) {
    if (count <= 0) return;
    allow_lds(generic_local_energy_pauli_kernel<V>, forward_lds(net));
    generic_local_energy_pauli_kernel<V><<<grid_for(count), 64, forward_lds(net), stream>>>(net, op, confs, count,
                                                                                           e_loc);
    ANN_LAUNCH_CHECK();
}

#define ANN_GENERIC_INSTANTIATE(V)                                                                                  \
    template void launch_generic_log_psi<V>(const V&, const uint64_t*, int, int64_t, cplx<double>, cplx<double>*, \
                                            double*, hipStream_t);                                                 \
    template void launch_generic_mc<V>(const V&, uint64_t*, int, int, int, int64_t, uint64_t, uint64_t,           \
                                       uint32_t*, hipStream_t);                                                    \
    template void launch_generic_local_energy<V>(const V&, const OpView&, const uint64_t*, int64_t,               \
                                                 cplx<double>*, hipStream_t);                                      \
    template void launch_generic_ok<V>(const V&, const uint64_t*, int, int64_t, const cplx<double>*,              \
                                       cplx<typename V::real_t>*, hipStream_t);

ANN_GENERIC_INSTANTIATE(CnnView<float>)
ANN_GENERIC_INSTANTIATE(CnnView<double>)
ANN_GENERIC_INSTANTIATE(MlpView<float>)
ANN_GENERIC_INSTANTIATE(MlpView<double>)
ANN_GENERIC_INSTANTIATE(PauliInput<MlpView<float>>)
ANN_GENERIC_INSTANTIATE(PauliInput<MlpView<double>>)
ANN_GENERIC_INSTANTIATE(ClassicalView<float>)
ANN_GENERIC_INSTANTIATE(ClassicalView<double>)
using ClassicalCnn32 = ClassicalView<float, CnnView<float>>;
using ClassicalCnn64 = ClassicalView<double, CnnView<double>>;
ANN_GENERIC_INSTANTIATE(ClassicalCnn32)
ANN_GENERIC_INSTANTIATE(ClassicalCnn64)

#define ANN_PAULI_INSTANTIATE(V)                                                                          \
    template void launch_generic_local_energy_pauli<V>(const V&, const PauliOpView&, const uint64_t*, int64_t, \
                                                       cplx<double>*, hipStream_t);

ANN_PAULI_INSTANTIATE(PauliInput<MlpView<float>>)
ANN_PAULI_INSTANTIATE(PauliInput<MlpView<double>>)
ANN_PAULI_INSTANTIATE(ClassicalView<float>)
ANN_PAULI_INSTANTIATE(ClassicalView<double>)

}  // namespace ann
