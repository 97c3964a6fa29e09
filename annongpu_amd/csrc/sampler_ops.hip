// sampler_ops.hip -- chain initialisation, SR vector plumbing and small fp64
// vector kernels used by the host runtime.
#include "ops.hpp"

namespace ann {

namespace {

// Random initial configurations on the Philox stream (Init_Policy.hpp:15-25 draws
// uniform bits; exchange sampling starts from a uniform configuration of the
// requested S_z sector).
__global__ void init_chains_kernel(uint64_t* confs, int num_chains, int words, int N, int move, int total_z,
                                   uint32_t call_index, uint32_t k0, uint32_t k1) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= num_chains) return;
    uint64_t w[kMaxWords] = {0, 0, 0, 0};
    if (move == MOVE_PAULI) {
        // a uniformly random Pauli string on the N sites (Init_Policy.hpp:33-39, PauliString.hpp:58-65)
        const u32x4 r = philox(uint32_t(c), 0u, call_index, 3u, k0, k1);
        const uint64_t m = N >= 64 ? ~uint64_t(0) : (uint64_t(1) << N) - 1;
        w[0] = (uint64_t(r.x) | (uint64_t(r.y) << 32)) & m;
        w[1] = (uint64_t(r.z) | (uint64_t(r.w) << 32)) & m;
    } else if (move == MOVE_FLIP) {
        for (int blk = 0; blk < (words + 1) / 2; blk++) {
            const u32x4 r = philox(uint32_t(c), uint32_t(blk), call_index, 1u, k0, k1);
            w[2 * blk] = uint64_t(r.x) | (uint64_t(r.y) << 32);
            if (2 * blk + 1 < words) w[2 * blk + 1] = uint64_t(r.z) | (uint64_t(r.w) << 32);
        }
        if (N % 64) w[N / 64] &= (uint64_t(1) << (N % 64)) - 1;
        for (int k = (N + 63) / 64; k < kMaxWords; k++) w[k] = 0;
    } else {
        const int n_up = (N + total_z) / 2;
        uint64_t down[kMaxWords] = {0, 0, 0, 0};
        for (int i = 0; i < N; i++) down[i >> 6] |= uint64_t(1) << (i & 63);
        u32x4 r = {0, 0, 0, 0};
        for (int k = 0; k < n_up; k++) {
            if ((k & 3) == 0) r = philox(uint32_t(c), uint32_t(k >> 2), call_index, 2u, k0, k1);
            const uint32_t x = (k & 3) == 0 ? r.x : (k & 3) == 1 ? r.y : (k & 3) == 2 ? r.z : r.w;
            const int pos = select_bit(down, words, range32(x, uint32_t(N - k)));
            down[pos >> 6] &= ~(uint64_t(1) << (pos & 63));
            w[pos >> 6] |= uint64_t(1) << (pos & 63);
        }
    }
    for (int k = 0; k < words; k++) confs[int64_t(c) * words + k] = w[k];
}

// v (reference parameter order) -> va [N], vb [Mpad], V [N][Mpad] (W part re-indexed
// onto visible units: V[conn(i,j)][j] += v_W(i,j)); one thread per V element
template <typename T>
__global__ void unpack_v_kernel(DenseView<T> net, const int* inv, const cplx<double>* v, cplx<double>* va,
                                cplx<double>* vb, cplx<T>* V) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t nV = int64_t(net.N) * net.Mpad;
    if (e < nV) {
        const int vis = int(e / net.Mpad), j = int(e % net.Mpad);
        double re = 0.0, im = 0.0;
        if (j < net.M) {
            if (net.layout == LAYOUT_RBM) {
                const cplx<double> x = v[int64_t(vis) * net.M + j];
                re = x.re;
                im = x.im;
            } else {
                // inv[vis][j]: the unique i with conn(i, j) = vis, -1 if none, -2 if several
                const int i0 = inv[e];
                if (i0 >= 0) {
                    const cplx<double> x = v[net.N + net.M + int64_t(i0) * net.M + j];
                    re = x.re;
                    im = x.im;
                } else if (i0 == -2) {
                    for (int i = 0; i < net.C; i++) {
                        if (int(net.conn[int64_t(i) * net.M + j]) == vis) {
                            const cplx<double> x = v[net.N + net.M + int64_t(i) * net.M + j];
                            re += x.re;
                            im += x.im;
                        }
                    }
                }
            }
        }
    V[e] = {T(re), T(im)};
        return;
    }
    const int64_t e2 = e - nV;
    if (e2 < net.Mpad) {
        const int j = int(e2);
        if (vb) vb[j] = (net.layout == LAYOUT_DEEP1 && j < net.M) ? v[net.N + j] : cplx<double>{0, 0};
    } else if (e2 < int64_t(net.Mpad) + net.N) {
        const int i = int(e2 - net.Mpad);
        if (va) va[i] = net.layout == LAYOUT_DEEP1 ? v[i] : cplx<double>{0, 0};
    }
}

// y (reference order) from Ya, Yb, Yw (+ shift * v)
__global__ void pack_y_kernel(int layout, int N, int M, int Mpad, int C, const uint32_t* conn,
                              const cplx<double>* Ya, const cplx<double>* Yb, const cplx<double>* Yw,
                              const cplx<double>* v, double shift, int P, cplx<double>* y) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P) return;
    cplx<double> r;
    if (layout == LAYOUT_RBM) {
        const int i = k / M, j = k % M;
        r = Yw[int64_t(i) * Mpad + j];
    } else if (k < N) {
        r = Ya[k];
    } else if (k < N + M) {
        r = Yb[k - N];
    } else {
        const int q = k - N - M;
        const int j = q % M;
        r = Yw[int64_t(conn[q]) * Mpad + j];
    }
    if (v) {
        r.re += shift * v[k].re;
        r.im += shift * v[k].im;
    }
    y[k] = r;
}

// bf16x3 fragments of a [N][Mpad] complex fp32 matrix (layout of theta_mfma_kernel)
__device__ inline uint16_t bf16_rn(float x) {
    const uint32_t b = __float_as_uint(x);
    return uint16_t((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}
__global__ void pack_fragments_kernel(const cplx<float>* V, int N, int Mpad, int KT, int NT, uint16_t* out) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int ncol = 2 * Mpad;
    if (e >= int64_t(KT) * 16 * ncol) return;
    const int v = int(e / ncol), col = int(e % ncol);
    float r = 0.f;
    if (v < N) {
        const cplx<float> x = V[int64_t(v) * Mpad + col / 2];
        r = (col & 1) ? x.im : x.re;
    }
    const int kt = v / 16, kk = v % 16, nt = col / 16, cc = col % 16;
    const int lane = (kk / 4) * 16 + cc, el = kk % 4;
    for (int sp = 0; sp < 3; sp++) {
        const uint16_t h = bf16_rn(r);
        r -= __uint_as_float(uint32_t(h) << 16);
        out[((((int64_t(sp) * KT + kt) * NT + nt) * 64 + lane) * 4) + el] = h;
    }
}

constexpr int kDot = 1024;
__global__ __launch_bounds__(kDot) void cdot_kernel(const cplx<double>* a, const cplx<double>* b, int64_t n,
                                                    int conj_a, double* out) {
    __shared__ double sr[kDot], si[kDot];
    double re = 0, im = 0;
    for (int64_t i = threadIdx.x; i < n; i += kDot) {
        const double ar = a[i].re, ai = conj_a ? -a[i].im : a[i].im;
        re += ar * b[i].re - ai * b[i].im;
        im += ar * b[i].im + ai * b[i].re;
    }
    sr[threadIdx.x] = re;
    si[threadIdx.x] = im;
    __syncthreads();
    for (int o = kDot / 2; o > 0; o >>= 1) {
        if (int(threadIdx.x) < o) {
            sr[threadIdx.x] += sr[threadIdx.x + o];
            si[threadIdx.x] += si[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sr[0];
        out[1] = si[0];
    }
}

// fused conjugated dot products: out[0..1] = <a, b>, out[2..3] = <c, d> (if c != nullptr)
constexpr int kDotBlocks = 64;
__global__ __launch_bounds__(256) void dot2_partial_kernel(const cplx<double>* a, const cplx<double>* b,
                                                           const cplx<double>* c, const cplx<double>* d,
                                                           int64_t n, int conj_a, double* partial) {
    __shared__ double sh[4][256];
    double acc[4] = {0, 0, 0, 0};
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const double ar = a[i].re, ai = conj_a ? -a[i].im : a[i].im;
        acc[0] += ar * b[i].re - ai * b[i].im;
        acc[1] += ar * b[i].im + ai * b[i].re;
        if (c) {
            const double cr = c[i].re, ci = conj_a ? -c[i].im : c[i].im;
            acc[2] += cr * d[i].re - ci * d[i].im;
            acc[3] += cr * d[i].im + ci * d[i].re;
        }
    }
    for (int k = 0; k < 4; k++) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (int(threadIdx.x) < o)
            for (int k = 0; k < 4; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 4) partial[blockIdx.x * 4 + threadIdx.x] = sh[threadIdx.x][0];
}
__global__ void dot2_final_kernel(const double* partial, int blocks, double* out) {
    const int k = threadIdx.x;
    if (k >= 4) return;
    double acc = 0.0;
    for (int b = 0; b < blocks; b++) acc += partial[b * 4 + k];
    out[k] = acc;
}

// CG: alpha = rz / pAp (0 on breakdown); x += alpha p; r -= alpha Ap
__global__ void cg_update_dev_kernel(cplx<double>* x, cplx<double>* r, const cplx<double>* p,
                                     const cplx<double>* Ap, const double* rz, const double* pAp, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double alpha = pAp[0] > 0.0 ? rz[0] / pAp[0] : 0.0;
    x[i].re += alpha * p[i].re;
    x[i].im += alpha * p[i].im;
    r[i].re -= alpha * Ap[i].re;
    r[i].im -= alpha * Ap[i].im;
}
// p = z + (rz_new / rz_old) p
__global__ void cg_dir_dev_kernel(cplx<double>* p, const cplx<double>* z, const double* rz_new,
                                  const double* rz_old, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double beta = rz_old[0] != 0.0 ? rz_new[0] / rz_old[0] : 0.0;
    p[i] = {z[i].re + beta * p[i].re, z[i].im + beta * p[i].im};
}

// u'_s = w_s * (u_s - c)   with c = (re, im) read from device memory
__global__ void center_scale_kernel(cplx<double>* u, const double* w, const double* c, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u[i] = {w[i] * (u[i].re - c[0]), w[i] * (u[i].im - c[1])};
}

// x += alpha p ; r -= alpha Ap   (alpha real)
__global__ void cg_step_kernel(cplx<double>* x, cplx<double>* r, const cplx<double>* p, const cplx<double>* Ap,
                               double alpha, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i].re += alpha * p[i].re;
    x[i].im += alpha * p[i].im;
    r[i].re -= alpha * Ap[i].re;
    r[i].im -= alpha * Ap[i].im;
}
// p = r + beta p
__global__ void cg_dir_kernel(cplx<double>* p, const cplx<double>* r, double beta, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    p[i] = {r[i].re + beta * p[i].re, r[i].im + beta * p[i].im};
}

// Jacobi preconditioner: z = r / (diag + shift)
__global__ void precond_kernel(cplx<double>* z, const cplx<double>* r, const double* diag, double shift, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = diag[i] + shift;
    z[i] = {r[i].re / d, r[i].im / d};
}

// diag of S for a dense net: a part wsum - |<O_a>|^2, b and W parts D_j - |<O_k>|^2
// (wsum read from the device)
__global__ void sr_diag_kernel(int layout, int N, int M, int Mpad, const double* wsum_p, const cplx<double>* Dj,
                               const cplx<double>* Obar, int P, double* diag) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P) return;
    const double wsum = *wsum_p;
    double d;
    if (layout == LAYOUT_RBM) {
        d = Dj[k % M].re;
    } else if (k < N) {
        d = wsum;
    } else if (k < N + M) {
        d = Dj[k - N].re;
    } else {
        d = Dj[(k - N - M) % M].re;
    }
    const cplx<double> o = Obar[k];
    d -= o.re * o.re + o.im * o.im;
    diag[k] = d > 0.0 ? d : 0.0;
    (void)Mpad;
}

// S += alpha * conj(o) o^T
__global__ void rank1_kernel(cplx<double>* S, const cplx<double>* o, int P, double alpha) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= int64_t(P) * P) return;
    const int k = int(e / P), kp = int(e % P);
    const cplx<double> a = o[k], b = o[kp];
    S[e].re += alpha * (a.re * b.re + a.im * b.im);
    S[e].im += alpha * (a.re * b.im - a.im * b.re);
}

__global__ void add_diag_kernel(cplx<double>* S, int P, double shift) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P) S[int64_t(i) * P + i].re += shift;
}

// symmetric diagonal scaling of a Hermitian matrix with a positive diagonal:
// dinv[i] = 1 / sqrt(A_ii) (1 where A_ii is not positive), A <- D^-1/2 A D^-1/2
__global__ void diag_rsqrt_kernel(const cplx<double>* A, int P, double* dinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const double d = A[int64_t(i) * P + i].re;
    dinv[i] = d > 0.0 ? 1.0 / sqrt(d) : 1.0;
}

__global__ void scale_sym_kernel(cplx<double>* A, int P, const double* dinv) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= int64_t(P) * P) return;
    const int i = int(e / P), j = int(e % P);
    const double f = dinv[i] * dinv[j];
    A[e].re *= f;
    A[e].im *= f;
}

__global__ void scale_vec_kernel(cplx<double>* x, int P, const double* dinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    x[i].re *= dinv[i];
    x[i].im *= dinv[i];
}

template <typename Tin, typename Tout>
__global__ void convert_kernel(const cplx<Tin>* a, cplx<Tout>* b, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = {Tout(a[i].re), Tout(a[i].im)};
}

__global__ void fill_kernel(double* a, double v, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// coef_s = w_s (e_s - E) with E = (sums[0], sums[1]) read from the device
__global__ void weight_coef_centered_dev_kernel(const double* w, const cplx<double>* e, const double* sums, int64_t n,
                                                cplx<double>* coef) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double er = sums[0], ei = sums[1];
    coef[i] = {w[i] * (e[i].re - er), w[i] * (e[i].im - ei)};
}

// coef_s = w_s * e_s (complex), or w_s when e == nullptr
__global__ void weight_coef_kernel(const double* w, const cplx<double>* e, cplx<double> shift, int64_t n,
                                   cplx<double>* coef) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    coef[i] = e ? cplx<double>{w[i] * (e[i].re - shift.re), w[i] * (e[i].im - shift.im)} : cplx<double>{w[i], 0.0};
}

inline unsigned blocks_for(int64_t n, int b = 256) { return unsigned((n + b - 1) / b); }

}  // namespace

void launch_init_chains(uint64_t* confs, int num_chains, int words, int N, int move, int total_z,
                        uint32_t call_index, uint64_t seed, hipStream_t stream) {
    init_chains_kernel<<<blocks_for(num_chains, 64), 64, 0, stream>>>(confs, num_chains, words, N, move, total_z,
                                                                      call_index, uint32_t(seed),
                                                                      uint32_t(seed >> 32));
    ANN_LAUNCH_CHECK();
}

template <typename T>
void launch_unpack_v(const DenseView<T>& net, const int* inv, const cplx<double>* v, cplx<double>* va,
                     cplx<double>* vb, cplx<T>* V, hipStream_t stream) {
    const int64_t n = int64_t(net.N) * net.Mpad + net.Mpad + net.N;
    unpack_v_kernel<T><<<blocks_for(n), 256, 0, stream>>>(net, inv, v, va, vb, V);
    ANN_LAUNCH_CHECK();
}
template void launch_unpack_v<float>(const DenseView<float>&, const int*, const cplx<double>*, cplx<double>*,
                                     cplx<double>*, cplx<float>*, hipStream_t);
template void launch_unpack_v<double>(const DenseView<double>&, const int*, const cplx<double>*, cplx<double>*,
                                      cplx<double>*, cplx<double>*, hipStream_t);

void launch_pack_y(int layout, int N, int M, int Mpad, int C, const uint32_t* conn, const cplx<double>* Ya,
                   const cplx<double>* Yb, const cplx<double>* Yw, const cplx<double>* v, double shift, int P,
                   cplx<double>* y, hipStream_t stream) {
    pack_y_kernel<<<blocks_for(P), 256, 0, stream>>>(layout, N, M, Mpad, C, conn, Ya, Yb, Yw, v, shift, P, y);
    ANN_LAUNCH_CHECK();
}

void launch_pack_fragments(const cplx<float>* V, int N, int Mpad, int npad, uint16_t* out, hipStream_t stream) {
    const int KT = npad / 16, NT = 2 * Mpad / 16;
    const int64_t n = int64_t(npad) * 2 * Mpad;
    pack_fragments_kernel<<<blocks_for(n), 256, 0, stream>>>(V, N, Mpad, KT, NT, out);
    ANN_LAUNCH_CHECK();
}

void launch_cdot(const cplx<double>* a, const cplx<double>* b, int64_t n, int conj_a, double* out2,
                 hipStream_t stream) {
    cdot_kernel<<<1, kDot, 0, stream>>>(a, b, n, conj_a, out2);
    ANN_LAUNCH_CHECK();
}

void launch_dot2(const cplx<double>* a, const cplx<double>* b, const cplx<double>* c, const cplx<double>* d,
                 int64_t n, int conj_a, double* scratch, double* out4, hipStream_t stream) {
    dot2_partial_kernel<<<kDotBlocks, 256, 0, stream>>>(a, b, c, d, n, conj_a, scratch);
    ANN_LAUNCH_CHECK();
    dot2_final_kernel<<<1, 64, 0, stream>>>(scratch, kDotBlocks, out4);
    ANN_LAUNCH_CHECK();
}

void launch_cg_update_dev(cplx<double>* x, cplx<double>* r, const cplx<double>* p, const cplx<double>* Ap,
                          const double* rz, const double* pAp, int64_t n, hipStream_t stream) {
    cg_update_dev_kernel<<<blocks_for(n), 256, 0, stream>>>(x, r, p, Ap, rz, pAp, n);
    ANN_LAUNCH_CHECK();
}

void launch_cg_dir_dev(cplx<double>* p, const cplx<double>* z, const double* rz_new, const double* rz_old, int64_t n,
                       hipStream_t stream) {
    cg_dir_dev_kernel<<<blocks_for(n), 256, 0, stream>>>(p, z, rz_new, rz_old, n);
    ANN_LAUNCH_CHECK();
}

void launch_center_scale(cplx<double>* u, const double* w, const double* c2, int64_t n, hipStream_t stream) {
    center_scale_kernel<<<blocks_for(n), 256, 0, stream>>>(u, w, c2, n);
    ANN_LAUNCH_CHECK();
}

void launch_cg_step(cplx<double>* x, cplx<double>* r, const cplx<double>* p, const cplx<double>* Ap, double alpha,
                    int64_t n, hipStream_t stream) {
    cg_step_kernel<<<blocks_for(n), 256, 0, stream>>>(x, r, p, Ap, alpha, n);
    ANN_LAUNCH_CHECK();
}

void launch_cg_dir(cplx<double>* p, const cplx<double>* r, double beta, int64_t n, hipStream_t stream) {
    cg_dir_kernel<<<blocks_for(n), 256, 0, stream>>>(p, r, beta, n);
    ANN_LAUNCH_CHECK();
}

void launch_precond(cplx<double>* z, const cplx<double>* r, const double* diag, double shift, int64_t n,
                    hipStream_t stream) {
    precond_kernel<<<blocks_for(n), 256, 0, stream>>>(z, r, diag, shift, n);
    ANN_LAUNCH_CHECK();
}

void launch_sr_diag(int layout, int N, int M, int Mpad, const double* wsum, const cplx<double>* Dj,
                    const cplx<double>* Obar, int P, double* diag, hipStream_t stream) {
    sr_diag_kernel<<<blocks_for(P), 256, 0, stream>>>(layout, N, M, Mpad, wsum, Dj, Obar, P, diag);
    ANN_LAUNCH_CHECK();
}

void launch_rank1(cplx<double>* S, const cplx<double>* o, int P, double alpha, hipStream_t stream) {
    rank1_kernel<<<blocks_for(int64_t(P) * P), 256, 0, stream>>>(S, o, P, alpha);
    ANN_LAUNCH_CHECK();
}

void launch_add_diag(cplx<double>* S, int P, double shift, hipStream_t stream) {
    add_diag_kernel<<<blocks_for(P), 256, 0, stream>>>(S, P, shift);
    ANN_LAUNCH_CHECK();
}

void launch_equilibrate(cplx<double>* A, int P, double* dinv, hipStream_t stream) {
    diag_rsqrt_kernel<<<blocks_for(P), 256, 0, stream>>>(A, P, dinv);
    ANN_LAUNCH_CHECK();
    scale_sym_kernel<<<blocks_for(int64_t(P) * P), 256, 0, stream>>>(A, P, dinv);
    ANN_LAUNCH_CHECK();
}

void launch_scale_vec(cplx<double>* x, int P, const double* dinv, hipStream_t stream) {
    scale_vec_kernel<<<blocks_for(P), 256, 0, stream>>>(x, P, dinv);
    ANN_LAUNCH_CHECK();
}

template <typename Tin, typename Tout>
void launch_convert(const cplx<Tin>* a, cplx<Tout>* b, int64_t n, hipStream_t stream) {
    if (n <= 0) return;
    convert_kernel<Tin, Tout><<<blocks_for(n), 256, 0, stream>>>(a, b, n);
    ANN_LAUNCH_CHECK();
}
template void launch_convert<float, double>(const cplx<float>*, cplx<double>*, int64_t, hipStream_t);
template void launch_convert<double, double>(const cplx<double>*, cplx<double>*, int64_t, hipStream_t);
template void launch_convert<double, float>(const cplx<double>*, cplx<float>*, int64_t, hipStream_t);

void launch_fill(double* a, double v, int64_t n, hipStream_t stream) {
    if (n <= 0) return;
    fill_kernel<<<blocks_for(n), 256, 0, stream>>>(a, v, n);
    ANN_LAUNCH_CHECK();
}

void launch_weight_coef(const double* w, const cplx<double>* e, int64_t n, cplx<double>* coef, hipStream_t stream) {
    if (n <= 0) return;
    weight_coef_kernel<<<blocks_for(n), 256, 0, stream>>>(w, e, cplx<double>{0.0, 0.0}, n, coef);
    ANN_LAUNCH_CHECK();
}

void launch_weight_coef_centered_dev(const double* w, const cplx<double>* e, const double* sums, int64_t n,
                                     cplx<double>* coef, hipStream_t stream) {
    if (n <= 0) return;
    weight_coef_centered_dev_kernel<<<blocks_for(n), 256, 0, stream>>>(w, e, sums, n, coef);
    ANN_LAUNCH_CHECK();
}

void launch_weight_coef_centered(const double* w, const cplx<double>* e, cplx<double> E, int64_t n,
                                 cplx<double>* coef, hipStream_t stream) {
    if (n <= 0) return;
    weight_coef_kernel<<<blocks_for(n), 256, 0, stream>>>(w, e, E, n, coef);
    ANN_LAUNCH_CHECK();
}

}  // namespace ann
