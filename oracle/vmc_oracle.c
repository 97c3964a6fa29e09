/*
 * vmc_oracle.c -- fp64 CPU restatement of ANNonGPU's variational Monte Carlo path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product and the timed CPU
 * baseline.  Nothing in annongpu_amd/ links or loads this file.
 *
 * It follows the reference's host (one chain per thread) semantics:
 *   - configurations: include/basis/Spins.h:65-118 (bit i set <=> s_i = +1)
 *   - Pauli strings:   include/basis/PauliString.hpp:133-143,193-208,242-254
 *   - activations:     include/quantum_state/psi_functions.hpp:40-48,102-116
 *   - PsiRBM:          include/quantum_state/PsiRBM.hpp:69-176
 *   - PsiDeep:         include/quantum_state/PsiDeep.hpp:138-445,
 *                      source/quantum_state/PsiDeep.cu:142-183 (parameter order)
 *   - PsiCNN:          include/quantum_state/PsiCNN.hpp:97-266,
 *                      include/quantum_state/detail/Convolve.hpp:112-173
 *   - local energy:    include/operator/Operator.hpp:36-121
 *   - Monte Carlo:     include/ensembles/MonteCarlo.hpp:55-177 (Metropolis rule)
 *   - TDVP statistics: source/network_functions/TDVP.cu.template:77-126,181-302
 * The random stream is Philox4x32-10, keyed per ensemble and countered per chain
 * and proposal, so the GPU sampler can replay the same trajectories.
 */
#include "vmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cd;

/* ------------------------------------------------------------------ RNG */

static inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k[0];
    const uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
        philox_round(c, k);
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

static inline void draw(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                        uint32_t out[4]) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t ctr[4] = {c0, c1, c2, c3};
    orc_philox4x32_10(ctr, key, out);
}

/* ---------------------------------------------------------- activations */

cd orc_logcosh(cd z, uint32_t layer) {
    const cd z2 = z * z;
    const cd z4 = z2 * z2;
    if (layer == 0u) return 0.5 * z2 - (1.0 / 12.0) * z4 + (1.0 / 45.0) * z4 * z2;
    return z - (1.0 / 3.0) * z2 * z + (2.0 / 15.0) * z4 * z;
}

cd orc_tanh(cd z, uint32_t layer) {
    const cd z2 = z * z;
    const cd z4 = z2 * z2;
    if (layer == 0u) return z - (1.0 / 3.0) * z2 * z + (2.0 / 15.0) * z4 * z;
    return 1.0 - z2 + (2.0 / 3.0) * z4;
}

/* --------------------------------------------------------------- basis */

static inline double spin_at(const uint64_t* conf, int i) {
    return ((conf[i >> 6] >> (i & 63)) & 1u) ? 1.0 : -1.0;
}

void orc_pauli_apply(int32_t words, const uint64_t* a, const uint64_t* b,
                     const uint64_t* conf, uint64_t* conf_out, cd* factor) {
    int num_y = 0, num_neg = 0;
    for (int w = 0; w < words; w++) {
        num_y += __builtin_popcountll(~a[w] & b[w]);          /* sigma_y = (0,1) */
        num_neg += __builtin_popcountll(~conf[w] & b[w]);     /* Z or Y on a down spin */
        conf_out[w] = conf[w] ^ (a[w] ^ b[w]);                /* X or Y flips */
    }
    cd f = 1.0;
    if ((num_y & 3) > 1) f = -f;
    if (num_y & 1) f *= -I;
    if (num_neg & 1) f = -f;
    *factor = f;
}

/* ------------------------------------------------------- wave function */

/* dense first-layer matrix: theta_j = b_j + sum_v Wd[v*size+j] * s_v */
static cd* first_layer_dense(const orc_psi* psi, int* width) {
    if (psi->kind == ORC_RBM) {
        *width = psi->M;
        cd* Wd = (cd*)malloc(sizeof(cd) * (size_t)psi->N * psi->M);
        memcpy(Wd, psi->W, sizeof(cd) * (size_t)psi->N * psi->M);
        return Wd;
    }
    const int size = psi->size[0], conn = psi->conn[0];
    *width = size;
    cd* Wd = (cd*)calloc((size_t)psi->N * size, sizeof(cd));
    for (int i = 0; i < conn; i++)
        for (int j = 0; j < size; j++)
            Wd[(size_t)psi->connections[0][i * size + j] * size + j] += psi->weights[0][i * size + j];
    return Wd;
}

static void compute_theta(const orc_psi* psi, const uint64_t* conf, cd* theta) {
    if (psi->kind == ORC_RBM) {
        for (int j = 0; j < psi->M; j++) {
            cd t = 0.0;
            for (int i = 0; i < psi->N; i++) t += psi->W[i * psi->M + j] * spin_at(conf, i);
            theta[j] = t;
        }
        return;
    }
    const int size = psi->size[0], conn = psi->conn[0];
    for (int j = 0; j < size; j++) {
        cd t = 0.0;
        for (int i = 0; i < conn; i++)
            t += psi->weights[0][i * size + j] * spin_at(conf, (int)psi->connections[0][i * size + j]);
        theta[j] = t + psi->biases[0][j];
    }
}

/* forward pass of PsiDeep given first-layer angles; optionally records the
 * pre-activations of every layer (PsiDeep.hpp:171-215). */
static cd deep_forward(const orc_psi* psi, const cd* theta, cd** z_rec) {
    int prev = psi->size[0];
    cd* act = (cd*)malloc(sizeof(cd) * 4096);
    cd* nxt = (cd*)malloc(sizeof(cd) * 4096);
    for (int j = 0; j < prev; j++) {
        act[j] = orc_logcosh(theta[j], 0);
        if (z_rec) z_rec[0][j] = theta[j];
    }
    for (int l = 1; l < psi->num_layers; l++) {
        const int size = psi->size[l], conn = psi->conn[l];
        for (int j = 0; j < size; j++) {
            cd z = 0.0;
            for (int i = 0; i < conn; i++)
                z += psi->weights[l][i * size + j] * act[psi->connections[l][i * size + j]];
            z += psi->biases[l][j];
            if (z_rec) z_rec[l][j] = z;
            nxt[j] = orc_logcosh(z, (uint32_t)l);
        }
        memcpy(act, nxt, sizeof(cd) * size);
        prev = size;
    }
    cd r = 0.0;
    for (int j = 0; j < prev; j++) r += act[j] * psi->final_weights[j];
    free(act); free(nxt);
    return r;
}

static cd log_psi_from_theta(const orc_psi* psi, const cd* theta) {
    if (psi->kind == ORC_RBM) {
        cd r = 0.0;
        for (int j = 0; j < psi->M; j++) r += orc_logcosh(theta[j], 0) * psi->final_weight;
        return psi->log_prefactor + r;
    }
    if (psi->num_layers == 1) {
        cd r = 0.0;
        for (int j = 0; j < psi->size[0]; j++) r += orc_logcosh(theta[j], 0) * psi->final_weights[j];
        return psi->log_prefactor + r;
    }
    return psi->log_prefactor + deep_forward(psi, theta, NULL);
}

/* PsiCNN (dim 3) neighbour of site idx at connection (k,i,j) -- Convolve.hpp:122-148 */
static inline int cnn_neighbour(const orc_psi* psi, int idx, int k, int i, int j) {
    const int e0 = psi->extent[0], e1 = psi->extent[1], e2 = psi->extent[2];
    const int page_size = e1 * e2;
    const int page = idx / page_size, row = (idx % page_size) / e2, col = (idx % page_size) % e2;
    return ((page + k) % e0) * page_size + ((row + i) % e1) * e2 + ((col + j) % e2);
}

static int cnn_link_offset(const orc_psi* psi, int layer, int c_i, int c_j) {
    int off = 0;
    for (int l = 0; l < layer; l++) {
        const int prev = l > 0 ? psi->channels[l - 1] : 1;
        const int vol = psi->connectivity[l][0] * psi->connectivity[l][1] * psi->connectivity[l][2];
        off += prev * psi->channels[l] * psi->num_sym * vol;
    }
    const int vol = psi->connectivity[layer][0] * psi->connectivity[layer][1] * psi->connectivity[layer][2];
    return off + (c_i * psi->channels[layer] + c_j) * psi->num_sym * vol;
}

/* PsiCNN forward (PsiCNN.hpp:97-160); angles[l] = [channels[l] * N] pre-activations */
static cd cnn_forward(const orc_psi* psi, const uint64_t* conf, cd** angles) {
    const int N = psi->N;
    int max_ch = 1;
    for (int l = 0; l < psi->cnn_layers; l++) if (psi->channels[l] > max_ch) max_ch = psi->channels[l];
    cd* in = (cd*)malloc(sizeof(cd) * (size_t)max_ch * N);
    cd* out = (cd*)malloc(sizeof(cd) * (size_t)max_ch * N);
    for (int j = 0; j < N; j++) in[j] = spin_at(conf, j);
    cd result = 0.0;
    for (int l = 0; l < psi->cnn_layers; l++) {
        const int prev = l > 0 ? psi->channels[l - 1] : 1;
        const int* cv = psi->connectivity[l];
        const int vol = cv[0] * cv[1] * cv[2];
        for (int c_j = 0; c_j < psi->channels[l]; c_j++) {
            for (int x = 0; x < N; x++) {
                cd a = 0.0;
                for (int c_i = 0; c_i < prev; c_i++) {
                    const cd* w = psi->params + cnn_link_offset(psi, l, c_i, c_j)
                                  + psi->symmetry_classes[x] * vol;
                    int ci = 0;
                    for (int k = 0; k < cv[0]; k++)
                        for (int i = 0; i < cv[1]; i++)
                            for (int j = 0; j < cv[2]; j++, ci++)
                                a += w[ci] * in[c_i * N + cnn_neighbour(psi, x, k, i, j)];
                }
                if (angles) angles[l][c_j * N + x] = a;
                out[c_j * N + x] = orc_logcosh(a, (uint32_t)l);
            }
        }
        if (l < psi->cnn_layers - 1) {
            memcpy(in, out, sizeof(cd) * (size_t)psi->channels[l] * N);
        } else {
            for (int j = 0; j < psi->channels[l] * N; j++) result += out[j] * psi->final_factor;
        }
    }
    free(in); free(out);
    return result;
}

cd orc_log_psi_s(const orc_psi* psi, const uint64_t* conf) {
    if (psi->kind == ORC_CNN) return psi->log_prefactor + cnn_forward(psi, conf, NULL);
    const int width = psi->kind == ORC_RBM ? psi->M : psi->size[0];
    cd* theta = (cd*)malloc(sizeof(cd) * width);
    compute_theta(psi, conf, theta);
    const cd r = log_psi_from_theta(psi, theta);
    free(theta);
    return r;
}

void orc_psi_O_k(const orc_psi* psi, const uint64_t* conf, cd* O) {
    const int N = psi->N;
    memset(O, 0, sizeof(cd) * psi->num_params);
    if (psi->kind == ORC_RBM) {
        /* PsiRBM.hpp:159-176 */
        cd* theta = (cd*)malloc(sizeof(cd) * psi->M);
        compute_theta(psi, conf, theta);
        for (int j = 0; j < psi->M; j++) {
            const cd g = psi->final_weight * orc_tanh(theta[j], 0);
            for (int i = 0; i < N; i++) O[i * psi->M + j] = g * spin_at(conf, i);
        }
        free(theta);
        return;
    }
    if (psi->kind == ORC_DEEP) {
        /* PsiDeep.hpp:345-445 with ENABLE_NETWORK_BASES (CMakeLists.txt:169-172) */
        const int L = psi->num_layers;
        cd* z[ORC_MAX_LAYERS];
        for (int l = 0; l < L; l++) z[l] = (cd*)malloc(sizeof(cd) * psi->size[l]);
        cd* theta = (cd*)malloc(sizeof(cd) * psi->size[0]);
        compute_theta(psi, conf, theta);
        deep_forward(psi, theta, z);
        for (int i = 0; i < N; i++) O[i] = spin_at(conf, i);
        int begin[ORC_MAX_LAYERS];
        int off = N;
        for (int l = 0; l < L; l++) { begin[l] = off; off += psi->size[l] + psi->conn[l] * psi->size[l]; }
        cd* delta = (cd*)malloc(sizeof(cd) * 4096);
        cd* ndelta = (cd*)malloc(sizeof(cd) * 4096);
        for (int j = 0; j < psi->size[L - 1]; j++)
            delta[j] = psi->final_weights[j] * orc_tanh(z[L - 1][j], (uint32_t)(L - 1));
        for (int l = L - 1; l >= 0; l--) {
            const int size = psi->size[l], conn = psi->conn[l];
            if (l < L - 1) {
                /* back-propagate through layer l+1 (rhs connections of layer l) */
                const int nsize = psi->size[l + 1], nconn = psi->conn[l + 1];
                for (int i = 0; i < size; i++) ndelta[i] = 0.0;
                for (int i = 0; i < nconn; i++)
                    for (int j = 0; j < nsize; j++)
                        ndelta[psi->connections[l + 1][i * nsize + j]] +=
                            psi->weights[l + 1][i * nsize + j] * delta[j];
                for (int i = 0; i < size; i++) delta[i] = ndelta[i] * orc_tanh(z[l][i], (uint32_t)l);
            }
            for (int j = 0; j < size; j++) {
                O[begin[l] + j] += delta[j];
                for (int i = 0; i < conn; i++) {
                    const int src = (int)psi->connections[l][i * size + j];
                    const cd input = l == 0 ? (cd)spin_at(conf, src) : orc_logcosh(z[l - 1][src], (uint32_t)(l - 1));
                    O[begin[l] + size + i * size + j] += delta[j] * input;
                }
            }
        }
        for (int l = 0; l < L; l++) free(z[l]);
        free(theta); free(delta); free(ndelta);
        return;
    }
    /* PsiCNN.hpp:183-266 */
    cd* angles[ORC_MAX_CNN_LAYERS];
    int max_ch = 1;
    for (int l = 0; l < psi->cnn_layers; l++) {
        angles[l] = (cd*)malloc(sizeof(cd) * (size_t)psi->channels[l] * N);
        if (psi->channels[l] > max_ch) max_ch = psi->channels[l];
    }
    cnn_forward(psi, conf, angles);
    cd* grad_out = (cd*)malloc(sizeof(cd) * (size_t)max_ch * N);   /* d logpsi / d output act */
    cd* grad_pre = (cd*)malloc(sizeof(cd) * (size_t)max_ch * N);   /* d logpsi / d pre-activation */
    const int last = psi->cnn_layers - 1;
    for (int j = 0; j < psi->channels[last] * N; j++) grad_out[j] = psi->final_factor;
    for (int l = last; l >= 0; l--) {
        const int prev = l > 0 ? psi->channels[l - 1] : 1;
        const int* cv = psi->connectivity[l];
        const int vol = cv[0] * cv[1] * cv[2];
        for (int c = 0; c < psi->channels[l]; c++)
            for (int x = 0; x < N; x++)
                grad_pre[c * N + x] = grad_out[c * N + x] * orc_tanh(angles[l][c * N + x], (uint32_t)l);
        for (int j = 0; j < prev * N; j++) grad_out[j] = 0.0;
        for (int c_i = 0; c_i < prev; c_i++) {
            for (int c_j = 0; c_j < psi->channels[l]; c_j++) {
                const int link = cnn_link_offset(psi, l, c_i, c_j);
                for (int x = 0; x < N; x++) {
                    const int cls = (int)psi->symmetry_classes[x];
                    int ci = 0;
                    for (int k = 0; k < cv[0]; k++)
                        for (int i = 0; i < cv[1]; i++)
                            for (int j = 0; j < cv[2]; j++, ci++) {
                                const int src = cnn_neighbour(psi, x, k, i, j);
                                const cd input = l == 0 ? (cd)spin_at(conf, src)
                                                        : orc_logcosh(angles[l - 1][c_i * N + src], (uint32_t)(l - 1));
                                O[link + cls * vol + ci] += grad_pre[c_j * N + x] * input;
                                if (l > 0)
                                    grad_out[c_i * N + src] += psi->params[link + cls * vol + ci] * grad_pre[c_j * N + x];
                            }
                }
            }
        }
    }
    for (int l = 0; l < psi->cnn_layers; l++) free(angles[l]);
    free(grad_out); free(grad_pre);
}

/* ------------------------------------------------------------ operator */

cd orc_local_energy(const orc_operator* op, const orc_psi* psi, const uint64_t* conf, cd log_psi) {
    cd result = 0.0;
    uint64_t prime[ORC_MAX_WORDS];
    for (int n = 0; n < op->num_strings; n++) {
        cd factor;
        orc_pauli_apply(op->words, op->a + (size_t)n * op->words, op->b + (size_t)n * op->words,
                        conf, prime, &factor);
        factor *= op->coefficients[n];
        int diag = 1;
        for (int w = 0; w < op->words; w++) diag &= (prime[w] == conf[w]);
        if (diag) {
            result += factor;
        } else {
            result += factor * cexp(orc_log_psi_s(psi, prime) - log_psi);
        }
    }
    return result;
}

/* E_loc with the reference's incremental angle update (Operator.hpp:60-75:
 * update_input_units, log_psi_s, revert) for wave functions with a dense first
 * layer; theta holds the angles of conf. */
static cd local_energy_incremental(const orc_operator* op, const orc_psi* psi, const cd* Wd, int width,
                                   const uint64_t* conf, const cd* theta, cd log_psi, cd* scratch) {
    cd result = 0.0;
    uint64_t prime[ORC_MAX_WORDS];
    for (int n = 0; n < op->num_strings; n++) {
        cd factor;
        orc_pauli_apply(op->words, op->a + (size_t)n * op->words, op->b + (size_t)n * op->words,
                        conf, prime, &factor);
        factor *= op->coefficients[n];
        int diag = 1;
        for (int w = 0; w < op->words; w++) diag &= (prime[w] == conf[w]);
        if (diag) { result += factor; continue; }
        memcpy(scratch, theta, sizeof(cd) * width);
        for (int w = 0; w < op->words; w++) {
            uint64_t m = prime[w] ^ conf[w];
            while (m) {
                const int i = w * 64 + __builtin_ctzll(m);
                m &= m - 1;
                const double si = spin_at(conf, i);
                for (int h = 0; h < width; h++) scratch[h] += -2.0 * si * Wd[(size_t)i * width + h];
            }
        }
        result += factor * cexp(log_psi_from_theta(psi, scratch) - log_psi);
    }
    return result;
}

/* --------------------------------------------------------- enumeration */

void orc_exact_log_psi_vector(const orc_psi* psi, int32_t num_sites, cd* out) {
    const int64_t n = (int64_t)1 << num_sites;
    #pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < n; idx++) {
        uint64_t conf[ORC_MAX_WORDS] = {(uint64_t)idx, 0, 0, 0};
        out[idx] = orc_log_psi_s(psi, conf);
    }
}

/* ---------------------------------------------------------- statistics */

void orc_tdvp_accumulate(const orc_operator* op, const orc_psi* psi, int64_t num_samples,
                         const uint64_t* confs, const double* weights,
                         cd* E, double* E2, cd* O_mean, cd* F, cd* S,
                         cd* E_loc_out, cd* O_samples_out, int32_t num_threads) {
    const int P = psi->num_params;
    const int words = psi->words;
    cd* O_all = O_samples_out ? O_samples_out : (cd*)malloc(sizeof(cd) * (size_t)num_samples * P);
    cd* El = E_loc_out ? E_loc_out : (cd*)malloc(sizeof(cd) * (size_t)num_samples);
#ifdef _OPENMP
    if (num_threads > 0) omp_set_num_threads(num_threads);
#endif
    int width = 0;
    cd* Wd = psi->kind == ORC_CNN ? NULL : first_layer_dense(psi, &width);
    #pragma omp parallel
    {
        cd* theta = Wd ? (cd*)malloc(sizeof(cd) * width) : NULL;
        cd* scratch = Wd ? (cd*)malloc(sizeof(cd) * width) : NULL;
        #pragma omp for schedule(dynamic, 16)
        for (int64_t s = 0; s < num_samples; s++) {
            const uint64_t* conf = confs + (size_t)s * words;
            if (Wd) {
                compute_theta(psi, conf, theta);
                const cd lp = log_psi_from_theta(psi, theta);
                El[s] = op ? local_energy_incremental(op, psi, Wd, width, conf, theta, lp, scratch) : 0.0;
            } else {
                const cd lp = orc_log_psi_s(psi, conf);
                El[s] = op ? orc_local_energy(op, psi, conf, lp) : 0.0;
            }
            orc_psi_O_k(psi, conf, O_all + (size_t)s * P);
        }
        free(theta);
        free(scratch);
    }
    free(Wd);
    cd e = 0.0; double e2 = 0.0;
    for (int64_t s = 0; s < num_samples; s++) {
        e += weights[s] * El[s];
        e2 += weights[s] * (creal(El[s]) * creal(El[s]) + cimag(El[s]) * cimag(El[s]));
    }
    *E = e; *E2 = e2;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < P; k++) {
        cd o = 0.0, f = 0.0;
        for (int64_t s = 0; s < num_samples; s++) {
            const cd ok = O_all[(size_t)s * P + k];
            o += weights[s] * ok;
            f += weights[s] * El[s] * conj(ok);
        }
        O_mean[k] = o;
        F[k] = f - e * conj(o);
    }
    if (S) {
        #pragma omp parallel for schedule(dynamic, 4)
        for (int k = 0; k < P; k++) {
            for (int kp = 0; kp < P; kp++) {
                cd acc = 0.0;
                for (int64_t s = 0; s < num_samples; s++)
                    acc += weights[s] * conj(O_all[(size_t)s * P + k]) * O_all[(size_t)s * P + kp];
                S[(size_t)k * P + kp] = acc - conj(O_mean[k]) * O_mean[kp];
            }
        }
    }
    if (!O_samples_out) free(O_all);
    if (!E_loc_out) free(El);
}

/* --------------------------------------------------------- Monte Carlo */

/* integer in [0, n) from a 32-bit random word: (x * n) >> 32 (Lemire) */
static inline uint32_t range32(uint32_t x, uint32_t n) { return (uint32_t)(((uint64_t)x * n) >> 32); }

static int select_set_bit(const uint64_t* mask, int words, uint32_t rank) {
    for (int w = 0; w < words; w++) {
        const uint32_t c = (uint32_t)__builtin_popcountll(mask[w]);
        if (rank < c) {
            uint64_t m = mask[w];
            for (uint32_t r = 0; r < rank; r++) m &= m - 1;
            return w * 64 + __builtin_ctzll(m);
        }
        rank -= c;
    }
    return -1;
}

static void init_configuration(const orc_psi* psi, const orc_mc_params* p, uint32_t chain, uint64_t* conf) {
    const int N = psi->N, words = psi->words;
    for (int w = 0; w < ORC_MAX_WORDS; w++) conf[w] = 0;
    if (p->move == ORC_MOVE_FLIP) {
        /* Init_Policy<Spins>: uniformly random bits (Init_Policy.hpp:15-25) */
        for (int blk = 0; blk < (words + 1) / 2; blk++) {
            uint32_t r[4];
            draw(p->seed, chain, (uint32_t)blk, p->call_index, 1u, r);
            conf[2 * blk] = (uint64_t)r[0] | ((uint64_t)r[1] << 32);
            if (2 * blk + 1 < words) conf[2 * blk + 1] = (uint64_t)r[2] | ((uint64_t)r[3] << 32);
        }
        if (N % 64) conf[N / 64] &= (((uint64_t)1) << (N % 64)) - 1u;
        for (int w = (N + 63) / 64; w < ORC_MAX_WORDS; w++) conf[w] = 0;
        return;
    }
    /* exchange moves: uniform configuration with n_up = (N + total_z) / 2 */
    const int n_up = (N + p->total_z) / 2;
    uint64_t down[ORC_MAX_WORDS] = {0, 0, 0, 0};
    for (int i = 0; i < N; i++) down[i >> 6] |= ((uint64_t)1) << (i & 63);
    uint32_t r[4] = {0, 0, 0, 0};
    for (int k = 0; k < n_up; k++) {
        if ((k & 3) == 0) draw(p->seed, chain, (uint32_t)(k >> 2), p->call_index, 2u, r);
        const int pos = select_set_bit(down, words, range32(r[k & 3], (uint32_t)(N - k)));
        down[pos >> 6] &= ~(((uint64_t)1) << (pos & 63));
        conf[pos >> 6] |= ((uint64_t)1) << (pos & 63);
    }
}

typedef struct {
    cd* theta;
    cd* theta_new;
    const cd* Wd;
    int width;
} chain_state;

static cd state_log_psi(const orc_psi* psi, const uint64_t* conf, chain_state* st, int use_new) {
    if (psi->kind == ORC_CNN) return orc_log_psi_s(psi, conf);
    return log_psi_from_theta(psi, use_new ? st->theta_new : st->theta);
}

void orc_mc_run(const orc_psi* psi, const orc_mc_params* p, uint64_t* chain_confs,
                uint64_t* sample_confs, cd* sample_log_psi,
                uint64_t* acceptances, uint64_t* rejections, int32_t num_threads) {
    const int N = psi->N, words = psi->words;
    int width = 0;
    cd* Wd = psi->kind == ORC_CNN ? NULL : first_layer_dense(psi, &width);
    uint64_t acc_total = 0, rej_total = 0;
#ifdef _OPENMP
    if (num_threads > 0) omp_set_num_threads(num_threads);
#endif
    #pragma omp parallel for schedule(dynamic, 1) reduction(+:acc_total, rej_total)
    for (int c = 0; c < p->num_chains; c++) {
        uint64_t conf[ORC_MAX_WORDS] = {0, 0, 0, 0}, next[ORC_MAX_WORDS];
        if (p->init_random) init_configuration(psi, p, (uint32_t)c, conf);
        else for (int w = 0; w < words; w++) conf[w] = chain_confs[(size_t)c * words + w];
        chain_state st = {NULL, NULL, Wd, width};
        if (Wd) {
            st.theta = (cd*)malloc(sizeof(cd) * width);
            st.theta_new = (cd*)malloc(sizeof(cd) * width);
            compute_theta(psi, conf, st.theta);
        }
        cd log_psi = state_log_psi(psi, conf, &st, 0);
        uint64_t t = p->counter;
        const int64_t total = (int64_t)p->num_thermalization_sweeps * N
                              + (int64_t)p->steps_per_chain * p->num_sweeps * N;
        int64_t sample_at = (int64_t)p->num_thermalization_sweeps * N + (int64_t)p->num_sweeps * N;
        int step = 0;
        for (int64_t q = 0; q <= total; q++) {
            if (q == sample_at && step < p->steps_per_chain) {
                const size_t idx = (size_t)step * p->num_chains + c;   /* MonteCarlo.hpp:108 */
                for (int w = 0; w < words; w++) sample_confs[idx * words + w] = conf[w];
                sample_log_psi[idx] = log_psi;
                step++;
                sample_at += (int64_t)p->num_sweeps * N;
            }
            if (q == total) break;
            uint32_t r[4];
            draw(p->seed, (uint32_t)c, (uint32_t)t, (uint32_t)(t >> 32), 0u, r);
            t++;
            int i = (int)range32(r[0], (uint32_t)N), j = -1;
            if (p->move == ORC_MOVE_EXCHANGE) {
                uint64_t opp[ORC_MAX_WORDS];
                const int up = spin_at(conf, i) > 0;
                int n_opp = 0;
                for (int w = 0; w < words; w++) {
                    uint64_t m = up ? ~conf[w] : conf[w];
                    if (w == words - 1 && (N % 64)) m &= (((uint64_t)1) << (N % 64)) - 1u;
                    opp[w] = m;
                    n_opp += __builtin_popcountll(m);
                }
                if (n_opp == 0) { rej_total++; continue; }
                j = select_set_bit(opp, words, range32(r[1], (uint32_t)n_opp));
            }
            for (int w = 0; w < words; w++) next[w] = conf[w];
            next[i >> 6] ^= ((uint64_t)1) << (i & 63);
            if (j >= 0) next[j >> 6] ^= ((uint64_t)1) << (j & 63);
            if (Wd) {
                const double si = spin_at(conf, i), sj = j >= 0 ? spin_at(conf, j) : 0.0;
                for (int h = 0; h < width; h++) {
                    cd d = -2.0 * si * Wd[(size_t)i * width + h];
                    if (j >= 0) d += -2.0 * sj * Wd[(size_t)j * width + h];
                    st.theta_new[h] = st.theta[h] + d;
                }
            }
            const cd next_log_psi = state_log_psi(psi, next, &st, 1);
            const double dre2 = 2.0 * (creal(next_log_psi) - creal(log_psi));
            const double u = ((double)r[2] + 0.5) * (1.0 / 4294967296.0);
            if (dre2 > 0.0 || u <= exp(dre2)) {         /* MonteCarlo.hpp:158-160 */
                for (int w = 0; w < words; w++) conf[w] = next[w];
                log_psi = next_log_psi;
                if (Wd) { cd* tmp = st.theta; st.theta = st.theta_new; st.theta_new = tmp; }
                acc_total++;
            } else {
                rej_total++;
            }
        }
        for (int w = 0; w < words; w++) chain_confs[(size_t)c * words + w] = conf[w];
        free(st.theta); free(st.theta_new);
    }
    free(Wd);
    *acceptances = acc_total;
    *rejections = rej_total;
}

/* ------------------------------------------------------------- solvers */

int orc_cholesky_solve(int32_t n, const cd* S, double shift, const cd* F, cd* x) {
    cd* L = (cd*)malloc(sizeof(cd) * (size_t)n * n);
    for (size_t k = 0; k < (size_t)n * n; k++) L[k] = S[k];
    for (int i = 0; i < n; i++) L[(size_t)i * n + i] += shift;
    /* A = L L^H, lower triangle in place (row-major) */
    for (int j = 0; j < n; j++) {
        double d = creal(L[(size_t)j * n + j]);
        for (int k = 0; k < j; k++) {
            const cd v = L[(size_t)j * n + k];
            d -= creal(v) * creal(v) + cimag(v) * cimag(v);
        }
        if (!(d > 0.0)) { free(L); return -1; }
        d = sqrt(d);
        L[(size_t)j * n + j] = d;
        #pragma omp parallel for schedule(static)
        for (int i = j + 1; i < n; i++) {
            cd acc = L[(size_t)i * n + j];
            for (int k = 0; k < j; k++) acc -= L[(size_t)i * n + k] * conj(L[(size_t)j * n + k]);
            L[(size_t)i * n + j] = acc / d;
        }
    }
    cd* y = (cd*)malloc(sizeof(cd) * n);
    for (int i = 0; i < n; i++) {
        cd acc = F[i];
        for (int k = 0; k < i; k++) acc -= L[(size_t)i * n + k] * y[k];
        y[i] = acc / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
        cd acc = y[i];
        for (int k = i + 1; k < n; k++) acc -= conj(L[(size_t)k * n + i]) * x[k];
        x[i] = acc / L[(size_t)i * n + i];
    }
    free(L); free(y);
    return 0;
}

static void dense_matvec(int n, const cd* S, double shift, const cd* v, cd* y) {
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        cd acc = shift * v[i];
        for (int k = 0; k < n; k++) acc += S[(size_t)i * n + k] * v[k];
        y[i] = acc;
    }
}

int orc_cg_solve(int32_t n, const cd* S, double shift, const cd* F, cd* x,
                 double rtol, int32_t max_iter, int32_t* iterations) {
    cd* r = (cd*)malloc(sizeof(cd) * n);
    cd* p = (cd*)malloc(sizeof(cd) * n);
    cd* Ap = (cd*)malloc(sizeof(cd) * n);
    double bnorm = 0.0, rr = 0.0;
    for (int i = 0; i < n; i++) {
        x[i] = 0.0; r[i] = F[i]; p[i] = F[i];
        rr += creal(F[i] * conj(F[i]));
    }
    bnorm = sqrt(rr);
    int it = 0;
    while (it < max_iter && sqrt(rr) > rtol * bnorm) {
        dense_matvec(n, S, shift, p, Ap);
        cd pAp = 0.0;
        for (int i = 0; i < n; i++) pAp += conj(p[i]) * Ap[i];
        const double alpha = rr / creal(pAp);
        double rr_new = 0.0;
        for (int i = 0; i < n; i++) {
            x[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
            rr_new += creal(r[i] * conj(r[i]));
        }
        const double beta = rr_new / rr;
        for (int i = 0; i < n; i++) p[i] = r[i] + beta * p[i];
        rr = rr_new;
        it++;
    }
    *iterations = it;
    free(r); free(p); free(Ap);
    return sqrt(rr) <= rtol * bnorm ? 0 : 1;
}

/* (S + shift) x = F by Jacobi-preconditioned CG with the matrix-free product
 * S v = sum_s w_s conj(Oc_s) (Oc_s . v), Oc_s = O_s - <O>, on the stored O_k samples
 * [ns][P]: the product a caller of the reference's TDVP.S_dot_vector
 * (source/network_functions/TDVP.cu.template:336-443) iterates on, restated on the
 * CPU for the baseline and for checks of the device solve.  Returns the iterations;
 * *res gets the final relative residual of the recursion. */
static void sv_matfree(int64_t ns, int P, const cd* O, const double* w, const cd* Ob, double shift,
                       const cd* v, cd* u, cd* y) {
    cd ov = 0.0;
    for (int k = 0; k < P; k++) ov += Ob[k] * v[k];
    #pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < ns; s++) {
        const cd* row = O + (size_t)s * P;
        cd acc = 0.0;
        for (int k = 0; k < P; k++) acc += row[k] * v[k];
        u[s] = w[s] * (acc - ov);
    }
    cd su = 0.0;
    for (int64_t s = 0; s < ns; s++) su += u[s];
    #pragma omp parallel
    {
        const int nt = omp_get_num_threads(), id = omp_get_thread_num();
        const int k0 = (int)((int64_t)P * id / nt), k1 = (int)((int64_t)P * (id + 1) / nt);
        for (int k = k0; k < k1; k++) y[k] = shift * v[k] - conj(Ob[k]) * su;
        for (int64_t s = 0; s < ns; s++) {
            const cd* row = O + (size_t)s * P;
            const cd us = u[s];
            for (int k = k0; k < k1; k++) y[k] += conj(row[k]) * us;
        }
    }
}

int orc_sr_cg_matfree(int64_t ns, int32_t P, const cd* O_samples, const double* w, const cd* O_mean,
                      const cd* F, double shift, double rtol, int32_t max_iter, cd* x, double* res,
                      int32_t num_threads) {
#ifdef _OPENMP
    if (num_threads > 0) omp_set_num_threads(num_threads);
#endif
    cd* r = (cd*)malloc(sizeof(cd) * P);
    cd* z = (cd*)malloc(sizeof(cd) * P);
    cd* p = (cd*)malloc(sizeof(cd) * P);
    cd* Ap = (cd*)malloc(sizeof(cd) * P);
    cd* u = (cd*)malloc(sizeof(cd) * (size_t)ns);
    double* dinv = (double*)malloc(sizeof(double) * P);
    /* Jacobi: diagonal of S + shift */
    for (int k = 0; k < P; k++) dinv[k] = shift;
    for (int64_t s = 0; s < ns; s++)
        for (int k = 0; k < P; k++) {
            const cd d = O_samples[(size_t)s * P + k] - O_mean[k];
            dinv[k] += w[s] * (creal(d) * creal(d) + cimag(d) * cimag(d));
        }
    double fn = 0.0, rz = 0.0, rr = 0.0;
    for (int k = 0; k < P; k++) {
        dinv[k] = 1.0 / dinv[k];
        x[k] = 0.0;
        r[k] = F[k];
        z[k] = dinv[k] * r[k];
        p[k] = z[k];
        fn += creal(F[k] * conj(F[k]));
        rz += creal(conj(r[k]) * z[k]);
    }
    rr = fn;
    int it = 0;
    while (it < max_iter && sqrt(rr) > rtol * sqrt(fn)) {
        sv_matfree(ns, P, O_samples, w, O_mean, shift, p, u, Ap);
        cd pAp = 0.0;
        for (int k = 0; k < P; k++) pAp += conj(p[k]) * Ap[k];
        const double alpha = rz / creal(pAp);
        double rz_new = 0.0;
        rr = 0.0;
        for (int k = 0; k < P; k++) {
            x[k] += alpha * p[k];
            r[k] -= alpha * Ap[k];
            z[k] = dinv[k] * r[k];
            rz_new += creal(conj(r[k]) * z[k]);
            rr += creal(r[k] * conj(r[k]));
        }
        const double beta = rz_new / rz;
        for (int k = 0; k < P; k++) p[k] = z[k] + beta * p[k];
        rz = rz_new;
        it++;
    }
    if (res) *res = sqrt(rr / fn);
    free(r); free(z); free(p); free(Ap); free(u); free(dinv);
    return it;
}
