/*
 * proposal_oracle.c -- fp64 CPU restatement of the learned block proposal.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP proposal network (forward,
 * backward, Adam) and for the block-move proposal probabilities.  Nothing in
 * annongpu_amd/ links or loads this file.
 *
 * The reference has no learned proposal; the definitions restated here are the
 * ones the HIP kernels implement:
 *   network   x_i = s_i for sites outside the block B, 0 inside;
 *             h = tanh(W1 x + b1)  (H units),  l = W2 h + b2  (N logits);
 *             parameters packed as [W1 (H x N) | b1 (H) | W2 (N x H) | b2 (N)]
 *   loss      mean over samples of sum_{i in B} softplus(l_i) - y_i l_i,
 *             y_i = 1 if s_i = +1 (pseudo-likelihood of the block given the rest)
 *   Adam      Kingma & Ba (2015), bias-corrected first and second moments
 *   proposal  the block's new pattern p (bit j = spin of site B_j up) is drawn
 *             among the patterns with the current number of up spins with
 *             probability exp(sum_{j in p} l_{B_j}) / Z  (conditional Poisson)
 */
#include "vmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static int spin_up(const uint64_t* conf, int i) { return (int)((conf[i >> 6] >> (i & 63)) & 1u); }

static void forward(int N, int H, const double* p, const double* x, double* h, double* l) {
    const double* W1 = p;
    const double* b1 = W1 + (size_t)H * N;
    const double* W2 = b1 + H;
    const double* b2 = W2 + (size_t)N * H;
    for (int k = 0; k < H; k++) {
        double a = b1[k];
        for (int i = 0; i < N; i++) a += W1[(size_t)k * N + i] * x[i];
        h[k] = tanh(a);
    }
    for (int o = 0; o < N; o++) {
        double a = b2[o];
        for (int k = 0; k < H; k++) a += W2[(size_t)o * H + k] * h[k];
        l[o] = a;
    }
}

static void inputs(int N, const uint64_t* conf, const int32_t* block, int bs, double* x) {
    for (int i = 0; i < N; i++) x[i] = spin_up(conf, i) ? 1.0 : -1.0;
    for (int j = 0; j < bs; j++) x[block[j]] = 0.0;
}

void orc_mlp_logits(int32_t N, int32_t H, const double* params, const uint64_t* conf,
                    const int32_t* block, int32_t bs, double* logits) {
    double* x = (double*)malloc(sizeof(double) * N);
    double* h = (double*)malloc(sizeof(double) * H);
    inputs(N, conf, block, bs, x);
    forward(N, H, params, x, h, logits);
    free(x);
    free(h);
}

static double softplus(double z) { return z > 0 ? z + log1p(exp(-z)) : log1p(exp(z)); }
static double sigmoid(double z) { return 1.0 / (1.0 + exp(-z)); }

double orc_mlp_loss_grad(int32_t N, int32_t H, const double* params, int64_t count,
                         const uint64_t* confs, int32_t words, const int32_t* blocks, int32_t bs,
                         double* grad) {
    const size_t P = (size_t)H * N + H + (size_t)N * H + N;
    const double* W2 = params + (size_t)H * N + H;
    double* gW1 = grad;
    double* gb1 = gW1 + (size_t)H * N;
    double* gW2 = gb1 + H;
    double* gb2 = gW2 + (size_t)N * H;
    memset(grad, 0, sizeof(double) * P);
    double* x = (double*)malloc(sizeof(double) * N);
    double* h = (double*)malloc(sizeof(double) * H);
    double* l = (double*)malloc(sizeof(double) * N);
    double* dl = (double*)malloc(sizeof(double) * N);
    double* da = (double*)malloc(sizeof(double) * H);
    double loss = 0.0;
    const double inv = 1.0 / (double)count;
    for (int64_t s = 0; s < count; s++) {
        const uint64_t* conf = confs + s * words;
        const int32_t* blk = blocks + s * bs;
        inputs(N, conf, blk, bs, x);
        forward(N, H, params, x, h, l);
        for (int o = 0; o < N; o++) dl[o] = 0.0;
        for (int j = 0; j < bs; j++) {
            const int o = blk[j];
            const double y = spin_up(conf, o) ? 1.0 : 0.0;
            loss += inv * (softplus(l[o]) - y * l[o]);
            dl[o] = inv * (sigmoid(l[o]) - y);
        }
        for (int k = 0; k < H; k++) da[k] = 0.0;
        for (int o = 0; o < N; o++) {
            if (dl[o] == 0.0) continue;
            gb2[o] += dl[o];
            for (int k = 0; k < H; k++) {
                gW2[(size_t)o * H + k] += dl[o] * h[k];
                da[k] += dl[o] * W2[(size_t)o * H + k];
            }
        }
        for (int k = 0; k < H; k++) {
            const double d = da[k] * (1.0 - h[k] * h[k]);
            gb1[k] += d;
            for (int i = 0; i < N; i++) gW1[(size_t)k * N + i] += d * x[i];
        }
    }
    free(x); free(h); free(l); free(dl); free(da);
    return loss;
}

void orc_adam_step(int64_t n, double* params, double* m, double* v, const double* grad, double lr,
                   double beta1, double beta2, double eps, int64_t t) {
    const double c1 = 1.0 - pow(beta1, (double)t), c2 = 1.0 - pow(beta2, (double)t);
    for (int64_t k = 0; k < n; k++) {
        m[k] = beta1 * m[k] + (1.0 - beta1) * grad[k];
        v[k] = beta2 * v[k] + (1.0 - beta2) * grad[k] * grad[k];
        params[k] -= lr * (m[k] / c1) / (sqrt(v[k] / c2) + eps);
    }
}

/* log q(p_new | context) for the block pattern p_new, given the logits of the
 * block's sites; the normalisation runs over the patterns with popcount k */
double orc_block_log_q(const double* block_logits, int32_t bs, uint32_t p_new, int32_t k) {
    double emax = -INFINITY;
    for (uint32_t p = 0; p < (1u << bs); p++) {
        if (__builtin_popcount(p) != k) continue;
        double e = 0.0;
        for (int j = 0; j < bs; j++) if (p >> j & 1u) e += block_logits[j];
        if (e > emax) emax = e;
    }
    double z = 0.0, enew = 0.0;
    for (uint32_t p = 0; p < (1u << bs); p++) {
        if (__builtin_popcount(p) != k) continue;
        double e = 0.0;
        for (int j = 0; j < bs; j++) if (p >> j & 1u) e += block_logits[j];
        z += exp(e - emax);
        if (p == p_new) enew = e;
    }
    return enew - emax - log(z);
}
