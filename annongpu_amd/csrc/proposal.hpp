// proposal.hpp -- the learned block proposal: a two-layer tanh network trained
// on the sampler's own configurations (fp32 training kernels) and a
// Metropolis-Hastings block sweep that uses it (proposal.hip).
//
// Network: x_i = s_i outside the block B, 0 inside; h = tanh(W1 x + b1);
// l = W2 h + b2.  Parameters (fp32) are packed [W1 (H x N) | b1 | W2 (N x H) | b2].
// A block move redraws the spins of the sites B_0..B_{bs-1} among the patterns
// with the same number of up spins, pattern p having probability
// exp(sum_{j: bit j of p} l_{B_j}) / Z.  The logits depend only on the spins
// outside B, which the move leaves unchanged, so the reverse move has the same
// Z and the proposal ratio q(s|s') / q(s'|s) is exp(e_old - e_new).
#pragma once

#include "kernels.hpp"

namespace ann {

constexpr int kMaxBlock = 8;          // patterns of a block: 2^8 = 4 per lane
constexpr float kFixScale = 1048576.f; // first-layer fixed point: 2^20 per unit

// inference form of the network, read by the block sweep
struct BlockNet {
    int N, H, bs, num_blocks;
    const int* blocks;        // [num_blocks][bs] sites
    const int32_t* W1q;       // [N][H] first layer, fixed point (site-major)
    const int32_t* b1q;       // [H]
    const uint16_t* W2h;      // [N][H] second layer, bf16 (row of output o)
    const float* b2;          // [N]
    int exclusive;            // propose only patterns other than the current one
};

// quantise the fp32 parameters into the inference form
void launch_block_net_pack(const float* params, int N, int H, int32_t* W1q, int32_t* b1q, uint16_t* W2h, float* b2,
                           hipStream_t stream);

// num_proposals block moves per chain on Philox stream 3 (counter = counter + q);
// accepted[c] counts the moves that changed the configuration, bstat[c][0] the
// redraws of the current pattern (nothing to evaluate), bstat[c][1] the flipped
// sites whose rows of W entered a psi ratio
template <typename T>
void launch_block_sweep(const DenseView<T>& net, const BlockNet& bn, uint64_t* chain_confs, cplx<T>* theta,
                        int num_chains, int words, int64_t num_proposals, uint64_t counter, uint64_t seed,
                        uint32_t* accepted, uint32_t* bstat, hipStream_t stream);

// ------------------------------------------------------------------ training
struct TrainBatch {
    int Bt;                   // rows of the batch
    int64_t* sample;          // [Bt] indices into the configuration array
    int* sites;               // [Bt][bs] block sites of each row
    float* X;                 // [Bt][N] inputs
    float* Y;                 // [Bt][N] targets (1 = up)
    float* Mask;              // [Bt][N] 1 on the block's sites
    float* Hd;                // [Bt][H] hidden activations
    float* A2;                // [Bt][N] logits, then d loss / d logits
    float* dH;                // [Bt][H] d loss / d hidden pre-activations
    double* loss_part;        // [Bt] per-row losses
    float* gpart;             // [kGradSplits][N H + max(N, H)] split reductions of the gradient
};
constexpr int kGradSplits = 32;

// rows of a random batch: sample = range32(x, count), block = range32(y, num_blocks)
// from Philox((row, step, 0, 4), seed)
void launch_draw_batch(const TrainBatch& b, int64_t count, const int* blocks, int num_blocks, int bs,
                       uint64_t step, uint64_t seed, hipStream_t stream);
void launch_build_inputs(const TrainBatch& b, const uint64_t* confs, int words, int N, int bs, hipStream_t stream);
// loss (mean over rows of the block's summed softplus(l) - y l) and the gradient,
// packed like the parameters; count_norm = number of rows in the mean
void launch_loss_grad(const TrainBatch& b, const float* params, int N, int H, double count_norm, float* grad,
                      hipStream_t stream);
void launch_adam(float* params, float* m, float* v, const float* grad, int64_t n, float lr, float beta1,
                 float beta2, float eps, int64_t t, hipStream_t stream);

}  // namespace ann
