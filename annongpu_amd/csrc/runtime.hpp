// runtime.hpp -- host-side objects behind the C ABI: device buffers, the stream,
// per-kernel hipEvent timers, wave functions, operators, ensembles and TDVP state.
#pragma once

#include "cnn.hpp"
#include "ops.hpp"
#include "proposal.hpp"

#include <complex>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace ann {

using cd = std::complex<double>;

// the library's HIP stream (non-blocking); every copy and launch goes through it
hipStream_t lib_stream();

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        release();
        p = o.p;
        n = o.n;
        o.p = nullptr;
        o.n = 0;
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void ensure(size_t m) {
        if (m <= n && p) return;
        release();
        ANN_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * (m > 0 ? m : 1)));
        n = m;
    }
    // ordered after the work already queued on the library stream; returns when done
    void upload(const T* h, size_t m) {
        ensure(m);
        if (m) {
            ANN_HIP_CHECK(hipMemcpyAsync(p, h, sizeof(T) * m, hipMemcpyHostToDevice, lib_stream()));
            ANN_HIP_CHECK(hipStreamSynchronize(lib_stream()));
        }
    }
    void download(T* h, size_t m) const {
        if (m) {
            ANN_HIP_CHECK(hipMemcpyAsync(h, p, sizeof(T) * m, hipMemcpyDeviceToHost, lib_stream()));
            ANN_HIP_CHECK(hipStreamSynchronize(lib_stream()));
        }
    }
};

class Runtime {
public:
    static Runtime& get();
    hipStream_t stream();
    // two more streams (default priority) for work that overlaps the library stream;
    // while a SideStream guard lives, stream() returns the one it names
    hipStream_t side(int k = 0);
    hipStream_t override_ = nullptr;
    void set_device(int device);
    void sync();
    // per-kernel-class device timers (hipEvents on the library stream); a class
    // may be timed on a sample of its launches (`untimed` counts the others)
    bool profiling = false;
    void begin(const char* name);
    void end();
    void collect();
    std::map<std::string, std::pair<double, int64_t>> totals;   // timed ms, timed launches
    std::map<std::string, int64_t> untimed;

private:
    hipStream_t stream_ = nullptr;
    hipStream_t side_[2] = {nullptr, nullptr};
    int device_ = 0;
    struct Pending {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> pool_;
    std::string open_name_;
    hipEvent_t open_a_ = nullptr;
    hipEvent_t event();
};

// launches inside the guard go to the side stream (no synchronisation may happen there)
struct SideStream {
    explicit SideStream(int k = 0) { Runtime::get().override_ = Runtime::get().side(k); }
    ~SideStream() { Runtime::get().override_ = nullptr; }
    SideStream(const SideStream&) = delete;
    SideStream& operator=(const SideStream&) = delete;
};

// device time of the enclosed launches under `name` (nothing when profiling is
// off); timed = false only counts the launch (sampled timing of launches in tight
// loops, where the events themselves would stall the queue)
struct ScopedTimer {
    explicit ScopedTimer(const char* name, bool timed = true) {
        if (!Runtime::get().profiling) return;
        if (timed) {
            Runtime::get().begin(name);
            on = true;
        } else {
            Runtime::get().untimed[name]++;
        }
    }
    ~ScopedTimer() {
        if (on) Runtime::get().end();
    }
    bool on = false;
};

// page-locked host buffer (hipHostMalloc), for asynchronous device-to-host reads
template <typename T>
struct PinnedBuf {
    T* p = nullptr;
    size_t n = 0;
    PinnedBuf() = default;
    PinnedBuf(PinnedBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    PinnedBuf& operator=(PinnedBuf&& o) noexcept {
        if (this != &o) {
            if (p) (void)hipHostFree(p);
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    void ensure(size_t m) {
        if (m <= n && p) return;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
        ANN_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&p), sizeof(T) * (m > 0 ? m : 1), hipHostMallocDefault));
        n = m;
    }
};

// a HIP event created on first use (no timing) and destroyed with its owner
struct OwnedEvent {
    hipEvent_t e = nullptr;
    OwnedEvent() = default;
    OwnedEvent(const OwnedEvent&) = delete;
    OwnedEvent& operator=(const OwnedEvent&) = delete;
    OwnedEvent(OwnedEvent&& o) noexcept : e(o.e) { o.e = nullptr; }
    OwnedEvent& operator=(OwnedEvent&& o) noexcept {
        if (this != &o) {
            if (e) (void)hipEventDestroy(e);
            e = o.e;
            o.e = nullptr;
        }
        return *this;
    }
    ~OwnedEvent() {
        if (e) (void)hipEventDestroy(e);
    }
    hipEvent_t get() {
        if (!e) ANN_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        return e;
    }
};

enum PsiKind : int { PSI_RBM = 0, PSI_DEEP = 1, PSI_CNN = 2, PSI_CLASSICAL = 3 };

struct Psi {
    int kind = PSI_RBM;
    int num_sites = 0, N = 0, num_params = 0, words = 1;
    cd log_prefactor = 0.0;
    int precision = 0;  // ANN_FP64 / ANN_FP32
    // dense nets (PsiRBM, PsiDeep with one hidden layer)
    int M = 0, C = 0, Mpad = 0, layout = LAYOUT_RBM;
    std::vector<cd> W, a, b, f;
    std::vector<uint32_t> conn;
    // multi-layer PsiDeep (host description)
    int num_layers = 1;
    std::vector<int> sizes, conns;
    std::vector<std::vector<uint32_t>> conn_l;
    std::vector<std::vector<cd>> w_l, b_l;
    // PsiCNN
    int extent[3] = {1, 1, 1};
    int cnn_layers = 0;
    std::vector<int> channels;
    std::vector<int> connectivity;  // [L][3]
    std::vector<uint32_t> sym;
    int num_sym = 1;
    std::vector<cd> params;  // PsiCNN parameters; the classical parameters of PsiClassical
    double final_factor = 1.0;
    // PsiClassical: H_local[n] holds the strings cl_begin[n] .. cl_begin[n + 1] - 1 (a and b
    // words, coefficients).  ref is the network reference (a PsiDeep), null for the fully
    // polarised one (log psi_ref = 0, no parameters); with cl_order == 2 it enters log psi
    // and its parameters follow the classical ones.
    int cl_order = 1;
    std::vector<int> cl_begin;
    std::vector<uint64_t> cl_a, cl_b;   // [strings][kMaxWords]
    std::vector<cd> cl_coef;
    cd ref_log_prefactor = 0.0;
    std::unique_ptr<Psi> ref;
    DevBuf<int> cl_begin_d;
    DevBuf<uint64_t> cl_a_d, cl_b_d;
    DevBuf<double> cl_coef_s_d, cl_coef_p_d;   // spin basis (times (-i)^{#Y}), Pauli basis

    // device state: dirty asks for the full upload; params_dirty (after the first upload,
    // dense nets) only for the parameters, staged in page-locked memory and repacked
    // on the device
    bool dirty = true;
    bool params_dirty = false;
    PinnedBuf<cplx<double>> stage;
    DevBuf<cplx<double>> params_d;
    OwnedEvent stage_ev;   // the staged copy has left the host buffer
    int conn_shift = -1;   // PsiDeep: connection i of unit j is site (j shift + i) mod N (else -1)
    bool real_f = true;
    DevBuf<cplx<double>> Wd64, b64, f64;
    DevBuf<cplx<float>> Wd32, b32, f32;
    DevBuf<uint32_t> conn_d;
    DevBuf<uint16_t> wfrag;
    int npad = 0;
    DevBuf<cplx<double>> zero64;
    DevBuf<cplx<float>> zero32;
    // PsiCNN device state
    DevBuf<cplx<double>> cp64;
    DevBuf<cplx<float>> cp32;
    DevBuf<int> nbr_d;
    DevBuf<uint32_t> sym_d;
    DevBuf<uint64_t> cons_d;
    // multi-layer PsiDeep device state (layers concatenated; offsets per layer)
    DevBuf<uint32_t> mlp_cn;
    DevBuf<cplx<double>> mlp_w64, mlp_b64;
    DevBuf<cplx<float>> mlp_w32, mlp_b32;
    DevBuf<int> mlp_rb, mlp_ri;
    std::vector<int> mlp_cn_off, mlp_w_off, mlp_b_off, mlp_rb_off, mlp_ri_off;

    bool mlp_stale = true;     // the layer tables of a one-layer PsiDeep's generic view
    bool dense() const { return kind == PSI_RBM || (kind == PSI_DEEP && num_layers == 1); }
    // a PsiDeep over Pauli strings: 3 input units per site
    bool pauli_net() const { return kind == PSI_DEEP && N == 3 * num_sites; }
    bool ref_enters() const { return kind == PSI_CLASSICAL && cl_order == 2 && ref != nullptr; }
    void finalize();           // derived sizes after construction
    void sync_device();        // upload parameters when dirty
    DenseView<double> view64();
    DenseView<float> view32();
    CnnView<double> cnn64();
    CnnView<float> cnn32();
    MlpView<double> mlp64();
    MlpView<float> mlp32();
    ClassicalView<double> classical64(int basis);
    ClassicalView<float> classical32(int basis);
    // the same with a PsiCNN reference (spin basis)
    ClassicalView<double, CnnView<double>> classical_cnn64();
    ClassicalView<float, CnnView<float>> classical_cnn32();
    bool cnn_ref() const { return kind == PSI_CLASSICAL && ref && ref->kind == PSI_CNN; }
    void upload_mlp_tables();
    std::vector<cd> get_params() const;
    void set_params(const cd* p);
    std::unique_ptr<Psi> clone() const;
};

struct DeviceOp {
    int words = 0;
    int num_groups = 0;
    DevBuf<uint64_t> group_flip, string_zy;
    DevBuf<int> group_begin;
    DevBuf<double> string_coef;
    OpView view() const {
        return {num_groups, words, group_flip.p, group_begin.p, string_zy.p, string_coef.p};
    }
};

// the strings as they act on Pauli-string configurations (at most 64 sites)
struct DevicePauliOp {
    int num_strings = 0;
    DevBuf<uint64_t> ab;
    DevBuf<double> coef;
    PauliOpView view() const { return {num_strings, ab.p, coef.p}; }
};

struct Operator {
    int num_strings = 0;
    std::vector<uint64_t> a, b;  // [n][kMaxWords]
    std::vector<cd> coef;
    std::map<int, std::unique_ptr<DeviceOp>> compiled;
    std::unique_ptr<DevicePauliOp> pauli_compiled;
    const DeviceOp& device(int words);
    const DevicePauliOp& device_pauli();
    int max_site() const;
};

struct Samples {
    int64_t count = 0;
    int words = 0, Mpad = 0, precision = 0;
    int basis = BASIS_SPINS;
    DevBuf<uint64_t> confs;
    DevBuf<double> w;
    bool w_uniform = false;      // every sample weighs 1 / count (Monte Carlo)
    DevBuf<cplx<double>> log_psi;
    DevBuf<cplx<float>> th32;
    DevBuf<cplx<double>> th64;
    DevBuf<cplx<double>> e_loc;
};

// learned block proposal: network parameters with Adam state, the inference
// form read by the block sweep, and the training batch workspace
struct Proposal {
    int N = 0, H = 0, bs = 0, num_blocks = 0;
    uint64_t seed = 0;
    int64_t num_params() const { return int64_t(H) * N + H + int64_t(N) * H + N; }
    DevBuf<int> blocks;
    DevBuf<float> params, m, v, grad;
    int64_t adam_t = 0;
    uint64_t train_step = 0;
    bool packed = false;
    DevBuf<int32_t> W1q, b1q;
    DevBuf<uint16_t> W2h;
    DevBuf<float> b2;
    // batch workspace (grow-only)
    DevBuf<int64_t> bsample;
    DevBuf<int> bsites;
    DevBuf<float> X, Y, Mask, Hd, A2, dH;
    DevBuf<double> loss_part;
    DevBuf<float> gpart;
    TrainBatch batch(int Bt);
    BlockNet view();
    double last_loss = 0.0;
};

enum EnsembleKind : int { ENS_MC = 0, ENS_EXACT = 1, ENS_LIST = 2 };

struct Ensemble {
    int kind = ENS_MC;
    int basis = BASIS_SPINS;
    uint32_t num_sites = 0;
    uint32_t num_samples = 0, num_sweeps = 0, num_therm = 0, num_chains = 0;
    int move = MOVE_FLIP, total_z = 0;
    uint64_t seed = 0, counter = 0;
    uint32_t call_index = 0;
    bool persistent = false, chains_valid = false;
    int chain_words = 0, chain_N = 0;
    DevBuf<uint64_t> chains;
    DevBuf<uint32_t> acc;
    DevBuf<cplx<float>> chain_th32;
    DevBuf<cplx<double>> chain_th64;
    uint64_t accepted = 0, rejected = 0;
    // per-chain accepted proposals of the last sampling call, copied to the host
    // asynchronously and counted when the acceptance is asked for
    PinnedBuf<uint32_t> acc_host;
    uint64_t acc_total = 0;
    bool acc_pending = false;
    void settle_acceptance();
    // learned block proposal (not owned); block sweeps per sample and plain sweeps per sample
    Proposal* proposal = nullptr;
    uint32_t plain_sweeps = 1;
    bool proposal_exclusive = false;   // block moves never redraw the current pattern
    // block-move counters of the last call: per chain [null redraws, rows evaluated],
    // summed on the host when asked for, and the block proposals made
    DevBuf<uint32_t> bstat;
    PinnedBuf<uint32_t> bstat_host;
    uint64_t block_proposals = 0, null_moves = 0, rows_evaluated = 0;
    bool bstat_pending = false;
    Samples samples;
    // caller-given configurations [num_samples][kMaxWords] and weights (ENS_LIST)
    std::vector<uint64_t> list_confs;
    std::vector<double> list_w;
    uint64_t num_steps() const {
        if (kind != ENS_EXACT) return num_samples;
        return uint64_t(1) << (basis == BASIS_PAULIS ? 2 * num_sites : num_sites);
    }
};


struct Tdvp {
    uint32_t P = 0;
    cd E = 0.0;
    double E2 = 0.0, wsum = 0.0;
    std::vector<cd> O, F;
    bool has_S = false, has_samples = false;
    DevBuf<cplx<double>> S, Obar, Fd;
    DevBuf<cplx<double>> chol, chol_work;   // Cholesky factor of the scaled S + shift and its panel workspace
    DevBuf<double> chol_dinv;               // the diagonal scaling of that factor
    // structured O of the last evaluation (dense nets)
    int precision = 0, words = 0, N = 0, M = 0, Mpad = 0, C = 0, layout = 0, npad = 0;
    int64_t count = 0;
    DevBuf<uint32_t> conn;
    DevBuf<uint64_t> confs;
    DevBuf<double> w;
    DevBuf<cplx<double>> e_loc;   // per-sample local energies of the last evaluation
    DevBuf<cplx<float>> g32;
    DevBuf<cplx<double>> g64;
    DevBuf<double> diag;   // diagonal of S (Jacobi preconditioner)
    // S.v workspace (grow-only)
    DevBuf<cplx<double>> va, vb, u, c1, Yw, Yb, Ya;
    DevBuf<cplx<float>> V32, SV32, zero32;
    DevBuf<cplx<double>> V64, SV64, zero64;
    DevBuf<uint16_t> frag;       // bf16x3 fragments of v (fp32 mode)
    DevBuf<cplx<float>> upart;   // per column pair partial O.v (fp32 mode)
    DevBuf<cplx<float>> u32;     // w (O.v - <O>.v) per sample (fp32 mode)
    DevBuf<float> spin_part;     // spin GEMM partials (fp32 mode)
    DevBuf<float> ss_partial;    // workspace of the TDVP statistics' spin GEMM (fp32 mode)
    DevBuf<double> ss_full;
    DevBuf<cplx<float>> ss_c32;
    DevBuf<double> opart, ypart; // dot-product partials of the fused matvec
    DevBuf<uint64_t> spin_bits;  // transposed spin bits of the samples (S.v in fp32 mode, K-FAC)
    bool w_uniform = false;      // the samples' weights are all equal (Monte Carlo)
    bool bits_valid = false;
    int64_t matvecs = 0;         // S.v products so far (every 16th is timed when profiling)
    PinnedBuf<double> rr_host;   // CG residual partials and ||F||^2 (page-locked)
    PinnedBuf<cplx<double>> x_host;   // the solution on its way to the caller
    PinnedBuf<int> flag_host;    // two slots of the CG convergence flag, read behind events
    int* flag_dev = nullptr;     // device address of flag_host (written by the check kernel)
    bool inv_unique = true;      // PsiDeep: every (site, unit) carries at most one connection
    int conn_shift = -1;         // PsiDeep: connection i of unit j is site (j shift + i) mod N (else -1)
    // K-FAC preconditioner (kfac.hip)
    DevBuf<double> kf_spart, kf_C, kf_Ci, kf_gam;
    bool kfac_hint = false;      // the last solve used K-FAC: the next evaluation computes the factor moments
    bool kf_ready = false;       // kf_C / kf_D hold the moments of the current samples
    double kfac_shift = 0.0;     // shift of the last K-FAC solve
    bool kf_inverted = false;    // kf_Ci / kf_D hold the inverses for the current samples and kf_inv_shift
    double kf_inv_shift = 0.0;
    OwnedEvent fork_ev, join_ev, join2_ev;   // the evaluation's side-stream fork and joins
    // E, E2, wsum (sums_host) and O, F (from Obar, Fd) of the last evaluation are read
    // to the host on first use
    bool sums_pending = false, of_pending = false;
    void settle();
    void fetch_host_vectors();
    std::vector<uint32_t> conn_host;   // connections of the uploaded tables (PsiDeep)
    int conn_mpad = 0, conn_n = 0;
    DevBuf<cplx<double>> kf_gpart, kf_D;
    WParamMap wmap() const { return {layout == LAYOUT_RBM ? 1 : 0, N, M, C, conn_shift}; }
    // while a CG solve runs with its vectors in site-major order (circulant PsiDeep,
    // C = N, fp32 S.v), the S.v products read and write that order and <O> in it
    bool solve_sm = false;
    DevBuf<cplx<double>> obar_sm;
    WParamMap order() const {
        WParamMap m = wmap();
        m.sm = solve_sm ? 1 : 0;
        return m;
    }
    DevBuf<int> kf_idx;
    DevBuf<int> inv;            // inverse gather table [N][Mpad] (PsiDeep)
    DevBuf<double> dot_scratch;
    DevBuf<double> sums;           // E.re, E.im, E2, sum w of the last evaluation
    PinnedBuf<double> sums_host;
    DevBuf<cplx<double>> coef;     // per-sample weights of the statistics
    // materialised O [count][P] of wave functions without a structured O (PsiCNN)
    bool generic = false;
    DevBuf<cplx<float>> Og32;
    DevBuf<cplx<double>> Og64;
    DevBuf<cplx<double>> part;
    // CG state of tdvp_solve (grow-only) and its two flag events
    DevBuf<cplx<double>> cg_x, cg_r, cg_r2, cg_z, cg_p, cg_p2, cg_Ap;
    DevBuf<double> cg_part, cg_apart, cg_scratch, cg_fnorm2;
    DevBuf<int> cg_done;
    OwnedEvent cg_ev[2];
};

// core operations
void sample(Psi& psi, Ensemble& ens);
void compute_theta(Psi& psi, const uint64_t* confs, int64_t count, Samples& out, int64_t offset);
void log_psi_of(Psi& psi, const uint64_t* confs, int64_t count, Samples& s, double* w);
// size the buffers of `count` samples of psi in a basis (Basis); throws when psi cannot
// read that basis
void prepare_samples(Psi& psi, Samples& s, int64_t count, int basis = BASIS_SPINS);
// sum_s coef_s op(O_s) in parameter order (op: 0 identity, 1 conjugate), any wave function
void ok_weighted_sum(Psi& psi, Samples& s, const cplx<double>* coef, int conj, cplx<double>* out);
void local_energies(Psi& psi, Operator& op, Samples& s);
// ExpectationValue::exp_sigma_z (network_functions/ExpectationValue.cu.template:52-77): the
// weighted sum over the samples of psi of exp(Operator::fast_local_energy(s))
cd expectation_exp_fast(Operator& op, Psi& psi, Ensemble& ens);
// ExpectationValue(op, psi, psi_sampling, ensemble) (ExpectationValue.cu.template:130-170):
// <op> in psi from the samples of psi_sampling weighted by |psi / psi_sampling|^2
cd expectation_reweighted(Operator& op, Psi& psi, Psi& psi_sampling, Ensemble& ens);
void tdvp_eval(Tdvp& t, Operator& op, Psi& psi, Ensemble& ens, bool with_S);
// y = (S + shift) v on the device; when the matvec also produced the block partials
// of v^H y, *ypart / *yparts describe them ([yparts][2]), else they are null / 0
void tdvp_S_dot(Tdvp& t, const cplx<double>* v_dev, double shift, cplx<double>* y_dev,
                const double** ypart = nullptr, int* yparts = nullptr);
// method 0 Cholesky, 1 Jacobi-preconditioned CG, 3 K-FAC-preconditioned CG (dense
// nets), the CG methods from x0 (zero when x0 is null)
int tdvp_solve(Tdvp& t, double shift, int method, double rtol, int max_iter, cd* x, int* iters, double* res,
               const cd* x0 = nullptr);
// O_k of every sample of the last evaluation, uncentred, [count][P] in parameter order (host)
void tdvp_O_samples(Tdvp& t, cd* out);

// HilbertSpaceDistance.distance / gradient (HilbertSpaceDistance.cu.template:113-165):
// the distance between psi' and U psi on psi's samples; with grad, also its gradient
// with respect to psi''s parameters (nu: the power of the distance it is divided by)
// KullbackLeibler (include/network_functions/KullbackLeibler.hpp): psi a classical network,
// psi_prime the network whose ensemble is sampled; the mean deviation of one call shifts
// the deviations of the next (KullbackLeibler.cu.template:191-197)
struct Kl {
    int P = 0;
    double log_psi_scale = 1.0;
    cd last_mean_deviation = 0.0;
    double total_weight = 0.0;
    cd mean_deviation = 0.0;
};
// mode 0: the value; 1: and the gradient (P values); 2: and the noise estimate
double kullback_leibler(Kl& kl, Psi& psi, Psi& psi_prime, Ensemble& ens, double threshold, int mode, double nu,
                        cd* grad, double* noise);
double hilbert_space_distance(Psi& psi, Psi& psi_prime, Operator& op, bool unitary, Ensemble& ens, double nu,
                              cd* grad);

// proposal training: Adam steps on random (sample, block) rows of the given
// device configurations; returns the loss of the last step
double proposal_train(Proposal& p, const uint64_t* confs, int64_t count, int words, int steps, int batch, double lr);
// loss and gradient on explicit rows (host configurations and block sites)
double proposal_loss_grad(Proposal& p, const uint64_t* confs, int64_t count, int words, const int* sites,
                          std::vector<float>& grad);
void proposal_adam_step(Proposal& p, const float* grad, double lr);

}  // namespace ann
