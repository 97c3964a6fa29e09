// capi.cpp -- the extern "C" boundary declared in include/annongpu.h.
//
// Every function catches C++ exceptions and turns them into a status code plus a
// thread-local message (the reference throws std::runtime_error from CUDA_CHECK,
// include/types.h:139-151, surfaced by pybind11 as RuntimeError).
#include "annongpu.h"
#include "runtime.hpp"

#include <cmath>
#include <random>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

struct ann_psi_s {
    ann::Psi impl;
};
struct ann_operator_s {
    ann::Operator impl;
};
struct ann_ensemble_s {
    ann::Ensemble impl;
};
struct ann_tdvp_s {
    ann::Tdvp impl;
};
struct ann_kl_s {
    ann::Kl impl;
};
struct ann_proposal_s {
    ann::Proposal impl;
};

namespace {

thread_local std::string g_error;

template <typename F>
int guard(F&& f) {
    try {
        f();
        return ANN_OK;
    } catch (const ann::HipError& e) {
        g_error = e.what();
        return ANN_E_DEVICE;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return ANN_E_INVALID;
    } catch (const std::logic_error& e) {
        g_error = e.what();
        return ANN_E_UNSUPPORTED;
    } catch (const std::exception& e) {
        g_error = e.what();
        return ANN_E_DEVICE;
    }
}

using ann::cd;

inline const cd* C(const double* p) { return reinterpret_cast<const cd*>(p); }
inline cd* C(double* p) { return reinterpret_cast<cd*>(p); }

void need(const void* p, const char* what) {
    if (!p) throw std::invalid_argument(std::string("null pointer: ") + what);
}

ann::Samples& single_sample(ann::Psi& psi, const uint64_t* conf, ann::Samples& s, int basis = ann::BASIS_SPINS) {
    psi.sync_device();
    ann::prepare_samples(psi, s, 1, basis);
    if (basis == ann::BASIS_PAULIS) {
        const uint64_t m = psi.num_sites >= 64 ? ~uint64_t(0) : (uint64_t(1) << psi.num_sites) - 1;
        if ((conf[0] | conf[1]) & ~m) throw std::invalid_argument("a Pauli string has sites beyond num_sites");
    }
    s.confs.upload(conf, s.words);
    ann::log_psi_of(psi, s.confs.p, 1, s, nullptr);
    return s;
}

std::vector<cd> host_log_psi(ann::Samples& s) {
    std::vector<cd> lp(s.count);
    ann::Runtime::get().sync();
    s.log_psi.download(reinterpret_cast<ann::cplx<double>*>(lp.data()), s.count);
    return lp;
}

std::vector<double> host_weights(ann::Samples& s) {
    std::vector<double> w(s.count);
    s.w.download(w.data(), s.count);
    return w;
}

void structured_sum(ann::Psi& psi, ann::Samples& s, const ann::cplx<double>* coef, int conj_g, cd* host_out) {
    ann::DevBuf<ann::cplx<double>> out;
    out.ensure(psi.num_params);
    ann::ok_weighted_sum(psi, s, coef, conj_g, out.p);
    out.download(reinterpret_cast<ann::cplx<double>*>(host_out), psi.num_params);
}

}  // namespace

extern "C" {

const char* ann_last_error(void) { return g_error.c_str(); }

int ann_set_device(int device) {
    return guard([&] { ann::Runtime::get().set_device(device); });
}

int ann_profiler_start(void) {
    return guard([&] { ann::Runtime::get().profiling = true; });
}
int ann_profiler_stop(void) {
    return guard([&] {
        ann::Runtime::get().sync();
        ann::Runtime::get().profiling = false;
    });
}
int ann_profiler_query(const char* name, double* ms, int64_t* launches) {
    return guard([&] {
        need(name, "name");
        ann::Runtime::get().sync();
        auto& t = ann::Runtime::get().totals;
        auto& u = ann::Runtime::get().untimed;
        auto it = t.find(name);
        auto iu = u.find(name);
        const double timed_ms = it == t.end() ? 0.0 : it->second.first;
        const int64_t timed = it == t.end() ? 0 : it->second.second;
        const int64_t all = timed + (iu == u.end() ? 0 : iu->second);
        // classes timed on a sample of their launches are scaled to all launches
        if (ms) *ms = timed > 0 ? timed_ms * double(all) / double(timed) : 0.0;
        if (launches) *launches = all;
    });
}
int ann_profiler_reset(void) {
    return guard([&] {
        ann::Runtime::get().sync();
        ann::Runtime::get().totals.clear();
        ann::Runtime::get().untimed.clear();
    });
}

// ------------------------------------------------------------- wave functions

int ann_psi_create_rbm(ann_psi* out, int32_t N, int32_t M, const double* W, const double* final_weight,
                       const double* log_prefactor) {
    return guard([&] {
        need(out, "out");
        need(W, "W");
        if (N <= 0 || M <= 0) throw std::invalid_argument("N and M must be positive");
        auto p = new ann_psi_s;
        try {
            auto& q = p->impl;
            q.kind = ann::PSI_RBM;
            q.num_sites = N;
            q.N = N;
            q.M = M;
            q.W.assign(C(W), C(W) + size_t(N) * M);
            q.f = {final_weight ? *C(final_weight) : cd(1.0)};
            q.log_prefactor = log_prefactor ? *C(log_prefactor) : cd(0.0);
            q.finalize();
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}

int ann_psi_create_deep(ann_psi* out, int32_t num_sites, int32_t N, const double* a, int32_t num_layers,
                        const int32_t* sizes, const int32_t* conns, const uint32_t* const* connections,
                        const double* const* weights, const double* const* biases, const double* final_weights,
                        const double* log_prefactor) {
    return guard([&] {
        need(out, "out");
        need(a, "a");
        need(sizes, "sizes");
        need(conns, "conns");
        if (num_layers < 1 || num_layers > 4) throw std::invalid_argument("PsiDeep supports 1..4 hidden layers");
        if (N != num_sites && !(N == 3 * num_sites && num_sites <= 64))
            throw std::invalid_argument("N must be num_sites (spins) or 3 num_sites (Pauli strings, <= 64 sites)");
        auto p = new ann_psi_s;
        try {
            auto& q = p->impl;
            q.kind = ann::PSI_DEEP;
            q.num_sites = num_sites;
            q.N = N;
            q.a.assign(C(a), C(a) + N);
            q.num_layers = num_layers;
            int prev = N;
            for (int l = 0; l < num_layers; l++) {
                if (sizes[l] <= 0 || conns[l] <= 0) throw std::invalid_argument("empty layer");
                need(connections[l], "connections");
                need(weights[l], "weights");
                need(biases[l], "biases");
                q.sizes.push_back(sizes[l]);
                q.conns.push_back(conns[l]);
                const size_t n = size_t(sizes[l]) * conns[l];
                q.conn_l.emplace_back(connections[l], connections[l] + n);
                for (uint32_t c : q.conn_l.back())
                    if (int(c) >= prev) throw std::invalid_argument("connection index out of range");
                q.w_l.emplace_back(C(weights[l]), C(weights[l]) + n);
                q.b_l.emplace_back(C(biases[l]), C(biases[l]) + sizes[l]);
                prev = sizes[l];
            }
            need(final_weights, "final_weights");
            q.f.assign(C(final_weights), C(final_weights) + sizes[num_layers - 1]);
            q.log_prefactor = log_prefactor ? *C(log_prefactor) : cd(0.0);
            q.finalize();
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}

int ann_psi_create_cnn(ann_psi* out, const int32_t* extent, int32_t num_layers, const int32_t* channels,
                       const int32_t* connectivity, const uint32_t* symmetry_classes, int32_t num_params,
                       const double* params, double final_factor, const double* log_prefactor) {
    return guard([&] {
        need(out, "out");
        need(extent, "extent");
        need(channels, "channels");
        need(connectivity, "connectivity");
        need(symmetry_classes, "symmetry_classes");
        need(params, "params");
        if (num_layers < 1 || num_layers > 3) throw std::invalid_argument("PsiCNN supports 1..3 layers");
        auto p = new ann_psi_s;
        try {
            auto& q = p->impl;
            q.kind = ann::PSI_CNN;
            int n = 1;
            for (int d = 0; d < 3; d++) {
                if (extent[d] <= 0) throw std::invalid_argument("extent must be positive");
                q.extent[d] = extent[d];
                n *= extent[d];
            }
            q.num_sites = n;
            q.N = n;
            q.cnn_layers = num_layers;
            q.channels.assign(channels, channels + num_layers);
            q.connectivity.assign(connectivity, connectivity + 3 * num_layers);
            for (int l = 0; l < num_layers; l++)
                for (int d = 0; d < 3; d++)
                    if (connectivity[3 * l + d] <= 0 || connectivity[3 * l + d] > extent[d])
                        throw std::invalid_argument("connectivity must lie in 1..extent");
            q.sym.assign(symmetry_classes, symmetry_classes + n);
            uint32_t mx = 0;
            for (uint32_t c : q.sym) mx = std::max(mx, c);
            q.num_sym = int(mx) + 1;
            int expect = 0, prev = 1;
            for (int l = 0; l < num_layers; l++) {
                expect += prev * channels[l] * q.num_sym * connectivity[3 * l] * connectivity[3 * l + 1] *
                          connectivity[3 * l + 2];
                prev = channels[l];
            }
            if (expect != num_params) throw std::invalid_argument("PsiCNN parameter count mismatch");
            q.params.assign(C(params), C(params) + num_params);
            q.final_factor = final_factor;
            q.log_prefactor = log_prefactor ? *C(log_prefactor) : cd(0.0);
            q.finalize();
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}

int ann_psi_create_classical(ann_psi* out, int32_t num_sites, int32_t num_ops, const ann_operator* ops,
                             const double* params, ann_psi ref, int32_t order, const double* log_prefactor,
                             const double* ref_log_prefactor) {
    return guard([&] {
        need(out, "out");
        if (num_sites <= 0 || num_sites > 64 * ann::kMaxWords) throw std::invalid_argument("num_sites must be 1..256");
        if (num_ops < 0) throw std::invalid_argument("num_ops must not be negative");
        if (num_ops > 0) {
            need(ops, "ops");
            need(params, "params");
        }
        if (order != 1 && order != 2) throw std::invalid_argument("order must be 1 or 2");
        if (ref) {
            if (ref->impl.kind != ann::PSI_DEEP && ref->impl.kind != ann::PSI_CNN)
                throw std::invalid_argument("the network reference must be a PsiCNN or a PsiDeep");
            if (ref->impl.num_sites != num_sites) throw std::invalid_argument("psi_ref has another num_sites");
        }
        auto p = new ann_psi_s;
        try {
            auto& q = p->impl;
            q.kind = ann::PSI_CLASSICAL;
            q.num_sites = num_sites;
            q.N = num_sites;
            q.cl_order = order;
            q.cl_begin.push_back(0);
            for (int n = 0; n < num_ops; n++) {
                need(ops[n], "ops[n]");
                const ann::Operator& o = ops[n]->impl;
                if (o.max_site() >= num_sites) throw std::invalid_argument("H_local acts on sites beyond num_sites");
                for (int m = 0; m < o.num_strings; m++) {
                    for (int w = 0; w < ann::kMaxWords; w++) {
                        q.cl_a.push_back(o.a[size_t(m) * ann::kMaxWords + w]);
                        q.cl_b.push_back(o.b[size_t(m) * ann::kMaxWords + w]);
                    }
                    q.cl_coef.push_back(o.coef[m]);
                }
                q.cl_begin.push_back(int(q.cl_coef.size()));
            }
            if (num_ops > 0) q.params.assign(C(params), C(params) + num_ops);
            if (ref) q.ref = ref->impl.clone();
            q.log_prefactor = log_prefactor ? *C(log_prefactor) : cd(0.0);
            if (ref_log_prefactor) {
                q.ref_log_prefactor = *C(ref_log_prefactor);
                if (q.ref) q.ref->log_prefactor = q.ref_log_prefactor;
            }
            q.finalize();
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}

int ann_psi_classical_info(ann_psi psi, int32_t* order, int32_t* num_ops, int32_t* has_network_ref) {
    return guard([&] {
        need(psi, "psi");
        const auto& q = psi->impl;
        if (q.kind != ann::PSI_CLASSICAL) throw std::invalid_argument("not a PsiClassical");
        if (order) *order = q.cl_order;
        if (num_ops) *num_ops = int32_t(q.params.size());
        if (has_network_ref) *has_network_ref = q.ref ? 1 : 0;
    });
}

int ann_psi_classical_ref(ann_psi psi, ann_psi* out) {
    return guard([&] {
        need(psi, "psi");
        need(out, "out");
        const auto& q = psi->impl;
        if (q.kind != ann::PSI_CLASSICAL || !q.ref) throw std::invalid_argument("no network reference");
        auto c = q.ref->clone();
        auto p = new ann_psi_s;
        p->impl = std::move(*c);
        *out = p;
    });
}

int ann_psi_get_ref_log_prefactor(ann_psi psi, double* value2) {
    return guard([&] {
        need(psi, "psi");
        need(value2, "value");
        const auto& q = psi->impl;
        if (q.kind != ann::PSI_CLASSICAL) throw std::invalid_argument("not a PsiClassical");
        *C(value2) = q.ref ? q.ref->log_prefactor : q.ref_log_prefactor;
    });
}

int ann_psi_set_ref_log_prefactor(ann_psi psi, const double* value2) {
    return guard([&] {
        need(psi, "psi");
        need(value2, "value");
        auto& q = psi->impl;
        if (q.kind != ann::PSI_CLASSICAL) throw std::invalid_argument("not a PsiClassical");
        q.ref_log_prefactor = *C(value2);
        if (q.ref) q.ref->log_prefactor = q.ref_log_prefactor;
    });
}

int ann_psi_copy(ann_psi src, ann_psi* out) {
    return guard([&] {
        need(src, "src");
        need(out, "out");
        auto q = src->impl.clone();
        auto p = new ann_psi_s;
        p->impl = std::move(*q);
        *out = p;
    });
}

int ann_psi_destroy(ann_psi psi) {
    return guard([&] { delete psi; });
}

int ann_psi_info(ann_psi psi, int32_t* num_params, int32_t* N, int32_t* num_sites) {
    return guard([&] {
        need(psi, "psi");
        if (num_params) *num_params = psi->impl.num_params;
        if (N) *N = psi->impl.N;
        if (num_sites) *num_sites = psi->impl.num_sites;
    });
}

int ann_psi_get_params(ann_psi psi, double* params) {
    return guard([&] {
        need(psi, "psi");
        need(params, "params");
        auto p = psi->impl.get_params();
        std::memcpy(params, p.data(), sizeof(cd) * p.size());
    });
}

int ann_psi_set_params(ann_psi psi, const double* params) {
    return guard([&] {
        need(psi, "psi");
        need(params, "params");
        psi->impl.set_params(C(params));
    });
}

int ann_psi_get_log_prefactor(ann_psi psi, double* value2) {
    return guard([&] {
        need(psi, "psi");
        *C(value2) = psi->impl.log_prefactor;
    });
}

int ann_psi_set_log_prefactor(ann_psi psi, const double* value2) {
    return guard([&] {
        need(psi, "psi");
        psi->impl.log_prefactor = *C(value2);
    });
}

int ann_psi_set_precision(ann_psi psi, int32_t precision) {
    return guard([&] {
        need(psi, "psi");
        if (precision != ANN_FP64 && precision != ANN_FP32) throw std::invalid_argument("unknown precision");
        psi->impl.precision = precision;
    });
}

// ------------------------------------------------------------------ operator

int ann_operator_create(ann_operator* out, int32_t num_strings, int32_t words, const uint64_t* a, const uint64_t* b,
                        const double* coefficients) {
    return guard([&] {
        need(out, "out");
        if (num_strings < 0 || words < 1 || words > ann::kMaxWords)
            throw std::invalid_argument("bad operator dimensions");
        if (num_strings > 0) {
            need(a, "a");
            need(b, "b");
            need(coefficients, "coefficients");
        }
        auto p = new ann_operator_s;
        auto& o = p->impl;
        o.num_strings = num_strings;
        o.a.assign(size_t(num_strings) * ann::kMaxWords, 0);
        o.b.assign(size_t(num_strings) * ann::kMaxWords, 0);
        for (int n = 0; n < num_strings; n++)
            for (int w = 0; w < words; w++) {
                o.a[size_t(n) * ann::kMaxWords + w] = a[size_t(n) * words + w];
                o.b[size_t(n) * ann::kMaxWords + w] = b[size_t(n) * words + w];
            }
        o.coef.assign(C(coefficients), C(coefficients) + num_strings);
        *out = p;
    });
}

int ann_operator_destroy(ann_operator op) {
    return guard([&] { delete op; });
}

// ----------------------------------------------------------------- ensembles

int ann_monte_carlo_create(ann_ensemble* out, uint32_t num_samples, uint32_t num_sweeps,
                           uint32_t num_thermalization_sweeps, uint32_t num_markov_chains, int32_t move,
                           int32_t total_z, uint64_t seed) {
    return guard([&] {
        need(out, "out");
        if (num_markov_chains == 0) throw std::invalid_argument("num_markov_chains must be positive");
        if (move != ANN_MOVE_FLIP && move != ANN_MOVE_EXCHANGE) throw std::invalid_argument("unknown move");
        auto p = new ann_ensemble_s;
        auto& e = p->impl;
        e.kind = ann::ENS_MC;
        e.num_samples = num_samples;
        e.num_sweeps = num_sweeps;
        e.num_therm = num_thermalization_sweeps;
        e.num_chains = num_markov_chains;
        e.move = move;
        e.total_z = total_z;
        e.seed = seed;
        *out = p;
    });
}

int ann_monte_carlo_set_persistent(ann_ensemble ens, int32_t persistent) {
    return guard([&] {
        need(ens, "ens");
        ens->impl.persistent = persistent != 0;
        if (!persistent) ens->impl.chains_valid = false;
    });
}

int ann_exact_summation_create(ann_ensemble* out, uint32_t num_sites) {
    return guard([&] {
        need(out, "out");
        if (num_sites == 0 || num_sites > 26) throw std::invalid_argument("exact summation needs 1..26 sites");
        auto p = new ann_ensemble_s;
        p->impl.kind = ann::ENS_EXACT;
        p->impl.num_sites = num_sites;
        *out = p;
    });
}

int ann_sample_list_create(ann_ensemble* out, uint32_t num_sites, const uint64_t* confs, int64_t count,
                           int32_t words, const double* weights) {
    return guard([&] {
        need(out, "out");
        need(confs, "confs");
        if (num_sites == 0 || num_sites > 64 * ann::kMaxWords) throw std::invalid_argument("num_sites must be 1..256");
        if (count <= 0 || count > int64_t(UINT32_MAX)) throw std::invalid_argument("count must be positive");
        if (words < int32_t((num_sites + 63) / 64) || words > ann::kMaxWords)
            throw std::invalid_argument("words does not fit num_sites");
        std::vector<uint64_t> list(size_t(count) * ann::kMaxWords, 0);
        for (int64_t i = 0; i < count; i++)
            for (int w = 0; w < words; w++) {
                const uint64_t x = confs[size_t(i) * words + w];
                const int lo = 64 * w;
                const uint64_t allowed = int(num_sites) >= lo + 64 ? ~uint64_t(0)
                                         : int(num_sites) <= lo ? 0 : (uint64_t(1) << (num_sites - lo)) - 1;
                if (x & ~allowed) throw std::invalid_argument("a configuration has bits beyond num_sites");
                list[size_t(i) * ann::kMaxWords + w] = x;
            }
        std::vector<double> wt(size_t(count), 1.0 / double(count));
        if (weights)
            for (int64_t i = 0; i < count; i++) {
                if (!std::isfinite(weights[i]) || weights[i] < 0.0) throw std::invalid_argument("weights must be >= 0");
                wt[size_t(i)] = weights[i];
            }
        auto p = new ann_ensemble_s;
        p->impl.kind = ann::ENS_LIST;
        p->impl.num_sites = num_sites;
        p->impl.num_samples = uint32_t(count);
        p->impl.list_confs = std::move(list);
        p->impl.list_w = std::move(wt);
        *out = p;
    });
}

int ann_exact_summation_paulis_create(ann_ensemble* out, uint32_t num_sites) {
    return guard([&] {
        need(out, "out");
        if (num_sites == 0 || num_sites > 13)
            throw std::invalid_argument("exact summation over Pauli strings needs 1..13 sites");
        auto p = new ann_ensemble_s;
        p->impl.kind = ann::ENS_EXACT;
        p->impl.basis = ann::BASIS_PAULIS;
        p->impl.num_sites = num_sites;
        *out = p;
    });
}

int ann_monte_carlo_paulis_create(ann_ensemble* out, uint32_t num_samples, uint32_t num_sweeps,
                                  uint32_t num_thermalization_sweeps, uint32_t num_markov_chains, uint64_t seed) {
    return guard([&] {
        need(out, "out");
        if (num_markov_chains == 0) throw std::invalid_argument("num_markov_chains must be positive");
        auto p = new ann_ensemble_s;
        auto& e = p->impl;
        e.kind = ann::ENS_MC;
        e.basis = ann::BASIS_PAULIS;
        e.num_samples = num_samples;
        e.num_sweeps = num_sweeps;
        e.num_therm = num_thermalization_sweeps;
        e.num_chains = num_markov_chains;
        e.move = ann::MOVE_PAULI;
        e.seed = seed;
        *out = p;
    });
}

int ann_pauli_sample_list_create(ann_ensemble* out, uint32_t num_sites, const uint64_t* ab, int64_t count,
                                 const double* weights) {
    return guard([&] {
        need(out, "out");
        need(ab, "ab");
        if (num_sites == 0 || num_sites > 64) throw std::invalid_argument("num_sites must be 1..64");
        if (count <= 0 || count > int64_t(UINT32_MAX)) throw std::invalid_argument("count must be positive");
        const uint64_t m = num_sites >= 64 ? ~uint64_t(0) : (uint64_t(1) << num_sites) - 1;
        std::vector<uint64_t> list(size_t(count) * ann::kMaxWords, 0);
        for (int64_t i = 0; i < count; i++) {
            const uint64_t a = ab[2 * i], b = ab[2 * i + 1];
            if ((a | b) & ~m) throw std::invalid_argument("a Pauli string has sites beyond num_sites");
            list[size_t(i) * ann::kMaxWords] = a;
            list[size_t(i) * ann::kMaxWords + 1] = b;
        }
        std::vector<double> wt(size_t(count), 1.0 / double(count));
        if (weights)
            for (int64_t i = 0; i < count; i++) {
                if (!std::isfinite(weights[i]) || weights[i] < 0.0) throw std::invalid_argument("weights must be >= 0");
                wt[size_t(i)] = weights[i];
            }
        auto p = new ann_ensemble_s;
        p->impl.kind = ann::ENS_LIST;
        p->impl.basis = ann::BASIS_PAULIS;
        p->impl.num_sites = num_sites;
        p->impl.num_samples = uint32_t(count);
        p->impl.list_confs = std::move(list);
        p->impl.list_w = std::move(wt);
        *out = p;
    });
}

int ann_ensemble_destroy(ann_ensemble ens) {
    return guard([&] { delete ens; });
}

int ann_ensemble_num_steps(ann_ensemble ens, uint64_t* n) {
    return guard([&] {
        need(ens, "ens");
        *n = ens->impl.num_steps();
    });
}

int ann_ensemble_acceptance(ann_ensemble ens, uint64_t* accepted, uint64_t* rejected) {
    return guard([&] {
        need(ens, "ens");
        ens->impl.settle_acceptance();
        if (accepted) *accepted = ens->impl.accepted;
        if (rejected) *rejected = ens->impl.rejected;
    });
}

int ann_ensemble_block_stats(ann_ensemble ens, uint64_t* proposals, uint64_t* null_moves, uint64_t* rows_evaluated) {
    return guard([&] {
        need(ens, "ens");
        ens->impl.settle_acceptance();
        if (proposals) *proposals = ens->impl.block_proposals;
        if (null_moves) *null_moves = ens->impl.null_moves;
        if (rows_evaluated) *rows_evaluated = ens->impl.rows_evaluated;
    });
}

int ann_ensemble_last_samples(ann_ensemble ens, uint64_t* confs, int64_t capacity, int64_t* count, int32_t* words) {
    return guard([&] {
        need(ens, "ens");
        auto& s = ens->impl.samples;
        if (count) *count = s.count;
        if (words) *words = s.words;
        if (confs) {
            if (capacity < s.count * s.words) throw std::invalid_argument("buffer too small");
            s.confs.download(confs, size_t(s.count) * s.words);
        }
    });
}

// --------------------------------------------------------- network functions

int ann_log_psi_s(ann_psi psi, const uint64_t* conf, double* out2) {
    return guard([&] {
        need(psi, "psi");
        need(conf, "conf");
        ann::Samples s;
        single_sample(psi->impl, conf, s);
        auto lp = host_log_psi(s);
        *C(out2) = lp[0];
    });
}

int ann_psi_O_k(ann_psi psi, const uint64_t* conf, double* out) {
    return guard([&] {
        need(psi, "psi");
        need(conf, "conf");
        auto& q = psi->impl;
        ann::Samples s;
        single_sample(q, conf, s);
        ann::DevBuf<ann::cplx<double>> coef;
        const ann::cplx<double> one = {1.0, 0.0};
        coef.upload(&one, 1);
        structured_sum(q, s, coef.p, 0, C(out));
    });
}

int ann_log_psi_s_pauli(ann_psi psi, const uint64_t* ab, double* out2) {
    return guard([&] {
        need(psi, "psi");
        need(ab, "ab");
        ann::Samples s;
        single_sample(psi->impl, ab, s, ann::BASIS_PAULIS);
        auto lp = host_log_psi(s);
        *C(out2) = lp[0];
    });
}

int ann_psi_O_k_pauli(ann_psi psi, const uint64_t* ab, double* out) {
    return guard([&] {
        need(psi, "psi");
        need(ab, "ab");
        auto& q = psi->impl;
        ann::Samples s;
        single_sample(q, ab, s, ann::BASIS_PAULIS);
        ann::DevBuf<ann::cplx<double>> coef;
        const ann::cplx<double> one = {1.0, 0.0};
        coef.upload(&one, 1);
        structured_sum(q, s, coef.p, 0, C(out));
    });
}

int ann_psi_O_k_vector(ann_psi psi, ann_ensemble ens, double* out) {
    return guard([&] {
        need(psi, "psi");
        need(ens, "ens");
        auto& q = psi->impl;
        ann::sample(q, ens->impl);
        auto& s = ens->impl.samples;
        ann::DevBuf<ann::cplx<double>> coef;
        std::vector<ann::cplx<double>> ones(s.count, {1.0, 0.0});
        coef.upload(ones.data(), ones.size());
        structured_sum(q, s, coef.p, 0, C(out));
    });
}

int ann_log_psi(ann_psi psi, ann_ensemble ens, double* out2) {
    return guard([&] {
        need(psi, "psi");
        need(ens, "ens");
        ann::sample(psi->impl, ens->impl);
        auto& s = ens->impl.samples;
        auto lp = host_log_psi(s);
        auto w = host_weights(s);
        cd acc = 0.0;
        for (int64_t i = 0; i < s.count; i++) acc += w[i] * lp[i];
        *C(out2) = acc;
    });
}

int ann_psi_vector(ann_psi psi, ann_ensemble ens, double* out) {
    return guard([&] {
        need(psi, "psi");
        need(ens, "ens");
        ann::sample(psi->impl, ens->impl);
        auto lp = host_log_psi(ens->impl.samples);
        for (size_t i = 0; i < lp.size(); i++) C(out)[i] = std::exp(lp[i]);
    });
}

int ann_log_psi_vector(ann_psi psi, ann_ensemble ens, double* out) {
    return guard([&] {
        need(psi, "psi");
        need(ens, "ens");
        ann::sample(psi->impl, ens->impl);
        auto lp = host_log_psi(ens->impl.samples);
        std::memcpy(out, lp.data(), sizeof(cd) * lp.size());
    });
}

int ann_psi_norm(ann_psi psi, ann_ensemble ens, double* out) {
    return guard([&] {
        need(psi, "psi");
        need(ens, "ens");
        if (ens->impl.kind != ann::ENS_EXACT) throw std::invalid_argument("psi_norm needs an ExactSummation");
        ann::sample(psi->impl, ens->impl);
        auto lp = host_log_psi(ens->impl.samples);
        double acc = 0.0;
        for (const cd& z : lp) acc += std::exp(2.0 * z.real());
        *out = std::sqrt(acc);
    });
}

int ann_apply_operator(ann_psi psi, ann_operator op, ann_ensemble ens, double* out) {
    return guard([&] {
        need(psi, "psi");
        need(op, "op");
        need(ens, "ens");
        ann::sample(psi->impl, ens->impl);
        auto& s = ens->impl.samples;
        ann::local_energies(psi->impl, op->impl, s);
        auto lp = host_log_psi(s);
        std::vector<cd> e(s.count);
        s.e_loc.download(reinterpret_cast<ann::cplx<double>*>(e.data()), s.count);
        for (int64_t i = 0; i < s.count; i++) C(out)[i] = std::exp(lp[i]) * e[i];
    });
}

int ann_activation_function(const double* x2, int32_t layer, double* out2) {
    return guard([&] {
        const ann::cplx<double> z = {x2[0], x2[1]};
        const ann::cplx<double> r = ann::act(z, layer);
        out2[0] = r.re;
        out2[1] = r.im;
    });
}

static void expectation_core(ann::Operator& op, ann::Psi& psi, ann::Ensemble& ens, double h[4]) {
    ann::sample(psi, ens);
    auto& s = ens.samples;
    ann::local_energies(psi, op, s);
    ann::DevBuf<double> sums;
    sums.ensure(4);
    ann::launch_weighted_sums(s.e_loc.p, s.w.p, s.count, sums.p, ann::Runtime::get().stream());
    ann::Runtime::get().sync();
    sums.download(h, 4);
}

int ann_expectation_value(ann_operator op, ann_psi psi, ann_ensemble ens, double* out2) {
    return guard([&] {
        need(op, "op");
        need(psi, "psi");
        need(ens, "ens");
        double h[4];
        expectation_core(op->impl, psi->impl, ens->impl, h);
        out2[0] = h[0];
        out2[1] = h[1];
    });
}

int ann_expectation_fluctuation(ann_operator op, ann_psi psi, ann_ensemble ens, double* sigma, double* value2) {
    return guard([&] {
        need(op, "op");
        need(psi, "psi");
        need(ens, "ens");
        double h[4];
        expectation_core(op->impl, psi->impl, ens->impl, h);
        // ExpectationValue.cu.template:409-412
        *sigma = std::sqrt(h[2] - (h[0] * h[0] + h[1] * h[1]));
        value2[0] = h[0];
        value2[1] = h[1];
    });
}

int ann_expectation_exp_sigma_z(ann_operator op, ann_psi psi, ann_ensemble ens, double* value2) {
    return guard([&] {
        need(op, "op");
        need(psi, "psi");
        need(ens, "ens");
        need(value2, "value");
        *C(value2) = ann::expectation_exp_fast(op->impl, psi->impl, ens->impl);
    });
}

int ann_expectation_value_reweighted(ann_operator op, ann_psi psi, ann_psi psi_sampling, ann_ensemble ens,
                                     double* value2) {
    return guard([&] {
        need(op, "op");
        need(psi, "psi");
        need(psi_sampling, "psi_sampling");
        need(ens, "ens");
        need(value2, "value");
        *C(value2) = ann::expectation_reweighted(op->impl, psi->impl, psi_sampling->impl, ens->impl);
    });
}

int ann_expectation_gradient(ann_operator op, ann_psi psi, ann_ensemble ens, double* gradient, double* value2) {
    return guard([&] {
        need(op, "op");
        need(psi, "psi");
        need(ens, "ens");
        need(gradient, "gradient");
        auto& q = psi->impl;
        double h[4];
        expectation_core(op->impl, q, ens->impl, h);
        auto& s = ens->impl.samples;
        const cd E(h[0], h[1]);
        ann::DevBuf<ann::cplx<double>> coef;
        coef.ensure(s.count);
        std::vector<cd> g1(q.num_params), g2(q.num_params);
        hipStream_t st = ann::Runtime::get().stream();
        ann::launch_weight_coef(s.w.p, s.e_loc.p, s.count, coef.p, st);
        structured_sum(q, s, coef.p, 1, g1.data());
        ann::launch_weight_coef(s.w.p, nullptr, s.count, coef.p, st);
        structured_sum(q, s, coef.p, 1, g2.data());
        // ExpectationValue.cu.template:467-469
        for (int k = 0; k < q.num_params; k++) C(gradient)[k] = g1[k] - E * g2[k];
        value2[0] = h[0];
        value2[1] = h[1];
    });
}

int ann_local_energies(ann_operator op, ann_psi psi, ann_ensemble ens, double* e_loc, int64_t capacity) {
    return guard([&] {
        need(op, "op");
        need(psi, "psi");
        need(ens, "ens");
        ann::sample(psi->impl, ens->impl);
        auto& s = ens->impl.samples;
        if (capacity < s.count) throw std::invalid_argument("buffer too small");
        ann::local_energies(psi->impl, op->impl, s);
        ann::Runtime::get().sync();
        s.e_loc.download(reinterpret_cast<ann::cplx<double>*>(e_loc), s.count);
    });
}

// ---------------------------------------------- local energies without allocation

namespace {

// Angles of the configurations of ann_local_energies_noalloc.  Each GPU the library
// runs on holds its own copy of this array until the process ends; host code reaches
// it only through the address hipGetSymbolAddress returns for the current device.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// the device address of `scratch` on the current device, looked up once per device
void* scratch_address() {
    int dev = -1;
    ANN_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw ann::HipError("device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw ann::HipError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

}  // namespace

int ann_local_energies_noalloc(ann_operator op, ann_psi psi, const uint64_t* confs, int64_t count, double* e_loc) {
    return guard([&] {
        need(op, "op");
        need(psi, "psi");
        need(confs, "confs");
        need(e_loc, "e_loc");
        auto& q = psi->impl;
        if (!q.dense()) throw std::logic_error("ann_local_energies_noalloc needs a PsiRBM or one-layer PsiDeep");
        if (count <= 0) throw std::invalid_argument("count must be positive");
        const size_t per = size_t(q.Mpad) * (q.precision == 0 ? sizeof(ann::cplx<double>) : sizeof(ann::cplx<float>));
        if (size_t(count) * per > kNoAllocScratchBytes)
            throw std::invalid_argument("ann_local_energies_noalloc: count * hidden units exceeds its " +
                                        std::to_string(kNoAllocScratchBytes >> 20) + " MiB scratch (at most " +
                                        std::to_string(kNoAllocScratchBytes / per) + " configurations here)");
        // the parameters and the operator tables must already be on the device: putting
        // them there allocates, which the ordinary entry points do on first use
        if (q.dirty || (q.params_dirty && q.params_d.n < size_t(q.num_params)))
            throw std::logic_error("the wave function's parameters are not on the device yet: make one ordinary "
                                   "call with it (e.g. ann_log_psi_s) first");
        if (!op->impl.compiled.count(q.words))
            throw std::logic_error("the operator's device tables are not built yet: make one ordinary local "
                                   "energy call with it first");
        if (op->impl.max_site() >= q.num_sites) throw std::invalid_argument("operator acts on sites beyond num_sites");
        void* th = scratch_address();
        const ann::DeviceOp& d = *op->impl.compiled.at(q.words);
        hipStream_t st = ann::lib_stream();
        auto* out = reinterpret_cast<ann::cplx<double>*>(e_loc);
        if (q.precision == 0) {
            auto v = q.view64();
            auto* t = static_cast<ann::cplx<double>*>(th);
            ann::launch_theta_valu<double>(v, confs, q.words, count, t, st);
            ann::launch_local_energy<double>(v, d.view(), confs, t, count, out, st);
        } else {
            auto v = q.view32();
            auto* t = static_cast<ann::cplx<float>*>(th);
            ann::launch_theta_mfma(v, q.wfrag.p, q.npad, confs, q.words, count, t, st);
            ann::launch_local_energy<float>(v, d.view(), confs, t, count, out, st);
        }
        ANN_HIP_CHECK(hipStreamSynchronize(st));
    });
}

int ann_hilbert_space_distance(ann_psi psi, ann_psi psi_prime, ann_operator op, int32_t is_unitary, ann_ensemble ens,
                               double nu, double* gradient, double* distance) {
    return guard([&] {
        need(psi, "psi");
        need(psi_prime, "psi_prime");
        need(op, "op");
        need(ens, "ens");
        need(distance, "distance");
        *distance = ann::hilbert_space_distance(psi->impl, psi_prime->impl, op->impl, is_unitary != 0, ens->impl, nu,
                                                gradient ? C(gradient) : nullptr);
    });
}

// ---------------------------------------------------------------------- TDVP

int ann_kl_create(ann_kl* out, uint32_t num_params) {
    return guard([&] {
        need(out, "out");
        auto p = new ann_kl_s;
        p->impl.P = int(num_params);
        *out = p;
    });
}

int ann_kl_destroy(ann_kl kl) {
    return guard([&] { delete kl; });
}

int ann_kl_eval(ann_kl kl, ann_psi psi, ann_psi psi_prime, ann_ensemble ens, double threshold, double nu,
                int32_t mode, double* gradient, double* noise, double* value) {
    return guard([&] {
        need(kl, "kl");
        need(psi, "psi");
        need(psi_prime, "psi_prime");
        need(ens, "ens");
        need(value, "value");
        if (mode < 0 || mode > 2) throw std::invalid_argument("mode must be 0, 1 or 2");
        if (psi->impl.kind != ann::PSI_CLASSICAL) throw std::invalid_argument("psi must be a PsiClassical");
        *value = ann::kullback_leibler(kl->impl, psi->impl, psi_prime->impl, ens->impl, threshold, mode, nu,
                                       gradient ? C(gradient) : nullptr, noise);
    });
}

int ann_kl_info(ann_kl kl, double* total_weight, double* mean_deviation2) {
    return guard([&] {
        need(kl, "kl");
        if (total_weight) *total_weight = kl->impl.total_weight;
        if (mean_deviation2) *C(mean_deviation2) = kl->impl.mean_deviation;
    });
}

int ann_kl_get_log_psi_scale(ann_kl kl, double* scale) {
    return guard([&] {
        need(kl, "kl");
        need(scale, "scale");
        *scale = kl->impl.log_psi_scale;
    });
}

int ann_kl_set_log_psi_scale(ann_kl kl, double scale) {
    return guard([&] {
        need(kl, "kl");
        if (!std::isfinite(scale)) throw std::invalid_argument("log_psi_scale must be finite");
        kl->impl.log_psi_scale = scale;
    });
}

int ann_tdvp_create(ann_tdvp* out, uint32_t num_params) {
    return guard([&] {
        need(out, "out");
        auto p = new ann_tdvp_s;
        p->impl.P = num_params;
        *out = p;
    });
}

int ann_tdvp_destroy(ann_tdvp tdvp) {
    return guard([&] { delete tdvp; });
}

int ann_tdvp_eval(ann_tdvp tdvp, ann_operator op, ann_psi psi, ann_ensemble ens, int32_t with_S) {
    return guard([&] {
        need(tdvp, "tdvp");
        need(op, "op");
        need(psi, "psi");
        need(ens, "ens");
        ann::tdvp_eval(tdvp->impl, op->impl, psi->impl, ens->impl, with_S != 0);
    });
}

int ann_tdvp_get_S(ann_tdvp tdvp, double* S) {
    return guard([&] {
        need(tdvp, "tdvp");
        if (!tdvp->impl.has_S) throw std::logic_error("no S matrix: call eval first");
        tdvp->impl.S.download(reinterpret_cast<ann::cplx<double>*>(S), size_t(tdvp->impl.P) * tdvp->impl.P);
    });
}

int ann_tdvp_get_F(ann_tdvp tdvp, double* F) {
    return guard([&] {
        need(tdvp, "tdvp");
        tdvp->impl.fetch_host_vectors();
        std::memcpy(F, tdvp->impl.F.data(), sizeof(cd) * tdvp->impl.F.size());
    });
}

int ann_tdvp_get_O(ann_tdvp tdvp, double* O) {
    return guard([&] {
        need(tdvp, "tdvp");
        tdvp->impl.fetch_host_vectors();
        std::memcpy(O, tdvp->impl.O.data(), sizeof(cd) * tdvp->impl.O.size());
    });
}

int ann_tdvp_get_energy(ann_tdvp tdvp, double* E2c, double* E2_local, double* var_H) {
    return guard([&] {
        need(tdvp, "tdvp");
        auto& t = tdvp->impl;
        t.settle();
        if (E2c) *C(E2c) = t.E;
        if (E2_local) *E2_local = t.E2;
        if (var_H) *var_H = t.E2 - std::norm(t.E);   // TDVP.hpp:69-71
    });
}

int ann_tdvp_get_samples(ann_tdvp tdvp, uint64_t* confs, double* weights, double* e_local, int64_t capacity,
                         int64_t* count, int32_t* words, double* total_weight) {
    return guard([&] {
        need(tdvp, "tdvp");
        auto& t = tdvp->impl;
        if (!t.has_samples) throw std::logic_error("TDVP has not been evaluated");
        t.settle();
        const int64_t n = t.count;
        if (count) *count = n;
        if (words) *words = t.words;
        if (total_weight) *total_weight = t.wsum;
        if ((confs || weights || e_local) && capacity < n) throw std::invalid_argument("buffer too small");
        if (confs) t.confs.download(confs, size_t(n) * t.words);
        if (weights) t.w.download(weights, size_t(n));
        if (e_local) t.e_loc.download(reinterpret_cast<ann::cplx<double>*>(e_local), size_t(n));
    });
}

int ann_tdvp_get_O_samples(ann_tdvp tdvp, double* out, int64_t capacity) {
    return guard([&] {
        need(tdvp, "tdvp");
        need(out, "out");
        auto& t = tdvp->impl;
        if (!t.has_samples) throw std::logic_error("TDVP has not been evaluated");
        if (capacity < t.count * int64_t(t.P)) throw std::invalid_argument("buffer too small");
        ann::tdvp_O_samples(t, C(out));
    });
}

int ann_tdvp_S_dot_vector(ann_tdvp tdvp, const double* v, double* out) {
    return guard([&] {
        need(tdvp, "tdvp");
        auto& t = tdvp->impl;
        ann::DevBuf<ann::cplx<double>> dv, dy;
        dv.upload(reinterpret_cast<const ann::cplx<double>*>(v), t.P);
        dy.ensure(t.P);
        ann::tdvp_S_dot(t, dv.p, 0.0, dy.p);
        ann::Runtime::get().sync();
        dy.download(reinterpret_cast<ann::cplx<double>*>(out), t.P);
    });
}

int ann_tdvp_solve(ann_tdvp tdvp, double shift, int32_t method, double rtol, int32_t max_iter, double* x,
                   int32_t* iterations, double* residual) {
    int rc = ANN_OK;
    const int g = guard([&] {
        // arguments are checked before the handle, so that they fail the same way with or without a GPU
        if (method < ANN_SOLVE_CHOLESKY || method > ANN_SOLVE_CG_KFAC_WARM) throw std::invalid_argument("unknown method");
        if (!std::isfinite(shift)) throw std::invalid_argument("shift must be finite");
        // the K-FAC damping is 0.3 sqrt(shift) times the factor scale: a shift <= 0 leaves the
        // spin covariance undamped, and under S_z-conserving moves it has the exact null vector
        // (1, ..., 1), so its inverse does not exist
        if ((method == ANN_SOLVE_CG_KFAC || method == ANN_SOLVE_CG_KFAC_WARM) && !(shift > 0.0))
            throw std::invalid_argument("the K-FAC preconditioned solve needs shift > 0");
        need(tdvp, "tdvp");
        need(x, "x");
        int it = 0;
        double res = 0.0;
        std::vector<cd> x0;
        const bool warm = method == ANN_SOLVE_CG_WARM || method == ANN_SOLVE_CG_KFAC_WARM;
        if (warm) x0.assign(C(x), C(x) + tdvp->impl.P);
        const int inner = method == ANN_SOLVE_CG_WARM ? 1 : method == ANN_SOLVE_CG_KFAC_WARM ? 3 : method;
        const int r = ann::tdvp_solve(tdvp->impl, shift, inner, rtol, max_iter, C(x), &it, &res,
                                      x0.empty() ? nullptr : x0.data());
        if (iterations) *iterations = it;
        if (residual) *residual = res;
        if (r == 1) {
            g_error = "S + shift is not positive definite";
            rc = ANN_E_NUMERIC;
        } else if (r == 2) {
            g_error = "CG did not reach the requested tolerance";
            rc = ANN_E_NUMERIC;
        }
    });
    return g != ANN_OK ? g : rc;
}

// ---------------------------------------------------------- learned proposal

int ann_proposal_create(ann_proposal* out, int32_t num_sites, int32_t hidden, int32_t num_blocks,
                        int32_t block_size, const int32_t* block_sites, uint64_t seed) {
    return guard([&] {
        need(out, "out");
        need(block_sites, "block_sites");
        if (num_sites < 1 || num_sites > 64 * ann::kMaxWords) throw std::invalid_argument("1..256 sites");
        if (hidden < 1 || hidden > 256) throw std::invalid_argument("hidden must be 1..256");
        if (block_size < 1 || block_size > ann::kMaxBlock) throw std::invalid_argument("block_size must be 1..8");
        if (num_blocks < 1) throw std::invalid_argument("num_blocks must be positive");
        for (int64_t r = 0; r < int64_t(num_blocks) * block_size; r++)
            if (block_sites[r] < 0 || block_sites[r] >= num_sites) throw std::invalid_argument("block site out of range");
        for (int b = 0; b < num_blocks; b++)
            for (int i = 0; i < block_size; i++)
                for (int j = 0; j < i; j++)
                    if (block_sites[b * block_size + i] == block_sites[b * block_size + j])
                        throw std::invalid_argument("repeated site in a block");
        auto p = new ann_proposal_s;
        try {
            auto& q = p->impl;
            q.N = num_sites;
            q.H = hidden;
            q.bs = block_size;
            q.num_blocks = num_blocks;
            q.seed = seed;
            q.blocks.upload(block_sites, size_t(num_blocks) * block_size);
            // W ~ U(-1/sqrt(fan_in), 1/sqrt(fan_in)), biases 0
            std::vector<float> h(q.num_params(), 0.0f);
            std::mt19937_64 gen(seed);
            std::uniform_real_distribution<double> u1(-1.0 / std::sqrt(double(num_sites)), 1.0 / std::sqrt(double(num_sites)));
            std::uniform_real_distribution<double> u2(-1.0 / std::sqrt(double(hidden)), 1.0 / std::sqrt(double(hidden)));
            const size_t nW1 = size_t(hidden) * num_sites;
            for (size_t k = 0; k < nW1; k++) h[k] = float(u1(gen));
            for (size_t k = 0; k < size_t(num_sites) * hidden; k++) h[nW1 + hidden + k] = float(u2(gen));
            q.params.upload(h.data(), h.size());
            std::vector<float> zero(h.size(), 0.0f);
            q.m.upload(zero.data(), zero.size());
            q.v.upload(zero.data(), zero.size());
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}

int ann_proposal_destroy(ann_proposal p) {
    return guard([&] { delete p; });
}

int ann_proposal_num_params(ann_proposal p, int64_t* n) {
    return guard([&] {
        need(p, "p");
        need(n, "n");
        *n = p->impl.num_params();
    });
}

int ann_proposal_get_params(ann_proposal p, float* params) {
    return guard([&] {
        need(p, "p");
        need(params, "params");
        p->impl.params.download(params, p->impl.num_params());
    });
}

int ann_proposal_set_params(ann_proposal p, const float* params) {
    return guard([&] {
        need(p, "p");
        need(params, "params");
        p->impl.params.upload(params, p->impl.num_params());
        p->impl.packed = false;
    });
}

int ann_proposal_loss_grad(ann_proposal p, const uint64_t* confs, int64_t count, int32_t words,
                           const int32_t* sites, double* loss, float* grad) {
    return guard([&] {
        need(p, "p");
        need(confs, "confs");
        need(sites, "sites");
        std::vector<float> g;
        const double l = ann::proposal_loss_grad(p->impl, confs, count, words, sites, g);
        if (loss) *loss = l;
        if (grad) std::memcpy(grad, g.data(), sizeof(float) * g.size());
    });
}

int ann_proposal_adam_step(ann_proposal p, const float* grad, double lr) {
    return guard([&] {
        need(p, "p");
        need(grad, "grad");
        ann::proposal_adam_step(p->impl, grad, lr);
    });
}

int ann_proposal_train(ann_proposal p, ann_ensemble ens, int32_t steps, int32_t batch, double lr, double* loss) {
    return guard([&] {
        need(p, "p");
        need(ens, "ens");
        auto& s = ens->impl.samples;
        if (s.count <= 0) throw std::invalid_argument("the ensemble holds no samples yet");
        const double l = ann::proposal_train(p->impl, s.confs.p, s.count, s.words, steps, batch, lr);
        if (loss) *loss = l;
    });
}

int ann_monte_carlo_set_proposal(ann_ensemble ens, ann_proposal p, uint32_t plain_sweeps) {
    return guard([&] {
        need(ens, "ens");
        if (ens->impl.kind != ann::ENS_MC) throw std::invalid_argument("proposals apply to Monte Carlo ensembles");
        ens->impl.proposal = p ? &p->impl : nullptr;
        ens->impl.plain_sweeps = plain_sweeps;
    });
}

int ann_monte_carlo_set_proposal_mode(ann_ensemble ens, int32_t exclude_current) {
    return guard([&] {
        need(ens, "ens");
        if (ens->impl.kind != ann::ENS_MC) throw std::invalid_argument("proposals apply to Monte Carlo ensembles");
        ens->impl.proposal_exclusive = exclude_current != 0;
    });
}

}  // extern "C"
