// runtime.cpp -- host runtime: parameter upload, sampling driver, local energies,
// TDVP statistics and the SR solve.  Every device call goes through the kernels
// in dense_net.hip / linalg.hip / sampler_ops.hip on one HIP stream.
#include "runtime.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <type_traits>

namespace ann {

// ------------------------------------------------------------------ runtime

Runtime& Runtime::get() {
    static Runtime rt;
    return rt;
}

hipStream_t Runtime::side(int k) {
    if (k < 0 || k > 1) throw std::logic_error("side stream index out of range");
    const hipStream_t main = stream();
    (void)main;
    if (!side_[k]) ANN_HIP_CHECK(hipStreamCreateWithFlags(&side_[k], hipStreamNonBlocking));
    return side_[k];
}

hipStream_t Runtime::stream() {
    if (override_) return override_;
    if (!stream_) {
        ANN_HIP_CHECK(hipSetDevice(device_));
        ANN_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        // keep stream-ordered allocations cached in the default pool
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, device_) == hipSuccess) {
            uint64_t threshold = ~uint64_t(0);
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold);
        }
    }
    return stream_;
}

void Runtime::set_device(int device) {
    if (stream_) {
        ANN_HIP_CHECK(hipStreamSynchronize(stream_));
        collect();
        for (auto e : pool_) (void)hipEventDestroy(e);
        pool_.clear();
        ANN_HIP_CHECK(hipStreamDestroy(stream_));
        stream_ = nullptr;
        for (auto& sd : side_) {
            if (!sd) continue;
            ANN_HIP_CHECK(hipStreamSynchronize(sd));
            ANN_HIP_CHECK(hipStreamDestroy(sd));
            sd = nullptr;
        }
    }
    device_ = device;
    ANN_HIP_CHECK(hipSetDevice(device_));
}

void Runtime::sync() {
    ANN_HIP_CHECK(hipStreamSynchronize(stream()));
    collect();
}

hipEvent_t Runtime::event() {
    if (!pool_.empty()) {
        hipEvent_t e = pool_.back();
        pool_.pop_back();
        return e;
    }
    hipEvent_t e;
    ANN_HIP_CHECK(hipEventCreate(&e));
    return e;
}

void Runtime::begin(const char* name) {
    open_name_ = name;
    open_a_ = event();
    ANN_HIP_CHECK(hipEventRecord(open_a_, stream()));
}

void Runtime::end() {
    hipEvent_t b = event();
    ANN_HIP_CHECK(hipEventRecord(b, stream()));
    pending_.push_back({open_name_, open_a_, b});
    open_a_ = nullptr;
}

void Runtime::collect() {
    for (auto& p : pending_) {
        ANN_HIP_CHECK(hipEventSynchronize(p.b));
        float ms = 0.f;
        ANN_HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
        auto& t = totals[p.name];
        t.first += ms;
        t.second += 1;
        pool_.push_back(p.a);
        pool_.push_back(p.b);
    }
    pending_.clear();
}

hipStream_t lib_stream() { return Runtime::get().stream(); }

static hipStream_t S() { return Runtime::get().stream(); }

// --------------------------------------------------------------------- Psi

// the wave functions a configuration basis admits: on Pauli strings PsiClassical and a
// PsiDeep with three input units per site (PsiDeep.hpp:283-305); such a PsiDeep reads no
// spins
static void check_basis(const Psi& psi, int basis) {
    if (basis == BASIS_PAULIS) {
        if (psi.num_sites > 64) throw std::invalid_argument("Pauli-string configurations hold at most 64 sites");
        if (psi.cnn_ref()) throw std::logic_error("a PsiClassical with a PsiCNN reference reads spins only");
        if (psi.kind == PSI_CLASSICAL || psi.pauli_net()) return;
        throw std::logic_error("on Pauli strings: PsiClassical, or a PsiDeep with 3 input units per site");
    }
    if (psi.pauli_net()) throw std::logic_error("this PsiDeep reads Pauli strings (3 input units per site)");
}

// the kernels without a structured O serve every wave function but a dense net on spins
static bool generic_path(const Psi& psi, int basis) {
    return !psi.dense() || basis != BASIS_SPINS || psi.pauli_net();
}

// f(view) with the generic view of the wave function in its precision and basis
template <typename F>
static void with_generic(Psi& psi, F&& f, int basis = BASIS_SPINS) {
    check_basis(psi, basis);
    const bool d = psi.precision == 0;
    if (psi.kind == PSI_CNN) {
        if (d)
            f(psi.cnn64());
        else
            f(psi.cnn32());
    } else if (psi.cnn_ref()) {
        if (d)
            f(psi.classical_cnn64());
        else
            f(psi.classical_cnn32());
    } else if (psi.kind == PSI_CLASSICAL) {
        if (d)
            f(psi.classical64(basis));
        else
            f(psi.classical32(basis));
    } else if (basis == BASIS_PAULIS) {
        if (d)
            f(PauliInput<MlpView<double>>{psi.mlp64(), psi.num_sites, psi.num_params});
        else
            f(PauliInput<MlpView<float>>{psi.mlp32(), psi.num_sites, psi.num_params});
    } else {
        if (d)
            f(psi.mlp64());
        else
            f(psi.mlp32());
    }
}

static int units_per_lane(int M) {
    const int need = (M + 63) / 64;
    for (int u : {1, 2, 4, 7, 8, 16})
        if (need <= u) return u;
    throw std::invalid_argument("hidden layer wider than 1024 units is not supported");
}

void Psi::finalize() {
    words = std::max(1, (num_sites + 63) / 64);
    if (num_sites > 64 * kMaxWords) throw std::invalid_argument("at most 256 sites are supported");
    if (kind == PSI_RBM) {
        layout = LAYOUT_RBM;
        C = N;
        num_params = N * M;
        Mpad = 64 * units_per_lane(M);
    } else if (kind == PSI_DEEP) {
        num_params = N;
        for (int l = 0; l < num_layers; l++) num_params += sizes[l] + conns[l] * sizes[l];
        if (num_layers == 1) {
            layout = LAYOUT_DEEP1;
            M = sizes[0];
            C = conns[0];
            W = w_l[0];
            b = b_l[0];
            conn = conn_l[0];
            Mpad = 64 * units_per_lane(M);
            for (uint32_t c : conn)
                if (int(c) >= N) throw std::invalid_argument("PsiDeep connection index out of range");
        }
    } else if (kind == PSI_CLASSICAL) {
        num_params = int(params.size()) + (ref_enters() ? ref->num_params : 0);
    } else {
        num_params = int(params.size());
    }
    dirty = true;
    mlp_stale = true;
}

// a complex host vector into an fp64 and an fp32 device copy
static void up_both(DevBuf<cplx<double>>& d64, DevBuf<cplx<float>>& d32, const std::vector<cd>& h) {
    d64.upload(reinterpret_cast<const cplx<double>*>(h.data()), h.size());
    std::vector<cplx<float>> t(h.size());
    for (size_t i = 0; i < h.size(); i++) t[i] = {float(h[i].real()), float(h[i].imag())};
    d32.upload(t.data(), t.size());
}

// the layer tables of the generic PsiDeep view: every layer's connections, weights and
// biases and, for layers l >= 1, the weights reading each unit of layer l - 1
void Psi::upload_mlp_tables() {
    std::vector<uint32_t> cn;
    std::vector<cd> wv, bv;
    std::vector<int> rb, ri;
    mlp_cn_off.assign(num_layers, 0);
    mlp_w_off.assign(num_layers, 0);
    mlp_b_off.assign(num_layers, 0);
    mlp_rb_off.assign(num_layers, 0);
    mlp_ri_off.assign(num_layers, 0);
    for (int l = 0; l < num_layers; l++) {
        const int Ml = sizes[l], Cl = conns[l];
        mlp_cn_off[l] = int(cn.size());
        mlp_w_off[l] = int(wv.size());
        mlp_b_off[l] = int(bv.size());
        cn.insert(cn.end(), conn_l[l].begin(), conn_l[l].end());
        wv.insert(wv.end(), w_l[l].begin(), w_l[l].end());
        bv.insert(bv.end(), b_l[l].begin(), b_l[l].end());
        mlp_rb_off[l] = int(rb.size());
        mlp_ri_off[l] = int(ri.size());
        if (l > 0) {
            const int prev = sizes[l - 1];
            std::vector<std::vector<int>> by_unit(prev);
            for (int i = 0; i < Cl; i++)
                for (int j = 0; j < Ml; j++) by_unit[conn_l[l][size_t(i) * Ml + j]].push_back(i * Ml + j);
            int e = 0;
            for (int u = 0; u < prev; u++) {
                rb.push_back(e);
                ri.insert(ri.end(), by_unit[u].begin(), by_unit[u].end());
                e += int(by_unit[u].size());
            }
            rb.push_back(e);
        }
    }
    mlp_cn.upload(cn.data(), cn.size());
    up_both(mlp_w64, mlp_w32, wv);
    up_both(mlp_b64, mlp_b32, bv);
    if (rb.empty()) rb.push_back(0);
    if (ri.empty()) ri.push_back(0);
    mlp_rb.upload(rb.data(), rb.size());
    mlp_ri.upload(ri.data(), ri.size());
}

void Psi::sync_device() {
    if (!dirty && params_dirty) {
        // parameters only: one asynchronous copy and a repack on the device
        params_d.ensure(size_t(num_params));
        ANN_HIP_CHECK(hipMemcpyAsync(params_d.p, stage.p, sizeof(cplx<double>) * num_params, hipMemcpyHostToDevice,
                                     S()));
        ANN_HIP_CHECK(hipEventRecord(stage_ev.get(), S()));
        launch_net_repack(layout, N, M, Mpad, C, conn_shift, params_d.p, kind == PSI_DEEP ? conn_d.p : nullptr, npad,
                          Wd64.p, Wd32.p, b64.p, b32.p, wfrag.p, S());
        params_dirty = false;
        return;
    }
    if (!dirty) return;
    params_dirty = false;
    if (kind == PSI_CNN) {
        // parameters, neighbour tables (Convolve.hpp:122-148) and symmetry classes
        cp64.upload(reinterpret_cast<const cplx<double>*>(params.data()), params.size());
        std::vector<cplx<float>> p32(params.size());
        for (size_t i = 0; i < params.size(); i++) p32[i] = {float(params[i].real()), float(params[i].imag())};
        cp32.upload(p32.data(), p32.size());
        const int e0 = extent[0], e1 = extent[1], e2 = extent[2], page = e1 * e2;
        std::vector<int> nbr;
        for (int l = 0; l < cnn_layers; l++) {
            const int* cv = &connectivity[3 * l];
            for (int x = 0; x < N; x++) {
                const int pg = x / page, row = (x % page) / e2, col = x % e2;
                for (int k = 0; k < cv[0]; k++)
                    for (int i = 0; i < cv[1]; i++)
                        for (int j = 0; j < cv[2]; j++)
                            nbr.push_back(((pg + k) % e0) * page + ((row + i) % e1) * e2 + (col + j) % e2);
            }
        }
        nbr_d.upload(nbr.data(), nbr.size());
        // consumers of every site in every layer (the receptive fields of the local updates)
        std::vector<uint64_t> cons(size_t(cnn_layers) * N * kMaxWords, 0);
        size_t off = 0;
        for (int l = 0; l < cnn_layers; l++) {
            const int* cv = &connectivity[3 * l];
            const int vol = cv[0] * cv[1] * cv[2];
            for (int x = 0; x < N; x++)
                for (int k = 0; k < vol; k++) {
                    const int y = nbr[off + size_t(x) * vol + k];
                    cons[(size_t(l) * N + y) * kMaxWords + (x >> 6)] |= uint64_t(1) << (x & 63);
                }
            off += size_t(N) * vol;
        }
        cons_d.upload(cons.data(), cons.size());
        sym_d.upload(sym.data(), sym.size());
        dirty = false;
        return;
    }
    if (kind == PSI_CLASSICAL) {
        // classical parameters, the strings of every H_local[n] with their coefficients
        // in both bases, and the reference network
        up_both(cp64, cp32, params);
        std::vector<int> begin(cl_begin.begin(), cl_begin.end());
        cl_begin_d.upload(begin.data(), begin.size());
        const size_t ns = cl_coef.size();
        std::vector<uint64_t> a0(cl_a), b0(cl_b);
        if (a0.empty()) a0.assign(kMaxWords, 0);
        if (b0.empty()) b0.assign(kMaxWords, 0);
        cl_a_d.upload(a0.data(), a0.size());
        cl_b_d.upload(b0.data(), b0.size());
        std::vector<double> cs(2 * std::max<size_t>(ns, 1), 0.0), cp(2 * std::max<size_t>(ns, 1), 0.0);
        for (size_t m = 0; m < ns; m++) {
            // (-i)^{#Y} of the string on spins (PauliString.hpp:189-206)
            int ny = 0;
            for (int k = 0; k < kMaxWords; k++)
                ny += __builtin_popcountll(~cl_a[m * kMaxWords + k] & cl_b[m * kMaxWords + k]);
            static const cd phase[4] = {cd(1, 0), cd(0, -1), cd(-1, 0), cd(0, 1)};
            const cd c = cl_coef[m] * phase[ny & 3];
            cs[2 * m] = c.real();
            cs[2 * m + 1] = c.imag();
            cp[2 * m] = cl_coef[m].real();
            cp[2 * m + 1] = cl_coef[m].imag();
        }
        cl_coef_s_d.upload(cs.data(), cs.size());
        cl_coef_p_d.upload(cp.data(), cp.size());
        dirty = false;
        return;
    }
    if (!dense()) {
        // multi-layer PsiDeep: every layer's tables, and the final weights
        upload_mlp_tables();
        up_both(f64, f32, f);
        dirty = false;
        return;
    }
    std::vector<cd> Wd(size_t(N) * Mpad, 0.0), bd(Mpad, 0.0), fd(Mpad, 0.0);
    if (kind == PSI_RBM) {
        for (int v = 0; v < N; v++)
            for (int j = 0; j < M; j++) Wd[size_t(v) * Mpad + j] = W[size_t(v) * M + j];
        for (int j = 0; j < M; j++) fd[j] = f[0];
    } else {
        for (int i = 0; i < C; i++)
            for (int j = 0; j < M; j++) Wd[size_t(conn[size_t(i) * M + j]) * Mpad + j] += W[size_t(i) * M + j];
        for (int j = 0; j < M; j++) {
            bd[j] = b[j];
            fd[j] = f[j];
        }
    }
    real_f = true;
    for (int j = 0; j < M; j++) real_f = real_f && fd[j].imag() == 0.0;
    auto up64 = [](DevBuf<cplx<double>>& d, const std::vector<cd>& h) {
        d.upload(reinterpret_cast<const cplx<double>*>(h.data()), h.size());
    };
    auto up32 = [](DevBuf<cplx<float>>& d, const std::vector<cd>& h) {
        std::vector<cplx<float>> t(h.size());
        for (size_t i = 0; i < h.size(); i++) t[i] = {float(h[i].real()), float(h[i].imag())};
        d.upload(t.data(), t.size());
    };
    up64(Wd64, Wd);
    up64(b64, bd);
    up64(f64, fd);
    up32(Wd32, Wd);
    up32(b32, bd);
    up32(f32, fd);
    std::vector<cd> z(Mpad, 0.0);
    up64(zero64, z);
    up32(zero32, z);
    conn_shift = -1;
    if (kind == PSI_DEEP) {
        conn_d.upload(conn.data(), conn.size());
        // circulant layout: connection i of unit j is site (j d + i) mod N
        const int d = M > 1 ? int(conn[1]) : 0;
        bool circ = C <= N;
        for (int i = 0; i < C && circ; i++)
            for (int j = 0; j < M; j++)
                if (int(conn[size_t(i) * M + j]) != int((int64_t(j) * d + i) % N)) {
                    circ = false;
                    break;
                }
        conn_shift = circ ? d : -1;
    }
    std::vector<uint16_t> frag;
    pack_theta_fragments(reinterpret_cast<const double*>(Wd.data()), N, Mpad, 3, frag, npad);
    wfrag.upload(frag.data(), frag.size());
    dirty = false;
}

DenseView<double> Psi::view64() {
    sync_device();
    return {N, M, Mpad, C, layout, num_params, real_f ? 1 : 0, Wd64.p, b64.p, f64.p,
            kind == PSI_DEEP ? conn_d.p : nullptr};
}

DenseView<float> Psi::view32() {
    sync_device();
    return {N, M, Mpad, C, layout, num_params, real_f ? 1 : 0, Wd32.p, b32.p, f32.p,
            kind == PSI_DEEP ? conn_d.p : nullptr};
}

template <typename T>
static CnnView<T> make_cnn_view(const Psi& q, const cplx<T>* p, const int* nbr, const uint32_t* sym) {
    CnnView<T> v{};
    v.N = q.N;
    v.L = q.cnn_layers;
    v.nsym = q.num_sym;
    v.num_params = q.num_params;
    v.max_ch = 1;
    int link = 0, noff = 0, zoff = 0, prev = 1;
    for (int l = 0; l < q.cnn_layers; l++) {
        const int* cv = &q.connectivity[3 * l];
        v.ch[l] = q.channels[l];
        v.vol[l] = cv[0] * cv[1] * cv[2];
        v.link_off[l] = link;
        v.nbr_off[l] = noff;
        v.zoff[l] = zoff;
        link += prev * v.ch[l] * v.nsym * v.vol[l];
        noff += q.N * v.vol[l];
        zoff += v.ch[l] * q.N;
        v.max_ch = std::max(v.max_ch, v.ch[l]);
        prev = v.ch[l];
    }
    v.zoff[q.cnn_layers] = zoff;
    v.final_factor = T(q.final_factor);
    v.params = p;
    v.nbr = nbr;
    v.sym = sym;
    v.cons = q.cons_d.p;
    return v;
}

template <typename T>
static MlpView<T> make_mlp_view(const Psi& q, const cplx<T>* w, const cplx<T>* b, const cplx<T>* f) {
    MlpView<T> v{};
    v.N = q.N;
    v.L = q.num_layers;
    v.num_params = q.num_params;
    int z = 0, par = q.N;
    for (int l = 0; l < q.num_layers; l++) {
        v.size[l] = q.sizes[l];
        v.conn[l] = q.conns[l];
        v.zoff[l] = z;
        v.poff[l] = par;
        z += q.sizes[l];
        par += q.sizes[l] + q.sizes[l] * q.conns[l];
        v.cn[l] = q.mlp_cn.p + q.mlp_cn_off[l];
        v.W[l] = w + q.mlp_w_off[l];
        v.b[l] = b + q.mlp_b_off[l];
        v.rb[l] = q.mlp_rb.p + q.mlp_rb_off[l];
        v.ri[l] = q.mlp_ri.p + q.mlp_ri_off[l];
    }
    v.zoff[q.num_layers] = z;
    v.f = f;
    return v;
}

// a one-layer PsiDeep builds the generic tables on first use after each change (its
// final weights are the dense net's, zero-padded beyond M)
static void ensure_mlp(Psi& q) {
    if (q.kind != PSI_DEEP) throw std::logic_error("not a PsiDeep");
    q.sync_device();
    if (q.dense() && q.mlp_stale) {
        q.upload_mlp_tables();
        q.mlp_stale = false;
    }
}

MlpView<double> Psi::mlp64() {
    ensure_mlp(*this);
    return make_mlp_view<double>(*this, mlp_w64.p, mlp_b64.p, f64.p);
}

MlpView<float> Psi::mlp32() {
    ensure_mlp(*this);
    return make_mlp_view<float>(*this, mlp_w32.p, mlp_b32.p, f32.p);
}

template <typename T, typename R = MlpView<T>>
static ClassicalView<T, R> make_classical_view(Psi& q, int basis, const cplx<T>* params) {
    ClassicalView<T, R> v{};
    v.N = q.num_sites;
    v.basis = basis;
    v.num_ops = int(q.params.size());
    v.num_params = q.num_params;
    v.has_ref = q.ref_enters() ? 1 : 0;
    v.op_begin = q.cl_begin_d.p;
    v.sa = q.cl_a_d.p;
    v.sb = q.cl_b_d.p;
    v.coef = basis == BASIS_PAULIS ? q.cl_coef_p_d.p : q.cl_coef_s_d.p;
    v.params = params;
    return v;
}

ClassicalView<double> Psi::classical64(int basis) {
    sync_device();
    auto v = make_classical_view<double>(*this, basis, cp64.p);
    if (v.has_ref) {
        ref->precision = precision;
        v.ref = ref->mlp64();
        v.ref_lp = {ref->log_prefactor.real(), ref->log_prefactor.imag()};
    }
    return v;
}

ClassicalView<float> Psi::classical32(int basis) {
    sync_device();
    auto v = make_classical_view<float>(*this, basis, cp32.p);
    if (v.has_ref) {
        ref->precision = precision;
        v.ref = ref->mlp32();
        v.ref_lp = {float(ref->log_prefactor.real()), float(ref->log_prefactor.imag())};
    }
    return v;
}

ClassicalView<double, CnnView<double>> Psi::classical_cnn64() {
    sync_device();
    auto v = make_classical_view<double, CnnView<double>>(*this, BASIS_SPINS, cp64.p);
    if (v.has_ref) {
        ref->precision = precision;
        v.ref = ref->cnn64();
        v.ref_lp = {ref->log_prefactor.real(), ref->log_prefactor.imag()};
    }
    return v;
}

ClassicalView<float, CnnView<float>> Psi::classical_cnn32() {
    sync_device();
    auto v = make_classical_view<float, CnnView<float>>(*this, BASIS_SPINS, cp32.p);
    if (v.has_ref) {
        ref->precision = precision;
        v.ref = ref->cnn32();
        v.ref_lp = {float(ref->log_prefactor.real()), float(ref->log_prefactor.imag())};
    }
    return v;
}

CnnView<double> Psi::cnn64() {
    sync_device();
    return make_cnn_view<double>(*this, cp64.p, nbr_d.p, sym_d.p);
}

CnnView<float> Psi::cnn32() {
    sync_device();
    return make_cnn_view<float>(*this, cp32.p, nbr_d.p, sym_d.p);
}

std::vector<cd> Psi::get_params() const {
    std::vector<cd> p;
    p.reserve(num_params);
    if (kind == PSI_RBM) {
        p = W;
    } else if (kind == PSI_DEEP) {
        p.insert(p.end(), a.begin(), a.end());
        for (int l = 0; l < num_layers; l++) {
            p.insert(p.end(), b_l[l].begin(), b_l[l].end());
            p.insert(p.end(), w_l[l].begin(), w_l[l].end());
        }
    } else {
        p = params;
        if (ref_enters()) {
            const auto r = ref->get_params();
            p.insert(p.end(), r.begin(), r.end());
        }
    }
    return p;
}

void Psi::set_params(const cd* p) {
    if (kind == PSI_RBM) {
        W.assign(p, p + num_params);
    } else if (kind == PSI_DEEP) {
        size_t o = 0;
        a.assign(p, p + N);
        o += N;
        for (int l = 0; l < num_layers; l++) {
            b_l[l].assign(p + o, p + o + sizes[l]);
            o += sizes[l];
            w_l[l].assign(p + o, p + o + size_t(conns[l]) * sizes[l]);
            o += size_t(conns[l]) * sizes[l];
        }
        if (num_layers == 1) {
            W = w_l[0];
            b = b_l[0];
        }
    } else if (kind == PSI_CLASSICAL) {
        const size_t nc = params.size();
        params.assign(p, p + nc);
        if (ref_enters()) ref->set_params(p + nc);
    } else {
        params.assign(p, p + num_params);
    }
    mlp_stale = true;
    if (dense() && !dirty) {
        // the previous staged copy must have left the page-locked buffer
        if (stage_ev.e) ANN_HIP_CHECK(hipEventSynchronize(stage_ev.e));
        stage.ensure(size_t(num_params));
        std::copy(p, p + num_params, reinterpret_cast<cd*>(stage.p));
        params_dirty = true;
    } else {
        dirty = true;
    }
}

std::unique_ptr<Psi> Psi::clone() const {
    auto q = std::make_unique<Psi>();
    q->kind = kind;
    q->num_sites = num_sites;
    q->N = N;
    q->num_params = num_params;
    q->words = words;
    q->log_prefactor = log_prefactor;
    q->precision = precision;
    q->M = M;
    q->C = C;
    q->Mpad = Mpad;
    q->layout = layout;
    q->W = W;
    q->a = a;
    q->b = b;
    q->f = f;
    q->conn = conn;
    q->num_layers = num_layers;
    q->sizes = sizes;
    q->conns = conns;
    q->conn_l = conn_l;
    q->w_l = w_l;
    q->b_l = b_l;
    std::copy(extent, extent + 3, q->extent);
    q->cnn_layers = cnn_layers;
    q->channels = channels;
    q->connectivity = connectivity;
    q->sym = sym;
    q->num_sym = num_sym;
    q->params = params;
    q->final_factor = final_factor;
    q->cl_order = cl_order;
    q->cl_begin = cl_begin;
    q->cl_a = cl_a;
    q->cl_b = cl_b;
    q->cl_coef = cl_coef;
    q->ref_log_prefactor = ref_log_prefactor;
    if (ref) q->ref = ref->clone();
    q->dirty = true;
    return q;
}

// ---------------------------------------------------------------- operator

int Operator::max_site() const {
    int m = -1;
    for (int n = 0; n < num_strings; n++)
        for (int w = 0; w < kMaxWords; w++) {
            const uint64_t x = a[size_t(n) * kMaxWords + w] | b[size_t(n) * kMaxWords + w];
            if (x) m = std::max(m, w * 64 + 63 - __builtin_clzll(x));
        }
    return m;
}

const DeviceOp& Operator::device(int words) {
    auto it = compiled.find(words);
    if (it != compiled.end()) return *it->second;
    auto d = std::make_unique<DeviceOp>();
    d->words = words;
    // group strings by flip mask (X|Y bits = a ^ b), first-occurrence order
    std::vector<std::vector<int>> groups;
    std::vector<std::vector<uint64_t>> masks;
    for (int n = 0; n < num_strings; n++) {
        std::vector<uint64_t> m(words);
        for (int w = 0; w < words; w++) m[w] = a[size_t(n) * kMaxWords + w] ^ b[size_t(n) * kMaxWords + w];
        size_t g = 0;
        for (; g < masks.size(); g++)
            if (masks[g] == m) break;
        if (g == masks.size()) {
            masks.push_back(m);
            groups.emplace_back();
        }
        groups[g].push_back(n);
    }
    std::vector<uint64_t> flip, zy;
    std::vector<int> begin;
    std::vector<double> scoef;   // per string: coefficient * (-i)^{#Y}, interleaved
    for (size_t g = 0; g < groups.size(); g++) {
        begin.push_back(int(scoef.size() / 2));
        flip.insert(flip.end(), masks[g].begin(), masks[g].end());
        for (int n : groups[g]) {
            int ny = 0;
            for (int w = 0; w < kMaxWords; w++)
                ny += __builtin_popcountll(~a[size_t(n) * kMaxWords + w] & b[size_t(n) * kMaxWords + w]);
            // (-i)^{#Y}  (PauliString.hpp:193-208)
            cd ph = 1.0;
            if ((ny & 3) > 1) ph = -ph;
            if (ny & 1) ph *= cd(0.0, -1.0);
            const cd c = coef[n] * ph;
            scoef.push_back(c.real());
            scoef.push_back(c.imag());
            for (int w = 0; w < words; w++) zy.push_back(b[size_t(n) * kMaxWords + w]);
        }
    }
    begin.push_back(int(scoef.size() / 2));
    d->num_groups = int(groups.size());
    d->group_flip.upload(flip.data(), flip.size());
    d->group_begin.upload(begin.data(), begin.size());
    d->string_zy.upload(zy.data(), zy.size());
    d->string_coef.upload(scoef.data(), scoef.size());
    auto& ref = *d;
    compiled[words] = std::move(d);
    return ref;
}

const DevicePauliOp& Operator::device_pauli() {
    if (pauli_compiled) return *pauli_compiled;
    if (max_site() >= 64) throw std::invalid_argument("Pauli-string configurations hold at most 64 sites");
    auto d = std::make_unique<DevicePauliOp>();
    d->num_strings = num_strings;
    std::vector<uint64_t> ab(2 * size_t(std::max(num_strings, 1)), 0);
    std::vector<double> c(2 * size_t(std::max(num_strings, 1)), 0.0);
    for (int n = 0; n < num_strings; n++) {
        ab[2 * n] = a[size_t(n) * kMaxWords];
        ab[2 * n + 1] = b[size_t(n) * kMaxWords];
        c[2 * n] = coef[n].real();
        c[2 * n + 1] = coef[n].imag();
    }
    d->ab.upload(ab.data(), ab.size());
    d->coef.upload(c.data(), c.size());
    pauli_compiled = std::move(d);
    return *pauli_compiled;
}

// ------------------------------------------------------------------ sampling

void compute_theta(Psi& psi, const uint64_t* confs, int64_t count, Samples& out, int64_t offset) {
    ScopedTimer timer("theta");
    if (psi.precision == 0) {
        launch_theta_valu<double>(psi.view64(), confs, psi.words, count, out.th64.p + offset * psi.Mpad, S());
    } else {
        auto v = psi.view32();
        launch_theta_mfma(v, psi.wfrag.p, psi.npad, confs, psi.words, count, out.th32.p + offset * psi.Mpad, S());
    }
}

static void require_dense(const Psi& psi) {
    if (!psi.dense())
        throw std::logic_error("this operation needs a PsiRBM or a one-layer PsiDeep (a structured O_k)");
}

void prepare_samples(Psi& psi, Samples& s, int64_t count, int basis) {
    check_basis(psi, basis);
    const bool gen = generic_path(psi, basis);
    s.count = count;
    s.w_uniform = false;
    s.basis = basis;
    s.words = basis == BASIS_PAULIS ? 2 : psi.words;
    s.Mpad = gen ? 0 : psi.Mpad;
    s.precision = psi.precision;
    s.confs.ensure(size_t(count) * s.words);
    s.w.ensure(count);
    s.log_psi.ensure(count);
    if (gen) return;
    if (psi.precision == 0)
        s.th64.ensure(size_t(count) * psi.Mpad);
    else
        s.th32.ensure(size_t(count) * psi.Mpad);
}

static int64_t exact_count(const Psi& psi, const Ensemble& ens) {
    if (int(ens.num_sites) != psi.num_sites)
        throw std::invalid_argument("ExactSummation num_sites does not match the wave function");
    if (ens.basis == BASIS_PAULIS) {
        if (ens.num_sites > 13) throw std::invalid_argument("exact summation over Pauli strings is limited to 13 sites");
        return int64_t(1) << (2 * ens.num_sites);
    }
    if (ens.num_sites > 26) throw std::invalid_argument("exact summation is limited to 26 sites");
    return int64_t(1) << ens.num_sites;
}

// configurations of an exact summation in index order
static void enumerate_into(const Ensemble& ens, Samples& s, int64_t count) {
    if (ens.basis == BASIS_PAULIS)
        launch_enumerate_paulis(s.confs.p, 0, count, S());
    else
        launch_enumerate_confs(s.confs.p, 0, count, s.words, S());
}

// words per chain configuration and the number of sites a move picks from
static void chain_shape(const Psi& psi, const Ensemble& ens, int* words, int* N) {
    *words = ens.basis == BASIS_PAULIS ? 2 : psi.words;
    *N = ens.basis == BASIS_PAULIS ? psi.num_sites : psi.N;
}

// chain set-up shared by the samplers: (re)initialises the chains unless they
// persist from the previous call; returns the thermalisation proposals per chain
static int64_t start_chains(const Psi& psi, Ensemble& ens, int64_t* steps) {
    const int B = int(ens.num_chains);
    if (B <= 0) throw std::invalid_argument("num_markov_chains must be positive");
    *steps = ens.num_samples / ens.num_chains;
    if (*steps <= 0) throw std::invalid_argument("num_samples must be at least num_markov_chains");
    int words = 0, N = 0;
    chain_shape(psi, ens, &words, &N);
    const bool reinit = !ens.persistent || !ens.chains_valid || ens.chain_words != words || ens.chain_N != N;
    ens.chains.ensure(size_t(B) * words);
    ens.acc.ensure(B);
    ANN_HIP_CHECK(hipMemsetAsync(ens.acc.p, 0, sizeof(uint32_t) * B, S()));
    if (reinit) {
        if (ens.move == MOVE_EXCHANGE && ((N + ens.total_z) % 2 != 0 || std::abs(ens.total_z) > N))
            throw std::invalid_argument("total_z incompatible with the number of sites");
        launch_init_chains(ens.chains.p, B, words, N, ens.move, ens.total_z, ens.call_index, ens.seed, S());
        ens.call_index++;
    }
    return reinit ? int64_t(ens.num_therm) * N : 0;
}

static void finish_chains(const Psi& psi, Ensemble& ens, int64_t therm, int64_t steps, int64_t per_step) {
    const int B = int(ens.num_chains);
    ens.chains_valid = true;
    chain_shape(psi, ens, &ens.chain_words, &ens.chain_N);
    // no synchronisation here: the counters are read when the acceptance is asked for
    ens.acc_host.ensure(size_t(B));
    ANN_HIP_CHECK(hipMemcpyAsync(ens.acc_host.p, ens.acc.p, sizeof(uint32_t) * B, hipMemcpyDeviceToHost, S()));
    ens.acc_total = uint64_t(B) * uint64_t(therm + steps * per_step);
    ens.acc_pending = true;
}

void Tdvp::settle() {
    if (!sums_pending) return;
    Runtime::get().sync();
    E = cd(sums_host.p[0], sums_host.p[1]);
    E2 = sums_host.p[2];
    wsum = sums_host.p[3];
    sums_pending = false;
}

void Tdvp::fetch_host_vectors() {
    settle();
    if (!of_pending) return;
    O.resize(P);
    F.resize(P);
    Obar.download(reinterpret_cast<cplx<double>*>(O.data()), P);
    Fd.download(reinterpret_cast<cplx<double>*>(F.data()), P);
    of_pending = false;
}

void Ensemble::settle_acceptance() {
    if (bstat_pending) {
        Runtime::get().sync();
        null_moves = rows_evaluated = 0;
        for (size_t c = 0; c < size_t(num_chains); c++) {
            null_moves += bstat_host.p[2 * c];
            rows_evaluated += bstat_host.p[2 * c + 1];
        }
        bstat_pending = false;
    }
    if (!acc_pending) return;
    Runtime::get().sync();
    uint64_t total_acc = 0;
    for (size_t c = 0; c < size_t(num_chains); c++) total_acc += acc_host.p[c];
    accepted = total_acc;
    rejected = acc_total - total_acc;   // redraws of the current block pattern did not move the chain
    acc_pending = false;
}

// log psi (and |psi|^2 weights when w is given) of configurations, any wave function
void log_psi_of(Psi& psi, const uint64_t* confs, int64_t count, Samples& s, double* w) {
    const cplx<double> lp = {psi.log_prefactor.real(), psi.log_prefactor.imag()};
    if (generic_path(psi, s.basis)) {
        ScopedTimer timer("log_psi");
        with_generic(psi, [&](const auto& v) {
            launch_generic_log_psi(v, confs, s.words, count, lp, s.log_psi.p, w, S());
        }, s.basis);
        return;
    }
    require_dense(psi);
    compute_theta(psi, confs, count, s, 0);
    if (psi.precision == 0)
        launch_net_log_psi<double>(psi.view64(), s.th64.p, count, lp, s.log_psi.p, w, S());
    else
        launch_net_log_psi<float>(psi.view32(), s.th32.p, count, lp, s.log_psi.p, w, S());
}

// caller-given configurations (ENS_LIST): uploaded as they are, with their weights
static void sample_list(Psi& psi, Ensemble& ens) {
    if (int(ens.num_sites) != psi.num_sites)
        throw std::invalid_argument("the sample list was built for another number of sites");
    Samples& s = ens.samples;
    const int64_t count = ens.num_samples;
    prepare_samples(psi, s, count, ens.basis);
    const int words = s.words;
    std::vector<uint64_t> packed(size_t(count) * words);
    for (int64_t i = 0; i < count; i++)
        for (int w = 0; w < words; w++) packed[size_t(i) * words + w] = ens.list_confs[size_t(i) * kMaxWords + w];
    s.confs.upload(packed.data(), packed.size());
    s.w.upload(ens.list_w.data(), size_t(count));
    s.w_uniform = std::all_of(ens.list_w.begin(), ens.list_w.end(), [&](double x) { return x == ens.list_w[0]; });
    log_psi_of(psi, s.confs.p, count, s, nullptr);
}

// PsiCNN, multi-layer PsiDeep, PsiClassical and every wave function on Pauli strings:
// Metropolis with a full forward pass per proposal (the reference keeps no incremental
// state for them, include/quantum_state/PsiCNN.hpp:177-181; PsiCNN uses local updates)
static void sample_generic(Psi& psi, Ensemble& ens) {
    check_basis(psi, ens.basis);
    if (psi.kind == PSI_DEEP) ensure_mlp(psi);
    psi.sync_device();
    if (ens.kind == ENS_LIST) return sample_list(psi, ens);
    Samples& s = ens.samples;
    if (ens.kind == ENS_EXACT) {
        const int64_t count = exact_count(psi, ens);
        prepare_samples(psi, s, count, ens.basis);
        enumerate_into(ens, s, count);
        log_psi_of(psi, s.confs.p, count, s, s.w.p);
        return;
    }
    if (ens.proposal) throw std::logic_error("the learned block proposal needs a PsiRBM or one-layer PsiDeep");
    const int B = int(ens.num_chains);
    int64_t steps = 0;
    const int64_t therm = start_chains(psi, ens, &steps);
    const int64_t count = steps * B;
    prepare_samples(psi, s, count, ens.basis);
    int words = 0, sites = 0;
    chain_shape(psi, ens, &words, &sites);
    for (int64_t step = 0; step < steps; step++) {
        const int64_t nprop = (step == 0 ? therm : 0) + int64_t(ens.num_sweeps) * sites;
        {
            ScopedTimer timer("mc_sweep");
            with_generic(psi, [&](const auto& v) {
                launch_generic_mc(v, ens.chains.p, words, B, ens.move, nprop, ens.counter, ens.seed, ens.acc.p, S());
            }, ens.basis);
        }
        ens.counter += uint64_t(nprop);
        ANN_HIP_CHECK(hipMemcpyAsync(s.confs.p + step * B * words, ens.chains.p, sizeof(uint64_t) * B * words,
                                     hipMemcpyDeviceToDevice, S()));
    }
    launch_fill(s.w.p, 1.0 / double(ens.num_samples), count, S());
    s.w_uniform = true;
    log_psi_of(psi, s.confs.p, count, s, nullptr);
    finish_chains(psi, ens, therm, steps, int64_t(ens.num_sweeps) * sites);
}

void sample(Psi& psi, Ensemble& ens) {
    if (generic_path(psi, ens.basis)) return sample_generic(psi, ens);
    require_dense(psi);
    psi.sync_device();
    if (ens.kind == ENS_LIST) return sample_list(psi, ens);
    Samples& s = ens.samples;
    const int words = psi.words;
    if (ens.kind == ENS_EXACT) {
        const int64_t count = exact_count(psi, ens);
        prepare_samples(psi, s, count);
        launch_enumerate_confs(s.confs.p, 0, count, words, S());
        log_psi_of(psi, s.confs.p, count, s, s.w.p);
        return;
    }
    const int B = int(ens.num_chains);
    int64_t steps = 0;
    const int64_t therm = start_chains(psi, ens, &steps);
    const int64_t count = steps * B;
    prepare_samples(psi, s, count);
    const int N = psi.N;
    // fresh angles of the chains
    Samples chain_view;
    if (psi.precision == 0) {
        ens.chain_th64.ensure(size_t(B) * psi.Mpad);
        chain_view.th64.p = ens.chain_th64.p;
    } else {
        ens.chain_th32.ensure(size_t(B) * psi.Mpad);
        chain_view.th32.p = ens.chain_th32.p;
    }
    compute_theta(psi, ens.chains.p, B, chain_view, 0);
    // with a learned proposal a sample interval is num_sweeps block sweeps
    // (ceil(N / block size) block moves each) plus plain_sweeps plain sweeps
    Proposal* prop = ens.proposal;
    BlockNet bn{};
    int64_t plain_per_step = int64_t(ens.num_sweeps) * N, block_per_step = 0;
    if (prop) {
        if (prop->N != N) throw std::invalid_argument("the proposal was built for another number of sites");
        bn = prop->view();
        bn.exclusive = ens.proposal_exclusive ? 1 : 0;
        plain_per_step = int64_t(ens.plain_sweeps) * N;
        block_per_step = int64_t(ens.num_sweeps) * ((N + prop->bs - 1) / prop->bs);
        ens.bstat.ensure(size_t(2) * B);
        ANN_HIP_CHECK(hipMemsetAsync(ens.bstat.p, 0, sizeof(uint32_t) * 2 * B, S()));
    }
    for (int64_t step = 0; step < steps; step++) {
        const int64_t nprop = (step == 0 ? therm : 0) + plain_per_step;
        if (nprop > 0) {
            ScopedTimer timer("mc_sweep");
            if (psi.precision == 0)
                launch_mc_sweep<double>(psi.view64(), ens.chains.p, ens.chain_th64.p, B, words, ens.move, nprop,
                                        ens.counter, ens.seed, ens.acc.p, S());
            else
                launch_mc_sweep<float>(psi.view32(), ens.chains.p, ens.chain_th32.p, B, words, ens.move, nprop,
                                       ens.counter, ens.seed, ens.acc.p, S());
        }
        ens.counter += uint64_t(nprop);
        if (block_per_step > 0) {
            ScopedTimer timer("block_sweep");
            if (psi.precision == 0)
                launch_block_sweep<double>(psi.view64(), bn, ens.chains.p, ens.chain_th64.p, B, words, block_per_step,
                                           ens.counter, ens.seed, ens.acc.p, ens.bstat.p, S());
            else
                launch_block_sweep<float>(psi.view32(), bn, ens.chains.p, ens.chain_th32.p, B, words, block_per_step,
                                          ens.counter, ens.seed, ens.acc.p, ens.bstat.p, S());
            ens.counter += uint64_t(block_per_step);
        }
        ANN_HIP_CHECK(hipMemcpyAsync(s.confs.p + step * B * words, ens.chains.p, sizeof(uint64_t) * B * words,
                                     hipMemcpyDeviceToDevice, S()));
        compute_theta(psi, ens.chains.p, B, s, step * B);
        // restart the next sweep from exact angles (after the last step the next call
        // recomputes the chains' angles anyway)
        if (step + 1 == steps) break;
        if (psi.precision == 0)
            ANN_HIP_CHECK(hipMemcpyAsync(ens.chain_th64.p, s.th64.p + step * B * psi.Mpad,
                                         sizeof(cplx<double>) * B * psi.Mpad, hipMemcpyDeviceToDevice, S()));
        else
            ANN_HIP_CHECK(hipMemcpyAsync(ens.chain_th32.p, s.th32.p + step * B * psi.Mpad,
                                         sizeof(cplx<float>) * B * psi.Mpad, hipMemcpyDeviceToDevice, S()));
    }
    chain_view.th64.p = nullptr;
    chain_view.th32.p = nullptr;
    launch_fill(s.w.p, 1.0 / double(ens.num_samples), count, S());
    s.w_uniform = true;
    const cplx<double> lp = {psi.log_prefactor.real(), psi.log_prefactor.imag()};
    if (psi.precision == 0)
        launch_net_log_psi<double>(psi.view64(), s.th64.p, count, lp, s.log_psi.p, nullptr, S());
    else
        launch_net_log_psi<float>(psi.view32(), s.th32.p, count, lp, s.log_psi.p, nullptr, S());
    finish_chains(psi, ens, therm, steps, plain_per_step + block_per_step);
    ens.block_proposals = uint64_t(B) * uint64_t(steps * block_per_step);
    if (block_per_step > 0) {
        ens.bstat_host.ensure(size_t(2) * B);
        ANN_HIP_CHECK(hipMemcpyAsync(ens.bstat_host.p, ens.bstat.p, sizeof(uint32_t) * 2 * B, hipMemcpyDeviceToHost,
                                     S()));
        ens.bstat_pending = true;
    } else {
        ens.bstat_pending = false;
        ens.null_moves = ens.rows_evaluated = 0;
    }
}

void local_energies(Psi& psi, Operator& op, Samples& s) {
    if (op.max_site() >= psi.num_sites) throw std::invalid_argument("operator acts on sites beyond num_sites");
    s.e_loc.ensure(s.count);
    if (s.basis == BASIS_PAULIS) {
        const DevicePauliOp& d = op.device_pauli();
        ScopedTimer timer("local_energy");
        with_generic(psi, [&](const auto& v) {
            using V = std::decay_t<decltype(v)>;
            using T = typename V::real_t;
            if constexpr (std::is_same<V, ClassicalView<T>>::value || std::is_same<V, PauliInput<MlpView<T>>>::value)
                launch_generic_local_energy_pauli(v, d.view(), s.confs.p, s.count, s.e_loc.p, S());
            else
                throw std::logic_error("this wave function does not read Pauli strings");
        }, s.basis);
        return;
    }
    const DeviceOp& d = op.device(psi.words);
    ScopedTimer timer("local_energy");
    if (generic_path(psi, s.basis)) {
        with_generic(psi, [&](const auto& v) {
            launch_generic_local_energy(v, d.view(), s.confs.p, s.count, s.e_loc.p, S());
        });
        return;
    }
    if (psi.precision == 0)
        launch_local_energy<double>(psi.view64(), d.view(), s.confs.p, s.th64.p, s.count, s.e_loc.p, S());
    else
        launch_local_energy<float>(psi.view32(), d.view(), s.confs.p, s.th32.p, s.count, s.e_loc.p, S());
}

// ----------------------------------------------------------------------- TDVP

namespace {

template <typename T>
struct Prec;
template <>
struct Prec<double> {
    static DevBuf<cplx<double>>& th(Samples& s) { return s.th64; }
    static DevBuf<cplx<double>>& g(Tdvp& t) { return t.g64; }
    static DenseView<double> view(Psi& p) { return p.view64(); }
};
template <>
struct Prec<float> {
    static DevBuf<cplx<float>>& th(Samples& s) { return s.th32; }
    static DevBuf<cplx<float>>& g(Tdvp& t) { return t.g32; }
    static DenseView<float> view(Psi& p) { return p.view32(); }
};

// g = f tanh(theta) of `count` samples into g, whose rows are padded to
// sr_count_pad(count) with zeros (the matrix-core kernels read whole 256-sample rounds)
template <typename T>
void g_factor_padded(const DenseView<T>& net, const cplx<T>* theta, int64_t count, DevBuf<cplx<T>>& g) {
    const int64_t pad = sr_count_pad(count);
    g.ensure(size_t(pad) * net.Mpad);
    launch_g_factor<T>(net, theta, count, g.p, S());
    if (pad > count)
        ANN_HIP_CHECK(hipMemsetAsync(g.p + size_t(count) * net.Mpad, 0, sizeof(cplx<T>) * (pad - count) * net.Mpad, S()));
}

// weighted sums of the structured O, packed in parameter order: out = sum_s coef_s O_s (or conj O_s)
template <typename T>
void structured_sums(Tdvp& t, const DenseView<T>& net, const uint64_t* confs, int words, const cplx<T>* g,
                     int conj_g, const cplx<double>* coef, int64_t count, cplx<double>* out, double shift,
                     const cplx<double>* v) {
    t.Yw.ensure(size_t(net.N) * net.Mpad);
    t.Yb.ensure(net.Mpad);
    t.Ya.ensure(net.N);
    ScopedTimer timer("spin_sums");
    if constexpr (sizeof(T) == 4) {
        // the transposed spin bits of this evaluation's samples are built once and
        // shared with the S.v products and the K-FAC set-up; the workspace persists
        const SpinPlan plan = spin_plan(net.N, net.M, net.Mpad, count, true);
        if (!t.bits_valid) {
            t.spin_bits.ensure(size_t(plan.bit_blocks) * plan.vpad);
            launch_spin_bits(confs, words, count, net.N, plan.vpad, t.spin_bits.p, S());
            t.bits_valid = true;
        }
        t.ss_partial.ensure(size_t(plan.chunks) * plan.n);
        t.ss_full.ensure(size_t(plan.n));
        t.ss_c32.ensure(size_t(sr_count_pad(count)));
        const SpinSumsWork ws{t.spin_bits.p, t.ss_partial.p, t.ss_full.p, t.ss_c32.p};
        launch_spin_sums_bf16(net, confs, words, g, conj_g, coef, count, t.Yw.p, t.Yb.p, t.Ya.p, S(), &ws);
    } else
        launch_spin_weighted_sums<T>(net, confs, words, g, conj_g, coef, count, t.Yw.p, t.Yb.p, t.Ya.p, S());
    if (out)
        launch_pack_y(net.layout, net.N, net.M, net.Mpad, net.C, net.conn, t.Ya.p, t.Yb.p, t.Yw.p, v, shift,
                      net.num_params, out, S());
}

// K-FAC factor buffers of the Tdvp's current sizes
KfacWork kfac_work(Tdvp& t) {
    t.kf_spart.ensure(kfac_spart_elems(t.N, t.count));
    t.kf_gpart.ensure(kfac_gpart_elems(t.M, t.count));
    t.kf_C.ensure(size_t(t.N) * t.N);
    t.kf_Ci.ensure(size_t(t.N) * t.N);
    t.kf_D.ensure(kfac_dblock_elems(t.M));
    t.kf_gam.ensure(2);
    return {t.kf_spart.p, t.kf_gpart.p, t.kf_C.p, t.kf_Ci.p, t.kf_D.p, t.kf_gam.p};
}

// The statistics that do not need E_loc (<O>, the diagonal sums of S) and, when the
// last solve of this Tdvp used K-FAC, the factor moments of the new samples run on the
// side stream while the local energies run on the library stream.
template <typename T>
void tdvp_eval_t(Tdvp& t, Operator& op, Psi& psi, Ensemble& ens, bool with_S) {
    sample(psi, ens);
    t.bits_valid = false;   // new samples: their transposed spin bits are built on first use
    t.kf_ready = false;
    t.kf_inverted = false;
    Samples& s = ens.samples;
    const int64_t n = s.count;
    const int P = psi.num_params;
    auto net = Prec<T>::view(psi);
    auto& g = Prec<T>::g(t);
    {
        ScopedTimer timer("o_k");
        g_factor_padded<T>(net, Prec<T>::th(s).p, n, g);
    }
    // the evaluation's description of the structured O (for S.v, CG and K-FAC); the
    // connection tables are uploaded when the connections change
    t.precision = psi.precision;
    t.words = s.words;
    t.N = psi.N;
    t.M = psi.M;
    t.Mpad = psi.Mpad;
    t.C = psi.C;
    t.layout = psi.layout;
    t.npad = psi.npad;
    t.count = n;
    if (psi.kind != PSI_DEEP) {
        t.conn_shift = -1;
        t.conn_host.clear();
    } else if (t.conn_host != psi.conn || t.conn_mpad != psi.Mpad || t.conn_n != psi.N) {
        t.conn_host = psi.conn;
        t.conn_mpad = psi.Mpad;
        t.conn_n = psi.N;
        t.conn.upload(psi.conn.data(), psi.conn.size());
        std::vector<int> inv(size_t(psi.N) * psi.Mpad, -1);
        for (int i = 0; i < psi.C; i++)
            for (int j = 0; j < psi.M; j++) {
                int& slot = inv[size_t(psi.conn[size_t(i) * psi.M + j]) * psi.Mpad + j];
                slot = slot == -1 ? i : -2;
            }
        t.inv.upload(inv.data(), inv.size());
        t.inv_unique = std::find(inv.begin(), inv.end(), -2) == inv.end();
        // circulant layout: connection i of unit j is site (j d + i) mod N
        const int d = psi.M > 1 ? int(psi.conn[1]) : 0;
        bool circ = psi.C <= psi.N;
        for (int i = 0; i < psi.C && circ; i++)
            for (int j = 0; j < psi.M; j++)
                if (int(psi.conn[size_t(i) * psi.M + j]) != int((int64_t(j) * d + i) % psi.N)) {
                    circ = false;
                    break;
                }
        t.conn_shift = circ ? d : -1;
    }
    DevBuf<cplx<double>>& coef = t.coef;
    coef.ensure(n);
    t.Obar.ensure(P);
    t.Fd.ensure(P);
    t.diag.ensure(P);
    t.sums.ensure(4);
    launch_weight_coef(s.w.p, nullptr, n, coef.p, S());
    const bool pre_kfac = t.kfac_hint && psi.N <= kKfacMaxSites && psi.M <= kKfacMaxUnits &&
                          (psi.layout == LAYOUT_RBM || t.inv_unique);
    KfacWork kw{};
    if (pre_kfac) {
        kw = kfac_work(t);
        t.kf_idx.ensure(size_t(t.N) * t.M);
    }
    if (sizeof(T) == 4 || pre_kfac) {
        // the transposed spin bits (fp32 statistics and S.v, the K-FAC spin covariance)
        // are shared by both streams' kernels: built before the fork
        const SpinPlan plan = spin_plan(net.N, net.M, net.Mpad, n, true);
        t.spin_bits.ensure(size_t(plan.bit_blocks) * plan.vpad);
        launch_spin_bits(s.confs.p, s.words, n, net.N, plan.vpad, t.spin_bits.p, S());
        t.bits_valid = true;
    }
    ANN_HIP_CHECK(hipEventRecord(t.fork_ev.get(), S()));
    {
        SideStream side(0);
        ANN_HIP_CHECK(hipStreamWaitEvent(S(), t.fork_ev.get(), 0));
        structured_sums<T>(t, net, s.confs.p, s.words, g.p, 0, coef.p, n, t.Obar.p, 0.0, nullptr);
        // diagonal of S for the Jacobi preconditioner: sum_s w |g_sj|^2 - |<O_k>|^2
        structured_sums<T>(t, net, s.confs.p, s.words, g.p, 2, coef.p, n, nullptr, 0.0, nullptr);
        ANN_HIP_CHECK(hipEventRecord(t.join_ev.get(), S()));
    }
    if (pre_kfac) {
        // the K-FAC set-up on a third stream: its few-block, latency-bound kernels run
        // beside both the local energies and the wide statistics kernels
        SideStream side(1);
        ANN_HIP_CHECK(hipStreamWaitEvent(S(), t.fork_ev.get(), 0));
        {
            ScopedTimer timer("kfac_setup");
            const SpinPlan plan = spin_plan(t.N, t.M, t.Mpad, n, t.layout == LAYOUT_DEEP1);
            launch_kfac_moments(s.confs.p, s.words, t.spin_bits.p, plan.vpad, s.w_uniform,
                                t.precision != 0 ? t.g32.p : nullptr, t.precision != 0 ? nullptr : t.g64.p, t.Mpad,
                                s.w.p, n, t.N, t.M, kw, S());
            // and the inverses for the shift of that solve (used when the next solve
            // asks for the same shift)
            launch_kfac_index(t.layout, t.N, t.M, t.Mpad, t.layout == LAYOUT_DEEP1 ? t.inv.p : nullptr, t.kf_idx.p,
                              S());
            launch_kfac_finish(t.N, t.M, t.kfac_shift, kw, S());
            t.kf_inverted = true;
            t.kf_inv_shift = t.kfac_shift;
        }
        ANN_HIP_CHECK(hipEventRecord(t.join2_ev.get(), S()));
    }
    local_energies(psi, op, s);
    // E, E2 and sum w stay on the device until the end of the evaluation
    launch_weighted_sums(s.e_loc.p, s.w.p, n, t.sums.p, S());
    ANN_HIP_CHECK(hipStreamWaitEvent(S(), t.join_ev.get(), 0));
    if (pre_kfac) ANN_HIP_CHECK(hipStreamWaitEvent(S(), t.join2_ev.get(), 0));
    launch_sr_diag(net.layout, net.N, net.M, net.Mpad, t.sums.p + 3, t.Yb.p, t.Obar.p, P, t.diag.p, S());
    // F = sum w conj(O) E_loc - E conj(<O>) (TDVP.cu.template:300) = sum w (E_loc - E) conj(O)
    launch_weight_coef_centered_dev(s.w.p, s.e_loc.p, t.sums.p, n, coef.p, S());
    structured_sums<T>(t, net, s.confs.p, s.words, g.p, 1, coef.p, n, t.Fd.p, 0.0, nullptr);
    // E, E2 and sum w go to page-locked memory; the host reads them, and <O> and F,
    // when they are asked for (the evaluation returns without synchronising, so a
    // solve that follows is queued behind it at once)
    t.sums_host.ensure(4);
    ANN_HIP_CHECK(hipMemcpyAsync(t.sums_host.p, t.sums.p, sizeof(double) * 4, hipMemcpyDeviceToHost, S()));
    t.sums_pending = true;
    t.of_pending = true;
    t.confs.ensure(size_t(n) * s.words);
    ANN_HIP_CHECK(hipMemcpyAsync(t.confs.p, s.confs.p, sizeof(uint64_t) * n * s.words, hipMemcpyDeviceToDevice, S()));
    t.w.ensure(n);
    ANN_HIP_CHECK(hipMemcpyAsync(t.w.p, s.w.p, sizeof(double) * n, hipMemcpyDeviceToDevice, S()));
    t.e_loc.ensure(n);
    ANN_HIP_CHECK(hipMemcpyAsync(t.e_loc.p, s.e_loc.p, sizeof(cplx<double>) * n, hipMemcpyDeviceToDevice, S()));
    t.generic = false;
    t.w_uniform = s.w_uniform;
    t.has_samples = true;
    t.has_S = false;
    if (with_S) {
        DevBuf<cplx<T>> O;
        O.ensure(size_t(n) * P);
        {
            ScopedTimer timer("o_k");
            launch_materialize_ok<T>(net, s.confs.p, s.words, g.p, n, t.Obar.p, O.p, S());
        }
        t.S.ensure(size_t(P) * P);
        {
            ScopedTimer timer("gram");
            launch_gram<T>(O.p, s.w.p, n, P, t.S.p, S());
        }
        // reference S = sum w conj(O) O - conj(<O>) <O>; with centred O this needs
        // (1 - sum w) conj(<O>) <O> added back (zero for normalised weights)
        t.settle();
        if (std::abs(1.0 - t.wsum) > 0.0) launch_rank1(t.S.p, t.Obar.p, P, 1.0 - t.wsum, S());
        Runtime::get().sync();
        t.has_S = true;
    }
}

template <typename T>
DevBuf<cplx<T>>& materialised(Tdvp& t) {
    if constexpr (sizeof(T) == 8)
        return t.Og64;
    else
        return t.Og32;
}

// O_k of a wave function on the generic path for `count` configurations into O [count][P]
template <typename T>
void generic_ok(Psi& psi, const uint64_t* confs, int words, int basis, int64_t count, cplx<T>* O) {
    ScopedTimer timer("o_k");
    with_generic(psi, [&](const auto& v) {
        using V = std::decay_t<decltype(v)>;
        if constexpr (std::is_same<typename V::real_t, T>::value)
            launch_generic_ok(v, confs, words, count, nullptr, O, S());
        else
            throw std::logic_error("precision mismatch");
    }, basis);
}

// TDVP statistics from a materialised O (TDVP.cu.template:77-126,181-302)
template <typename T>
void tdvp_eval_generic(Tdvp& t, Operator& op, Psi& psi, Ensemble& ens, bool with_S) {
    sample(psi, ens);
    Samples& s = ens.samples;
    local_energies(psi, op, s);
    const int64_t n = s.count;
    const int P = psi.num_params;
    DevBuf<double> sums;
    sums.ensure(4);
    launch_weighted_sums(s.e_loc.p, s.w.p, n, sums.p, S());
    auto& O = materialised<T>(t);
    O.ensure(size_t(n) * P);
    generic_ok<T>(psi, s.confs.p, s.words, s.basis, n, O.p);
    DevBuf<cplx<double>> coef, sq;
    coef.ensure(n);
    sq.ensure(P);
    t.Obar.ensure(P);
    t.Fd.ensure(P);
    t.diag.ensure(P);
    t.part.ensure(size_t(col_sum_partials(n)) * P);
    {
        ScopedTimer timer("spin_sums");
        launch_weight_coef(s.w.p, nullptr, n, coef.p, S());
        launch_ok_colsum<T>(O.p, n, P, coef.p, 0, 0.0, nullptr, t.part.p, t.Obar.p, S());
        launch_ok_colsum<T>(O.p, n, P, coef.p, 2, 0.0, nullptr, t.part.p, sq.p, S());
        launch_diag_from_sums(sq.p, t.Obar.p, P, t.diag.p, S());
    }
    double h[4];
    ANN_HIP_CHECK(hipMemcpyAsync(h, sums.p, sizeof(h), hipMemcpyDeviceToHost, S()));
    Runtime::get().sync();
    t.E = cd(h[0], h[1]);
    t.E2 = h[2];
    t.wsum = h[3];
    t.sums_pending = false;
    t.of_pending = false;
    {
        // F = sum w (E_loc - E) conj(O)  (TDVP.cu.template:300 without the cancellation)
        ScopedTimer timer("spin_sums");
        launch_weight_coef_centered(s.w.p, s.e_loc.p, {t.E.real(), t.E.imag()}, n, coef.p, S());
        launch_ok_colsum<T>(O.p, n, P, coef.p, 1, 0.0, nullptr, t.part.p, t.Fd.p, S());
    }
    t.O.resize(P);
    t.F.resize(P);
    t.Obar.download(reinterpret_cast<cplx<double>*>(t.O.data()), P);
    t.Fd.download(reinterpret_cast<cplx<double>*>(t.F.data()), P);
    t.generic = true;
    t.precision = psi.precision;
    t.words = s.words;
    t.N = psi.N;
    t.count = n;
    t.w.ensure(n);
    ANN_HIP_CHECK(hipMemcpyAsync(t.w.p, s.w.p, sizeof(double) * n, hipMemcpyDeviceToDevice, S()));
    t.confs.ensure(size_t(n) * s.words);
    ANN_HIP_CHECK(hipMemcpyAsync(t.confs.p, s.confs.p, sizeof(uint64_t) * n * s.words, hipMemcpyDeviceToDevice, S()));
    t.e_loc.ensure(n);
    ANN_HIP_CHECK(hipMemcpyAsync(t.e_loc.p, s.e_loc.p, sizeof(cplx<double>) * n, hipMemcpyDeviceToDevice, S()));
    t.w_uniform = s.w_uniform;
    t.has_samples = true;
    t.has_S = false;
    if (with_S) {
        t.S.ensure(size_t(P) * P);
        {
            ScopedTimer timer("gram");
            launch_gram<T>(O.p, s.w.p, n, P, t.S.p, S());
        }
        // S = sum w conj(O) O - conj(<O>) <O>
        launch_rank1(t.S.p, t.Obar.p, P, -1.0, S());
        t.has_S = true;
    }
    Runtime::get().sync();
}

// y = S.v + shift v = sum_s conj(O_s) w_s (O_s.v - <O>.v) + shift v
template <typename T>
void tdvp_S_dot_generic(Tdvp& t, const cplx<double>* v, double shift, cplx<double>* y) {
    const int P = int(t.P);
    auto& O = materialised<T>(t);
    t.u.ensure(t.count);
    t.c1.ensure(2);
    t.dot_scratch.ensure(256);
    t.part.ensure(size_t(col_sum_partials(t.count)) * P);
    ScopedTimer timer("sr_matvec");
    launch_ok_rowdot<T>(O.p, t.count, P, v, t.u.p, S());
    launch_dot2(t.Obar.p, v, nullptr, nullptr, P, 0, t.dot_scratch.p, reinterpret_cast<double*>(t.c1.p), S());
    launch_center_scale(t.u.p, t.w.p, reinterpret_cast<const double*>(t.c1.p), t.count, S());
    launch_ok_colsum<T>(O.p, t.count, P, t.u.p, 1, shift, v, t.part.p, y, S());
}

template <typename T>
struct Ws;
template <>
struct Ws<float> {
    static DevBuf<cplx<float>>& V(Tdvp& t) { return t.V32; }
    static DevBuf<cplx<float>>& SV(Tdvp& t) { return t.SV32; }
    static DevBuf<cplx<float>>& zero(Tdvp& t) { return t.zero32; }
};
template <>
struct Ws<double> {
    static DevBuf<cplx<double>>& V(Tdvp& t) { return t.V64; }
    static DevBuf<cplx<double>>& SV(Tdvp& t) { return t.SV64; }
    static DevBuf<cplx<double>>& zero(Tdvp& t) { return t.zero64; }
};

// y = (S + shift) v for the structured O of a dense net.  The fp32 path runs the
// fused bf16 matrix-core pipeline and also leaves the partials of v^H y (*ypart).
template <typename T>
void tdvp_S_dot_t(Tdvp& t, const cplx<double>* v, double shift, cplx<double>* y, const double** ypart,
                  int* yparts, const CgDir* dir = nullptr) {
    const int P = int(t.P);
    DenseView<T> net{};
    net.N = t.N;
    net.M = t.M;
    net.Mpad = t.Mpad;
    net.C = t.C;
    net.layout = t.layout;
    net.num_params = P;
    net.conn = t.layout == LAYOUT_DEEP1 ? t.conn.p : nullptr;
    auto& g = Prec<T>::g(t);
    t.va.ensure(t.N);
    t.vb.ensure(t.Mpad);
    t.u.ensure(t.count);
    if constexpr (sizeof(T) == 4) {
        const bool with_coef = t.layout == LAYOUT_DEEP1;
        const SpinPlan plan = spin_plan(t.N, t.M, t.Mpad, t.count, with_coef);
        t.frag.ensure(sr_frag_elems(t.N, t.Mpad));
        t.upart.ensure(size_t(sr_ov_pairs(t.M)) * t.count);
        t.u32.ensure(size_t(sr_count_pad(t.count)));
        t.spin_part.ensure(size_t(plan.chunks) * plan.n);
        t.opart.ensure(256);
        t.ypart.ensure(512);
        if (!t.bits_valid) {
            // transposed spins of the samples, reused by every matvec until the next eval
            t.spin_bits.ensure(size_t(plan.bit_blocks) * plan.vpad);
            launch_spin_bits(t.confs.p, t.words, t.count, t.N, plan.vpad, t.spin_bits.p, S());
            t.bits_valid = true;
        }
        const SrWorkspace ws{t.frag.p, t.va.p, t.vb.p, t.upart.p, t.u32.p, t.spin_part.p, t.opart.p, t.ypart.p};
        const WParamMap wm = t.order();
        const cplx<double>* obar = t.solve_sm ? t.obar_sm.p : t.Obar.p;
        // timing events between the launches stall the queue for microseconds, so only
        // every 16th product is timed (averages per launch stay exact)
        const bool timed = t.matvecs++ % 16 == 0;
        {
            ScopedTimer timer("sr_pack", timed);
            launch_sr_pack(net, wm, t.inv.p, v, dir ? *dir : CgDir{}, obar, ws, S());
        }
        {
            ScopedTimer timer("sr_ov", timed);
            launch_sr_ov32(net, t.confs.p, t.words, t.count, g.p, ws, S());
        }
        {
            ScopedTimer timer("sr_finish", timed);
            launch_sr_finish(net, t.w.p, t.confs.p, t.words, t.count, ws, S());
        }
        {
            ScopedTimer timer("spin_gemm", timed);
            launch_spin_gemm32(net, t.spin_bits.p, g.p, 1, ws.u, with_coef ? 1 : 0, t.count, plan, ws.spin_part,
                               S());
        }
        {
            ScopedTimer timer("sr_y", timed);
            launch_sr_y_bf16(net, wm, v, shift, plan, ws, y, S());
        }
        if (ypart) *ypart = t.ypart.p;
        if (yparts) *yparts = sr_y_parts(net, wm);
        return;
    } else {
        auto& V = Ws<T>::V(t);
        auto& SV = Ws<T>::SV(t);
        auto& zero = Ws<T>::zero(t);
        t.c1.ensure(2);
        t.dot_scratch.ensure(256);
        V.ensure(size_t(t.N) * t.Mpad);
        SV.ensure(size_t(t.count) * t.Mpad);
        if (zero.n < size_t(t.Mpad)) {
            zero.ensure(t.Mpad);
            ANN_HIP_CHECK(hipMemsetAsync(zero.p, 0, sizeof(cplx<T>) * t.Mpad, S()));
        }
        {
            ScopedTimer timer("sr_matvec");
            launch_unpack_v<T>(net, t.inv.p, v, t.va.p, t.vb.p, V.p, S());
            DenseView<T> vnet = net;
            vnet.Wd = V.p;
            vnet.b = zero.p;
            const cplx<double>* vb = t.layout == LAYOUT_DEEP1 ? t.vb.p : nullptr;
            const cplx<double>* va = t.layout == LAYOUT_DEEP1 ? t.va.p : nullptr;
            launch_theta_valu<T>(vnet, t.confs.p, t.words, t.count, SV.p, S());
            launch_structured_Ov<T>(net, t.confs.p, t.words, g.p, SV.p, vb, va, t.count, t.u.p, S());
            launch_dot2(t.Obar.p, v, nullptr, nullptr, P, 0, t.dot_scratch.p, reinterpret_cast<double*>(t.c1.p),
                        S());
            launch_center_scale(t.u.p, t.w.p, reinterpret_cast<const double*>(t.c1.p), t.count, S());
        }
        structured_sums<T>(t, net, t.confs.p, t.words, g.p, 1, t.u.p, t.count, y, shift, v);
        if (ypart) *ypart = nullptr;
        if (yparts) *yparts = 0;
    }
}

}  // namespace

namespace {

template <typename T>
void ok_weighted_sum_t(Psi& psi, Samples& s, const cplx<double>* coef, int conj, cplx<double>* out) {
    const int P = psi.num_params;
    if (generic_path(psi, s.basis)) {
        // materialise O in row chunks of at most 256 MiB, accumulating into out
        const int64_t rows = std::max<int64_t>(1, (int64_t(256) << 20) / (int64_t(sizeof(cplx<T>)) * P));
        DevBuf<cplx<T>> O;
        DevBuf<cplx<double>> part;
        O.ensure(size_t(std::min(rows, s.count)) * P);
        part.ensure(size_t(col_sum_partials(std::min(rows, s.count))) * P);
        ANN_HIP_CHECK(hipMemsetAsync(out, 0, sizeof(cplx<double>) * P, S()));
        for (int64_t r0 = 0; r0 < s.count; r0 += rows) {
            const int64_t m = std::min(rows, s.count - r0);
            generic_ok<T>(psi, s.confs.p + r0 * s.words, s.words, s.basis, m, O.p);
            launch_ok_colsum<T>(O.p, m, P, coef + r0, conj, 1.0, out, part.p, out, S());
        }
        Runtime::get().sync();
        return;
    }
    require_dense(psi);
    DenseView<T> net = Prec<T>::view(psi);
    DevBuf<cplx<T>> g;
    g_factor_padded<T>(net, Prec<T>::th(s).p, s.count, g);
    DevBuf<cplx<double>> Yw, Yb, Ya;
    Yw.ensure(size_t(psi.N) * psi.Mpad);
    Yb.ensure(psi.Mpad);
    Ya.ensure(psi.N);
    if constexpr (sizeof(T) == 4)
        launch_spin_sums_bf16(net, s.confs.p, s.words, g.p, conj, coef, s.count, Yw.p, Yb.p, Ya.p, S());
    else
        launch_spin_weighted_sums<T>(net, s.confs.p, s.words, g.p, conj, coef, s.count, Yw.p, Yb.p, Ya.p, S());
    launch_pack_y(net.layout, net.N, net.M, net.Mpad, net.C, net.conn, Ya.p, Yb.p, Yw.p, nullptr, 0.0,
                  psi.num_params, out, S());
    Runtime::get().sync();
}

}  // namespace

void ok_weighted_sum(Psi& psi, Samples& s, const cplx<double>* coef, int conj, cplx<double>* out) {
    if (psi.precision == 0)
        ok_weighted_sum_t<double>(psi, s, coef, conj, out);
    else
        ok_weighted_sum_t<float>(psi, s, coef, conj, out);
}

namespace {

// out_i[k] = sum_s coef_i[s] op_i(O[s][k]) for several coefficient vectors at once, O
// materialised in row chunks of at most 256 MiB (op: identity, conjugate or |.|^2)
template <typename T>
void ok_column_sums_t(Psi& psi, Samples& s, int nsum, const cplx<double>* const* coef, const int* mode,
                      cplx<double>* const* out) {
    const int P = psi.num_params;
    const bool gen = generic_path(psi, s.basis);
    DevBuf<cplx<T>> g;
    DenseView<T> net{};
    if (!gen) {
        net = Prec<T>::view(psi);
        g_factor_padded<T>(net, Prec<T>::th(s).p, s.count, g);
    }
    const int64_t rows = std::max<int64_t>(1, (int64_t(256) << 20) / (int64_t(sizeof(cplx<T>)) * P));
    DevBuf<cplx<T>> O;
    DevBuf<cplx<double>> part;
    O.ensure(size_t(std::min(rows, s.count)) * P);
    part.ensure(size_t(col_sum_partials(std::min(rows, s.count))) * P);
    for (int i = 0; i < nsum; i++) ANN_HIP_CHECK(hipMemsetAsync(out[i], 0, sizeof(cplx<double>) * P, S()));
    for (int64_t r0 = 0; r0 < s.count; r0 += rows) {
        const int64_t m = std::min(rows, s.count - r0);
        if (gen)
            generic_ok<T>(psi, s.confs.p + r0 * s.words, s.words, s.basis, m, O.p);
        else
            launch_materialize_ok<T>(net, s.confs.p + r0 * s.words, s.words, g.p + r0 * net.Mpad, m, nullptr, O.p,
                                     S());
        for (int i = 0; i < nsum; i++)
            launch_ok_colsum<T>(O.p, m, P, coef[i] + r0, mode[i], 1.0, out[i], part.p, out[i], S());
    }
    Runtime::get().sync();
}

}  // namespace

cd expectation_exp_fast(Operator& op, Psi& psi, Ensemble& ens) {
    if (op.max_site() >= psi.num_sites) throw std::invalid_argument("operator acts on sites beyond num_sites");
    sample(psi, ens);
    Samples& s = ens.samples;
    DevBuf<cplx<double>> e;
    DevBuf<double> sums;
    e.ensure(s.count);
    sums.ensure(4);
    if (s.basis == BASIS_PAULIS)
        launch_exp_fast_energy_pauli(op.device_pauli().view(), s.confs.p, s.count, e.p, S());
    else
        launch_exp_fast_energy(op.device(s.words).view(), s.confs.p, s.count, e.p, S());
    launch_weighted_sums(e.p, s.w.p, s.count, sums.p, S());
    double h[4];
    ANN_HIP_CHECK(hipMemcpyAsync(h, sums.p, sizeof(h), hipMemcpyDeviceToHost, S()));
    Runtime::get().sync();
    return {h[0], h[1]};
}

cd expectation_reweighted(Operator& op, Psi& psi, Psi& psi_sampling, Ensemble& ens) {
    if (psi.num_sites != psi_sampling.num_sites) throw std::invalid_argument("psi and psi_sampling differ in num_sites");
    if (op.max_site() >= psi.num_sites) throw std::invalid_argument("operator acts on sites beyond num_sites");
    // samples and log psi of psi_sampling; log psi and E_loc of psi on the same configurations
    sample(psi_sampling, ens);
    Samples& s = ens.samples;
    const int64_t n = s.count;
    psi.sync_device();
    Samples sc;
    prepare_samples(psi, sc, n, s.basis);
    ANN_HIP_CHECK(hipMemcpyAsync(sc.confs.p, s.confs.p, sizeof(uint64_t) * n * s.words, hipMemcpyDeviceToDevice,
                                 S()));
    log_psi_of(psi, sc.confs.p, n, sc, nullptr);
    local_energies(psi, op, sc);
    DevBuf<double> wr, sums;
    wr.ensure(n);
    sums.ensure(4);
    launch_reweight(s.w.p, sc.log_psi.p, s.log_psi.p, n, wr.p, S());
    launch_weighted_sums(sc.e_loc.p, wr.p, n, sums.p, S());
    double h[4];
    ANN_HIP_CHECK(hipMemcpyAsync(h, sums.p, sizeof(h), hipMemcpyDeviceToHost, S()));
    Runtime::get().sync();
    if (!(h[3] > 0.0)) throw std::runtime_error("ExpectationValue: the reweighted samples carry no weight");
    return cd(h[0], h[1]) / h[3];
}

double kullback_leibler(Kl& kl, Psi& psi, Psi& psi_prime, Ensemble& ens, double threshold, int mode, double nu,
                        cd* grad, double* noise) {
    if (psi.num_sites != psi_prime.num_sites) throw std::invalid_argument("psi and psi_prime differ in num_sites");
    if (mode > 0 && psi_prime.num_params != kl.P)
        throw std::invalid_argument("KullbackLeibler num_params does not match psi_prime");
    if (!std::isfinite(threshold)) throw std::invalid_argument("threshold must be finite");
    // samples of psi' with log psi'; log psi of the classical network on the same configurations
    sample(psi_prime, ens);
    Samples& s = ens.samples;
    const int64_t n = s.count;
    psi.sync_device();
    Samples sc;
    prepare_samples(psi, sc, n, s.basis);
    ANN_HIP_CHECK(hipMemcpyAsync(sc.confs.p, s.confs.p, sizeof(uint64_t) * n * s.words, hipMemcpyDeviceToDevice,
                                 S()));
    log_psi_of(psi, sc.confs.p, n, sc, nullptr);
    DevBuf<cplx<double>> c_w, c_d, c_d2;
    DevBuf<double> sums;
    c_w.ensure(n);
    c_d.ensure(n);
    c_d2.ensure(n);
    sums.ensure(6);
    const cplx<double> last = {kl.last_mean_deviation.real(), kl.last_mean_deviation.imag()};
    launch_kl_terms(sc.log_psi.p, s.log_psi.p, s.w.p, n, kl.log_psi_scale, last, threshold, c_w.p, c_d.p, c_d2.p,
                    sums.p, S());
    double h[6];
    ANN_HIP_CHECK(hipMemcpyAsync(h, sums.p, sizeof(h), hipMemcpyDeviceToHost, S()));
    Runtime::get().sync();
    const double W = h[0];
    if (!(W > 0.0)) throw std::runtime_error("KullbackLeibler: the samples carry no weight");
    kl.total_weight = W;
    kl.mean_deviation = cd(h[1], h[2]) / W;
    kl.last_mean_deviation = kl.mean_deviation;   // KullbackLeibler.cu.template:191-197
    const cd d = cd(h[3], h[4]) / W;
    const double d2 = h[5] / W;
    const double value = std::sqrt(std::max(1e-8, d2 - std::norm(d)));
    if (mode == 0) return value;
    const int P = psi_prime.num_params;
    const int nsum = mode == 2 ? 7 : 2;
    std::vector<DevBuf<cplx<double>>> outs(nsum);
    std::vector<cplx<double>*> op(nsum);
    for (int i = 0; i < nsum; i++) {
        outs[i].ensure(P);
        op[i] = outs[i].p;
    }
    // 0: <O>, 1: <dev conj(O)>; noise: 2: <|dev|^2 |O|^2>, 3: <dev O>, 4: <dev |O|^2>,
    // 5: <|dev|^2 O>, 6: <|O|^2> (all weighted, deviation terms masked)
    const cplx<double>* coef[7] = {c_w.p, c_d.p, c_d2.p, c_d.p, c_d.p, c_d2.p, c_w.p};
    const int modes[7] = {0, 1, 2, 0, 2, 0, 2};
    if (psi_prime.precision == 0)
        ok_column_sums_t<double>(psi_prime, s, nsum, coef, modes, op.data());
    else
        ok_column_sums_t<float>(psi_prime, s, nsum, coef, modes, op.data());
    std::vector<std::vector<cd>> hs(nsum, std::vector<cd>(P));
    for (int i = 0; i < nsum; i++) {
        outs[i].download(reinterpret_cast<cplx<double>*>(hs[i].data()), P);
        for (cd& x : hs[i]) x /= W;
    }
    const double factor = std::pow(value, nu);
    const double steps = double(ens.num_steps());
    for (int k = 0; k < P; k++) {
        const cd Ok = hs[0][k], dOc = hs[1][k];
        if (grad) grad[k] = (dOc - d * std::conj(Ok)) / factor;
        if (mode == 2 && noise) {
            // KullbackLeibler.cu.template:300-312
            const double d2Ok2 = hs[2][k].real();
            const cd dOk = hs[3][k], dOk2 = hs[4][k], d2Ok = hs[5][k];
            const double Ok2 = hs[6][k].real();
            const double var = d2Ok2 - std::norm(dOc) +
                               2.0 * (dOk * std::conj(d) * std::conj(Ok) + 2.0 * std::conj(dOc) * d * std::conj(Ok) -
                                      d2Ok * std::conj(Ok) - dOk2 * std::conj(d))
                                         .real() +
                               d2 * std::norm(Ok) + std::norm(d) * Ok2 - 4.0 * std::norm(d) * std::norm(Ok);
            noise[k] = std::sqrt(var / steps) / factor;
        }
    }
    return value;
}

double hilbert_space_distance(Psi& psi, Psi& psi_prime, Operator& op, bool unitary, Ensemble& ens, double nu,
                              cd* grad) {
    if (psi.num_sites != psi_prime.num_sites) throw std::invalid_argument("psi and psi_prime differ in num_sites");
    sample(psi, ens);
    Samples& s = ens.samples;
    local_energies(psi, op, s);
    const int64_t n = s.count;
    // log psi' and (for its O_k) the angles of psi' on the same configurations
    psi_prime.sync_device();
    Samples sp;
    prepare_samples(psi_prime, sp, n, s.basis);
    ANN_HIP_CHECK(hipMemcpyAsync(sp.confs.p, s.confs.p, sizeof(uint64_t) * n * s.words, hipMemcpyDeviceToDevice,
                                 S()));
    log_psi_of(psi_prime, sp.confs.p, n, sp, nullptr);
    DevBuf<cplx<double>> c_omega, c_ratio;
    DevBuf<double> sums;
    c_omega.ensure(n);
    c_ratio.ensure(n);
    sums.ensure(4);
    launch_hsd_terms(s.e_loc.p, s.log_psi.p, sp.log_psi.p, s.w.p, n, unitary, c_omega.p, c_ratio.p, sums.p, S());
    double h[4];
    ANN_HIP_CHECK(hipMemcpyAsync(h, sums.p, sizeof(h), hipMemcpyDeviceToHost, S()));
    Runtime::get().sync();
    const cd omega(h[0], h[1]);
    const double u = std::norm(omega), v = h[3] * h[2];
    const double distance = std::sqrt(std::max(1.0 - u / v, 1e-8));
    if (grad) {
        const int P = psi_prime.num_params;
        DevBuf<cplx<double>> a, b;
        a.ensure(P);
        b.ensure(P);
        ok_weighted_sum(psi_prime, sp, c_omega.p, 1, a.p);   // sum w omega conj(O'_k)
        ok_weighted_sum(psi_prime, sp, c_ratio.p, 1, b.p);   // sum w ratio conj(O'_k)
        std::vector<cd> ha(P), hb(P);
        a.download(reinterpret_cast<cplx<double>*>(ha.data()), P);
        b.download(reinterpret_cast<cplx<double>*>(hb.data()), P);
        const double prefactor = std::pow(distance, nu);
        for (int k = 0; k < P; k++) {
            const cd uk = std::conj(omega) * ha[k];
            const cd vk = h[3] * hb[k];
            grad[k] = -(uk * v - u * vk) / (v * v) / prefactor;
        }
    }
    return distance;
}

void tdvp_eval(Tdvp& t, Operator& op, Psi& psi, Ensemble& ens, bool with_S) {
    if (uint32_t(psi.num_params) != t.P) throw std::invalid_argument("TDVP num_params does not match the wave function");
    if (generic_path(psi, ens.basis)) {
        if (psi.precision == 0)
            tdvp_eval_generic<double>(t, op, psi, ens, with_S);
        else
            tdvp_eval_generic<float>(t, op, psi, ens, with_S);
        return;
    }
    require_dense(psi);
    t.generic = false;
    if (psi.precision == 0)
        tdvp_eval_t<double>(t, op, psi, ens, with_S);
    else
        tdvp_eval_t<float>(t, op, psi, ens, with_S);
}

namespace {
template <typename T>
void o_samples_t(Tdvp& t, cd* out) {
    const int64_t n = t.count;
    const int P = int(t.P);
    std::vector<cplx<T>> host;
    if (t.generic) {
        auto& O = materialised<T>(t);
        host.resize(size_t(n) * P);
        O.download(host.data(), host.size());
        for (size_t i = 0; i < host.size(); i++) out[i] = cd(host[i].re, host[i].im);
        return;
    }
    // the structured O (g and the spins) of the evaluation, materialised in slices of samples
    const DenseView<T> net{t.N, t.M, t.Mpad, t.C, t.layout, P, 1, nullptr, nullptr, nullptr,
                           t.layout == LAYOUT_DEEP1 ? t.conn.p : nullptr};
    const cplx<T>* g = Prec<T>::g(t).p;
    const int64_t slice = std::max<int64_t>(1, std::min<int64_t>(n, (int64_t(1) << 27) / std::max(P, 1)));
    DevBuf<cplx<T>> O;
    O.ensure(size_t(slice) * P);
    host.resize(size_t(slice) * P);
    for (int64_t s0 = 0; s0 < n; s0 += slice) {
        const int64_t m = std::min(slice, n - s0);
        launch_materialize_ok<T>(net, t.confs.p + s0 * t.words, t.words, g + s0 * t.Mpad, m, nullptr, O.p, S());
        O.download(host.data(), size_t(m) * P);
        for (size_t i = 0; i < size_t(m) * P; i++) out[size_t(s0) * P + i] = cd(host[i].re, host[i].im);
    }
}
}  // namespace

void tdvp_O_samples(Tdvp& t, cd* out) {
    if (!t.has_samples) throw std::logic_error("TDVP has not been evaluated");
    if (t.precision == 0)
        o_samples_t<double>(t, out);
    else
        o_samples_t<float>(t, out);
}

void tdvp_S_dot(Tdvp& t, const cplx<double>* v, double shift, cplx<double>* y, const double** ypart, int* yparts) {
    if (ypart) *ypart = nullptr;
    if (yparts) *yparts = 0;
    if (!t.has_samples) throw std::logic_error("TDVP.eval must run before S_dot_vector");
    if (t.generic) {
        if (t.precision == 0)
            tdvp_S_dot_generic<double>(t, v, shift, y);
        else
            tdvp_S_dot_generic<float>(t, v, shift, y);
        return;
    }
    if (t.precision == 0)
        tdvp_S_dot_t<double>(t, v, shift, y, ypart, yparts);
    else
        tdvp_S_dot_t<float>(t, v, shift, y, ypart, yparts);
}

const cplx<double>* tdvp_S_dot_dir(Tdvp& t, const CgDir& dir, double shift, cplx<double>* y, const double** ypart,
                                   int* yparts) {
    if (!t.has_samples) throw std::logic_error("TDVP.eval must run before S_dot_vector");
    if (!t.generic && t.precision != 0) {
        // the fp32 structured matvec forms the direction inside its first kernel
        tdvp_S_dot_t<float>(t, dir.p_new, shift, y, ypart, yparts, &dir);
        return dir.p_new;
    }
    cplx<double>* p = const_cast<cplx<double>*>(dir.p_old);
    launch_pcg_dir(p, dir.z, int64_t(t.P), dir.newpart, dir.oldpart, dir.done, S());
    tdvp_S_dot(t, p, shift, y, ypart, yparts);
    return p;
}

int tdvp_solve(Tdvp& t, double shift, int method, double rtol, int max_iter, cd* x_out, int* iters, double* res,
               const cd* x0) {
    const int P = int(t.P);
    if (method == 0) {
        DevBuf<cplx<double>> x;
        x.ensure(P);
        if (!t.has_S) throw std::logic_error("Cholesky solve needs TDVP.eval with the S matrix");
        // the factor and its workspace persist in the TDVP object (4.5 GB at P = 16,704)
        DevBuf<cplx<double>>& A = t.chol;
        A.ensure(size_t(P) * P);
        t.chol_work.ensure(cholesky_work_elems());
        DevBuf<int> info;
        info.ensure(1);
        ANN_HIP_CHECK(hipMemsetAsync(info.p, 0, sizeof(int), S()));
        ANN_HIP_CHECK(hipMemcpyAsync(A.p, t.S.p, sizeof(cplx<double>) * P * P, hipMemcpyDeviceToDevice, S()));
        {
            ScopedTimer timer("cholesky");
            launch_add_diag(A.p, P, shift, S());
            // factor the diagonally scaled matrix (same solution, far smaller condition
            // number when the O_k of different parameters differ in scale)
            DevBuf<double>& dinv = t.chol_dinv;
            dinv.ensure(P);
            launch_equilibrate(A.p, P, dinv.p, S());
            cholesky_lower(A.p, P, info.p, t.chol_work.p, S());
            ANN_HIP_CHECK(hipMemcpyAsync(x.p, t.Fd.p, sizeof(cplx<double>) * P, hipMemcpyDeviceToDevice, S()));
            launch_scale_vec(x.p, P, dinv.p, S());
            cholesky_solve_inplace(A.p, P, x.p, S());
            launch_scale_vec(x.p, P, dinv.p, S());
        }
        int hinfo = 0;
        ANN_HIP_CHECK(hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, S()));
        Runtime::get().sync();
        if (hinfo != 0) return 1;
        x.download(reinterpret_cast<cplx<double>*>(x_out), P);
        // residual || (S + shift) x - F || / || F ||
        t.fetch_host_vectors();
        DevBuf<cplx<double>> y;
        y.ensure(P);
        launch_dense_matvec(t.S.p, x.p, P, shift, y.p, S());
        std::vector<cd> hy(P);
        y.download(reinterpret_cast<cplx<double>*>(hy.data()), P);
        double num = 0, den = 0;
        for (int k = 0; k < P; k++) {
            num += std::norm(hy[k] - t.F[k]);
            den += std::norm(t.F[k]);
        }
        *res = den > 0 ? std::sqrt(num / den) : std::sqrt(num);
        *iters = 1;
        return 0;
    }
    // conjugate gradients with a Jacobi (diagonal) preconditioner.  Vectors, dot
    // products and the CG scalars stay on the device (fp64) as per-block partial
    // sums; the host reads the residual norm every `check` iterations only.
    if (!t.has_samples) throw std::logic_error("CG solve needs TDVP.eval first");
    if (max_iter < 0) throw std::invalid_argument("max_iter must not be negative");
    // the solve's vectors live in the Tdvp (no allocation, hence no device-wide
    // synchronisation, per solve); the host synchronises once, at the end
    DevBuf<cplx<double>>& x = t.cg_x;
    DevBuf<cplx<double>>& r = t.cg_r;
    DevBuf<cplx<double>>& z = t.cg_z;
    DevBuf<cplx<double>>& p = t.cg_p;
    DevBuf<cplx<double>>& p2 = t.cg_p2;
    DevBuf<cplx<double>>& Ap = t.cg_Ap;
    DevBuf<double>& part = t.cg_part;
    DevBuf<double>& apart = t.cg_apart;
    DevBuf<double>& fnorm2 = t.cg_fnorm2;
    x.ensure(P);
    r.ensure(P);
    z.ensure(P);
    p.ensure(P);
    p2.ensure(P);
    Ap.ensure(P);
    t.cg_scratch.ensure(256);
    // partials of (r^H z, -, r^H r, -) per iteration, and of p^H Ap
    const size_t stride = size_t(kCgParts) * 4;
    part.ensure(size_t(max_iter + 1) * stride);
    apart.ensure(size_t(kCgParts) * 2);
    fnorm2.ensure(4);
    launch_dot2(t.Fd.p, t.Fd.p, nullptr, nullptr, P, 1, t.cg_scratch.p, fnorm2.p, S());
    // method 1: Jacobi preconditioner; method 3: K-FAC on the W parameters of a dense
    // net (the biases keep Jacobi), factors from this evaluation's samples
    const bool kfac = method == 3;
    t.kfac_hint = kfac;
    if (kfac) t.kfac_shift = shift;
    // K-FAC with the fp32 S.v of a circulant PsiDeep with C = N keeps the solve's
    // vectors in site-major order (the S.v and K-FAC kernels then read and write whole
    // rows of 64 units); F, x0 and <O> are converted on entry and x on exit
    const bool sm = kfac && t.precision != 0 && !t.generic && t.layout == LAYOUT_DEEP1 && t.conn_shift >= 0 &&
                    t.C == t.N && int64_t(t.N) + t.M + int64_t(t.N) * t.M == P;
    struct OrderGuard {
        Tdvp& t;
        ~OrderGuard() { t.solve_sm = false; }
    } order_guard{t};
    t.solve_sm = sm;
    const cplx<double>* F = t.Fd.p;
    if (sm) {
        t.obar_sm.ensure(P);
        launch_site_major(t.Obar.p, t.N, t.M, t.conn_shift, true, t.obar_sm.p, S());
        launch_site_major(t.Fd.p, t.N, t.M, t.conn_shift, true, z.p, S());
        F = z.p;
    }
    if (x0) {
        // warm start: r = F - (S + shift) x0
        if (sm) {
            Ap.upload(reinterpret_cast<const cplx<double>*>(x0), P);
            launch_site_major(Ap.p, t.N, t.M, t.conn_shift, true, x.p, S());
        } else {
            x.upload(reinterpret_cast<const cplx<double>*>(x0), P);
        }
        tdvp_S_dot(t, x.p, shift, Ap.p);
        launch_cg_dir(Ap.p, F, -1.0, P, S());
        ANN_HIP_CHECK(hipMemcpyAsync(r.p, Ap.p, sizeof(cplx<double>) * P, hipMemcpyDeviceToDevice, S()));
    } else {
        ANN_HIP_CHECK(hipMemsetAsync(x.p, 0, sizeof(cplx<double>) * P, S()));
        ANN_HIP_CHECK(hipMemcpyAsync(r.p, F, sizeof(cplx<double>) * P, hipMemcpyDeviceToDevice, S()));
    }
    const double* diag = t.diag.p != nullptr && (method == 1 || method == 3) ? t.diag.p : nullptr;
    int64_t jac_end = P;
    KfacView kv{};
    if (kfac) {
        if (t.generic) throw std::invalid_argument("the K-FAC preconditioner needs a PsiRBM or one-layer PsiDeep");
        if (t.layout == LAYOUT_DEEP1 && !t.inv_unique)
            throw std::invalid_argument("the K-FAC preconditioner needs distinct connection sites per hidden unit");
        t.kf_idx.ensure(size_t(t.N) * t.M);
        KfacWork kw = kfac_work(t);
        {
            ScopedTimer timer("kfac_setup");
            // The factors of this evaluation's samples: computed and inverted during
            // the evaluation when the previous solve of this Tdvp also used K-FAC (and
            // reused when its shift is asked for again), else here
            if (!(t.kf_inverted && t.kf_inv_shift == shift)) {
                if (!t.kf_ready) {
                    // the transposed spin bits (shared with the fp32 S.v) give the spin covariance
                    const SpinPlan plan = spin_plan(t.N, t.M, t.Mpad, t.count, t.layout == LAYOUT_DEEP1);
                    if (!t.bits_valid) {
                        t.spin_bits.ensure(size_t(plan.bit_blocks) * plan.vpad);
                        launch_spin_bits(t.confs.p, t.words, t.count, t.N, plan.vpad, t.spin_bits.p, S());
                        t.bits_valid = true;
                    }
                    launch_kfac_moments(t.confs.p, t.words, t.spin_bits.p, plan.vpad, t.w_uniform,
                                        t.precision != 0 ? t.g32.p : nullptr, t.precision != 0 ? nullptr : t.g64.p,
                                        t.Mpad, t.w.p, t.count, t.N, t.M, kw, S());
                }
                // the moments are used once: the inversion overwrites D
                t.kf_ready = false;
                launch_kfac_index(t.layout, t.N, t.M, t.Mpad, t.layout == LAYOUT_DEEP1 ? t.inv.p : nullptr,
                                  t.kf_idx.p, S());
                launch_kfac_finish(t.N, t.M, shift, kw, S());
                t.kf_inverted = true;
                t.kf_inv_shift = shift;
            }
        }
        jac_end = t.layout == LAYOUT_DEEP1 ? int64_t(t.N) + t.M : 0;
        kv = {t.N, t.M, t.kf_idx.p, t.order(), t.kf_Ci.p, t.kf_D.p};
    }
    // r is double-buffered under K-FAC (every workgroup of its step reads r at many
    // parameters while their owners write r - alpha Ap)
    DevBuf<cplx<double>>& r2 = t.cg_r2;
    cplx<double>* r_cur = r.p;
    if (kfac) {
        r2.ensure(P);
        KfacStep a{};
        a.r_in = r.p;
        a.z = z.p;
        a.p_init = p.p;
        a.diag = diag;
        a.shift = shift;
        a.jac_end = jac_end;
        a.newpart = part.p;
        launch_kfac_step(kv, a, S());
    } else {
        launch_pcg_init(r.p, z.p, p.p, diag, shift, P, jac_end, part.p, S());
    }
    // Convergence is decided on the device, against (rtol ||F||)^2, at the start and
    // every `check` iterations (a flag that freezes x, r, z, p from then on); the host
    // reads the flag behind an event one check late, so the stream never drains, and
    // the result is the iterate of the first converged check.
    DevBuf<int>& done = t.cg_done;
    done.ensure(2);
    ANN_HIP_CHECK(hipMemsetAsync(done.p, 0, sizeof(int) * 2, S()));
    launch_pcg_check(part.p, fnorm2.p, rtol, 0, done.p, S());
    if (!t.flag_host.p) {
        t.flag_host.ensure(4);
        ANN_HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&t.flag_dev), t.flag_host.p, 0));
    }
    hipEvent_t ev[2] = {t.cg_ev[0].get(), t.cg_ev[1].get()};
    const int check = 4;
    int it = 0, pending = -1;   // slot of the last issued, not yet read flag copy
    bool stop = false;
    // the direction update of iteration k happens in the first kernel of its matvec
    // (p_k = z + beta p_{k-1} into the other of two buffers)
    const cplx<double>* pk = p.p;
    cplx<double>* spare = p2.p;
    while (!stop && it < max_iter) {
        const double* cur = part.p + size_t(it) * stride;
        double* nxt = part.p + size_t(it + 1) * stride;
        const double* ypart = nullptr;
        int yparts = 0;
        if (it == 0) {
            tdvp_S_dot(t, pk, shift, Ap.p, &ypart, &yparts);
        } else {
            CgDir dir;
            dir.z = z.p;
            dir.p_old = pk;
            dir.newpart = cur;
            dir.oldpart = cur - stride;
            dir.parts = kCgParts;
            dir.done = done.p;
            dir.p_new = spare;
            const cplx<double>* pn = tdvp_S_dot_dir(t, dir, shift, Ap.p, &ypart, &yparts);
            if (pn == spare) {
                spare = const_cast<cplx<double>*>(pk);
                pk = pn;
            }
        }
        if (!ypart) {
            launch_pcg_pap(pk, Ap.p, P, apart.p, S());
            ypart = apart.p;
            yparts = kCgParts;
        }
        if (kfac) {
            ScopedTimer timer("kfac_step", (it & 15) == 0);
            KfacStep a{};
            a.x = x.p;
            a.r_in = r_cur;
            a.r_out = r_cur == r.p ? r2.p : r.p;
            a.z = z.p;
            a.p = pk;
            a.Ap = Ap.p;
            a.diag = diag;
            a.shift = shift;
            a.jac_end = jac_end;
            a.rzpart = cur;
            a.apart = ypart;
            a.a_parts = yparts;
            a.newpart = nxt;
            a.done = done.p;
            launch_kfac_step(kv, a, S());
            r_cur = a.r_out;
        } else {
            launch_pcg_step(x.p, r.p, z.p, pk, Ap.p, diag, shift, P, jac_end, cur, ypart, yparts, nxt, done.p, S());
        }
        it++;
        if (it % check == 0 || it == max_iter) {
            const int slot = (it / check) & 1;
            // the check writes the flag straight into the page-locked slot (no copy)
            launch_pcg_check(part.p + size_t(it) * stride, fnorm2.p, rtol, it, done.p, S(), t.flag_dev + 2 * slot);
            ANN_HIP_CHECK(hipEventRecord(ev[slot], S()));
            if (pending >= 0) {
                ANN_HIP_CHECK(hipEventSynchronize(ev[pending]));
                stop = t.flag_host.p[2 * pending] != 0;
            }
            pending = slot;
        }
    }
    // residual, iteration and ||F||^2 of the result, read with x (one synchronisation)
    t.rr_host.ensure(4);
    launch_pcg_final(part.p, it, done.p, fnorm2.p, t.cg_scratch.p, S());
    ANN_HIP_CHECK(hipMemcpyAsync(t.rr_host.p, t.cg_scratch.p, sizeof(double) * 4, hipMemcpyDeviceToHost, S()));
    // x through page-locked memory (one fast copy, then a host copy)
    const cplx<double>* xr = x.p;
    if (sm) {
        launch_site_major(x.p, t.N, t.M, t.conn_shift, false, z.p, S());
        xr = z.p;
    }
    t.x_host.ensure(size_t(P));
    ANN_HIP_CHECK(hipMemcpyAsync(t.x_host.p, xr, sizeof(cplx<double>) * P, hipMemcpyDeviceToHost, S()));
    Runtime::get().sync();
    std::memcpy(x_out, t.x_host.p, sizeof(cplx<double>) * P);
    const double rr = t.rr_host.p[0];
    const double bnorm = std::sqrt(t.rr_host.p[2]);
    const double tol2 = (rtol * bnorm) * (rtol * bnorm);
    const bool converged = t.rr_host.p[3] != 0.0;
    *iters = int(t.rr_host.p[1]);
    *res = bnorm > 0 ? std::sqrt(rr) / bnorm : 0.0;
    return converged || rr <= tol2 ? 0 : 2;
}

// ------------------------------------------------------------ learned proposal

TrainBatch Proposal::batch(int Bt) {
    bsample.ensure(Bt);
    bsites.ensure(size_t(Bt) * bs);
    X.ensure(size_t(Bt) * N);
    Y.ensure(size_t(Bt) * N);
    Mask.ensure(size_t(Bt) * N);
    Hd.ensure(size_t(Bt) * H);
    A2.ensure(size_t(Bt) * N);
    dH.ensure(size_t(Bt) * H);
    loss_part.ensure(Bt);
    gpart.ensure(size_t(kGradSplits) * (size_t(N) * H + std::max(N, H)));
    return {Bt, bsample.p, bsites.p, X.p, Y.p, Mask.p, Hd.p, A2.p, dH.p, loss_part.p, gpart.p};
}

BlockNet Proposal::view() {
    if (!packed) {
        W1q.ensure(size_t(N) * 256);
        b1q.ensure(256);
        W2h.ensure(size_t(N) * 256);
        b2.ensure(N);
        launch_block_net_pack(params.p, N, H, W1q.p, b1q.p, W2h.p, b2.p, S());
        packed = true;
    }
    return {N, H, bs, num_blocks, blocks.p, W1q.p, b1q.p, W2h.p, b2.p, 0};
}

static double batch_loss(Proposal& p, int Bt) {
    std::vector<double> h(Bt);
    p.loss_part.download(h.data(), Bt);
    double s = 0.0;
    for (double x : h) s += x;
    return s;
}

double proposal_train(Proposal& p, const uint64_t* confs, int64_t count, int words, int steps, int batch, double lr) {
    if (count <= 0) throw std::invalid_argument("no configurations to train on");
    if (batch <= 0 || steps < 0) throw std::invalid_argument("batch must be positive");
    if (words != (p.N + 63) / 64) throw std::invalid_argument("configurations do not match the proposal's sites");
    TrainBatch b = p.batch(batch);
    p.grad.ensure(p.num_params());
    for (int it = 0; it < steps; it++) {
        ScopedTimer timer("proposal_train");
        launch_draw_batch(b, count, p.blocks.p, p.num_blocks, p.bs, p.train_step++, p.seed, S());
        launch_build_inputs(b, confs, words, p.N, p.bs, S());
        launch_loss_grad(b, p.params.p, p.N, p.H, double(batch), p.grad.p, S());
        launch_adam(p.params.p, p.m.p, p.v.p, p.grad.p, p.num_params(), float(lr), 0.9f, 0.999f, 1e-8f, ++p.adam_t,
                    S());
    }
    p.packed = false;
    Runtime::get().sync();
    p.last_loss = steps > 0 ? batch_loss(p, batch) : 0.0;
    return p.last_loss;
}

double proposal_loss_grad(Proposal& p, const uint64_t* confs, int64_t count, int words, const int* sites,
                          std::vector<float>& grad) {
    if (count <= 0 || count > (int64_t(1) << 24)) throw std::invalid_argument("1..2^24 rows");
    if (words != (p.N + 63) / 64) throw std::invalid_argument("configurations do not match the proposal's sites");
    for (int64_t r = 0; r < count * p.bs; r++)
        if (sites[r] < 0 || sites[r] >= p.N) throw std::invalid_argument("block site out of range");
    const int Bt = int(count);
    TrainBatch b = p.batch(Bt);
    DevBuf<uint64_t> dconfs;
    dconfs.upload(confs, size_t(count) * words);
    std::vector<int64_t> idx(Bt);
    for (int r = 0; r < Bt; r++) idx[r] = r;
    p.bsample.upload(idx.data(), Bt);
    p.bsites.upload(sites, size_t(Bt) * p.bs);
    p.grad.ensure(p.num_params());
    launch_build_inputs(b, dconfs.p, words, p.N, p.bs, S());
    launch_loss_grad(b, p.params.p, p.N, p.H, double(Bt), p.grad.p, S());
    Runtime::get().sync();
    grad.resize(p.num_params());
    p.grad.download(grad.data(), grad.size());
    return batch_loss(p, Bt);
}

void proposal_adam_step(Proposal& p, const float* grad, double lr) {
    p.grad.upload(grad, p.num_params());
    launch_adam(p.params.p, p.m.p, p.v.p, p.grad.p, p.num_params(), float(lr), 0.9f, 0.999f, 1e-8f, ++p.adam_t, S());
    p.packed = false;
    Runtime::get().sync();
}

}  // namespace ann
