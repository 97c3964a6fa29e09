"""ctypes front-end of the fp64 CPU oracle (oracle/vmc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke() check and by the
CPU-baseline leg of bench.py.  The product (annongpu_amd) never imports it.

The builders take the same constructor arguments as the reference's Python
classes (pyANNonGPU/main.cpp.template:71-204) so a test can build the oracle and
the HIP wave function from one set of numbers.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvmc_oracle.so")

MAX_WORDS = 4
MAX_LAYERS = 4
MAX_CNN_LAYERS = 3
RBM, DEEP, CNN = 0, 1, 2
MOVE_FLIP, MOVE_EXCHANGE = 0, 1



class c_cplx(ctypes.Structure):
    """C99 double complex (passed and returned like a struct of two doubles)."""
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]

P = ctypes.c_void_p


class OrcPsi(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("N", ctypes.c_int32), ("num_params", ctypes.c_int32),
        ("words", ctypes.c_int32), ("log_prefactor", c_cplx),
        ("M", ctypes.c_int32), ("pad0", ctypes.c_int32), ("W", P), ("final_weight", c_cplx),
        ("num_layers", ctypes.c_int32), ("size", ctypes.c_int32 * MAX_LAYERS),
        ("conn", ctypes.c_int32 * MAX_LAYERS), ("pad1", ctypes.c_int32),
        ("connections", P * MAX_LAYERS), ("weights", P * MAX_LAYERS), ("biases", P * MAX_LAYERS),
        ("input_weights", P), ("final_weights", P),
        ("extent", ctypes.c_int32 * 3), ("cnn_layers", ctypes.c_int32),
        ("channels", ctypes.c_int32 * MAX_CNN_LAYERS),
        ("connectivity", (ctypes.c_int32 * 3) * MAX_CNN_LAYERS), ("num_sym", ctypes.c_int32),
        ("symmetry_classes", P), ("params", P), ("final_factor", ctypes.c_double),
    ]


class OrcOperator(ctypes.Structure):
    _fields_ = [("num_strings", ctypes.c_int32), ("words", ctypes.c_int32),
                ("a", P), ("b", P), ("coefficients", P)]


class OrcMCParams(ctypes.Structure):
    _fields_ = [("num_chains", ctypes.c_int32), ("steps_per_chain", ctypes.c_int32),
                ("num_sweeps", ctypes.c_int32), ("num_thermalization_sweeps", ctypes.c_int32),
                ("move", ctypes.c_int32), ("total_z", ctypes.c_int32),
                ("seed", ctypes.c_uint64), ("counter", ctypes.c_uint64),
                ("call_index", ctypes.c_uint32), ("init_random", ctypes.c_int32)]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.orc_log_psi_s.restype = c_cplx
        _lib.orc_local_energy.restype = c_cplx
        _lib.orc_logcosh.restype = c_cplx
        _lib.orc_tanh.restype = c_cplx
        _lib.orc_logcosh.argtypes = [c_cplx, ctypes.c_uint32]
        _lib.orc_tanh.argtypes = [c_cplx, ctypes.c_uint32]
        _lib.orc_local_energy.argtypes = [P, P, P, c_cplx]
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _cplx(z):
    z = complex(z)
    return c_cplx(z.real, z.imag)


def _from_cplx(c):
    return complex(c.re, c.im)


def words_for(n):
    return max(1, (n + 63) // 64)


class OraclePsi:
    """Keeps the numpy arrays alive for as long as the C struct points to them."""

    def __init__(self):
        self.s = OrcPsi()
        self._keep = []

    def keep(self, a):
        self._keep.append(a)
        return _ptr(a)

    @property
    def ref(self):
        return ctypes.byref(self.s)

    @property
    def num_params(self):
        return self.s.num_params

    @property
    def N(self):
        return self.s.N


def rbm(W, final_weight, log_prefactor=0.0):
    """PsiRBM(W, final_weight, log_prefactor) -- include/quantum_state/PsiRBM.hpp:225-243"""
    W = np.ascontiguousarray(W, dtype=np.complex128)
    o = OraclePsi()
    N, M = W.shape
    o.s.kind, o.s.N, o.s.M, o.s.num_params, o.s.words = RBM, N, M, N * M, words_for(N)
    o.s.W = o.keep(W)
    o.s.final_weight = _cplx(final_weight)
    o.s.log_prefactor = _cplx(log_prefactor)
    return o


def deep(num_sites, a, b_list, connections_list, W_list, final_weights, log_prefactor=0.0):
    """PsiDeep(num_sites, a, b, connections, W, final_weights, log_prefactor) --
    include/quantum_state/PsiDeep.hpp:504-578, parameter order source/quantum_state/PsiDeep.cu:142-183"""
    o = OraclePsi()
    a = np.ascontiguousarray(a, dtype=np.complex128)
    N = a.shape[0]
    o.s.kind, o.s.N, o.s.words = DEEP, N, words_for(num_sites)
    o.s.num_layers = len(W_list)
    num_params = N
    for l, (b, c, w) in enumerate(zip(b_list, connections_list, W_list)):
        b = np.ascontiguousarray(b, dtype=np.complex128)
        c = np.ascontiguousarray(c, dtype=np.uint32)
        w = np.ascontiguousarray(w, dtype=np.complex128)
        assert c.shape == w.shape and w.shape[1] == b.shape[0]
        o.s.size[l] = b.shape[0]
        o.s.conn[l] = w.shape[0]
        o.s.biases[l] = o.keep(b)
        o.s.connections[l] = o.keep(c)
        o.s.weights[l] = o.keep(w)
        num_params += b.size + w.size
    o.s.input_weights = o.keep(a)
    o.s.final_weights = o.keep(np.ascontiguousarray(final_weights, dtype=np.complex128))
    o.s.num_params = num_params
    o.s.log_prefactor = _cplx(log_prefactor)
    return o


def cnn(extent, num_channels_list, connectivity_list, symmetry_classes, params, final_factor,
        log_prefactor=0.0):
    """PsiCNN(extent, num_channels_list, connectivity_list, symmetry_classes, params,
    final_factor, log_prefactor) -- include/quantum_state/PsiCNN.hpp:308-340 (dim = 3)"""
    o = OraclePsi()
    extent = list(extent)
    assert len(extent) == 3
    N = int(np.prod(extent))
    params = np.ascontiguousarray(params, dtype=np.complex128)
    sym = np.ascontiguousarray(symmetry_classes, dtype=np.uint32)
    o.s.kind, o.s.N, o.s.words = CNN, N, words_for(N)
    for d in range(3):
        o.s.extent[d] = extent[d]
    o.s.cnn_layers = len(num_channels_list)
    for l, (ch, conn) in enumerate(zip(num_channels_list, connectivity_list)):
        o.s.channels[l] = int(ch)
        for d in range(3):
            o.s.connectivity[l][d] = int(conn[d])
    o.s.num_sym = len(set(sym.tolist()))
    o.s.symmetry_classes = o.keep(sym)
    o.s.params = o.keep(params)
    o.s.num_params = params.size
    o.s.final_factor = float(final_factor)
    o.s.log_prefactor = _cplx(log_prefactor)
    return o


class OracleOperator:
    def __init__(self, a, b, coefficients, words):
        self.a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, words)
        self.b = np.ascontiguousarray(b, dtype=np.uint64).reshape(-1, words)
        self.c = np.ascontiguousarray(coefficients, dtype=np.complex128)
        self.s = OrcOperator(len(self.c), words, _ptr(self.a), _ptr(self.b), _ptr(self.c))

    @property
    def ref(self):
        return ctypes.byref(self.s)


def conf_words(idx_or_bits, words):
    """Integer bit pattern -> little-endian uint64 words."""
    x = int(idx_or_bits)
    return np.array([(x >> (64 * w)) & 0xFFFFFFFFFFFFFFFF for w in range(MAX_WORDS)], dtype=np.uint64)


def log_psi_s(psi, conf):
    conf = np.ascontiguousarray(conf, dtype=np.uint64)
    return _from_cplx(lib().orc_log_psi_s(psi.ref, _ptr(conf)))


def psi_O_k(psi, conf):
    conf = np.ascontiguousarray(conf, dtype=np.uint64)
    out = np.zeros(psi.num_params, dtype=np.complex128)
    lib().orc_psi_O_k(psi.ref, _ptr(conf), _ptr(out))
    return out


def local_energy(op, psi, conf, log_psi):
    conf = np.ascontiguousarray(conf, dtype=np.uint64)
    return _from_cplx(lib().orc_local_energy(op.ref, psi.ref, _ptr(conf), _cplx(log_psi)))


def pauli_apply(a, b, conf, words):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    conf = np.ascontiguousarray(conf, dtype=np.uint64)
    out = np.zeros(MAX_WORDS, dtype=np.uint64)
    f = np.zeros(1, dtype=np.complex128)
    lib().orc_pauli_apply(ctypes.c_int32(words), _ptr(a), _ptr(b), _ptr(conf), _ptr(out), _ptr(f))
    return out, complex(f[0])


def activation(z, layer, derivative=False):
    fn = lib().orc_tanh if derivative else lib().orc_logcosh
    return _from_cplx(fn(_cplx(z), layer))


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_ptr(c), _ptr(k), _ptr(out))
    return out


def exact_log_psi_vector(psi, num_sites):
    out = np.zeros(1 << num_sites, dtype=np.complex128)
    lib().orc_exact_log_psi_vector(psi.ref, ctypes.c_int32(num_sites), _ptr(out))
    return out


def tdvp_accumulate(op, psi, confs, weights, with_S=True, num_threads=0):
    """Returns dict(E, E2, O, F, S, E_loc, O_samples) following
    source/network_functions/TDVP.cu.template:77-126,181-302."""
    confs = np.ascontiguousarray(confs, dtype=np.uint64)
    weights = np.ascontiguousarray(weights, dtype=np.float64)
    ns = weights.shape[0]
    Pn = psi.num_params
    E = np.zeros(1, dtype=np.complex128)
    E2 = np.zeros(1, dtype=np.float64)
    O = np.zeros(Pn, dtype=np.complex128)
    F = np.zeros(Pn, dtype=np.complex128)
    S = np.zeros((Pn, Pn), dtype=np.complex128) if with_S else None
    El = np.zeros(ns, dtype=np.complex128)
    Os = np.zeros((ns, Pn), dtype=np.complex128)
    lib().orc_tdvp_accumulate(op.ref if op is not None else None, psi.ref, ctypes.c_int64(ns),
                              _ptr(confs), _ptr(weights), _ptr(E), _ptr(E2), _ptr(O), _ptr(F),
                              _ptr(S) if with_S else None, _ptr(El), _ptr(Os),
                              ctypes.c_int32(num_threads))
    return dict(E=complex(E[0]), E2=float(E2[0]), O=O, F=F, S=S, E_loc=El, O_samples=Os)


def exact_confs(num_sites, words):
    n = 1 << num_sites
    confs = np.zeros((n, words), dtype=np.uint64)
    confs[:, 0] = np.arange(n, dtype=np.uint64)
    return confs


def mc_run(psi, num_chains, steps_per_chain, num_sweeps, num_thermalization_sweeps, move=MOVE_FLIP,
           total_z=0, seed=0, counter=0, call_index=0, chain_confs=None, num_threads=0):
    """Metropolis chains of include/ensembles/MonteCarlo.hpp:55-177 on the Philox stream."""
    words = psi.s.words
    init_random = chain_confs is None
    chains = np.zeros((num_chains, words), dtype=np.uint64) if init_random else \
        np.ascontiguousarray(chain_confs, dtype=np.uint64).copy()
    p = OrcMCParams(num_chains, steps_per_chain, num_sweeps, num_thermalization_sweeps, move, total_z,
                    seed, counter, call_index, 1 if init_random else 0)
    ns = num_chains * steps_per_chain
    samples = np.zeros((ns, words), dtype=np.uint64)
    lp = np.zeros(ns, dtype=np.complex128)
    acc = ctypes.c_uint64(0)
    rej = ctypes.c_uint64(0)
    lib().orc_mc_run(psi.ref, ctypes.byref(p), _ptr(chains), _ptr(samples), _ptr(lp),
                     ctypes.byref(acc), ctypes.byref(rej), ctypes.c_int32(num_threads))
    return dict(samples=samples, log_psi=lp, chains=chains, acceptances=acc.value,
                rejections=rej.value)


def cholesky_solve(S, shift, F):
    S = np.ascontiguousarray(S, dtype=np.complex128)
    F = np.ascontiguousarray(F, dtype=np.complex128)
    x = np.zeros_like(F)
    rc = lib().orc_cholesky_solve(ctypes.c_int32(F.size), _ptr(S), ctypes.c_double(shift), _ptr(F), _ptr(x))
    if rc != 0:
        raise np.linalg.LinAlgError("matrix not positive definite")
    return x


def cg_solve(S, shift, F, rtol=1e-10, max_iter=1000):
    S = np.ascontiguousarray(S, dtype=np.complex128)
    F = np.ascontiguousarray(F, dtype=np.complex128)
    x = np.zeros_like(F)
    it = ctypes.c_int32(0)
    rc = lib().orc_cg_solve(ctypes.c_int32(F.size), _ptr(S), ctypes.c_double(shift), _ptr(F), _ptr(x),
                            ctypes.c_double(rtol), ctypes.c_int32(max_iter), ctypes.byref(it))
    return x, it.value, rc


def sr_cg_matfree(O_samples, weights, O_mean, F, shift, rtol=1e-6, max_iter=200, num_threads=0):
    """(S + shift) x = F by Jacobi CG on the matrix-free S = sum_s w_s conj(Oc_s) Oc_s^T"""
    O_samples = np.ascontiguousarray(O_samples, dtype=np.complex128)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    O_mean = np.ascontiguousarray(O_mean, dtype=np.complex128)
    F = np.ascontiguousarray(F, dtype=np.complex128)
    ns, Pn = O_samples.shape
    x = np.zeros(Pn, dtype=np.complex128)
    res = ctypes.c_double(0.0)
    f = lib().orc_sr_cg_matfree
    f.restype = ctypes.c_int32
    it = f(ctypes.c_int64(ns), ctypes.c_int32(Pn), _ptr(O_samples), _ptr(w), _ptr(O_mean), _ptr(F),
           ctypes.c_double(shift), ctypes.c_double(rtol), ctypes.c_int32(max_iter), _ptr(x), ctypes.byref(res),
           ctypes.c_int32(num_threads))
    return x, int(it), res.value


# ------------------------------------------------------------ learned proposal
# (oracle/proposal_oracle.c: no counterpart in the reference)

def mlp_num_params(N, H):
    return H * N + H + N * H + N


def mlp_logits(N, H, params, conf, block):
    """logits of the proposal network for conf with the block's sites masked"""
    params = np.ascontiguousarray(params, dtype=np.float64)
    conf = np.ascontiguousarray(conf, dtype=np.uint64)
    block = np.ascontiguousarray(block, dtype=np.int32)
    out = np.zeros(N, dtype=np.float64)
    lib().orc_mlp_logits(ctypes.c_int32(N), ctypes.c_int32(H), _ptr(params), _ptr(conf), _ptr(block),
                         ctypes.c_int32(block.size), _ptr(out))
    return out


def mlp_loss_grad(N, H, params, confs, blocks):
    """(loss, gradient) of the block pseudo-likelihood loss, mean over samples"""
    params = np.ascontiguousarray(params, dtype=np.float64)
    confs = np.ascontiguousarray(confs, dtype=np.uint64)
    if confs.ndim == 1:
        confs = confs[:, None]
    blocks = np.ascontiguousarray(blocks, dtype=np.int32)
    grad = np.zeros(params.size, dtype=np.float64)
    f = lib().orc_mlp_loss_grad
    f.restype = ctypes.c_double
    loss = f(ctypes.c_int32(N), ctypes.c_int32(H), _ptr(params), ctypes.c_int64(confs.shape[0]), _ptr(confs),
             ctypes.c_int32(confs.shape[1]), _ptr(blocks), ctypes.c_int32(blocks.shape[1]), _ptr(grad))
    return loss, grad


def adam_step(params, m, v, grad, lr, t, beta1=0.9, beta2=0.999, eps=1e-8):
    """one Adam step in place on float64 arrays"""
    for a in (params, m, v):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    grad = np.ascontiguousarray(grad, dtype=np.float64)
    lib().orc_adam_step(ctypes.c_int64(params.size), _ptr(params), _ptr(m), _ptr(v), _ptr(grad),
                        ctypes.c_double(lr), ctypes.c_double(beta1), ctypes.c_double(beta2),
                        ctypes.c_double(eps), ctypes.c_int64(t))


def block_log_q(block_logits, p_new, k):
    """log probability of block pattern p_new among the patterns with k up spins"""
    bl = np.ascontiguousarray(block_logits, dtype=np.float64)
    f = lib().orc_block_log_q
    f.restype = ctypes.c_double
    return f(_ptr(bl), ctypes.c_int32(bl.size), ctypes.c_uint32(p_new), ctypes.c_int32(k))
