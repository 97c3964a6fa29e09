"""Python classes and functions with the names and argument order of the
reference's pyANNonGPU bindings (pyANNonGPU/main.cpp.template:65-541) and
wrappers (pyANNonGPU/PsiRBM.py, PsiDeep.py, PsiCNN.py), implemented on the
C ABI of libannongpu.so.

The reference's ``gpu`` flags are accepted and ignored: every computation runs
on the MI355X; there is no CPU execution path in this package.
"""
import ctypes
import json

import numpy as np

from . import _lib
from ._lib import lib, check, ptr, ref, as_c128, c2, handle
from .json_numpy import NumpyEncoder, NumpyDecoder
from .pauli import PauliExpression

MAX_WORDS = 4


def _words_for(n):
    return max(1, (int(n) + 63) // 64)


# ----------------------------------------------------------------- basis

class Spins:
    """Spin configuration (include/basis/Spins.h:20-414): bit i set <=> s_i = +1.

    ``Spins(configuration, num_spins)`` masks to num_spins bits, like the
    reference's Spins_t<1>(configuration, num_spins) (Spins.h:278-285); any
    number of sites up to 256 is accepted here.
    """

    __slots__ = ("configuration",)

    def __init__(self, configuration=0, num_spins=256):
        c = int(configuration)
        if num_spins < 256:
            c &= (1 << int(num_spins)) - 1
        self.configuration = c

    @staticmethod
    def enumerate(index):
        return Spins(int(index))

    @staticmethod
    def from_array(arr):
        c = 0
        for i, s in enumerate(arr):
            if s > 0:
                c |= 1 << i
        return Spins(c)

    def roll(self, shift, N):
        shift %= N
        mask = (1 << N) - 1
        c = self.configuration & mask
        return Spins(((c << shift) | (c >> (N - shift))) & mask, N)

    def array(self, num_spins):
        return np.array([1.0 if (self.configuration >> i) & 1 else -1.0 for i in range(num_spins)])

    def flip(self, position):
        return Spins(self.configuration ^ (1 << position))

    def words(self, n=MAX_WORDS):
        return np.array([(self.configuration >> (64 * w)) & 0xFFFFFFFFFFFFFFFF for w in range(n)],
                        dtype=np.uint64)

    def __eq__(self, other):
        return isinstance(other, Spins) and other.configuration == self.configuration

    def __hash__(self):
        return hash(self.configuration)

    def __repr__(self):
        return f"Spins({self.configuration:#x})"


def _conf_words(spins):
    if isinstance(spins, Spins):
        return spins.words()
    if isinstance(spins, (int, np.integer)):
        return Spins(int(spins)).words()
    a = np.ascontiguousarray(spins, dtype=np.uint64).ravel()
    out = np.zeros(MAX_WORDS, dtype=np.uint64)
    out[:a.size] = a
    return out


class PauliString:
    """Pauli-string configuration (include/basis/PauliString.hpp:18-278): bit i of ``a``
    and ``b`` give site i's type a_i + 2 b_i (0 = identity, 1 = X, 2 = Y, 3 = Z); at most
    64 sites."""

    __slots__ = ("a", "b")

    def __init__(self, a=0, b=0):
        self.a, self.b = int(a), int(b)

    @staticmethod
    def enumerate(index):
        """PauliString.hpp:30-35: a from the even bits of the index, b from the odd bits"""
        index = int(index)
        a = b = 0
        for k in range(32):
            a |= ((index >> (2 * k)) & 1) << k
            b |= ((index >> (2 * k + 1)) & 1) << k
        return PauliString(a, b)

    def __getitem__(self, i):
        return ((self.a >> i) & 1) | (((self.b >> i) & 1) << 1)

    def set_at(self, i, kind):
        """PauliString.hpp:37-56"""
        bit = 1 << i
        self.a = (self.a & ~bit) | (bit if kind & 1 else 0)
        self.b = (self.b & ~bit) | (bit if kind & 2 else 0)

    def roll(self, shift, N):
        shift %= N
        mask = (1 << N) - 1

        def rot(x):
            x &= mask
            return ((x << shift) | (x >> (N - shift))) & mask

        return PauliString(rot(self.a), rot(self.b))

    def array(self, num_sites):
        return np.array([self[i] for i in range(num_sites)], dtype=np.uint32)

    def words(self):
        return np.array([self.a, self.b], dtype=np.uint64)

    def __eq__(self, other):
        return isinstance(other, PauliString) and (self.a, self.b) == (other.a, other.b)

    def __hash__(self):
        return hash((self.a, self.b))

    def __repr__(self):
        return f"PauliString({self.a:#x}, {self.b:#x})"


# ---------------------------------------------------------- wave functions

class _Psi:
    _h = None

    def __del__(self):
        if self._h is not None and self._h.value:
            try:
                lib.ann_psi_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def _info(self):
        p, n, s = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(lib.ann_psi_info(self._h, ref(p), ref(n), ref(s)))
        return p.value, n.value, s.value

    @property
    def num_params(self):
        return self._info()[0]

    @property
    def params(self):
        out = np.zeros(self.num_params, dtype=np.complex128)
        check(lib.ann_psi_get_params(self._h, ptr(out)))
        return out

    @params.setter
    def params(self, value):
        v = as_c128(value).ravel()
        if v.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters, got {v.size}")
        check(lib.ann_psi_set_params(self._h, ptr(v)))
        self._on_params(v)

    def _on_params(self, v):
        pass

    @property
    def log_prefactor(self):
        out = np.zeros(2)
        check(lib.ann_psi_get_log_prefactor(self._h, ptr(out)))
        return complex(out[0], out[1])

    @log_prefactor.setter
    def log_prefactor(self, value):
        check(lib.ann_psi_set_log_prefactor(self._h, ptr(c2(value))))

    @property
    def precision(self):
        return self._precision

    @precision.setter
    def precision(self, value):
        code = {"fp64": _lib.FP64, "fp32": _lib.FP32}[value]
        check(lib.ann_psi_set_precision(self._h, code))
        self._precision = value

    @property
    def symmetric(self):
        return False

    @property
    def gpu(self):
        return True

    def copy(self):
        h = handle()
        check(lib.ann_psi_copy(self._h, ref(h)))
        c = object.__new__(type(self))
        c.__dict__.update(self.__dict__)
        c._h = h
        return c

    def __pos__(self):
        return self.copy()

    # reference: psi._vector(ensemble) / psi.vector property -> bound method
    def _vector(self, ensemble):
        return psi_vector(self, ensemble)

    @property
    def vector(self):
        return self._vector

    def norm(self, exact_summation):
        out = ctypes.c_double()
        check(lib.ann_psi_norm(self._h, exact_summation._h, ref(out)))
        return out.value

    def normalize(self, exact_summation):
        """PsiRBM.py:39-40: log_prefactor -= log(norm)"""
        self.log_prefactor -= np.log(self.norm(exact_summation))

    def calibrate(self, ensemble):
        """PsiRBM.py:43-49"""
        if type(ensemble).__name__.startswith("ExactSummation"):
            self.normalize(ensemble)
            self.log_prefactor -= 1j * log_psi(self, ensemble).imag
        else:
            self.log_prefactor = 0
            self.log_prefactor -= log_psi(self, ensemble)


class PsiRBM(_Psi):
    """PsiRBM(W, final_weight, log_prefactor, gpu) -- include/quantum_state/PsiRBM.hpp:225-243.

    log psi(s) = log_prefactor + final_weight * sum_j lc(theta_j), theta_j = sum_i W_ij s_i.
    """

    def __init__(self, W, final_weight=1.0, log_prefactor=0.0, gpu=True):
        W = as_c128(W)
        if W.ndim != 2:
            raise ValueError("W must be a 2d array [N, M]")
        self._N, self._M = W.shape
        self._final_weight = complex(final_weight)
        self._precision = "fp64"
        self._h = handle()
        check(lib.ann_psi_create_rbm(ref(self._h), self._N, self._M, ptr(W), ptr(c2(final_weight)),
                                     ptr(c2(log_prefactor))))

    @property
    def N(self):
        return self._N

    @property
    def M(self):
        return self._M

    @property
    def num_sites(self):
        return self._N

    @property
    def W(self):
        return self.params.reshape(self._N, self._M)

    @property
    def final_weight(self):
        return self._final_weight

    def to_json(self):
        obj = dict(type="PsiRBM", W=self.W, final_weight=self.final_weight,
                   log_prefactor_re=self.log_prefactor.real, log_prefactor_im=self.log_prefactor.imag)
        return json.loads(json.dumps(obj, cls=NumpyEncoder))

    @staticmethod
    def from_json(json_obj, gpu=True):
        obj = json.loads(json.dumps(json_obj, cls=NumpyEncoder), cls=NumpyDecoder)
        return PsiRBM(obj["W"], obj["final_weight"], obj["log_prefactor_re"] + 1j * obj["log_prefactor_im"], gpu)


class PsiDeep(_Psi):
    """PsiDeep(num_sites, a, b_list, connections_list, W_list, final_weights, log_prefactor, gpu)
    -- include/quantum_state/PsiDeep.hpp:504-578.

    Parameter order [a | b_1 | W_1 | b_2 | W_2 ...] (source/quantum_state/PsiDeep.cu:245-270).
    As in the reference, log psi does not contain the a.s term (PsiDeep.hpp:260-262)
    although O_k reports s_i for the a parameters.
    """

    def __init__(self, num_sites, a, b, connections, W, final_weights, log_prefactor=0.0, gpu=True):
        self._num_sites = int(num_sites)
        a = as_c128(a).ravel()
        self._N = a.size
        self._layers = []
        sizes, conns, c_ptrs, w_ptrs, b_ptrs = [], [], [], [], []
        keep = []
        for bl, cl, wl in zip(b, connections, W):
            bl = as_c128(bl).ravel()
            cl = np.ascontiguousarray(cl, dtype=np.uint32)
            wl = as_c128(wl)
            if cl.shape != wl.shape or wl.shape[1] != bl.size:
                raise ValueError("connections / W / b shapes do not match")
            sizes.append(bl.size)
            conns.append(wl.shape[0])
            keep += [bl, cl, wl]
            c_ptrs.append(cl.ctypes.data)
            w_ptrs.append(wl.ctypes.data)
            b_ptrs.append(bl.ctypes.data)
            self._layers.append((bl.size, wl.shape[0], cl.copy()))
        L = len(sizes)
        fw = as_c128(final_weights).ravel()
        sizes_a = np.array(sizes, dtype=np.int32)
        conns_a = np.array(conns, dtype=np.int32)
        cp = (ctypes.c_void_p * L)(*c_ptrs)
        wp = (ctypes.c_void_p * L)(*w_ptrs)
        bp = (ctypes.c_void_p * L)(*b_ptrs)
        self._final_weights = fw.copy()
        self._precision = "fp64"
        self._h = handle()
        check(lib.ann_psi_create_deep(ref(self._h), self._num_sites, self._N, ptr(a), L, ptr(sizes_a), ptr(conns_a),
                                      ctypes.cast(cp, ctypes.c_void_p), ctypes.cast(wp, ctypes.c_void_p),
                                      ctypes.cast(bp, ctypes.c_void_p), ptr(fw), ptr(c2(log_prefactor))))
        del keep

    @property
    def N(self):
        return self._N

    @property
    def num_sites(self):
        return self._num_sites

    def _split(self):
        p = self.params
        a = p[:self._N]
        o = self._N
        bs, Ws = [], []
        for size, conn, _ in self._layers:
            bs.append(p[o:o + size])
            o += size
            Ws.append(p[o:o + conn * size].reshape(conn, size))
            o += conn * size
        return a, bs, Ws

    @property
    def a(self):
        return self._split()[0]

    input_weights = a

    @property
    def b(self):
        return self._split()[1]

    @property
    def W(self):
        return self._split()[2]

    @property
    def connections(self):
        return [c.copy() for _, _, c in self._layers]

    @property
    def final_weights(self):
        return self._final_weights.copy()

    @property
    def first_layer_params_slice(self):
        return slice(self._N, self._N + self.W[0].size + self.b[0].size)

    def to_json(self):
        obj = dict(type="PsiDeep", num_sites=self.num_sites, a=self.a, b=self.b, connections=self.connections,
                   W=self.W, final_weights=self.final_weights, log_prefactor_re=self.log_prefactor.real,
                   log_prefactor_im=self.log_prefactor.imag)
        return json.loads(json.dumps(obj, cls=NumpyEncoder))

    @staticmethod
    def from_json(json_obj, gpu=True):
        obj = json.loads(json.dumps(json_obj, cls=NumpyEncoder), cls=NumpyDecoder)
        return PsiDeep(obj["num_sites"], obj["a"], obj["b"], obj["connections"], obj["W"], obj["final_weights"],
                       obj["log_prefactor_re"] + 1j * obj["log_prefactor_im"], gpu)


class PsiCNN(_Psi):
    """PsiCNN(extent, num_channels_list, connectivity_list, symmetry_classes, params,
    final_factor, log_prefactor, gpu) -- include/quantum_state/PsiCNN.hpp:308-340 (dim = 3).

    Extents of fewer than three dimensions are padded with leading 1s.
    """

    def __init__(self, extent, num_channels_list, connectivity_list, symmetry_classes, params, final_factor,
                 log_prefactor=0.0, gpu=True):
        extent = [int(e) for e in extent]
        pad = 3 - len(extent)
        if pad < 0:
            raise ValueError("at most 3 dimensions")
        self._extent = [1] * pad + extent
        self._channels = np.array(num_channels_list, dtype=np.int32).ravel()
        conn = np.array(connectivity_list, dtype=np.int32).reshape(len(self._channels), -1)
        conn = np.concatenate([np.ones((conn.shape[0], pad), dtype=np.int32), conn], axis=1)
        self._connectivity = np.ascontiguousarray(conn)
        N = int(np.prod(self._extent))
        if symmetry_classes is None:
            symmetry_classes = np.zeros(N)
        self._sym = np.ascontiguousarray(symmetry_classes, dtype=np.uint32).ravel()
        self._final_factor = float(final_factor)
        params = as_c128(params).ravel()
        self._precision = "fp64"
        ext = np.array(self._extent, dtype=np.int32)
        self._h = handle()
        check(lib.ann_psi_create_cnn(ref(self._h), ptr(ext), len(self._channels), ptr(self._channels),
                                     ptr(self._connectivity), ptr(self._sym), params.size, ptr(params),
                                     self._final_factor, ptr(c2(log_prefactor))))

    @property
    def N(self):
        return int(np.prod(self._extent))

    num_sites = N

    @property
    def extent(self):
        return list(self._extent)

    @property
    def dim(self):
        return 3

    @property
    def num_channels_list(self):
        # unsigned, as the reference's property (MAIN:180)
        return self._channels.astype(np.uint32)

    @property
    def connectivity_list(self):
        return self._connectivity.astype(np.uint32)

    @property
    def symmetry_classes(self):
        return self._sym.copy()

    @property
    def num_symmetry_classes(self):
        return len(set(self._sym.tolist()))

    @property
    def final_factor(self):
        return self._final_factor

    def init_gradient(self, num_steps):
        """Kept for interface compatibility (PsiCNN.cpp:250-254); storage is managed internally."""

    def channel_link(self, layer, i, j):
        """PsiCNN.py:67-83"""
        ns = self.num_symmetry_classes
        offset = 0
        for l in range(layer):
            prev = self._channels[l - 1] if l > 0 else 1
            offset += self._channels[l] * prev * ns * int(np.prod(self._connectivity[l]))
        c = int(np.prod(self._connectivity[layer]))
        offset += (i * self._channels[layer] + j) * ns * c
        return self.params[offset:offset + ns * c]

    def to_json(self):
        obj = dict(type="PsiCNN", extent=self.extent, num_channels_list=self.num_channels_list,
                   connectivity_list=self.connectivity_list, symmetry_classes=self.symmetry_classes,
                   params=self.params, final_factor=self.final_factor, log_prefactor_re=self.log_prefactor.real,
                   log_prefactor_im=self.log_prefactor.imag)
        return json.loads(json.dumps(obj, cls=NumpyEncoder))

    @staticmethod
    def from_json(json_obj, gpu=True):
        obj = json.loads(json.dumps(json_obj, cls=NumpyEncoder), cls=NumpyDecoder)
        return PsiCNN(obj["extent"], obj["num_channels_list"], obj["connectivity_list"], obj["symmetry_classes"],
                      obj["params"], obj["final_factor"], obj["log_prefactor_re"] + 1j * obj["log_prefactor_im"],
                      gpu)


class PsiFullyPolarized:
    """PsiFullyPolarized(num_sites, log_prefactor) -- include/quantum_state/PsiFullyPolarized.hpp:
    the reference state of the classical networks: log psi = 0 on every configuration (its
    log_prefactor is stored but, as in the reference, does not enter log psi), no parameters."""

    def __init__(self, num_sites, log_prefactor=0.0):
        self.num_sites = int(num_sites)
        self.log_prefactor = complex(log_prefactor)
        self.num_params = 0
        self._h = handle()
        check(lib.ann_psi_create_classical(ref(self._h), self.num_sites, 0, None, None, None, 1, None, None))
        self._precision = "fp64"

    def __del__(self):
        if self._h is not None and self._h.value:
            try:
                lib.ann_psi_destroy(self._h)
            except Exception:
                pass
            self._h = None

    @property
    def params(self):
        return np.zeros(0, dtype=np.complex128)

    @property
    def gpu(self):
        return True

    def copy(self):
        return PsiFullyPolarized(self.num_sites, self.log_prefactor)

    def __pos__(self):
        return self.copy()

    def vector(self, ensemble):
        return psi_vector(self, ensemble)

    def to_json(self):
        """pyANNonGPU/PsiFullyPolarized.py:4-10"""
        return dict(type="PsiFullyPolarized", num_sites=self.num_sites, log_prefactor_re=self.log_prefactor.real,
                    log_prefactor_im=self.log_prefactor.imag)

    @staticmethod
    def from_json(obj):
        return PsiFullyPolarized(obj["num_sites"], obj["log_prefactor_re"] + 1j * obj["log_prefactor_im"])


class PsiClassical(_Psi):
    """PsiClassical (include/quantum_state/PsiClassical.hpp:48-160, source/quantum_state/
    PsiClassical.cu:9-61), the reference's PsiClassicalFP_1/_2 and PsiClassicalANN_1/_2:

        log psi(s) = log_prefactor + sum_n params[n] L_n(s)  [+ log psi_ref(s), order 2],

    L_n(s) = sum over the strings of H_local[n] of c_m <s|P_m|s'> (Operator::fast_local_energy,
    operator/Operator.hpp:123-135), on spin configurations or Pauli strings.  psi_ref is a
    PsiFullyPolarized (log psi_ref = 0, no parameters), a PsiCNN (the reference's
    PsiClassicalANN, PsiClassical.hpp:245-256; spin configurations) or a PsiDeep (also on
    Pauli strings); with order 2 a network reference enters log psi and its parameters follow
    the classical ones.
    """

    def __init__(self, num_sites, H_local, params, psi_ref, log_prefactor=0.0, gpu=True, order=1):
        self._num_sites = int(num_sites)
        ops = [h if isinstance(h, Operator) else Operator(h) for h in H_local]
        self._H_local = ops
        p = as_c128(params).ravel()
        if p.size < len(ops):
            raise ValueError(f"expected {len(ops)} classical parameters, got {p.size}")
        p = np.ascontiguousarray(p[:len(ops)])
        hs = (ctypes.c_void_p * max(len(ops), 1))(*[o._h.value for o in ops])
        if psi_ref is None:
            psi_ref = PsiFullyPolarized(num_sites, 0.0)
        self._ref_template = psi_ref
        net = None if isinstance(psi_ref, PsiFullyPolarized) else psi_ref
        if net is not None and not isinstance(net, (PsiDeep, PsiCNN)):
            raise TypeError("psi_ref must be a PsiFullyPolarized, a PsiCNN or a PsiDeep")
        self._precision = "fp64"
        self._h = handle()
        check(lib.ann_psi_create_classical(ref(self._h), self._num_sites, len(ops), ctypes.cast(hs, ctypes.c_void_p),
                                           ptr(p), net._h if net is not None else None, int(order),
                                           ptr(c2(log_prefactor)), ptr(c2(psi_ref.log_prefactor))))

    @property
    def num_sites(self):
        return self._num_sites

    @property
    def N(self):
        return self._num_sites

    @property
    def H_local(self):
        return list(self._H_local)

    def _cinfo(self):
        o, n, r = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(lib.ann_psi_classical_info(self._h, ref(o), ref(n), ref(r)))
        return o.value, n.value, r.value

    @property
    def order(self):
        return self._cinfo()[0]

    @property
    def ref_log_prefactor(self):
        out = np.zeros(2)
        check(lib.ann_psi_get_ref_log_prefactor(self._h, ptr(out)))
        return complex(out[0], out[1])

    @ref_log_prefactor.setter
    def ref_log_prefactor(self, value):
        check(lib.ann_psi_set_ref_log_prefactor(self._h, ptr(c2(value))))

    @property
    def psi_ref(self):
        """a copy of the reference (writing to it has no effect, MAIN:236-239)"""
        t = self._ref_template
        if isinstance(t, PsiFullyPolarized):
            return PsiFullyPolarized(self._num_sites, self.ref_log_prefactor)
        h = handle()
        check(lib.ann_psi_classical_ref(self._h, ref(h)))
        c = object.__new__(type(t))
        c.__dict__.update(t.__dict__)
        c._h = h
        c._precision = "fp64"
        return c

    def update_psi_ref_kernel(self):
        """MAIN:255; the reference is held by the library and needs no update"""

    def calibrate(self, ensemble):
        """MAIN:274-281 (MonteCarloSpins): psi_ref.log_prefactor -= log_psi(psi_ref),
        log_prefactor -= log_psi(psi).  On an exact summation both are normalised instead
        (and the mean phase of log psi removed), as the other wave functions do."""
        exact = type(ensemble).__name__.startswith("ExactSummation")
        if not isinstance(self._ref_template, PsiFullyPolarized):
            r = self.psi_ref
            if exact:
                self.ref_log_prefactor = self.ref_log_prefactor - np.log(r.norm(ensemble))
            else:
                self.ref_log_prefactor = self.ref_log_prefactor - log_psi(r, ensemble)
        if exact:
            _Psi.calibrate(self, ensemble)
        else:
            self.log_prefactor -= log_psi(self, ensemble)

    def to_json(self, ansatz=None):
        """vCN_to_json (pyANNonGPU/PsiClassical.py:7-25); ansatz defaults to the H_local expressions"""
        ansatz = [o.expr for o in self._H_local] if ansatz is None else list(ansatz)
        pref = self.psi_ref
        obj = dict(type="PsiClassical", num_sites=self.num_sites, order=self.order,
                   ansatz=[a.to_json() for a in ansatz], params=self.params[:len(self._H_local)],
                   psi_ref=pref.to_json(), log_prefactor_re=self.log_prefactor.real,
                   log_prefactor_im=self.log_prefactor.imag)
        return json.loads(json.dumps(obj, cls=NumpyEncoder))

    @staticmethod
    def from_json(json_obj, gpu=True):
        """vCN_from_json (pyANNonGPU/PsiClassical.py:28-67)"""
        obj = json.loads(json.dumps(json_obj, cls=NumpyEncoder), cls=NumpyDecoder)
        r = obj["psi_ref"]
        readers = {"PsiFullyPolarized": lambda o: PsiFullyPolarized.from_json(o),
                   "PsiDeep": lambda o: PsiDeep.from_json(o, gpu), "PsiCNN": lambda o: PsiCNN.from_json(o, gpu)}
        if r["type"] not in readers:
            raise ValueError(f"unknown reference network type {r['type']!r}")
        psi_ref = readers[r["type"]](r)
        H_local = [Operator(PauliExpression.from_json(h)) for h in obj["ansatz"]]
        return PsiClassical(obj["num_sites"], H_local, obj["params"], psi_ref,
                            obj["log_prefactor_re"] + 1j * obj["log_prefactor_im"], gpu, order=obj["order"])


def PsiClassicalFP_1(num_sites, H_local, params, psi_ref, log_prefactor=0.0, gpu=True):
    return PsiClassical(num_sites, H_local, params, psi_ref, log_prefactor, gpu, order=1)


def PsiClassicalFP_2(num_sites, H_local, params, psi_ref, log_prefactor=0.0, gpu=True):
    return PsiClassical(num_sites, H_local, params, psi_ref, log_prefactor, gpu, order=2)


def PsiClassicalANN_1(num_sites, H_local, params, psi_ref, log_prefactor=0.0, gpu=True):
    return PsiClassical(num_sites, H_local, params, psi_ref, log_prefactor, gpu, order=1)


def PsiClassicalANN_2(num_sites, H_local, params, psi_ref, log_prefactor=0.0, gpu=True):
    return PsiClassical(num_sites, H_local, params, psi_ref, log_prefactor, gpu, order=2)


def new_classical_network(num_sites, order, H_local, symmetric=False, distance="max", params=0,
                          psi_ref="fully polarized", use_super_operator=False, gpu=False):
    """pyANNonGPU/new_classical_network.py:6-70.  H_local: a PauliExpression (one local
    operator per term, as iterating the reference's PauliExpression yields) or a list of
    expressions.  The symmetric (PSI_CLASSICAL_SYMMETRIC) and SuperOperator variants are
    not built."""
    assert order in (1, 2)
    if symmetric or use_super_operator:
        raise NotImplementedError("the symmetric and SuperOperator classical networks are not built")
    if isinstance(H_local, PauliExpression):
        H_local = H_local.term_list()
    H_local = list(H_local)
    log_prefactor = np.log(1 / 2 ** (num_sites / 2))
    if isinstance(params, (int, float)) and params == 0:
        params = np.zeros(len(H_local), dtype=complex)
    if isinstance(psi_ref, str):
        return PsiClassical(num_sites, H_local, params, PsiFullyPolarized(num_sites, log_prefactor), log_prefactor,
                            gpu, order=order)
    return PsiClassical(num_sites, H_local, params, psi_ref, 0.0, gpu, order=order)


# ------------------------------------------------------------------ operator

class Operator:
    """Operator(PauliExpression, gpu) -- source/operator/Operator.cpp:18-39."""

    def __init__(self, expr, gpu=True):
        # a PauliExpression, or any iterable of ((a, b), coefficient) in the
        # reference's (a, b) string encoding (include/basis/PauliString.hpp:37-56)
        try:
            terms = [((int(ab[0]), int(ab[1])), complex(c)) for ab, c in expr]
        except (TypeError, ValueError) as e:
            raise TypeError("Operator expects a PauliExpression or ((a, b), coefficient) pairs") from e
        n = len(terms)
        words = MAX_WORDS
        a = np.zeros((max(n, 1), words), dtype=np.uint64)
        b = np.zeros((max(n, 1), words), dtype=np.uint64)
        c = np.zeros(max(n, 1), dtype=np.complex128)
        for k, ((ma, mb), coef) in enumerate(terms):
            if max(ma.bit_length(), mb.bit_length()) > 64 * words:
                raise ValueError("Pauli string beyond 256 sites")
            for w in range(words):
                a[k, w] = (ma >> (64 * w)) & 0xFFFFFFFFFFFFFFFF
                b[k, w] = (mb >> (64 * w)) & 0xFFFFFFFFFFFFFFFF
            c[k] = coef
        self.expr = expr
        self._h = handle()
        check(lib.ann_operator_create(ref(self._h), n, words, ptr(a), ptr(b), ptr(c)))
        self.num_strings = n

    def __del__(self):
        if self._h is not None and self._h.value:
            try:
                lib.ann_operator_destroy(self._h)
            except Exception:
                pass
            self._h = None


# ----------------------------------------------------------------- ensembles

class _Ensemble:
    _h = None

    def __del__(self):
        if self._h is not None and self._h.value:
            try:
                lib.ann_ensemble_destroy(self._h)
            except Exception:
                pass
            self._h = None

    @property
    def gpu(self):
        return True

    @property
    def num_steps(self):
        n = ctypes.c_uint64()
        check(lib.ann_ensemble_num_steps(self._h, ref(n)))
        return n.value

    def last_samples(self):
        """Configurations of the last sampling call, shape [num_steps, words] (uint64)."""
        cnt, words = ctypes.c_int64(), ctypes.c_int32()
        check(lib.ann_ensemble_last_samples(self._h, None, 0, ref(cnt), ref(words)))
        out = np.zeros((cnt.value, words.value), dtype=np.uint64)
        check(lib.ann_ensemble_last_samples(self._h, ptr(out), out.size, ref(cnt), ref(words)))
        return out


class MonteCarloSpins(_Ensemble):
    """MonteCarloSpins(num_samples, num_sweeps, num_thermalization_sweeps, num_markov_chains, gpu)
    -- include/ensembles/MonteCarlo.hpp:249-259.

    Additive keywords: ``move`` ('flip' or 'exchange'), ``total_z`` (2 S_z of the
    sector used by exchange moves), ``seed`` (Philox key), ``persistent`` (keep the
    chains between calls instead of re-initialising and re-thermalising them).
    """

    def __init__(self, num_samples, num_sweeps, num_thermalization_sweeps, num_markov_chains, gpu=True,
                 move="flip", total_z=0, seed=0, persistent=False):
        mv = {"flip": _lib.MOVE_FLIP, "exchange": _lib.MOVE_EXCHANGE}[move]
        self.move = move
        self._h = handle()
        check(lib.ann_monte_carlo_create(ref(self._h), int(num_samples), int(num_sweeps),
                                         int(num_thermalization_sweeps), int(num_markov_chains), mv, int(total_z),
                                         int(seed) & 0xFFFFFFFFFFFFFFFF))
        self.persistent = persistent

    @property
    def persistent(self):
        return self._persistent

    @persistent.setter
    def persistent(self, value):
        check(lib.ann_monte_carlo_set_persistent(self._h, 1 if value else 0))
        self._persistent = bool(value)

    @property
    def acceptance_rate(self):
        a, r = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib.ann_ensemble_acceptance(self._h, ref(a), ref(r)))
        tot = a.value + r.value
        return a.value / tot if tot else float("nan")


    @property
    def block_stats(self):
        """block moves of the last sampling call with a learned proposal: proposals made,
        redraws of the current pattern (null moves; acceptance_rate counts them as not
        accepted, since they do not move the chain) and flipped sites whose rows of W
        entered a psi ratio"""
        p, n, r = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib.ann_ensemble_block_stats(self._h, ref(p), ref(n), ref(r)))
        return dict(proposals=p.value, null_moves=n.value, rows_evaluated=r.value)

    def set_proposal(self, proposal, plain_sweeps=1, exclude_current=False):
        """Attach a trained BlockProposal (None detaches).  Each sample interval then
        holds num_sweeps block sweeps plus ``plain_sweeps`` plain sweeps.  With
        ``exclude_current`` a block move never redraws the current pattern (Metropolized
        Gibbs form of the proposal, exact under the corrected Metropolis-Hastings ratio)."""
        check(lib.ann_monte_carlo_set_proposal(self._h, proposal._h if proposal is not None else None,
                                               int(plain_sweeps)))
        check(lib.ann_monte_carlo_set_proposal_mode(self._h, 1 if exclude_current else 0))
        self.exclude_current = bool(exclude_current)
        self._proposal = proposal      # keeps the network alive while attached
        self.plain_sweeps = int(plain_sweeps)

    @property
    def proposal(self):
        return getattr(self, "_proposal", None)


class BlockProposal:
    """Learned S_z-preserving block proposal (additive; the reference has none).

    ``blocks`` is a [num_blocks][block_size] array of site indices (block_size <= 8).
    The network N -> hidden -> N (tanh) sees the configuration with the block's
    spins masked and scores the block's sites; a move redraws the block among the
    patterns with the same number of up spins.  Training (``fit``) minimises the
    block pseudo-likelihood on samples the sampler itself produced.
    """

    def __init__(self, num_sites, blocks, hidden=256, seed=0):
        blocks = np.ascontiguousarray(blocks, dtype=np.int32)
        if blocks.ndim != 2:
            raise ValueError("blocks must be [num_blocks][block_size]")
        self.num_sites = int(num_sites)
        self.hidden = int(hidden)
        self.blocks = blocks.copy()
        self._h = handle()
        check(lib.ann_proposal_create(ref(self._h), self.num_sites, self.hidden, blocks.shape[0], blocks.shape[1],
                                      ptr(blocks), int(seed) & 0xFFFFFFFFFFFFFFFF))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib.ann_proposal_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def num_params(self):
        n = ctypes.c_int64()
        check(lib.ann_proposal_num_params(self._h, ref(n)))
        return n.value

    @property
    def params(self):
        out = np.zeros(self.num_params, dtype=np.float32)
        check(lib.ann_proposal_get_params(self._h, ptr(out)))
        return out

    @params.setter
    def params(self, value):
        v = np.ascontiguousarray(value, dtype=np.float32).ravel()
        if v.size != self.num_params:
            raise ValueError("expected %d parameters" % self.num_params)
        check(lib.ann_proposal_set_params(self._h, ptr(v)))

    def loss_grad(self, confs, sites):
        """loss and gradient on explicit rows: confs [count][words] (or [count]), sites [count][block_size]"""
        c = np.ascontiguousarray(confs, dtype=np.uint64)
        if c.ndim == 1:
            c = c[:, None]
        st = np.ascontiguousarray(sites, dtype=np.int32)
        grad = np.zeros(self.num_params, dtype=np.float32)
        loss = ctypes.c_double()
        check(lib.ann_proposal_loss_grad(self._h, ptr(c), c.shape[0], c.shape[1], ptr(st), ref(loss), ptr(grad)))
        return loss.value, grad

    def adam_step(self, grad, lr):
        g = np.ascontiguousarray(grad, dtype=np.float32)
        check(lib.ann_proposal_adam_step(self._h, ptr(g), float(lr)))

    def fit(self, ensemble, steps, batch=4096, lr=1e-3):
        """Adam on random (sample, block) rows of the ensemble's last samples; returns the last batch loss"""
        loss = ctypes.c_double()
        check(lib.ann_proposal_train(self._h, ensemble._h, int(steps), int(batch), float(lr), ref(loss)))
        return loss.value


def rect_blocks(L, shape=(2, 4), both_orientations=True):
    """all placements of a shape[0] x shape[1] rectangle on a periodic L[0] x L[1]
    lattice (site r * L[1] + c), optionally also the transposed rectangle"""
    Lx, Ly = (L, L) if isinstance(L, int) else L
    shapes = [tuple(shape)]
    if both_orientations and shape[0] != shape[1]:
        shapes.append((shape[1], shape[0]))
    blocks = []
    for a, b in shapes:
        for r in range(Lx):
            for c in range(Ly):
                blocks.append([((r + i) % Lx) * Ly + (c + j) % Ly for i in range(a) for j in range(b)])
    return np.array(blocks, dtype=np.int32)


def chain_blocks(N, size=8):
    """all windows of ``size`` consecutive sites of a periodic chain"""
    return np.array([[(i + j) % N for j in range(size)] for i in range(N)], dtype=np.int32)


class ExactSummationSpins(_Ensemble):
    """ExactSummationSpins(num_sites, gpu) -- source/ensembles/ExactSummation.cu:16-25."""

    def __init__(self, num_sites, gpu=True):
        self._h = handle()
        check(lib.ann_exact_summation_create(ref(self._h), int(num_sites)))
        self.num_sites = int(num_sites)


class ExactSummationPaulis(_Ensemble):
    """ExactSummationPaulis(num_sites, gpu) -- include/ensembles/ExactSummation.hpp:125: the
    4^N Pauli strings in PauliString.enumerate order."""

    def __init__(self, num_sites, gpu=True):
        self._h = handle()
        check(lib.ann_exact_summation_paulis_create(ref(self._h), int(num_sites)))
        self.num_sites = int(num_sites)


class MonteCarloPaulis(_Ensemble):
    """MonteCarloPaulis(num_samples, num_sweeps, num_thermalization_sweeps, num_markov_chains, gpu)
    -- include/ensembles/MonteCarlo.hpp:265-283: Metropolis moves that set a random site to a
    random Pauli type (policies/Update_Policy.hpp:41-56), num_sweeps x num_sites moves
    between samples.  Additive: ``seed``."""

    def __init__(self, num_samples, num_sweeps, num_thermalization_sweeps, num_markov_chains, gpu=True, seed=0):
        self._h = handle()
        check(lib.ann_monte_carlo_paulis_create(ref(self._h), int(num_samples), int(num_sweeps),
                                                int(num_thermalization_sweeps), int(num_markov_chains),
                                                int(seed) & 0xFFFFFFFFFFFFFFFF))

    @property
    def acceptance_rate(self):
        a, r = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib.ann_ensemble_acceptance(self._h, ref(a), ref(r)))
        tot = a.value + r.value
        return a.value / tot if tot else float("nan")


class PauliSampleList(_Ensemble):
    """Additive: an ensemble of given Pauli strings (``PauliString`` objects or rows (a, b))
    with weights (default 1 / count)."""

    def __init__(self, num_sites, strings, weights=None, gpu=True):
        rows = [(s.a, s.b) if isinstance(s, PauliString) else (int(s[0]), int(s[1])) for s in strings]
        arr = np.ascontiguousarray(np.array(rows, dtype=np.uint64).reshape(-1, 2))
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64).ravel()
        if w is not None and w.size != arr.shape[0]:
            raise ValueError("one weight per configuration")
        self._h = handle()
        check(lib.ann_pauli_sample_list_create(ref(self._h), int(num_sites), ptr(arr), arr.shape[0], ptr(w)))
        self.num_sites = int(num_sites)


class SampleList(_Ensemble):
    """Additive (no counterpart in the reference): an ensemble of given configurations.

    ``configurations`` are integers, ``Spins`` or rows of uint64 words (bit i set
    <=> s_i = +1); ``weights`` default to 1 / count.  Every network function treats
    them as one sampling call's samples, so two wave functions, two precisions or a
    run and a re-run see exactly the same samples.
    """

    def __init__(self, num_sites, configurations, weights=None, gpu=True):
        words = _words_for(num_sites)
        confs = np.asarray(configurations)
        if confs.dtype == np.uint64 and confs.ndim == 2:
            arr = np.zeros((confs.shape[0], words), dtype=np.uint64)
            k = min(words, confs.shape[1])
            arr[:, :k] = confs[:, :k]
        else:
            arr = np.stack([_conf_words(c)[:words] for c in configurations]) if len(configurations) else \
                np.zeros((0, words), dtype=np.uint64)
        arr = np.ascontiguousarray(arr, dtype=np.uint64)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64).ravel()
        if w is not None and w.size != arr.shape[0]:
            raise ValueError("one weight per configuration")
        self._h = handle()
        check(lib.ann_sample_list_create(ref(self._h), int(num_sites), ptr(arr), arr.shape[0], words, ptr(w)))
        self.num_sites = int(num_sites)


# --------------------------------------------------------- network functions

def log_psi_s(psi, spins):
    """PsiVector.cu.template:101-133 (spins or a PauliString)"""
    out = np.zeros(2)
    if isinstance(spins, PauliString):
        check(lib.ann_log_psi_s_pauli(psi._h, ptr(spins.words()), ptr(out)))
    else:
        check(lib.ann_log_psi_s(psi._h, ptr(_conf_words(spins)), ptr(out)))
    return complex(out[0], out[1])


def psi_O_k(psi, spins):
    """PsiOkVector.cu.template:40-74 (spins or a PauliString)"""
    out = np.zeros(psi.num_params, dtype=np.complex128)
    if isinstance(spins, PauliString):
        check(lib.ann_psi_O_k_pauli(psi._h, ptr(spins.words()), ptr(out)))
    else:
        check(lib.ann_psi_O_k(psi._h, ptr(_conf_words(spins)), ptr(out)))
    return out


def psi_O_k_vector(psi, ensemble):
    """PsiOkVector.cu.template:8-37 (unweighted sum over the ensemble)"""
    out = np.zeros(psi.num_params, dtype=np.complex128)
    check(lib.ann_psi_O_k_vector(psi._h, ensemble._h, ptr(out)))
    return out


def log_psi(psi, ensemble):
    """PsiVector.cu.template:73-99 (sum_s w_s log psi(s))"""
    out = np.zeros(2)
    check(lib.ann_log_psi(psi._h, ensemble._h, ptr(out)))
    return complex(out[0], out[1])


def psi_vector(psi, ensemble):
    out = np.zeros(ensemble.num_steps, dtype=np.complex128)
    check(lib.ann_psi_vector(psi._h, ensemble._h, ptr(out)))
    return out


def log_psi_vector(psi, ensemble):
    out = np.zeros(ensemble.num_steps, dtype=np.complex128)
    check(lib.ann_log_psi_vector(psi._h, ensemble._h, ptr(out)))
    return out


def apply_operator(psi, op, ensemble):
    """ApplyOperator.cu.template:9-44: psi(s) * E_loc(s) = (H psi)(s)"""
    out = np.zeros(ensemble.num_steps, dtype=np.complex128)
    check(lib.ann_apply_operator(psi._h, op._h, ensemble._h, ptr(out)))
    return out


def local_energies(op, psi, ensemble):
    """Additive: E_loc(s) for every sample of one sampling call."""
    out = np.zeros(ensemble.num_steps, dtype=np.complex128)
    check(lib.ann_local_energies(op._h, psi._h, ensemble._h, ptr(out), out.size))
    return out


def activation_function(x, layer=0):
    """main.cpp.template:534-536"""
    out = np.zeros(2)
    check(lib.ann_activation_function(ptr(c2(x)), int(layer), ptr(out)))
    return complex(out[0], out[1])


def setDevice(device):
    check(lib.ann_set_device(int(device)))


def start_profiling():
    check(lib.ann_profiler_start())


def stop_profiling():
    check(lib.ann_profiler_stop())


def profiler_query(name):
    ms, n = ctypes.c_double(), ctypes.c_int64()
    check(lib.ann_profiler_query(name.encode(), ref(ms), ref(n)))
    return ms.value, n.value


def profiler_reset():
    check(lib.ann_profiler_reset())


class ExpectationValue:
    """ExpectationValue(gpu) -- network_functions/ExpectationValue.hpp:21-42"""

    def __init__(self, gpu=True):
        pass

    def __call__(self, operator_, psi, *args):
        """(operator, psi, ensemble), (operator list, psi, ensemble), or (operator, psi,
        psi_sampling, ensemble): <operator> in psi from the samples of psi_sampling, weighted
        by |psi / psi_sampling|^2 (ExpectationValue.hpp:51-59, ExpectationValue.cu.template:
        130-172; the reference divides by an accumulator it never fills, here the weights'
        sum)"""
        if len(args) == 2:
            psi_sampling, ensemble = args
            out = np.zeros(2)
            check(lib.ann_expectation_value_reweighted(operator_._h, psi._h, psi_sampling._h, ensemble._h, ptr(out)))
            return complex(out[0], out[1])
        if len(args) != 1:
            raise TypeError("ExpectationValue(operator, psi, [psi_sampling,] ensemble)")
        ensemble = args[0]
        if isinstance(operator_, (list, tuple)):
            return np.array([self(op, psi, ensemble) for op in operator_])
        out = np.zeros(2)
        check(lib.ann_expectation_value(operator_._h, psi._h, ensemble._h, ptr(out)))
        return complex(out[0], out[1])

    def exp_sigma_z(self, operator_, psi, ensemble):
        """ExpectationValue.cu.template:52-77: the weighted sum over the samples of psi of
        exp(Operator::fast_local_energy(s)), the matrix-element sum of the operator's strings"""
        out = np.zeros(2)
        check(lib.ann_expectation_exp_sigma_z(operator_._h, psi._h, ensemble._h, ptr(out)))
        return complex(out[0], out[1])

    def fluctuation(self, operator_, psi, ensemble):
        sigma = ctypes.c_double()
        out = np.zeros(2)
        check(lib.ann_expectation_fluctuation(operator_._h, psi._h, ensemble._h, ref(sigma), ptr(out)))
        return sigma.value, complex(out[0], out[1])

    def gradient(self, operator_, psi, ensemble):
        g = np.zeros(psi.num_params, dtype=np.complex128)
        out = np.zeros(2)
        check(lib.ann_expectation_gradient(operator_._h, psi._h, ensemble._h, ptr(g), ptr(out)))
        return g, complex(out[0], out[1])

    def gradient_with_noise(self, operator_, psi, ensemble):
        """MAIN:427 binds this name to ExpectationValue::gradient (the noise variant's body is
        commented out in ExpectationValue.cu.template:278-337): the same (gradient, value)"""
        return self.gradient(operator_, psi, ensemble)


class HilbertSpaceDistance:
    """HilbertSpaceDistance(num_params, gpu) -- network_functions/HilbertSpaceDistance.hpp,
    bound at MAIN:436-441.  distance(psi', U psi) on psi's samples and its gradient with
    respect to psi''s parameters."""

    def __init__(self, num_params, gpu=True):
        self.num_params = int(num_params)

    def __call__(self, psi, psi_prime, operator_, is_unitary, spin_ensemble):
        d = ctypes.c_double()
        check(lib.ann_hilbert_space_distance(psi._h, psi_prime._h, operator_._h, 1 if is_unitary else 0,
                                             spin_ensemble._h, 1.0, None, ref(d)))
        return d.value

    def gradient(self, psi, psi_prime, operator_, is_unitary, spin_ensemble, nu):
        """(gradient, distance); the gradient of the distance itself for nu = 1"""
        g = np.zeros(psi_prime.num_params, dtype=np.complex128)
        d = ctypes.c_double()
        check(lib.ann_hilbert_space_distance(psi._h, psi_prime._h, operator_._h, 1 if is_unitary else 0,
                                             spin_ensemble._h, float(nu), ptr(g), ref(d)))
        return g, d.value


class KullbackLeibler:
    """KullbackLeibler(num_params, gpu) -- network_functions/KullbackLeibler.hpp:14-117, bound at
    MAIN:445-462.  psi is a classical network, psi_prime the network whose ensemble is
    sampled and whose parameters the gradient is taken with respect to."""

    def __init__(self, num_params, gpu=True):
        self.num_params = int(num_params)
        self._h = handle()
        check(lib.ann_kl_create(ref(self._h), self.num_params))

    def __del__(self):
        if self._h is not None and self._h.value:
            try:
                lib.ann_kl_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def _eval(self, psi, psi_prime, ensemble, threshold, nu, mode):
        g = np.zeros(psi_prime.num_params, dtype=np.complex128) if mode > 0 else None
        n = np.zeros(psi_prime.num_params, dtype=np.float64) if mode == 2 else None
        v = ctypes.c_double()
        check(lib.ann_kl_eval(self._h, psi._h, psi_prime._h, ensemble._h, float(threshold), float(nu), int(mode),
                              ptr(g), ptr(n), ref(v)))
        return g, n, v.value

    def __call__(self, psi, psi_prime, ensemble, threshold=0.0):
        """KullbackLeibler.cu.template:171-196"""
        return self._eval(psi, psi_prime, ensemble, threshold, 1.0, 0)[2]

    def gradient(self, psi, psi_prime, ensemble, nu, threshold=0.0):
        """(gradient, value) -- KullbackLeibler.cu.template:199-242, gradient_py"""
        g, _, v = self._eval(psi, psi_prime, ensemble, threshold, nu, 1)
        return g, v

    def gradient_with_noise(self, psi, psi_prime, ensemble, nu, threshold=0.0):
        """(gradient, noise, value) -- KullbackLeibler.cu.template:245-316"""
        return self._eval(psi, psi_prime, ensemble, threshold, nu, 2)

    def _info(self):
        w = ctypes.c_double()
        m = np.zeros(2)
        check(lib.ann_kl_info(self._h, ref(w), ptr(m)))
        return w.value, complex(m[0], m[1])

    @property
    def total_weight(self):
        return self._info()[0]

    @property
    def mean_deviation(self):
        return self._info()[1]

    @property
    def log_psi_scale(self):
        x = ctypes.c_double()
        check(lib.ann_kl_get_log_psi_scale(self._h, ref(x)))
        return x.value

    @log_psi_scale.setter
    def log_psi_scale(self, value):
        check(lib.ann_kl_set_log_psi_scale(self._h, float(value)))


class TDVP:
    """TDVP(num_params, gpu) -- network_functions/TDVP.hpp:16-101.

    Additive: ``solve(shift, method='cholesky'|'cg')`` for (S + shift I) x = F.
    """

    def __init__(self, num_params, gpu=True):
        self.num_params = int(num_params)
        self._h = handle()
        check(lib.ann_tdvp_create(ref(self._h), self.num_params))

    def __del__(self):
        if self._h is not None and self._h.value:
            try:
                lib.ann_tdvp_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def eval(self, op, psi, ensemble):
        check(lib.ann_tdvp_eval(self._h, op._h, psi._h, ensemble._h, 1))

    def eval_F(self, op, psi, ensemble):
        check(lib.ann_tdvp_eval(self._h, op._h, psi._h, ensemble._h, 0))

    @property
    def S_matrix(self):
        P = self.num_params
        out = np.zeros((P, P), dtype=np.complex128)
        check(lib.ann_tdvp_get_S(self._h, ptr(out)))
        return out

    @property
    def F_vector(self):
        out = np.zeros(self.num_params, dtype=np.complex128)
        check(lib.ann_tdvp_get_F(self._h, ptr(out)))
        return out

    @property
    def O_k_vector(self):
        out = np.zeros(self.num_params, dtype=np.complex128)
        check(lib.ann_tdvp_get_O(self._h, ptr(out)))
        return out

    def _energy(self):
        e = np.zeros(2)
        e2, var = ctypes.c_double(), ctypes.c_double()
        check(lib.ann_tdvp_get_energy(self._h, ptr(e), ref(e2), ref(var)))
        return complex(e[0], e[1]), e2.value, var.value

    @property
    def E_local(self):
        return self._energy()[0]

    @property
    def E2_local(self):
        return self._energy()[1]

    @property
    def var_H(self):
        return self._energy()[2]

    # threshold gates the reference's evaluation against a reference wave function
    # (eval_with_psi_ref, PsiClassical); the evaluations here do not read it
    threshold = -1e6

    def _samples(self):
        cnt, words, tw = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_double()
        check(lib.ann_tdvp_get_samples(self._h, None, None, None, 0, ref(cnt), ref(words), ref(tw)))
        return cnt.value, words.value, tw.value

    @property
    def total_weight(self):
        """sum of the sample weights of the last evaluation (MAIN:471)"""
        return self._samples()[2]

    @property
    def E_local_samples(self):
        """local energy of every sample of the last evaluation (MAIN:473)"""
        n, _, _ = self._samples()
        out = np.zeros(n, dtype=np.complex128)
        check(lib.ann_tdvp_get_samples(self._h, None, None, ptr(out), n, None, None, None))
        return out

    @property
    def samples(self):
        """(configurations [count][words] uint64, weights [count]) of the last evaluation"""
        n, words, _ = self._samples()
        confs = np.zeros((n, words), dtype=np.uint64)
        w = np.zeros(n)
        check(lib.ann_tdvp_get_samples(self._h, ptr(confs), ptr(w), None, n, None, None, None))
        return confs, w

    @property
    def O_k_samples(self):
        """uncentred O_k of every sample of the last evaluation, [count][num_params] (MAIN:472;
        the reference returns it flattened)"""
        n, _, _ = self._samples()
        out = np.zeros((n, self.num_params), dtype=np.complex128)
        check(lib.ann_tdvp_get_O_samples(self._h, ptr(out), out.size))
        return out

    def S_dot_vector(self, vec, ensemble=None):
        v = as_c128(vec).ravel()
        out = np.zeros(self.num_params, dtype=np.complex128)
        check(lib.ann_tdvp_S_dot_vector(self._h, ptr(v), ptr(out)))
        return out

    def solve(self, shift=1e-3, method="cholesky", rtol=1e-6, max_iter=500, raise_on_failure=True, x0=None):
        """(S + shift) x = F in fp64 by Cholesky of the dense S ("cholesky"), or by CG on
        the matrix-free S with a Jacobi ("cg") or Kronecker-factored ("cg_kfac": spin x
        hidden-unit covariance factors, dense nets of up to 128 sites) preconditioner;
        ``x0`` starts CG from a previous solution"""
        m = {"cholesky": _lib.SOLVE_CHOLESKY, "cg": _lib.SOLVE_CG, "cg_kfac": _lib.SOLVE_CG_KFAC}[method]
        x = np.zeros(self.num_params, dtype=np.complex128)
        if x0 is not None:
            if method == "cholesky":
                raise ValueError("a starting vector applies to CG only")
            x[:] = np.asarray(x0, dtype=np.complex128).ravel()
            m = _lib.SOLVE_CG_WARM if method == "cg" else _lib.SOLVE_CG_KFAC_WARM
        it, res = ctypes.c_int32(), ctypes.c_double()
        rc = lib.ann_tdvp_solve(self._h, float(shift), m, float(rtol), int(max_iter), ptr(x), ref(it), ref(res))
        self.last_solve = dict(iterations=it.value, residual=res.value)
        if rc != 0 and (raise_on_failure or rc != _lib.ANN_E_NUMERIC):
            check(rc)
        return x
