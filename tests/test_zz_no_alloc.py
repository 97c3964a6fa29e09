"""ann_local_energies_noalloc (device pointers in and out, angles in the library's
static scratch array, no allocation) against ann_local_energies on the same
configurations: bit for bit, on two calls in a row, the second with fewer
configurations, in both precisions; plus its size limit and preconditions."""
import ctypes

import numpy as np
import pytest

import annongpu_amd as A
from annongpu_amd import _lib

pytestmark = pytest.mark.gpu

# the HIP runtime the library itself is linked against (device memory for the test)
_hip = ctypes.CDLL("libamdhip64.so.7")
_hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
_hip.hipFree.argtypes = [ctypes.c_void_p]
_hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
H2D, D2H = 1, 2

L, N, M = 10, 100, 400
SCRATCH = 16 << 20


class DeviceArray:
    def __init__(self, nbytes):
        self.p = ctypes.c_void_p()
        assert _hip.hipMalloc(ctypes.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def __del__(self):
        if self.p:
            _hip.hipFree(self.p)
            self.p = None


def sz0_confs(count, seed):
    rng = np.random.default_rng(seed)
    out = np.zeros((count, 2), dtype=np.uint64)
    for r in range(count):
        for i in rng.permutation(N)[: N // 2]:
            out[r, i >> 6] |= np.uint64(1 << (int(i) & 63))
    return out


def noalloc(op, psi, confs):
    count = confs.shape[0]
    dc = DeviceArray(confs.nbytes)
    de = DeviceArray(16 * count)
    assert _hip.hipMemcpy(dc.p, confs.ctypes.data, confs.nbytes, H2D) == 0
    rc = _lib.lib.ann_local_energies_noalloc(op._h, psi._h, dc.p, count, de.p)
    _lib.check(rc)
    out = np.zeros(count, dtype=np.complex128)
    assert _hip.hipMemcpy(out.ctypes.data, de.p, out.nbytes, D2H) == 0
    return out


@pytest.mark.parametrize("precision,counts", [("fp32", (4096, 1000)), ("fp64", (2048, 500))])
def test_noalloc_local_energies_match(precision, counts):
    psi = A.seeded_rbm_deep(N, M, seed=3, sigma=0.05)
    psi.precision = precision
    op = A.Operator(A.hamiltonians.heisenberg_square(L))
    confs = sz0_confs(max(counts), 7)
    A.local_energies(op, psi, A.SampleList(N, confs[:8]))   # uploads the net and the operator tables
    for k, count in enumerate(counts):
        c = np.ascontiguousarray(confs[:count][::-1] if k else confs[:count])
        ref = A.local_energies(op, psi, A.SampleList(N, c))
        got = noalloc(op, psi, c)
        assert np.array_equal(got, ref), (count, np.abs(got - ref).max())


def test_noalloc_limits_and_preconditions():
    op = A.Operator(A.hamiltonians.heisenberg_square(L))
    fresh = A.seeded_rbm_deep(N, M, seed=4)
    fresh.precision = "fp32"
    confs = sz0_confs(16, 1)
    # parameters never uploaded: refused (uploading would allocate)
    with pytest.raises(_lib.AnnError):
        noalloc(op, fresh, confs)
    A.log_psi_s(fresh, A.Spins(0, N))
    # operator tables not built yet: refused
    with pytest.raises(_lib.AnnError):
        noalloc(op, fresh, confs)
    A.local_energies(op, fresh, A.SampleList(N, confs))
    per = 448 * 8    # padded hidden units x complex float
    too_many = SCRATCH // per + 1
    with pytest.raises(ValueError, match="scratch"):
        rc = _lib.lib.ann_local_energies_noalloc(op._h, fresh._h, ctypes.c_void_p(1), too_many, ctypes.c_void_p(1))
        _lib.check(rc)
    # at the limit itself it runs
    c = sz0_confs(SCRATCH // per, 2)
    assert np.array_equal(noalloc(op, fresh, c), A.local_energies(op, fresh, A.SampleList(N, c)))
