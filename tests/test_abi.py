"""CPU checks of the boundary: the library loads and exports every entry point
declared in include/annongpu.h, the Python bindings cover them, and host-side
validation rejects bad input without touching the GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "annongpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ann_[A-Za-z0-9_]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(ROOT, "annongpu_amd", "libannongpu.so"))
    names = declared_symbols()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing


def test_python_binding_covers_the_header():
    from annongpu_amd import _lib
    assert set(declared_symbols()) - {"ann_last_error"} == set(_lib.symbols())


def test_host_validation_errors():
    import annongpu_amd as A
    with pytest.raises(ValueError):
        A.PsiRBM(np.zeros((0, 4)))
    with pytest.raises(ValueError):
        A.MonteCarloSpins(10, 1, 1, 0)
    with pytest.raises(ValueError):
        A.ExactSummationSpins(40)
    # connection index beyond the visible layer
    with pytest.raises(ValueError):
        A.PsiDeep(4, np.zeros(4), [np.zeros(4)], [np.full((4, 4), 7)], [np.zeros((4, 4))], np.ones(4))
    psi = A.PsiRBM(np.zeros((3, 5)) + 0.1, 2.0, 0.5j)
    assert psi.num_params == 15
    assert psi.log_prefactor == 0.5j
    p = np.arange(15) * (1 + 1j)
    psi.params = p
    assert np.array_equal(psi.params, p)
    with pytest.raises(ValueError):
        psi.params = np.zeros(3)


def test_params_roundtrip_and_json():
    import annongpu_amd as A
    psi = A.seeded_rbm_deep(6, 12, 0)
    js = psi.to_json()
    psi2 = A.PsiDeep.from_json(js)
    assert np.array_equal(psi.params, psi2.params)
    assert psi.num_params == 6 + 12 + 6 * 12
    assert np.array_equal(psi2.connections[0], psi.connections[0])
    cnn = A.seeded_cnn([4, 4], [(2, [3, 3]), (2, [3, 3])], 1)
    cnn2 = A.PsiCNN.from_json(cnn.to_json())
    assert np.array_equal(cnn.params, cnn2.params)
    assert cnn.extent == [1, 4, 4]


def test_activation_function_matches_oracle():
    import annongpu_amd as A
    import pyoracle as orc
    for z in (0.1 + 0.2j, -0.4 + 0.05j):
        for layer in (0, 1):
            assert A.activation_function(z, layer) == pytest.approx(orc.activation(z, layer), rel=1e-14)


def test_kfac_solve_rejects_non_positive_shift():
    """The K-FAC damping scales with sqrt(shift); shift <= 0 would invert the singular
    spin covariance of S_z-conserving samples, so the boundary refuses it up front."""
    from annongpu_amd import _lib
    x = np.zeros(4, dtype=np.complex128)
    it, res = ctypes.c_int32(), ctypes.c_double()
    for method in (_lib.SOLVE_CG_KFAC, _lib.SOLVE_CG_KFAC_WARM):
        for shift in (0.0, -1e-3, float("nan")):
            rc = _lib.lib.ann_tdvp_solve(None, shift, method, 1e-6, 10, x.ctypes.data,
                                         ctypes.byref(it), ctypes.byref(res))
            assert rc == _lib.ANN_E_INVALID
            assert b"shift" in _lib.lib.ann_last_error()
    rc = _lib.lib.ann_tdvp_solve(None, float("inf"), _lib.SOLVE_CG, 1e-6, 10, x.ctypes.data, None, None)
    assert rc == _lib.ANN_E_INVALID and b"shift" in _lib.lib.ann_last_error()
    # a valid shift with no handle fails on the handle, not on the shift
    rc = _lib.lib.ann_tdvp_solve(None, 1e-3, _lib.SOLVE_CG_KFAC, 1e-6, 10, x.ctypes.data, None, None)
    assert rc == _lib.ANN_E_INVALID and b"tdvp" in _lib.lib.ann_last_error()
