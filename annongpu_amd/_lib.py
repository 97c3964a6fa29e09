"""Loader of libannongpu.so, the gfx950 HIP library behind include/annongpu.h.

The library is built in-tree (``make -C annongpu_amd``).  There is no CPU
fallback: if the shared library is missing, importing the package raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libannongpu.so")

ANN_OK, ANN_E_INVALID, ANN_E_DEVICE, ANN_E_UNSUPPORTED, ANN_E_NUMERIC = 0, 1, 2, 3, 4
FP64, FP32 = 0, 1
MOVE_FLIP, MOVE_EXCHANGE, MOVE_PAULI = 0, 1, 2
SOLVE_CHOLESKY, SOLVE_CG, SOLVE_CG_WARM, SOLVE_CG_KFAC, SOLVE_CG_KFAC_WARM = 0, 1, 2, 3, 4

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()' "
        "or make -C annongpu_amd)")

lib = ctypes.CDLL(LIB_PATH)
lib.ann_last_error.restype = ctypes.c_char_p

P = ctypes.c_void_p
i32, u32, i64, u64, f64 = ctypes.c_int32, ctypes.c_uint32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double

_SIGNATURES = {
    "ann_set_device": [ctypes.c_int],
    "ann_profiler_start": [],
    "ann_profiler_stop": [],
    "ann_profiler_query": [ctypes.c_char_p, P, P],
    "ann_profiler_reset": [],
    "ann_psi_create_rbm": [P, i32, i32, P, P, P],
    "ann_psi_create_deep": [P, i32, i32, P, i32, P, P, P, P, P, P, P],
    "ann_psi_create_cnn": [P, P, i32, P, P, P, i32, P, f64, P],
    "ann_psi_copy": [P, P],
    "ann_psi_destroy": [P],
    "ann_psi_info": [P, P, P, P],
    "ann_psi_get_params": [P, P],
    "ann_psi_set_params": [P, P],
    "ann_psi_get_log_prefactor": [P, P],
    "ann_psi_set_log_prefactor": [P, P],
    "ann_psi_set_precision": [P, i32],
    "ann_operator_create": [P, i32, i32, P, P, P],
    "ann_operator_destroy": [P],
    "ann_monte_carlo_create": [P, u32, u32, u32, u32, i32, i32, u64],
    "ann_monte_carlo_set_persistent": [P, i32],
    "ann_exact_summation_create": [P, u32],
    "ann_ensemble_destroy": [P],
    "ann_ensemble_num_steps": [P, P],
    "ann_ensemble_acceptance": [P, P, P],
    "ann_ensemble_last_samples": [P, P, i64, P, P],
    "ann_log_psi_s": [P, P, P],
    "ann_psi_O_k": [P, P, P],
    "ann_psi_O_k_vector": [P, P, P],
    "ann_log_psi": [P, P, P],
    "ann_psi_vector": [P, P, P],
    "ann_log_psi_vector": [P, P, P],
    "ann_psi_norm": [P, P, P],
    "ann_apply_operator": [P, P, P, P],
    "ann_activation_function": [P, i32, P],
    "ann_expectation_value": [P, P, P, P],
    "ann_expectation_fluctuation": [P, P, P, P, P],
    "ann_expectation_gradient": [P, P, P, P, P],
    "ann_expectation_exp_sigma_z": [P, P, P, P],
    "ann_expectation_value_reweighted": [P, P, P, P, P],
    "ann_local_energies": [P, P, P, P, i64],
    "ann_tdvp_create": [P, u32],
    "ann_tdvp_destroy": [P],
    "ann_tdvp_eval": [P, P, P, P, i32],
    "ann_tdvp_get_S": [P, P],
    "ann_tdvp_get_F": [P, P],
    "ann_tdvp_get_O": [P, P],
    "ann_tdvp_get_energy": [P, P, P, P],
    "ann_tdvp_S_dot_vector": [P, P, P],
    "ann_tdvp_solve": [P, f64, i32, f64, i32, P, P, P],
    "ann_proposal_create": [P, i32, i32, i32, i32, P, u64],
    "ann_proposal_destroy": [P],
    "ann_proposal_num_params": [P, P],
    "ann_proposal_get_params": [P, P],
    "ann_proposal_set_params": [P, P],
    "ann_proposal_loss_grad": [P, P, i64, i32, P, P, P],
    "ann_proposal_adam_step": [P, P, f64],
    "ann_proposal_train": [P, P, i32, i32, f64, P],
    "ann_monte_carlo_set_proposal": [P, P, u32],
    "ann_monte_carlo_set_proposal_mode": [P, i32],
    "ann_sample_list_create": [P, u32, P, i64, i32, P],
    "ann_ensemble_block_stats": [P, P, P, P],
    "ann_tdvp_get_samples": [P, P, P, P, i64, P, P, P],
    "ann_tdvp_get_O_samples": [P, P, i64],
    "ann_local_energies_noalloc": [P, P, P, i64, P],
    "ann_hilbert_space_distance": [P, P, P, i32, P, f64, P, P],
    "ann_psi_create_classical": [P, i32, i32, P, P, P, i32, P, P],
    "ann_psi_classical_info": [P, P, P, P],
    "ann_psi_classical_ref": [P, P],
    "ann_psi_get_ref_log_prefactor": [P, P],
    "ann_psi_set_ref_log_prefactor": [P, P],
    "ann_exact_summation_paulis_create": [P, u32],
    "ann_monte_carlo_paulis_create": [P, u32, u32, u32, u32, u64],
    "ann_pauli_sample_list_create": [P, u32, P, i64, P],
    "ann_log_psi_s_pauli": [P, P, P],
    "ann_psi_O_k_pauli": [P, P, P],
    "ann_kl_create": [P, u32],
    "ann_kl_destroy": [P],
    "ann_kl_eval": [P, P, P, P, f64, f64, i32, P, P, P],
    "ann_kl_info": [P, P, P],
    "ann_kl_get_log_psi_scale": [P, P],
    "ann_kl_set_log_psi_scale": [P, f64],
}

for _name, _args in _SIGNATURES.items():
    _fn = getattr(lib, _name)
    _fn.argtypes = _args
    _fn.restype = ctypes.c_int


class AnnError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


def check(rc):
    if rc != ANN_OK:
        msg = lib.ann_last_error().decode(errors="replace")
        if rc == ANN_E_INVALID:
            raise ValueError(msg)
        raise AnnError(rc, msg)


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def ref(x):
    return ctypes.byref(x)


def as_c128(x, shape=None):
    a = np.ascontiguousarray(x, dtype=np.complex128)
    if shape is not None:
        a = a.reshape(shape)
    return a


def c2(z):
    z = complex(z)
    return np.array([z.real, z.imag], dtype=np.float64)


def handle():
    return ctypes.c_void_p()


def symbols():
    """Names of every entry point bound here (they mirror include/annongpu.h)."""
    return list(_SIGNATURES)
