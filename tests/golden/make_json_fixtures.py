"""Writes the JSON fixtures in the reference's on-disk format (tests/golden/*.json).

The reference stores wave functions with ``to_json`` (pyANNonGPU/PsiDeep.py:7-23,
pyANNonGPU/PsiCNN.py:7-23) through its numpy encoder (pyANNonGPU/json_numpy.py:5-21):
an array becomes {"type": "ndarray", "dtype": <numpy dtype name>, "complex": bool} with
"real"/"imag" lists for complex arrays and "data" otherwise.  The reference ships no
such file, so this script spells the format out by hand (it does not use the package's
encoder) for seeded numbers; run it from the repository root to regenerate them.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def nd(x, dtype):
    x = np.asarray(x)
    if np.iscomplexobj(x):
        return {"type": "ndarray", "dtype": dtype, "complex": True,
                "real": np.round(x.real, 6).tolist(), "imag": np.round(x.imag, 6).tolist()}
    return {"type": "ndarray", "dtype": dtype, "complex": False, "data": x.tolist()}


def cg(rng, shape, s):
    return s * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def psideep():
    rng = np.random.default_rng(2024)
    N, M, C = 6, [6, 3], [3, 2]
    conns = [np.array([[(j + i) % N for j in range(M[0])] for i in range(C[0])], dtype=np.uint32),
             np.array([[(2 * j + i) % M[0] for j in range(M[1])] for i in range(C[1])], dtype=np.uint32)]
    obj = {"type": "PsiDeep", "num_sites": N,
           "a": nd(cg(rng, N, 0.1), "complex128"),
           "b": [nd(cg(rng, m, 0.1), "complex128") for m in M],
           "connections": [nd(c, "uint32") for c in conns],
           "W": [nd(cg(rng, (c, m), 0.2), "complex128") for c, m in zip(C, M)],
           "final_weights": nd(1 + cg(rng, M[-1], 0.2), "complex128"),
           "log_prefactor_re": 0.25, "log_prefactor_im": -0.125}
    return obj


def psicnn():
    rng = np.random.default_rng(2025)
    extent = [1, 3, 3]
    channels = [2, 1]
    connectivity = [[1, 3, 3], [1, 2, 2]]
    num_params = 1 * 2 * 9 + 2 * 1 * 4
    obj = {"type": "PsiCNN", "extent": extent,
           "num_channels_list": nd(np.array(channels), "uint32"),
           "connectivity_list": nd(np.array(connectivity), "uint32"),
           "symmetry_classes": nd(np.zeros(9, dtype=int), "uint32"),
           "params": nd(cg(rng, num_params, 0.15), "complex128"),
           "final_factor": 2.0, "log_prefactor_re": -0.5, "log_prefactor_im": 0.75}
    return obj


if __name__ == "__main__":
    for name, obj in (("psideep_reference_format.json", psideep()), ("psicnn_reference_format.json", psicnn())):
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(obj, f, indent=1)
        print("wrote", name)
