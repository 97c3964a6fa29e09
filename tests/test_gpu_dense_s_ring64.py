"""Dense-S workload at its full size: 1D Heisenberg ring N = 64 (192 Pauli strings),
one-layer PsiDeep with M = 256 (P = 16,704), 8,192 exchange chains in the S_z = 0
sector, fp32 E_loc and O_k, the dense S = <O* O> - <O*><O> built on fp64 matrix cores
and the fp64 blocked Cholesky solve of (S + 1e-3) x = F.

Checks: S is Hermitian with a real non-negative diagonal; the matrix-free S.v equals
the dense S applied to v (2e-5, the fp32 structured product against the fp64 Gram of
the fp32 O); the Cholesky solution's residual in the dense S is below 1e-8; and, on a
256-sample SampleList, S and F equal the oracle's (numpy Gram of the oracle's O_k
samples, fp32 device mode: 1e-4 relative to the norm; fp64: 1e-10).
"""
import numpy as np
import pytest

import pyoracle as orc
from helpers import cgauss, oracle_operator
import annongpu_amd as A

pytestmark = pytest.mark.gpu

N, M, SHIFT = 64, 256, 1e-3


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


@pytest.fixture(scope="module")
def ring():
    psi = A.seeded_rbm_deep(N, M, seed=2, sigma=0.05, final_weight=1.0)
    H = A.hamiltonians.heisenberg_chain(N)
    return psi, H, A.Operator(H)


def test_dense_S_and_cholesky_full_size(ring):
    psi, H, op = ring
    psi = psi.copy()
    psi.precision = "fp32"
    P = psi.num_params
    assert P == 16704
    ens = A.MonteCarloSpins(8192, 64, 20, 8192, move="exchange", seed=12)
    tdvp = A.TDVP(P)
    tdvp.eval(op, psi, ens)
    S = tdvp.S_matrix
    F = tdvp.F_vector
    d = np.diag(S)
    assert np.all(d.real >= -1e-12) and np.abs(d.imag).max() <= 1e-12 * np.abs(d.real).max()
    rng = np.random.default_rng(0)
    idx = rng.integers(0, P, (4096, 2))
    assert np.allclose(S[idx[:, 0], idx[:, 1]], S[idx[:, 1], idx[:, 0]].conj(), rtol=0, atol=1e-12 * np.abs(d).max())
    for _ in range(2):
        v = cgauss(rng, P, 1.0)
        assert rel(tdvp.S_dot_vector(v), S @ v) < 2e-5
    x = tdvp.solve(SHIFT, "cholesky")
    r = S @ x + SHIFT * x - F
    assert np.linalg.norm(r) / np.linalg.norm(F) < 1e-8
    assert tdvp.last_solve["residual"] < 1e-8


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_dense_S_matches_oracle_on_shared_samples(ring, precision):
    psi, H, op = ring
    psi = psi.copy()
    psi.precision = "fp32"
    mc = A.MonteCarloSpins(256, 8, 20, 256, move="exchange", seed=3)
    A.log_psi(psi, mc)
    confs = mc.last_samples()
    psi.precision = precision
    tdvp = A.TDVP(psi.num_params)
    tdvp.eval(op, psi, A.SampleList(N, confs))
    w = np.full(len(confs), 1.0 / len(confs))
    o = orc.deep(N, psi.a, psi.b, psi.connections, psi.W, psi.final_weights, 0.0)
    ref = orc.tdvp_accumulate(oracle_operator(H, 1), o, confs, w, with_S=False)
    Oc = ref["O_samples"] - ref["O"][None, :]
    S_ref = (Oc.conj().T * w) @ Oc
    tol = 1e-4 if precision == "fp32" else 1e-10
    assert rel(tdvp.F_vector, ref["F"]) < tol
    assert rel(tdvp.S_matrix, S_ref) < tol
    x = tdvp.solve(SHIFT, "cholesky")
    assert tdvp.last_solve["residual"] < 1e-8
    if precision == "fp64":
        # rank-256 S: (S + shift)^-1 F by the Woodbury identity, in fp64
        U = np.sqrt(w)[:, None] * Oc
        G = U @ U.conj().T + SHIFT * np.eye(len(w))
        x_ref = (ref["F"] - U.conj().T @ np.linalg.solve(G, U @ ref["F"])) / SHIFT
        assert rel(x, x_ref) < 1e-7
