"""Kronecker-factored preconditioner of the CG solve (needs an MI355X).

The preconditioner changes how CG reaches the solution, not the solution: every
test compares the K-FAC-preconditioned solve with a dense solve of the same
(S + shift) x = F (numpy on the oracle's S, or the library's dense S), and checks
the true residual through the matrix-free S.v.
"""
import numpy as np
import pytest

import pyoracle as orc
from helpers import random_deep, oracle_operator, cgauss
import annongpu_amd as A
from annongpu_amd import hamiltonians as ham

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-10, "fp32": 1e-5}


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("shape", [(8, 16), (12, 6)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_kfac_cg_matches_dense_solve_deep(precision, shape):
    """PsiDeep (biases by Jacobi, W by the Kronecker factors) on ExactSummation: the
    solution equals numpy's solve of the oracle's S + shift.  Connection shifts 1
    (8 sites, 16 units) and 2 (12 sites, 6 units); with C = N the fp32 solve runs in
    site-major order."""
    N, M = shape
    spec = random_deep(N, [M], [N], seed=6, sigma=0.1)
    o = spec.oracle()
    psi = A.PsiDeep(*spec.args())
    psi.precision = precision
    ens = A.ExactSummationSpins(N)
    psi.normalize(ens)
    o.s.log_prefactor = orc._cplx(psi.log_prefactor)
    H = ham.transverse_ising_chain(N, 1.0, 0.8)
    tdvp = A.TDVP(psi.num_params)
    tdvp.eval_F(A.Operator(H), psi, ens)
    w = np.exp(2 * orc.exact_log_psi_vector(o, N).real)
    ref = orc.tdvp_accumulate(oracle_operator(H, 1), o, orc.exact_confs(N, 1), w)
    shift = 1e-3
    x_ref = np.linalg.solve(ref["S"] + shift * np.eye(psi.num_params), ref["F"])
    tol = 1e5 * TOL[precision]
    x = tdvp.solve(shift, "cg_kfac", rtol=1e-10, max_iter=2000)
    assert rel(x, x_ref) < tol
    assert tdvp.last_solve["residual"] <= 1e-10
    cold = tdvp.last_solve["iterations"]
    # warm start from the solution: converged at the first check in fp64; in fp32 the
    # restarted residual carries the fp32 error of S.v and needs some iterations
    tdvp.solve(shift, "cg_kfac", rtol=1e-8, max_iter=2000, x0=x)
    assert tdvp.last_solve["iterations"] <= (1 if precision == "fp64" else cold // 2)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_kfac_cg_matches_dense_solve_rbm(precision):
    """PsiRBM (every parameter is a W parameter) on MC samples: the solution equals
    the dense solve with the library's S built from the same samples."""
    N, M = 12, 24
    rng = np.random.default_rng(4)
    W = 0.2 * (rng.standard_normal((N, M)) + 1j * rng.standard_normal((N, M)))
    psi = A.PsiRBM(W, 1.0)
    psi.precision = precision
    ens = A.MonteCarloSpins(4096, 2, 10, 1024, seed=3)
    tdvp = A.TDVP(psi.num_params)
    tdvp.eval(A.Operator(ham.heisenberg_chain(N)), psi, ens)
    S, F = tdvp.S_matrix, tdvp.F_vector
    shift = 1e-2
    x_ref = np.linalg.solve(S + shift * np.eye(psi.num_params), F)
    # fp32 S.v is accurate to ~1e-7: its solves are asked for 1e-6
    x = tdvp.solve(shift, "cg_kfac", rtol=1e-10 if precision == "fp64" else 1e-6, max_iter=3000)
    assert rel(x, x_ref) < (1e-7 if precision == "fp64" else 1e-3)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_kfac_cg_converges_faster_than_jacobi(precision):
    """Heisenberg 8x8 (N = 64), complex RBM with biases alpha = 4 (P = 16,704) on 8,192
    MC samples after 15 SR steps: K-FAC reaches rtol 1e-2 (true residual checked
    through S.v) in at most half of Jacobi's iterations.  (At this tolerance the two
    solutions of the ill-conditioned system differ in its small-eigenvalue directions,
    so x itself is not compared.)"""
    L = 8
    N, M = L * L, 4 * L * L
    psi = A.seeded_rbm_deep(N, M, 3, sigma=0.01)
    psi.precision = precision
    op = A.Operator(ham.heisenberg_square(L))
    ens = A.MonteCarloSpins(8192, 4, 20, 8192, move="exchange", seed=5, persistent=True)
    tdvp = A.TDVP(psi.num_params)
    for _ in range(15):
        tdvp.eval_F(op, psi, ens)
        x = tdvp.solve(1e-2, "cg_kfac", rtol=1e-3, max_iter=300, raise_on_failure=False)
        d = 0.05 * x
        psi.params = psi.params - d * min(1.0, 0.02 / np.abs(d).max())
    tdvp.eval_F(op, psi, ens)
    shift = 1e-3
    F = tdvp.F_vector
    x_j = tdvp.solve(shift, "cg", rtol=1e-2, max_iter=1500, raise_on_failure=False)
    it_j, res_j = tdvp.last_solve["iterations"], tdvp.last_solve["residual"]
    x_k = tdvp.solve(shift, "cg_kfac", rtol=1e-2, max_iter=1500)
    it_k = tdvp.last_solve["iterations"]
    r = tdvp.S_dot_vector(x_k) + shift * x_k - F
    assert np.linalg.norm(r) <= 1.05e-2 * np.linalg.norm(F)
    assert it_k <= it_j // 2, (it_k, it_j)
    if res_j <= 1e-2:
        r = tdvp.S_dot_vector(x_j) + shift * x_j - F
        assert np.linalg.norm(r) <= 1.05e-2 * np.linalg.norm(F)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_kfac_cg_general_connections(precision):
    """PsiDeep with shuffled (non-circulant, distinct) connections: the K-FAC step reads
    the index table; the solution equals the dense solve with the library's S"""
    from test_gpu_parity import shuffled_deep
    psi = shuffled_deep(24, 48, 10, 21, repeated=False)
    psi.precision = precision
    ens = A.MonteCarloSpins(2048, 2, 10, 512, seed=4)
    tdvp = A.TDVP(psi.num_params)
    tdvp.eval(A.Operator(ham.heisenberg_chain(24)), psi, ens)
    S, F = tdvp.S_matrix, tdvp.F_vector
    shift = 1e-2
    x_ref = np.linalg.solve(S + shift * np.eye(psi.num_params), F)
    x = tdvp.solve(shift, "cg_kfac", rtol=1e-10 if precision == "fp64" else 1e-6, max_iter=3000)
    assert rel(x, x_ref) < (1e-7 if precision == "fp64" else 1e-3)


def test_kfac_rejects_generic_nets():
    psi = A.seeded_cnn([8], [(2, [3]), (1, [3])], 1, sigma=0.1)
    ens = A.ExactSummationSpins(8)
    tdvp = A.TDVP(psi.num_params)
    tdvp.eval_F(A.Operator(ham.transverse_ising_chain(8, 1.0, 0.5)), psi, ens)
    with pytest.raises(ValueError):
        tdvp.solve(1e-3, "cg_kfac", rtol=1e-6)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_kfac_moments_precomputed_during_eval(precision):
    """After a K-FAC solve the next evaluation of the same TDVP computes the factors and
    their inverses for that solve's shift on a second stream while the local energies
    run; its statistics and solves equal bit for bit those of a TDVP that computes the
    factors in the solve, for the same shift and for another one."""
    N, M = 16, 32
    psi = A.seeded_rbm_deep(N, M, 3, sigma=0.05)
    psi.precision = precision
    op = A.Operator(ham.heisenberg_chain(N))
    t1 = A.TDVP(psi.num_params)
    t1.eval_F(op, psi, A.MonteCarloSpins(2048, 2, 10, 512, move="exchange", seed=4))
    t1.solve(1e-3, "cg_kfac", rtol=1e-6, max_iter=500)
    t2 = A.TDVP(psi.num_params)
    t1.eval_F(op, psi, A.MonteCarloSpins(2048, 2, 10, 512, move="exchange", seed=9))
    t2.eval_F(op, psi, A.MonteCarloSpins(2048, 2, 10, 512, move="exchange", seed=9))
    assert t1.E_local == t2.E_local
    assert np.array_equal(t1.F_vector, t2.F_vector)
    x1 = t1.solve(1e-3, "cg_kfac", rtol=1e-6, max_iter=500)
    x2 = t2.solve(1e-3, "cg_kfac", rtol=1e-6, max_iter=500)
    assert t1.last_solve["iterations"] == t2.last_solve["iterations"]
    assert np.array_equal(x1, x2)
    # a second solve with the same shift reuses the inverses; another shift recomputes
    # the factors from the samples
    x3 = t1.solve(1e-3, "cg_kfac", rtol=1e-6, max_iter=500)
    assert np.array_equal(x1, x3)
    x4 = t1.solve(3e-3, "cg_kfac", rtol=1e-6, max_iter=500)
    x5 = t2.solve(3e-3, "cg_kfac", rtol=1e-6, max_iter=500)
    assert np.array_equal(x4, x5)
