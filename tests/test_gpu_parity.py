"""HIP path vs the fp64 oracle on the same inputs (needs an MI355X).

Tolerances: FP64 device mode 1e-10 relative; FP32 device mode (bf16x3 angle GEMM,
fp32 sweeps / E_loc / O_k, fp64 reductions) 1e-5 relative on log psi and
energies, 1e-4 relative-to-norm on vectors of O_k statistics.
"""
import math

import numpy as np
import pytest

import pyoracle as orc
from helpers import (random_deep, conf_array, oracle_operator, expr_dense, cgauss, DeepSpec)
import annongpu_amd as A
from annongpu_amd import hamiltonians as ham

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-10, "fp32": 1e-5}


def hip_deep(spec, precision="fp64"):
    psi = A.PsiDeep(*spec.args())
    psi.precision = precision
    return psi


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", [(16, 32, 16), (10, 20, 5), (64, 256, 64), (100, 400, 100), (130, 70, 130)])
def test_log_psi_s(precision, shape):
    N, M, C = shape
    spec = random_deep(N, [M], [C], seed=N + M, sigma=0.05, log_prefactor=0.3 - 0.2j)
    o = spec.oracle()
    psi = hip_deep(spec, precision)
    rng = np.random.default_rng(1)
    for _ in range(5):
        idx = int(rng.integers(0, 2 ** 62)) | (int(rng.integers(0, 2 ** 62)) << 62) | (int(rng.integers(0, 2 ** 20)) << 124)
        idx &= (1 << N) - 1
        ref = orc.log_psi_s(o, conf_array(idx))
        got = A.log_psi_s(psi, A.Spins(idx, N))
        assert abs(got - ref) <= TOL[precision] * max(1.0, abs(ref))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_log_psi_s_rbm(precision):
    rng = np.random.default_rng(2)
    N, M = 24, 48
    W = cgauss(rng, (N, M), 0.1)
    o = orc.rbm(W, 1.5 - 0.5j, 0.1)
    psi = A.PsiRBM(W, 1.5 - 0.5j, 0.1)
    psi.precision = precision
    for idx in rng.integers(0, 2 ** N, 6):
        ref = orc.log_psi_s(o, conf_array(int(idx)))
        assert abs(A.log_psi_s(psi, int(idx)) - ref) <= TOL[precision] * max(1, abs(ref))
        assert rel(A.psi_O_k(psi, int(idx)), orc.psi_O_k(o, conf_array(int(idx)))) < 10 * TOL[precision]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", [(12, 24, 12), (8, 16, 4), (70, 140, 70)])
def test_psi_O_k(precision, shape):
    N, M, C = shape
    spec = random_deep(N, [M], [C], seed=3, sigma=0.05)
    o = spec.oracle()
    psi = hip_deep(spec, precision)
    for idx in (0, 5, (1 << N) - 3):
        ref = orc.psi_O_k(o, conf_array(idx))
        got = A.psi_O_k(psi, A.Spins(idx, N))
        assert rel(got, ref) < 10 * TOL[precision]


def _tfim16_rbm(precision, seed=1):
    psi = A.seeded_rbm_deep(16, 32, seed)
    psi.precision = precision
    o = orc.deep(16, psi.a, psi.b, psi.connections, psi.W, psi.final_weights, 0.0)
    return psi, o


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_exact_summation_tfim16_rbm(precision):
    """TFIM ring N=16, complex RBM alpha=2: psi vector, norm, E, sigma, gradient."""
    psi, o = _tfim16_rbm(precision)
    ens = A.ExactSummationSpins(16)
    H = ham.transverse_ising_chain(16, 1.0, 1.0)
    op = A.Operator(H)
    lps_ref = orc.exact_log_psi_vector(o, 16)
    lps = A.log_psi_vector(psi, ens)
    assert np.max(np.abs(lps - lps_ref)) < TOL[precision] * 10
    norm_ref = math.sqrt(np.sum(np.exp(2 * lps_ref.real)))
    assert psi.norm(ens) == pytest.approx(norm_ref, rel=TOL[precision] * 10)
    psi.normalize(ens)
    o.s.log_prefactor = orc._cplx(psi.log_prefactor)
    oop = oracle_operator(H, 1)
    w = np.exp(2 * orc.exact_log_psi_vector(o, 16).real)
    ref = orc.tdvp_accumulate(oop, o, orc.exact_confs(16, 1), w, with_S=False)
    E = A.ExpectationValue()(op, psi, ens)
    assert abs(E - ref["E"]) < TOL[precision] * abs(ref["E"])
    sigma, E2 = A.ExpectationValue().fluctuation(op, psi, ens)
    sigma_ref = math.sqrt(ref["E2"] - abs(ref["E"]) ** 2)
    assert sigma == pytest.approx(sigma_ref, rel=100 * TOL[precision], abs=1e-6)
    g, _ = A.ExpectationValue().gradient(op, psi, ens)
    Om = ref["O"]
    g_ref = ref["F"]   # gradient = sum w conj(O) E - E sum w conj(O) = F (unnormalised weights cancel)
    assert rel(g, g_ref) < 100 * TOL[precision]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_apply_operator(precision):
    N = 10
    spec = random_deep(N, [20], [N], seed=5, sigma=0.1)
    o = spec.oracle()
    psi = hip_deep(spec, precision)
    H = ham.heisenberg_chain(N) + 0.3 * A.sigma_y(2) * A.sigma_x(7) + 0.2 * A.sigma_x(4)
    ens = A.ExactSummationSpins(N)
    got = A.apply_operator(psi, A.Operator(H), ens)
    vec = np.exp(orc.exact_log_psi_vector(o, N))
    assert rel(got, expr_dense(H, N) @ vec) < 10 * TOL[precision]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_tdvp_exact(precision):
    """S, F, <O>, E, var_H of TDVP.eval on ExactSummation vs the oracle."""
    N = 8
    spec = random_deep(N, [16], [N], seed=6, sigma=0.1)
    o = spec.oracle()
    psi = hip_deep(spec, precision)
    ens = A.ExactSummationSpins(N)
    psi.normalize(ens)
    o.s.log_prefactor = orc._cplx(psi.log_prefactor)
    H = ham.transverse_ising_chain(N, 1.0, 0.8)
    tdvp = A.TDVP(psi.num_params)
    tdvp.eval(A.Operator(H), psi, ens)
    w = np.exp(2 * orc.exact_log_psi_vector(o, N).real)
    ref = orc.tdvp_accumulate(oracle_operator(H, 1), o, orc.exact_confs(N, 1), w)
    tol = 100 * TOL[precision]
    assert abs(tdvp.E_local - ref["E"]) < tol * abs(ref["E"])
    assert rel(tdvp.O_k_vector, ref["O"]) < tol
    assert rel(tdvp.F_vector, ref["F"]) < tol
    assert rel(tdvp.S_matrix, ref["S"]) < tol
    assert tdvp.var_H == pytest.approx(ref["E2"] - abs(ref["E"]) ** 2, rel=tol, abs=1e-8)
    v = cgauss(np.random.default_rng(0), psi.num_params, 1.0)
    assert rel(tdvp.S_dot_vector(v), ref["S"] @ v) < tol
    shift = 1e-3
    x_ref = np.linalg.solve(ref["S"] + shift * np.eye(psi.num_params), ref["F"])
    x_ch = tdvp.solve(shift, "cholesky")
    assert rel(x_ch, x_ref) < 1e3 * tol
    x_cg = tdvp.solve(shift, "cg", rtol=1e-10, max_iter=2000)
    assert rel(x_cg, x_ref) < 1e3 * tol
    cold = tdvp.last_solve["iterations"]
    # warm starts: from the solution (at most a few iterations) and from a perturbed solution
    x_w = tdvp.solve(shift, "cg", rtol=1e-8, max_iter=2000, x0=x_cg)
    assert tdvp.last_solve["iterations"] <= (1 if precision == "fp64" else cold // 2)
    assert rel(x_w, x_ref) < 1e3 * tol
    x_w = tdvp.solve(shift, "cg", rtol=1e-10, max_iter=2000, x0=x_cg + 0.1 * np.abs(x_cg).max())
    assert rel(x_w, x_ref) < 1e3 * tol


def test_tdvp_unnormalised_weights():
    """reference S = sum w conj(O) O - conj(<O>) <O> with raw exact weights"""
    N = 6
    spec = random_deep(N, [12], [N], seed=7, sigma=0.1, log_prefactor=0.4)
    o = spec.oracle()
    psi = hip_deep(spec)
    ens = A.ExactSummationSpins(N)
    H = ham.transverse_ising_chain(N)
    tdvp = A.TDVP(psi.num_params)
    tdvp.eval(A.Operator(H), psi, ens)
    w = np.exp(2 * orc.exact_log_psi_vector(o, N).real)
    ref = orc.tdvp_accumulate(oracle_operator(H, 1), o, orc.exact_confs(N, 1), w)
    assert rel(tdvp.S_matrix, ref["S"]) < 1e-9
    assert rel(tdvp.F_vector, ref["F"]) < 1e-9


@pytest.mark.parametrize("move", ["flip", "exchange"])
def test_mc_trajectories_match_oracle(move):
    """fp64 device sampler replays the oracle's Philox stream: identical samples."""
    N = 20
    spec = random_deep(N, [40], [N], seed=8, sigma=0.2)
    o = spec.oracle()
    psi = hip_deep(spec, "fp64")
    chains, steps, sweeps, therm = 64, 3, 2, 5
    ens = A.MonteCarloSpins(chains * steps, sweeps, therm, chains, move=move, seed=99)
    A.log_psi(psi, ens)    # one sampling call
    got = ens.last_samples()[:, 0]
    ref = orc.mc_run(o, chains, steps, sweeps, therm, move=orc.MOVE_EXCHANGE if move == "exchange" else orc.MOVE_FLIP,
                     seed=99)
    assert np.array_equal(got, ref["samples"][:, 0])
    assert ens.acceptance_rate == pytest.approx(ref["acceptances"] / (ref["acceptances"] + ref["rejections"]))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_mc_energy_tfim16_rbm(precision):
    """MCMC energy within 3 standard errors of exact summation (TFIM N=16, 1024 chains x 200 sweeps)."""
    psi, o = _tfim16_rbm(precision, seed=11)
    H = ham.transverse_ising_chain(16, 1.0, 1.0)
    op = A.Operator(H)
    ex = A.ExactSummationSpins(16)
    psi.normalize(ex)
    E_exact = A.ExpectationValue()(op, psi, ex).real
    ens = A.MonteCarloSpins(1024 * 200, 1, 20, 1024, seed=3)
    el = A.local_energies(op, psi, ens).real.reshape(200, 1024)
    se = el.mean(axis=0).std(ddof=1) / math.sqrt(1024)
    assert abs(el.mean() - E_exact) < 3 * se + 1e-12


def test_exchange_conserves_sz():
    N = 64
    psi = A.seeded_rbm_deep(N, 256, 2)
    psi.precision = "fp32"
    ens = A.MonteCarloSpins(512, 4, 2, 512, move="exchange", seed=4)
    A.log_psi(psi, ens)
    s = ens.last_samples()
    ups = np.array([bin(int(x)).count("1") for x in s[:, 0]])
    assert np.all(ups == N // 2)
    assert 0.0 < ens.acceptance_rate <= 1.0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_structured_matvec_matches_dense_gram(precision):
    """matrix-free S.v (two spin GEMMs) equals the dense Gram S = O^H W O applied to v,
    on MC samples at a size the oracle cannot cover (P = 8384)."""
    N, M = 64, 128
    psi = A.seeded_rbm_deep(N, M, 9, sigma=0.05)
    psi.precision = precision
    ens = A.MonteCarloSpins(2048, 2, 5, 512, move="exchange", seed=5)
    tdvp = A.TDVP(psi.num_params)
    tdvp.eval(A.Operator(A.hamiltonians.heisenberg_chain(N)), psi, ens)
    S = tdvp.S_matrix
    assert np.allclose(S, S.conj().T, atol=1e-12 * np.abs(S).max())
    rng = np.random.default_rng(1)
    for _ in range(2):
        v = cgauss(rng, psi.num_params, 1.0)
        assert rel(tdvp.S_dot_vector(v), S @ v) < (1e-10 if precision == "fp64" else 2e-5)
    # Jacobi preconditioner = diagonal of S: (S e_k)_k for a few k
    for k in (0, 5, 64, 64 + 7, 64 + 128 + 300, psi.num_params - 1):
        e = np.zeros(psi.num_params, dtype=complex)
        e[k] = 1
        assert tdvp.S_dot_vector(e)[k].real == pytest.approx(S[k, k].real, rel=1e-4, abs=1e-9)
    x = tdvp.solve(1e-3, "cg", rtol=1e-5, max_iter=3000)
    assert rel((S + 1e-3 * np.eye(psi.num_params)) @ x, tdvp.F_vector) < 2e-5
    assert tdvp.last_solve["iterations"] < 3000


def shuffled_deep(N, M, C, seed, repeated):
    """one-layer PsiDeep whose connections are not the circulant 1d layout: per unit a
    random choice of C sites, distinct (repeated=False) or with repeats (several
    connections of a unit on one site, repeated=True)"""
    rng = np.random.default_rng(seed)
    conn = np.stack([rng.choice(N, C, replace=repeated) for _ in range(M)], axis=1).astype(np.uint32)
    a, b = cgauss(rng, N, 0.05), cgauss(rng, M, 0.05)
    W = cgauss(rng, (C, M), 0.05)
    return A.PsiDeep(N, a, [b], [conn], [W], np.ones(M), 0.0)


@pytest.mark.parametrize("repeated", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_structured_matvec_general_connections(precision, repeated):
    """S.v through the connection tables (a PsiDeep whose connections are not the
    circulant layout) equals the dense Gram matrix applied to v"""
    N, M, C = 32, 64, 12
    psi = shuffled_deep(N, M, C, 17, repeated)
    psi.precision = precision
    ens = A.MonteCarloSpins(1024, 2, 5, 256, move="exchange", seed=8)
    tdvp = A.TDVP(psi.num_params)
    tdvp.eval(A.Operator(A.hamiltonians.heisenberg_chain(N)), psi, ens)
    S = tdvp.S_matrix
    rng = np.random.default_rng(2)
    for _ in range(2):
        v = cgauss(rng, psi.num_params, 1.0)
        assert rel(tdvp.S_dot_vector(v), S @ v) < (1e-10 if precision == "fp64" else 2e-5)
    x = tdvp.solve(1e-3, "cg", rtol=1e-6, max_iter=3000)
    assert rel((S + 1e-3 * np.eye(psi.num_params)) @ x, tdvp.F_vector) < 1e-5


@pytest.mark.parametrize("kind", ["rbm", "deep", "deep_repeated", "deep_circulant"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_params_update_matches_fresh_upload(precision, kind):
    """New parameters of a net already on the device (staged and repacked there by a
    kernel) give bit for bit the log psi, O_k, energy and F of a net first uploaded
    with them (host packing)"""
    N, M, C = 20, 40, 8

    def make():
        if kind == "rbm":
            return A.PsiRBM(cgauss(np.random.default_rng(1), (N, M), 0.1), 1.0 - 0.3j, 0.1)
        if kind == "deep_circulant":
            return A.seeded_rbm_deep(N, M, 1, sigma=0.1)
        return shuffled_deep(N, M, C, 1, kind == "deep_repeated")

    a = make()
    a.precision = precision
    idx = [int(i) for i in np.random.default_rng(3).integers(0, 2 ** N, 4)]
    A.log_psi_s(a, idx[0])                      # first, full upload
    rng = np.random.default_rng(4)
    for _ in range(2):
        newp = a.params + cgauss(rng, a.num_params, 0.02)
        a.params = newp                          # staged update
        b = make()
        b.params = newp                          # never uploaded: full upload
        b.precision = precision
        for i in idx:
            assert A.log_psi_s(a, i) == A.log_psi_s(b, i)
            assert np.array_equal(A.psi_O_k(a, i), A.psi_O_k(b, i))
        op = A.Operator(A.hamiltonians.heisenberg_chain(N))
        res = []
        for psi in (a, b):
            ens = A.MonteCarloSpins(512, 2, 10, 128, move="exchange", seed=9)
            tdvp = A.TDVP(psi.num_params)
            tdvp.eval_F(op, psi, ens)
            res.append((tdvp.E_local, tdvp.F_vector))
        assert res[0][0] == res[1][0]
        assert np.array_equal(res[0][1], res[1][1])
