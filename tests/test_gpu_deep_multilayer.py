"""PsiDeep with several hidden layers on the HIP path (PsiDeep.hpp:171-266 forward,
345-445 O_k back-propagation) against the fp64 oracle, which restates the same
forward and back-propagation: the reference's own fixture (3 sites, layers
[18, 9, 3], connectivities [3, 2, 3], test/conftest.py:25-34), a two-layer net at
N = 64, exact TDVP statistics, E_loc through apply_operator against the dense
Hamiltonian, and a Monte Carlo energy within 4 standard errors of exact summation.

Tolerances: fp64 device mode 1e-10 relative, fp32 1e-5 (log psi, energies) and
1e-4 relative to the norm on P-vectors.
"""
import math

import numpy as np
import pytest

import pyoracle as orc
from helpers import random_deep, conf_array, oracle_operator, expr_dense, cgauss
import annongpu_amd as A
from annongpu_amd import hamiltonians as ham

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-10, "fp32": 1e-5}


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_reference_fixture_log_psi_and_O_k(precision):
    psi = A.new_deep_neural_network(3, 3, [18, 9, 3], [3, 2, 3], a=0, noise=1e-3, rng=np.random.default_rng(4))
    psi.precision = precision
    o = orc.deep(3, psi.a, psi.b, psi.connections, psi.W, psi.final_weights, 0.0)
    ex = A.ExactSummationSpins(3)
    ref = orc.exact_log_psi_vector(o, 3)
    assert np.max(np.abs(A.log_psi_vector(psi, ex) - ref)) < TOL[precision] * max(1.0, np.abs(ref).max())
    for idx in range(8):
        got = A.psi_O_k(psi, A.Spins(idx, 3))
        assert rel(got, orc.psi_O_k(o, conf_array(idx))) < 10 * TOL[precision]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_two_layer_net_at_64_sites(precision):
    spec = random_deep(64, [128, 64], [64, 32], seed=12, sigma=0.05, log_prefactor=0.1 - 0.2j)
    psi = A.PsiDeep(*spec.args())
    psi.precision = precision
    o = spec.oracle()
    rng = np.random.default_rng(3)
    for _ in range(4):
        idx = int(rng.integers(0, 2 ** 63)) | (int(rng.integers(0, 2)) << 63)
        ref = orc.log_psi_s(o, conf_array(idx))
        assert abs(A.log_psi_s(psi, A.Spins(idx, 64)) - ref) <= TOL[precision] * max(1.0, abs(ref))
        assert rel(A.psi_O_k(psi, A.Spins(idx, 64)), orc.psi_O_k(o, conf_array(idx))) < 10 * TOL[precision]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_tdvp_exact_multilayer(precision):
    N = 8
    spec = random_deep(N, [16, 8], [N, 8], seed=6, sigma=0.1)
    psi = A.PsiDeep(*spec.args())
    psi.precision = precision
    o = spec.oracle()
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
    v = cgauss(np.random.default_rng(0), psi.num_params, 1.0)
    assert rel(tdvp.S_dot_vector(v), ref["S"] @ v) < tol
    x = tdvp.solve(1e-3, "cholesky")
    x_ref = np.linalg.solve(ref["S"] + 1e-3 * np.eye(psi.num_params), ref["F"])
    assert rel(x, x_ref) < 1e3 * tol
    x = tdvp.solve(1e-3, "cg", rtol=1e-10, max_iter=3000)
    assert rel(x, x_ref) < 1e3 * tol
    # the Kronecker-factored preconditioner needs a one-layer net
    with pytest.raises(Exception):
        tdvp.solve(1e-3, "cg_kfac")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_apply_operator_multilayer(precision):
    N = 9
    spec = random_deep(N, [18, 9], [N, 6], seed=8, sigma=0.1)
    psi = A.PsiDeep(*spec.args())
    psi.precision = precision
    H = ham.heisenberg_chain(N) + 0.3 * A.sigma_y(2) * A.sigma_x(7) + 0.2 * A.sigma_x(4)
    ens = A.ExactSummationSpins(N)
    got = A.apply_operator(psi, A.Operator(H), ens)
    vec = np.exp(orc.exact_log_psi_vector(spec.oracle(), N))
    assert rel(got, expr_dense(H, N) @ vec) < 10 * TOL[precision]


@pytest.mark.parametrize("move", ["flip", "exchange"])
def test_mc_energy_multilayer(move):
    N = 12
    spec = random_deep(N, [24, 12], [N, 12], seed=9, sigma=0.1)
    psi = A.PsiDeep(*spec.args())
    psi.precision = "fp32"
    H = ham.heisenberg_chain(N)
    op = A.Operator(H)
    o = spec.oracle()
    if move == "flip":
        ex = A.ExactSummationSpins(N)
        psi.normalize(ex)
        E_exact = A.ExpectationValue()(op, psi, ex).real
    else:   # exact sum over the S_z = 0 sector
        from itertools import combinations
        confs = np.array([sum(1 << i for i in c) for c in combinations(range(N), N // 2)], dtype=np.uint64)
        lp = np.array([orc.log_psi_s(o, orc.conf_words(int(x), 4)) for x in confs])
        wt = np.exp(2 * (lp.real - lp.real.max()))
        wt /= wt.sum()
        E_exact = orc.tdvp_accumulate(oracle_operator(H, 1), o, confs[:, None], wt, with_S=False)["E"].real
    chains, steps = 1024, 100
    ens = A.MonteCarloSpins(chains * steps, 1, 20, chains, move=move, seed=3)
    el = A.local_energies(op, psi, ens).real.reshape(steps, chains)
    se = el.mean(axis=0).std(ddof=1) / math.sqrt(chains)
    assert abs(el.mean() - E_exact) < 4 * se + 1e-12, (el.mean(), E_exact, se)
    if move == "exchange":
        s = ens.last_samples()[:, 0]
        assert np.all([bin(int(x)).count("1") == N // 2 for x in s[:200]])
