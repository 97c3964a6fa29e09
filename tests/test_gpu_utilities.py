"""Free functions of the Python API on the HIP path against the fp64 oracle:
psi_O_k_vector (PsiOkVector.cu.template:8-37: unweighted sum of O_k over the
ensemble), psi_vector and the value of log_psi (PsiVector.cu.template:44-99:
exp(log psi) per configuration; sum_s w_s log psi(s) with the ensemble's weights,
the unnormalised |psi|^2 for ExactSummation), calibrate (pyANNonGPU/PsiRBM.py:43-49)
the operator-list form of ExpectationValue.__call__ (main.cpp.template:424), exp_sigma_z
(:428) against the dense operator's row sums, the reweighted form (:431) against <H> of the
wave function it reweights to, and gradient_with_noise (:427).
"""
import cmath
import math

import numpy as np
import pytest

import pyoracle as orc
from helpers import random_deep, oracle_operator
import annongpu_amd as A
from annongpu_amd import hamiltonians as ham

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-10, "fp32": 1e-5}


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def _net(precision, N=10, seed=3, log_prefactor=0.2 - 0.4j):
    spec = random_deep(N, [20], [N], seed=seed, sigma=0.1, log_prefactor=log_prefactor)
    psi = A.PsiDeep(*spec.args())
    psi.precision = precision
    return spec, psi


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_psi_vector_and_log_psi_exact(precision):
    N = 10
    spec, psi = _net(precision, N)
    o = spec.oracle()
    ex = A.ExactSummationSpins(N)
    lp = orc.exact_log_psi_vector(o, N)
    assert rel(A.psi_vector(psi, ex), np.exp(lp)) < 10 * TOL[precision]
    w = np.exp(2 * lp.real)
    ref = complex(np.sum(w * lp))
    assert abs(A.log_psi(psi, ex) - ref) < 10 * TOL[precision] * abs(ref)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_psi_O_k_vector(precision):
    N = 8
    spec, psi = _net(precision, N, seed=5)
    o = spec.oracle()
    ex = A.ExactSummationSpins(N)
    confs = orc.exact_confs(N, 1)
    ref = orc.tdvp_accumulate(None, o, confs, np.ones(len(confs)), with_S=False)["O_samples"].sum(axis=0)
    assert rel(A.psi_O_k_vector(psi, ex), ref) < 10 * TOL[precision]
    # Monte Carlo: the unweighted sum over the call's samples
    mc = A.MonteCarloSpins(512, 2, 5, 128, seed=4)
    got = A.psi_O_k_vector(psi, mc)
    s = mc.last_samples()
    ref = orc.tdvp_accumulate(None, o, s, np.ones(len(s)), with_S=False)["O_samples"].sum(axis=0)
    assert rel(got, ref) < 10 * TOL[precision]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_psi_vector_and_log_psi_on_samples(precision):
    N = 70
    spec = random_deep(N, [140], [N], seed=9, sigma=0.05, log_prefactor=0.1j)
    psi = A.PsiDeep(*spec.args())
    psi.precision = precision
    o = spec.oracle()
    rng = np.random.default_rng(2)
    confs = rng.integers(0, 2 ** 63, (300, 2), dtype=np.int64).astype(np.uint64)
    confs[:, 1] &= np.uint64((1 << (N - 64)) - 1)
    w = rng.random(300)
    ens = A.SampleList(N, confs, weights=w)
    lp = np.array([orc.log_psi_s(o, np.concatenate([c, np.zeros(2, np.uint64)])) for c in confs])
    assert np.max(np.abs(A.log_psi_vector(psi, ens) - lp)) < TOL[precision] * np.abs(lp).max()
    assert rel(A.psi_vector(psi, ens), np.exp(lp)) < 10 * TOL[precision]
    ref = complex(np.sum(w * lp))
    assert abs(A.log_psi(psi, ens) - ref) < 10 * TOL[precision] * abs(ref)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_calibrate(precision):
    N = 10
    spec, psi = _net(precision, N, seed=7)
    ex = A.ExactSummationSpins(N)
    # exact: normalised, and the |psi|^2-weighted mean phase of log psi removed
    lp0 = psi.log_prefactor
    lp = orc.exact_log_psi_vector(spec.oracle(), N) - spec.log_prefactor
    norm = math.sqrt(np.sum(np.exp(2 * lp.real)))
    lp_n = lp - math.log(norm)
    phase = float(np.sum(np.exp(2 * lp_n.real) * lp_n.imag))
    psi.calibrate(ex)
    expect = -math.log(norm) - 1j * phase
    assert abs(psi.log_prefactor - expect) < 10 * TOL[precision] * max(1.0, abs(expect)), (psi.log_prefactor, lp0)
    assert psi.norm(ex) == pytest.approx(1.0, rel=10 * TOL[precision])
    assert abs(A.log_psi(psi, ex).imag) < 10 * TOL[precision]
    # sampled: log_prefactor = -<log psi> over the samples (with log_prefactor 0)
    confs = orc.exact_confs(N, 1)[::7]
    sl = A.SampleList(N, confs)
    psi.calibrate(sl)
    ref = -np.mean(lp[::7])
    assert abs(psi.log_prefactor - ref) < 10 * TOL[precision] * max(1.0, abs(ref))
    assert abs(A.log_psi(psi, sl)) < 10 * TOL[precision] * max(1.0, abs(ref))


def test_expectation_value_of_operator_list():
    N = 8
    spec, psi = _net("fp64", N, seed=11)
    ex = A.ExactSummationSpins(N)
    psi.normalize(ex)
    ops = [ham.transverse_ising_chain(N, 1.0, 0.5), A.sigma_z(0) * A.sigma_z(3), A.sigma_x(2) + 0.5 * A.sigma_y(5)]
    got = A.ExpectationValue()([A.Operator(h) for h in ops], psi, ex)
    o = spec.oracle()
    o.s.log_prefactor = orc._cplx(psi.log_prefactor)
    w = np.exp(2 * orc.exact_log_psi_vector(o, N).real)
    for g, h in zip(got, ops):
        ref = orc.tdvp_accumulate(oracle_operator(h, 1), o, orc.exact_confs(N, 1), w, with_S=False)["E"]
        assert abs(g - ref) < 1e-10 * max(1.0, abs(ref))
    assert cmath.isfinite(got[0])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_exp_sigma_z(precision):
    """ExpectationValue.exp_sigma_z (ExpectationValue.cu.template:52-77): sum_s w_s
    exp(L(s)) with L(s) the row sum of the operator's dense matrix (the matrix-element sum
    of its strings at s, Operator.hpp:123-135) and w the normalised |psi|^2"""
    import classical_oracle as co
    from helpers import expr_dense
    N = 6
    spec, psi = _net(precision, N, seed=21)
    ex = A.ExactSummationSpins(N)
    psi.normalize(ex)
    H = A.sigma_z(0) * A.sigma_z(1) + 0.3 * A.sigma_x(2) - 0.2j * A.sigma_y(3) * A.sigma_x(4) + 0.1 * A.sigma_y(5)
    got = A.ExpectationValue().exp_sigma_z(A.Operator(H), psi, ex)
    o = spec.oracle()
    o.s.log_prefactor = orc._cplx(psi.log_prefactor)
    w = np.exp(2 * orc.exact_log_psi_vector(o, N).real)
    L = expr_dense(H, N).sum(axis=1)
    assert all(co.fast_local_energy(list(H), s, "spins") == pytest.approx(L[s]) for s in range(2 ** N))
    ref = w @ np.exp(L)
    assert abs(got - ref) < 10 * TOL[precision] * abs(ref)


def test_exp_sigma_z_on_pauli_strings():
    import classical_oracle as co
    N = 3
    H = A.sigma_z(0) * A.sigma_x(1) + 0.4 * A.sigma_y(2) + 0.25 * A.sigma_x(0) * A.sigma_y(1)
    psi = A.new_classical_network(N, 1, A.sigma_x(0) * A.sigma_z(1) + A.sigma_y(2))
    psi.params = np.array([0.3 - 0.1j, -0.2 + 0.05j])
    ex = A.ExactSummationPaulis(N)
    psi.normalize(ex)
    got = A.ExpectationValue().exp_sigma_z(A.Operator(H), psi, ex)
    strings = [co.enumerate_pauli(i) for i in range(4 ** N)]
    lps = A.log_psi_vector(psi, ex)
    w = np.exp(2 * lps.real)
    ref = sum(wi * np.exp(co.fast_local_energy(list(H), x, "paulis")) for wi, x in zip(w, strings))
    assert abs(got - ref) < 1e-10 * abs(ref)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_expectation_value_reweighted(precision):
    """ExpectationValue(op, psi, psi_sampling, ensemble) (ExpectationValue.cu.template:130-172)
    on an exact summation over psi_sampling: the |psi / psi_sampling|^2 reweighting turns the
    samples of psi_sampling into those of psi, so the result is <H> in psi exactly"""
    from helpers import expr_dense
    N = 6
    _, psi_s = _net(precision, N, seed=31)
    H_cl = A.sigma_z(0) * A.sigma_x(1) + 1.2 * A.sigma_y(2) + 0.7 * A.sigma_x(3) * A.sigma_x(4)
    psi = A.new_classical_network(N, 1, H_cl)
    psi.params = np.array([0.2 + 0.1j, -0.3j, 0.15])
    psi.precision = precision
    ex = A.ExactSummationSpins(N)
    H = ham.transverse_ising_chain(N, 1.0, 0.7)
    got = A.ExpectationValue()(A.Operator(H), psi, psi_s, ex)
    v = A.psi_vector(psi, ex)
    ref = (v.conj() @ expr_dense(H, N) @ v) / (v.conj() @ v)
    assert abs(got - ref) < 10 * TOL[precision] * abs(ref)
    # with psi_sampling = psi the reweighting is the identity
    same = A.ExpectationValue()(A.Operator(H), psi, psi, ex)
    psi.normalize(ex)
    assert abs(same - A.ExpectationValue()(A.Operator(H), psi, ex)) < 10 * TOL[precision] * abs(ref)


def test_gradient_with_noise_is_the_gradient():
    N = 6
    _, psi = _net("fp64", N, seed=41)
    ex = A.ExactSummationSpins(N)
    op = A.Operator(ham.transverse_ising_chain(N, 1.0, 0.5))
    g1, e1 = A.ExpectationValue().gradient(op, psi, ex)
    g2, e2 = A.ExpectationValue().gradient_with_noise(op, psi, ex)
    assert np.array_equal(g1, g2) and e1 == e2
