"""HilbertSpaceDistance on the HIP path (HilbertSpaceDistance.cu.template:16-165),
checked the way the reference's own tests check it (test/test_HilbertSpaceDistance.py:
the distance against dense state vectors, the gradient against central finite
differences of the distance) and, for the generator form (is_unitary = False, where
exp(E_loc) replaces E_loc), against the dense exponential of a diagonal generator,
for which exp(E_loc(s)) = (e^G psi)(s) / psi(s) holds exactly.
"""
import math

import numpy as np
import pytest

import pyoracle as orc
from helpers import random_deep, expr_dense, oracle_operator
import annongpu_amd as A
from annongpu_amd import hamiltonians as ham

pytestmark = pytest.mark.gpu

N = 8


def net(seed, precision="fp64", sigma=0.15):
    spec = random_deep(N, [16], [N], seed=seed, sigma=sigma)
    psi = A.PsiDeep(*spec.args())
    psi.precision = precision
    return spec, psi


def dense_distance(psi_vec, psi_prime_vec, U):
    """sqrt(1 - |<psi'|U psi>|^2 / (<psi'|psi'> <U psi|U psi>))"""
    up = U @ psi_vec
    ov = np.vdot(psi_prime_vec, up)
    return math.sqrt(max(1.0 - abs(ov) ** 2 / (np.vdot(psi_prime_vec, psi_prime_vec).real * np.vdot(up, up).real),
                         1e-8))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_distance_unitary_against_dense_vectors(precision):
    _, psi = net(1, precision)
    _, psi1 = net(2, precision)
    ex = A.ExactSummationSpins(N)
    psi.calibrate(ex)
    H = ham.heisenberg_chain(N) + 0.3 * ham.transverse_ising_chain(N, 1.0, 0.7)
    U = 1 + 1j * H * 1e-2
    hsd = A.HilbertSpaceDistance(psi.num_params)
    Ud = expr_dense(U, N)
    tol = 1e-9 if precision == "fp64" else 2e-5
    # psi' = psi (the reference's test_distance1) and psi' another net
    for other in (psi, psi1):
        got = hsd(psi, other, A.Operator(U), True, ex)
        ref = dense_distance(psi.vector(ex), other.vector(ex), Ud)
        assert got == pytest.approx(ref, rel=tol, abs=1e-7), (got, ref)


def test_distance_generator_against_dense_exponential():
    """is_unitary = False: omega = exp(E_loc + conj(log psi' - log psi)); for a diagonal
    generator G, exp(E_loc) is the ratio of e^G psi to psi"""
    _, psi = net(3)
    _, psi1 = net(4)
    ex = A.ExactSummationSpins(N)
    G = 0.05j * sum((A.sigma_z(i) * A.sigma_z((i + 1) % N) for i in range(N)), A.PauliExpression()) \
        + 0.02 * A.sigma_z(3)
    Gd = expr_dense(G, N)
    assert np.allclose(Gd, np.diag(np.diag(Gd)))
    eG = np.diag(np.exp(np.diag(Gd)))
    hsd = A.HilbertSpaceDistance(psi.num_params)
    got = hsd(psi, psi1, A.Operator(G), False, ex)
    ref = dense_distance(psi.vector(ex), psi1.vector(ex), eG)
    assert got == pytest.approx(ref, rel=1e-9, abs=1e-7)


@pytest.mark.parametrize("unitary", [True, False])
def test_gradient_matches_finite_differences(unitary):
    """the reference's test_gradient1: gradient (nu = 1) = d distance / d Re(p_k)
    + i d distance / d Im(p_k), central differences with eps 1e-6"""
    _, psi = net(5)
    ex = A.ExactSummationSpins(N)
    psi.calibrate(ex)
    psi1 = psi.copy()
    psi1.params = psi.params + 0.05 * np.random.default_rng(1).standard_normal(psi.num_params)
    op = A.Operator(A.sigma_x(2) * A.sigma_y(5) + 0.5 * A.sigma_z(1) if unitary
                    else 0.1j * A.sigma_x(2) + 0.05 * A.sigma_z(1) * A.sigma_z(6))
    hsd = A.HilbertSpaceDistance(psi.num_params)
    grad, dist = hsd.gradient(psi, psi1, op, unitary, ex, 1.0)
    assert dist == pytest.approx(hsd(psi, psi1, op, unitary, ex), rel=1e-12)
    p0 = psi1.params.copy()
    eps = 1e-6

    def diff(delta):
        psi1.params = p0 + delta
        plus = hsd(psi, psi1, op, unitary, ex)
        psi1.params = p0 - delta
        minus = hsd(psi, psi1, op, unitary, ex)
        return (plus - minus) / (2 * eps)

    ref = np.zeros(psi.num_params, dtype=complex)
    for k in range(psi.num_params):
        d = np.zeros(psi.num_params, dtype=complex)
        d[k] = eps
        ref[k] = diff(d)
        d[k] = 1j * eps
        ref[k] += 1j * diff(d)
    psi1.params = p0
    # the visible biases a enter O_k (as s_i) but not log psi, a quirk the reference
    # shares (PsiDeep.hpp:239,261 leave a.s out of log psi, :353-362 emit O_k = s_i): the distance does
    # not move with them, so only the other parameters are compared
    assert np.all(ref[:N] == 0)
    assert np.allclose(grad[N:], ref[N:], rtol=1e-4, atol=1e-6), np.abs(grad[N:] - ref[N:]).max()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_distance_on_samples_matches_oracle(precision):
    """weighted samples of a SampleList: the HIP terms against the oracle's E_loc and
    log psi restated in numpy (the per-sample formulas of the reference)"""
    Nl = 70
    spec = random_deep(Nl, [140], [Nl], seed=9, sigma=0.05)
    spec1 = random_deep(Nl, [140], [Nl], seed=10, sigma=0.05)
    psi, psi1 = A.PsiDeep(*spec.args()), A.PsiDeep(*spec1.args())
    psi.precision = psi1.precision = precision
    rng = np.random.default_rng(4)
    confs = rng.integers(0, 2 ** 63, (400, 2), dtype=np.int64).astype(np.uint64)
    confs[:, 1] &= np.uint64((1 << (Nl - 64)) - 1)
    w = rng.random(400)
    H = ham.heisenberg_chain(Nl)
    U = 1 + 0.01j * H
    sl = A.SampleList(Nl, confs, weights=w)
    hsd = A.HilbertSpaceDistance(psi1.num_params)
    grad, got = hsd.gradient(psi, psi1, A.Operator(U), True, sl, 1.0)
    o, o1 = spec.oracle(), spec1.oracle()
    ref = orc.tdvp_accumulate(oracle_operator(U, 2), o, confs, w, with_S=False)
    ref1 = orc.tdvp_accumulate(None, o1, confs, w, with_S=False)
    lp = np.array([orc.log_psi_s(o, np.concatenate([c, np.zeros(2, np.uint64)])) for c in confs])
    lp1 = np.array([orc.log_psi_s(o1, np.concatenate([c, np.zeros(2, np.uint64)])) for c in confs])
    om = np.exp(np.conj(lp1 - lp)) * ref["E_loc"]
    ratio = np.exp(2 * (lp1.real - lp.real))
    A_ = np.sum(w * om)
    u = abs(A_) ** 2
    nn = np.sum(w * np.abs(ref["E_loc"]) ** 2)
    v = nn * np.sum(w * ratio)
    d_ref = math.sqrt(max(1 - u / v, 1e-8))
    tol = 1e-9 if precision == "fp64" else 1e-5
    assert got == pytest.approx(d_ref, rel=tol)
    Oc = ref1["O_samples"].conj()
    uk = np.conj(A_) * ((w * om) @ Oc)
    vk = nn * ((w * ratio) @ Oc)
    g_ref = -(uk * v - u * vk) / v ** 2 / d_ref
    assert np.linalg.norm(grad - g_ref) / np.linalg.norm(g_ref) < (1e-9 if precision == "fp64" else 1e-4)
