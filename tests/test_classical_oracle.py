"""Pins of oracle/classical_oracle.py against what the mathematics fixes (CPU only):

- the Pauli-string phase of PauliString::apply (PauliString.hpp:257-276) against
  Kronecker products of Pauli matrices: it is the coefficient of sigma(x) sigma(P),
  the configuration times the string;
- the spin-basis matrix element (PauliString.hpp:242-254) against the dense matrix;
- L_n = Operator::fast_local_energy (Operator.hpp:123-135): on spins the row sum of the
  dense operator matrix, on Pauli strings the sum of the Pauli-expansion coefficients of
  sigma(x) H_n;
- the Pauli-basis local energy: E_loc(x) psi(x) is the coefficient of sigma(x) in H Psi,
  Psi = sum_y psi(y) sigma(y) (left multiplication by H in the Pauli-string basis);
- PauliString::enumerate and the three input units per site;
- the KullbackLeibler value against the closed form of the reference's test
  (test/test_KullbackLeibler.py:18-44), its gradient against central differences and
  its noise formula (KullbackLeibler.cu.template:300-312) against the weighted variance of
  the per-sample term it estimates.
"""
import itertools

import numpy as np
import pytest

import classical_oracle as co
from helpers import pauli_dense, expr_dense, cgauss
import annongpu_amd as A


def all_strings(n):
    return [co.enumerate_pauli(i) for i in range(4 ** n)]


def expand(M, n):
    """Pauli-expansion coefficients {(a, b): tr(sigma(a, b) M) / 2^n}"""
    return {(a, b): np.trace(pauli_dense(a, b, n) @ M) / 2 ** n for a, b in all_strings(n)}


@pytest.mark.parametrize("n", [1, 2, 3])
def test_pauli_phase_is_configuration_times_string(n):
    for (pa, pb), (xa, xb) in itertools.product(all_strings(n), repeat=2):
        lhs = pauli_dense(xa, xb, n) @ pauli_dense(pa, pb, n)
        rhs = co.pauli_phase(pa, pb, xa, xb) * pauli_dense(pa ^ xa, pb ^ xb, n)
        assert np.allclose(lhs, rhs), (pa, pb, xa, xb)


def test_pauli_phase_is_not_string_times_configuration():
    # a plausible slip (the other product order) differs: X . Y = iZ but Y . X = -iZ
    assert co.pauli_phase(1, 0, 0, 1) == pytest.approx(-1j)


def test_spin_matrix_element_against_dense():
    n = 3
    for a, b in all_strings(n):
        M = pauli_dense(a, b, n)
        for s in range(2 ** n):
            sp = s ^ (a ^ b)    # the flipped configuration (X, Y sites)
            assert co.spin_matrix_element(a, b, s) == pytest.approx(M[s, sp]), (a, b, s)


def test_fast_local_energy_spins_is_the_row_sum():
    n = 4
    H = A.sigma_z(0) * A.sigma_z(1) + 0.7 * A.sigma_x(0) - 0.3j * A.sigma_y(2) * A.sigma_x(3) + 0.2 * A.sigma_y(1)
    rows = expr_dense(H, n).sum(axis=1)
    for s in range(2 ** n):
        assert co.fast_local_energy(list(H), s, "spins") == pytest.approx(rows[s])


def test_fast_local_energy_paulis_is_the_expansion_sum():
    n = 2
    H = A.sigma_z(0) * A.sigma_y(1) + 1.1 * A.sigma_x(0) - 0.4j * A.sigma_y(1)
    Hd = expr_dense(H, n)
    for xa, xb in all_strings(n):
        coeffs = expand(pauli_dense(xa, xb, n) @ Hd, n)
        assert co.fast_local_energy(list(H), (xa, xb), "paulis") == pytest.approx(sum(coeffs.values()))


def test_pauli_local_energy_is_left_multiplication():
    n = 3
    rng = np.random.default_rng(3)
    strings = all_strings(n)
    lpsi = {s: complex(v) for s, v in zip(strings, cgauss(rng, len(strings), 0.5))}
    Psi = sum(np.exp(lpsi[s]) * pauli_dense(*s, n) for s in strings)
    H = A.sigma_z(0) * A.sigma_z(1) + 0.6 * A.sigma_x(2) + 0.3 * A.sigma_y(0) * A.sigma_x(1) + 0.25 * A.identity()
    coeffs = expand(expr_dense(H, n) @ Psi, n)
    for s in strings:
        e = co.pauli_local_energy(list(H), lambda c: lpsi[c], s)
        assert e * np.exp(lpsi[s]) == pytest.approx(coeffs[s], abs=1e-12), s


def test_enumerate_and_units():
    for idx in range(4 ** 3):
        a, b = co.enumerate_pauli(idx)
        types = [((idx >> (2 * i)) & 1) | (((idx >> (2 * i + 1)) & 1) << 1) for i in range(3)]
        assert [((a >> i) & 1) | (((b >> i) & 1) << 1) for i in range(3)] == types
        u = co.pauli_units(a, b, 3)
        for i, t in enumerate(types):
            block = (u >> (3 * i)) & 7
            assert block == (0 if t == 0 else 1 << (t - 1))
    # the package's PauliString agrees with the oracle's enumeration
    for idx in (0, 5, 77, 4 ** 5 - 1):
        p = A.PauliString.enumerate(idx)
        assert (p.a, p.b) == co.enumerate_pauli(idx)


def _kl_case(seed, n=40, P=5):
    rng = np.random.default_rng(seed)
    lp = cgauss(rng, n, 0.5)
    lpp = cgauss(rng, n, 0.5)
    wp = rng.random(n) + 0.1
    O = cgauss(rng, (n, P), 1.0)
    return lp, lpp, wp, O


def test_kl_value_closed_form():
    lp, lpp, wp, _ = _kl_case(1)
    value, mean_dev, W = co.kl_value(lp, lpp, wp)
    p = np.exp(2 * (lp.real - lpp.real)) * wp
    p /= p.sum()
    d = lpp - lp
    ref = np.sqrt(p @ np.abs(d) ** 2 - abs(p @ d) ** 2)
    assert value == pytest.approx(ref, rel=1e-12)
    assert mean_dev == pytest.approx(p @ d, rel=1e-12)
    # the shift by the last mean deviation leaves the value (a variance) unchanged
    assert co.kl_value(lp, lpp, wp, last_mean=0.3 - 0.2j)[0] == pytest.approx(value, rel=1e-12)


def test_kl_gradient_finite_differences():
    """d value / d Re p_k + i d value / d Im p_k with log psi'(s) = l0(s) + sum_k p_k O_k(s)
    (a model that is linear in its parameters, whose O_k are exactly O)"""
    lp, lpp0, wp, O = _kl_case(2)
    P = O.shape[1]
    params = np.zeros(P, dtype=complex)

    def value(p):
        lpp = lpp0 + O @ p
        # the samples follow psi': wp carries |psi'|^2 up to a constant
        w = wp * np.exp(2 * (lpp.real - lpp0.real))
        return co.kl_value(lp, lpp, w)[0]

    g, v = co.kl_gradient(lp, lpp0, wp, O, 1.0)
    assert v == pytest.approx(value(params), rel=1e-12)
    eps = 1e-6
    for k in range(P):
        e = np.zeros(P, dtype=complex)
        e[k] = eps
        ref = (value(e) - value(-e)) / (2 * eps)
        e[k] = 1j * eps
        ref += 1j * (value(e) - value(-e)) / (2 * eps)
        assert g[k] == pytest.approx(ref, rel=1e-5, abs=1e-8), k


@pytest.mark.parametrize("threshold", [0.0, 0.5])
def test_kl_noise_is_the_variance_of_the_sample_term(threshold):
    lp, lpp, wp, O = _kl_case(4, n=200)
    w, dev, mask, _ = co.kl_terms(lp, lpp, wp, threshold)
    W = w.sum()
    p = w / W
    m = mask.astype(float)
    d = np.sum(p * dev * m)
    d2 = np.sum(p * np.abs(dev) ** 2 * m)
    Ok = p @ O
    # the reference's expansion (KullbackLeibler.cu.template:300-312)
    dOc = (p * dev * m) @ O.conj()
    d2Ok2 = (p * np.abs(dev) ** 2 * m) @ np.abs(O) ** 2
    dOk = (p * dev * m) @ O
    dOk2 = (p * dev * m) @ np.abs(O) ** 2
    d2Ok = (p * np.abs(dev) ** 2 * m) @ O
    Ok2 = p @ np.abs(O) ** 2
    var = (d2Ok2 - np.abs(dOc) ** 2 + 2 * (dOk * np.conj(d) * Ok.conj() + 2 * dOc.conj() * d * Ok.conj()
                                           - d2Ok * Ok.conj() - dOk2 * np.conj(d)).real
           + d2 * np.abs(Ok) ** 2 + abs(d) ** 2 * Ok2 - 4 * abs(d) ** 2 * np.abs(Ok) ** 2)
    # the per-sample term whose weighted variance it is
    X = (m * dev)[:, None] * O.conj() - (m * dev)[:, None] * Ok.conj()[None, :] - d * O.conj()
    mean = p @ X
    direct = p @ np.abs(X - mean[None, :]) ** 2
    assert np.allclose(var, direct, rtol=1e-10, atol=1e-12)
