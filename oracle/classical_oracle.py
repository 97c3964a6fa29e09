"""numpy restatement of the reference's classical networks, Pauli-string basis and
KullbackLeibler terms.

TEST INFRASTRUCTURE ONLY: imported by tests/.  The product (annongpu_amd) never
imports it.

Restated from:
  PauliString::apply on Pauli strings     include/basis/PauliString.hpp:257-276
  PauliString::apply on spins             include/basis/PauliString.hpp:189-206,242-254
  PauliString::enumerate / network units  include/basis/PauliString.hpp:30-35,84-90
  Operator::fast_local_energy             include/operator/Operator.hpp:123-135
  Operator::nth_local_energy              include/operator/Operator.hpp:36-84
  PsiClassical log psi / O_k              include/quantum_state/PsiClassical.hpp:79-141
  KullbackLeibler terms                   source/network_functions/KullbackLeibler.cu.template:16-112,171-316

How each function is pinned (tests/test_classical_oracle.py): the Pauli-string phase
against Kronecker products of Pauli matrices; L_n on spins against the row sums of
the dense operator matrix and on Pauli strings against the sum of the expansion
coefficients of sigma(x) H_n; the Pauli-basis local energy against the expansion of
H Psi in Pauli strings; the KullbackLeibler value against the reference test's
closed form and its noise against the weighted variance of the per-sample term it
estimates.
"""
import numpy as np


def popcount(x):
    return bin(int(x)).count("1")


def pauli_phase(pa, pb, xa, xb):
    """coefficient of PauliString(pa, pb).apply(PauliString(xa, xb)) (PauliString.hpp:257-276)"""
    px, py, pz = pa & ~pb, ~pa & pb, pa & pb
    xx, xy, xz = xa & ~xb, ~xa & xb, xa & xb
    f = -1.0 if popcount((px & xz) | (py & xx) | (pz & xy)) & 1 else 1.0
    ne = popcount((pa | pb) & (xa | xb) & ((pa ^ xa) | (pb ^ xb)))
    if (ne & 3) > 1:
        f = -f
    return f * (-1j if ne & 1 else 1.0)


def spin_matrix_element(a, b, s):
    """<s|P|s'> of the string (a, b) on the spin configuration s (PauliString.hpp:242-254)"""
    ny = popcount(~a & b)
    f = 1.0 + 0j
    if (ny & 3) > 1:
        f = -f
    if ny & 1:
        f *= -1j
    if popcount(~s & b) & 1:   # b = Z | Y bits
        f = -f
    return f


def fast_local_energy(terms, conf, basis):
    """sum_m c_m (matrix-element coefficient of string m on conf) (Operator.hpp:123-135);
    terms: [((a, b), c)], conf: int (spins) or (a, b) (Pauli strings)"""
    tot = 0j
    for (a, b), c in terms:
        if basis == "paulis":
            tot += c * pauli_phase(a, b, conf[0], conf[1])
        else:
            tot += c * spin_matrix_element(a, b, conf)
    return tot


def pauli_units(xa, xb, n):
    """the 3 n input units of a Pauli string as a spin bit pattern (PauliString.hpp:84-90):
    unit 3 i + t - 1 is +1 iff site i holds Pauli type t"""
    bits = 0
    for i in range(n):
        t = ((xa >> i) & 1) | (((xb >> i) & 1) << 1)
        if t:
            bits |= 1 << (3 * i + t - 1)
    return bits


def enumerate_pauli(index):
    """PauliString::enumerate: a from the even bits, b from the odd bits of the index"""
    a = b = 0
    for k in range(32):
        a |= ((index >> (2 * k)) & 1) << k
        b |= ((index >> (2 * k + 1)) & 1) << k
    return a, b


class Classical:
    """PsiClassical: H_local as lists of ((a, b), c), params, log_prefactor, and an optional
    reference log psi function that enters with order 2 (PsiClassical.hpp:79-141)"""

    def __init__(self, H_local, params, log_prefactor=0.0, ref_log_psi=None, ref_O_k=None, order=1):
        self.H = [list(h) for h in H_local]
        self.params = np.asarray(params, dtype=complex)
        self.lp = complex(log_prefactor)
        self.ref_log_psi = ref_log_psi
        self.ref_O_k = ref_O_k
        self.order = order

    def L(self, conf, basis):
        return np.array([fast_local_energy(h, conf, basis) for h in self.H])

    def log_psi(self, conf, basis):
        r = self.lp + self.params @ self.L(conf, basis)
        if self.order > 1 and self.ref_log_psi is not None:
            r += self.ref_log_psi(conf)
        return r

    def O_k(self, conf, basis):
        o = self.L(conf, basis)
        if self.order > 1 and self.ref_O_k is not None:
            o = np.concatenate([o, self.ref_O_k(conf)])
        return o


def pauli_local_energy(terms, log_psi_fn, conf):
    """E_loc on a Pauli string (Operator.hpp:36-84 with PauliString::apply on Pauli strings):
    identity strings add c, the others c phase psi(x') / psi(x), x' = (a ^ x.a, b ^ x.b)"""
    xa, xb = conf
    l0 = log_psi_fn(conf)
    e = 0j
    for (a, b), c in terms:
        if a == 0 and b == 0:
            e += c
            continue
        e += c * pauli_phase(a, b, xa, xb) * np.exp(log_psi_fn((xa ^ a, xb ^ b)) - l0)
    return e


def kl_terms(lp, lpp, wp, threshold=0.0, last_mean=0.0, scale=1.0):
    """per-sample KullbackLeibler quantities (KullbackLeibler.cu.template:38-56): the
    reweighted weights, deviations and mask"""
    lp = scale * np.asarray(lp)
    lpp = np.asarray(lpp)
    w = np.exp(2 * (lp.real - lpp.real)) * np.asarray(wp)
    dev = lpp - lp - last_mean
    mask = np.abs(dev) ** 2 > threshold ** 2
    return w, dev, mask, lpp - lp


def kl_value(lp, lpp, wp, threshold=0.0, last_mean=0.0, scale=1.0):
    """KullbackLeibler::value (KullbackLeibler.cu.template:171-196)"""
    w, dev, mask, g = kl_terms(lp, lpp, wp, threshold, last_mean, scale)
    W = w.sum()
    d = np.sum(w * dev * mask) / W
    d2 = np.sum(w * np.abs(dev) ** 2 * mask) / W
    return np.sqrt(max(1e-8, d2 - abs(d) ** 2)), np.sum(w * g) / W, W


def kl_gradient(lp, lpp, wp, O, nu, threshold=0.0, last_mean=0.0, scale=1.0):
    """KullbackLeibler::gradient (KullbackLeibler.cu.template:199-242); O [samples][P] of psi'"""
    w, dev, mask, _ = kl_terms(lp, lpp, wp, threshold, last_mean, scale)
    W = w.sum()
    d = np.sum(w * dev * mask) / W
    d2 = np.sum(w * np.abs(dev) ** 2 * mask) / W
    value = np.sqrt(max(1e-8, d2 - abs(d) ** 2))
    Ok = (w @ O) / W
    dOc = ((w * dev * mask) @ O.conj()) / W
    return (dOc - d * Ok.conj()) / value ** nu, value
