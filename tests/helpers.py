"""Textbook references shared by the tests: dense Pauli matrices built from
Kronecker products, dense Hamiltonians, numpy forward passes, and builders
that create the same wave function in the oracle and in the HIP library."""
import numpy as np

import pyoracle as orc

# single-site matrices in the basis (|bit 0> = spin down, |bit 1> = spin up)
# with sigma_z |up> = +|up>  (include/basis/Spins.h:104-108)
I2 = np.eye(2, dtype=complex)
X2 = np.array([[0, 1], [1, 0]], dtype=complex)
Y2 = np.array([[0, 1j], [-1j, 0]], dtype=complex)   # Y|up> = i|down>, Y|down> = -i|up>
Z2 = np.array([[-1, 0], [0, 1]], dtype=complex)
SINGLE = {0: I2, 1: X2, 2: Y2, 3: Z2}


def pauli_dense(a, b, n):
    """Dense matrix of the Pauli string (a, b) on n sites; site n-1 is the most
    significant bit of the basis index."""
    m = np.ones((1, 1), dtype=complex)
    for site in reversed(range(n)):
        t = ((a >> site) & 1) | (((b >> site) & 1) << 1)
        m = np.kron(m, SINGLE[t])
    return m


def expr_dense(expr, n):
    H = np.zeros((2 ** n, 2 ** n), dtype=complex)
    for (a, b), c in expr:
        H += c * pauli_dense(a, b, n)
    return H


def operator_arrays(expr, words=4):
    terms = list(expr)
    a = np.zeros((len(terms), words), dtype=np.uint64)
    b = np.zeros((len(terms), words), dtype=np.uint64)
    c = np.zeros(len(terms), dtype=np.complex128)
    for k, ((ma, mb), coef) in enumerate(terms):
        for w in range(words):
            a[k, w] = (ma >> (64 * w)) & 0xFFFFFFFFFFFFFFFF
            b[k, w] = (mb >> (64 * w)) & 0xFFFFFFFFFFFFFFFF
        c[k] = coef
    return a, b, c


def oracle_operator(expr, words):
    a, b, c = operator_arrays(expr, words)
    return orc.OracleOperator(a, b, c, words)


def spins_of(idx, n):
    return np.array([1.0 if (idx >> i) & 1 else -1.0 for i in range(n)])


def conf_array(idx):
    return orc.conf_words(idx, 4)


def cgauss(rng, shape, sigma):
    return sigma * rng.standard_normal(shape) + 1j * sigma * rng.standard_normal(shape)


class DeepSpec:
    """Arguments of PsiDeep(num_sites, a, b, connections, W, final_weights, log_prefactor)."""

    def __init__(self, num_sites, a, b, connections, W, final_weights, log_prefactor=0.0):
        self.num_sites = num_sites
        self.a = np.asarray(a, dtype=complex)
        self.b = [np.asarray(x, dtype=complex) for x in b]
        self.connections = [np.asarray(x, dtype=np.uint32) for x in connections]
        self.W = [np.asarray(x, dtype=complex) for x in W]
        self.final_weights = np.asarray(final_weights, dtype=complex)
        self.log_prefactor = log_prefactor

    def args(self):
        return (self.num_sites, self.a, self.b, self.connections, self.W, self.final_weights, self.log_prefactor)

    def oracle(self):
        return orc.deep(*self.args())

    def params(self):
        p = [self.a]
        for b, w in zip(self.b, self.W):
            p += [b, w.ravel()]
        return np.concatenate(p)


def connections_1d(N, M, C):
    out = []
    for n, m, c in zip([N] + list(M[:-1]), M, C):
        d = 1 if m > n else (n // m if n % m == 0 else c)
        out.append(np.array([[(j * d + i) % n for j in range(m)] for i in range(c)], dtype=np.uint32))
    return out


def random_deep(N, M, C, seed, sigma=0.1, final=1.0, log_prefactor=0.0):
    rng = np.random.default_rng(seed)
    conns = connections_1d(N, M, C)
    a = cgauss(rng, N, sigma)
    b = [cgauss(rng, m, sigma) for m in M]
    W = [cgauss(rng, (c, m), sigma) for c, m in zip(C, M)]
    fw = final * (1 + 0.3 * cgauss(rng, M[-1], 1.0)) if final is not None else np.ones(M[-1])
    return DeepSpec(N, a, b, conns, W, fw, log_prefactor)


def logcosh_poly(z, layer):
    z2 = z * z
    z4 = z2 * z2
    if layer == 0:
        return 0.5 * z2 - z4 / 12 + z4 * z2 / 45
    return z - z2 * z / 3 + 2 * z4 * z / 15


def deep_log_psi_numpy(spec, idx):
    """numpy forward pass over the gather tables (independent of the oracle)."""
    act = spins_of(idx, spec.num_sites).astype(complex)
    for l, (b, c, w) in enumerate(zip(spec.b, spec.connections, spec.W)):
        z = (w * act[c]).sum(axis=0) + b
        act = logcosh_poly(z, l)
    return spec.log_prefactor + act @ spec.final_weights
