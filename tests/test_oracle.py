"""Pins the fp64 CPU oracle against known answers, closed forms, textbook dense
linear algebra and finite differences (the reference's own test strategy:
test/test_Psi.py, test_operator.py, test_ExpectationValue.py)."""
import cmath
import math

import numpy as np
import pytest

import pyoracle as orc
from helpers import (pauli_dense, expr_dense, oracle_operator, conf_array, spins_of, random_deep,
                     deep_log_psi_numpy, logcosh_poly, cgauss, connections_1d, DeepSpec)
from annongpu_amd.pauli import sigma_x, sigma_y, sigma_z, PauliExpression
from annongpu_amd import hamiltonians as ham


# ------------------------------------------------------------------ RNG

def test_philox_known_answers():
    """Random123 known-answer vectors for Philox4x32-10."""
    assert list(orc.philox([0, 0, 0, 0], [0, 0])) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert list(orc.philox([0xffffffff] * 4, [0xffffffff] * 2)) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert list(orc.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0])) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


# ----------------------------------------------------------- activations

@pytest.mark.parametrize("z", [0.05 + 0.02j, -0.03 + 0.04j, 0.01 - 0.07j])
def test_activation_polynomials_are_taylor_series(z):
    """lc(z) = log cosh z + O(z^8); tanh polynomial = tanh z + O(z^7); derivative chain."""
    assert abs(orc.activation(z, 0) - cmath.log(cmath.cosh(z))) < 20 * abs(z) ** 8
    assert abs(orc.activation(z, 0, True) - cmath.tanh(z)) < 20 * abs(z) ** 7
    assert abs(orc.activation(z, 1) - cmath.tanh(z)) < 20 * abs(z) ** 7
    assert abs(orc.activation(z, 1, True) - 1 / cmath.cosh(z) ** 2) < 20 * abs(z) ** 6
    for layer in (0, 1):
        h = 1e-5
        fd = (orc.activation(z + h, layer) - orc.activation(z - h, layer)) / (2 * h)
        assert abs(fd - orc.activation(z, layer, True)) < 1e-9


# ------------------------------------------------------- Pauli strings

def _random_strings(rng, n, count):
    out = []
    for _ in range(count):
        types = rng.integers(0, 4, n)
        a = sum(int(t & 1) << i for i, t in enumerate(types))
        b = sum(int(t >> 1) << i for i, t in enumerate(types))
        out.append((a, b))
    return out


def test_pauli_apply_matches_kronecker_products():
    """<s|P|s'> of PauliString::apply equals the dense Kronecker-product matrix,
    entry by entry, for random strings on 5 sites (all 32 basis states)."""
    rng = np.random.default_rng(0)
    n = 5
    for a, b in _random_strings(rng, n, 40) + [(0, 0), (1, 0), (0, 1), (1, 1)]:
        D = pauli_dense(a, b, n)
        for s in range(2 ** n):
            out, f = orc.pauli_apply(conf_array(a), conf_array(b), conf_array(s), 1)
            sp = int(out[0])
            row = D[s].copy()
            assert row[sp] == pytest.approx(f)
            row[sp] = 0
            assert np.all(row == 0)


def test_pauli_apply_multiword():
    """strings and configurations spanning several 64-bit words"""
    rng = np.random.default_rng(1)
    for _ in range(50):
        n = 150
        s = int(rng.integers(0, 2 ** 62)) | (int(rng.integers(0, 2 ** 62)) << 64) | (int(rng.integers(0, 2 ** 20)) << 128)
        sites = rng.choice(n, 4, replace=False)
        types = rng.integers(1, 4, 4)
        expr = PauliExpression.identity()
        for site, t in zip(sites, types):
            expr = expr * PauliExpression.single(int(site), int(t))
        ((a, b), c), = list(expr)
        out, f = orc.pauli_apply(conf_array(a), conf_array(b), conf_array(s), 3)
        # closed form per site: X flips; Y flips with factor (i if down->... ) ; Z sign
        exp = complex(c) ** 0
        sp = s
        for site, t in zip(sites, types):
            up = (s >> int(site)) & 1
            if t == 1:
                sp ^= 1 << int(site)
            elif t == 2:
                sp ^= 1 << int(site)
                exp *= (-1j if up else 1j)     # <s|Y|s'> with s' the flipped state
            else:
                exp *= 1 if up else -1
        got = sum(int(x) << (64 * w) for w, x in enumerate(out))
        assert got == sp
        assert f == pytest.approx(exp)


def test_pauli_products_follow_the_algebra():
    X, Y, Z = sigma_x(0), sigma_y(0), sigma_z(0)
    assert X * Y == 1j * Z
    assert Y * Z == 1j * X
    assert Z * X == 1j * Y
    assert Y * X == -1j * Z
    assert X * X == PauliExpression.identity()
    n = 3
    rng = np.random.default_rng(2)
    for _ in range(20):
        e1 = PauliExpression()
        e2 = PauliExpression()
        for _ in range(3):
            e1 = e1 + complex(*rng.standard_normal(2)) * PauliExpression.single(int(rng.integers(n)), int(rng.integers(1, 4)))
            e2 = e2 + complex(*rng.standard_normal(2)) * PauliExpression.single(int(rng.integers(n)), int(rng.integers(1, 4)))
        assert np.allclose(expr_dense(e1 * e2, n), expr_dense(e1, n) @ expr_dense(e2, n))


def test_hamiltonian_closed_forms():
    """Ground-state energies: critical TFIM ring E0 = -2/sin(pi/2N) (Jordan-Wigner);
    Heisenberg rings of 4 and 6 sites (Pauli units): -8 and -11.2111."""
    N = 8
    E0 = np.linalg.eigvalsh(expr_dense(ham.transverse_ising_chain(N, 1.0, 1.0), N))[0]
    assert E0 == pytest.approx(-2.0 / math.sin(math.pi / (2 * N)), rel=1e-10)
    assert np.linalg.eigvalsh(expr_dense(ham.heisenberg_chain(4), 4))[0] == pytest.approx(-8.0, rel=1e-10)
    assert np.linalg.eigvalsh(expr_dense(ham.heisenberg_chain(6), 6))[0] == pytest.approx(4 * -2.8027756377, rel=1e-8)
    # 2x2 square with periodic wrap is a 4-ring with doubled bonds: spectrum = 2 x ring
    assert len(ham.square_bonds(4)) == 32
    assert len(ham.square_bonds(10)) == 200


# -------------------------------------------------------- wave functions

def test_rbm_log_psi_numpy():
    rng = np.random.default_rng(3)
    N, M = 6, 9
    W = cgauss(rng, (N, M), 0.2)
    fw, lp = 0.7 - 0.2j, 0.3 + 0.1j
    psi = orc.rbm(W, fw, lp)
    for idx in rng.integers(0, 2 ** N, 10):
        s = spins_of(int(idx), N)
        ref = lp + fw * logcosh_poly(s @ W, 0).sum()
        assert orc.log_psi_s(psi, conf_array(int(idx))) == pytest.approx(ref, rel=1e-13, abs=1e-14)


@pytest.mark.parametrize("shape", [(6, [12], [6]), (6, [6], [3]), (3, [18, 9, 3], [3, 2, 3]), (2, [2], [2])])
def test_deep_log_psi_numpy(shape):
    N, M, C = shape
    spec = random_deep(N, M, C, seed=4, sigma=0.2, log_prefactor=0.2 - 0.4j)
    psi = spec.oracle()
    for idx in range(2 ** N):
        assert orc.log_psi_s(psi, conf_array(idx)) == pytest.approx(deep_log_psi_numpy(spec, idx), rel=1e-12, abs=1e-13)


def test_deep_log_psi_omits_input_weights():
    """PsiDeep.hpp:260-262: a.s is not part of log psi"""
    spec = random_deep(4, [8], [4], seed=5)
    psi1 = spec.oracle()
    spec.a = spec.a + 3.0
    psi2 = spec.oracle()
    for idx in range(16):
        assert orc.log_psi_s(psi1, conf_array(idx)) == orc.log_psi_s(psi2, conf_array(idx))


def _nd_convolve(inputs, weights):
    """periodic forward-shifted correlation (test/test_Psi.py:30-47)"""
    out = np.zeros(inputs.shape, dtype=complex)
    for shift in np.ndindex(*inputs.shape):
        rolled = np.roll(inputs, [-s for s in shift], axis=tuple(range(inputs.ndim)))
        out[shift] = np.sum(rolled[tuple(slice(0, w) for w in weights.shape)] * weights)
    return out


def _cnn_spec(extent, layers, seed, sigma=0.1):
    rng = np.random.default_rng(seed)
    n = 0
    prev = 1
    for ch, conn in layers:
        n += prev * ch * int(np.prod(conn))
        prev = ch
    return dict(extent=extent, channels=[l[0] for l in layers], conn=[l[1] for l in layers],
                params=cgauss(rng, n, sigma), final_factor=1.3)


def _cnn_oracle(spec, lp=0.0):
    N = int(np.prod(spec["extent"]))
    return orc.cnn(spec["extent"], spec["channels"], spec["conn"], np.zeros(N), spec["params"], spec["final_factor"], lp)


def _cnn_numpy(spec, idx, lp=0.0):
    ext = spec["extent"]
    N = int(np.prod(ext))
    acts = [spins_of(idx, N).reshape(ext).astype(complex)]
    off = 0
    prev = 1
    for l, (ch, conn) in enumerate(zip(spec["channels"], spec["conn"])):
        vol = int(np.prod(conn))
        new = []
        for cj in range(ch):
            z = np.zeros(ext, dtype=complex)
            for ci in range(prev):
                w = spec["params"][off + (ci * ch + cj) * vol: off + (ci * ch + cj + 1) * vol].reshape(conn)
                z += _nd_convolve(acts[ci], w)
            new.append(logcosh_poly(z, l))
        off += prev * ch * vol
        prev = ch
        acts = new
    return lp + spec["final_factor"] * sum(a.sum() for a in acts)


@pytest.mark.parametrize("extent,layers", [
    ([2, 4, 4], [(2, [2, 4, 4]), (4, [2, 2, 2])]),          # test/conftest.py:41-45
    ([1, 3, 4], [(3, [1, 3, 3]), (2, [1, 2, 3])]),
])
def test_cnn_log_psi_numpy(extent, layers):
    spec = _cnn_spec(extent, layers, seed=6)
    psi = _cnn_oracle(spec, 0.1j)
    rng = np.random.default_rng(7)
    N = int(np.prod(extent))
    for idx in rng.integers(0, 2 ** N, 6):
        assert orc.log_psi_s(psi, conf_array(int(idx))) == pytest.approx(_cnn_numpy(spec, int(idx), 0.1j), rel=1e-11)


def test_cnn_translation_invariance():
    spec = _cnn_spec([1, 4, 4], [(2, [1, 3, 3]), (2, [1, 3, 3])], seed=8)
    psi = _cnn_oracle(spec)
    rng = np.random.default_rng(9)
    for idx in rng.integers(0, 2 ** 16, 4):
        s = spins_of(int(idx), 16).reshape(4, 4)
        v0 = orc.log_psi_s(psi, conf_array(int(idx)))
        for sh in [(1, 0), (0, 1), (2, 3)]:
            r = np.roll(s, sh, axis=(0, 1)).ravel()
            j = sum(1 << i for i, x in enumerate(r) if x > 0)
            assert orc.log_psi_s(psi, conf_array(j)) == pytest.approx(v0, rel=1e-12)


def _fd_O_k(psi_builder, params, idx, eps=1e-5):
    out = np.zeros(params.size, dtype=complex)
    for k in range(params.size):
        p = params.copy()
        p[k] += eps
        lp_p = orc.log_psi_s(psi_builder(p), conf_array(idx))
        p[k] -= 2 * eps
        lp_m = orc.log_psi_s(psi_builder(p), conf_array(idx))
        out[k] = (lp_p - lp_m) / (2 * eps)
    return out


def test_O_k_finite_differences_rbm():
    rng = np.random.default_rng(10)
    N, M = 4, 6
    W = cgauss(rng, (N, M), 0.3)
    build = lambda p: orc.rbm(p.reshape(N, M), 1.4, 0.0)
    for idx in (3, 9, 14):
        assert np.allclose(orc.psi_O_k(build(W.ravel()), conf_array(idx)), _fd_O_k(build, W.ravel(), idx),
                           rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("shape", [(4, [8], [4]), (3, [6, 3], [3, 2]), (3, [18, 9, 3], [3, 2, 3])])
def test_O_k_finite_differences_deep(shape):
    N, M, C = shape
    spec = random_deep(N, M, C, seed=11, sigma=0.3)

    def build(p):
        o = N
        b, W = [], []
        for m, c in zip(M, C):
            b.append(p[o:o + m])
            o += m
            W.append(p[o:o + c * m].reshape(c, m))
            o += c * m
        return DeepSpec(N, p[:N], b, spec.connections, W, spec.final_weights).oracle()

    p0 = spec.params()
    for idx in (1, 5):
        fd = _fd_O_k(build, p0, idx)
        fd[:N] = spins_of(idx, N)       # quirk: O_k of a is s_i although log psi has no a.s
        assert np.allclose(orc.psi_O_k(build(p0), conf_array(idx)), fd, rtol=1e-6, atol=1e-8)


def test_O_k_finite_differences_cnn():
    spec = _cnn_spec([1, 3, 4], [(2, [1, 2, 3]), (2, [1, 2, 2])], seed=12, sigma=0.3)

    def build(p):
        s = dict(spec)
        s["params"] = p
        return _cnn_oracle(s)

    for idx in (77, 2049):
        fd = _fd_O_k(build, spec["params"], idx)
        assert np.allclose(orc.psi_O_k(build(spec["params"]), conf_array(idx)), fd, rtol=1e-6, atol=1e-8)


# ---------------------------------------------------------- local energy

@pytest.mark.parametrize("hname", ["tfim", "heis", "mixed"])
def test_apply_operator_is_H_psi(hname):
    """psi(s) E_loc(s) = (H psi)(s)   (test/test_operator.py:6-26)"""
    N = 6
    if hname == "tfim":
        H = ham.transverse_ising_chain(N, 1.0, 0.7)
    elif hname == "heis":
        H = ham.heisenberg_chain(N)
    else:
        H = 0.3 * sigma_y(2) + sigma_x(0) * sigma_z(3) * sigma_y(5) - 0.5j * sigma_y(1) * sigma_y(4) + 0.25
    spec = random_deep(N, [12], [N], seed=13, sigma=0.1)
    psi = spec.oracle()
    op = oracle_operator(H, 1)
    lps = orc.exact_log_psi_vector(psi, N)
    vec = np.exp(lps)
    HD = expr_dense(H, N)
    got = np.array([vec[s] * orc.local_energy(op, psi, conf_array(s), lps[s]) for s in range(2 ** N)])
    assert np.allclose(got, HD @ vec, rtol=1e-11, atol=1e-13)


# ----------------------------------------------------------- statistics

def test_tdvp_exact_statistics_dense():
    """E, E2, <O>, F, S of the oracle vs dense linear algebra on the 2^N vector."""
    N = 6
    spec = random_deep(N, [6], [N], seed=14, sigma=0.3)
    psi = spec.oracle()
    H = ham.transverse_ising_chain(N, 1.0, 0.9)
    op = oracle_operator(H, 1)
    lps = orc.exact_log_psi_vector(psi, N)
    w = np.exp(2 * lps.real)
    confs = orc.exact_confs(N, 1)
    r = orc.tdvp_accumulate(op, psi, confs, w)
    vec = np.exp(lps)
    El = (expr_dense(H, N) @ vec) / vec
    O = np.array([orc.psi_O_k(psi, conf_array(s)) for s in range(2 ** N)])
    E = np.sum(w * El)
    assert r["E"] == pytest.approx(E, rel=1e-12)
    assert r["E2"] == pytest.approx(np.sum(w * abs(El) ** 2), rel=1e-12)
    Om = w @ O
    assert np.allclose(r["O"], Om, rtol=1e-12)
    assert np.allclose(r["F"], (w * El) @ O.conj() - E * Om.conj(), rtol=1e-10, atol=1e-12)
    S = (O.conj().T * w) @ O - np.outer(Om.conj(), Om)
    assert np.allclose(r["S"], S, rtol=1e-10, atol=1e-12)


def test_solvers_match_numpy():
    rng = np.random.default_rng(15)
    P = 40
    A = cgauss(rng, (80, P), 1.0)
    S = A.conj().T @ A / 80
    F = cgauss(rng, P, 1.0)
    x_ref = np.linalg.solve(S + 1e-3 * np.eye(P), F)
    assert np.allclose(orc.cholesky_solve(S, 1e-3, F), x_ref, rtol=1e-9)
    x, it, rc = orc.cg_solve(S, 1e-3, F, rtol=1e-12, max_iter=2000)
    assert rc == 0 and np.allclose(x, x_ref, rtol=1e-7)


# ------------------------------------------------------------ Monte Carlo

def _exact_energy(psi, H, N):
    lps = orc.exact_log_psi_vector(psi, N)
    vec = np.exp(lps)
    vec = vec / np.linalg.norm(vec)
    return np.vdot(vec, expr_dense(H, N) @ vec).real, np.abs(vec) ** 2


@pytest.mark.parametrize("move", [orc.MOVE_FLIP, orc.MOVE_EXCHANGE])
def test_mc_energy_within_standard_errors(move):
    N = 8
    spec = random_deep(N, [16], [N], seed=16, sigma=0.5)
    psi = spec.oracle()
    H = ham.heisenberg_chain(N) if move == orc.MOVE_EXCHANGE else ham.transverse_ising_chain(N, 1.0, 0.8)
    op = oracle_operator(H, 1)
    r = orc.mc_run(psi, num_chains=256, steps_per_chain=40, num_sweeps=2, num_thermalization_sweeps=20, move=move,
                   seed=123)
    w = np.full(len(r["samples"]), 1.0 / len(r["samples"]))
    acc = orc.tdvp_accumulate(op, psi, r["samples"], w, with_S=False)
    el = acc["E_loc"].real.reshape(40, 256)
    chain_means = el.mean(axis=0)
    se = chain_means.std(ddof=1) / math.sqrt(256)
    if move == orc.MOVE_EXCHANGE:
        # exact energy restricted to the S_z = 0 sector
        lps = orc.exact_log_psi_vector(psi, N)
        vec = np.exp(lps)
        sz = np.array([bin(i).count("1") * 2 - N for i in range(2 ** N)])
        vec[sz != 0] = 0
        vec /= np.linalg.norm(vec)
        E = np.vdot(vec, expr_dense(H, N) @ vec).real
        ups = [bin(int(c)).count("1") for c in r["samples"][:, 0]]
        assert set(ups) == {N // 2}
    else:
        E, _ = _exact_energy(psi, H, N)
    assert abs(el.mean() - E) < 4 * se + 1e-12


def test_mc_visits_follow_psi_squared():
    """chi-square of visit frequencies against |psi|^2 (single flips, 5 sites)"""
    N = 5
    spec = random_deep(N, [5], [N], seed=17, sigma=0.25)
    psi = spec.oracle()
    r = orc.mc_run(psi, num_chains=400, steps_per_chain=50, num_sweeps=3, num_thermalization_sweeps=10,
                   move=orc.MOVE_FLIP, seed=7)
    counts = np.bincount(r["samples"][:, 0].astype(np.int64), minlength=2 ** N)
    _, p = _exact_energy(psi, ham.transverse_ising_chain(N), N)
    expected = p * counts.sum()
    chi2 = np.sum((counts - expected) ** 2 / expected)
    assert chi2 < 2 ** N + 6 * math.sqrt(2 * 2 ** N)


def test_mc_counter_continuation():
    """two calls with persistent chains equal one call of twice the length"""
    N = 6
    psi = random_deep(N, [6], [N], seed=18, sigma=0.4).oracle()
    r1 = orc.mc_run(psi, 8, 2, 3, 0, seed=5)
    r_a = orc.mc_run(psi, 8, 1, 3, 0, seed=5)
    r_b = orc.mc_run(psi, 8, 1, 3, 0, seed=5, counter=3 * N, chain_confs=r_a["chains"])
    assert np.array_equal(r1["samples"][8:], r_b["samples"])


def test_matrix_free_sr_cg_matches_dense_solve():
    """the CPU baseline's matrix-free SR solve equals numpy's dense solve of
    (Oc^H diag(w) Oc + shift) x = F, Oc = O - <O>"""
    rng = np.random.default_rng(3)
    ns, P = 60, 40
    O = rng.standard_normal((ns, P)) + 1j * rng.standard_normal((ns, P))
    w = rng.random(ns)
    w /= w.sum()
    Ob = w @ O
    F = rng.standard_normal(P) + 1j * rng.standard_normal(P)
    Oc = O - Ob
    S = (Oc.conj().T * w) @ Oc
    x_ref = np.linalg.solve(S + 0.05 * np.eye(P), F)
    x, it, res = orc.sr_cg_matfree(O, w, Ob, F, 0.05, rtol=1e-12, max_iter=500)
    assert res <= 1e-12 and it <= 500
    assert np.linalg.norm(x - x_ref) < 1e-9 * np.linalg.norm(x_ref)
    # a dropped centring term would be caught: the uncentred system has another solution
    S_unc = (O.conj().T * w) @ O
    assert np.linalg.norm(np.linalg.solve(S_unc + 0.05 * np.eye(P), F) - x_ref) > 1e-3 * np.linalg.norm(x_ref)
