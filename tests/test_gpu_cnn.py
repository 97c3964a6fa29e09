"""PsiCNN on the HIP path vs the fp64 oracle (needs an MI355X).

Tolerances as in test_gpu_parity: fp64 device mode 1e-10 relative, fp32 device
mode 1e-5 relative on log psi / energies and 1e-4 on vectors of O_k statistics.
"""
import numpy as np
import pytest

import pyoracle as orc
from helpers import conf_array, oracle_operator, cgauss
import annongpu_amd as A
from annongpu_amd import hamiltonians as ham

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-10, "fp32": 1e-5}


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def make_cnn(extent, layers, seed, sigma=0.1, sym=None, final_factor=1.3, lp=0.0, precision="fp64"):
    """the same PsiCNN on the device and in the oracle"""
    rng = np.random.default_rng(seed)
    N = int(np.prod(extent))
    sym = np.zeros(N, dtype=np.uint32) if sym is None else np.asarray(sym, dtype=np.uint32)
    nsym = len(set(sym.tolist()))
    n = 0
    prev = 1
    for ch, conn in layers:
        n += prev * ch * nsym * int(np.prod(conn))
        prev = ch
    params = cgauss(rng, n, sigma)
    chans = [l[0] for l in layers]
    conns = [l[1] for l in layers]
    psi = A.PsiCNN(extent, chans, conns, sym, params, final_factor, lp)
    psi.precision = precision
    o = orc.cnn(extent, chans, conns, sym, params, final_factor, lp)
    return psi, o


SHAPES = [
    ([1, 4, 4], [(3, [1, 3, 3]), (2, [1, 3, 3])]),
    ([2, 4, 4], [(2, [2, 4, 4]), (4, [2, 2, 2])]),       # shape of test/conftest.py:41-45
    ([1, 1, 10], [(2, [1, 1, 3]), (3, [1, 1, 2]), (2, [1, 1, 3])]),
    ([1, 12, 12], [(8, [1, 3, 3]), (8, [1, 3, 3])]),     # TFIM 12x12 network
]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", range(len(SHAPES)))
def test_cnn_log_psi_and_O_k(precision, shape):
    extent, layers = SHAPES[shape]
    psi, o = make_cnn(extent, layers, seed=shape, lp=0.2 - 0.1j, precision=precision)
    N = psi.N
    rng = np.random.default_rng(3)
    for _ in range(4):
        idx = 0
        for k in range(0, N, 60):
            idx |= int(rng.integers(0, 2 ** min(60, N - k))) << k
        ref = orc.log_psi_s(o, conf_array(idx))
        got = A.log_psi_s(psi, A.Spins(idx, N))
        assert abs(got - ref) <= TOL[precision] * max(1.0, abs(ref))
        assert rel(A.psi_O_k(psi, A.Spins(idx, N)), orc.psi_O_k(o, conf_array(idx))) < 10 * TOL[precision]


def test_cnn_symmetry_classes():
    """two sublattice classes (checkerboard) on a 4x4 lattice"""
    sym = [(i + j) % 2 for i in range(4) for j in range(4)]
    psi, o = make_cnn([1, 4, 4], [(2, [1, 3, 3]), (2, [1, 2, 2])], seed=11, sym=sym)
    for idx in (0, 0x1234, 0xBEEF):
        assert abs(A.log_psi_s(psi, idx) - orc.log_psi_s(o, conf_array(idx))) < 1e-10
        assert rel(A.psi_O_k(psi, idx), orc.psi_O_k(o, conf_array(idx))) < 1e-10


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_cnn_tdvp_exact(precision):
    """E, <O>, F, S, S.v and both solves on ExactSummation (TFIM 3x4)"""
    psi, o = make_cnn([1, 3, 4], [(3, [1, 3, 3]), (2, [1, 3, 3])], seed=5, precision=precision)
    N = psi.N
    ens = A.ExactSummationSpins(N)
    psi.normalize(ens)
    o.s.log_prefactor = orc._cplx(psi.log_prefactor)
    H = ham.transverse_ising(ham.square_bonds([3, 4]), N, 1.0, 3.04)
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
    shift = 1e-3
    x_ref = np.linalg.solve(ref["S"] + shift * np.eye(psi.num_params), ref["F"])
    assert rel(tdvp.solve(shift, "cholesky"), x_ref) < 1e3 * tol
    assert rel(tdvp.solve(shift, "cg", rtol=1e-10, max_iter=2000), x_ref) < 1e3 * tol


@pytest.mark.parametrize("move", ["flip", "exchange"])
def test_cnn_mc_trajectories_match_oracle(move):
    """fp64 device sampler replays the oracle's Philox stream for PsiCNN"""
    psi, o = make_cnn([1, 4, 4], [(2, [1, 3, 3]), (2, [1, 3, 3])], seed=8, sigma=0.3)
    chains, steps, sweeps, therm = 32, 3, 2, 3
    ens = A.MonteCarloSpins(chains * steps, sweeps, therm, chains, move=move, seed=17)
    A.log_psi(psi, ens)
    got = ens.last_samples()[:, 0]
    ref = orc.mc_run(o, chains, steps, sweeps, therm,
                     move=orc.MOVE_EXCHANGE if move == "exchange" else orc.MOVE_FLIP, seed=17)
    assert np.array_equal(got, ref["samples"][:, 0])
    assert ens.acceptance_rate == pytest.approx(ref["acceptances"] / (ref["acceptances"] + ref["rejections"]))


def test_cnn_tfim12x12_samples_match_oracle():
    """TFIM 12x12, 2 conv layers 3x3 x 8 channels (N = 144, 3 words): E_loc and O_k
    of device MC samples equal the oracle's on the same samples."""
    psi, o = make_cnn([1, 12, 12], [(8, [1, 3, 3]), (8, [1, 3, 3])], seed=4, sigma=0.05, precision="fp32")
    N = psi.N
    H = ham.transverse_ising_square(12, 1.0, 3.04)
    op = A.Operator(H)
    ens = A.MonteCarloSpins(256, 1, 2, 128, seed=6)
    tdvp = A.TDVP(psi.num_params)
    tdvp.eval(op, psi, ens)
    confs = ens.last_samples()
    assert confs.shape == (256, 3)
    w = np.full(256, 1 / 256)
    ref = orc.tdvp_accumulate(oracle_operator(H, 3), o, confs, w, with_S=False)
    el = A.local_energies(op, psi, ens)      # fresh samples from the continued stream
    assert el.shape == (256,)
    assert abs(tdvp.E_local - ref["E"]) < 1e-5 * abs(ref["E"])
    assert rel(tdvp.O_k_vector, ref["O"]) < 1e-4
    assert rel(tdvp.F_vector, ref["F"]) < 1e-4
    v = cgauss(np.random.default_rng(2), psi.num_params, 1.0)
    Oc = ref["O_samples"]
    Sv = (Oc.conj().T @ (w * (Oc @ v))) - ref["O"].conj() * (ref["O"] @ v)
    assert rel(tdvp.S_dot_vector(v), Sv) < 1e-4


def _oracle_rk4(spec_fn, p0, H, N, dt, steps, shift):
    """RK4 of d p / dt = -i (S + shift)^-1 F with oracle S and F on normalised exact weights"""
    oop = oracle_operator(H, 1)
    confs = orc.exact_confs(N, 1)

    def rhs(p):
        o = spec_fn(p)
        lp = orc.exact_log_psi_vector(o, N)
        w = np.exp(2 * (lp.real - lp.real.max()))
        w /= w.sum()
        ref = orc.tdvp_accumulate(oop, o, confs, w)
        return -1j * np.linalg.solve(ref["S"] + shift * np.eye(p.size), ref["F"])

    p = p0.copy()
    out = []
    for _ in range(steps):
        k1 = rhs(p)
        k2 = rhs(p + 0.5 * dt * k1)
        k3 = rhs(p + 0.5 * dt * k2)
        k4 = rhs(p + dt * k3)
        p = p + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        out.append(p.copy())
    return out


def test_rk4_quench_matches_oracle_rk4():
    """real-time RK4 of a PsiCNN after a quench (TFIM 3x3, h = 3.04) with exact sums:
    the device integrator follows the oracle's RK4 parameter trajectory"""
    extent, layers = [1, 3, 3], [(2, [1, 3, 3]), (2, [1, 2, 2])]
    psi, o = make_cnn(extent, layers, seed=31, sigma=0.05, final_factor=1.0)
    N = psi.N
    H = ham.transverse_ising(ham.square_bonds([3, 3]), N, 1.0, 3.04)
    p0 = psi.params.copy()
    sym = np.zeros(N, dtype=np.uint32)

    def spec_fn(p):
        return orc.cnn(extent, [l[0] for l in layers], [l[1] for l in layers], sym, p, 1.0, 0.0)

    dt, steps, shift = 0.01, 3, 1e-3
    ref = _oracle_rk4(spec_fn, p0, H, N, dt, steps, shift)
    rk = A.RK4(psi, A.Operator(H), A.ExactSummationSpins(N), dt, shift=shift, method="cholesky")
    for k in range(steps):
        p = rk.step()
        assert rel(p, ref[k]) < 1e-9
    assert rk.evaluations == 4 * steps
    # energy at the start of each step is (nearly) conserved by real-time TDVP
    E = np.array(rk.energies).real[::4]
    assert np.ptp(E) < 5e-3 * abs(E[0])


@pytest.mark.parametrize("move", ["flip", "exchange"])
def test_cnn_local_update_sampler_matches_full_forward(move):
    """the local-update sampler (receptive fields only) against exact sums: a 4x4 TFIM
    CNN with a 3x3 and a 2x2 layer, energies within 4 standard errors of exact
    summation (flip: all states; exchange: the S_z = 0 sector)"""
    from itertools import combinations
    psi, o = make_cnn([1, 4, 4], [(3, [1, 3, 3]), (2, [1, 2, 2])], seed=21, sigma=0.2, precision="fp32")
    N = 16
    H = ham.transverse_ising_square(4, 1.0, 1.5)
    op = A.Operator(H)
    if move == "flip":
        confs = orc.exact_confs(N, 1)
    else:
        confs = np.array([[sum(1 << i for i in c)] for c in combinations(range(N), N // 2)], dtype=np.uint64)
    lp = np.array([orc.log_psi_s(o, orc.conf_words(int(x), 4)) for x in confs[:, 0]])
    wt = np.exp(2 * (lp.real - lp.real.max()))
    wt /= wt.sum()
    E_ex = orc.tdvp_accumulate(oracle_operator(H, 1), o, confs, wt, with_S=False)["E"].real
    chains, steps = 1024, 64
    ens = A.MonteCarloSpins(chains * steps, 1, 20, chains, move=move, seed=13)
    el = A.local_energies(op, psi, ens).real.reshape(steps, chains)
    se = el.mean(axis=0).std(ddof=1) / np.sqrt(chains)
    assert abs(el.mean() - E_ex) < 4 * se, (el.mean(), E_ex, se)
    # the local-update E_loc of these samples equals the oracle's
    s = ens.last_samples()[:512]
    ref = orc.tdvp_accumulate(oracle_operator(H, 1), o, s, np.full(len(s), 1.0 / len(s)), with_S=False)
    got = A.local_energies(op, psi, A.SampleList(N, s))
    assert rel(got, ref["E_loc"]) < 1e-5


def test_cnn_tfim12x12_rk4_quench_at_size():
    """the quench workload at its size: TFIM 12x12 (h = 3.04) from the near-uniform state
    (all CNN weights ~1e-3, the h = infinity ground state up to that noise), 2 layers of
    3x3 kernels with 8 channels (648 parameters), 16,384 single-flip chains, RK4 with
    dt = 0.005 for 3 steps and the fp64 Cholesky solve of S + 1e-3.  Real-time TDVP
    conserves the energy: E stays within 5 standard errors of its start; the parameters
    stay finite and every solve is exact to 1e-8."""
    psi, o = make_cnn([1, 12, 12], [(8, [1, 3, 3]), (8, [1, 3, 3])], seed=44, sigma=1e-3, final_factor=1.0,
                      precision="fp32")
    assert psi.num_params == 648
    H = ham.transverse_ising_square(12, 1.0, 3.04)
    op = A.Operator(H)
    chains = 16384
    ens = A.MonteCarloSpins(chains, 1, 50, chains, move="flip", seed=3, persistent=True)
    rk = A.RK4(psi, op, ens, 0.005, shift=1e-3, method="cholesky")
    residuals = []
    for _ in range(3):
        p = rk.step()
        residuals.append(rk.tdvp.last_solve["residual"])
        assert np.all(np.isfinite(p))
    assert max(residuals) < 1e-8
    E = np.array(rk.energies).real
    el = A.local_energies(op, psi, ens).real
    se = el.std() / np.sqrt(chains)
    # the h = infinity state has E = -h N exactly
    assert abs(E[0] + 3.04 * 144) < 5 * se + 1e-2
    assert np.ptp(E) < 10 * se, (E, se)
