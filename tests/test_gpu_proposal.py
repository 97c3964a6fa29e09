"""Learned block proposal on the HIP path (needs an MI355X).

* forward/backward of the proposal network and the Adam step against the fp64
  CPU oracle (fp32 device arithmetic: 1e-5 relative);
* the Metropolis-Hastings block sampler leaves |psi|^2 invariant: energies and a
  correlation function within 4 standard errors of exact sums over the S_z = 0
  sector, with a deliberately non-uniform (untrained) and a trained network;
* what training gains: on an SR-optimised Heisenberg ring, the trained network's
  block moves are accepted far more often and decorrelate the local energy in
  fewer sweeps than the untrained network's.
"""
import math
from itertools import combinations

import numpy as np
import pytest

import pyoracle as orc
from helpers import oracle_operator
import annongpu_amd as A
from annongpu_amd import hamiltonians as ham

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _rows(N, count, bs, seed):
    rng = np.random.default_rng(seed)
    words = (N + 63) // 64
    confs = rng.integers(0, 2 ** 62, (count, words), dtype=np.int64).astype(np.uint64)
    if N % 64:
        confs[:, -1] &= np.uint64((1 << (N % 64)) - 1)
    sites = np.array([rng.choice(N, bs, replace=False) for _ in range(count)], dtype=np.int32)
    return confs, sites


@pytest.mark.parametrize("N,H,bs,count", [(20, 32, 4, 64), (100, 256, 8, 300), (130, 96, 6, 70)])
def test_loss_grad_matches_oracle(N, H, bs, count):
    blocks = A.chain_blocks(N, bs)
    prop = A.BlockProposal(N, blocks, hidden=H, seed=3)
    rng = np.random.default_rng(N)
    params = (0.3 * rng.standard_normal(prop.num_params)).astype(np.float32)
    prop.params = params
    assert np.array_equal(prop.params, params)
    confs, sites = _rows(N, count, bs, seed=N + 1)
    loss, grad = prop.loss_grad(confs, sites)
    loss_ref, grad_ref = orc.mlp_loss_grad(N, H, params.astype(np.float64), confs, sites)
    assert loss == pytest.approx(loss_ref, rel=1e-5)
    assert rel(grad, grad_ref) < 1e-5
    # each block of the gradient separately (W1, b1, W2, b2)
    cuts = np.cumsum([0, H * N, H, N * H, N])
    for a, b in zip(cuts[:-1], cuts[1:]):
        assert rel(grad[a:b], grad_ref[a:b]) < 1e-4


def test_adam_matches_oracle():
    N, H = 24, 64
    prop = A.BlockProposal(N, A.chain_blocks(N, 4), hidden=H, seed=1)
    p0 = prop.params.astype(np.float64)
    m = np.zeros_like(p0)
    v = np.zeros_like(p0)
    p = p0.copy()
    rng = np.random.default_rng(2)
    for t in (1, 2, 3):
        g = rng.standard_normal(p.size).astype(np.float32)
        prop.adam_step(g, 1e-2)
        orc.adam_step(p, m, v, g.astype(np.float64), 1e-2, t)
    assert np.max(np.abs(prop.params - p)) < 1e-6


def _sector_exact(psi_oracle, N, H, words=1):
    """exact |psi|^2-weighted E and <s_0 s_1> over the S_z = 0 sector"""
    confs = np.array([sum(1 << i for i in c) for c in combinations(range(N), N // 2)], dtype=np.uint64)
    lp = np.array([orc.log_psi_s(psi_oracle, orc.conf_words(int(x), 4)) for x in confs])
    w = np.exp(2 * (lp.real - lp.real.max()))
    w /= w.sum()
    ref = orc.tdvp_accumulate(oracle_operator(H, words), psi_oracle, confs[:, None], w, with_S=False)
    s0s1 = np.array([(1 if x & 1 else -1) * (1 if x & 2 else -1) for x in confs.tolist()])
    return ref["E"].real, float(w @ s0s1)


def _chain_stats(values, chains):
    """mean and standard error from per-chain means (chains are independent)"""
    v = np.asarray(values).reshape(-1, chains)
    cm = v.mean(axis=0)
    return cm.mean(), cm.std(ddof=1) / math.sqrt(chains)


@pytest.mark.parametrize("precision,scale,exclusive", [("fp64", 3.0, False), ("fp32", 3.0, False),
                                                       ("fp32", 0.0, False), ("fp32", None, False),
                                                       ("fp64", 3.0, True), ("fp32", 0.0, True)])
def test_block_sampler_preserves_psi_squared(precision, scale, exclusive):
    """block moves alone (no plain sweeps after plain thermalisation) sample |psi|^2
    restricted to S_z = 0; scale 0 is the uniform proposal (proposal ratio 1) and
    scale None the plain exchange sampler (the control)"""
    N = 12
    psi = A.seeded_rbm_deep(N, 24, 21, sigma=0.1)
    psi.precision = precision
    o = orc.deep(N, psi.a, psi.b, psi.connections, psi.W, psi.final_weights, 0.0)
    H = ham.heisenberg_chain(N)
    E_ex, c_ex = _sector_exact(o, N, H)
    chains, steps = 2048, 48
    ens = A.MonteCarloSpins(chains * steps, 2, 10, chains, move="exchange", seed=8)
    if scale is not None:
        prop = A.BlockProposal(N, A.chain_blocks(N, 4), hidden=64, seed=5)
        # a strongly non-uniform network: a wrong proposal ratio would bias the result
        prop.params = scale * prop.params
        ens.set_proposal(prop, plain_sweeps=0, exclude_current=exclusive)
    el = A.local_energies(A.Operator(H), psi, ens).real
    s = ens.last_samples()[:, 0].astype(np.int64)
    assert np.all([bin(int(x)).count("1") == N // 2 for x in s[:64]])
    E, se = _chain_stats(el, chains)
    assert abs(E - E_ex) < 4 * se + 1e-9, (E, E_ex, se)
    s0s1 = np.where((s & 1) == (s >> 1 & 1), 1.0, -1.0)
    c, sec = _chain_stats(s0s1, chains)
    assert abs(c - c_ex) < 4 * sec, (c, c_ex, sec)
    assert 0.0 < ens.acceptance_rate < 1.0


@pytest.mark.parametrize("exclusive", [False, True])
@pytest.mark.parametrize("precision,scale", [("fp32", 1.0), ("fp32", 3.0), ("fp64", 3.0), ("fp32", "trained")])
def test_plaquette_block_sampler_preserves_psi_squared(precision, scale, exclusive):
    """the benchmark's block shape: 2x4 and 4x2 plaquettes (8 sites, so every lane of the
    pattern draw holds patterns and the shared high-bit energies span bits 2..7) with the
    benchmark's 256 hidden units, on a 4x4 Heisenberg torus.  E and <s_0 s_1> within 4
    standard errors of exact sums over the 12,870 states of the S_z = 0 sector, for the
    network as initialised, scaled to be strongly non-uniform, and trained on plain
    samples; redraws of the current pattern are counted apart from accepted moves.
    exclusive: the Metropolized Gibbs form (the current pattern is never proposed)"""
    L, N = 4, 16
    psi = A.seeded_rbm_deep(N, 32, 21, sigma=0.1)   # participation ratio ~1,500 of 12,870 states
    psi.precision = precision
    o = orc.deep(N, psi.a, psi.b, psi.connections, psi.W, psi.final_weights, 0.0)
    H = ham.heisenberg_square(L)
    E_ex, c_ex = _sector_exact(o, N, H)
    blocks = A.rect_blocks(L, (2, 4))
    assert blocks.shape == (32, 8) and all(len(set(b)) == 8 for b in blocks.tolist())
    prop = A.BlockProposal(N, blocks, hidden=256, seed=5)
    if scale == "trained":
        data = A.MonteCarloSpins(1024 * 32, 2, 20, 1024, move="exchange", seed=2)
        A.log_psi(psi, data)
        prop.fit(data, steps=300, batch=2048, lr=3e-3)
    else:
        prop.params = scale * prop.params
    chains, steps = 2048, 48
    ens = A.MonteCarloSpins(chains * steps, 2, 10, chains, move="exchange", seed=8)
    ens.set_proposal(prop, plain_sweeps=0, exclude_current=exclusive)
    el = A.local_energies(A.Operator(H), psi, ens).real
    s = ens.last_samples()[:, 0].astype(np.int64)
    assert np.all([bin(int(x)).count("1") == N // 2 for x in s[:256]])
    E, se = _chain_stats(el, chains)
    assert abs(E - E_ex) < 4 * se + 1e-9, (E, E_ex, se)
    s0s1 = np.where((s & 1) == (s >> 1 & 1), 1.0, -1.0)
    c, sec = _chain_stats(s0s1, chains)
    assert abs(c - c_ex) < 4 * sec, (c, c_ex, sec)
    st = ens.block_stats
    assert st["proposals"] == chains * steps * 2 * 2     # 2 block sweeps of ceil(16 / 8) moves per sample
    assert 0 <= st["null_moves"] < st["proposals"]
    if exclusive:   # only blocks whose spins are all equal have no other pattern
        assert st["null_moves"] < 0.05 * st["proposals"]
    moved = ens.acceptance_rate * st["proposals"]
    assert 0.0 < ens.acceptance_rate < 1.0
    # every evaluated proposal flips at least 2 sites (S_z is conserved) and at most 8
    evaluated = st["proposals"] - st["null_moves"]
    assert 2 * evaluated <= st["rows_evaluated"] <= 8 * evaluated
    assert moved <= evaluated
    print(f"\n{precision} scale {scale} exclusive {exclusive}: E {E:.4f} +- {se:.4f} (exact {E_ex:.4f}), acceptance "
          f"{ens.acceptance_rate:.3f}, null redraws {st['null_moves'] / st['proposals']:.3f}, "
          f"rows per evaluated move {st['rows_evaluated'] / max(evaluated, 1):.2f}")


@pytest.mark.parametrize("sampler", ["plain", "learned"])
def test_persistent_chains_do_not_drift(sampler):
    """persistent chains run for about 27,000 sweeps: the energy of the first window (10
    thermalisation sweeps) and of the last 8,000 sweeps both within 4 standard errors of the exact
    S_z = 0 sum, for the plain exchange sampler and for the trained 2x4-plaquette proposal
    (Metropolized Gibbs form).  A sampler whose error grows with the length of a run (a
    random stream reused across calls, drifting angles) would fail the last window."""
    L, N = 4, 16
    psi = A.seeded_rbm_deep(N, 32, 21, sigma=0.1)
    psi.precision = "fp32"
    o = orc.deep(N, psi.a, psi.b, psi.connections, psi.W, psi.final_weights, 0.0)
    H = ham.heisenberg_square(L)
    E_ex, _ = _sector_exact(o, N, H)
    op = A.Operator(H)
    chains, steps = 2048, 40
    ens = A.MonteCarloSpins(chains * steps, 2, 10, chains, move="exchange", seed=12, persistent=True)
    if sampler == "learned":
        prop = A.BlockProposal(N, A.rect_blocks(L, (2, 4)), hidden=256, seed=5)
        data = A.MonteCarloSpins(1024 * 32, 2, 20, 1024, move="exchange", seed=2)
        A.log_psi(psi, data)
        prop.fit(data, steps=300, batch=2048, lr=3e-3)
        ens.set_proposal(prop, plain_sweeps=0, exclude_current=True)
    early = A.local_energies(op, psi, ens).real
    E0, se0 = _chain_stats(early, chains)
    for _ in range(240):   # 80 sweeps per call
        A.local_energies(op, psi, ens)
    late = np.concatenate([A.local_energies(op, psi, ens).real for _ in range(100)])
    E1, se1 = _chain_stats(late, chains)
    print(f"\n{sampler}: exact {E_ex:.5f}; sweeps 10-90 {E0:.5f} +- {se0:.5f}; "
          f"sweeps 19,290-27,290 {E1:.5f} +- {se1:.5f}")
    assert abs(E0 - E_ex) < 4 * se0, (E0, E_ex, se0)
    assert abs(E1 - E_ex) < 4 * se1, (E1, E_ex, se1)


def _optimise_ring(N=12, M=24, steps=50, lr=0.02):
    """imaginary-time SR on the Heisenberg ring with exact sums (the polynomial
    activations diverge if the weights grow much further than 50 such steps take them)"""
    psi = A.seeded_rbm_deep(N, M, 4, sigma=0.05)
    psi.precision = "fp64"
    op = A.Operator(ham.heisenberg_chain(N))
    ex = A.ExactSummationSpins(N)
    tdvp = A.TDVP(psi.num_params)
    for _ in range(steps):
        psi.normalize(ex)
        tdvp.eval(op, psi, ex)
        x = tdvp.solve(1e-3, "cholesky")
        psi.params = psi.params - lr * x
    psi.normalize(ex)
    tdvp.eval(op, psi, ex)
    return psi, op, tdvp.E_local.real


def _tau_int(series):
    """integrated autocorrelation time of a [T][chains] series (Sokal window, c = 5)"""
    x = series - series.mean(axis=0)
    T = x.shape[0]
    var = (x * x).mean()
    tau = 0.5
    for t in range(1, T // 2):
        rho = (x[:-t] * x[t:]).mean() / var
        tau += rho
        if t >= 5 * tau:
            break
    return tau


def _staggered_sq(confs, N):
    """squared staggered magnetisation of a ring, per sample"""
    bits = (confs[..., None] >> np.arange(N, dtype=np.uint64)) & np.uint64(1)
    s = 2.0 * bits.astype(np.float64) - 1.0
    return (s * (-1.0) ** np.arange(N)).sum(axis=-1) ** 2


def test_trained_proposal_gains_over_untrained():
    """12-site Heisenberg ring after 50 exact SR steps; blocks of 6 consecutive sites.
    Trained on plain-sampler configurations, the network's block moves are accepted
    more often and decorrelate the staggered magnetisation and E_loc faster (per
    block sweep) than the same network untrained; energies agree."""
    N = 12
    psi, op, E_opt = _optimise_ring(N)
    assert E_opt < -15.0      # ground state of the 12-site ring: -21.55 (Pauli units)
    chains, T = 1024, 200
    # training data: plain exchange samples
    data = A.MonteCarloSpins(chains * 64, 2, 20, chains, move="exchange", seed=1)
    A.log_psi(psi, data)
    blocks = A.chain_blocks(N, 6)
    trained = A.BlockProposal(N, blocks, hidden=64, seed=2)
    untrained = A.BlockProposal(N, blocks, hidden=64, seed=2)
    rows = np.arange(4096)
    loss0, _ = trained.loss_grad(data.last_samples()[:4096], blocks[rows % N])
    loss1 = trained.fit(data, steps=400, batch=2048, lr=3e-3)
    assert loss1 < 0.9 * loss0
    res = {}
    for name, prop in (("trained", trained), ("untrained", untrained)):
        ens = A.MonteCarloSpins(chains * T, 1, 5, chains, move="exchange", seed=9)
        ens.set_proposal(prop, plain_sweeps=0)
        el = A.local_energies(op, psi, ens).real.reshape(T, chains)
        stag = _staggered_sq(ens.last_samples()[:, 0].reshape(T, chains), N)
        res[name] = dict(acc=ens.acceptance_rate, tau_E=_tau_int(el), tau_stag=_tau_int(stag), E=el.mean())
    t, u = res["trained"], res["untrained"]
    assert t["acc"] > 1.5 * u["acc"], res
    assert t["tau_stag"] < 0.7 * u["tau_stag"], res
    assert t["tau_E"] < u["tau_E"], res
    assert abs(t["E"] - u["E"]) < 0.01 * abs(u["E"]), res
