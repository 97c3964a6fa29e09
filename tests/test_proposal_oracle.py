"""Pins the CPU oracle of the learned block proposal (oracle/proposal_oracle.c)
with finite differences, closed forms and exact enumeration; no GPU needed."""
import math
from itertools import combinations

import numpy as np
import pytest

import pyoracle as orc


def _random_problem(N=7, H=5, count=4, bs=3, seed=0):
    rng = np.random.default_rng(seed)
    params = 0.4 * rng.standard_normal(orc.mlp_num_params(N, H))
    confs = rng.integers(0, 2 ** N, count).astype(np.uint64)
    blocks = np.array([rng.choice(N, bs, replace=False) for _ in range(count)], dtype=np.int32)
    return params, confs, blocks


def test_loss_grad_matches_finite_differences():
    N, H = 7, 5
    params, confs, blocks = _random_problem(N, H)
    loss, grad = orc.mlp_loss_grad(N, H, params, confs, blocks)
    h = 1e-6
    fd = np.zeros_like(params)
    for k in range(params.size):
        p = params.copy()
        p[k] += h
        lp, _ = orc.mlp_loss_grad(N, H, p, confs, blocks)
        p[k] -= 2 * h
        lm, _ = orc.mlp_loss_grad(N, H, p, confs, blocks)
        fd[k] = (lp - lm) / (2 * h)
    assert np.max(np.abs(fd - grad)) < 1e-8 * max(1.0, np.max(np.abs(grad)))
    assert np.count_nonzero(grad) > params.size // 2


def test_loss_at_zero_parameters_is_bs_log2():
    N, H, bs = 9, 4, 4
    _, confs, blocks = _random_problem(N, H, count=6, bs=bs, seed=1)
    loss, grad = orc.mlp_loss_grad(N, H, np.zeros(orc.mlp_num_params(N, H)), confs, blocks)
    assert loss == pytest.approx(bs * math.log(2), rel=1e-14)
    # d loss / d b2_o = mean over samples of [o in block] (1/2 - y_o)
    gb2 = grad[-N:]
    ref = np.zeros(N)
    for c, blk in zip(confs, blocks):
        for o in blk:
            ref[o] += (0.5 - ((int(c) >> int(o)) & 1)) / len(confs)
    assert np.allclose(gb2, ref, atol=1e-15)


def test_logits_closed_form_and_block_masking():
    N, H = 6, 3
    rng = np.random.default_rng(2)
    params = rng.standard_normal(orc.mlp_num_params(N, H))
    W1 = params[:H * N].reshape(H, N)
    b1 = params[H * N:H * N + H]
    W2 = params[H * N + H:H * N + H + N * H].reshape(N, H)
    b2 = params[-N:]
    block = np.array([1, 4], dtype=np.int32)
    for conf in (0b101101, 0b010010, 0b111111):
        x = np.array([1.0 if conf >> i & 1 else -1.0 for i in range(N)])
        x[block] = 0.0
        ref = W2 @ np.tanh(W1 @ x + b1) + b2
        got = orc.mlp_logits(N, H, params, np.array([conf], dtype=np.uint64), block)
        assert np.allclose(got, ref, atol=1e-14)
        # the block's own spins do not enter
        other = conf ^ (1 << 1) ^ (1 << 4)
        assert np.array_equal(orc.mlp_logits(N, H, params, np.array([other], dtype=np.uint64), block), got)


def test_adam_closed_form():
    rng = np.random.default_rng(3)
    g1, g2 = rng.standard_normal(10), rng.standard_normal(10)
    p = rng.standard_normal(10)
    p0 = p.copy()
    m, v = np.zeros(10), np.zeros(10)
    lr, b1, b2, eps = 1e-2, 0.9, 0.999, 1e-8
    orc.adam_step(p, m, v, g1, lr, 1, b1, b2, eps)
    # first step: m_hat = g, v_hat = g^2
    assert np.allclose(p, p0 - lr * g1 / (np.abs(g1) + eps), atol=1e-15)
    p1 = p.copy()
    orc.adam_step(p, m, v, g2, lr, 2, b1, b2, eps)
    mh = (b1 * (1 - b1) * g1 + (1 - b1) * g2) / (1 - b1 ** 2)
    vh = (b2 * (1 - b2) * g1 ** 2 + (1 - b2) * g2 ** 2) / (1 - b2 ** 2)
    assert np.allclose(p, p1 - lr * mh / (np.sqrt(vh) + eps), atol=1e-14)


def test_block_proposal_is_a_distribution():
    rng = np.random.default_rng(4)
    bs = 8
    logits = 2.0 * rng.standard_normal(bs)
    for k in range(bs + 1):
        pats = [sum(1 << j for j in c) for c in combinations(range(bs), k)]
        q = np.array([math.exp(orc.block_log_q(logits, p, k)) for p in pats])
        assert q.sum() == pytest.approx(1.0, rel=1e-12)
        e = np.array([sum(logits[j] for j in range(bs) if p >> j & 1) for p in pats])
        assert np.allclose(q, np.exp(e) / np.exp(e).sum(), rtol=1e-12)
    # equal logits: uniform over the C(8, 4) = 70 patterns
    assert orc.block_log_q(np.full(bs, 0.7), 0b00001111, 4) == pytest.approx(-math.log(70), rel=1e-12)
    # two sites, one up: logistic in the logit difference
    assert math.exp(orc.block_log_q(np.array([0.3, -1.1]), 0b01, 1)) == \
        pytest.approx(1 / (1 + math.exp(-1.1 - 0.3)), rel=1e-12)
