"""Network constructors with the reference's signatures and initialisation
(pyANNonGPU/new_RBM.py:13-29, new_neural_network.py:31-131,
new_convolutional_network.py:31-82), plus seeded complex-Gaussian builders used
for the benchmark configurations.

The reference draws noise from numpy's global RandomState; the same calls are
made here, so seeding ``np.random.seed`` reproduces its networks.  An optional
``rng`` (numpy Generator) can be passed instead.
"""
import math

import numpy as np

from .api import PsiRBM, PsiDeep, PsiCNN


def _real_noise(shape, rng=None):
    r = rng.random(shape) if rng is not None else np.random.random_sample(shape)
    return 2 * r - 1


def _complex_noise(shape, rng=None):
    return _real_noise(shape, rng) + 1j * _real_noise(shape, rng)


def _noise_vector(shape, real=True, rng=None):
    return _real_noise(shape, rng) if real else _complex_noise(shape, rng)


def _prod(xs):
    r = 1
    for x in xs:
        r *= x
    return r


def new_RBM(N, M, initial_value=(0.01 + 1j * math.pi / 4), noise=1e-4, gpu=True, final_weight=10, rng=None):
    """new_RBM.py:13-29"""
    assert M >= N
    W = noise * _complex_noise((N, M), rng)
    for alpha in range(M // N):
        W[:, alpha * N:(alpha + 1) * N] += initial_value * np.eye(N)
    return PsiRBM(W, final_weight, 0.0, gpu)


def _delta(n, m, c):
    if m > n:
        return 1
    if n % m == 0:
        return n // m
    return c


def deep_connections(N, M, C):
    """Gather tables of new_neural_network.py:85-104 (1d): unit j of a layer reads
    (j * delta + i) % n for i < c."""
    out = []
    for n, m, c in zip([N] + list(M[:-1]), M, C):
        d = _delta(n, m, c)
        out.append(np.array([[(j * d + i) % n for j in range(m)] for i in range(c)], dtype=np.uint32))
    return out


def new_deep_neural_network(num_sites, N, M, C, initial_value=(0.01 + 1j * math.pi / 4), a=0, noise=1e-4,
                            gpu=True, noise_modulation="auto", final_weights=10, rng=None):
    """new_neural_network.py:31-131 (one-dimensional layouts)."""
    if isinstance(N, (list, tuple)):
        raise NotImplementedError("multi-dimensional PsiDeep layouts are not built")
    for n, m, c in zip([N] + M[:-1], M, C):
        assert (m * c) % n == 0
        assert c <= n
    a = a * np.ones(N) if isinstance(a, (float, int, complex)) else np.array(a)
    is_real = (complex(initial_value).imag == 0)
    b = [noise * _noise_vector(m, is_real, rng) for m in M]
    w = noise * _noise_vector((C[0], M[0]), is_real, rng)
    w[C[0] // 2, :] += initial_value
    W = [w]
    if noise_modulation == "auto":
        noise_modulation = [math.sqrt(6 / (c + next_c)) for c, next_c in zip(C[1:], C[2:] + [1])]
    for c, m, next_c, nm in zip(C[1:], M[1:], C[2:] + [1], noise_modulation):
        W.append(math.sqrt(6 / (c + next_c)) * _real_noise((c, m), rng) + noise * _noise_vector((c, m), is_real, rng))
    connections = deep_connections(N, M, C)
    if isinstance(final_weights, (float, int)):
        final_weights = final_weights * np.ones(M[-1])
    return PsiDeep(num_sites, a, b, connections, W, final_weights, 0.0, gpu)


def new_convolutional_network(L, layers, initial_value=(0.01 + 1j * math.pi / 4), noise=1e-4, final_factor=10,
                              symmetry_classes=None, real=False, gpu=True, rng=None):
    """new_convolutional_network.py:31-82"""
    num_channels_list = np.array([l[0] for l in layers])
    connectivity_list = np.array([l[1] for l in layers])
    if symmetry_classes is None:
        symmetry_classes = np.zeros(_prod(L))
    num_sym = len(set(np.asarray(symmetry_classes).tolist()))
    params = []
    for layer, (num_channels, nd_conn) in enumerate(layers):
        for c, l in zip(nd_conn, L):
            assert c <= l
        connectivity = _prod(nd_conn)
        num_prev = layers[layer - 1][0] if layer > 0 else 1
        for _ in range(num_channels * num_prev):
            for _ in range(num_sym):
                link = noise * _noise_vector(connectivity, real, rng)
                if layer == 0:
                    link[connectivity // 2] = initial_value.real if real else initial_value
                else:
                    link = link + math.sqrt(6 / (connectivity * num_prev + connectivity * num_channels)) * \
                        _real_noise(connectivity, rng)
                params += list(link)
    return PsiCNN(L, num_channels_list, connectivity_list, symmetry_classes, np.array(params), final_factor, 0, gpu)


# -------------------------------------------------------- seeded builders

def complex_gaussian(rng, shape, sigma):
    """re and im independently N(0, sigma^2)"""
    return sigma * rng.standard_normal(shape) + 1j * sigma * rng.standard_normal(shape)


def seeded_rbm_deep(N, M, seed, sigma=0.01, final_weight=1.0, gpu=True):
    """Fully connected one-hidden-layer PsiDeep ("complex RBM with biases") with
    a, b, W drawn complex-Gaussian(sigma) from numpy.random.default_rng(seed),
    connections of new_neural_network.py with C = N."""
    rng = np.random.default_rng(seed)
    a = complex_gaussian(rng, N, sigma)
    b = complex_gaussian(rng, M, sigma)
    W = complex_gaussian(rng, (N, M), sigma)
    conn = deep_connections(N, [M], [N])[0]
    return PsiDeep(N, a, [b], [conn], [W], final_weight * np.ones(M), 0.0, gpu)


def seeded_rbm(N, M, seed, sigma=0.01, final_weight=1.0, gpu=True):
    """PsiRBM (no biases) with complex-Gaussian(sigma) weights."""
    rng = np.random.default_rng(seed)
    return PsiRBM(complex_gaussian(rng, (N, M), sigma), final_weight, 0.0, gpu)


def seeded_cnn(L, layers, seed, sigma=0.01, final_factor=1.0, gpu=True):
    """PsiCNN with every parameter complex-Gaussian(sigma); L and the kernel
    shapes are 1-3 dimensional (padded to 3 as the reference's dim = 3 type)."""
    rng = np.random.default_rng(seed)
    n = 0
    prev = 1
    for ch, conn in layers:
        n += prev * ch * _prod(conn)
        prev = ch
    params = complex_gaussian(rng, n, sigma)
    return PsiCNN(L, [l[0] for l in layers], [l[1] for l in layers], None, params, final_factor, 0.0, gpu)
