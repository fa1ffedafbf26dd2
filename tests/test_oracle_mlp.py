"""The oracle's network of the learned coarse correction (FP64, no reference counterpart).

Pinned by: central finite differences of its own forward pass (exact mode, no BF16 rounding),
the zero-last-layer identity (output = b3 for any input), linearity of the gradient in dy, and
BF16 emulation idempotence (inputs that are already BF16 values are left unchanged).
"""

import numpy as np
import pytest

import oracle as O

P = O.MLP_PARAMS
W1, B1 = 0, 128 * 64
W2, B2 = B1 + 128, B1 + 128 + 128 * 128
W3 = B2 + 128
B3 = W3 + 64 * 128


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(P, dtype=np.float32)
    p[W1:B1] = rng.uniform(-1, 1, B1 - W1) / 8.0 * scale
    p[B1:W2] = rng.uniform(-0.1, 0.1, W2 - B1)
    p[W2:B2] = rng.uniform(-1, 1, B2 - W2) / np.sqrt(128) * scale
    p[B2:W3] = rng.uniform(-0.1, 0.1, W3 - B2)
    p[W3:B3] = rng.uniform(-1, 1, B3 - W3) / np.sqrt(128) * scale
    p[B3:] = rng.uniform(-0.1, 0.1, P - B3)
    return p


def test_zero_last_layer_gives_bias():
    p = random_params(1)
    p[W3:B3] = 0.0
    x = np.random.default_rng(2).standard_normal((5, 64)).astype(np.float32)
    y = O.mlp_forward(p, x, emulate_bf16=False)
    assert np.array_equal(y, np.tile(p[B3:].astype(np.float64), (5, 1)))


def test_gradient_matches_central_differences():
    # FP64 exact mode; perturb the float32 parameters through a float64 copy via scale tricks:
    # the oracle takes float32 parameters, so difference quotients use steps that are exact in
    # float32 around the chosen values (h = 2^-12 relative to O(1) entries)
    p = random_params(3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((7, 64)).astype(np.float32)
    dy = rng.standard_normal((7, 64)).astype(np.float32)
    g = O.mlp_grad(p, x, dy, emulate_bf16=False)
    idx = np.concatenate([rng.choice(np.arange(W1, B1), 6, replace=False), rng.choice(np.arange(B1, W2), 3, replace=False),
                          rng.choice(np.arange(W2, B2), 6, replace=False), rng.choice(np.arange(B2, W3), 3, replace=False),
                          rng.choice(np.arange(W3, B3), 6, replace=False), rng.choice(np.arange(B3, P), 3, replace=False)])

    def L(pp):
        return float(np.sum(dy.astype(np.float64) * O.mlp_forward(pp, x, emulate_bf16=False)))

    h = 2.0 ** -12
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += np.float32(h)
        pm[i] -= np.float32(h)
        fd = (L(pp) - L(pm)) / (float(pp[i]) - float(pm[i]))
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_gradient_is_linear_in_dy():
    p = random_params(5)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((4, 64)).astype(np.float32)
    a = rng.standard_normal((4, 64)).astype(np.float32)
    b = rng.standard_normal((4, 64)).astype(np.float32)
    ga = O.mlp_grad(p, x, a, emulate_bf16=True)
    gb = O.mlp_grad(p, x, b, emulate_bf16=True)
    gab = O.mlp_grad(p, x, (a + b).astype(np.float32), emulate_bf16=True)
    # a + b is rounded to float32 once; the tolerance covers that rounding
    assert np.abs(gab - (ga + gb)).max() <= 1e-5 * np.abs(ga + gb).max()


def test_bf16_emulation_leaves_bf16_values_alone():
    rng = np.random.default_rng(7)
    p = random_params(8)
    # round parameters and inputs to BF16 values: emulation must then equal exact mode up to
    # the rounding of the hidden activations only
    pb = (p.view(np.uint32) & np.uint32(0xFFFF0000)).view(np.float32)
    x = (rng.standard_normal((3, 64)).astype(np.float32).view(np.uint32) & np.uint32(0xFFFF0000)).view(np.float32)
    ye = O.mlp_forward(pb, x, emulate_bf16=True)
    yx = O.mlp_forward(pb, x, emulate_bf16=False)
    assert np.abs(ye - yx).max() <= 2e-2 * np.abs(yx).max()
    assert not np.array_equal(ye, yx)  # hidden activations are rounded in emulation
