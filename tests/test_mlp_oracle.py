"""The surrogate MLP's CPU oracle (oracle/mlp_oracle.py), pinned without a GPU:
its exact gradient against central finite differences, its bfloat16 rounding
against hand-worked values, the BF16-emulating gradient against the exact one,
and one Adam step against the closed form of a first step."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as M  # noqa: E402


def _params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    W1 = np.zeros((M.N_H, M.N_INP))
    W1[:, :M.N_IN] = rng.uniform(-1, 1, (M.N_H, M.N_IN)) * np.sqrt(6 / 134) * scale
    W2 = rng.uniform(-1, 1, (M.N_H, M.N_H)) * np.sqrt(6 / 256) * scale
    W3 = np.zeros((M.N_OUTP, M.N_H))
    W3[:2] = rng.uniform(-1, 1, (2, M.N_H)) * np.sqrt(6 / 130) * scale
    b1 = rng.uniform(-0.1, 0.1, M.N_H)
    b2 = rng.uniform(-0.1, 0.1, M.N_H)
    b3 = np.zeros(M.N_OUTP)
    b3[:2] = rng.uniform(-0.1, 0.1, 2)
    return M.pack(W1, b1, W2, b2, W3, b3)


def _data(n, seed=1):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, M.N_IN))
    t = np.stack([np.sin(2 * x[:, 0]) + x[:, 1] ** 2, x[:, 2] * x[:, 3] - 0.3 * x[:, 5]], axis=1)
    return x, t


def test_layout_size():
    assert M.NPARAM == 20752
    p = _params()
    assert np.array_equal(M.pack(*M.unpack(p)), p)


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -7 + 2 ** -8, -2.5, 3.0e-3, 0.0])
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1.
    # 1 + 3*2^-9 is above halfway -> 1 + 2^-7.  1 + 2^-7 + 2^-8 ties to even -> 1 + 2^-6.
    exp = np.array([1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0030059814453125, 0.0])
    assert np.array_equal(M.bf16(vals), exp)


@pytest.mark.parametrize("seed", [0, 3])
def test_exact_gradient_matches_finite_differences(seed):
    p = _params(seed)
    x, t = _data(37, seed)
    ow = (1.0, 0.5)
    _, g = M.loss_and_grad(p, x, t, ow, emulate_bf16=False)
    rng = np.random.default_rng(seed)
    idx = np.concatenate([
        rng.integers(M.OFF_W1, M.OFF_B1, 8), rng.integers(M.OFF_B1, M.OFF_W2, 4),
        rng.integers(M.OFF_W2, M.OFF_B2, 8), rng.integers(M.OFF_B2, M.OFF_W3, 4),
        M.OFF_W3 + rng.integers(0, 256, 6), [M.OFF_B3, M.OFF_B3 + 1],
    ])
    idx = [i for i in idx if not (i < M.OFF_B1 and (i % M.N_INP) >= M.N_IN)]
    fd = M.finite_difference(p, x, t, idx, h=1e-6, out_weight=ow)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-6, atol=1e-9)
    # padding receives no gradient
    W1g, _, _, _, W3g, b3g = M.unpack(g)
    assert not W1g[:, M.N_IN:].any() and not W3g[2:].any() and not b3g[2:].any()


def test_emulated_gradient_close_to_exact():
    p = _params(5)
    x, t = _data(128, 5)
    le, ge = M.loss_and_grad(p, x, t, emulate_bf16=True)
    lx, gx = M.loss_and_grad(p, x, t, emulate_bf16=False)
    assert abs(le - lx) <= 2e-2 * lx
    for a, b in zip(M.unpack(ge), M.unpack(gx)):
        assert np.linalg.norm(a - b) <= 3e-2 * np.linalg.norm(b) + 1e-12


def test_descent_direction():
    p = _params(2)
    x, t = _data(64, 2)
    l0, g = M.loss_and_grad(p, x, t, emulate_bf16=False)
    l1, _ = M.loss_and_grad(p - 1e-3 * g / np.linalg.norm(g), x, t, emulate_bf16=False)
    assert l1 < l0


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(0)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    p1, m1, v1 = M.adam_step(p, np.zeros(50), np.zeros(50), g, 1, 1e-2, eps=1e-8)
    # first step: m_hat = g, v_hat = g^2, so the step is lr * g / (|g| + eps)
    np.testing.assert_allclose(p1, p - 1e-2 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    np.testing.assert_allclose(m1, 0.1 * g)
    np.testing.assert_allclose(v1, 0.001 * g * g)


def test_split_bf16_carries_16_bits():
    rng = np.random.default_rng(4)
    a = rng.uniform(-1, 1, 10000)
    s = M.split_bf16(a)
    assert np.max(np.abs(s - a) / np.abs(a)) <= 2.0 ** -16
    assert np.max(np.abs(M.bf16(a) - a) / np.abs(a)) > 2.0 ** -10
