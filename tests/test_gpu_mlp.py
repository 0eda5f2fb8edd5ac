"""MFMA kernels of the surrogate MLP against the CPU oracle (needs an MI355X).

Bars: the forward pass and every gradient tensor within 2e-3 (norm-relative)
of the oracle that rounds to bfloat16 where the kernels do (the remaining
difference is FP32 accumulation order and tanhf), within 5e-2 of the exact
FP64 gradient; one Adam step within 1e-5 of the oracle's step from the same
gradient; training deterministic bit for bit and reducing the loss.
"""
import os
import sys

import numpy as np
import pytest

from conftest import announce

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sur():
    from elasticscattering_amd import surrogate

    return surrogate


def _data(n, seed=1):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, M.N_IN)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) + x[:, 1] ** 2, x[:, 2] * x[:, 3] - 0.3 * x[:, 5]], axis=1).astype(np.float32)
    return x, t


def _rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30)


def test_forward_matches_oracle(sur):
    announce("mlp forward vs oracle, 3 members, 1000 rows")
    mlp = sur.MLP(members=3, seed=5)
    p = mlp.params
    x, _ = _data(1000, 2)
    y = mlp.predict(x)
    assert y.shape == (3, 1000, 2)
    for m in range(3):
        yo, _ = M.forward(p[m].astype(np.float64), x.astype(np.float64), emulate_bf16=True)
        assert _rel(y[m], yo) <= 2e-3, m
        yx, _ = M.forward(p[m].astype(np.float64), x.astype(np.float64), emulate_bf16=False)
        assert _rel(y[m], yx) <= 5e-2, m
    mlp.close()


@pytest.mark.parametrize("nrows", [64, 200, 1])
def test_gradients_match_oracle(sur, nrows):
    announce(f"mlp gradients vs oracle, nrows={nrows}")
    mlp = sur.MLP(members=2, seed=9)
    x, t = _data(500, 3)
    mlp.set_data(x, t)
    # move away from the initialisation so every layer carries gradient
    mlp.train(epochs=3, batch=100, lr=3e-3, lr_final=3e-3)
    p = mlp.params
    rows = np.random.default_rng(nrows).choice(500, size=nrows, replace=False).astype(np.int32)
    ow = (1.0, 0.7)
    g, loss = mlp.gradients(rows, ow)
    xs, ts = x[rows].astype(np.float64), t[rows].astype(np.float64)
    for m in range(2):
        lo, go = M.loss_and_grad(p[m].astype(np.float64), xs, ts, ow, emulate_bf16=True)
        lx, gx = M.loss_and_grad(p[m].astype(np.float64), xs, ts, ow, emulate_bf16=False)
        assert abs(loss[m] - lo) <= 2e-3 * lo + 1e-7
        for name, a, b, c in zip(("W1", "b1", "W2", "b2", "W3", "b3"), M.unpack(g[m]), M.unpack(go), M.unpack(gx)):
            assert _rel(a, b) <= 2e-3 + (2e-2 if nrows == 1 else 0.0), (m, name, _rel(a, b))
            assert _rel(a, c) <= 5e-2, (m, name, _rel(a, c))
        W1g, _, _, _, W3g, b3g = M.unpack(g[m])
        assert not W1g[:, M.N_IN:].any() and not W3g[2:].any() and not b3g[2:].any()
    mlp.close()


def test_adam_step_matches_oracle(sur):
    announce("one Adam step vs oracle")
    mlp = sur.MLP(members=2, seed=3)
    x, t = _data(128, 4)
    mlp.set_data(x, t)
    p0 = mlp.params.astype(np.float64)
    # a full-batch epoch is one step; its rows are a permutation of all rows,
    # so the gradient equals the one over rows 0..127 up to summation order
    g, _ = mlp.gradients(np.arange(128, dtype=np.int32))
    mlp.train(epochs=1, batch=128, lr=1e-3, lr_final=1e-5)
    p1 = mlp.params.astype(np.float64)
    for m in range(2):
        po, _, _ = M.adam_step(p0[m], np.zeros(M.NPARAM), np.zeros(M.NPARAM), g[m].astype(np.float64), 1, 1e-3)
        moved = np.abs(g[m]) > 1e-6 * np.abs(g[m]).max()
        np.testing.assert_allclose(p1[m][moved], po[moved], rtol=0, atol=2e-6)
        assert np.array_equal(p1[m][~moved & (g[m] == 0)], p0[m][~moved & (g[m] == 0)])
    mlp.close()


def test_training_reduces_loss_and_is_deterministic(sur):
    announce("training 4 members, 2 runs")
    x, t = _data(4096, 7)
    runs = []
    for _ in range(2):
        mlp = sur.MLP(members=4, seed=21)
        mlp.set_data(x, t)
        hist = mlp.train(epochs=40, batch=256, lr=3e-3, lr_final=1e-4, shuffle_seed=5)
        runs.append((hist, mlp.params))
        secs, steps = mlp.last_train_time()
        assert secs > 0 and steps == 40 * 16
        mlp.close()
    (h0, p0), (h1, p1) = runs
    assert np.array_equal(h0, h1) and np.array_equal(p0, p1)
    assert np.all(h0[-1] < h0[0] / 10.0), (h0[0], h0[-1])
    members_differ = [not np.array_equal(p0[0], p0[m]) for m in range(1, 4)]
    assert all(members_differ)


def test_forward_is_smooth_along_one_input(sur):
    """The forward pass carries each activation as a BF16 pair: along a fine
    sweep of one input the output's second differences stay at the level of
    the function's own curvature, far below one BF16 step."""
    announce("mlp forward smoothness along a 400-point sweep")
    mlp = sur.MLP(members=1, seed=8)
    x = np.zeros((400, M.N_IN), dtype=np.float32)
    x[:, 0] = np.linspace(-1, 1, 400)
    x[:, 1:] = np.array([0.3, -0.2, 0.5, 0.1, -0.7])
    y = mlp.predict(x)[0].astype(np.float64)
    yo, _ = M.forward(mlp.params[0].astype(np.float64), x.astype(np.float64), emulate_bf16=True)
    d2 = np.abs(np.diff(y, 2, axis=0)).max()
    d2o = np.abs(np.diff(yo, 2, axis=0)).max()
    assert d2 <= 3 * d2o + 1e-5, (d2, d2o)
    mlp.close()
