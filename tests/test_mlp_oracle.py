"""Pins of the surrogate MLP oracle (oracle/mlp_oracle.py): central finite differences of
the loss for every parameter tensor, the tanh-GELU closed form, and the feature transform."""
import math
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402


def test_gelu_closed_form_and_derivative():
    z = np.linspace(-5, 5, 101)
    ref = np.array([0.5 * v * (1 + math.tanh(math.sqrt(2 / math.pi) * (v + 0.044715 * v ** 3))) for v in z])
    assert np.allclose(M.gelu(z), ref, atol=1e-15)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.allclose(M.gelu_grad(z), fd, atol=1e-8)
    # GELU(z) -> z for large z, -> 0 for very negative z
    assert abs(M.gelu(np.array([8.0]))[0] - 8.0) < 1e-9 and abs(M.gelu(np.array([-8.0]))[0]) < 1e-9


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init_params(1)
    for k in p:
        p[k] = p[k] + 0.05 * rng.standard_normal(p[k].shape)
    f = rng.standard_normal((7, M.IN))
    t = rng.standard_normal((7, M.OUT))
    _, g = M.loss_and_grads(p, f, t)
    h = 1e-6
    for name in ["W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4"]:
        flat = p[name].reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp, _ = M.loss_and_grads(p, f, t)
            flat[idx] = old - h
            lm, _ = M.loss_and_grads(p, f, t)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(g[name].reshape(-1)[idx] - fd) <= 1e-6 * max(1.0, abs(fd)), (name, idx)
    # the unused outputs 9..15 carry no loss
    assert np.all(g["W4"][:, M.NOUT:] == 0) and np.all(g["b4"][M.NOUT:] == 0)


def test_features():
    x = np.array([[273.15, 85000.0, 1.0, 0.01, 1e-3, 0.0, 0.0, 2e-4, 0.0, 5e4, 0.0, 3.0, 250.0, 2.0]])
    f = M.features(x)[0]
    assert f.shape == (M.IN,)
    assert f[0] == (273.15 - 250) / 30 and abs(f[1] - 0.25) < 1e-12 and abs(f[2] - 0.3) < 1e-12
    assert abs(f[3] - (math.log(0.01 + 1e-8) + 12) / 4) < 1e-12
    assert abs(f[5] - (math.log(1e-8) + 12) / 4) < 1e-12  # qr = 0
    assert abs(f[9] - (math.log(5e4 + 1) - 5) / 5) < 1e-12
    assert f[11] == 0.3 and f[12] == -0.5 and abs(f[13] + 0.6) < 1e-12 and f[14] == 0 and f[15] == 0


def test_teacher_rows_get_one_order_whatever_order_the_gpu_appended_them():
    """Training draws batches by row index: the same set of teacher rows in any order must reach it in one order."""
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))
    from outbreaklab_amd.surrogate import canonical_order

    rng = np.random.default_rng(5)
    rows = rng.standard_normal((5000, 24)).astype(np.float32)
    rows[100] = rows[7]  # identical rows are interchangeable
    a = canonical_order(rows)
    for seed in (1, 2):
        b = canonical_order(rows[np.random.default_rng(seed).permutation(len(rows))])
        assert np.array_equal(a, b)
    assert np.array_equal(np.sort(a, axis=0), np.sort(rows, axis=0))  # a permutation of the rows
