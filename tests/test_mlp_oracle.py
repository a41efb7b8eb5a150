"""The CPU checker of the contact-seed network (oracle/mlp_ref.py), pinned
independently of the GPU kernels it checks:
  * bfloat16 rounding against torch's own float32 -> bfloat16 conversion;
  * the exact-mode gradient against central finite differences;
  * the features against closed forms (frame invariance, pure rotations);
  * the seed-angle map against the solve's normal parametrisation;
  * one Adam step against its closed form."""
import numpy as np
import pytest
import torch

from oracle import mlp_ref as M


def test_bf16_rounding_matches_torch():
    g = np.random.default_rng(0)
    x = np.concatenate([g.normal(size=5000) * 10.0 ** g.integers(-20, 20, 5000),
                        [1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 0.0]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    np.testing.assert_array_equal(M.bf16(x), ref)
    assert M.bf16(np.float32(1.0 + 2 ** -8)) == 1.0          # tie to even
    assert M.bf16(np.float32(1.0 + 3 * 2 ** -9)) == 1.0 + 2 ** -7


def _targets(g, n):
    t = g.normal(size=(n, 3))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    return np.concatenate([t, (g.random((n, 1)) < 0.6).astype(float)], axis=1)


def test_loss_closed_forms():
    """Contact rows: squared distance of the normal plus the logistic loss;
    other rows: the logistic loss only (softplus(z) = log(1 + e^z))."""
    y = np.array([[0.0, 1.0, 0.0, 2.0], [5.0, 5.0, 5.0, -1.0]])
    T = np.array([[1.0, 0.0, 0.0, 1.0], [0.0, 0.0, 1.0, 0.0]])
    np.testing.assert_allclose(M.pair_loss(y, T),
                               [2.0 + np.log1p(np.exp(2.0)) - 2.0, np.log1p(np.exp(-1.0))],
                               rtol=1e-14)


def test_exact_gradient_matches_finite_differences():
    g = np.random.default_rng(1)
    theta = M.init_params(3).astype(np.float64)
    theta += 0.05 * g.normal(size=theta.size)                # non-zero biases too
    X = g.normal(size=(40, M.NIN))
    T = _targets(g, 40)
    grad, _ = M.backward(theta, X, T, mode="exact")
    idx = np.concatenate([g.choice(theta.size, 200, replace=False),
                          np.arange(M.NPARAM - 4, M.NPARAM)])  # b3 included
    h = 1e-6
    for k in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (M.loss(tp, X, T) - M.loss(tm, X, T)) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, grad[k])


def test_bf16_mode_is_close_to_exact():
    g = np.random.default_rng(2)
    theta = M.init_params(4)
    X = g.normal(size=(256, M.NIN))
    T = _targets(g, 256)
    ge, ye = M.backward(theta, X, T, mode="exact")
    gb, yb = M.backward(theta, X, T, mode="bf16")
    assert np.abs(yb - ye).max() <= 0.05 * np.abs(ye).max()
    cos = ge @ gb / np.linalg.norm(ge) / np.linalg.norm(gb)
    assert cos > 0.999


def _rand_q(g, n):
    q = g.normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def test_features_frame_invariance_and_closed_forms():
    g = np.random.default_rng(5)
    n = 50
    b1, b2 = g.normal(size=(n, 3)), g.normal(size=(n, 3))
    q1, q2 = _rand_q(g, n), _rand_q(g, n)
    a1, a2 = g.uniform(0.1, 1.0, (n, 3)), g.uniform(0.1, 1.0, (n, 3))
    F = M.features(b1, q1, a1, b2, q2, a2)
    # a global rotation of both particles and their centres changes nothing
    qg = _rand_q(g, 1).repeat(n, axis=0)
    Rg = M.rotm(qg)
    F2 = M.features(np.einsum("nij,nj->ni", Rg, b1), M.qmul(qg, q1), a1,
                    np.einsum("nij,nj->ni", Rg, b2), M.qmul(qg, q2), a2)
    np.testing.assert_allclose(F2, F, atol=1e-12)
    # identity orientations: the offset itself, no relative rotation
    I = np.tile([1.0, 0, 0, 0], (n, 1))
    F3 = M.features(b1, I, a1, b2, I, a2)
    s = a1.max(1) + a2.max(1)
    np.testing.assert_allclose(F3[:, 0:3], (b2 - b1) / s[:, None], atol=1e-14)
    np.testing.assert_allclose(F3[:, 9:12], 0.0, atol=1e-14)
    # relative rotation by angle t about z -> rotation vector (0, 0, t)
    t = g.uniform(0.0, 3.0, n)
    qz = np.stack([np.cos(t / 2), 0 * t, 0 * t, np.sin(t / 2)], axis=1)
    F4 = M.features(b1, q1, a1, b2, M.qmul(q1, qz), a2)
    np.testing.assert_allclose(F4[:, 9:12], np.stack([0 * t, 0 * t, t], 1), atol=1e-12)


def test_seed_angles_reproduce_the_normal():
    g = np.random.default_rng(6)
    q1 = _rand_q(g, 100)
    y = g.normal(size=(100, 4))
    x = M.seed_angles(q1, y)
    cv = np.stack([np.cos(x[:, 0]) * np.cos(x[:, 1]), np.sin(x[:, 0]) * np.cos(x[:, 1]),
                   np.sin(x[:, 1])], 1)
    n = y[:, :3] / np.linalg.norm(y[:, :3], axis=1, keepdims=True)
    np.testing.assert_allclose(M.target(q1, cv), n, atol=1e-12)


def test_adam_first_step_closed_form():
    g = np.random.default_rng(7)
    theta = g.normal(size=10).astype(np.float32)
    grad = g.normal(size=10).astype(np.float32)
    t1, m, v = M.adam(theta, np.zeros(10, np.float32), np.zeros(10, np.float32), grad, 1, 1e-3)
    np.testing.assert_allclose(t1, theta - 1e-3 * grad / (np.abs(grad) + 1e-8), rtol=1e-5, atol=1e-7)
