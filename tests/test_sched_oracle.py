"""Pins the scheduler-network oracle on the CPU: analytic gradient (with and
without row weights) against central finite differences, the row weighting
against its defining identities, forward pass against hand-built networks,
and the initialisation and Adam step against their definitions."""
import numpy as np

from oracle import mlp


def _rand(seed, n=7):
    rng = np.random.default_rng(seed)
    w = rng.normal(0, 0.4, mlp.P)
    X = rng.uniform(0, 1, (n, mlp.F))
    T = rng.normal(0, 1, (n, mlp.A))
    return w, X, T


def test_gradient_matches_finite_differences():
    w, X, T = _rand(3)
    loss, g = mlp.loss_grad(w, X, T)
    h = 1e-6
    rng = np.random.default_rng(0)
    idx = np.concatenate([np.arange(0, mlp.P, 37), rng.integers(0, mlp.P, 40), [mlp.P - 1]])
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (mlp.loss_grad(wp, X, T)[0] - mlp.loss_grad(wm, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # the loss is the defined mean squared error
    y = mlp.forward(w, X)
    assert np.isclose(loss, 0.5 * ((y - T) ** 2).sum() / len(X))


def test_weighted_gradient_matches_finite_differences():
    w, X, T = _rand(4, n=9)
    wr = np.random.default_rng(1).uniform(0.1, 3.0, len(X))
    loss, g = mlp.loss_grad(w, X, T, wr)
    h = 1e-6
    idx = np.concatenate([np.arange(0, mlp.P, 41), [mlp.P - 1]])
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (mlp.loss_grad(wp, X, T, wr)[0] - mlp.loss_grad(wm, X, T, wr)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_row_weight_identities():
    """Integer weights equal repeated rows; a common factor changes nothing;
    a zero weight drops its row; all-ones weights are the unweighted loss."""
    w, X, T = _rand(6, n=5)
    k = np.array([1, 3, 0, 2, 1])
    lw, gw = mlp.loss_grad(w, X, T, k)
    lr_, gr = mlp.loss_grad(w, np.repeat(X, k, axis=0), np.repeat(T, k, axis=0))
    assert np.isclose(lw, lr_) and np.allclose(gw, gr, rtol=1e-12, atol=1e-14)
    l7, g7 = mlp.loss_grad(w, X, T, 7.0 * k)
    assert np.isclose(lw, l7) and np.allclose(gw, g7, rtol=1e-12, atol=1e-14)
    l1, g1 = mlp.loss_grad(w, X, T, np.ones(len(X)))
    l0, g0 = mlp.loss_grad(w, X, T)
    assert np.isclose(l1, l0) and np.allclose(g1, g0, rtol=1e-12, atol=1e-14)


def test_row_weights_are_time_at_stake():
    cost = np.array([[4.0, 2.0, 9.0], [10.0, 30.0, 20.0], [6.0, 6.0, 6.0]])
    # smallest costs 2, 10, 6 with mean 6
    assert np.allclose(mlp.row_weights(cost), [2 / 6, 10 / 6, 1.0])
    assert np.isclose(mlp.row_weights(cost).mean(), 1.0)


def test_planted_mistakes_are_caught():
    """A gradient with a dropped ReLU mask or a transposed weight fails the
    finite-difference check, so the check can tell."""
    w, X, T = _rand(5)
    _, g = mlp.loss_grad(w, X, T)
    W1, b1, W2, b2, W3, b3 = mlp.unpack(w)
    Xf = np.asarray(X)
    z1 = Xf @ W1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0)
    dy = (h2 @ W3.T + b3 - T) / len(X)
    d2_nomask = dy @ W3                       # ReLU derivative dropped
    gW2_bad = d2_nomask.T @ h1
    off = mlp.F * mlp.H + mlp.H
    assert not np.allclose(gW2_bad.ravel(), g[off:off + mlp.H * mlp.H], atol=1e-9)
    d2 = (dy @ W3) * (z2 > 0)
    d1_transposed = (d2 @ W2.T) * (z1 > 0)    # wrong orientation of W2
    gW1_bad = d1_transposed.T @ Xf
    assert not np.allclose(gW1_bad.ravel(), g[:mlp.F * mlp.H], atol=1e-9)


def test_forward_hand_built_network():
    # W1 = [I; 0], W2 = I, W3 = 2*[I 0], b3 = 0..A-1: y_a = 2*relu(x_a) + a
    w = np.zeros(mlp.P)
    for j in range(mlp.F):
        w[j * mlp.F + j] = 1.0
    off_W2 = mlp.F * mlp.H + mlp.H
    for j in range(mlp.H):
        w[off_W2 + j * mlp.H + j] = 1.0
    off_W3 = off_W2 + mlp.H * mlp.H + mlp.H
    for a in range(mlp.A):
        w[off_W3 + a * mlp.H + a] = 2.0
    w[-mlp.A:] = np.arange(mlp.A)
    x = np.array([[0.5, -1.0, 0.25, 3.0, -2.0, 1.0, 0, 0, 0, 0, 0, 0]])
    y = mlp.forward(w, x)
    assert np.allclose(y[0], 2 * np.maximum(x[0, :mlp.A], 0) + np.arange(mlp.A))
    assert mlp.argmin_actions(w, x)[0] == int(np.argmin(y[0]))


def test_init_and_adam_definitions():
    w = mlp.init_params(7)
    assert w.dtype == np.float32 and w.shape == (mlp.P,)
    W1, b1, W2, b2, W3, b3 = mlp.unpack(w)
    for W, fi, fo in [(W1, mlp.F, mlp.H), (W2, mlp.H, mlp.H), (W3, mlp.H, mlp.A)]:
        lim = np.sqrt(6.0 / (fi + fo))
        assert np.abs(W).max() <= lim and np.abs(W).max() > 0.8 * lim
        assert abs(W.mean()) < 0.2 * lim
    assert not b1.any() and not b2.any() and not b3.any()
    assert not np.array_equal(mlp.init_params(7), mlp.init_params(8))
    # Adam's first step moves each parameter by lr * sign(g) (bias-corrected)
    g = np.array([0.5, -2.0, 1e-3])
    w1, m, v = mlp.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01)
    assert np.allclose(w1, -0.01 * np.sign(g), rtol=1e-4)


def test_row_targets_are_centred_logs():
    c = np.array([[1.0, 2.0, 4.0, 8.0, 16.0, 32.0]])
    t = mlp.row_targets(c)
    assert np.allclose(t.sum(), 0) and np.allclose(np.diff(t[0]), np.log(2))
