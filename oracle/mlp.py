"""CPU oracle of the learned scheduler's network (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use
this module, and only as the checker of the engine's HIP kernels.

The network has no counterpart in the reference (its s–t visit always
expands both frontiers top-down, bernaschi/ST_CONN_ONE_ATOMIC/current-st/
make_stcon.c:960-1200); it is restated here in float64 numpy:

  h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = W3 h2 + b3
  L  = 0.5 * sum_r w_r * sum_a (y[r,a] - target[r,a])^2 / sum_r w_r

(w_r = 1 unless row weights are given; row_weights() gives the
time-at-stake weighting), with the flat parameter layout W1[32][12] b1[32]
W2[32][32] b2[32] W3[9][32] b3[9]. Pinning: the analytic gradient, weighted
and not, is checked against central finite differences in float64
(tests/test_sched_oracle.py); the weighting against its defining
identities (integer weights equal repeated rows, a common factor changes
nothing, a zero weight drops the row); the forward pass
against hand-built networks with known outputs, and the initialisation and
Adam step against their published definitions (Glorot & Bengio 2010 uniform
bound sqrt(6/(fan_in+fan_out)); Kingma & Ba 2015 with bias correction).
"""
import numpy as np

F, H, A = 12, 32, 9
P = F * H + H + H * H + H + A * H + A
_M64 = (1 << 64) - 1


def unpack(w):
    w = np.asarray(w, np.float64)
    o = 0
    W1 = w[o:o + H * F].reshape(H, F); o += H * F
    b1 = w[o:o + H]; o += H
    W2 = w[o:o + H * H].reshape(H, H); o += H * H
    b2 = w[o:o + H]; o += H
    W3 = w[o:o + A * H].reshape(A, H); o += A * H
    b3 = w[o:o + A]; o += A
    assert o == P
    return W1, b1, W2, b2, W3, b3


def forward(w, X):
    W1, b1, W2, b2, W3, b3 = unpack(w)
    X = np.asarray(X, np.float64)
    h1 = np.maximum(X @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ W3.T + b3


def loss_grad(w, X, T, weight=None):
    """Loss and gradient (flat, float64) of
    0.5*sum_r w_r*sum_a (y-T)^2 / sum_r w_r (weight None: all w_r = 1)."""
    W1, b1, W2, b2, W3, b3 = unpack(w)
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    wr = np.ones(len(X)) if weight is None else np.asarray(weight, np.float64)
    z1 = X @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ W3.T + b3
    d = y - T
    loss = 0.5 * float((wr[:, None] * d * d).sum()) / wr.sum()
    dy = d * (wr / wr.sum())[:, None]
    gW3 = dy.T @ h2
    gb3 = dy.sum(0)
    d2 = (dy @ W3) * (z2 > 0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (z1 > 0)
    gW1 = d1.T @ X
    gb1 = d1.sum(0)
    return loss, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])


def row_targets(cost):
    """Training targets of traced rows: log(cost) minus the row mean."""
    lc = np.log(np.asarray(cost, np.float64))
    return lc - lc.mean(axis=1, keepdims=True)


def row_weights(cost):
    """Time-at-stake row weights: each row's smallest cost over the
    actions divided by the mean of those over the rows."""
    m = np.asarray(cost, np.float64).min(axis=1)
    return m / m.mean()


def init_params(seed):
    """Seeded initialisation: Glorot-uniform weights from a SplitMix64
    stream (state advanced by the golden-ratio increment, output mixed with
    the 30/27/31 shift-multiply finaliser), zero biases."""
    st = (seed * 0x9E3779B97F4A7C15 + 1) & _M64
    w = np.zeros(P, np.float32)

    def uni():
        nonlocal st
        st = (st + 0x9E3779B97F4A7C15) & _M64
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        z ^= z >> 31
        return (z >> 11) * (1.0 / 9007199254740992.0)

    def fill(off, fan_out, fan_in):
        lim = np.sqrt(6.0 / (fan_in + fan_out))
        for i in range(fan_out * fan_in):
            w[off + i] = np.float32((2.0 * uni() - 1.0) * lim)

    fill(0, H, F)
    fill(H * F + H, H, H)
    fill(H * F + H + H * H + H, A, H)
    return w


def adam_step(w, g, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (step t, 1-based); returns new (w, m, v)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** t)
    vh = v / (1 - beta2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def argmin_actions(w, X):
    """The action the network picks per row (first index of the minimum)."""
    return np.argmin(forward(w, X), axis=1)
