"""CPU checks that pin the network oracle (oracle/cnn_oracle.py).

The network has no reference counterpart; its oracle is pinned by closed forms
of a single convolution (identity and shift kernels, linearity, the adjoint
identity <conv(x), g> = <x, conv^T(g)>) and by central finite differences of
the full loss, which its backward pass must match.
"""
import numpy as np

from oracle import cnn_oracle as C


def test_single_conv_closed_forms():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 4, 5, 6, 3))
    # centre tap = identity
    W = np.zeros((27, 3, 3))
    W[13] = np.eye(3)
    assert np.allclose(C.conv(x, W, np.zeros(3)), x)
    # tap (dz, dy, dx) = (0, 0, 1) reads the +x neighbour, zero at the far face
    W = np.zeros((27, 3, 3))
    W[C.OFFSETS.index((0, 0, 1))] = np.eye(3)
    y = C.conv(x, W, np.zeros(3))
    assert np.allclose(y[..., :-1, :], x[..., 1:, :]) and np.allclose(y[..., -1, :], 0)
    # bias only
    assert np.allclose(C.conv(x, np.zeros((27, 2, 3)), np.array([1.5, -2.0]))[..., 1], -2.0)
    # adjoint identity
    W = rng.normal(size=(27, 4, 3))
    g = rng.normal(size=(2, 4, 5, 6, 4))
    lhs = np.sum(C.conv(x, W, np.zeros(4)) * g)
    rhs = np.sum(x * C.conv_input_grad(g, W))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    # weight gradient = derivative of <conv(x; W), g> in W
    gw = C.conv_weight_grad(g, x)
    i = (7, 2, 1)
    E = np.zeros_like(W)
    E[i] = 1.0
    assert abs(np.sum(C.conv(x, E, np.zeros(4)) * g) - gw[i]) < 1e-10


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    ch = [3, 4, 4, 1]
    params = C.init_params(2, ch)
    params = [(W, rng.normal(0, 0.1, b.shape)) for W, b in params]
    x = rng.normal(size=(2, 4, 3, 5, 3))
    t = rng.normal(size=(2, 4, 3, 5))
    loss, grads = C.loss_and_grad(params, x, t)
    flat = C.pack(params)
    gflat = C.pack(grads)
    h = 1e-6
    for k in rng.choice(flat.size, 40, replace=False):
        fp, fm = flat.copy(), flat.copy()
        fp[k] += h
        fm[k] -= h
        lp, _ = C.loss_and_grad(C.unpack(fp, ch), x, t)
        lm, _ = C.loss_and_grad(C.unpack(fm, ch), x, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gflat[k]) < 1e-6 * max(1.0, abs(fd)), (k, fd, gflat[k])


def test_dilated_conv_closed_forms():
    """Dilation d moves every tap to d * offset: the +x tap reads x + d (zero in
    the last d planes); the adjoint identity still holds."""
    rng = np.random.default_rng(3)
    x = rng.normal(size=(1, 5, 6, 9, 2))
    for d in (2, 3, 8):
        W = np.zeros((27, 2, 2))
        W[C.OFFSETS.index((0, 0, 1))] = np.eye(2)
        y = C.conv(x, W, np.zeros(2), dil=d)
        if d < 9:
            assert np.allclose(y[..., :9 - d, :], x[..., d:, :])
        assert np.allclose(y[..., max(0, 9 - d):, :], 0)
        W = np.zeros((27, 2, 2))
        W[C.OFFSETS.index((-1, 0, 0))] = np.eye(2)
        y = C.conv(x, W, np.zeros(2), dil=d)
        if d < 5:
            assert np.allclose(y[:, d:], x[:, :5 - d])
        assert np.allclose(y[:, :min(d, 5)], 0)
        W = rng.normal(size=(27, 3, 2))
        g = rng.normal(size=(1, 5, 6, 9, 3))
        lhs = np.sum(C.conv(x, W, np.zeros(3), dil=d) * g)
        rhs = np.sum(x * C.conv_input_grad(g, W, dil=d))
        assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_dilated_backward_matches_finite_differences():
    rng = np.random.default_rng(4)
    ch = [3, 4, 4, 1]
    dil = [1, 2, 3]
    params = C.init_params(5, ch)
    params = [(W, rng.normal(0, 0.1, b.shape)) for W, b in params]
    x = rng.normal(size=(1, 5, 4, 7, 3))
    t = rng.normal(size=(1, 5, 4, 7))
    loss, grads = C.loss_and_grad(params, x, t, dil=dil)
    flat, gflat = C.pack(params), C.pack(grads)
    h = 1e-6
    for k in rng.choice(flat.size, 40, replace=False):
        fp, fm = flat.copy(), flat.copy()
        fp[k] += h
        fm[k] -= h
        lp, _ = C.loss_and_grad(C.unpack(fp, ch), x, t, dil=dil)
        lm, _ = C.loss_and_grad(C.unpack(fm, ch), x, t, dil=dil)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gflat[k]) < 1e-6 * max(1.0, abs(fd)), (k, fd, gflat[k])


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -3.140625, 0.0])
    r = C.bf16_round(v)
    # bfloat16 spacing at 1 is 2^-7: 1 + 2^-8 is the midpoint, the even mantissa (1.0) wins
    assert r[0] == 1.0 and r[1] == 1.0
    assert r[3] == 1.0                       # below the midpoint: rounds down
    assert r[2] == 1.0 + 2 ** -7             # above the midpoint: rounds up
    assert r[4] == -3.140625 and r[5] == 0.0
