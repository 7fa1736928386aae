"""CPU FP64 oracle of the learned flux correction network (test infrastructure only).

The network has no counterpart in the reference: it is pinned by central
finite differences of this oracle's own forward pass (the backward pass is
checked against them in tests/test_cnn_oracle.py) and by closed forms of a
single 3x3x3 convolution.  With ``bf16=True`` the forward/backward emulate the
GPU kernels' rounding points (weights, network input and hidden activations
rounded to bfloat16, round-to-nearest-even; the loss gradient rounded to
bfloat16; accumulation kept in FP64), so the GPU's FP32-accumulated results
can be compared with it.

Layout (same as allium_amd/csrc/cnn.hip):
  activations  [B, Z, Y, X, C]  (channels last)
  weights      W[l] [27, CO, CI] with offset o = (dz+1)*9 + (dy+1)*3 + (dx+1)
  biases       b[l] [CO]
  network      5 layers 3 -> 32 -> 32 -> 32 -> 32 -> 1, ReLU after the first four,
               zero padding ("same" size); layer l's taps optionally dilated
               by dil[l] (offsets dil * (-1, 0, 1) per axis)
  loss         mean over voxels of (out - target)^2
"""
import numpy as np

OFFSETS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
CHANNELS = [3, 32, 32, 32, 32, 1]


def bf16_round(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = ((u + 0x7FFF + lsb) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def param_sizes(channels=CHANNELS):
    """(W shape, b shape) per layer, in the packed parameter order W0, b0, W1, b1, ..."""
    return [((27, channels[l + 1], channels[l]), (channels[l + 1],)) for l in range(len(channels) - 1)]


def pack(params):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in params])


def unpack(flat, channels=CHANNELS):
    out, i = [], 0
    for ws, bs in param_sizes(channels):
        nw, nb = int(np.prod(ws)), int(np.prod(bs))
        out.append((flat[i:i + nw].reshape(ws), flat[i + nw:i + nw + nb].reshape(bs)))
        i += nw + nb
    return out


def shifted(x, dz, dy, dx):
    """y[..., z, y, x, :] = x[..., z+dz, y+dy, x+dx, :] with zeros outside."""
    B, Z, Y, X, C = x.shape
    p = max(abs(dz), abs(dy), abs(dx), 1)
    xp = np.pad(x, ((0, 0), (p, p), (p, p), (p, p), (0, 0)))
    return xp[:, p + dz:p + dz + Z, p + dy:p + dy + Y, p + dx:p + dx + X, :]


def conv(x, W, b, dil=1):
    """out[v, co] = b[co] + sum_o sum_ci W[o, co, ci] x[v + dil * off(o), ci]."""
    out = np.zeros(x.shape[:4] + (W.shape[1],)) + b
    for o, (dz, dy, dx) in enumerate(OFFSETS):
        out += np.einsum("bzyxc,oc->bzyxo", shifted(x, dil * dz, dil * dy, dil * dx), W[o], optimize=True)
    return out


def conv_input_grad(g, W, dil=1):
    """Adjoint of conv in its input: gx[v, ci] = sum_o sum_co W[o, co, ci] g[v - dil * off(o), co]."""
    out = np.zeros(g.shape[:4] + (W.shape[2],))
    for o, (dz, dy, dx) in enumerate(OFFSETS):
        out += np.einsum("bzyxo,oc->bzyxc", shifted(g, -dil * dz, -dil * dy, -dil * dx), W[o], optimize=True)
    return out


def conv_weight_grad(g, x, dil=1):
    """gW[o, co, ci] = sum_v g[v, co] x[v + dil * off(o), ci]."""
    return np.stack([np.einsum("bzyxo,bzyxc->oc", g, shifted(x, dil * dz, dil * dy, dil * dx), optimize=True)
                     for (dz, dy, dx) in OFFSETS])


def forward(params, x, bf16=False, dil=None):
    """x [B, Z, Y, X, 3] -> (out [B, Z, Y, X], cache); dil: per-layer dilation (default 1)."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    dil = dil or [1] * len(params)
    a = r(x)
    acts, zs = [a], []
    L = len(params)
    for l, (W, b) in enumerate(params):
        z = conv(a, r(W), b, dil[l])
        zs.append(z)
        if l < L - 1:
            a = r(np.maximum(z, 0.0))
            acts.append(a)
    return zs[-1][..., 0], {"acts": acts, "zs": zs}


def loss_and_grad(params, x, target, bf16=False, dil=None):
    """Mean squared error and its gradient with respect to every parameter."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    dil = dil or [1] * len(params)
    out, cache = forward(params, x, bf16, dil)
    n = out.size
    diff = out - target
    loss = float(np.sum(diff * diff) / n)
    g = r(2.0 * diff / n)[..., None]   # gradient of the last pre-activation
    grads = [None] * len(params)
    for l in range(len(params) - 1, -1, -1):
        W, b = params[l]
        a_in = cache["acts"][l]
        grads[l] = (conv_weight_grad(g, a_in, dil[l]), g.sum(axis=(0, 1, 2, 3)))
        if l > 0:
            gin = conv_input_grad(g, r(W), dil[l])
            g = r(gin * (cache["acts"][l] > 0))
    return loss, grads


def init_params(seed=0, channels=CHANNELS):
    """He-normal weights, zero biases (the GPU handle uses its own seeded init)."""
    rng = np.random.default_rng(seed)
    out = []
    for ws, bs in param_sizes(channels):
        fan_in = ws[0] * ws[2]
        out.append((rng.normal(0.0, np.sqrt(2.0 / fan_in), ws), np.zeros(bs)))
    return out
