"""Learned low-order flux correction for source iteration (include/allium_cnn.h).

No counterpart in the reference: the network stands in for diffusion
synthetic acceleration (src/transport_op.hpp:38-59).  Inputs on the element
grid are the latest element-averaged scalar-flux change divided by its RMS,
log sigma_t and the scattering ratio; the output, times the same RMS, is an
element-constant correction added to the scalar flux.  It is trained on
(change, converged-minus-current flux) pairs from plain source iterations.
"""
import ctypes
import time

import numpy as np

from . import _lib
from ._lib import check


class DeviceArray32:
    """A float32 array in device memory."""

    def __init__(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        self.shape, self.n = a.shape, a.size
        lib = _lib.load()
        p = ctypes.c_void_p()
        check(lib.allium_device_malloc(ctypes.byref(p), 4 * max(1, self.n)))
        self.ptr = p
        self.upload(a)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.size != self.n:
            raise ValueError("size mismatch")
        check(_lib.load().allium_memcpy_h2d(self.ptr, a.ctypes.data_as(ctypes.c_void_p), 4 * self.n))

    def __del__(self):
        try:
            if self.ptr is not None and self.ptr.value:
                _lib.load().allium_device_free(self.ptr)
        except Exception:
            pass


class FluxCorrector:
    """The correction network on the GPU (BF16 MFMA kernels, FP32 master weights)."""

    def __init__(self, seed=0, dilation=None):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        check(self._lib.allium_cnn_create(ctypes.c_ulonglong(seed), ctypes.byref(h)))
        self._h = h
        self.nparams = int(self._lib.allium_cnn_num_params())
        if dilation is not None:
            self.set_dilation(dilation)

    def set_dilation(self, dilation):
        """Per-layer dilation of the 3x3x3 taps (5 values in 1..8)."""
        d = (ctypes.c_int * 5)(*[int(v) for v in dilation])
        check(self._lib.allium_cnn_set_dilation(self._h, d))

    def dilation(self):
        d = (ctypes.c_int * 5)()
        check(self._lib.allium_cnn_get_dilation(self._h, d))
        return tuple(d)

    @property
    def handle(self):
        return self._h

    def params(self):
        out = np.empty(self.nparams, dtype=np.float32)
        check(self._lib.allium_cnn_get_params(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters")
        check(self._lib.allium_cnn_set_params(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def forward(self, x):
        """x [B, Z, Y, X, 3] (host) -> out [B, Z, Y, X]."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        B, Z, Y, X, _ = x.shape
        dx = DeviceArray32(x)
        dout = DeviceArray32(np.zeros((B, Z, Y, X), np.float32))
        check(self._lib.allium_cnn_forward_device(self._h, B, Z, Y, X, dx.ptr, dout.ptr, None))
        out = np.empty((B, Z, Y, X), np.float32)
        check(self._lib.allium_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), dout.ptr, 4 * out.size))
        return out

    def gradient(self, x, target):
        """(loss, gradient packed like the parameters) for one batch."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        B, Z, Y, X, _ = x.shape
        dx, dt = DeviceArray32(x), DeviceArray32(target)
        g = np.empty(self.nparams, dtype=np.float32)
        loss = ctypes.c_double()
        check(self._lib.allium_cnn_gradient_device(self._h, B, Z, Y, X, dx.ptr, dt.ptr,
                                                   g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                   ctypes.byref(loss)))
        return loss.value, g

    def train_step(self, dx, dt, shape, lr):
        """One Adam step on device batches; returns the batch loss before the step."""
        B, Z, Y, X = shape
        loss = ctypes.c_double()
        check(self._lib.allium_cnn_train_step_device(self._h, B, Z, Y, X, dx.ptr, dt.ptr, ctypes.c_float(lr),
                                                     ctypes.byref(loss)))
        return loss.value

    def __del__(self):
        try:
            if self._h:
                self._lib.allium_cnn_destroy(self._h)
        except Exception:
            pass


def pairs_from_history(hist, phi_avg, sigma_t, c, shape, iters):
    """Training pairs from one source iteration.

    hist[k-1] = (element-averaged scalar flux after iteration k, the change
    r_k iteration k made before any correction), phi_avg the converged flux.
    For each k in ``iters``: input (r_k over its RMS, log sigma_t, c), target
    (phi* minus the uncorrected iterate phi_{k-1} + r_k) over the same RMS."""
    X, T = [], []
    logs = np.log(sigma_t).reshape(shape)
    cc = np.asarray(c, dtype=np.float64).reshape(shape)
    prev = np.zeros_like(phi_avg)
    for k in range(1, len(hist) + 1):
        state, r = hist[k - 1, 0], hist[k - 1, 1]
        if k in iters:
            s = np.sqrt(np.mean(r * r))
            if s > 0:
                X.append(np.stack([(r / s).reshape(shape), logs, cc], axis=-1))
                T.append(((phi_avg - (prev + r)) / s).reshape(shape))
        prev = state
    return X, T


def generate_training_data(n=32, media=200, seed0=1000, iters=(1, 2, 3, 4, 6, 8, 11, 15), log_sd=1.5, c=0.9999,
                           sn=8, maxit=400, log=None, accel=None):
    """Run GPU source iteration (plain, or with ``accel``'s correction so
    that the pairs follow the corrected iterates) on seeded media and
    collect pairs."""
    import allium_amd as A
    from allium_amd import problems as P

    Xs, Ts = [], []
    shape = (n, n, n)
    t0 = time.perf_counter()
    Linv = q = None
    for i in range(media):
        cfg = P.problem_3d_hetero(n=n, seed=seed0 + i, log_sd=log_sd, c_const=c)
        if Linv is None:
            # one sweep handle for every medium: only sigma_t changes (the source
            # Q = 1 does not depend on it), and the launch strategy is tuned once
            Linv = A.InverseAdvectionOperator(A.Mesh(cfg["nel"]), 1, sn, cfg["sigma_t"])
            Linv.set_launch(-1)
            q = Linv.form_source(Q=1.0)
        else:
            Linv.AssembleLocalMatrices(cfg["sigma_t"])
        hist = np.zeros((maxit, 2, n ** 3))
        phi, _, it, _ = A.source_iteration(Linv, cfg["sigma_s"], q, abstol=1e-8, maxit=maxit, want_psi=False,
                                           history=hist, accel=accel, patience=3)
        phi_avg = phi.reshape(-1, 8).mean(axis=1)
        X, T = pairs_from_history(hist[:it - 1], phi_avg, cfg["sigma_t"].reshape(-1), cfg["c"], shape, set(iters))
        Xs += X
        Ts += T
        if log and (i + 1) % 25 == 0:
            log(f"  media {i + 1}/{media}: {len(Xs)} pairs, {time.perf_counter() - t0:.1f}s")
    return np.stack(Xs).astype(np.float32), np.stack(Ts).astype(np.float32)


# dilation of the 3x3x3 taps per layer (receptive field 33 cells) and the
# iterations whose (change, converged - iterate) pairs train the network:
# the asymptotic phase of source iteration is where most iterations go
DILATION = (1, 2, 4, 8, 1)
ITERS = (1, 2, 3, 4, 6, 8, 11, 15, 20, 28, 40, 55)


def fit(n=32, media=60, rounds=2, steps=1500, seed0=1000, sn=8, log=None, seed=1, maxit=1000, dilation=DILATION,
        iters=ITERS):
    """Train a corrector on plain-iteration pairs, then for ``rounds`` more
    rounds add pairs from source iterations that use the current network (the
    iterates the correction actually meets) and train again.  Returns
    (corrector, info)."""
    t0 = time.perf_counter()
    X, T = generate_training_data(n=n, media=media, seed0=seed0, sn=sn, maxit=maxit, iters=iters)
    net = FluxCorrector(seed, dilation=dilation)
    L = train(net, X, T, steps=steps, batch=8, lr=1e-3, seed=seed)
    info = {"rounds": [{"pairs": int(X.shape[0]), "loss": float(np.mean(L[-50:]))}]}
    if log:
        log(f"  round 0: {X.shape[0]} pairs, loss {info['rounds'][-1]['loss']:.3f}")
    for r in range(1, rounds + 1):
        X2, T2 = generate_training_data(n=n, media=media, seed0=seed0 + 1000 * r, sn=sn, maxit=maxit, accel=net,
                                        iters=iters)
        X, T = np.concatenate([X, X2]), np.concatenate([T, T2])
        L = train(net, X, T, steps=steps, batch=8, lr=5e-4, seed=seed + r)
        info["rounds"].append({"pairs": int(X.shape[0]), "loss": float(np.mean(L[-50:]))})
        if log:
            log(f"  round {r}: {X.shape[0]} pairs, loss {info['rounds'][-1]['loss']:.3f}")
    info["seconds"] = time.perf_counter() - t0
    info["media_per_round"], info["grid"], info["steps_per_round"] = media, n, steps
    info["dilation"], info["iterations_sampled"] = list(net.dilation()), list(iters)
    return net, info


def train(corr, X, T, steps=1500, batch=8, lr=1e-3, seed=0, log=None):
    """Adam on random minibatches (whole samples); returns the loss history."""
    rng = np.random.default_rng(seed)
    shape = X.shape[1:4]
    losses = []
    dx = DeviceArray32(np.zeros((batch,) + X.shape[1:], np.float32))
    dt = DeviceArray32(np.zeros((batch,) + T.shape[1:], np.float32))
    for s in range(steps):
        idx = rng.choice(X.shape[0], batch, replace=False)
        dx.upload(X[idx])
        dt.upload(T[idx])
        lr_s = lr * (0.1 if s > 0.8 * steps else 1.0)
        losses.append(corr.train_step(dx, dt, (batch,) + shape, lr_s))
        if log and (s + 1) % 250 == 0:
            log(f"  step {s + 1}: loss {np.mean(losses[-50:]):.4f}")
    return np.array(losses)


def conv_flops(shape, batch=1):
    """Multiply-add FLOPs (2 per MAC) of one forward pass on a (Z, Y, X) grid."""
    vox = batch * int(np.prod(shape))
    ch = [3, 32, 32, 32, 32, 1]
    return 2 * vox * 27 * sum(ch[l] * ch[l + 1] for l in range(5))


def main(argv=None):
    """Time the network kernels: training steps on batches of 32^3 grids and
    inference on an n^3 grid, with the MFMA rate they reach."""
    import argparse

    from . import synchronize
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--n", type=int, default=64, help="inference grid n^3")
    ap.add_argument("--batch", type=int, default=8, help="training batch of 32^3 grids")
    ap.add_argument("--train-steps", type=int, default=20)
    ap.add_argument("--infer", type=int, default=20)
    a = ap.parse_args(argv)
    rng = np.random.default_rng(0)
    net = FluxCorrector(0)
    tshape = (a.batch, 32, 32, 32)
    dx = DeviceArray32(rng.normal(size=tshape + (3,)).astype(np.float32))
    dt = DeviceArray32(rng.normal(size=tshape).astype(np.float32))
    net.train_step(dx, dt, tshape, 1e-3)
    synchronize()
    t = time.perf_counter()
    for _ in range(a.train_steps):
        net.train_step(dx, dt, tshape, 1e-3)
    synchronize()
    t_train = (time.perf_counter() - t) / a.train_steps
    ishape = (1, a.n, a.n, a.n)
    ix = DeviceArray32(rng.normal(size=ishape + (3,)).astype(np.float32))
    iout = DeviceArray32(np.zeros(ishape, np.float32))
    lib = net._lib
    check(lib.allium_cnn_forward_device(net.handle, 1, a.n, a.n, a.n, ix.ptr, iout.ptr, None))
    synchronize()
    t = time.perf_counter()
    for _ in range(a.infer):
        check(lib.allium_cnn_forward_device(net.handle, 1, a.n, a.n, a.n, ix.ptr, iout.ptr, None))
    synchronize()
    t_inf = (time.perf_counter() - t) / a.infer
    f_inf = conv_flops((a.n,) * 3)
    f_train = 3 * conv_flops((32, 32, 32), a.batch)   # forward + backward-data + weight gradients
    print({"train_step_ms": t_train * 1e3, "train_tflops": f_train / t_train / 1e12,
           "inference_ms": t_inf * 1e3, "inference_tflops": f_inf / t_inf / 1e12,
           "inference_grid": a.n, "train_batch": a.batch}, flush=True)


if __name__ == "__main__":
    main()
