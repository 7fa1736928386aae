"""GPU checks of the learned flux correction network (allium_amd/csrc/cnn.hip).

The BF16 MFMA forward and backward kernels are compared with the FP64 CPU
oracle (oracle/cnn_oracle.py) run with the same bfloat16 rounding points;
the remaining differences come from FP32 versus FP64 accumulation order.
Tolerances: forward 5e-3 and per-layer gradients 2e-2 relative (L2).
"""
import numpy as np
import pytest

import allium_amd as A
from allium_amd import cnn
from allium_amd import problems as P
from oracle import cnn_oracle as C

pytestmark = pytest.mark.gpu


def _random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    params = C.init_params(seed)
    return [(W * scale, rng.normal(0, 0.05, b.shape)) for W, b in params]


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.mark.parametrize("shape", [(2, 5, 6, 7), (1, 4, 4, 33)])
def test_forward_matches_oracle(shape):
    print(f"cnn forward {shape}", flush=True)
    rng = np.random.default_rng(3)
    params = _random_params(4)
    net = cnn.FluxCorrector(0)
    net.set_params(C.pack(params).astype(np.float32))
    # the device keeps FP32 master weights: the oracle uses the same FP32 values
    p32 = C.unpack(C.pack(params).astype(np.float32).astype(np.float64))
    x = rng.normal(size=shape + (3,)).astype(np.float32)
    ref, _ = C.forward(p32, x.astype(np.float64), bf16=True)
    got = net.forward(x)
    assert _rel(got, ref) < 5e-3


def test_gradients_match_oracle():
    print("cnn gradients", flush=True)
    rng = np.random.default_rng(5)
    params = _random_params(6)
    flat32 = C.pack(params).astype(np.float32)
    p32 = C.unpack(flat32.astype(np.float64))
    net = cnn.FluxCorrector(0)
    net.set_params(flat32)
    shape = (2, 4, 5, 6)
    x = rng.normal(size=shape + (3,)).astype(np.float32)
    t = rng.normal(size=shape).astype(np.float32)
    loss_ref, grads_ref = C.loss_and_grad(p32, x.astype(np.float64), t.astype(np.float64), bf16=True)
    loss, g = net.gradient(x, t)
    assert abs(loss - loss_ref) < 1e-2 * abs(loss_ref)
    got = C.unpack(g.astype(np.float64))
    for l, ((gW, gb), (rW, rb)) in enumerate(zip(got, grads_ref)):
        eW, eb = _rel(gW, rW), _rel(gb, rb)
        print(f"  layer {l}: weights {eW:.2e} bias {eb:.2e}")
        assert eW < 2e-2 and eb < 2e-2


DILATIONS = [(1, 2, 4, 8, 1), (2, 1, 3, 1, 2)]


@pytest.mark.parametrize("dil", DILATIONS)
def test_dilated_forward_and_gradients_match_oracle(dil):
    print(f"cnn dilation {dil}", flush=True)
    rng = np.random.default_rng(15)
    params = _random_params(16)
    flat32 = C.pack(params).astype(np.float32)
    p32 = C.unpack(flat32.astype(np.float64))
    net = cnn.FluxCorrector(0, dilation=dil)
    assert net.dilation() == dil
    net.set_params(flat32)
    shape = (2, 9, 10, 11)
    x = rng.normal(size=shape + (3,)).astype(np.float32)
    t = rng.normal(size=shape).astype(np.float32)
    ref, _ = C.forward(p32, x.astype(np.float64), bf16=True, dil=list(dil))
    assert _rel(net.forward(x), ref) < 5e-3
    loss_ref, grads_ref = C.loss_and_grad(p32, x.astype(np.float64), t.astype(np.float64), bf16=True, dil=list(dil))
    loss, g = net.gradient(x, t)
    assert abs(loss - loss_ref) < 1e-2 * abs(loss_ref)
    for l, ((gW, gb), (rW, rb)) in enumerate(zip(C.unpack(g.astype(np.float64)), grads_ref)):
        eW, eb = _rel(gW, rW), _rel(gb, rb)
        print(f"  layer {l}: weights {eW:.2e} bias {eb:.2e}")
        assert eW < 2e-2 and eb < 2e-2


def test_dilation_errors():
    net = cnn.FluxCorrector(0)
    assert net.dilation() == (1, 1, 1, 1, 1)
    with pytest.raises(A.AlliumError):
        net.set_dilation((1, 2, 9, 1, 1))
    with pytest.raises(A.AlliumError):
        net.set_dilation((0, 1, 1, 1, 1))


def test_training_reduces_loss():
    """Adam on the GPU fits a smoothed linear map of the first input channel."""
    print("cnn training", flush=True)
    rng = np.random.default_rng(7)
    shape = (4, 8, 8, 8)
    x = rng.normal(size=(32,) + shape[1:] + (3,)).astype(np.float32)
    t = (3.0 * x[..., 0] + 0.5 * np.roll(x[..., 0], 1, axis=3)).astype(np.float32)
    net = cnn.FluxCorrector(11)
    l0 = net.gradient(x[:4], t[:4])[0]
    losses = cnn.train(net, x, t, steps=300, batch=4, lr=2e-3)
    l1 = net.gradient(x[:4], t[:4])[0]
    print(f"  loss {l0:.3f} -> {l1:.3f}")
    assert np.all(np.isfinite(losses)) and l1 < 0.2 * l0


def test_bad_network_falls_back_and_converges():
    """A network whose correction is far off makes the residual grow; the
    source iteration switches it off and still reaches the plain solution."""
    print("cnn fallback", flush=True)
    cfg = P.problem_3d_hetero(n=12, seed=2, log_sd=1.0, c_const=0.99)
    Linv = A.InverseAdvectionOperator(A.Mesh(cfg["nel"]), 1, 4, cfg["sigma_t"])
    q = Linv.form_source(Q=1.0)
    phi0, _, it0, _ = A.source_iteration(Linv, cfg["sigma_s"], q, abstol=1e-10, maxit=2000, want_psi=False)
    params = _random_params(9)
    W, b = params[-1]
    params[-1] = (W * 200.0, b + 50.0)
    net = cnn.FluxCorrector(0)
    net.set_params(C.pack(params).astype(np.float32))
    phi1, _, it1, norms = A.source_iteration(Linv, cfg["sigma_s"], q, abstol=1e-10, maxit=2000, want_psi=False,
                                             accel=net)
    print(f"  plain {it0} its, with a bad network {it1} its")
    assert norms[-1] < 1e-10
    assert _rel(phi1, phi0) < 1e-8


def test_trained_network_reduces_iterations():
    """Train on pairs from seeded 16^3 media (c = 0.9999, log-sd 1.5): plain
    iterations first, then iterations that use the network; solve a held-out
    medium with and without the correction: same converged flux, fewer
    iterations with it."""
    print("cnn gain", flush=True)
    net, info = cnn.fit(n=16, media=24, rounds=2, steps=400, seed0=500, sn=4)
    print("  training:", info)
    cfg = P.problem_3d_hetero(n=16, seed=9999, log_sd=1.5, c_const=0.9999)
    Linv = A.InverseAdvectionOperator(A.Mesh(cfg["nel"]), 1, 4, cfg["sigma_t"])
    q = Linv.form_source(Q=1.0)
    phi0, _, it0, _ = A.source_iteration(Linv, cfg["sigma_s"], q, abstol=1e-8, maxit=2000, want_psi=False)
    phi1, _, it1, n1 = A.source_iteration(Linv, cfg["sigma_s"], q, abstol=1e-8, maxit=2000, want_psi=False,
                                          accel=net)
    print(f"  held-out medium: plain {it0} iterations, learned correction {it1}")
    assert n1[-1] < 1e-8
    assert _rel(phi1, phi0) < 1e-6
    assert it1 < 0.9 * it0
