"""GPU checks of two benchmark configurations at their own sizes:

* the uneven batch (12 independent 3D problems, jittered hexes, DG p = 2,
  S12, one launch per level of every instance): a small batch against the
  oracle at 1e-12, and the benchmark's batch (n^3, n drawn from {16..48})
  through properties that hold at any size: L(L^{-1} q) = q for every
  instance (test/test_sweep.cpp:289-361) and the particle balance
  absorption + leakage = sum_a w_a sum q (src/sweep.cpp:1400-1462);
* the learned-correction medium (64^3 hex, p = 1, S8, c = 0.9999, lognormal
  sigma_t with log-sd 1.5, Q = 1): the learned source iteration reaches plain
  SI's scalar flux in fewer iterations, and the FP32-refined source
  iteration reaches the FP64 one.
"""
import numpy as np
import pytest

import allium_amd as A
import oracle
from allium_amd import cnn
from allium_amd import problems as P
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def _instances(ns, seed, jitter=0.05):
    ops, verts, sigs = [], [], []
    for i, n in enumerate(ns):
        nel = (n, n, n)
        r = np.random.default_rng(seed * 1000 + i)
        v = P.jittered_vertices(nel, jitter, r)
        sig = P.lognormal_sigma(n ** 3, 1.0, r).reshape(1, -1)
        ops.append(A.InverseAdvectionOperator(A.Mesh(nel, vertices=v), 2, 12, sig))
        verts.append(v)
        sigs.append(sig)
    return ops, verts, sigs


def test_config3_small_batch_matches_oracle():
    """12 jittered (0.05h) p = 2 S12 instances of uneven size n^3, n in {3..6},
    swept in one batched launch per level: every instance's psi equals the
    oracle's to 1e-12."""
    print("config 3: small uneven batch vs oracle", flush=True)
    rng = np.random.default_rng(7)
    ns = [int(v) for v in rng.integers(3, 7, 12)]
    ops, verts, sigs = _instances(ns, seed=7)
    isos = [A.DeviceVector.from_host(np.ones(op.phi_size)) for op in ops]
    psis = [A.DeviceVector(op.size) for op in ops]
    A.SweepBatch(ops).Mult(None, psis, isos)
    om, w = ops[0].quadrature()
    worst = 0.0
    for n, op, v, sig, d in zip(ns, ops, verts, sigs, psis):
        pr = oracle.Problem(3, (n, n, n), v, 2, om, w, sig)
        ref = pr.sweep(np.ones(pr.size))
        worst = max(worst, rel_err(d.to_host(), ref))
    print(f"  meshes {ns}: worst rel err {worst:.2e}")
    assert worst < 1e-12


def test_config3_benchmark_batch_properties():
    """The benchmark's batch (seed 0: n = 44, 37, 32, 24, 26, 17, 18, 16, 21,
    42, 37, 46; 2.02e9 angular unknowns per batched sweep) with an isotropic
    unit source: L psi_i = q_i for every instance and its balance equals
    sum_a w_a sum q = 4 pi Ndof, both to 1e-11."""
    print("config 3: benchmark batch properties", flush=True)
    rng = np.random.default_rng(0)
    ns = [int(v) for v in rng.integers(16, 49, 12)]
    ops, _, _ = _instances(ns, seed=0)
    isos = [A.DeviceVector.from_host(np.ones(op.phi_size)) for op in ops]
    psis = [A.DeviceVector(op.size) for op in ops]
    A.SweepBatch(ops).Mult(None, psis, isos)
    print(f"  swept {sum(op.size for op in ops)} unknowns, meshes {ns}", flush=True)
    for op, d in zip(ops, psis):
        L = A.AdvectionOperator(op)
        y = A.DeviceVector(op.size)
        L.Mult(d, y)
        err = float(np.abs(y.to_host() - 1.0).max())
        del y
        bal = L.ComputeBalance(d)
        expect = 4.0 * np.pi * op.Ndof
        print(f"  n={op.mesh.nel[0]}: max |L psi - q| {err:.2e}, balance rel err {abs(bal - expect) / expect:.2e}",
              flush=True)
        assert err < 1e-11 and abs(bal - expect) < 1e-11 * expect


def test_config5_learned_and_fp32_refined_reach_plain_solution():
    """64^3, p = 1, S8, c = 0.9999, log-sd 1.5, Q = 1, to a 1e-8 FP64 residual:
    a network trained briefly (40 media of 32^3, one round of 800 steps) takes
    fewer source iterations than plain SI and reaches its scalar flux to 1e-9;
    the FP32-refined source iteration reaches the FP64 scalar flux to 1e-9 as
    well.  (Trained on 16^3 media, the network's effect on this medium ranges
    from 52 to 79 iterations with the seed; on 32^3 media, 55 to 69.)"""
    print("config 5: learned and FP32-refined vs plain", flush=True)
    net, info = cnn.fit(n=32, media=40, rounds=1, steps=800, seed0=500, sn=8)
    print(f"  trained in {info['seconds']:.1f}s", flush=True)
    cfg = P.problem_3d_hetero(n=64, seed=0, log_sd=1.5, c_const=0.9999)
    Linv = A.InverseAdvectionOperator(A.Mesh(cfg["nel"]), 1, 8, cfg["sigma_t"])
    q = A.DeviceVector.from_host(Linv.form_source(Q=1.0))
    phi0, _, it0, n0 = A.source_iteration(Linv, cfg["sigma_s"], q, abstol=1e-8, maxit=1000, want_psi=False)
    phi1, _, it1, n1 = A.source_iteration(Linv, cfg["sigma_s"], q, abstol=1e-8, maxit=1000, want_psi=False,
                                          accel=net)
    phi2, _, it2, n2 = A.source_iteration(Linv, cfg["sigma_s"], q, abstol=1e-8, maxit=1000, want_psi=False,
                                          precision=A.FP32_REFINED)
    e1 = float(np.linalg.norm(phi1 - phi0) / np.linalg.norm(phi0))
    e2 = float(np.linalg.norm(phi2 - phi0) / np.linalg.norm(phi0))
    print(f"  plain {it0} its, learned {it1} its (rel diff {e1:.1e}), FP32-refined {it2} its (rel diff {e2:.1e})")
    assert n0[it0 - 1] < 1e-8 and n1[it1 - 1] < 1e-8 and n2[it2 - 1] < 1e-8
    assert it1 < it0 and e1 < 1e-9 and e2 < 1e-9
