"""GPU parity of the HIP sweep against the CPU oracle, through the C ABI.

Bar: FP64 sweeps within 1e-12 relative of the oracle;
restated reference tests (psi == 1, L(L^-1 x) = x) on the GPU path; the
FP32-refined path within its stated bound.  Sizes are those the oracle runs in
seconds; the full 64^3 workload is covered by size-independent properties.
"""
import numpy as np
import pytest

import allium_amd as A
import oracle
from allium_amd import problems as P
from tests.helpers import ls_fixture, rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _pair(nel, p, sn, sigma_t, vertices=None, bc=None):
    dim = len(nel)
    v = P.cartesian_vertices(nel) if vertices is None else vertices
    m = A.Mesh(nel, vertices=v, bc=bc)
    Linv = A.InverseAdvectionOperator(m, p, sn, sigma_t)
    om, w = ls_fixture(sn, dim)
    pr = oracle.Problem(dim, nel, v, p, om, w, sigma_t, bc=bc)
    return Linv, pr


CASES = [
    # cfg1-style 2D: 32x32 quads, p=1, S8, sigma_t = 1
    ("2d_cfg1", (32, 32), 1, 8, None, 0.0),
    ("2d_p2_jitter", (12, 10), 2, 6, None, 0.2),
    ("2d_p0", (16, 16), 0, 4, None, 0.0),
    ("3d_hetero", (12, 12, 12), 1, 8, "lognormal", 0.0),
    ("3d_jitter", (8, 7, 6), 1, 6, "lognormal", 0.2),
    ("3d_p0", (10, 10, 10), 0, 4, None, 0.0),
    ("3d_reflect", (6, 6, 6), 1, 4, None, 0.0),
]


PERSISTENT_OFF = ("persistent kernel (launch mode 2) disabled: its first MI355X run did not terminate "
                  "and the cause is not identified")


@pytest.mark.parametrize("name,nel,p,sn,sig,jitter", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("mode", [0, 1, pytest.param(2, marks=pytest.mark.skip(reason=PERSISTENT_OFF)), 3, 4, 5])
def test_sweep_matches_oracle(name, nel, p, sn, sig, jitter, mode):
    print(f"case {name} mode {mode}", flush=True)
    dim = len(nel)
    ne = int(np.prod(nel))
    rng = np.random.default_rng(17)
    sigma = P.lognormal_sigma(ne, 1.0, rng).reshape(1, ne) if sig == "lognormal" else np.ones((1, ne))
    v = P.jittered_vertices(nel, jitter, rng) if jitter else None
    bc = (1, 0, 0, 0, 0, 0) if name == "3d_reflect" else None
    Linv, pr = _pair(nel, p, sn, sigma, vertices=v, bc=bc)
    if mode in (3, 4, 5) and bc is not None:
        with pytest.raises(A.AlliumError, match="reflective"):
            Linv.set_launch(mode)
        return
    if mode in (4, 5) and jitter:
        with pytest.raises(A.AlliumError, match="uniform geometry"):
            Linv.set_launch(mode)
        return
    Linv.set_launch(mode)
    q = pr.form_source(Q=1.0) + rng.uniform(0, 0.1, pr.size)
    ref = pr.sweep(q)
    got = Linv.Mult(q, np.empty_like(q))
    assert rel_err(got, ref) < FP64_TOL
    # in place (Mult(psi, psi), src/transport_op.hpp:28)
    buf = q.copy()
    Linv.Mult(buf, buf)
    assert rel_err(buf, ref) < FP64_TOL


@pytest.mark.parametrize("lanes", [0, 1, 2, 4])
@pytest.mark.parametrize("block", [64, 256])
@pytest.mark.parametrize("name", ["2d_cfg1", "3d_hetero", "3d_jitter", "3d_reflect"])
def test_lanes_per_cell_match_oracle(name, lanes, block):
    """Every (lanes per cell, workgroup size) variant the autotuner may pick."""
    print(f"case {name} lanes {lanes} block {block}", flush=True)
    _, nel, p, sn, sig, jitter = [c for c in CASES if c[0] == name][0]
    ne = int(np.prod(nel))
    rng = np.random.default_rng(31)
    sigma = P.lognormal_sigma(ne, 1.0, rng).reshape(1, ne) if sig == "lognormal" else np.ones((1, ne))
    v = P.jittered_vertices(nel, jitter, rng) if jitter else None
    bc = (1, 0, 0, 0, 0, 0) if name == "3d_reflect" else None
    Linv, pr = _pair(nel, p, sn, sigma, vertices=v, bc=bc)
    if lanes == 0 and jitter:
        with pytest.raises(A.AlliumError, match="uniform geometry"):
            Linv.set_tuning(0, block, lanes)
        return
    Linv.set_tuning(0, block, lanes)
    q = pr.form_source(Q=1.0) + rng.uniform(0, 0.1, pr.size)
    ref = pr.sweep(q)
    assert rel_err(Linv.Mult(q, np.empty_like(q)), ref) < FP64_TOL
    Linv.set_precision(A.FP32_REFINED)
    err32 = rel_err(Linv.Mult(q, np.empty_like(q)), ref)
    print(f"  fp32 rel err {err32:.2e}")
    assert err32 < 1e-4


@pytest.mark.parametrize("block", [64, 256, 512])
def test_direction_sweeps_block_sizes(block):
    nel = (14, 13, 12)
    ne = 14 * 13 * 12
    rng = np.random.default_rng(41)
    sigma = P.lognormal_sigma(ne, 1.0, rng).reshape(1, ne)
    # jitter 0.1h keeps every direction's sweep graph acyclic (0.2h does not)
    for verts in (None, P.jittered_vertices(nel, 0.1, rng)):
        Linv, pr = _pair(nel, 1, 8, sigma, vertices=verts)
        Linv.set_tuning(3, block, 1)
        q = pr.form_source(Q=1.0)
        assert rel_err(Linv.Mult(q, np.empty_like(q)), pr.sweep(q)) < FP64_TOL
        Linv.set_precision(A.FP32_REFINED)
        assert rel_err(Linv.Mult(q, np.empty_like(q)), pr.sweep(q)) < 1e-4


BOX_TILINGS = [
    # (mesh, p, sn, box extents, directions per workgroup, threads)
    ((14, 13, 12), 1, 8, (4, 5, 3), 3, 64),
    ((14, 13, 12), 1, 8, (8, 8, 8), 5, 256),
    ((14, 13, 12), 1, 8, (16, 16, 16), 2, 128),
    ((14, 13, 12), 1, 8, (1, 1, 1), 10, 128),
    ((14, 13, 12), 1, 8, (8, 8, 8), 5, 1024),
    ((14, 13, 12), 1, 8, (6, 7, 5), 10, 512),
    ((33, 31), 1, 8, (16, 16), 10, 1024),
    ((7, 6, 5), 0, 6, (3, 2, 4), 4, 192),
    ((9, 7), 2, 8, (4, 3), 3, 64),
    ((33, 31), 1, 8, (16, 16), 10, 256),
    ((12, 10), 0, 4, (5, 2), 1, 64),
    ((20, 18, 17), 1, 8, (16, 16, 8), 1, 512),
    ((33, 31), 1, 8, (16, 16), 1, 256),
    ((14, 13, 12), 1, 8, (8, 8, 8), 1, 1024),
]


@pytest.mark.parametrize("nel,p,sn,box,nd,threads", BOX_TILINGS,
                         ids=[f"{'x'.join(map(str, c[0]))}_p{c[1]}_b{'x'.join(map(str, c[3]))}_d{c[4]}_t{c[5]}"
                              for c in BOX_TILINGS])
def test_box_tilings_match_oracle(nel, p, sn, box, nd, threads):
    """Box sweeps on ragged meshes (boxes that do not divide the mesh), every
    directions-per-workgroup split and workgroup size, FP64 and FP32."""
    print(f"box case {nel} p{p} S{sn} box {box} nd {nd} threads {threads}", flush=True)
    ne = int(np.prod(nel))
    rng = np.random.default_rng(53)
    sigma = P.lognormal_sigma(ne, 1.0, rng).reshape(1, ne)
    Linv, pr = _pair(nel, p, sn, sigma)
    q = pr.form_source(Q=1.0, inflow=np.ones((1, 2 * len(nel)))) + rng.uniform(0, 0.1, pr.size)
    ref = pr.sweep(q)
    for pf in ((0, 2, 3) if p == 1 else (0, 2)):
        if len(nel) == 3:
            Linv.set_box_tiling(*box, dirs_per_group=nd, threads=threads, prefetch=pf)
        else:
            Linv.set_box_tiling(box[0], box[1], 1, dirs_per_group=nd, threads=threads, prefetch=pf)
        for mode in (4, 5):
            Linv.set_launch(mode)
            assert rel_err(Linv.Mult(q, np.empty_like(q)), ref) < FP64_TOL
        buf = q.copy()
        Linv.Mult(buf, buf)   # in place: prefetched sources of later cells are not overwritten
        assert rel_err(buf, ref) < FP64_TOL
    Linv.set_precision(A.FP32_REFINED)
    assert rel_err(Linv.Mult(q, np.empty_like(q)), ref) < 1e-4


MOMENT_CASES = [BOX_TILINGS[1], BOX_TILINGS[6], BOX_TILINGS[7], BOX_TILINGS[8]]


@pytest.mark.parametrize("nel,p,sn,box,nd,threads", MOMENT_CASES,
                         ids=[f"{'x'.join(map(str, c[0]))}_p{c[1]}_b{'x'.join(map(str, c[3]))}"
                              for c in MOMENT_CASES])
def test_sweep_mult_moments(nel, p, sn, box, nd, threads):
    """allium_sweep_mult_moments_device (the sweep and DiscreteToMoment pair of
    TransportOperator::Mult) gives the psi of the sweep and the phi of the
    moment kernel bit for bit, for a psi-sized source, an isotropic source and
    both, in box and level launch modes."""
    print(f"sweep + moments {nel} p{p} S{sn} box {box}", flush=True)
    ne = int(np.prod(nel))
    rng = np.random.default_rng(61)
    sigma = P.lognormal_sigma(ne, 1.0, rng).reshape(1, ne)
    Linv, pr = _pair(nel, p, sn, sigma)
    q = pr.form_source(Q=1.0) + rng.uniform(0, 0.1, pr.size)
    iso = rng.uniform(0.5, 1.5, Linv.phi_size)
    dq, diso = A.DeviceVector.from_host(q), A.DeviceVector.from_host(iso)
    psi0, psi1 = A.DeviceVector(Linv.size), A.DeviceVector(Linv.size)
    phi0, phi1 = A.DeviceVector(Linv.phi_size), A.DeviceVector(Linv.phi_size)
    D = A.DiscreteToMoment(Linv)
    ref = pr.D(pr.sweep(q))
    if len(nel) == 3:
        Linv.set_box_tiling(*box, dirs_per_group=nd, threads=threads, prefetch=3 if p == 1 else 2)
    else:
        Linv.set_box_tiling(box[0], box[1], 1, dirs_per_group=nd, threads=threads, prefetch=2)
    mult_iso = A._lib.load().allium_sweep_mult_iso_device
    for mode in (4, 5, 0):
        Linv.set_launch(mode)
        for src, isov in ((dq, None), (None, diso), (dq, diso)):
            if isov is None:
                Linv.Mult(src, psi0)
            else:
                A.check(mult_iso(Linv._h, None if src is None else src.ptr, isov.ptr, psi0.ptr, None))
            D.Mult(psi0, phi0)
            Linv.MultMoments(src, psi1, phi1, iso=isov)
            A.synchronize()
            assert np.array_equal(psi0.to_host(), psi1.to_host()), (mode, src is None)
            assert np.array_equal(phi0.to_host(), phi1.to_host()), (mode, src is None)
        Linv.MultMoments(dq, psi1, phi1)
        assert rel_err(phi1.to_host(), ref) < FP64_TOL
    with pytest.raises(A.AlliumError):
        Linv.MultMoments(dq, psi1, psi1)


def test_box_tiling_errors():
    nel = (16, 16, 16)
    Linv, _ = _pair(nel, 1, 8, np.ones((1, 4096)))
    with pytest.raises(A.AlliumError, match="LDS"):
        Linv.set_box_tiling(16, 16, 16, dirs_per_group=10, threads=256)
    with pytest.raises(A.AlliumError, match="workgroup size"):
        Linv.set_box_tiling(8, 8, 8, dirs_per_group=5, threads=1280)
    Lp0, _ = _pair((8, 8, 8), 0, 8, np.ones((1, 512)))
    with pytest.raises(A.AlliumError, match="workgroup size"):
        Lp0.set_box_tiling(8, 8, 8, dirs_per_group=5, threads=512)
    Linv.set_box_tiling(8, 8, 8, dirs_per_group=5, threads=128)
    assert Linv.box_tiling() == {"bx": 8, "by": 8, "bz": 8, "dirs_per_group": 5, "threads": 128, "prefetch": 2}
    with pytest.raises(A.AlliumError, match="prefetch"):
        Linv.set_box_tiling(8, 8, 8, dirs_per_group=5, threads=128, prefetch=1)


def test_full_size_box_sweep_properties():
    """64^3 hex, p=1, S8, lognormal sigma_t: size-independent checks at the
    benchmark size.  The box sweep, the level waves and the direction sweeps
    agree; L (L^{-1} q) = q through the independently assembled streaming
    operator; particle balance equals the source (src/sweep.cpp:1400-1462)."""
    nel = (64, 64, 64)
    ne = 64 ** 3
    sigma = P.lognormal_sigma(ne, 1.0, np.random.default_rng(0)).reshape(1, ne)
    Linv = A.InverseAdvectionOperator(A.Mesh(nel), 1, 8, sigma)
    q = Linv.form_source(Q=1.0)
    dq = A.DeviceVector.from_host(q)
    out = {}
    for mode in (5, 0, 3):
        Linv.set_launch(mode)
        dpsi = A.DeviceVector(q.size)
        Linv.Mult(dq, dpsi)
        out[mode] = dpsi
    ref = out[0].to_host()
    for mode in (5, 3):
        assert rel_err(out[mode].to_host(), ref) < FP64_TOL
    L = A.AdvectionOperator(Linv)
    dLpsi = A.DeviceVector(q.size)
    L.Mult(out[5], dLpsi)
    assert rel_err(dLpsi.to_host(), q) < 1e-12
    total_source = 4 * np.pi * 1.0   # Q = 1 per steradian over the unit cube
    assert abs(L.ComputeBalance(out[5]) - total_source) < 1e-10 * total_source


def test_autotune_picks_a_valid_variant():
    nel = (16, 16, 16)
    ne = 4096
    rng = np.random.default_rng(5)
    sigma = P.lognormal_sigma(ne, 1.0, rng).reshape(1, ne)
    Linv, pr = _pair(nel, 1, 8, sigma)
    mode = Linv.set_launch(-1)
    t = Linv.tuning()
    print("autotune picked", t, Linv.box_tiling() if mode >= 4 else "")
    assert mode in (0, 1, 3, 4, 5) and t["lanes"] in (0, 1, 2, 4)
    q = pr.form_source(Q=1.0)
    assert rel_err(Linv.Mult(q, np.empty_like(q)), pr.sweep(q)) < FP64_TOL


def test_multigroup_time_absorption_matches_oracle():
    nel, ne, G = (6, 5, 4), 120, 3
    rng = np.random.default_rng(3)
    sigma = rng.uniform(0.1, 5.0, (G, ne))
    Linv, pr = _pair(nel, 1, 6, sigma)
    Linv.SetTimeAbsorption(7.5)
    pr.set_time_absorption(7.5)
    q = pr.form_source(Q=rng.uniform(0.5, 1.5, (G, ne)), inflow=np.ones((G, 6)))
    for mode in (0, 1, 3, 4, 5):
        Linv.set_launch(mode)
        assert rel_err(Linv.Mult(q, np.empty_like(q)), pr.sweep(q)) < FP64_TOL
    # new cross sections (AssembleLocalMatrices)
    sigma2 = rng.uniform(0.1, 5.0, (G, ne))
    Linv.AssembleLocalMatrices(sigma2)
    pr.set_sigma(sigma2)
    assert rel_err(Linv.Mult(q, np.empty_like(q)), pr.sweep(q)) < FP64_TOL


@pytest.mark.parametrize("nel,p", [((10, 10), 1), ((10, 10), 2), ((6, 6, 6), 1), ((10, 10, 10), 0)])
def test_constant_solution_gpu(nel, p):
    # test/test_sweep.cpp:7-108 on the GPU path (two groups, sigma = 0)
    dim = len(nel)
    ne = int(np.prod(nel))
    m = A.Mesh(nel)
    Linv = A.InverseAdvectionOperator(m, p, 4, np.zeros((2, ne)))
    q = Linv.form_source(Q=0.0, inflow=np.ones((2, 2 * dim)))
    psi = Linv.Mult(q, np.empty_like(q))
    assert np.abs(psi - 1.0).max() < 1e-10


@pytest.mark.parametrize("nel", [(3, 3), (3, 3, 3), (9, 8, 7)])
def test_L_Linv_gpu(nel):
    # test/test_sweep.cpp:289-361 through the GPU L and L^-1
    ne = int(np.prod(nel))
    m = A.Mesh(nel)
    Linv = A.InverseAdvectionOperator(m, 1, 4, np.vstack([np.ones(ne), 0.5 * np.ones(ne)]))
    L = A.AdvectionOperator(Linv)
    x = np.random.default_rng(12345).uniform(-1, 1, Linv.size)
    y = L.Mult(x, np.empty_like(x))
    z = Linv.Mult(y, y)
    assert np.linalg.norm(z - x) < 1e-12 * max(1.0, np.sqrt(x.size / 1000))


def test_apply_L_and_balance_match_oracle():
    nel = (7, 6, 5)
    ne = 210
    rng = np.random.default_rng(9)
    sigma = P.lognormal_sigma(ne, 1.0, rng).reshape(1, ne)
    v = P.jittered_vertices(nel, 0.2, rng)
    for verts in (None, v):
        Linv, pr = _pair(nel, 1, 6, sigma, vertices=verts)
        x = rng.normal(size=pr.size)
        L = A.AdvectionOperator(Linv)
        assert rel_err(L.Mult(x, np.empty_like(x)), pr.apply_L(x)) < 1e-13
        q = pr.form_source(Q=1.0)
        psi = pr.sweep(q)
        assert abs(L.ComputeBalance(psi) - pr.balance(psi)) < 1e-12 * abs(pr.balance(psi))
        assert abs(L.ComputeBalance(psi) - 4 * np.pi) < 1e-10


def test_device_pointer_api_and_moments():
    nel = (10, 9, 8)
    ne = 720
    rng = np.random.default_rng(21)
    sigma = P.lognormal_sigma(ne, 1.0, rng).reshape(1, ne)
    Linv, pr = _pair(nel, 1, 8, sigma)
    q = pr.form_source(Q=1.0)
    dq = A.DeviceVector.from_host(q)
    dpsi = A.DeviceVector(q.size)
    Linv.Mult(dq, dpsi)
    ref = pr.sweep(q)
    assert rel_err(dpsi.to_host(), ref) < FP64_TOL
    D = A.DiscreteToMoment(Linv)
    dphi = A.DeviceVector(Linv.phi_size)
    D.Mult(dpsi, dphi)
    assert rel_err(dphi.to_host(), pr.D(ref)) < 1e-14
    D.MultTranspose(dphi, dq)
    assert rel_err(dq.to_host(), pr.DT(pr.D(ref))) < 1e-14


def test_source_iteration_matches_oracle():
    # cfg1: 2D 32x32, p=1, S8, sigma_t = 1, c = 0.9, Q = 1, vacuum, SI to 1e-8
    d = P.problem_2d_cfg1()
    Linv, pr = _pair(d["nel"], 1, 8, d["sigma_t"])
    q = pr.form_source(Q=1.0)
    phi_o, psi_o, it_o, norms_o = pr.source_iteration(d["sigma_s"], q, abstol=1e-8, maxit=400)
    for mode in (0, 1, 3):
        Linv.set_launch(mode)
        phi, psi, it, norms = A.source_iteration(Linv, d["sigma_s"], q, abstol=1e-8, maxit=400)
        assert it == it_o
        np.testing.assert_allclose(norms, norms_o, rtol=1e-9, atol=1e-13 * norms_o[0])
        assert rel_err(phi, phi_o) < 1e-11
        assert rel_err(psi, psi_o) < 1e-11


def test_fp32_refined_sweep_error_bound():
    """FP32 cell assembly and LU (FP64 vectors in and out): error of one sweep
    vs the FP64 oracle on the 3D lognormal medium, printed and bounded."""
    nel = (12, 12, 12)
    ne = 1728
    rng = np.random.default_rng(23)
    sigma = P.lognormal_sigma(ne, 1.0, rng).reshape(1, ne)
    Linv, pr = _pair(nel, 1, 8, sigma)
    q = pr.form_source(Q=1.0)
    ref = pr.sweep(q)
    Linv.set_precision(A.FP32_REFINED)
    got = Linv.Mult(q, np.empty_like(q))
    err = rel_err(got, ref)
    print(f"fp32-refined sweep rel err {err:.3e}")
    assert err < 1e-5


def test_fp32_refined_source_iteration_reaches_fp64_solution():
    """Defect-correction SI with FP32 sweeps converges to the FP64 SI solution."""
    nel = (10, 10, 10)
    ne = 1000
    d = P.problem_3d_hetero(n=10, seed=4)
    Linv, pr = _pair(nel, 1, 6, d["sigma_t"])
    q = pr.form_source(Q=1.0)
    phi_o, psi_o, it_o, _ = pr.source_iteration(d["sigma_s"], q, abstol=1e-10, maxit=400)
    phi, psi, it, norms = A.source_iteration(Linv, d["sigma_s"], q, abstol=1e-10, maxit=400,
                                             precision=A.FP32_REFINED)
    assert norms[-1] < 1e-10
    assert rel_err(phi, phi_o) < 1e-9
    assert abs(it - it_o) <= 2


def test_fp32_refined_source_iteration_with_reflective_side():
    """With a reflective side the defect q + D^T Ms phi - L psi couples the
    reflected direction exactly as the sweep does, so the FP32-refined SI
    reaches the FP64 oracle's SI solution (6^3, p=1, S4, x-low reflective)."""
    print("fp32-refined SI, reflective side", flush=True)
    nel, ne = (6, 6, 6), 216
    rng = np.random.default_rng(5)
    sigma = P.lognormal_sigma(ne, 1.0, rng).reshape(1, ne)
    sigma_s = rng.uniform(0.5, 0.95, ne) * sigma[0]
    Linv, pr = _pair(nel, 1, 4, sigma, bc=(1, 0, 0, 0, 0, 0))
    q = pr.form_source(Q=1.0)
    phi_o, _, it_o, _ = pr.source_iteration(sigma_s, q, abstol=1e-10, maxit=400)
    phi, _, it, norms = A.source_iteration(Linv, sigma_s, q, abstol=1e-10, maxit=400, precision=A.FP32_REFINED)
    err = rel_err(phi, phi_o)
    print(f"  iterations {it} (oracle FP64 {it_o}), rel err {err:.2e}")
    assert norms[-1] < 1e-10 and err < 1e-9


@pytest.mark.parametrize("mode", [4, 5])
def test_fp32_refined_box_source_iteration_reaches_fp64_solution(mode):
    """With box launches the FP32-refined SI corrects the scattered angular flux in
    place inside the box sweep (the defect formed in FP64 from the new upwind traces
    and the cell's previous value, the correction solved in FP32): it reaches the
    FP64 oracle's SI solution (12^3 hex, p = 1, S8, lognormal medium)."""
    print(f"fp32-refined box SI, mode {mode}", flush=True)
    d = P.problem_3d_hetero(n=12, seed=6)
    Linv, pr = _pair((12, 12, 12), 1, 8, d["sigma_t"])
    Linv.set_box_tiling(4, 4, 4, dirs_per_group=2, threads=256, prefetch=2)
    Linv.set_launch(mode)
    q = pr.form_source(Q=1.0)
    phi_o, psi_o, it_o, _ = pr.source_iteration(d["sigma_s"], q, abstol=1e-10, maxit=400)
    phi, psi, it, norms = A.source_iteration(Linv, d["sigma_s"], q, abstol=1e-10, maxit=400,
                                             precision=A.FP32_REFINED)
    e_phi, e_psi = rel_err(phi, phi_o), rel_err(psi, psi_o)
    print(f"  iterations {it} (oracle FP64 {it_o}), phi rel err {e_phi:.2e}, psi rel err {e_psi:.2e}")
    assert norms[it - 1] < 1e-10 and e_phi < 1e-9 and e_psi < 1e-9
    assert abs(it - it_o) <= 2


def test_fp32_refined_source_iteration_rejects_fixup():
    nel, ne = (4, 4, 4), 64
    Linv, pr = _pair(nel, 1, 4, np.ones((1, ne)))
    Linv.SetFixupOperator("zero")
    q = pr.form_source(Q=1.0)
    with pytest.raises(A.AlliumError, match="fixup"):
        A.source_iteration(Linv, 0.5 * np.ones(ne), q, abstol=1e-8, maxit=50, precision=A.FP32_REFINED)


def test_errors_are_reported_not_aborted():
    m = A.Mesh((4, 4), bc=(1, 1, 0, 0))
    with pytest.raises(A.AlliumError, match="DAG"):
        A.InverseAdvectionOperator(m, 1, 4, np.ones((1, 16)))
    with pytest.raises(A.AlliumError, match="non-finite"):
        A.InverseAdvectionOperator(A.Mesh((4, 4)), 1, 4, np.full((1, 16), np.nan))
    with pytest.raises(A.AlliumError):
        A.InverseAdvectionOperator(A.Mesh((2, 2, 2)), 3, 4, np.ones((1, 8)))
    Linv = A.InverseAdvectionOperator(A.Mesh((4, 4)), 1, 4, np.ones((1, 16)))
    with pytest.raises(ValueError):
        Linv.Mult(np.zeros(3), np.zeros(3))
    with pytest.raises(A.AlliumError, match="disabled"):
        Linv.set_launch(2)
    with pytest.raises(A.AlliumError, match="bad tuning"):
        Linv.set_tuning(0, 512, 1)


P2_CASES = [
    # (mesh, jitter, sn)
    ((5, 4, 4), 0.1, 4),
    ((4, 4, 3), 0.0, 4),
    ((3, 3, 3), 0.1, 6),
]


@pytest.mark.parametrize("nel,jitter,sn", P2_CASES, ids=[f"{'x'.join(map(str, c[0]))}_j{c[1]}_S{c[2]}" for c in P2_CASES])
def test_p2_cooperative_sweep_matches_oracle(nel, jitter, sn):
    """3D p=2 (27 dofs): 32 lanes per cell; FP64 within 1e-12 of the oracle,
    FP32 LU within 1e-4; level launches and their hipGraph."""
    print(f"p2 case {nel} jitter {jitter} S{sn}", flush=True)
    ne = int(np.prod(nel))
    rng = np.random.default_rng(61)
    sigma = P.lognormal_sigma(ne, 1.0, rng).reshape(1, ne)
    v = P.jittered_vertices(nel, jitter, rng) if jitter else None
    Linv, pr = _pair(nel, 2, sn, sigma, vertices=v)
    q = pr.form_source(Q=1.0, inflow=np.ones((1, 6))) + rng.uniform(0, 0.1, pr.size)
    ref = pr.sweep(q)
    for mode in (0, 1):
        Linv.set_launch(mode)
        assert rel_err(Linv.Mult(q, np.empty_like(q)), ref) < FP64_TOL
    with pytest.raises(A.AlliumError, match="cooperative"):
        Linv.set_launch(3)
    Linv.set_precision(A.FP32_REFINED)
    e32 = rel_err(Linv.Mult(q, np.empty_like(q)), ref)
    print(f"  fp32 LU rel err {e32:.2e}")
    assert e32 < 1e-4


@pytest.mark.parametrize("jitter", [0.0, 0.1])
def test_p2_reflective_side_time_absorption_two_groups(jitter):
    """p=2 cooperative cells with a reflective low-x side (upwind traces from
    the reflected direction's own face nodes), time absorption (the time-mass
    row) and two groups, on uniform and jittered geometry: FP64 within 1e-12
    of the oracle in level launches and their hipGraph."""
    print(f"p2 reflective/tau/G=2 jitter {jitter}", flush=True)
    nel, ne = (4, 3, 3), 36
    rng = np.random.default_rng(91)
    sigma = np.vstack([P.lognormal_sigma(ne, 1.0, rng), P.lognormal_sigma(ne, 1.0, rng)])
    v = P.jittered_vertices(nel, jitter, rng) if jitter else None
    Linv, pr = _pair(nel, 2, 4, sigma, vertices=v, bc=(1, 0, 0, 0, 0, 0))
    Linv.SetTimeAbsorption(3.0)
    pr.set_time_absorption(3.0)
    q = pr.form_source(Q=1.0, inflow=np.ones((2, 6))) + rng.uniform(0, 0.1, pr.size)
    ref = pr.sweep(q)
    for mode in (0, 1):
        Linv.set_launch(mode)
        assert rel_err(Linv.Mult(q, np.empty_like(q)), ref) < FP64_TOL
    # FP32 LU: the DPP cells on uniform geometry, the LDS-exchange cells on jittered
    Linv.set_precision(A.FP32_REFINED)
    e32 = rel_err(Linv.Mult(q, np.empty_like(q)), ref)
    print(f"  fp32 LU rel err {e32:.2e}")
    assert e32 < 1e-4


def test_p2_batch_of_uneven_instances():
    """Three uneven jittered p=2 meshes swept in one launch per level: equal to
    sweeping each alone (same kernel) and to the oracle."""
    print("p2 batch", flush=True)
    rng = np.random.default_rng(71)
    ops, prs, qs = [], [], []
    for nel in ((4, 4, 4), (6, 3, 2), (3, 5, 4)):
        ne = int(np.prod(nel))
        sigma = P.lognormal_sigma(ne, 1.0, rng).reshape(1, ne)
        Linv, pr = _pair(nel, 2, 4, sigma, vertices=P.jittered_vertices(nel, 0.1, rng))
        ops.append(Linv)
        prs.append(pr)
        qs.append(pr.form_source(Q=1.0) + rng.uniform(0, 0.1, pr.size))
    batch = A.SweepBatch(ops)
    dq = [A.DeviceVector.from_host(q) for q in qs]
    dpsi = [A.DeviceVector(q.size) for q in qs]
    batch.Mult(dq, dpsi)
    for Linv, pr, q, d in zip(ops, prs, qs, dpsi):
        got = d.to_host()
        assert rel_err(got, pr.sweep(q)) < FP64_TOL
        assert np.array_equal(got, Linv.Mult(q, np.empty_like(q)))
    t, n = batch.timing()
    assert n == 1 and t > 0


def test_p2_fp32_source_iteration_reaches_fp64_solution():
    nel = (4, 4, 4)
    ne = 64
    rng = np.random.default_rng(81)
    sigma = P.lognormal_sigma(ne, 1.0, rng).reshape(1, ne)
    c = rng.uniform(0.5, 0.99, ne)
    Linv, pr = _pair(nel, 2, 4, sigma, vertices=P.jittered_vertices(nel, 0.1, rng))
    q = pr.form_source(Q=1.0)
    phi64, _, it64, _ = A.source_iteration(Linv, c * sigma[0], q, abstol=1e-11)
    phi32, _, it32, _ = A.source_iteration(Linv, c * sigma[0], q, abstol=1e-11, precision=A.FP32_REFINED)
    ref_phi, _, _, _ = pr.source_iteration(c * sigma[0], q, abstol=1e-11)
    assert rel_err(phi64, ref_phi) < 1e-10
    assert rel_err(phi32, ref_phi) < 1e-9


def test_sweep_timing_log_counts_every_sweep():
    """The "sweep" timing log and "sweeps" count (src/sweep.cpp:665-668): the
    per-sweep events are read lazily, so more sweeps than the event ring holds
    (64) are all counted and timed."""
    print("sweep timing log", flush=True)
    Linv = A.InverseAdvectionOperator(A.Mesh((8, 8, 8)), 1, 4, np.ones(512))
    src = A.DeviceVector.from_host(np.ones(Linv.size))
    psi = A.DeviceVector(Linv.size)
    t0, n0 = Linv.timing()
    for _ in range(150):
        Linv.Mult(src, psi)
    t1, n1 = Linv.timing()
    assert n1 - n0 == 150
    assert t1 > t0 and (t1 - t0) / 150 < 0.1
    t2, n2 = Linv.timing()
    assert (t2, n2) == (t1, n1)


def test_operator_getters():
    """GetTimeAbsorption / GetLumpingType / IsParallelBlockJacobi
    (src/sweep.hpp:105,109,127): 0 until SetTimeAbsorption, no lumping, no PBJ."""
    print("operator getters", flush=True)
    Linv = A.InverseAdvectionOperator(A.Mesh((3, 3, 3)), 1, 4, np.ones(27))
    assert Linv.GetTimeAbsorption() == 0.0
    Linv.SetTimeAbsorption(2.5)
    assert Linv.GetTimeAbsorption() == 2.5
    assert Linv.GetLumpingType() == 0 and not Linv.IsParallelBlockJacobi()


@pytest.mark.parametrize("dim", [2, 3])
def test_anisotropic_moments(dim):
    """DiscreteToMoment with dim + 1 moments (src/transport_op.cpp:3-89): for
    psi = c0 + Omega . v the level-symmetric set integrates exactly
    (sum w Omega = 0, sum w Omega Omega^T = (sum w / 3) I), so phi = (sum w) c0,
    J = (sum w / 3) v, and the transpose returns psi itself."""
    print(f"anisotropic moments {dim}d", flush=True)
    G = 2
    nel = (4,) * dim
    ne = 4 ** dim
    om, w = ls_fixture(6, dim)
    Linv = A.InverseAdvectionOperator(A.Mesh(nel), 1, (om, w), np.ones((G, ne)))
    No, Ndof = len(w), Linv.phi_size // G
    rng = np.random.default_rng(4)
    c0 = rng.normal(size=(G, Ndof))
    v = rng.normal(size=(G, dim, Ndof))
    psi = c0[:, None, :] + np.einsum("am,gms->gas", np.asarray(om)[:, :dim], v)
    D = A.DiscreteToMoment(Linv, dim + 1)
    dpsi = A.DeviceVector.from_host(psi.ravel())
    dmom = A.DeviceVector((dim + 1) * G * Ndof)
    D.Mult(dpsi, dmom)
    A.synchronize()
    mom = dmom.to_host().reshape(dim + 1, G, Ndof)
    wsum = float(np.sum(w))
    np.testing.assert_allclose(mom[0], wsum * c0, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(mom[1:].transpose(1, 0, 2), wsum / 3 * v, rtol=1e-12, atol=1e-12)
    dback = A.DeviceVector(psi.size)
    D.MultTranspose(dmom, dback)
    A.synchronize()
    np.testing.assert_allclose(dback.to_host(), psi.ravel(), rtol=1e-12, atol=1e-12)
    with pytest.raises(ValueError):
        A.DiscreteToMoment(Linv, 2 if dim == 3 else 4)
