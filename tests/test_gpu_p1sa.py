"""GPU checks of chive's block DSA preconditioners: "p1sa" (DSA(method="p1sa"),
exe/chive.cpp:608-619), the 2x2 P1 system of P1Discretization::GetOperator
(src/moment_discretization.cpp:400-490), and "mip" (DSA(method="block mip"),
exe/chive.cpp:635-662), the block interior-penalty system whose Schur
complement the reference solves.  The GPU solves both block systems by GMRES
preconditioned with element block-Jacobi around a MIP V-cycle.

The solve is compared with a direct sparse solve of the host-assembled P1
matrix (itself checked against the oracle's p1_matrix in test_dsa_oracle.py);
the accelerated source iteration with the oracle's (direct inner solve) on the
reference's thick-diffusion-limit deck and on a small heterogeneous 3D medium.
Tolerances: phi-part of the P1 solve 1e-9 relative (GMRES to 1e-12); fluxes
1e-8 relative; iteration counts equal the oracle's within one.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import allium_amd as A
import oracle
from allium_amd import problems as P
from oracle import mip_oracle as M
from tests.helpers import ls_fixture

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


CASES = [
    # dim, nel, p, jitter, sigma scale, c, reflective x-high side
    (2, (8, 8), 1, 0.0, 1e4, 1 - 1e-8, False),
    (2, (9, 7), 1, 0.2, 1.0, 0.99, True),
    (2, (10, 10), 2, 0.1, 1.0, 0.9, False),
    (3, (6, 6, 6), 1, 0.0, 1.0, 0.9999, False),
    (3, (7, 5, 6), 1, 0.15, 3.0, 0.99, True),
]


def _setup(dim, nel, p, jitter, scale, c, refl, seed=0):
    rng = np.random.default_rng(seed)
    vert = P.jittered_vertices(nel, jitter, rng) if jitter else P.cartesian_vertices(nel)
    ne = int(np.prod(nel))
    st = P.lognormal_sigma(ne, 1.0, rng) * scale
    ss = st * c
    bc = [0, 1 if refl else 0] + [0] * (2 * dim - 2)
    om, w = ls_fixture(4, dim)
    mesh = A.Mesh(nel, vertices=vert, bc=bc)
    return mesh, vert, bc, st, ss, om, w


def _phi_part_solve(mesh, p, st, ss, b):
    rp, col, val = A.host_p1_matrix(mesh, p, st, ss)
    Pm = sp.csc_matrix(sp.csr_matrix((val, col, rp), shape=(rp.size - 1,) * 2))
    n = (p + 1) ** mesh.dim
    full = np.zeros((mesh.Ne, mesh.dim + 1, n))
    full[:, mesh.dim, :] = b.reshape(mesh.Ne, n)
    return spla.spsolve(Pm, full.ravel()).reshape(mesh.Ne, mesh.dim + 1, n)[:, mesh.dim, :].ravel()


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}d-{'x'.join(map(str, c[1]))}-p{c[2]}" for c in CASES])
def test_p1_solve_matches_direct(case):
    dim, nel, p = case[:3]
    print(f"p1sa solve {case}", flush=True)
    mesh, vert, bc, st, ss, om, w = _setup(*case)
    Linv = A.InverseAdvectionOperator(mesh, p, (om, w), st)
    dsa = A.DSA(Linv, ss, method="p1sa", cg_reltol=1e-12, cg_maxit=3000)
    b = np.random.default_rng(5).normal(size=Linv.phi_size)
    x = dsa.solve(b)
    ref = _phi_part_solve(mesh, p, st, ss, b)
    info = dsa.info()
    print(f"  P1 unknowns {info['level_size']} GMRES its {info['last_cg_iterations']} "
          f"res {info['last_cg_norm']:.2e} rel {_rel(x, ref):.2e}")
    assert info["level_size"][0] == (dim + 1) * Linv.phi_size
    assert info["last_cg_iterations"] < 3000
    assert _rel(x, ref) < 1e-9


def test_p1sa_mult_device():
    """y = x + (P^{-1} [0; Ms x])_phi (DiffusionSyntheticAccelerationOperator::Mult
    around the extraction operators)."""
    print("p1sa mult", flush=True)
    case = (3, (6, 6, 6), 1, 0.1, 1.0, 0.99, False)
    mesh, vert, bc, st, ss, om, w = _setup(*case)
    Linv = A.InverseAdvectionOperator(mesh, 1, (om, w), st)
    dsa = A.DSA(Linv, ss, method="p1sa", cg_reltol=1e-12, cg_maxit=3000)
    pr = oracle.Problem(3, mesh.nel, vert, 1, om, w, st)
    x = np.random.default_rng(8).normal(size=Linv.phi_size)
    dx, dy = A.DeviceVector.from_host(x), A.DeviceVector(Linv.phi_size)
    dsa.Mult(dx, dy)
    A.synchronize()
    Pm = M.p1_matrix(3, mesh.nel, vert, 1, st, ss)
    ref = x + M.p1_solver(Pm, 3, 8)(pr.mass(ss, x))
    assert _rel(dy.to_host(), ref) < 1e-9


def test_tdl_deck_with_p1sa():
    """test/chive/tdl.lua with "p1sa": fewer than 12 iterations
    (test/chive/test_tdl.py's bound), the oracle's iterations and flux."""
    print("p1sa tdl", flush=True)
    eps = 1e-4
    nel = (8, 8)
    vert = P.cartesian_vertices(nel)
    om, w = ls_fixture(4, 2)
    st, ss = np.full(64, 1 / eps), np.full(64, 1 / eps - eps)
    Linv = A.InverseAdvectionOperator(A.Mesh(nel, vertices=vert), 1, (om, w), st)
    q = Linv.form_source(Q=eps)
    dsa = A.DSA(Linv, ss, method="p1sa", cg_reltol=1e-13, cg_maxit=3000)
    phi, _, it, norms = A.source_iteration(Linv, ss, q, abstol=1e-6, maxit=50, want_psi=False, accel=dsa)
    pr = oracle.Problem(2, nel, vert, 1, om, w, st)
    Pm = M.p1_matrix(2, nel, vert, 1, st, ss)
    phi_o, it_o, norms_o = M.dsa_source_iteration(pr, ss, pr.form_source(Q=eps), None, abstol=1e-6, maxit=50,
                                                  solve=M.p1_solver(Pm, 2, 4))
    print(f"  GPU {it} its, oracle {it_o}; last norms {norms[-1]:.3e} {norms_o[-1]:.3e}; "
          f"GMRES its per application {dsa.info()['cg_iterations'] / max(1, dsa.info()['applications']):.1f}")
    assert it < 12 and it == it_o
    np.testing.assert_allclose(norms, norms_o, rtol=1e-3, atol=1e-10 * norms_o[0])
    assert _rel(phi, phi_o) < 1e-8


def test_p1sa_si_matches_oracle_on_heterogeneous_medium():
    """6^3 lognormal medium (log-sd 1), c = 0.999, S4: the oracle's
    iterations (direct inner solve there) and plain SI's fixed point."""
    print("p1sa si hetero 6^3", flush=True)
    nel = (6, 6, 6)
    vert = P.cartesian_vertices(nel)
    rng = np.random.default_rng(3)
    st = P.lognormal_sigma(216, 1.0, rng) * 5.0
    ss = st * 0.999
    om, w = ls_fixture(4, 3)
    Linv = A.InverseAdvectionOperator(A.Mesh(nel, vertices=vert), 1, (om, w), st)
    q = Linv.form_source(Q=1.0)
    phi0, _, it0, _ = A.source_iteration(Linv, ss, q, abstol=1e-10, maxit=5000, want_psi=False)
    dsa = A.DSA(Linv, ss, method="p1sa", cg_reltol=1e-12, cg_maxit=3000)
    phi1, _, it1, n1 = A.source_iteration(Linv, ss, q, abstol=1e-10, maxit=5000, want_psi=False, accel=dsa)
    pr = oracle.Problem(3, nel, vert, 1, om, w, st)
    Pm = M.p1_matrix(3, nel, vert, 1, st, ss)
    phi_o, it_o, _ = M.dsa_source_iteration(pr, ss, pr.form_source(Q=1.0), None, abstol=1e-10, maxit=500,
                                            solve=M.p1_solver(Pm, 3, 8))
    print(f"  plain {it0}, P1SA {it1}, oracle P1SA {it_o}")
    assert abs(it1 - it_o) <= 1
    assert _rel(phi1, phi_o) < 1e-8 and _rel(phi1, phi0) < 1e-7
    assert it1 < it0


def test_p1sa_preconditions_gmres():
    """chive's gmres outer solver with the "p1sa" preconditioner reaches the
    fixed point of plain SI."""
    print("p1sa gmres", flush=True)
    nel = (6, 6, 6)
    st = np.full(216, 4.0)
    ss = st * 0.999
    Linv = A.InverseAdvectionOperator(A.Mesh(nel), 1, 4, st)
    q = Linv.form_source(Q=1.0)
    phi0, _, it0, _ = A.source_iteration(Linv, ss, q, abstol=1e-10, maxit=5000, want_psi=False)
    dsa = A.DSA(Linv, ss, method="p1sa", cg_reltol=1e-12, cg_maxit=3000)
    phi1, _, it1, _ = A.krylov_solve(Linv, ss, q, solver="gmres", abstol=1e-10, maxit=200, accel=dsa,
                                     want_psi=False)
    print(f"  plain SI {it0}, GMRES + P1SA {it1}")
    assert _rel(phi1, phi0) < 1e-7 and it1 < it0


def test_p1sa_errors():
    print("p1sa errors", flush=True)
    mesh = A.Mesh((4, 4, 4))
    st = np.ones(64)
    with pytest.raises(A.AlliumError):
        A.DSA(A.InverseAdvectionOperator(mesh, 0, 4, st), st * 0.9, method="p1sa")
    with pytest.raises(KeyError):
        A.DSA(A.InverseAdvectionOperator(mesh, 1, 4, st), st * 0.9, method="h1sa")


BLOCK_CASES = [c for c in CASES if c[2] == 1]


@pytest.mark.parametrize("case", BLOCK_CASES, ids=[f"{c[0]}d-{'x'.join(map(str, c[1]))}" for c in BLOCK_CASES])
def test_block_mip_solve_matches_schur_direct(case):
    """chive's block "mip" (DSA(method="block mip")): the phi-part of the GPU
    block solve equals a direct solve of the Schur complement Ma + D Mt^-1 D^T
    that the reference forms (FormSchurComplement, src/moment_discretization.cpp:248-264)."""
    dim, nel, p = case[:3]
    print(f"block mip solve {case}", flush=True)
    mesh, vert, bc, st, ss, om, w = _setup(*case)
    Linv = A.InverseAdvectionOperator(mesh, p, (om, w), st)
    dsa = A.DSA(Linv, ss, method="block mip", cg_reltol=1e-12, cg_maxit=3000)
    b = np.random.default_rng(6).normal(size=Linv.phi_size)
    x = dsa.solve(b)
    rp, col, val = A.host_block_mip_matrix(mesh, p, (om, w), st, ss)
    S = M.schur_complement(sp.csr_matrix((val, col, rp), shape=(rp.size - 1,) * 2), dim, (p + 1) ** dim)
    ref = spla.spsolve(sp.csc_matrix(S), b)
    info = dsa.info()
    print(f"  GMRES its {info['last_cg_iterations']} rel {_rel(x, ref):.2e}")
    assert info["last_cg_iterations"] < 3000 and _rel(x, ref) < 1e-9


def test_tdl_deck_with_block_mip():
    """chive.tdl.mip (test/chive/CMakeLists.txt:117-121): under 12 iterations,
    the oracle's iterations (Schur complement solved directly there) and flux."""
    print("block mip tdl", flush=True)
    eps = 1e-4
    nel = (8, 8)
    vert = P.cartesian_vertices(nel)
    om, w = ls_fixture(4, 2)
    st, ss = np.full(64, 1 / eps), np.full(64, 1 / eps - eps)
    Linv = A.InverseAdvectionOperator(A.Mesh(nel, vertices=vert), 1, (om, w), st)
    q = Linv.form_source(Q=eps)
    dsa = A.DSA(Linv, ss, method="block mip", cg_reltol=1e-13, cg_maxit=3000)
    phi, _, it, norms = A.source_iteration(Linv, ss, q, abstol=1e-6, maxit=50, want_psi=False, accel=dsa)
    pr = oracle.Problem(2, nel, vert, 1, om, w, st)
    Pm = M.block_ip_matrix(2, nel, vert, 1, st, ss, om, w)
    phi_o, it_o, norms_o = M.dsa_source_iteration(pr, ss, pr.form_source(Q=eps), None, abstol=1e-6, maxit=50,
                                                  solve=M.p1_solver(Pm, 2, 4))
    print(f"  GPU {it} its, oracle {it_o}; GMRES its per application "
          f"{dsa.info()['cg_iterations'] / max(1, dsa.info()['applications']):.1f}")
    assert it < 12 and it == it_o
    np.testing.assert_allclose(norms, norms_o, rtol=1e-3, atol=1e-10 * norms_o[0])
    assert _rel(phi, phi_o) < 1e-8


@pytest.mark.parametrize("case", BLOCK_CASES, ids=[f"{c[0]}d-{'x'.join(map(str, c[1]))}" for c in BLOCK_CASES])
def test_ldg_solve_matches_schur_direct(case):
    """chive's "ldgsa" (DSA(method="ldgsa")): the GPU block solve's phi-part
    equals a direct solve of the LDG Schur complement."""
    dim, nel, p = case[:3]
    print(f"ldgsa solve {case}", flush=True)
    mesh, vert, bc, st, ss, om, w = _setup(*case)
    Linv = A.InverseAdvectionOperator(mesh, p, (om, w), st)
    dsa = A.DSA(Linv, ss, method="ldgsa", cg_reltol=1e-12, cg_maxit=3000)
    b = np.random.default_rng(7).normal(size=Linv.phi_size)
    x = dsa.solve(b)
    rp, col, val = A.host_block_ldg_matrix(mesh, p, (om, w), st, ss)
    S = M.schur_complement(sp.csr_matrix((val, col, rp), shape=(rp.size - 1,) * 2), dim, (p + 1) ** dim)
    ref = spla.spsolve(sp.csc_matrix(S), b)
    info = dsa.info()
    print(f"  GMRES its {info['last_cg_iterations']} rel {_rel(x, ref):.2e}")
    assert info["last_cg_iterations"] < 3000 and _rel(x, ref) < 1e-9


def test_tdl_deck_with_ldgsa():
    """chive.tdl.ldgsa: under 12 iterations, the oracle's iterations and flux."""
    print("ldgsa tdl", flush=True)
    eps = 1e-4
    nel = (8, 8)
    vert = P.cartesian_vertices(nel)
    om, w = ls_fixture(4, 2)
    st, ss = np.full(64, 1 / eps), np.full(64, 1 / eps - eps)
    Linv = A.InverseAdvectionOperator(A.Mesh(nel, vertices=vert), 1, (om, w), st)
    q = Linv.form_source(Q=eps)
    dsa = A.DSA(Linv, ss, method="ldgsa", cg_reltol=1e-13, cg_maxit=3000)
    phi, _, it, norms = A.source_iteration(Linv, ss, q, abstol=1e-6, maxit=50, want_psi=False, accel=dsa)
    pr = oracle.Problem(2, nel, vert, 1, om, w, st)
    Pm = M.block_ldg_matrix(2, nel, vert, 1, st, ss, om, w)
    phi_o, it_o, norms_o = M.dsa_source_iteration(pr, ss, pr.form_source(Q=eps), None, abstol=1e-6, maxit=50,
                                                  solve=M.p1_solver(Pm, 2, 4))
    print(f"  GPU {it} its, oracle {it_o}; GMRES its per application "
          f"{dsa.info()['cg_iterations'] / max(1, dsa.info()['applications']):.1f}")
    assert it < 12 and it == it_o
    np.testing.assert_allclose(norms, norms_o, rtol=1e-3, atol=1e-10 * norms_o[0])
    assert _rel(phi, phi_o) < 1e-8
