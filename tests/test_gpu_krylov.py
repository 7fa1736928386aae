"""GPU checks of the Krylov outer solvers (allium_amd/csrc/krylov*.{hip,cpp}).

The GPU's GMRES / FGMRES against the oracle's restatement of mfem's
(oracle/krylov_oracle.py) on the oracle's transport operator: the same
iteration count, residual histories within 1e-8 of beta0 and solutions within
1e-9 relative (the operators agree to round-off, and GMRES's estimates are
stable functions of them).  Preconditioned forms with DSA against the
oracle's exact MIP solve (the GPU's CG runs to 1e-12): iteration counts within
one, solutions within 1e-8.  The learned preconditioner is checked for the
fixed point it reaches and for its switch-off rule.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import allium_amd as A
import oracle
from allium_amd import cnn
from allium_amd import problems as P
from oracle import cnn_oracle as C
from oracle import krylov_oracle as K
from oracle import mip_oracle as M
from tests.helpers import ls_fixture

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


CASES = [
    # dim, nel, p, jitter, sigma scale, c, S_N, kdim
    (3, (6, 6, 6), 1, 0.0, 2.0, 0.99, 4, 50),
    (3, (5, 4, 6), 1, 0.15, 2.0, 0.99, 4, 50),
    (3, (6, 6, 6), 1, 0.0, 2.0, 0.999, 4, 5),
    (2, (10, 8), 2, 0.2, 4.0, 0.99, 6, 50),
    (2, (12, 12), 1, 0.0, 4.0, 0.999, 8, 7),
    (3, (4, 4, 4), 2, 0.0, 2.0, 0.99, 4, 50),
]


def _problem(dim, nel, p, jitter, scale, c, order, seed=0):
    rng = np.random.default_rng(seed)
    vert = P.jittered_vertices(nel, jitter, rng) if jitter else P.cartesian_vertices(nel)
    ne = int(np.prod(nel))
    st = P.lognormal_sigma(ne, 1.0, rng) * scale
    ss = st * c
    om, w = ls_fixture(order, dim)
    Linv = A.InverseAdvectionOperator(A.Mesh(nel, vertices=vert), p, (om, w), st)
    pr = oracle.Problem(dim, nel, vert, p, om, w, st)
    return Linv, pr, st, ss, om, w, vert


@pytest.mark.parametrize("solver", ["gmres", "fgmres"])
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}d-{'x'.join(map(str, c[1]))}-p{c[2]}-k{c[7]}" for c in CASES])
def test_krylov_matches_oracle(case, solver):
    dim, nel, p, jitter, scale, c, order, kdim = case
    print(f"krylov {solver} {case}", flush=True)
    Linv, pr, st, ss, om, w, vert = _problem(dim, nel, p, jitter, scale, c, order)
    q = Linv.form_source(Q=1.0)
    phi, psi, it, norms = A.krylov_solve(Linv, ss, q, solver=solver, kdim=kdim, abstol=1e-10, maxit=300)
    T = K.transport_operator(pr, ss)
    b = pr.D(pr.sweep(q))
    f = K.gmres if solver == "gmres" else K.fgmres
    x, it_o, norms_o, conv = f(T, b, kdim=kdim, abstol=1e-10, maxit=300)
    print(f"  GPU {it} its, oracle {it_o}; rel {_rel(phi, x):.2e}")
    assert conv and it == it_o
    np.testing.assert_allclose(norms, norms_o, rtol=0, atol=1e-8 * norms_o[0])
    assert _rel(phi, x) < 1e-9
    # the final sweep gives the angular flux of the converged scalar flux
    psi_o = pr.sweep(q + pr.DT(pr.mass(ss, x)))
    assert _rel(psi, psi_o) < 1e-9


@pytest.mark.parametrize("solver", ["gmres", "fgmres"])
def test_krylov_with_dsa_matches_oracle(solver):
    """chive's "scalar mip" preconditioner on the Krylov solve (exe/chive.cpp:715)."""
    print(f"krylov {solver} + dsa", flush=True)
    Linv, pr, st, ss, om, w, vert = _problem(3, (6, 6, 6), 1, 0.0, 4.0, 0.999, 4, seed=2)
    q = Linv.form_source(Q=1.0)
    dsa = A.DSA(Linv, ss, cg_reltol=1e-13, cg_maxit=400)
    inner = np.zeros(300, dtype=np.int32)
    phi, _, it, norms = A.krylov_solve(Linv, ss, q, solver=solver, abstol=1e-10, maxit=300, accel=dsa,
                                       want_psi=False, inner_iterations=inner)
    Km = M.mip_matrix(3, (6, 6, 6), vert, 1, st, ss, om, w)
    solve = spla.factorized(sp.csc_matrix(Km))
    Mop = lambda v: v + solve(pr.mass(ss, v))
    T = K.transport_operator(pr, ss)
    b = pr.D(pr.sweep(q))
    f = K.gmres if solver == "gmres" else K.fgmres
    x, it_o, norms_o, conv = f(T, b, M=Mop, kdim=50, abstol=1e-10, maxit=300)
    _, it_plain, _, _ = K.gmres(T, b, kdim=50, abstol=1e-10, maxit=300)
    print(f"  GPU {it} its, oracle {it_o}, unpreconditioned {it_plain}; CG its {inner[:it]}")
    assert conv and abs(it - it_o) <= 1 and it < it_plain
    assert _rel(phi, x) < 1e-8
    assert np.all(inner[:it] > 0)
    np.testing.assert_allclose(norms[:3], norms_o[:3], rtol=1e-8)


def test_fgmres_learned_preconditioner_reaches_the_fixed_point():
    """A network (random weights here; trained ones in the benchmark) as the
    flexible preconditioner: same fixed point as source iteration."""
    print("fgmres learned", flush=True)
    cfg = P.problem_3d_hetero(n=12, seed=4, log_sd=1.0, c_const=0.999)
    Linv = A.InverseAdvectionOperator(A.Mesh(cfg["nel"]), 1, 4, cfg["sigma_t"])
    q = Linv.form_source(Q=1.0)
    phi0, _, it0, _ = A.source_iteration(Linv, cfg["sigma_s"], q, abstol=1e-10, maxit=3000, want_psi=False)
    net = cnn.FluxCorrector(3)
    phi1, _, it1, n1 = A.krylov_solve(Linv, cfg["sigma_s"], q, solver="fgmres", abstol=1e-10, maxit=300,
                                      accel=net, want_psi=False)
    steps = A.krylov_solve.last_accel_steps
    print(f"  SI {it0} its, FGMRES + network {it1} its ({steps} with the network)")
    assert n1[-1] <= 1e-10 and steps > 0
    assert _rel(phi1, phi0) < 1e-7


def test_fgmres_switches_off_a_bad_network():
    """A network whose corrections are useless stops being applied after
    `patience` slow steps; FGMRES then finishes unpreconditioned or with DSA."""
    print("fgmres bad network", flush=True)
    cfg = P.problem_3d_hetero(n=10, seed=2, log_sd=1.0, c_const=0.99)
    Linv = A.InverseAdvectionOperator(A.Mesh(cfg["nel"]), 1, 4, cfg["sigma_t"])
    q = Linv.form_source(Q=1.0)
    params = C.init_params(9)
    W, b = params[-1]
    params[-1] = (W * 200.0, b + 50.0)
    net = cnn.FluxCorrector(0)
    net.set_params(C.pack(params).astype(np.float32))
    phi0, _, it0, _ = A.krylov_solve(Linv, cfg["sigma_s"], q, solver="fgmres", abstol=1e-10, maxit=300,
                                     want_psi=False)
    dsa = A.DSA(Linv, cfg["sigma_s"])
    for fb in (None, dsa):
        phi1, _, it1, n1 = A.krylov_solve(Linv, cfg["sigma_s"], q, solver="fgmres", abstol=1e-10, maxit=300,
                                          accel=net, fallback=fb, patience=2, want_psi=False)
        steps = A.krylov_solve.last_accel_steps
        print(f"  plain FGMRES {it0}, bad network + fallback {fb is not None}: {it1} its, {steps} with the network")
        assert n1[-1] <= 1e-10 and steps < it1
        assert _rel(phi1, phi0) < 1e-8


def test_krylov_errors():
    print("krylov errors", flush=True)
    Linv, pr, st, ss, om, w, vert = _problem(3, (4, 4, 4), 1, 0.0, 1.0, 0.9, 4)
    q = Linv.form_source(Q=1.0)
    with pytest.raises(A.AlliumError):
        A.krylov_solve(Linv, ss, q, kdim=0)
    with pytest.raises(A.AlliumError):
        A.krylov_solve(Linv, ss, q, kdim=201)
    with pytest.raises(ValueError):
        A.krylov_solve(Linv, ss, q, solver="bicgstab")
    with pytest.raises(A.AlliumError):
        A.krylov_solve(Linv, ss, q, solver="gmres", accel=cnn.FluxCorrector(0))
    other = A.InverseAdvectionOperator(A.Mesh((4, 4, 4)), 1, 4, st)
    with pytest.raises(A.AlliumError):
        A.krylov_solve(Linv, ss, q, accel=A.DSA(other, ss))
    # maxit reached: the iterate is returned, not an error
    phi, _, it, norms = A.krylov_solve(Linv, ss, q, abstol=1e-30, maxit=3, want_psi=False)
    assert it == 3 and norms.size == 4 and np.all(np.isfinite(phi))


def test_device_entry_points_match_host():
    """allium_source_iteration_device / allium_krylov_solve_device with the
    source already in HBM give the host entry points' results bit for bit."""
    print("device entry points", flush=True)
    Linv, pr, st, ss, om, w, vert = _problem(3, (6, 6, 6), 1, 0.0, 2.0, 0.99, 4)
    q = Linv.form_source(Q=1.0)
    dq = A.DeviceVector.from_host(q)
    for f, kw in ((A.source_iteration, {}), (A.krylov_solve, {"solver": "gmres"}), (A.krylov_solve, {})):
        phi_h, psi_h, it_h, n_h = f(Linv, ss, q, abstol=1e-10, maxit=500, **kw)
        phi_d, psi_d, it_d, n_d = f(Linv, ss, dq, abstol=1e-10, maxit=500, **kw)
        assert it_h == it_d
        assert np.array_equal(n_h, n_d) and np.array_equal(phi_h, phi_d)
        assert np.array_equal(psi_h, psi_d.to_host())
