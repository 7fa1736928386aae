"""CPU pins of the Krylov oracle (oracle/krylov_oracle.py): mfem's GMRES(m)
and FGMRES(m) as chive uses them on the transport operator
(src/io.cpp:206-225, exe/chive.cpp:702-721).  No GPU compute is called."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
from allium_amd import problems as P
from oracle import krylov_oracle as K
from tests.helpers import ls_fixture


def _system(n=30, seed=0, spread=0.3):
    rng = np.random.default_rng(seed)
    A = np.eye(n) + spread * rng.normal(size=(n, n)) / np.sqrt(n)
    b = rng.normal(size=n)
    return A, b


@pytest.mark.parametrize("k", [1, 2, 3, 5, 8])
def test_iterate_is_minimum_residual_over_krylov_space(k):
    """x_k = argmin ||b - A x|| over span{b, Ab, ..., A^{k-1} b} (x_0 = 0)."""
    A, b = _system()
    x, it, norms, conv = K.gmres(lambda v: A @ v, b, kdim=50, maxit=k)
    assert not conv and it == k
    Kry = np.stack([np.linalg.matrix_power(A, j) @ b for j in range(k)], 1)
    Q, _ = np.linalg.qr(Kry)
    y, *_ = np.linalg.lstsq(A @ Q, b, rcond=None)
    x_ls = Q @ y
    assert np.linalg.norm(x - x_ls) < 1e-10 * np.linalg.norm(x_ls)
    # the rotated estimate after step k is the true minimum residual
    assert abs(norms[k] - np.linalg.norm(b - A @ x_ls)) < 1e-10 * np.linalg.norm(b)


def test_exact_in_n_steps():
    A, b = _system(n=12, seed=1, spread=1.0)
    x, it, norms, conv = K.gmres(lambda v: A @ v, b, kdim=50, abstol=1e-11 * np.linalg.norm(b), maxit=40)
    assert conv and it <= 12
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("kdim", [50, 5])
def test_history_matches_scipy(kdim):
    """The per-step residual history (relative to ||b||) equals SciPy's GMRES
    with the same restart length; the final iterates agree."""
    A, b = _system(n=40, seed=2, spread=0.9)
    hist = []
    xs, info = spla.gmres(A, b, rtol=1e-10, atol=0.0, restart=kdim, maxiter=400,
                          callback=lambda r: hist.append(r), callback_type="pr_norm")
    x, it, norms, conv = K.gmres(lambda v: A @ v, b, kdim=kdim, reltol=1e-10, maxit=400)
    assert conv and info == 0
    ours = norms[1:] / norms[0]
    m = min(len(hist), len(ours))
    assert m >= 10
    np.testing.assert_allclose(ours[:m], hist[:m], rtol=1e-6, atol=1e-13)
    assert np.linalg.norm(x - xs) < 1e-8 * np.linalg.norm(xs)


def test_restart_changes_the_iterates_but_not_the_solution():
    A, b = _system(n=40, seed=3, spread=0.9)
    x50, it50, _, c50 = K.gmres(lambda v: A @ v, b, kdim=50, reltol=1e-12, maxit=400)
    x4, it4, _, c4 = K.gmres(lambda v: A @ v, b, kdim=4, reltol=1e-12, maxit=400)
    assert c50 and c4 and it4 > it50
    ref = np.linalg.solve(A, b)
    for x in (x50, x4):
        assert np.linalg.norm(x - ref) < 1e-9 * np.linalg.norm(ref)


def test_left_preconditioned_is_gmres_on_MA():
    A, b = _system(n=30, seed=4, spread=0.8)
    d = 1.0 / np.random.default_rng(5).uniform(0.5, 2.0, 30)
    M = lambda v: d * v
    x, it, norms, conv = K.gmres(lambda v: A @ v, b, M=M, kdim=50, reltol=1e-10, maxit=100)
    MA = d[:, None] * A
    x2, it2, norms2, conv2 = K.gmres(lambda v: MA @ v, d * b, kdim=50, reltol=1e-10, maxit=100)
    assert conv and conv2 and it == it2
    np.testing.assert_allclose(norms, norms2, rtol=1e-9, atol=1e-12 * norms2[0])
    assert norms[0] == pytest.approx(np.linalg.norm(d * b))


def test_fgmres_with_fixed_preconditioner_is_right_preconditioned_gmres():
    """With a linear M, FGMRES minimises ||b - A M u|| over the Krylov space of
    A M; its residual estimates are true (unpreconditioned) residuals."""
    A, b = _system(n=30, seed=6, spread=0.8)
    d = np.random.default_rng(7).uniform(0.5, 2.0, 30)
    M = lambda v: d * v
    x, it, norms, conv = K.fgmres(lambda v: A @ v, b, M=M, kdim=50, reltol=1e-10, maxit=100)
    AM = A * d[None, :]
    u, it2, norms2, conv2 = K.gmres(lambda v: AM @ v, b, kdim=50, reltol=1e-10, maxit=100)
    assert conv and conv2 and it == it2
    np.testing.assert_allclose(norms, norms2, rtol=1e-9, atol=1e-12 * norms2[0])
    np.testing.assert_allclose(x, d * u, rtol=1e-9, atol=1e-12)
    assert norms[-1] == pytest.approx(np.linalg.norm(b - A @ x), rel=1e-6, abs=1e-14)


def test_fgmres_nonlinear_preconditioner_converges():
    """A preconditioner that is not linear (as a trained network is not)
    still gives the solution; the last estimate is the true residual."""
    A, b = _system(n=30, seed=8, spread=0.8)
    M = lambda v: v + 0.2 * np.tanh(v)
    x, it, norms, conv = K.fgmres(lambda v: A @ v, b, M=M, kdim=50, reltol=1e-10, maxit=100)
    assert conv
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-8, atol=1e-12)
    assert norms[-1] == pytest.approx(np.linalg.norm(b - A @ x), rel=1e-5, abs=1e-13)


@pytest.mark.parametrize("solver", ["gmres", "fgmres"])
def test_transport_operator_solution_is_the_si_fixed_point(solver):
    """GMRES on T = I - D L^{-1} D^T Ms with b = D L^{-1} q reaches source
    iteration's fixed point in fewer operator applications (c = 0.99)."""
    nel = (6, 6)
    vert = P.cartesian_vertices(nel)
    rng = np.random.default_rng(9)
    st = P.lognormal_sigma(36, 1.0, rng) * 4.0
    ss = st * 0.99
    om, w = ls_fixture(4, 2)
    pr = oracle.Problem(2, nel, vert, 1, om, w, st)
    q = pr.form_source(Q=1.0)
    b = pr.D(pr.sweep(q))
    T = K.transport_operator(pr, ss)
    phi_si, _, it_si, _ = pr.source_iteration(ss, q, abstol=1e-12, maxit=5000)
    f = K.gmres if solver == "gmres" else K.fgmres
    x, it, norms, conv = f(T, b, kdim=50, abstol=1e-12, maxit=500)
    assert conv and it < it_si / 3
    assert np.linalg.norm(x - phi_si) < 1e-9 * np.linalg.norm(phi_si)
    assert np.linalg.norm(b - T(x)) < 1e-10
