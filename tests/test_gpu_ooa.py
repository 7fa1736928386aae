"""Order of accuracy of the GPU sweep and source iteration against the PDE,
through the C ABI: the reference's exponential-solution sweep tests
(test/test_sweep.cpp:112-287) and manufactured-solution source iteration
(test/test_transport_mms.cpp:7-197), in 2D as the reference states them and
in 3D (level, cooperative p = 2 and box kernels), with the ψ-sized sources
formed on the host by allium_amd.fem and handed to the GPU."""
import numpy as np
import pytest

import allium_amd as A
from allium_amd.fem import BoxSpace
from tests import mms

pytestmark = pytest.mark.gpu

MU = 0.35002117   # first level-symmetric S4 direction's x cosine


def _gpu_exp_error(nel, p, om, exact):
    sig = (1.0, 0.5)
    sp, _, omega, w, st, inflow = mms.exponential_problem(nel, p, om, sig)
    Linv = A.InverseAdvectionOperator(A.Mesh(nel), p, (omega, w), st)
    q = Linv.form_source(inflow=inflow)
    psi = Linv.Mult(q, np.empty_like(q))
    return mms.exponential_error(sp, psi, omega[0], sig, exact)


@pytest.mark.parametrize("dim,p", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_exponential_solution_order(dim, p):
    """ExponentialSolution1Dp1/p2 as a slab of (10, 1[, 1]) and (20, 1[, 1])
    elements swept along (mu, 0[, 0]): order p + 1 within 0.1."""
    print(f"exponential OOA dim {dim} p {p}", flush=True)
    nel1, nel2 = (10,) + (1,) * (dim - 1), (20,) + (1,) * (dim - 1)
    om = (MU,) + (0.0,) * (dim - 1)
    e1, e2 = _gpu_exp_error(nel1, p, om, mms.exp_1d), _gpu_exp_error(nel2, p, om, mms.exp_1d)
    ooa = np.log2(e1 / e2)
    print(f"  errors {e1:.3e} {e2:.3e}, order {ooa:.3f}")
    assert abs(ooa - (p + 1)) < 0.1


@pytest.mark.parametrize("p,tol", [(1, 0.2), (2, 0.3)])
def test_exponential_diagonal_order(p, tol):
    """ExponentialSolution2Dp1/p2 on 50^2 and 100^2 quads (kink across the
    elements): order 1 within 0.2 (p = 1) and 0.3 (p = 2)."""
    print(f"exponential diagonal OOA p {p}", flush=True)
    e1 = _gpu_exp_error((50, 50), p, (MU, MU), mms.exp_diag)
    e2 = _gpu_exp_error((100, 100), p, (MU, MU), mms.exp_diag)
    ooa = np.log2(e1 / e2)
    print(f"  errors {e1:.3e} {e2:.3e}, order {ooa:.3f}")
    assert abs(ooa - 1.0) < tol


def _gpu_mms_error(nel, p, launch=None):
    dim = len(nel)
    om, w = A.level_symmetric(6, dim)
    Q, phi_ex = mms.mms_2d() if dim == 2 else mms.mms_3d()
    sp = BoxSpace(nel, p)
    Linv = A.InverseAdvectionOperator(A.Mesh(nel), p, 6, np.ones((1, sp.Ne)))
    if launch is not None:
        Linv.set_launch(launch)
    q = sp.project_source(Q, om).ravel() + Linv.form_source(inflow=np.full((1, 2 * dim), 2.0 / mms.FOURPI))
    phi, _, it, norms = A.source_iteration(Linv, np.full(sp.Ne, 0.25), q, abstol=1e-10, maxit=200)
    assert norms[it - 1] < 1e-10
    return sp.l2_error(phi, phi_ex)


@pytest.mark.parametrize("dim,p,n,launch", [(2, 1, 10, None), (2, 2, 10, None), (3, 1, 8, None), (3, 1, 8, 4),
                                            (3, 2, 6, None)])
def test_mms_source_iteration_order(dim, p, n, launch):
    """LinearTransport.MMS2Dp1/p2 (S6, sigma_t = 1, sigma_s = 0.25, inflow
    2/(4pi), SI to 1e-10 on n and 2n elements per axis) and its 3D analogue,
    also through the box sweeps (launch mode 4): order p + 1 within 0.2."""
    print(f"MMS OOA dim {dim} p {p} n {n} launch {launch}", flush=True)
    e1 = _gpu_mms_error((n,) * dim, p, launch)
    e2 = _gpu_mms_error((2 * n,) * dim, p, launch)
    ooa = np.log2(e1 / e2)
    print(f"  errors {e1:.3e} {e2:.3e}, order {ooa:.3f}")
    assert abs(ooa - (p + 1)) < 0.2
