"""Lumped sweep matrices on the GPU (InverseAdvectionOperator's `lump` bit set,
src/sweep.cpp:940-1022, src/lumping.cpp:3-12), through the C ABI: the
reference's LumpSweep and LumpedAdvectionOperator tests
(test/test_sweep.cpp:55-58,80-83,100-103,189-243,343-361) restated, and the
GPU sweep against the CPU oracle for every lump bit.  Lumped handles run the
level kernels; the box kernels decline them."""
import numpy as np
import pytest

import allium_amd as A
import oracle
from allium_amd import problems as P
from tests import mms
from tests.helpers import ls_fixture, rel_err

pytestmark = pytest.mark.gpu

MU = 0.35002117   # first level-symmetric S4 direction's x cosine


@pytest.mark.parametrize("nel", [(10, 10), (6, 6, 6)])
def test_lumped_constant_solution(nel):
    """LumpSweep.ConstantSolution2Dp1/3Dp1: sigma = 0, inflow 1 on every side,
    lump = 7 -> psi == 1 (to 1e-10), and the oracle's lumped sweep to 1e-12."""
    print(f"lumped constant solution {nel}", flush=True)
    dim, ne = len(nel), int(np.prod(nel))
    Linv = A.InverseAdvectionOperator(A.Mesh(nel), 1, 4, np.zeros((2, ne)), lump=7)
    assert Linv.GetLumpingType() == 7
    q = Linv.form_source(inflow=np.ones((2, 2 * dim)))
    psi = Linv.Mult(q, np.empty_like(q))
    assert np.abs(psi - 1.0).max() < 1e-10
    om, w = ls_fixture(4, dim)
    pr = oracle.Problem(dim, nel, P.cartesian_vertices(nel), 1, om, w, np.zeros((2, ne)), lump=7)
    assert rel_err(psi, pr.sweep(q)) < 1e-12


@pytest.mark.parametrize("nel", [(3, 3), (3, 3, 3)])
def test_lumped_L_Linv(nel):
    """LumpedAdvectionOperator.Operator2D/3D: L^-1 (L x) = x with lump = 7,
    sigma = (1, 0.5), through the GPU L and L^-1."""
    print(f"lumped L Linv {nel}", flush=True)
    ne = int(np.prod(nel))
    Linv = A.InverseAdvectionOperator(A.Mesh(nel), 1, 4, np.vstack([np.ones(ne), 0.5 * np.ones(ne)]), lump=7)
    L = A.AdvectionOperator(Linv)
    x = np.random.default_rng(12345).uniform(-1, 1, Linv.size)
    y = L.Mult(x, np.empty_like(x))
    z = Linv.Mult(y, y)
    assert np.linalg.norm(z - x) < 1e-12


@pytest.mark.parametrize("lump", [1, 2, 4, 5, 7])
@pytest.mark.parametrize("nel,p,jitter", [((8, 7, 6), 1, 0.1), ((12, 10), 1, 0.2), ((9, 8, 7), 0, 0.0),
                                          ((10, 9, 8), 1, 0.0)])
def test_lumped_sweep_matches_oracle(nel, p, jitter, lump):
    """Every lump bit on general and uniform geometry: the GPU sweep, L psi and
    the particle balance against the oracle's lumped matrices."""
    print(f"lumped sweep {nel} p {p} jitter {jitter} lump {lump}", flush=True)
    dim, ne = len(nel), int(np.prod(nel))
    rng = np.random.default_rng(31 + lump)
    sigma = P.lognormal_sigma(ne, 1.0, rng).reshape(1, ne)
    v = P.jittered_vertices(nel, jitter, rng) if jitter else P.cartesian_vertices(nel)
    Linv = A.InverseAdvectionOperator(A.Mesh(nel, vertices=v), p, 6, sigma, lump=lump)
    om, w = ls_fixture(6, dim)
    pr = oracle.Problem(dim, nel, v, p, om, w, sigma, lump=lump)
    q = pr.form_source(Q=1.0) + rng.uniform(0, 0.1, pr.size)
    ref = pr.sweep(q)
    # Lumping the gradient but not the face terms (lump = 2) breaks the
    # summation-by-parts pairing of the two, so the cell matrices lose their
    # positive-definite symmetric part and the scheme amplifies (|psi| ~ 70 for
    # sources ~ 0.1).  The GPU factors without pivoting, the oracle with partial
    # pivoting, and on these matrices the two agree to ~1e-11.
    tol = 1e-10 if lump == 2 else 1e-12
    assert rel_err(Linv.Mult(q, np.empty_like(q)), ref) < tol
    L = A.AdvectionOperator(Linv)
    x = rng.normal(size=pr.size)
    assert rel_err(L.Mult(x, np.empty_like(x)), pr.apply_L(x)) < 1e-13
    assert abs(L.ComputeBalance(ref) - pr.balance(ref)) < 1e-12 * abs(pr.balance(ref))


def _lumped_exp_error(nel, om, exact):
    sig = (1.0, 0.5)
    sp, _, omega, w, st, inflow = mms.exponential_problem(nel, 1, om, sig)
    Linv = A.InverseAdvectionOperator(A.Mesh(nel), 1, (omega, w), st, lump=7)
    q = Linv.form_source(inflow=inflow)
    psi = Linv.Mult(q, np.empty_like(q))
    return mms.exponential_error(sp, psi, omega[0], sig, exact)


def test_lumped_exponential_solution_order():
    """LumpSweep.ExponentialSolution1Dp1 (a slab of 10 and 20 elements swept
    along (mu, 0)): order 2 within 0.1; LumpSweep.ExponentialSolution2Dp1
    (50^2 and 100^2 quads, kink across the elements): order 1 within 0.2."""
    print("lumped exponential OOA", flush=True)
    e1, e2 = _lumped_exp_error((10, 1), (MU, 0.0), mms.exp_1d), _lumped_exp_error((20, 1), (MU, 0.0), mms.exp_1d)
    ooa = np.log2(e1 / e2)
    print(f"  slab errors {e1:.3e} {e2:.3e}, order {ooa:.3f}")
    assert abs(ooa - 2.0) < 0.1
    e1 = _lumped_exp_error((50, 50), (MU, MU), mms.exp_diag)
    e2 = _lumped_exp_error((100, 100), (MU, MU), mms.exp_diag)
    ooa = np.log2(e1 / e2)
    print(f"  diagonal errors {e1:.3e} {e2:.3e}, order {ooa:.3f}")
    assert abs(ooa - 1.0) < 0.2


def test_lumped_handles_decline_box_sweeps_and_autotune_to_levels():
    nel = (16, 16, 16)
    ne = int(np.prod(nel))
    Linv = A.InverseAdvectionOperator(A.Mesh(nel), 1, 4, np.ones((1, ne)), lump=7)
    for mode in (4, 5):
        with pytest.raises(A.AlliumError, match="unlumped"):
            Linv.set_launch(mode)
    Linv.set_launch(-1)
    assert Linv.tuning()["mode"] < 4
    om, w = ls_fixture(4, 3)
    pr = oracle.Problem(3, nel, P.cartesian_vertices(nel), 1, om, w, np.ones((1, ne)), lump=7)
    q = pr.form_source(Q=1.0)
    assert rel_err(Linv.Mult(q, np.empty_like(q)), pr.sweep(q)) < 1e-12


def test_lumped_p2_is_declined():
    with pytest.raises(A.AlliumError, match="fe_order 0 and 1"):
        A.InverseAdvectionOperator(A.Mesh((4, 4)), 2, 4, np.ones((1, 16)), lump=7)


def test_dsa_declines_lumped_handles():
    """chive assembles its DSA operators with the transport lumping
    (exe/chive.cpp:610-666); the library's DSA operators are unlumped, so a
    lumped handle is declined rather than preconditioned inconsistently."""
    Linv = A.InverseAdvectionOperator(A.Mesh((4, 4, 4)), 1, 4, np.ones((1, 64)), lump=1)
    with pytest.raises(A.AlliumError, match="unlumped"):
        A.DSA(Linv, 0.5 * np.ones(64))
