"""GPU checks of the negative-flux fixups (allium_sweep_set_fixup; src/fixup.cpp
applied after each cell solve, src/sweep.cpp:465-469) against the oracle's
(pinned in test_fixup_oracle.py).  Optically thick cells lit from one side
give negative DG values.  Tolerances: psi 1e-12 relative; the applied/failed
counts and the per-element monitor equal the oracle's."""
import numpy as np
import pytest

import allium_amd as A
import oracle
from allium_amd import problems as P
from tests.helpers import ls_fixture, rel_err

pytestmark = pytest.mark.gpu

CASES = [
    # dim, nel, p, sigma_t, jitter, fixup, psi_min, launch mode
    (2, (8, 8), 1, 50.0, 0.0, "zero", 0.0, 0),
    (2, (8, 8), 1, 50.0, 0.0, "zero and scale", 0.0, 4),
    (2, (8, 8), 1, 20.0, 0.0, "ryosuke", 0.0, 0),
    (2, (9, 7), 2, 50.0, 0.15, "zero and scale", 0.0, 0),
    (3, (6, 6, 6), 1, 30.0, 0.0, "zero", 0.0, 5),
    (3, (6, 6, 6), 1, 30.0, 0.0, "zero and scale", 0.0, -1),
    (3, (5, 6, 4), 1, 30.0, 0.2, "zero and scale", 0.0, 0),
    # p = 0 stays positive: a floor above its smallest values makes the fixup act
    (3, (6, 5, 7), 0, 30.0, 0.0, "zero and scale", 1e-6, 0),
]


def _problem(dim, nel, p, sig, jitter, G=1, seed=0):
    rng = np.random.default_rng(seed)
    vert = P.jittered_vertices(nel, jitter, rng) if jitter else P.cartesian_vertices(nel)
    om, w = ls_fixture(4, dim)
    ne = int(np.prod(nel))
    st = np.full((G, ne), sig) * rng.uniform(0.8, 1.2, (G, ne))
    Linv = A.InverseAdvectionOperator(A.Mesh(nel, vertices=vert), p, (om, w), st)
    pr = oracle.Problem(dim, nel, vert, p, om, w, st)
    inflow = np.zeros((G, 2 * dim))
    inflow[:, 0] = 1.0
    q = pr.form_source(Q=np.zeros((G, ne)), inflow=inflow)
    return Linv, pr, q, ne


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}d-{'x'.join(map(str, c[1]))}-p{c[2]}-{c[5].replace(' ', '_')}"
                                             f"-mode{c[7]}" for c in CASES])
def test_fixup_matches_oracle(case):
    dim, nel, p, sig, jitter, kind, mn, mode = case
    print(f"fixup {case}", flush=True)
    Linv, pr, q, ne = _problem(dim, nel, p, sig, jitter)
    Linv.set_launch(mode)
    plain = Linv.Mult(q, np.empty_like(q))
    assert plain.min() < mn - 1e-9
    mon_o = np.zeros(ne)
    pr.set_fixup(kind, mn, mon_o)
    ref = pr.sweep(q)
    mon = A.DeviceVector.from_host(np.zeros(ne))
    Linv.SetFixupOperator(kind, mn)
    Linv.SetFixupMonitorData(mon)
    assert Linv.IsFixupOn()
    psi = Linv.Mult(q, np.empty_like(q))
    counts, counts_o = Linv.fixup_counts(), pr.fixup_counts()
    print(f"  counts GPU {counts} oracle {counts_o}; rel {rel_err(psi, ref):.2e}")
    assert rel_err(psi, ref) < 1e-12 and psi.min() >= min(mn, 0.0)
    assert counts == counts_o[:2] and counts[0] > 0
    np.testing.assert_array_equal(mon.to_host(), mon_o)
    # a second sweep adds to the counts and the monitor
    Linv.Mult(q, np.empty_like(q))
    assert Linv.fixup_counts() == (2 * counts[0], 2 * counts[1])
    assert np.array_equal(mon.to_host(), 2 * mon_o)
    # UseFixup(false) gives the plain sweep back
    Linv.UseFixup(False)
    assert not Linv.IsFixupOn()
    assert rel_err(Linv.Mult(q, np.empty_like(q)), plain) < 1e-14


def test_fixup_multigroup_and_minimum():
    print("fixup multigroup / psi_min", flush=True)
    Linv, pr, q, ne = _problem(3, (6, 6, 6), 1, 30.0, 0.0, G=2, seed=3)
    for kind, mn in (("zero", 1e-4), ("zero and scale", 0.0)):
        mon_o = np.zeros(2 * ne)
        pr.set_fixup(kind, mn, mon_o)
        ref = pr.sweep(q)
        mon = A.DeviceVector.from_host(np.zeros(2 * ne))
        Linv.SetFixupOperator(kind, mn)
        Linv.SetFixupMonitorData(mon)
        psi = Linv.Mult(q, np.empty_like(q))
        assert rel_err(psi, ref) < 1e-12 and psi.min() >= mn
        assert Linv.fixup_counts() == pr.fixup_counts()[:2]
        np.testing.assert_array_equal(mon.to_host(), mon_o)


def test_source_iteration_with_fixup():
    """SI with the zero-and-scale fixup inside every sweep: the oracle's
    iterations and flux (the iteration is no longer linear)."""
    print("fixup SI", flush=True)
    Linv, pr, q, ne = _problem(2, (10, 10), 1, 40.0, 0.0)
    ss = np.full(ne, 40.0 * 0.9)
    Linv.SetFixupOperator("zero and scale")
    pr.set_fixup("zero and scale")
    phi, _, it, norms = A.source_iteration(Linv, ss, q, abstol=1e-11, maxit=500, want_psi=False)
    phi_o, _, it_o, norms_o = pr.source_iteration(ss, q, abstol=1e-11, maxit=500)
    print(f"  GPU {it} its, oracle {it_o}")
    assert it == it_o and rel_err(phi, phi_o) < 1e-10
    assert Linv.fixup_counts()[0] > 0


def test_fixup_errors():
    print("fixup errors", flush=True)
    Linv, pr, q, ne = _problem(3, (6, 6, 6), 1, 30.0, 0.0)
    Linv.SetFixupOperator("ryosuke")
    with pytest.raises(A.AlliumError, match="more than 3"):
        Linv.Mult(q, np.empty_like(q))
    Linv.SetFixupOperator("none")
    assert not Linv.IsFixupOn()
    with pytest.raises(KeyError):
        Linv.SetFixupOperator("local optimization")
    L2 = A.InverseAdvectionOperator(A.Mesh((3, 3, 3)), 2, 4, np.ones(27))
    with pytest.raises(A.AlliumError):
        L2.SetFixupOperator("zero")
