"""CPU pins of the oracle's negative-flux fixups (src/fixup.cpp, applied after
each local solve at src/sweep.cpp:465-469).  The reference's tests never switch
a fixup on (the one deck that names one has it commented out,
test/scallion/marshak_thin.lua:107-110), so the fixups are pinned by what
defines them, checked through the oracle's independent AdvectionOperator:

* every fixup leaves psi >= psi_min;
* "zero and scale" and "ryosuke" restore each cell's balance
  1^T (A psi_e - rhs) = 0, i.e. the per-cell sum of L psi - q vanishes; "zero"
  does not (it adds the clipped mass);
* cells that needed no fixup keep the plain solve, so without negative
  values the fixup changes nothing;
* the monitor counts every application, successful or failed;
* "ryosuke" supports at most 3 positive unknowns per cell (the reference
  aborts beyond); the oracle reports that case.
No GPU compute is called."""
import numpy as np
import pytest

import oracle
from allium_amd import problems as P
from tests.helpers import ls_fixture


def _thick(dim, nel, p, sig, order=4):
    """Optically thick cells lit from the x-low side: DG psi dips below zero."""
    vert = P.cartesian_vertices(nel)
    om, w = ls_fixture(order, dim)
    ne = int(np.prod(nel))
    pr = oracle.Problem(dim, nel, vert, p, om, w, np.full(ne, sig))
    inflow = np.zeros((1, 2 * dim))
    inflow[0, 0] = 1.0
    return pr, pr.form_source(Q=0.0, inflow=inflow), len(w), ne


def _cell_balance(pr, psi, q, na, ne):
    return (pr.apply_L(psi) - q).reshape(na, ne, pr.n).sum(-1)


@pytest.mark.parametrize("dim,nel,p,sig", [(2, (8, 8), 1, 50.0), (2, (8, 8), 2, 50.0), (3, (6, 6, 6), 1, 30.0)])
def test_zero_and_scale(dim, nel, p, sig):
    pr, q, na, ne = _thick(dim, nel, p, sig)
    plain = pr.sweep(q)
    assert plain.min() < -1e-3
    mon = np.zeros(ne)
    pr.set_fixup("zero and scale", 0.0, mon)
    psi = pr.sweep(q)
    applied, failed, err = pr.fixup_counts()
    assert applied > 0 and failed == 0 and not err
    assert psi.min() >= 0.0
    assert np.abs(_cell_balance(pr, psi, q, na, ne)).max() < 1e-13
    assert mon.sum() == applied + failed


@pytest.mark.parametrize("dim,nel,p,sig", [(2, (8, 8), 1, 50.0), (3, (6, 6, 6), 1, 30.0)])
def test_zero_clips_without_rebalancing(dim, nel, p, sig):
    pr, q, na, ne = _thick(dim, nel, p, sig)
    pr.set_fixup("zero", 0.0)
    psi = pr.sweep(q)
    applied, failed, err = pr.fixup_counts()
    assert applied > 0 and failed == 0
    assert psi.min() >= 0.0
    # clipping adds mass where it applied: positive cell residuals, none negative
    cb = _cell_balance(pr, psi, q, na, ne)
    assert cb.max() > 1e-4 and cb.min() > -1e-13
    # a minimum above the solution's clips everything to it
    pr.set_fixup("zero", 1e3)
    assert pr.sweep(q).min() == 1e3


def test_ryosuke_restores_balance_and_keeps_ratios():
    """2D p=1 (4 unknowns per cell, so at most 3 positive after clipping)."""
    pr, q, na, ne = _thick(2, (8, 8), 1, 20.0)
    mon = np.zeros(ne)
    pr.set_fixup("ryosuke", 0.0, mon)
    psi = pr.sweep(q)
    applied, failed, err = pr.fixup_counts()
    assert applied > 0 and not err
    assert psi.min() >= 0.0
    assert np.abs(_cell_balance(pr, psi, q, na, ne)).max() < 1e-13
    assert mon.sum() == applied + failed
    # the rebalanced positive unknowns keep the ratios of the plain local solve:
    # check on the first cell (in sweep order) that needed a fixup, whose
    # upwind data the fixup has not touched
    pr.set_fixup("none")
    plain = pr.sweep(q)
    lv = pr.levels()
    cells = psi.reshape(na, ne, 4)
    pcells = plain.reshape(na, ne, 4)
    changed = np.argwhere(np.abs(cells - pcells).max(-1) > 1e-14)
    first = min(changed, key=lambda ae: lv[ae[1], ae[0]])
    a, e = first
    assert lv[e, a] == min(lv[ee, aa] for aa, ee in changed)
    x0, x1 = pcells[a, e], cells[a, e]
    keep = x0 > 0
    assert 0 < keep.sum() < 4 and np.all(x1[~keep] == 0.0)
    r0, r1 = x0[keep] / x0[keep][0], x1[keep] / x1[keep][0]
    np.testing.assert_allclose(r1, r0, rtol=1e-12)


def test_ryosuke_reports_more_than_three_positive_unknowns():
    pr, q, na, ne = _thick(3, (6, 6, 6), 1, 30.0)
    pr.set_fixup("ryosuke", 0.0)
    pr.sweep(q)
    assert pr.fixup_counts()[2]


def test_fixup_changes_nothing_without_negative_values():
    nel = (6, 6)
    vert = P.cartesian_vertices(nel)
    om, w = ls_fixture(4, 2)
    pr = oracle.Problem(2, nel, vert, 1, om, w, np.full(36, 0.5))
    q = pr.form_source(Q=1.0)
    plain = pr.sweep(q)
    assert plain.min() > 0
    for kind in ("zero", "zero and scale", "ryosuke"):
        pr.set_fixup(kind, 0.0)
        assert np.array_equal(pr.sweep(q), plain)
        assert pr.fixup_counts() == (0, 0, False)
