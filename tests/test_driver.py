"""The chive-style driver (allium_amd/driver.py): deck validation on the CPU;
the reference's thick-diffusion-limit acceptance check on the GPU, with
source iteration and with the Krylov outer solvers its decks also name
(test/chive/mms.lua, test/chive/easy_cp.lua: type = "gmres")."""
import copy
import os

import numpy as np
import pytest
import yaml

import oracle
from allium_amd import driver
from allium_amd import problems as P
from tests.helpers import ls_fixture

DECK = os.path.join(os.path.dirname(os.path.abspath(__file__)), "decks", "tdl.yaml")


def _deck():
    with open(DECK) as fh:
        return yaml.safe_load(fh)


@pytest.mark.parametrize("edit,msg", [
    (lambda d: d["driver"]["solver"].update(type="bicgstab"), "not supported"),
    (lambda d: d["driver"]["solver"].update(type="gmres", kdim=0), "kdim"),
    (lambda d: d["sn"].update(type="abu shumays"), "level-symmetric"),
    (lambda d: d.update(material_map=["mat"] * 3), "one entry per element"),
    (lambda d: d.update(boundary_map=["vacuum"] * 3), "boundary_map"),
    (lambda d: d["boundary_conditions"]["vacuum"].update(type="periodic"), "not supported"),
    (lambda d: d["driver"].update(preconditioner={"type": "h1sa"}), "not supported"),
    (lambda d: d["driver"].update(preconditioner={"type": "p1sa", "solver": {"type": "cg"}}), "only direct"),
    (lambda d: d["driver"].update(lumping=9), "lumping is a bit set"),
])
def test_deck_errors(edit, msg):
    d = _deck()
    edit(d)
    with pytest.raises(ValueError, match=msg):
        driver.run(d)


@pytest.mark.gpu
def test_tdl_deck_outer_iterations():
    """test/chive/test_tdl.py asserts `outer iterations < 12` for the TDL deck
    with a DSA preconditioner; plain SI runs to max_iter there."""
    print("driver tdl", flush=True)
    d = _deck()
    plain = driver.run(copy.deepcopy(d))
    assert plain["outer iterations"] == 50 and not plain["converged"]
    d["driver"]["preconditioner"] = {"type": "scalar mip", "solver": {"type": "cg", "reltol": 1e-12,
                                                                      "max_iter": 400}}
    out = driver.run(d)
    print(yaml.safe_dump({k: out[k] for k in ("outer iterations", "inner iteration")}))
    assert out["converged"] and out["outer iterations"] < 12
    rows = out["transport iterations"]
    assert len(rows) == out["outer iterations"] and rows[0]["it"] == 1
    assert float(rows[-1]["norm"]) < 1e-6 and rows[-1]["inner solver"]["it"] > 0
    assert out["inner iteration"]["total"] == sum(r["inner solver"]["it"] for r in rows)
    assert np.isfinite(out["balance"])
    # TimingLog "sweep" and EventLog "sweeps": at least one sweep per iteration plus the
    # first and last (launch autotuning may add trial sweeps)
    assert out["event log"]["sweeps"] >= out["outer iterations"] + 2
    assert float(out["timing log"]["sweep"].split()[0]) > 0


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["p1sa", "mip", "ldgsa"])
def test_tdl_deck_direct_inner_solver(prec):
    """test/chive/CMakeLists.txt:117-121: the TDL deck with
    `preconditioner = {type = "<dsa>", solver = {type = "direct"}}` for p1sa,
    ldgsa and mip (the block interior-penalty system) needs fewer than 12 outer
    iterations."""
    print(f"driver tdl {prec} direct", flush=True)
    d = _deck()
    d["driver"]["preconditioner"] = {"type": prec, "solver": {"type": "direct"}}
    out = driver.run(d)
    print(yaml.safe_dump({k: out[k] for k in ("outer iterations", "inner iteration", "preconditioner")}))
    assert out["converged"] and out["outer iterations"] < 12
    assert out["preconditioner"]["type"] == prec


@pytest.mark.gpu
@pytest.mark.parametrize("solver", ["gmres", "fgmres"])
def test_krylov_decks(solver):
    """A gmres / fgmres deck reaches the fixed point source iteration reaches;
    with the scalar MIP preconditioner the TDL deck needs fewer than 12
    iterations, as chive's acceptance check asks of SLI there."""
    print(f"driver {solver}", flush=True)
    d = _deck()
    d["materials"]["mat"].update(total=4.0, scattering=3.96, source=1.0)
    d["driver"]["solver"] = {"type": "sli", "abstol": 1e-10, "max_iter": 3000}
    ref = driver.run(copy.deepcopy(d))
    d["driver"]["solver"] = {"type": solver, "abstol": 1e-10, "max_iter": 300, "kdim": 10}
    out = driver.run(copy.deepcopy(d))
    assert out["converged"] and ref["converged"]
    assert out["outer iterations"] < ref["outer iterations"] / 2
    assert out["scalar flux"]["l2"] == pytest.approx(ref["scalar flux"]["l2"], rel=1e-8)
    rows = out["transport iterations"]
    assert rows[0]["it"] == 0 and out["solver"]["kdim"] == 10
    t = _deck()
    t["driver"]["solver"] = {"type": solver, "abstol": 1e-6, "max_iter": 50}
    t["driver"]["preconditioner"] = {"type": "scalar mip", "solver": {"type": "cg", "reltol": 1e-12,
                                                                      "max_iter": 400}}
    out = driver.run(t)
    print(yaml.safe_dump({k: out[k] for k in ("outer iterations", "inner iteration")}))
    assert out["converged"] and out["outer iterations"] < 12
    assert out["inner iteration"]["total"] > 0


MMS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "decks", "mms.yaml")


def test_deck_expressions():
    """Deck expressions stand in for chive's Lua functions: parameters and
    definitions in order, numpy's math, no builtins."""
    with open(MMS) as fh:
        d = yaml.safe_load(fh)
    ex = driver._Expressions(d)
    X = np.array([[0.1, 0.2], [0.3, 0.7]])
    om = np.array([0.35002117, 0.35002117, 0.8688903])
    psi = ex.function(d["definitions"]["psi"], "psi")(X, om)
    x, y = X[:, 0], X[:, 1]
    mu, eta = om[0], om[1]
    ref = (np.sin(np.pi * x) * np.sin(np.pi * y) + 0.5 * (mu + eta) * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y)
           + 0.25 * (mu ** 2 + eta ** 2) * np.sin(3 * np.pi * x) * np.sin(3 * np.pi * y) + 2.0) / 4 / np.pi
    np.testing.assert_allclose(psi, ref, rtol=1e-15)
    assert ex.function(2.5, "c") == 2.5
    with pytest.raises(NameError):
        ex.function("open('x')", "bad")(X, om)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["mip", "ldgsa", "p1sa"])
def test_mms_deck_order_of_accuracy(prec):
    """test/chive/test_ooa.py on test/chive/mms.lua as chive.mms.{mip, ldgsa,
    p1sa} run it (test/chive/CMakeLists.txt:31-48): GMRES to 1e-10 with the
    preconditioner and a direct inner solve, num_elements {Ne, Ne} for Ne = 10
    and 20; the "L2 error" of the scalar flux and of the current both fall at
    order 2 within 0.2."""
    print(f"driver mms {prec}", flush=True)
    errs = []
    for ne in (10, 20):
        with open(MMS) as fh:
            d = yaml.safe_load(fh)
        d["mesh"]["num_elements"] = [ne, ne]
        d["driver"]["preconditioner"] = {"type": prec, "solver": {"type": "direct"}}
        out = yaml.safe_load(yaml.safe_dump(driver.run(d), sort_keys=False))   # through the printed YAML
        assert out["converged"]
        errs.append((out["L2 error"]["scalar flux"], out["L2 error"]["current"]))
    phi_ooa = np.log2(errs[0][0] / errs[1][0])
    J_ooa = np.log2(errs[0][1] / errs[1][1])
    print(f"  phi {phi_ooa:.3f} ({errs[0][0]:.3e}, {errs[1][0]:.3e}), J {J_ooa:.3f} ({errs[0][1]:.3e}, {errs[1][1]:.3e})")
    assert abs(phi_ooa - 2) < 0.2 and abs(J_ooa - 2) < 0.2


@pytest.mark.gpu
def test_lumped_deck_matches_oracle():
    """chive's `driver.lumping = 7` (exe/chive.cpp:532-568,594): a scattering
    deck with every matrix on the vertex rule, the scattering mass included,
    converges to the oracle's lumped source iteration and reports chive's
    "lumping type" keys."""
    print("driver lumped deck", flush=True)
    d = {"mesh": {"num_elements": [8, 8], "extents": [1, 1]},
         "materials": {"mat": {"total": 1.0, "scattering": 0.5, "source": 1.0}},
         "material_map": "mat", "boundary_conditions": {"vacuum": {"type": "inflow", "value": 0}},
         "boundary_map": "vacuum", "sn": {"type": "level symmetric", "order": 4},
         "driver": {"fe_order": 1, "lumping": 7, "solver": {"type": "sli", "abstol": 1e-12, "max_iter": 200}}}
    out = driver.run(d)
    assert out["converged"] and out["lumping type"] == {"mass": True, "gradient": True, "faces": True}
    om, w = ls_fixture(4, 2)
    pr = oracle.Problem(2, (8, 8), P.cartesian_vertices((8, 8)), 1, om, w, np.ones((1, 64)), lump=7)
    phi, _, it, _ = pr.source_iteration(0.5, pr.form_source(Q=1.0), abstol=1e-12, maxit=200)
    print(f"  driver {out['outer iterations']} its, oracle {it} its")
    assert out["outer iterations"] == it
    assert abs(out["scalar flux"]["l2"] - np.linalg.norm(phi)) < 1e-10 * np.linalg.norm(phi)
