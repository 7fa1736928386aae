"""chive-style steady one-group source iteration on the GPU.

Reads a deck with the structure of chive's Lua inputs (exe/chive.cpp; e.g.
test/chive/tdl.lua) written as YAML or a dict, runs the outer-solve path of
exe/chive.cpp:550-750 (FormTransportSource, InverseAdvectionOperator,
SLISolver, GMRESSolver or FGMRESSolver on TransportOperator as
io::CreateIterativeSolver builds them (src/io.cpp:179-230), optional
"scalar mip", block "mip", "ldgsa" or "p1sa" DSA preconditioner, final sweep) and reports with chive's YAML keys ("transport iterations" with
it / norm / inner solver, "outer iterations", "inner iteration" max / min /
avg / total; exe/chive.cpp:40-75,738-746), so scripts that parse chive's
output read this one too.

Deck keys (YAML):

    mesh: {num_elements: [8, 8], extents: [1, 1]}
    materials: {mat: {total: 1.0e4, scattering: 9999.9999, source: 1.0e-4}}
    material_map: mat                 # one name, or a name per element (x fastest)
    boundary_conditions: {vacuum: {type: inflow, value: 0}}
    boundary_map: vacuum              # one name, or 2*dim names (side = axis*2 + {low, high})
    sn: {type: level symmetric, order: 4}
    driver:
      fe_order: 1
      lumping: 0                       # chive's lumping bits (1 mass, 2 gradient, 4 faces); 7 lumps all
      solver: {type: sli, abstol: 1.0e-6, reltol: 0, max_iter: 50}   # or gmres / fgmres with kdim
      preconditioner: {type: scalar mip, kappa: 4, solver: {type: cg, reltol: 1.0e-10, max_iter: 200}}
      # or {type: mip} (the block system) / {type: p1sa, solver: {type: direct}}; a "direct"
      # inner solve is CG (scalar mip) or GMRES (mip, p1sa) to a relative residual of 1e-12

Material `source` and boundary `value` are per-steradian densities, as in
chive (exe/chive.cpp:236-247,294-306): a number, or an expression of the
position (x, y, z) and direction (mu, eta, xi) standing in for chive's Lua
functions, with numpy's sin, cos, exp, log, sqrt, ..., pi, the deck's
`parameters` and its `definitions` (named expressions, evaluated in order).
`scalar_flux_solution` (an expression of x, y, z) and `current_solution` (one
expression per axis) make the driver print chive's "L2 error" of the scalar
flux and the current J = sum_a w_a Omega_a psi_a (exe/chive.cpp:1036-1072).
Boundary type "reflective" maps to a reflective side.

Usage: ``python -m allium_amd.driver deck.yaml`` prints the output YAML.
"""
import argparse
import sys
import time

import numpy as np
import yaml

import allium_amd as A
from allium_amd import fem


_MATH = {k: getattr(np, k) for k in ("sin", "cos", "tan", "exp", "log", "sqrt", "abs", "tanh", "sinh", "cosh",
                                     "arctan", "arcsin", "arccos", "minimum", "maximum", "where", "pi", "e")}


class _Expressions:
    """Deck expressions of (x, y, z, mu, eta, xi) with the deck's parameters and
    definitions, evaluated on arrays of points (..., dim) and one direction."""

    def __init__(self, deck):
        self.params = {str(k): float(v) for k, v in (deck.get("parameters") or {}).items()}
        self.defs = [(str(k), compile(str(v), f"<definition {k}>", "eval"))
                     for k, v in (deck.get("definitions") or {}).items()]

    def function(self, value, what):
        """A constant (a number) or a callable f(X, omega)."""
        if not isinstance(value, str):
            return float(value)
        code = compile(value, f"<{what}>", "eval")

        def f(X, om=None):
            ns = dict(_MATH)
            ns.update(self.params)
            X = np.asarray(X, float)
            for d, nm in enumerate(("x", "y", "z")):
                ns[nm] = X[..., d] if d < X.shape[-1] else np.zeros(X.shape[:-1])
            om = np.zeros(3) if om is None else np.asarray(om, float)
            for d, nm in enumerate(("mu", "eta", "xi")):
                ns[nm] = float(om[d]) if d < om.size else 0.0
            for nm, c in self.defs:
                ns[nm] = eval(c, {"__builtins__": {}}, ns)
            return np.broadcast_to(np.asarray(eval(code, {"__builtins__": {}}, ns), float), X.shape[:-1])

        return f


def _per_element(value, ne, what):
    if isinstance(value, str):
        return [value] * ne
    if len(value) != ne:
        raise ValueError(f"{what} needs one entry per element ({ne})")
    return list(value)


def run(deck):
    """Solve the deck; returns the output mapping (chive's keys)."""
    mesh_d = deck["mesh"]
    nel = tuple(int(v) for v in mesh_d["num_elements"])
    dim = len(nel)
    ext = tuple(float(v) for v in mesh_d.get("extents", [1.0] * dim))
    ne = int(np.prod(nel))
    mats = deck["materials"]
    ex = _Expressions(deck)
    names = _per_element(deck.get("material_map", next(iter(mats))), ne, "material_map")
    sigma_t = np.array([float(mats[m]["total"]) for m in names])
    sigma_s = np.array([float(mats[m].get("scattering", 0.0)) for m in names])
    src_f = {m: ex.function(mats[m].get("source", 0.0), f"materials.{m}.source") for m in mats}
    Q = np.array([src_f[m] if not callable(src_f[m]) else 0.0 for m in names])
    bcs = deck.get("boundary_conditions", {"vacuum": {"type": "inflow", "value": 0.0}})
    bmap = deck.get("boundary_map", next(iter(bcs)))
    side_names = [bmap] * (2 * dim) if isinstance(bmap, str) else list(bmap)
    if len(side_names) != 2 * dim:
        raise ValueError(f"boundary_map needs one name or {2 * dim} names")
    bc, inflow, inflow_f = [], [], {}
    for side, nm in enumerate(side_names):
        spec = bcs[nm]
        typ = str(spec.get("type", "inflow")).lower()
        if typ in ("reflective", "reflect"):
            bc.append(A.BC_REFLECTIVE)
            inflow.append(0.0)
        elif typ in ("inflow", "vacuum"):
            bc.append(A.BC_INFLOW)
            v = ex.function(spec.get("value", 0.0), f"boundary_conditions.{nm}.value")
            if callable(v):
                inflow_f[side] = v
                inflow.append(0.0)
            else:
                inflow.append(v)
        else:
            raise ValueError(f"boundary condition type {typ} is not supported")
    sn = deck.get("sn", {"type": "level symmetric", "order": 4})
    if "level" not in str(sn.get("type", "level symmetric")).lower():
        raise ValueError("only level-symmetric quadrature is supported")
    drv = deck.get("driver", {})
    p = int(drv.get("fe_order", 1))
    lump = int(drv.get("lumping", 0))   # chive's bit set: 1 mass, 2 gradient, 4 faces
    if not 0 <= lump <= 7:
        raise ValueError("lumping is a bit set in 0..7 (1 mass, 2 gradient, 4 faces)")
    sol = drv.get("solver", {"type": "sli"})
    stype = str(sol.get("type", "sli")).lower()
    if stype not in ("sli", "gmres", "fgmres"):
        raise ValueError(f"solver {stype} is not supported (sli, gmres, fgmres)")
    abstol, reltol = float(sol.get("abstol", 0.0)), float(sol.get("reltol", 0.0))
    maxit = int(sol.get("max_iter", 50))
    kdim = int(sol.get("kdim", 50))
    if stype != "sli" and not 1 <= kdim <= 200:
        raise ValueError("kdim must be in 1..200")
    prec = drv.get("preconditioner")
    ptype = str(prec.get("type", "")).lower() if prec else None
    if prec and ptype not in ("scalar mip", "mip", "p1sa", "ldgsa"):
        raise ValueError(f"preconditioner {prec.get('type')} is not supported (scalar mip, mip, p1sa, ldgsa)")
    # chive's inner solver defaults to "direct"; p1sa takes nothing else (exe/chive.cpp:502-524,609)
    itype = str(prec.get("solver", {}).get("type", "direct")).lower() if prec else None
    if ptype == "p1sa" and itype not in ("direct", "superlu"):
        raise ValueError("only direct available for P1")

    out = {}
    t0 = time.perf_counter()
    mesh = A.Mesh(nel, extents=ext, bc=bc)
    Linv = A.InverseAdvectionOperator(mesh, p, int(sn.get("order", 4)), sigma_t.reshape(1, -1), lump=lump)
    out["lumping type"] = {"mass": bool(lump & 1), "gradient": bool(lump & 2), "faces": bool(lump & 4)}
    Linv.set_launch(-1)
    q = Linv.form_source(Q=Q.reshape(1, -1), inflow=np.array(inflow).reshape(1, -1))
    space = None
    funcs = [m for m in mats if callable(src_f[m])]
    if funcs or inflow_f or "scalar_flux_solution" in deck or "current_solution" in deck:
        space = fem.BoxSpace(nel, p, extents=ext)
    if funcs or inflow_f:
        # function sources and inflows: FormTransportSource's integrals on the host
        om, _ = Linv.quadrature()
        qa = q.reshape(len(om), ne, -1)
        for m in funcs:
            sel = np.array([nm == m for nm in names])
            qa[:, sel] += space.project_source(src_f[m], om).reshape(len(om), ne, -1)[:, sel]
        for side, f in inflow_f.items():
            qa += space.project_inflow(f, om, [side]).reshape(len(om), ne, -1)
    accel = None
    out["preconditioner"] = "none"
    if ptype == "p1sa":
        # the direct solve is emulated by GMRES to a relative residual of 1e-12
        accel = A.DSA(Linv, sigma_s, method="p1sa", cg_reltol=1e-12, cg_maxit=20000)
        out["preconditioner"] = {"type": "p1sa", "solver": {"type": itype}}
    elif ptype == "ldgsa":
        # chive's block LDG system (exe/chive.cpp:620-634), solved as for "mip"
        inner = prec.get("solver", {})
        direct = itype in ("direct", "superlu")
        accel = A.DSA(Linv, sigma_s, method="ldgsa", cg_reltol=1e-12 if direct else float(inner.get("reltol", 1e-10)),
                      cg_abstol=0.0 if direct else float(inner.get("abstol", 0.0)),
                      cg_maxit=20000 if direct else int(inner.get("max_iter", 20000)))
        out["preconditioner"] = {"type": "ldgsa", "solver": {"type": itype}}
    elif ptype == "mip":
        # chive's block interior-penalty system (exe/chive.cpp:635-662); the reference
        # solves its Schur complement (CG + BoomerAMG or direct), here GMRES solves the
        # block system, whose phi-part is that solution, to a relative residual of
        # the inner reltol (1e-12 for "direct")
        inner = prec.get("solver", {})
        rt = 1e-12 if itype in ("direct", "superlu") else float(inner.get("reltol", 1e-10))
        kappa = float(prec.get("kappa", (p + 1) ** 2))
        scale = bool(prec.get("scale_stabilization", True))
        lower = bool(prec.get("bound_stabilization_below", True))
        accel = A.DSA(Linv, sigma_s, method="block mip", kappa=kappa, bound_below=lower, scale_penalty=scale,
                      cg_reltol=rt, cg_abstol=float(inner.get("abstol", 0.0)),
                      cg_maxit=int(inner.get("max_iter", 20000)) if itype not in ("direct", "superlu") else 20000)
        out["preconditioner"] = {"type": "mip", "kappa": kappa, "scale stabilization": scale,
                                 "bound stabilization from below": lower, "solver": {"type": itype}}
    elif prec:
        inner = prec.get("solver", {})
        if itype in ("direct", "superlu"):
            inner = {"reltol": 1e-12, "max_iter": 1000}
        kappa = float(prec.get("kappa", (p + 1) ** 2))
        accel = A.DSA(Linv, sigma_s, kappa=kappa, bound_below=bool(prec.get("bound_stabilization_below", True)),
                      sym=float(prec.get("sigma", -1.0)), cg_reltol=float(inner.get("reltol", 1e-10)),
                      cg_abstol=float(inner.get("abstol", 0.0)), cg_maxit=int(inner.get("max_iter", 200)))
        out["preconditioner"] = {"type": "scalar mip", "kappa": kappa,
                                 "solver": {"type": "cg", "reltol": float(inner.get("reltol", 1e-10)),
                                            "abstol": float(inner.get("abstol", 0.0)),
                                            "max_iter": int(inner.get("max_iter", 200))}}
    out["setup time"] = f"{time.perf_counter() - t0:.3e} s"
    out["solver"] = {"type": stype, "abstol": abstol, "reltol": reltol, "max_iter": maxit}
    if stype != "sli":
        out["solver"]["kdim"] = kdim
    inner_its = np.zeros(maxit, dtype=np.int32)
    t1 = time.perf_counter()
    if stype == "sli":
        phi, psi, it, norms = A.source_iteration(Linv, sigma_s, q, abstol=abstol, reltol=reltol, maxit=maxit,
                                                 accel=accel, inner_iterations=inner_its)
        # SLISolver's monitor: iteration k reports the norm of step k
        steps = [(k + 1, norms[k]) for k in range(it)]
        converged = bool(norms[-1] < max(reltol * norms[0], abstol))
    else:
        phi, psi, it, norms = A.krylov_solve(Linv, sigma_s, q, solver=stype, kdim=kdim, abstol=abstol, reltol=reltol,
                                             maxit=maxit, accel=accel, inner_iterations=inner_its)
        # mfem's Krylov monitors report the initial residual as iteration 0
        steps = [(k, norms[k]) for k in range(norms.size)]
        converged = bool(norms[-1] <= max(reltol * norms[0], abstol))
    solve = time.perf_counter() - t1
    rows = []
    for k, nrm in steps:
        row = {"it": k, "norm": f"{nrm:.3e}"}
        if accel is not None and k >= 1:
            row["inner solver"] = {"it": int(inner_its[k - 1])}
        rows.append(row)
    out["transport iterations"] = rows
    out["outer iterations"] = int(it)
    if accel is not None:
        ii = inner_its[:max(1, min(it, maxit))]
        out["inner iteration"] = {"max": int(ii.max()), "min": int(ii.min()), "avg": float(ii.mean()),
                                  "total": int(ii.sum())}
    out["converged"] = converged
    # chive's TimingLog / EventLog (exe/chive.cpp:1286-1294): the sweeps' GPU time and count
    sweep_s, sweeps = Linv.timing()
    out["timing log"] = {"sweep": f"{sweep_s:.3e} s"}
    out["event log"] = {"sweeps": int(sweeps)}
    out["timings"] = {"setup": out["setup time"], "solve": {"total": f"{solve:.3e} s", "sweep": f"{sweep_s:.3e} s"}}
    out["scalar flux"] = {"l2": float(np.linalg.norm(phi)), "min": float(phi.min()), "max": float(phi.max())}
    out["balance"] = float(A.AdvectionOperator(Linv).ComputeBalance(psi))
    # chive's "L2 error" (exe/chive.cpp:1036-1072), three significant digits
    errs = {}
    if "scalar_flux_solution" in deck:
        f = ex.function(str(deck["scalar_flux_solution"]), "scalar_flux_solution")
        errs["scalar flux"] = float(f"{space.l2_error(phi, f):.3e}")
    if "current_solution" in deck:
        comps = deck["current_solution"]
        if len(comps) != dim:
            raise ValueError(f"current_solution needs {dim} expressions")
        om, w = Linv.quadrature()
        pa = psi.reshape(len(w), -1)
        e2 = 0.0
        for d in range(dim):
            Jd = (w * om[:, d]) @ pa
            e2 += space.l2_error(Jd, ex.function(str(comps[d]), f"current_solution[{d}]")) ** 2
        errs["current"] = float(f"{np.sqrt(e2):.3e}")
    if errs:
        out["L2 error"] = errs
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="chive-style source iteration on the GPU")
    ap.add_argument("deck", help="YAML deck")
    a = ap.parse_args(argv)
    with open(a.deck) as fh:
        deck = yaml.safe_load(fh)
    sys.stdout.write(yaml.safe_dump(run(deck), sort_keys=False))


if __name__ == "__main__":
    main()
