"""GPU checks of the multigroup linearized TRT step (allium_amd/csrc/trt.hip)
against the CPU oracle (oracle/trt_oracle.py): temperature and scalar flux
after each step within 1e-9 relative (FP64; the sweeps agree to 1e-12 and the
source iterations to their 1e-10 tolerance), and the radiative-equilibrium
invariant on the GPU."""
import numpy as np
import pytest

from allium_amd import trt
from oracle import trt_oracle as O
from tests.helpers import ls_fixture, rel_err

pytestmark = pytest.mark.gpu


def test_trt_steps_match_oracle():
    print("trt parity", flush=True)
    nel, G = (4, 3, 3), 4
    om, w = ls_fixture(4, 3)
    kw = dict(G=G, emin=1e-2, emax=1e1, coef=10.0, cv=1.0, dt=1e-3, T0=0.01, T_inflow=1.0, inflow_sides=1,
              si_reltol=1e-10)
    o = O.TRTOracle(nel, om, w, **kw)
    g = trt.TRT(nel, sn=4, groups=G, emin=1e-2, emax=1e1, opacity_coef=10.0, cv=1.0, dt=1e-3, T0=0.01, T_inflow=1.0,
                inflow_sides=1, si_reltol=1e-10)
    for step in range(3):
        o.step()
        info = g.step()
        eT = rel_err(g.temperature(), o.T.ravel())
        ep = rel_err(g.scalar_flux(), o.phi)
        print(f"  step {step + 1}: T err {eT:.2e}, phi err {ep:.2e}, its gpu {info['transport_iterations']} "
              f"oracle {o.iterations[-1]}, T in [{info['T_min']:.4g}, {info['T_max']:.4g}]")
        assert eT < 1e-9 and ep < 1e-9
        assert abs(info["transport_iterations"] - o.iterations[-1]) <= 1


def test_trt_keeps_radiative_equilibrium_on_gpu():
    print("trt equilibrium", flush=True)
    g = trt.TRT((3, 3, 2), sn=4, groups=4, emin=1e-3, emax=1e2, opacity_coef=0.1, T0=1.0, T_inflow=1.0,
                inflow_sides=63, dt=1e-2, si_reltol=1e-12)
    g.step()
    T = g.temperature()
    assert np.abs(T - 1.0).max() < 1e-9
    b, _ = trt.log_groups(1e-3, 1e2, 4)
    eq = np.diff([trt.planck_integrals(x)[0] for x in b])
    phi = g.scalar_flux().reshape(4, -1)
    assert np.allclose(phi[1:], eq[1:, None], rtol=1e-9, atol=0)


def test_trt_config4_steps():
    """32^3 hex, p=1, S6, 16 groups on [1e-2, 1e1], sigma_g = 10/(e_g^3 sqrt T),
    c_v = 1, T0 = 0.01, Planckian T = 1 inflow on x = 0, dt = 1e-3: two steps
    run, stay finite and heat the material from the inflow side (the DG p=1
    temperature may dip slightly below T0 ahead of the heating front: no
    negative-flux fixup is applied, as in the reference without one)."""
    print("trt config 4", flush=True)
    g = trt.TRT((32, 32, 32), sn=6, groups=16, si_reltol=1e-8)
    g.set_launch(-1)
    for _ in range(2):
        info = g.step()
        print("  ", {k: info[k] for k in ("step", "transport_iterations", "T_min", "T_max", "radiation_energy")})
    T = g.temperature().reshape(32, 32, 32, 8)
    assert np.isfinite(T).all() and T.min() > 0.0
    assert T[:, :, 0].mean() > T[:, :, -1].mean()


@pytest.mark.parametrize("implicit", [False, True])
def test_trt_outer_fixed_point_matches_oracle(implicit):
    """scallion's outer FixedPointIterationSolver on T (src/fixed_point.cpp:134-172)
    around the linearized step, with explicit or implicit opacities
    (OpacityUpdate, src/opacity_update.hpp): same passes as the oracle, same
    T and phi; the converged step differs from the single linearized pass."""
    print(f"trt outer fixed point, implicit opacity {implicit}", flush=True)
    nel, G = (4, 3, 3), 4
    om, w = ls_fixture(4, 3)
    kw = dict(G=G, emin=1e-2, emax=1e1, coef=10.0, cv=1.0, dt=1e-2, T0=0.01, T_inflow=1.0, inflow_sides=1,
              si_reltol=1e-12)
    outer = dict(outer_maxit=40, outer_abstol=1e-11, implicit_opacity=implicit)
    o = O.TRTOracle(nel, om, w, **kw, **outer)
    one = O.TRTOracle(nel, om, w, **kw)
    g = trt.TRT(nel, sn=4, groups=G, emin=1e-2, emax=1e1, opacity_coef=10.0, cv=1.0, dt=1e-2, T0=0.01, T_inflow=1.0,
                inflow_sides=1, si_reltol=1e-12, **outer)
    for step in range(3):
        o.step()
        one.step()
        info = g.step()
        eT = rel_err(g.temperature(), o.T.ravel())
        ep = rel_err(g.scalar_flux(), o.phi)
        print(f"  step {step + 1}: passes gpu {info['outer_iterations']} oracle {o.outer_iterations[-1]}, "
              f"last change {info['outer_residual']:.2e}; T err {eT:.2e}, phi err {ep:.2e}; "
              f"vs one pass {rel_err(o.T, one.T):.2e}")
        assert eT < 1e-9 and ep < 1e-9
        assert abs(info["outer_iterations"] - o.outer_iterations[-1]) <= 1 and info["outer_iterations"] > 2
        assert info["outer_residual"] < 1e-11
    assert rel_err(o.T, one.T) > 1e-6


def test_trt_fixup_only_in_the_final_sweep():
    """A sweep fixup set on the TRT handle is off for the transport solve and
    on for the final sweep of each pass (src/trt_linearized.cpp:236-238,284-289)."""
    print("trt fixup", flush=True)
    import ctypes
    from allium_amd import _lib
    nel, G = (4, 3, 3), 2
    om, w = ls_fixture(4, 3)
    kw = dict(G=G, emin=1e-2, emax=1e1, coef=10.0, cv=1.0, dt=1e-3, T0=0.01, T_inflow=1.0, inflow_sides=1,
              si_reltol=1e-12)
    o = O.TRTOracle(nel, om, w, **kw, fixup=("zero and scale", 0.0))
    g = trt.TRT(nel, sn=4, groups=G, emin=1e-2, emax=1e1, opacity_coef=10.0, cv=1.0, dt=1e-3, T0=0.01, T_inflow=1.0,
                inflow_sides=1, si_reltol=1e-12)
    lib = _lib.load()
    L = lib.allium_trt_sweep(g._h)
    _lib.check(lib.allium_sweep_set_fixup(L, 2, ctypes.c_double(0.0)))
    applied = ctypes.c_longlong()
    for step in range(2):
        o.step()
        g.step()
        assert rel_err(g.temperature(), o.T.ravel()) < 1e-9 and rel_err(g.scalar_flux(), o.phi) < 1e-9
    _lib.check(lib.allium_sweep_fixup_counts(L, ctypes.byref(applied), None))
    print(f"  fixup applications on the GPU: {applied.value}")
    on = ctypes.c_int()
    _lib.check(lib.allium_sweep_get_fixup(L, None, None, ctypes.byref(on)))
    assert on.value == 1


def _gpu_identities(e):
    """(R) final-sweep balance and (M) material update defects, relative."""
    dR = abs(e["absorption"] + e["time_absorption"] + e["leakage"] - e["source"] - e["emission_linear"]) / \
        (e["source"] + e["emission_linear"])
    dM = abs(e["material"] - e["floor"] - (e["absorption_si"] - e["gray_linear"])) / max(abs(e["absorption_si"]),
                                                                                          abs(e["gray_linear"]))
    return dR, dM


def test_trt_energy_terms_match_oracle():
    """The GPU's energy bookkeeping (allium_trt_energy) equals the oracle's
    terms, and scallion's balance value, over three off-equilibrium steps."""
    print("trt energy parity", flush=True)
    nel, G = (4, 3, 3), 4
    om, w = ls_fixture(4, 3)
    o = O.TRTOracle(nel, om, w, G=G, T0=0.01, T_inflow=1.0, inflow_sides=1, si_reltol=1e-12)
    g = trt.TRT(nel, sn=4, groups=G, T0=0.01, T_inflow=1.0, inflow_sides=1, si_reltol=1e-12)
    g.track_energy(True)
    for step in range(3):
        o.step()
        g.step()
        eo, eg = o.energy_terms(), g.energy_terms()
        for k in ("absorption", "time_absorption", "leakage", "source", "inflow", "emission_linear",
                  "absorption_si", "material", "floor", "emission"):
            assert abs(eg[k] - eo[k]) <= 1e-10 * max(abs(eo[k]), 1e-3 * eo["source"]), (step, k, eg[k], eo[k])
        assert abs(eg["balance"] - eo["balance"]) < 1e-9
        dR, dM = _gpu_identities(eg)
        print(f"  step {step + 1}: R {dR:.1e} M {dM:.1e} balance {eg['balance']:.3e}")
        assert dR < 1e-11 and dM < 1e-11


def test_trt_config4_conserves_energy_over_20_steps():
    """Config 4 (32^3, p=1, S6, 16 groups, T0 = 0.01, Planckian T = 1 inflow on
    x = 0, dt = 1e-3) for its 20 steps: each step's final-sweep balance and
    material update hold to rounding, and total energy,
      (E_mat + E_rad)^{n+1} - (E_mat + E_rad)^n = dt (inflow - leakage)
        + dt (emission-rule mismatch) + dt (source-iteration defect),
    holds to rounding.  The two remaining terms are the discretization's, not
    errors, and stay small: the mismatch of the rule 2p+1 group emission against
    the rule 2p+2 gray emission (the reference's two integrators; it grows to
    ~1e-4 of the inflow as the heating front steepens) below 1e-2 of the inflow,
    and the SI defect (phi of the source iteration, stopped at 1e-8 of its first
    change, against D psi of the final sweep) below 1e-4 of it."""
    print("trt config 4 energy", flush=True)
    g = trt.TRT((32, 32, 32), sn=6, groups=16, si_reltol=1e-8)
    g.set_launch(-1)
    g.track_energy(True)
    for step in range(20):
        info = g.step()
        e = g.energy_terms()
        dR, dM = _gpu_identities(e)
        lhs = e["material"] - e["floor"] + e["time_absorption"] - (e["source"] - e["inflow"]) + e["leakage"] - \
            e["inflow"]
        mismatch = e["emission_linear"] - e["gray_linear"]
        si = e["absorption_si"] - e["absorption"]
        dE = abs(lhs - mismatch - si) / e["source"]
        if step % 4 == 3 or step == 0:
            print(f"  step {step + 1}: R {dR:.1e} M {dM:.1e} E {dE:.1e} mismatch/inflow {mismatch / e['inflow']:.1e} "
                  f"si/inflow {si / e['inflow']:.1e} floored {info['floored']} floor/inflow "
                  f"{e['floor'] / e['inflow']:.1e} balance {e['balance']:.3e} its {info['transport_iterations']}",
                  flush=True)
        assert (info["floored"] == 0) == (e["floor"] == 0)
        assert dR < 1e-10 and dM < 1e-10 and dE < 1e-10
        assert abs(mismatch) < 1e-2 * e["inflow"] and abs(si) < 1e-4 * e["inflow"]
