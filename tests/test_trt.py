"""CPU checks of the TRT pieces: the Planck/Rosseland integrals (library and
oracle) against numerical quadrature of their integrands, the log-spaced
energy grid, and the oracle's radiative-equilibrium invariant."""
import numpy as np
import pytest
from scipy import integrate

from allium_amd import trt
from oracle import trt_oracle as O
from tests.helpers import ls_fixture

NORM = 15.0 / np.pi ** 4


def _quad_planck(x):
    f = lambda t: t ** 3 / np.expm1(t) if t > 0 else 0.0
    return NORM * integrate.quad(f, 0.0, x, epsabs=1e-15, epsrel=1e-13, limit=200)[0]


def _quad_rosseland(x):
    f = lambda t: t ** 4 * np.exp(-t) / (-np.expm1(-t)) ** 2 if t > 0 else 0.0
    return 0.25 * NORM * integrate.quad(f, 0.0, x, epsabs=1e-15, epsrel=1e-13, limit=200)[0]


@pytest.mark.parametrize("x", [1e-4, 1e-2, 0.3, 1.0, 1.9, 2.5, 4.0, 10.0, 35.0, 200.0])
def test_planck_integrals_match_quadrature(x):
    p, r = trt.planck_integrals(x)
    # the series/polylog switch is accurate to a few 1e-9 where the two
    # approximations cross (x ~ 2); the library and oracle agree to rounding
    assert abs(p - _quad_planck(x)) < 5e-9
    assert abs(r - _quad_rosseland(x)) < 5e-9
    assert abs(p - float(O.planck_integral(x))) < 1e-15
    assert abs(r - float(O.rosseland_integral(x))) < 1e-15


def test_planck_limits():
    assert trt.planck_integrals(0.0) == (0.0, 0.0)
    assert trt.planck_integrals(1e101)[0] == 1.0
    # small x: P ~ NORM x^3 / 3, R ~ NORM x^3 / 12
    p, r = trt.planck_integrals(1e-3)
    assert abs(p / (NORM * 1e-9 / 3) - 1) < 1e-3 and abs(r / (NORM * 1e-9 / 12) - 1) < 1e-3


def test_log_groups_match_geomspace():
    b, m = trt.log_groups(1e-2, 1e1, 16)
    assert np.allclose(b, np.geomspace(1e-2, 1e1, 17), rtol=1e-14)
    assert np.allclose(m, np.sqrt(b[:-1] * b[1:]), rtol=1e-15)
    ob, om = O.log_groups(1e-2, 1e1, 16)
    assert np.array_equal(b, ob) and np.array_equal(m, om)


def test_oracle_keeps_radiative_equilibrium():
    """Material and radiation in equilibrium at T0 = 1 with a Planckian inflow
    at T0 on every side stay in equilibrium through a linearized step.  The
    first group's emission covers [0, E_1] while the inflow covers
    [E_0, E_1] (src/planck.hpp EvalSpectrum versus
    src/phase_coefficient.cpp:72-76), so that group is left out; with
    E_0 = 1e-3 the energy this leaks moves T by ~5e-11."""
    om, w = ls_fixture(4, 3)
    for coef in (1e-3, 1e-1):
        o = O.TRTOracle((3, 3, 2), om, w, G=4, emin=1e-3, emax=1e2, coef=coef, T0=1.0, T_inflow=1.0,
                        inflow_sides=63, dt=1e-2, si_reltol=1e-12)
        o.step()
        assert np.abs(o.T - 1.0).max() < 1e-9
        eq = np.diff(O.planck_integral(o.bounds))
        phi = o.phi.reshape(o.G, -1)
        assert np.allclose(phi[1:], eq[1:, None], rtol=1e-9, atol=0)


def _fd_linear(f, T_star, dT, eps=1e-4):
    """f(T*) + f'(T*) dT summed, with the derivative by central differences."""
    return f(T_star).sum() + (f(T_star + eps * dT).sum() - f(T_star - eps * dT).sum()) / (2 * eps)


def _identities(o):
    """The energy identities of the oracle's last step; returns the relative
    defects of (R) the final sweep's balance, (M) the material update and (E)
    total energy, with the linearized emissions evaluated here by finite
    differences of the emission functions (independent of the step's Jacobians)."""
    e = o.energy_terms()
    dT = o.T_raw - o.T_star
    gl = _fd_linear(o.gray_emission, o.T_star, dT)
    el = _fd_linear(o.group_emission, o.T_star, dT)
    scale = e["source"] + e["emission_linear"]
    dR = abs(e["absorption"] + e["time_absorption"] + e["leakage"] - e["source"] - el) / scale
    dM = abs(e["material"] - e["floor"] - (e["absorption_si"] - gl)) / max(abs(e["absorption_si"]), abs(gl))
    # total energy: (E_mat^{n+1} - E_mat^n + E_rad^{n+1} - E_rad^n) / dt
    #   = inflow - leakage + (emission mismatch of the two rules) + (SI tolerance)
    lhs = e["material"] - e["floor"] + e["time_absorption"] - (e["source"] - e["inflow"]) + e["leakage"] - e["inflow"]
    rhs = (el - gl) + (e["absorption_si"] - e["absorption"])
    dE = abs(lhs - rhs) / scale
    return e, dR, dM, dE, el, gl


def test_oracle_linearized_step_conserves_energy():
    """One linearized backward-Euler step off equilibrium (T0 = 0.01, a T = 1
    Planckian inflow on x = 0): the final sweep's particle balance with the
    linearized emission, the material energy update and total energy all hold
    to the source-iteration tolerance, with the emissions' Jacobians checked
    by finite differences of the Planck group and gray emissions."""
    om, w = ls_fixture(4, 3)
    o = O.TRTOracle((4, 3, 3), om, w, G=4, emin=1e-2, emax=1e1, coef=10.0, cv=1.0, dt=1e-3, T0=0.01, T_inflow=1.0,
                    inflow_sides=1, si_reltol=1e-13)
    for _ in range(3):
        o.step()
        e, dR, dM, dE, el, gl = _identities(o)
        assert dR < 1e-11 and dM < 1e-7 and dE < 1e-7, (dR, dM, dE)
        assert abs(el - e["emission_linear"]) < 1e-7 * abs(el)
        # the step moves energy: the identities are not satisfied trivially
        assert abs(e["material"]) > 1e-3 * e["source"] and e["leakage"] > 0
        assert 0 <= e["balance"] < 1


def test_oracle_energy_identity_catches_wrong_heat_capacity_factor():
    """A wrong c_v/dt in the temperature update (planted in the oracle) breaks
    the material identity."""
    om, w = ls_fixture(4, 3)

    class Wrong(O.TRTOracle):
        def _pass(self, Tn, qpsi):
            self.cv *= 2.0
            try:
                super()._pass(Tn, qpsi)
            finally:
                self.cv /= 2.0

    o = Wrong((4, 3, 3), om, w, G=4, T0=0.01, T_inflow=1.0, inflow_sides=1, si_reltol=1e-13)
    o.step()
    _, dR, dM, dE, _, _ = _identities(o)
    assert dM > 1e-3 and dE > 1e-4


def test_oracle_energy_identity_catches_wrong_emission_derivative(monkeypatch):
    """The Rosseland group fractions replaced by the Planck ones (a wrong dB/dT,
    planted in the oracle) break the identities that use the finite-difference
    Jacobians."""
    om, w = ls_fixture(4, 3)
    monkeypatch.setattr(O, "rosseland_integral", O.planck_integral)
    o = O.TRTOracle((4, 3, 3), om, w, G=4, T0=0.01, T_inflow=1.0, inflow_sides=1, si_reltol=1e-13)
    o.step()
    monkeypatch.undo()
    _, dR, dM, dE, _, _ = _identities(o)
    assert max(dR, dM) > 1e-4
