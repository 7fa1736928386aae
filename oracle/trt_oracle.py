"""CPU oracle of the multigroup linearized backward-Euler TRT step (test infrastructure only).

Restates, in FP64 numpy on top of the C sweep oracle:
  * the normalized Planck/Rosseland integrals of src/planck.hpp:1-253 (the
    Taylor series with Bernoulli-number coefficients through x^21 and the
    9-term polylogarithm form of B. A. Clark, JCP 1987; the smaller is kept);
  * MultiGroupEnergyGrid::MakeLogSpaced (src/multigroup.cpp:40-59);
  * PlanckEmissionNFI / GrayPlanckEmissionNFI element integrals
    (src/trt_integrators.cpp:56-205) on the tensor Gauss rules of order 2p+1
    and 2p+2;
  * LinearizedTRTOperator::Mult (src/trt_linearized.cpp:230-312) with
    SLISolver (src/linalg.cpp:103-153) for the transport system and the
    scallion time loop (exe/scallion.cpp:1354-1555): time source
    inflow + (1/(c dt)) Mt psi^n, explicit opacity recomputed after each step.
Dimensionless units: c = 1 and the black-body emission is T^4.
Pinned in tests/test_trt.py by numerical quadrature of the Planck and
Rosseland integrands, by the radiative-equilibrium invariant, by central
finite differences of the group and gray emissions (the linearization's
Jacobians) and by the energy identities of one linearized backward-Euler step
(energy_terms): the final sweep's DG balance, the material energy update, and
their sum, total energy conservation up to the emission-rule mismatch and the
source-iteration tolerance.
"""
import numpy as np

from . import Problem, gauss_rule

NORM = 15.0 / np.pi ** 4


def planck_taylor(x):
    from math import factorial
    B = {0: 1.0, 1: -0.5, 2: 1.0 / 6, 4: -1.0 / 30, 6: 1.0 / 42, 8: -1.0 / 30, 10: 5.0 / 66, 12: -691.0 / 2730,
         14: 7.0 / 6, 16: -3617.0 / 510, 18: 43867.0 / 798}
    x = np.asarray(x, dtype=float)
    return NORM * sum(b / (factorial(n) * (n + 3)) * x ** (n + 3) for n, b in B.items())


def planck_polylog(x):
    x = np.asarray(x, dtype=float)
    e = np.exp(-x)
    li = [sum(e ** l / l ** a for l in range(1, 10)) for a in (1, 2, 3, 4)]
    return 1.0 - NORM * (6 * li[3] + x * (6 * li[2] + x * (3 * li[1] + x * li[0])))


def planck_integral(x):
    x = np.asarray(x, dtype=float)
    with np.errstate(over="ignore", invalid="ignore"):
        v = np.minimum(planck_taylor(x), planck_polylog(x))
    return np.where(x > 1e100, 1.0, np.where(x == 0.0, 0.0, v))


def rosseland_integral(x):
    x = np.asarray(x, dtype=float)
    p = planck_integral(x)
    with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
        r = p - 0.25 * NORM * x ** 4 / np.expm1(x)
    return np.where(x == 0.0, 0.0, np.where(x > 1.1579208923731620e77, p, r))


def log_groups(emin, emax, G):
    lo, d = np.log(emin), (np.log(emax) - np.log(emin)) / G
    b = np.array([emin] + [np.exp(lo + g * d) for g in range(1, G + 1)])
    return b, np.sqrt(b[:-1] * b[1:])


def spectrum(bounds, T, rosseland=False):
    """Group fractions at temperatures T (any shape) -> (..., G); the first
    group starts at 0 and the last extends to infinity (src/planck.hpp EvalSpectrum)."""
    F = rosseland_integral if rosseland else planck_integral
    T = np.asarray(T, dtype=float)
    G = len(bounds) - 1
    vals = [F(bounds[g + 1] / T) for g in range(G - 1)]
    out = np.empty(T.shape + (G,))
    prev = np.zeros_like(T)
    for g in range(G - 1):
        out[..., g] = vals[g] - prev
        prev = vals[g]
    out[..., G - 1] = 1.0 - prev
    return out


def hex_rule(order, h):
    """Trilinear basis values (nq, 8) and weights x volume of the tensor Gauss rule."""
    n = order // 2 + 1
    x, w = gauss_rule(n)
    pts, wts = [], []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                pts.append((x[i], x[j], x[k]))
                wts.append(w[i] * w[j] * w[k] * h[0] * h[1] * h[2])
    phi = np.array([[(px if i & 1 else 1 - px) * (py if i & 2 else 1 - py) * (pz if i & 4 else 1 - pz)
                     for i in range(8)] for (px, py, pz) in pts])
    return phi, np.array(wts)


class TRTOracle:
    """Linearized backward-Euler TRT on a uniform unit-box hex mesh, p = 1."""

    def __init__(self, nel, omega, weights, G=4, emin=1e-2, emax=1e1, coef=10.0, cv=1.0, dt=1e-3, T0=0.01,
                 T_inflow=1.0, inflow_sides=1, si_reltol=1e-10, si_abstol=0.0, si_maxit=500, warm_start=False,
                 outer_maxit=1, outer_abstol=0.0, outer_reltol=0.0, implicit_opacity=False, fixup=None):
        self.nel = tuple(nel)
        # (kind, psi_min): the sweep fixup, applied in the final sweep of each pass only
        self.fixup = fixup
        # FixedPointIterationSolver on T around the linearized step (src/fixed_point.cpp:134-172)
        self.outer = (max(1, int(outer_maxit)), outer_abstol, outer_reltol, bool(implicit_opacity))
        self.outer_iterations = []
        self.Ne = int(np.prod(nel))
        self.G, self.cv, self.dt, self.coef = G, cv, dt, coef
        self.si = (si_reltol, si_abstol, si_maxit, warm_start)
        self.bounds, self.mid = log_groups(emin, emax, G)
        h = [1.0 / n for n in nel]
        v = np.stack(np.meshgrid(*[np.linspace(0, 1, n + 1) for n in nel], indexing="ij"), -1)
        verts = v.transpose(2, 1, 0, 3).reshape(-1, 3)
        self.T = np.full((self.Ne, 8), T0)
        self.pr = Problem(3, nel, verts, 1, omega, weights, self.opacity(), bc=None)
        self.pr.set_time_absorption(1.0 / dt)
        self.wsum = float(np.sum(weights))
        _, _, Mt = self.pr.element_matrices(0)     # MassIntegrator-rule mass (time, heat capacity)
        self.Mm = Mt
        self.phi3, self.w3 = hex_rule(3, h)
        self.phi4, self.w4 = hex_rule(4, h)
        # unit MG mass (rule order 2p: 2 Gauss points per axis, as the order-3 rule)
        self.Mu = np.einsum("q,qi,qj->ij", self.w3, self.phi3, self.phi3)
        # initial radiation in equilibrium with T0 (group-bound Planck fractions)
        frac = planck_integral(self.bounds[1:] / T0) - planck_integral(self.bounds[:-1] / T0)
        No = len(weights)
        self.psi = np.repeat(frac * T0 ** 4 / (4 * np.pi), No * self.Ne * 8)
        fi = (planck_integral(self.bounds[1:] / T_inflow) - planck_integral(self.bounds[:-1] / T_inflow)) \
            * T_inflow ** 4 / (4 * np.pi) if T_inflow > 0 else np.zeros(G)
        inflow = np.zeros((G, 6))
        for s in range(6):
            if (inflow_sides >> s) & 1:
                inflow[:, s] = fi
        self.qin = self.pr.form_source(inflow=inflow)
        self.sigma = self.opacity()
        self.pr.set_sigma(self.sigma)
        self.iterations = []

    def opacity(self):
        tc = self.T.mean(axis=1)
        return self.coef / np.sqrt(tc)[None, :] / self.mid[:, None] ** 3

    def _apply_dB(self, u, dpw):
        """(G, Ne, 8): sum_q dpw[e, q, g] phi_q (phi_q . u_e)."""
        uq = u @ self.phi3.T                                      # (Ne, nq)
        return np.einsum("eqg,eq,qi->gei", dpw, uq, self.phi3)

    def _Ms(self, phi, Binv, dpw):
        t = np.einsum("g,gei->ei", np.ones(self.G), (self.sigma[:, :, None] * (phi.reshape(self.G, self.Ne, 8) @ self.Mu.T)))
        u = np.einsum("eij,ej->ei", Binv, t)
        return self._apply_dB(u, dpw)

    def step(self):
        """One time step: passes of LinearizedTRTOperator::Mult (src/trt_linearized.cpp:230-312)
        until the temperature change meets the outer tolerance (one pass by
        default); opacities after every pass (implicit) or once (explicit)."""
        pr = self.pr
        Tn, psin = self.T.copy(), self.psi.copy()
        qpsi = self.qin + (1.0 / self.dt) * (psin.reshape(-1, 8) @ self.Mm.T).ravel()
        self.Tn, self.psin, self.qpsi = Tn, psin, qpsi
        maxit, abstol, reltol, implicit = self.outer
        r0 = 0.0
        for k in range(1, maxit + 1):
            Tk = self.T.copy()
            self._pass(Tn, qpsi)
            if implicit:
                self.sigma = self.opacity()
                pr.set_sigma(self.sigma)
            if maxit == 1:
                break
            nrm = float(np.linalg.norm(self.T - Tk))
            if k == 1:
                r0 = max(abstol, nrm * reltol)
            if nrm < r0:
                break
        self.outer_iterations.append(k)
        if not implicit:
            self.sigma = self.opacity()
            pr.set_sigma(self.sigma)

    def _pass(self, Tn, qpsi):
        """LinearizedTRTOperator::Mult at the current (psi, T) with the previous
        time level's sources (Tn, qpsi)."""
        pr, G, Ne = self.pr, self.G, self.Ne
        cvdt = self.cv / self.dt
        T = self.T
        self.T_star, self.sigma_pass = T.copy(), self.sigma.copy()
        # the fixup is off for the transport solve (src/trt_linearized.cpp:236-238)
        pr.set_fixup("none")
        # linearization (rule of order 4 for the gray balance, 3 for the groups)
        Tq4 = T @ self.phi4.T
        sp4, sr4 = spectrum(self.bounds, Tq4), spectrum(self.bounds, Tq4, True)
        sig = self.sigma.T                                        # (Ne, G)
        gp = np.einsum("eqg,eg->eq", sp4, sig)
        gr = np.einsum("eqg,eg->eq", sr4, sig)
        A = cvdt * self.Mm[None] + np.einsum("q,eq,qi,qj->eij", self.w4, 4 * gr * Tq4 ** 3, self.phi4, self.phi4)
        resid = cvdt * (Tn - T) @ self.Mm.T - np.einsum("q,eq,qi->ei", self.w4, gp * Tq4 ** 4, self.phi4)
        Binv = np.linalg.inv(A)
        Tq3 = T @ self.phi3.T
        sp3, sr3 = spectrum(self.bounds, Tq3), spectrum(self.bounds, Tq3, True)
        em = np.einsum("q,eqg,eg,eq,qi->gei", self.w3, sp3, sig, Tq3 ** 4, self.phi3)
        dpw = np.einsum("q,eg,eqg,eq->eqg", self.w3, sig, sr3, 4 * Tq3 ** 3)
        t1 = np.einsum("eij,ej->ei", Binv, resid)
        em = em + self._apply_dB(t1, dpw)
        psi = pr.sweep(qpsi + pr.DT(em.ravel()))
        phisrc = pr.D(psi)
        reltol, abstol, maxit, warm = self.si
        phi = pr.D(self.psi) if warm else np.zeros_like(phisrc)
        r0 = None
        for it in range(1, maxit + 1):
            y = pr.D(pr.sweep(pr.DT(self._Ms(phi, Binv, dpw).ravel())))
            r = phisrc - phi + y
            phi = phi + r
            nrm = np.sqrt(np.dot(r, r))
            if it == 1:
                r0 = max(nrm * reltol, abstol)
            if nrm < r0:
                break
        self.iterations.append(it)
        em = em + self._Ms(phi, Binv, dpw)
        self.em_final = em
        if self.fixup is not None:
            pr.set_fixup(*self.fixup)
        self.psi = pr.sweep(qpsi + pr.DT(em.ravel()))
        t = resid + np.einsum("ge,gei->ei", self.sigma, phi.reshape(G, Ne, 8) @ self.Mu.T)
        self.T_raw = T + np.einsum("eij,ej->ei", Binv, t)
        self.T = np.maximum(self.T_raw, 1e-8)
        self.phi = phi

    # ---- emissions at a given temperature (the linearization's functions) ----
    def group_emission(self, T, sigma=None):
        """PlanckEmissionNFI on rule 3: (G, Ne, 8) int sigma_g b_g(T) T^4 phi_i."""
        sig = (self.sigma_pass if sigma is None else sigma).T
        Tq = np.asarray(T).reshape(self.Ne, 8) @ self.phi3.T
        sp = spectrum(self.bounds, Tq)
        return np.einsum("q,eqg,eg,eq,qi->gei", self.w3, sp, sig, Tq ** 4, self.phi3)

    def gray_emission(self, T, sigma=None):
        """GrayPlanckEmissionNFI on rule 4: (Ne, 8) int sum_g sigma_g b_g(T) T^4 phi_i."""
        sig = (self.sigma_pass if sigma is None else sigma).T
        Tq = np.asarray(T).reshape(self.Ne, 8) @ self.phi4.T
        sp = spectrum(self.bounds, Tq)
        return np.einsum("q,eqg,eg,eq,qi->ei", self.w4, sp, sig, Tq ** 4, self.phi4)

    def energy_terms(self):
        """Energy bookkeeping of the last step (all rates: energy per unit time),
        for the operator of its last pass (opacities sigma_pass, linearized at
        T_star).  With rowsums of the element masses:
          absorption       sum_g int sigma_g phi_g        phi = D psi^{n+1}
          time_absorption  (1/dt) int sum_g phi_g^{n+1}
          leakage          outflow of psi^{n+1} (AdvectionOperator::ComputeBalance
                           minus its absorption terms)
          source           sum D q_psi = inflow + (1/dt) int sum_g phi^n_g
          inflow           sum D q_in
          emission_linear  sum of the emission source of the final sweep (em + Ms phi)
          absorption_si    sum_g int sigma_g phi_g of the source iteration's phi
          material         int c_v (T^{n+1} - T^n) / dt
          floor            int c_v (T^{n+1} - T_raw) / dt, what the 1e-8 floor added to
                           the update T_raw = T* + Bt'^{-1}(M_sigma phi + resid)
          emission         sum of PlanckEmissionNFI at T^{n+1} (scallion's emission form)
          balance          scallion's ComputeBalance (exe/scallion.cpp:157-173):
                           |absorption + time_absorption + leakage - (emission + source)|
                           / (emission + source)
        """
        pr, G, Ne = self.pr, self.G, self.Ne
        tau = 1.0 / self.dt
        rMm, rMu = self.Mm.sum(axis=1), self.Mu.sum(axis=1)
        phif = pr.D(self.psi).reshape(G, Ne, 8)
        phin = pr.D(self.psin).reshape(G, Ne, 8)
        sig = self.sigma_pass
        absorption = float(np.einsum("ge,gei,i->", sig, phif, rMu))
        t_abs = tau * float(np.einsum("gei,i->", phif, rMm))
        # ComputeBalance with the step's operator: absorption (sigma + tau) + leakage
        pr.set_sigma(sig)
        lb = pr.balance(self.psi)
        pr.set_sigma(self.sigma)
        leakage = lb - absorption - t_abs
        source = float(pr.D(self.qpsi).sum())
        inflow = float(pr.D(self.qin).sum())
        phis = self.phi.reshape(G, Ne, 8)
        out = dict(absorption=absorption, time_absorption=t_abs, leakage=leakage, source=source, inflow=inflow,
                   emission_linear=float(self.em_final.sum()),
                   absorption_si=float(np.einsum("ge,gei,i->", sig, phis, rMu)),
                   material=self.cv / self.dt * float(np.einsum("ei,i->", self.T - self.Tn, rMm)),
                   floor=self.cv / self.dt * float(np.einsum("ei,i->", self.T - self.T_raw, rMm)),
                   emission=float(self.group_emission(self.T).sum()))
        out["balance"] = abs(lb - (out["emission"] + source)) / (out["emission"] + source)
        return out
