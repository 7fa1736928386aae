/*
 * allium_trt.h -- multigroup linearized backward-Euler thermal radiative
 * transfer on the MI355X sweep (scallion's "linearized" solver).
 *
 * One time step is LinearizedTRTOperator::Mult (src/trt_linearized.cpp:230-312)
 * driven by scallion's time loop (exe/scallion.cpp:1354-1555):
 *   - emission B(T) = sum_g int sigma_g b_g(T) T^4 phi_i (PlanckEmissionNFI,
 *     src/trt_integrators.cpp:56-118) and the energy balance
 *     Bt(T) = (c_v/dt) M T + sum_g int sigma_g b_g(T) T^4 phi_i
 *     (GrayPlanckEmissionNFI + MassIntegrator, exe/scallion.cpp:752-760),
 *     linearized at the current T with the Rosseland spectrum;
 *   - the transport system (I - D Linv D^T Ms) phi = D Linv(...) with the
 *     pseudo-scattering Ms = dB (dBt)^{-1} M_sigma, solved by source iteration
 *     (SLISolver, src/linalg.cpp:103-153);
 *   - a final sweep for psi, T += (dBt)^{-1}(M_sigma phi + resid), floored at 1e-8;
 *   - explicit opacities sigma_g = coef / (e_g^3 sqrt(T)) at the element centre
 *     and group midpoint e_g (recomputed after every step, exe/scallion.cpp:1462).
 * Units are dimensionless (radiation constant times light speed = 1, c = 1):
 * the time absorption is 1/dt and black-body emission is T^4.
 * Groups are log-spaced (MultiGroupEnergyGrid::MakeLogSpaced,
 * src/multigroup.cpp:40-59).  Uniform 3D meshes, p = 1.
 */
#ifndef ALLIUM_TRT_H
#define ALLIUM_TRT_H

#include "allium_sweep.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct allium_trt *allium_trt_t;

typedef struct {
    int groups;           /* G */
    double emin, emax;    /* photon energy range of the log-spaced groups */
    double opacity_coef;  /* sigma_g = opacity_coef / (e_g^3 sqrt(T)) */
    double cv;            /* heat capacity */
    double dt;            /* time step */
    double T0;            /* initial temperature; radiation starts in equilibrium with it */
    double T_inflow;      /* Planckian inflow temperature on the inflow sides */
    int inflow_sides;     /* bit s set: side s (axis*2 + hi) carries the inflow; others vacuum */
    double si_reltol;     /* source iteration (SLISolver) stops when ||phi_{k+1} - phi_k|| <  */
    double si_abstol;     /*   max(si_reltol * first norm, si_abstol) or after si_maxit */
    int si_maxit;
    int si_warm_start;    /* 0: start from phi = 0 (SLISolver without iterative_mode);
                             1: from D psi of the current iterate (psi^n on the first pass) */
    /* scallion's outer FixedPointIterationSolver on the temperature
     * (src/fixed_point.cpp:134-172, exe/scallion.cpp:830-861): each pass
     * relinearizes at the current T with the previous time level's sources, until
     * ||T_k - T_{k-1}||_2 < max(outer_reltol * first change, outer_abstol) or
     * outer_maxit passes.  0 or 1: one pass, the linearized solver alone. */
    int outer_maxit;
    double outer_abstol, outer_reltol;
    int implicit_opacity; /* 1: opacities re-evaluated at T after every pass (OpacityUpdate,
                             src/opacity_update.hpp, src/trt_linearized.cpp:311); needs
                             outer_maxit > 1.  0: once per time step (explicit) */
} allium_trt_desc;

typedef struct {
    int step;             /* steps taken */
    double time;
    int transport_iterations;   /* source iterations of the last step */
    double transport_residual;  /* final ||phi_{k+1} - phi_k|| of the last step */
    long floored;         /* temperature values floored at 1e-8 in the last step */
    double T_min, T_max;  /* over the temperature dofs after the last step */
    double radiation_energy;    /* sum_g int phi_g (dimensionless E = phi / c) */
    double sweep_seconds; /* GPU time in sweeps so far */
    long sweeps;
    int outer_iterations;       /* fixed-point passes of the last step */
    double outer_residual;      /* last ||T_k - T_{k-1}||_2 (0 with one pass) */
} allium_trt_info;

/* Energy bookkeeping of the last step, as rates (energy per unit time), for
 * the operator of its last pass (that pass's opacities, linearized at its
 * starting temperature T*).  Masses enter through their row sums:
 *   absorption       sum_g int sigma_g phi_g, phi = D psi^{n+1} of the final sweep
 *   time_absorption  (1/dt) int sum_g phi_g^{n+1}
 *   leakage          outflow of psi^{n+1}: AdvectionOperator::ComputeBalance
 *                    (src/sweep.cpp:1400-1462) minus the two absorptions
 *   source           sum D q_psi = inflow + (1/dt) int sum_g (D psi^n)_g
 *   inflow           sum D q_in
 *   emission_linear  sum of the emission source of the final sweep (em + Ms phi)
 *   absorption_si    sum_g int sigma_g phi_g of the source iteration's phi
 *   gray_linear      sum of the gray emission at T* plus its Jacobian times (T_raw - T*),
 *                    T_raw = T* + Bt'^{-1}(M_sigma phi + resid) the update before the floor
 *   material         int c_v (T^{n+1} - T^n) / dt
 *   floor            int c_v (T^{n+1} - T_raw) / dt: the energy the 1e-8 floor added
 *   emission         sum of PlanckEmissionNFI at T^{n+1} (scallion's emission form)
 *   balance          scallion's ComputeBalance (exe/scallion.cpp:157-173):
 *                    |absorption + time_absorption + leakage - (emission + source)|
 *                    / (emission + source)
 * A linearized backward-Euler step satisfies, to rounding:
 *   absorption + time_absorption + leakage = source + emission_linear  (final sweep)
 *   material - floor = absorption_si - gray_linear                   (temperature update)
 * and their sum is total energy conservation up to the floor, the two emission
 * rules' difference (emission_linear - gray_linear) and the source-iteration
 * tolerance (absorption_si - absorption). */
typedef struct {
    double absorption, time_absorption, leakage, source, inflow;
    double emission_linear, absorption_si, gray_linear, material, floor, emission, balance;
} allium_trt_energy;

int allium_trt_create(const allium_mesh_desc *mesh, int fe_order, const allium_quad_desc *quad,
                      const allium_trt_desc *desc, allium_trt_t *out);
void allium_trt_destroy(allium_trt_t h);
/* advance one linearized backward-Euler step */
int allium_trt_step(allium_trt_t h, allium_trt_info *info);
/* temperature dofs (Ne * n) and scalar flux (G * Ne * n, RIGHT layout) */
int allium_trt_temperature(allium_trt_t h, double *T);
int allium_trt_scalar_flux(allium_trt_t h, double *phi);
/* The state scallion restarts from (exe/scallion.cpp:604-631, src/restart.cpp):
 * the block vector x = [psi (psi_size), T (Ne * n)] with the step count and
 * time.  set_state recomputes the scalar flux and the opacities from it, so a
 * run continued from a saved state repeats the uninterrupted run bit for bit
 * (the time step stays the solver's; allium_trt_set_time_step changes it). */
int allium_trt_state_size(allium_trt_t h, long *size);
int allium_trt_get_state(allium_trt_t h, double *x, int *cycle, double *time);
int allium_trt_set_state(allium_trt_t h, const double *x, int cycle, double time);
/* Energy bookkeeping on (1) or off (0, the default): when on, each step also
 * computes the terms of allium_trt_energy (one extra pass over psi^n and two
 * over psi^{n+1}). */
int allium_trt_track_energy(allium_trt_t h, int on);
/* the terms of the last step; ALLIUM_ERR_ARG while the bookkeeping is off */
int allium_trt_energy_terms(allium_trt_t h, allium_trt_energy *out);
/* Change the time step from the next step on (scallion adopts a restart
 * file's time step, exe/scallion.cpp:616-674): the time absorption 1/dt, the
 * time source and c_v/dt follow it. */
int allium_trt_set_time_step(allium_trt_t h, double dt);
/* the sweep handle the solver uses (tuning, timing, fixups) */
allium_sweep_t allium_trt_sweep(allium_trt_t h);

/* host helpers: normalized Planck and Rosseland integrals over [0, x]
 * (src/planck.hpp:158-180) and the log-spaced group bounds (G + 1) and
 * geometric midpoints (G) */
int allium_planck_integrals(double x, double *planck, double *rosseland);
int allium_log_groups(double emin, double emax, int G, double *bounds, double *midpoints);

#ifdef __cplusplus
}
#endif

#endif
