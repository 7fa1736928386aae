/*
 * allium_dsa.h -- diffusion synthetic acceleration (DSA) of the source
 * iteration on the MI355X: the scalar modified-interior-penalty (MIP)
 * diffusion operator of chive's "scalar mip" preconditioner.
 *
 * ≙ DiffusionSyntheticAccelerationOperator (src/transport_op.hpp:38-59):
 *     y = x + K^{-1} Ms x
 * with Ms the scattering mass (exe/chive.cpp:593-599) and K the scalar MIP
 * diffusion matrix of InteriorPenaltyDiscretization::GetOperator
 * (src/moment_discretization.cpp:73-109) built as exe/chive.cpp:663-682 does:
 *   DiffusionIntegrator(1 / (3 sigma_t)) + MassIntegrator(sigma_t - sigma_s)
 *   + MIPDiffusionIntegrator on interior faces (src/moment_integrators.cpp:3-195):
 *       -<{D grad u.n},[v]> + sym <{D grad v.n},[u]>
 *       + <max(kappa {D/h}, alpha/2) [u],[v]>
 *   + BoundaryMassIntegrator(alpha) on inflow (Marshak) sides,
 *   alpha = ComputeAlpha(quad, x axis) (src/angular_quadrature.cpp:266-276).
 * The reference solves K with hypre BoomerAMG + CG or a direct solver; here K
 * is solved on the GPU by CG (mfem::CGSolver stopping rule) preconditioned by
 * one multigrid V-cycle: damped block-Jacobi on the DG level, the continuous
 * d-linear vertex space as first coarse level (Galerkin P^T K P) and
 * vertex-grid coarsening by two down to a dense coarsest solve.
 * One group only (chive's steady one-group SI).
 *
 * desc.method = ALLIUM_DSA_BLOCK_MIP selects chive's "mip" (the block interior-penalty
 * system whose Schur complement the reference solves), and ALLIUM_DSA_P1 chive's
 * "p1sa" preconditioner
 * (exe/chive.cpp:608-619): y = x + (P^{-1} [0; Ms x])_phi with P the 2x2 P1
 * system of allium_host_p1_matrix, which the reference solves directly.  Here
 * GMRES solves it, preconditioned by block-Jacobi around one V-cycle of the
 * MIP hierarchy above on the phi-part of the residual.
 */
#ifndef ALLIUM_DSA_H
#define ALLIUM_DSA_H

#include "allium_sweep.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct allium_dsa *allium_dsa_t;

typedef struct {
    double kappa;       /* penalty factor, InteriorPenaltyDiscretization::SetKappa semantics:
                           k < 0 means |k| (p+1)^2; chive's default is (p+1)^2 */
    int bound_below;    /* 1: penalty bounded below by alpha / 2 (SetPenaltyLowerBound) */
    double alpha;       /* Marshak coefficient; <= 0 selects ComputeAlpha over the quadrature */
    double sym;         /* sigma of MIPDiffusionIntegrator (-1: symmetric interior penalty) */
    double cg_reltol;   /* CG stops when r.z <= max(cg_reltol^2 r0.z0, cg_abstol^2) */
    double cg_abstol;
    int cg_maxit;
    int smooth;         /* damped Jacobi sweeps before and after each coarse correction */
    int method;         /* ALLIUM_DSA_MIP, or ALLIUM_DSA_P1: chive's "p1sa" (exe/chive.cpp:608-619),
                           the 2x2 P1 system of allium_host_p1_matrix solved by GMRES(40) with
                           element block-Jacobi from the right, to ||r|| <= max(cg_reltol ||b||,
                           cg_abstol) in at most cg_maxit steps; kappa, alpha, sym, smooth unused;
                           or ALLIUM_DSA_BLOCK_MIP: chive's "mip" (exe/chive.cpp:635-662), the 2x2
                           block system of allium_host_block_mip_matrix (kappa, bound_below, alpha,
                           scale_penalty as for it), solved like ALLIUM_DSA_P1 */
    int scale_penalty;  /* ALLIUM_DSA_BLOCK_MIP: penalty scaled by the mean of 1/(3 sigma_t)
                           (chive's scale_stabilization, default 1) */
} allium_dsa_desc;

#define ALLIUM_DSA_MIP 0
#define ALLIUM_DSA_P1 1
#define ALLIUM_DSA_BLOCK_MIP 2
/* chive's "ldgsa" (exe/chive.cpp:620-634): the block LDG system of
 * allium_host_block_ldg_matrix, solved like ALLIUM_DSA_P1 (alpha as for MIP) */
#define ALLIUM_DSA_LDG 3

typedef struct {
    int levels;                 /* DG level + vertex-grid levels */
    long long level_size[16];   /* unknowns per level */
    double omega[16];           /* smoother damping per level */
    int dense_coarsest;         /* 1: the coarsest level is solved with a dense inverse */
    int last_cg_iterations;     /* of the last application */
    double last_cg_norm;        /* sqrt(r.z) at the end of the last application */
    long applications;
    long cg_iterations;         /* summed over applications */
    double seconds;             /* GPU time in applications (event timing) */
} allium_dsa_info;

/* defaults: kappa -1 ((p+1)^2), bound_below 1, alpha 0 (ComputeAlpha), sym -1,
 * cg_reltol 1e-6, cg_abstol 0, cg_maxit 100, smooth 2, method ALLIUM_DSA_MIP, scale_penalty 1 */
void allium_dsa_default_desc(allium_dsa_desc *desc);

/* Build K on the mesh, quadrature and sigma_t of the sweep handle (G == 1)
 * with the element-constant scattering sigma_s[Ne]; the handle must outlive
 * the DSA object.  Setup (assembly, Galerkin coarse operators, smoother
 * damping by power iteration, coarsest inverse) runs on the host. */
int allium_dsa_create(allium_sweep_t h, const double *sigma_s, const allium_dsa_desc *desc, allium_dsa_t *out);
void allium_dsa_destroy(allium_dsa_t d);
/* ≙ DiffusionSyntheticAccelerationOperator::Mult: y = x + K^{-1} Ms x (device pointers, phi-sized) */
int allium_dsa_mult_device(allium_dsa_t d, const double *d_x, double *d_y, void *stream);
/* x = K^{-1} b by preconditioned CG from x = 0 (device pointers) */
int allium_dsa_solve_device(allium_dsa_t d, const double *d_b, double *d_x, void *stream);
/* the same with host pointers */
int allium_dsa_solve(allium_dsa_t d, const double *b, double *x);
int allium_dsa_get_info(allium_dsa_t d, allium_dsa_info *info);

/* Host only (no GPU): the MIP matrix K in CSR, rows and columns in the phi
 * layout.  Call with row_ptr == NULL to get *nnz, then with arrays of
 * Ndof + 1, *nnz and *nnz entries. */
int allium_host_mip_matrix(const allium_mesh_desc *mesh, int fe_order, const allium_quad_desc *quad,
                           const double *sigma_t, const double *sigma_s, const allium_dsa_desc *desc,
                           long long *nnz, long long *row_ptr, int *col, double *val);
/* Host only: chive's "p1sa" operator, P1Discretization::GetOperator
 * (src/moment_discretization.cpp:400-490) with alpha 0.5, in CSR over the
 * element-block layout [J_0 (n), .., J_{dim-1} (n), phi (n)] per element
 * ((dim + 1) Ndof rows).  2D p = 1, 2 and 3D p = 1. */
int allium_host_p1_matrix(const allium_mesh_desc *mesh, int fe_order, const double *sigma_t, const double *sigma_s,
                          long long *nnz, long long *row_ptr, int *col, double *val);
/* Host only: chive's "ldgsa" block operator, BlockLDGDiscretization::GetOperator
 * (src/moment_discretization.cpp:266-340) with desc->alpha (<= 0: ComputeAlpha),
 * in allium_host_p1_matrix's layout. */
int allium_host_block_ldg_matrix(const allium_mesh_desc *mesh, int fe_order, const allium_quad_desc *quad,
                                 const double *sigma_t, const double *sigma_s, const allium_dsa_desc *desc,
                                 long long *nnz, long long *row_ptr, int *col, double *val);
/* Host only: chive's "mip" block operator, BlockIPDiscretization::GetOperator
 * (src/moment_discretization.cpp:342-398) with the penalty options of desc
 * (kappa, bound_below, alpha, scale_penalty), in allium_host_p1_matrix's
 * layout.  Its Schur complement on phi is what the reference solves. */
int allium_host_block_mip_matrix(const allium_mesh_desc *mesh, int fe_order, const allium_quad_desc *quad,
                                 const double *sigma_t, const double *sigma_s, const allium_dsa_desc *desc,
                                 long long *nnz, long long *row_ptr, int *col, double *val);

#ifdef __cplusplus
}
#endif

#endif
