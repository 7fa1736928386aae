/*
 * allium_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of allium's DG upwind discrete-ordinates sweep
 * (InverseAdvectionOperator, src/sweep.cpp:10-244, 277-669, 924-1022,
 * 1135-1244), its streaming operator and particle balance
 * (AdvectionOperator, src/sweep.cpp:1246-1462), the discrete-to-moment maps
 * (src/transport_op.cpp:3-60) and the source iteration (SLISolver,
 * src/linalg.cpp:103-153, driven as in exe/chive.cpp:700-735).
 *
 * It is the checker for the HIP product path and the CPU baseline timed by
 * bench.py.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library never links or calls it.
 *
 * Meshes are structured (1D segments, 2D quads, 3D hexes) with arbitrary
 * vertex positions, element order x-fastest (mfem::Mesh::MakeCartesian*D with
 * sfc_ordering=false).  Basis: Gauss-Lobatto Lagrange (L2_FECollection with
 * BasisType::GaussLobatto), dofs lexicographic per element.  Vector layout
 * RIGHT: idx(g,a,s) = s + Ndof*(a + Nomega*g) (src/tvector.hpp:47-57).
 */
#ifndef ALLIUM_ORACLE_H
#define ALLIUM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_problem orc_problem;

/* bc[2*dim]: per boundary side (axis*2 + {0: low, 1: high}); 0 = inflow
 * (vacuum or prescribed), 1 = reflective.  lump: bit flags (1 mass, 2
 * gradient, 4 faces), src/lumping.hpp:10-18. */
orc_problem *orc_create(int dim, const int *nel, const double *vertices, int p,
                        int nomega, const double *omega, const double *weights,
                        int G, const double *sigma_t, const int *bc, int lump,
                        int *status);
void orc_destroy(orc_problem *pr);

/* out[0..7] = Ne, n (dofs/elem), Ndof, nodes, edges, levels, faces, psi size */
void orc_sizes(const orc_problem *pr, long *out);

/* frontier index of every node e*Nomega+a (Kahn frontiers, src/sweep.cpp:308-512) */
int orc_levels(const orc_problem *pr, int *level_of_node);

/* AssembleLocalMatrices with new sigma_t[G][Ne] (src/sweep.cpp:924-1001) */
int orc_set_sigma(orc_problem *pr, const double *sigma_t);
/* SetTimeAbsorption (src/sweep.cpp:1003-1022) */
int orc_set_time_absorption(orc_problem *pr, double tau);

/* Mult_Upwind: psi = L^{-1} source; source and psi may alias. Returns the
 * number of local solves whose pivot fell below MFEM's LinearSolve TOL. */
long orc_sweep(orc_problem *pr, const double *source, double *psi, int nthreads);

/* negative-flux fixup after each local solve (SetFixupOperator / UseFixup,
 * src/sweep.hpp:116-122; src/fixup.cpp): type 0 off, 1 zero, 2 zero and
 * scale, 3 ryosuke; monitor (G x Ne, may be NULL) counts the applications per
 * element (SetFixupMonitorData).  Resets the counts. */
int orc_set_fixup(orc_problem *pr, int type, double psi_min, double *monitor);
/* out[0] applied, out[1] failed (EventLog "fixup applied" / "fixup failed",
 * src/sweep.cpp:649-650), out[2] 1 if ryosuke met more than 3 positive
 * unknowns (where the reference aborts) */
int orc_fixup_counts(const orc_problem *pr, long *out);

/* AdvectionOperator::Mult and ComputeBalance */
int orc_apply_L(const orc_problem *pr, const double *psi, double *Lpsi);
double orc_balance(const orc_problem *pr, const double *psi);

/* FormTransportSource with per-steradian source Q[G][Ne] (element-constant)
 * and isotropic inflow[G][2*dim] per boundary side. */
int orc_form_source(const orc_problem *pr, const double *Q, const double *inflow,
                    double *source);

/* DiscreteToMoment (0th moment) and its transpose */
int orc_D(const orc_problem *pr, const double *psi, double *phi);
int orc_DT(const orc_problem *pr, const double *phi, double *psi);

/* block-diagonal MassIntegrator(coef) applied to each group block of x
 * (the scattering mass Ms of exe/chive.cpp:593-599) */
int orc_mass(const orc_problem *pr, const double *coef, const double *x, double *y);

/* chive source iteration: phi = D L^{-1}(Ms phi / 4pi + q), stopping when
 * ||phi_{k+1}-phi_k||_2 < max(reltol*||r_1||, abstol) or at maxit; psi gets
 * the final sweep.  Returns iterations; norms[0..it-1] holds the history. */
int orc_source_iteration(orc_problem *pr, const double *sigma_s, const double *source,
                         double abstol, double reltol, int maxit, double *phi,
                         double *psi, double *norms, int nthreads);

/* element and face matrices (column-major n x n) for known-answer tests */
int orc_element_matrices(const orc_problem *pr, int e, double *mass_g0, double *grad,
                         double *tmass);
int orc_face_blocks(const orc_problem *pr, int e1, int e2, double *blocks);
int orc_face_normal(const orc_problem *pr, int e1, int e2, double *nor);

/* exposed building blocks: GLL nodes and Gauss-Legendre rule on [0,1] */
int orc_gll_nodes(int p, double *x);
int orc_gauss_rule(int npts, double *x, double *w);

#ifdef __cplusplus
}
#endif
#endif
