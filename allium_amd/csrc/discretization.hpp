// Host-side discretization for the MI355X sweep: level-symmetric quadrature,
// structured quad/hex meshes, Gauss-Lobatto DG element/face matrices and the
// per-direction topological levels of the sweep graph.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace allium {

// Level-symmetric S_N set (LevelSymmetricQuadrature, src/angular_quadrature.cpp:98-147).
// omega is stored padded to 3 components per direction.
struct Quadrature {
    int dim = 0;
    std::vector<double> omega;  // size() x 3
    std::vector<double> w;
    double wsum = 0.0;
    int size() const { return (int)w.size(); }
};
bool level_symmetric(int order, int dim, Quadrature &q, std::string &err);

// Structured mesh: nel[d] elements per axis, elements and vertices x-fastest.
struct Mesh {
    int dim = 0;
    int nel[3] = {1, 1, 1};
    int nv[3] = {1, 1, 1};
    int Ne = 0;
    std::vector<double> vert;  // Nv x dim
    int bc[6] = {0, 0, 0, 0, 0, 0};
    bool uniform = false;      // every element the same axis-aligned box
    double h[3] = {0, 0, 0};
    // faces, grouped by axis; e1 owns the normal (lower element id on
    // interior faces), e2 = -1 on the boundary
    int Nf = 0;
    int foff[4] = {0, 0, 0, 0};
    int fext[3][3];
    std::vector<int> f_e1, f_e2, f_bside;
    std::vector<double> f_nor;  // Nf x 3, unit normal at the face centre, outward of e1

    void elem_coords(int e, int *c) const {
        c[0] = e % nel[0];
        c[1] = (e / nel[0]) % nel[1];
        c[2] = e / (nel[0] * nel[1]);
    }
    int face_of(int e, int side) const;  // side = 2*axis + {0 low, 1 high}
};
bool build_mesh(int dim, const int *nel, const double *vertices, const double *extents,
                const int *bc, Mesh &m, std::string &err);

// DG Gauss-Lobatto element data. Matrices are row-major n x n.
struct ElementData {
    int dim = 0, p = 0, n = 0, nf = 0, lump = 0;
    bool uniform = false;
    std::vector<double> Mu;     // unit mass (MGMassIntegrator rule, order 2p), per element
    std::vector<double> Mm;     // MassIntegrator rule (time mass, scattering mass), per element
    std::vector<double> Gd;     // per element: dim blocks of -int phi_j d_d phi_i
    std::vector<double> Fface;  // per face (uniform: per side): nf x nf face mass
    std::vector<int> side_nodes;  // 2*dim x nf local node ids of each side, face-lexicographic
};
bool assemble(const Mesh &m, int p, int lump, ElementData &ed, std::string &err);
// Gauss points of mfem::IntRules.Get(cube/square, order) on one uniform element:
// phi[q * n + i] = basis function i at point q, w[q] = weight times element volume
void uniform_element_rule(const Mesh &m, int p, int order, std::vector<double> &phi, std::vector<double> &w);

// Scalar interior-penalty (MIP) diffusion matrix of diffusion synthetic
// acceleration in element-block form (InteriorPenaltyDiscretization::GetOperator,
// src/moment_discretization.cpp:73-109, with MIPDiffusionIntegrator,
// src/moment_integrators.cpp:3-195):
//   K = DiffusionIntegrator(1/(3 sigma_t)) + MassIntegrator(sigma_a)
//     + interior faces: -<{D grad u.n},[v]> + sym <{D grad v.n},[u]> + <pen [u],[v]>,
//       pen = max(kappa {D/h}, lower) per face point
//     + inflow (Marshak) boundary faces: alpha <u, v>.
struct MipParams {
    double kappa = 4.0;   // penalty factor; 0 turns the penalty off (lower bound too)
    double lower = 0.0;   // penalty lower bound (chive: alpha / 2)
    double sym = -1.0;    // sigma of the integrator: -1 symmetric interior penalty
    double alpha = 0.5;   // Marshak boundary coefficient
    int lump = 0;         // bit 0 mass, bit 1 diffusion, bit 2 face terms on the corner rule
};
struct MipMatrix {
    int n = 0, nside = 0, nc = 0;
    std::vector<double> Kd;   // Ne x n x n, row-major
    std::vector<double> Ko;   // Ne x nside x n x n: rows of e, columns of nb[e][s]
    std::vector<int> nb;      // Ne x nside, -1 on the boundary
    std::vector<double> P;    // n x nc (nc = 2^dim): DG node values of the d-linear vertex functions
};
bool assemble_mip(const Mesh &m, int p, const double *sigma_t, const double *sigma_a, const MipParams &prm,
                  MipMatrix &K, std::string &err);
// chive's "mip" block operator (BlockIPDiscretization::GetOperator, src/moment_discretization.cpp:342-398):
// interior penalty max(kappa {D} area / det_1, limit) with {D} the mean of 1/(3 sigma_t) (scale_D)
// or 1, the vector mass alone in Mt, BoundaryMass(alpha) and no boundary term in D
struct BlockIpParams {
    double kappa;
    double limit;
    bool scale_D;
    // chive's "ldgsa" (BlockLDGDiscretization, :266-340) instead: penalty alpha/2
    // (kappa, limit, scale_D unused) and the one-sided LDG trace [psi] J_1.n in D
    bool ldg = false;
};
// chive's "p1sa" operator (P1Discretization::GetOperator, src/moment_discretization.cpp:400-490)
// in the same element-block storage: per element the unknowns [J_0 (n), .., J_{dim-1} (n), phi (n)],
// block size K.n = (dim + 1) n; K.P stays empty.  With ip, the "mip" block operator instead.
bool assemble_p1(const Mesh &m, int p, const double *sigma_t, const double *sigma_a, double alpha, MipMatrix &K,
                 std::string &err, const BlockIpParams *ip = nullptr);

// Sweep plan: level of every node e*Nomega+a and the cells grouped by level.
struct SweepPlan {
    int nlevels = 0;
    long nedges = 0;
    std::vector<int> level;        // Ne*Nomega
    std::vector<uint32_t> cells;   // packed (e << 10) | a, grouped by level, then angle, then e
    std::vector<long> level_ptr;   // nlevels+1
    std::vector<int> refl;         // Nomega x 2dim reflected angle or -1
    bool angle_independent = true; // no reflective coupling between directions
    // wave plan: per level, part and XCD bin, runs of 64 cells sharing one
    // direction (tails padded with kPad).  Parts are groups of octants that
    // never couple (4 when directions are independent, else 1); bins are
    // spatial 8-colour checkerboard blocks.
    static constexpr uint32_t kPad = 0xFFFFFFFFu;
    int nparts = 1;
    std::vector<uint32_t> wcells;
    std::vector<long> wlevel_base;   // nlevels+1, offset of each level in wcells
    std::vector<int> wpb;            // nlevels x (nparts*8+1): prefix of waves per (part, bin)
};
// fill the wave plan from level[] (bin = checkerboard colour of blocks of `block` elements)
void build_wave_plan(const Mesh &m, const Quadrature &q, SweepPlan &plan, int block, int nparts);
// Returns 0, or -2 (not a DAG) / -3 (missing reflected direction).
int build_plan(const Mesh &m, const Quadrature &q, SweepPlan &plan, std::string &err);

// Source vector (FormTransportSource) on the host, RIGHT layout.
void form_source(const Mesh &m, const ElementData &ed, const Quadrature &q, int G, const double *Q,
                 const double *inflow, double *source);

// small helpers
void parallel_for(long n, const std::function<void(long, long)> &body);

}  // namespace allium
