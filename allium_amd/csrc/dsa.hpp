// Device data and kernel launchers of diffusion synthetic acceleration: the
// scalar MIP diffusion matrix in element-block form, the vertex-grid levels of
// its multigrid preconditioner and the conjugate-gradient vector kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace allium {

// DG level: K x for element e, row i =
//   sum_j Kd[e](i,j) x[e,j] + sum_s sum_j Ko[e,s](i,j) x[nb[e,s],j]
// blocks are stored column-major (entry (i,j) at j*n + i) so that the rows of
// one element are read by consecutive lanes.
struct DgLevel {
    int dim, n, ns, nc;
    int Ne;
    int nel[3];
    const double *Kd;    // Ne x n x n
    const double *Ko;    // Ne x ns x n x n
    const int *nb;       // Ne x ns, -1 on the boundary
    const double *Dinv;  // Ne x n x n, inverse diagonal blocks
    const double *P;     // n x nc, node values of the d-linear vertex functions (row-major)
    double omega;        // damping of the block-Jacobi smoother
};

// Vertex-grid level: 3^dim-point stencil A[k * N + v] for vertex v and offset
// k = (dx+1) + 3 (dy+1) + 9 (dz+1); vertices x-fastest.
struct StLevel {
    int dim;
    int nv[3];
    long long N;
    const double *A;
    const double *dinv;  // 1 / A(v, v)
    double omega;
};

// x_out = x_in + omega Dinv (b - K x_in); x_in == nullptr means x_in = 0
hipError_t dg_jacobi(const DgLevel &L, const double *x_in, const double *b, double *x_out, hipStream_t s);
// r = b - K x
hipError_t dg_resid(const DgLevel &L, const double *x, const double *b, double *r, hipStream_t s);
// y = K x and per-block partial sums of x.y (dot_blocks(Ne * n) of them)
hipError_t dg_spmv_dot(const DgLevel &L, const double *x, double *y, double *partial, hipStream_t s);
// rv = P^T r (DG nodes -> vertices of the element grid)
hipError_t dg_restrict(const DgLevel &L, const double *r, double *rv, hipStream_t s);
// x += P xv
hipError_t dg_prolong_add(const DgLevel &L, const double *xv, double *x, hipStream_t s);

hipError_t st_jacobi(const StLevel &L, const double *x_in, const double *b, double *x_out, hipStream_t s);
hipError_t st_resid(const StLevel &L, const double *x, const double *b, double *r, hipStream_t s);
// coarse = P^T fine for the d-linear interpolation from a grid with every other vertex
hipError_t st_restrict(const StLevel &F, const StLevel &C, const double *rf, double *rc, hipStream_t s);
hipError_t st_prolong_add(const StLevel &F, const StLevel &C, const double *xc, double *xf, hipStream_t s);
// x = Ainv b with a dense N x N inverse
hipError_t dense_apply(long long N, const double *Ainv, const double *b, double *x, hipStream_t s);

// conjugate-gradient helpers; scalars live in device memory sc[]
int dot_blocks(long long n);
// number of partial sums dg_spmv_dot writes
int dg_blocks(const DgLevel &L);
hipError_t dot_partial(long long n, const double *a, const double *b, double *partial, hipStream_t s);
// sc[slot] = sum of nblk partials
hipError_t reduce_to(const double *partial, int nblk, double *sc, int slot, hipStream_t s);
// alpha = sc[inom] / sc[iden]; x += alpha p; r -= alpha q
hipError_t cg_update_xr(long long n, const double *sc, int inom, int iden, const double *p, const double *q,
                        double *x, double *r, hipStream_t s);
// beta = sc[inew] / sc[iold]; p = z + beta p
hipError_t cg_update_p(long long n, const double *sc, int inew, int iold, const double *z, double *p, hipStream_t s);
// x += c; partial sums of (r + c)^2 (the DSA-corrected change)
hipError_t dsa_finish(long long n, const double *r, const double *c, double *x, double *partial, hipStream_t s);

}  // namespace allium
