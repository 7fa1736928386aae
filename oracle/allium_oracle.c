/*
 * allium_oracle.c -- TEST INFRASTRUCTURE ONLY (see allium_oracle.h).
 *
 * A plain-C FP64 restatement of the reference's DG upwind sweep, written from
 * the behaviour of the reference sources cited at each function.  It stores
 * the same per-element / per-face dense matrices the reference stores and
 * re-assembles every local system per (element, angle, group), exactly as
 * Mult_Upwind does, so that it checks the GPU path and serves as the CPU
 * baseline.  Third-party arithmetic the reference delegates to MFEM
 * (LinearSolve, integrator quadrature orders, CalcOrtho) is restated here from
 * MFEM's published algorithms at the commit pinned by the reference
 * (scripts/linux/build.sh:52, MFEM 171eeb40).
 */
#include "allium_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXP 6
#define ORC_MAXN 64
#define ORC_LU_TOL 1e-9 /* mfem::LinearSolve default TOL */

struct orc_problem {
    int dim, p, q1, n, nf, lump;
    int nel[3], nv[3];
    int Ne, Ndof, Nomega, G;
    double *vert;
    double *omega, *w;
    double wsum;
    int bc[6];
    /* faces: e1 (owner of the normal), e2 (-1 on the boundary) */
    int Nf;
    int *f_e1, *f_e2, *f_axis, *f_side1, *f_bside;
    double *f_nor;
    double *F;   /* Nf x 4 blocks (00,01,10,11), each n x n column-major */
    int *e2f;    /* Ne x 2*dim */
    /* element matrices, column-major n x n */
    double *Mg;  /* G x Ne */
    double *Gd;  /* dim x Ne */
    double *Mt;  /* Ne, allocated by orc_set_time_absorption */
    double *Mm;  /* Ne, MassIntegrator rule, used for the scattering mass */
    double tau;
    int timedep;
    double *sigma;
    /* sweep graph (nodes e*Nomega + a) */
    long nnodes, nedges;
    long *in_ptr;
    int *in_src, *in_face;
    int *order;   /* nodes in frontier order */
    long *lvl_ptr;
    int nlev;
    int *lvl;
    /* negative-flux fixup (src/fixup.cpp): 0 off, 1 zero, 2 zero and scale, 3 ryosuke */
    int fixup;
    double fix_min;
    long fix_applied, fix_failed;
    int fix_error;     /* ryosuke with more than 3 positive unknowns */
    double *fix_monitor;   /* G x Ne counts, or NULL */
};

/* ------------------------------------------------------------------------- */
/* 1D rules on [0,1]                                                         */
/* ------------------------------------------------------------------------- */

static void legendre(int n, double t, double *P, double *dP)
{
    double p0 = 1.0, p1 = t;
    if (n == 0) { *P = 1.0; *dP = 0.0; return; }
    for (int k = 2; k <= n; k++) {
        double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1; p1 = p2;
    }
    *P = p1;
    *dP = n * (t * p1 - p0) / (t * t - 1.0);
}

/* Gauss-Legendre points/weights mapped to [0,1], ascending (the rule
 * mfem::IntRules.Get(SEGMENT, order) returns for npts = order/2 + 1). */
int orc_gauss_rule(int npts, double *x, double *w)
{
    if (npts < 1 || npts > 16) return -1;
    for (int i = 0; i < npts; i++) {
        double t = cos(M_PI * (i + 0.75) / (npts + 0.5));
        double P, dP;
        for (int it = 0; it < 100; it++) {
            legendre(npts, t, &P, &dP);
            double dt = P / dP;
            t -= dt;
            if (fabs(dt) < 1e-17) break;
        }
        legendre(npts, t, &P, &dP);
        double wt = 2.0 / ((1.0 - t * t) * dP * dP);
        x[npts - 1 - i] = 0.5 * (1.0 + t);
        w[npts - 1 - i] = 0.5 * wt;
    }
    return 0;
}

/* Gauss-Lobatto nodes on [0,1] (interior nodes are roots of P'_p). p=0 gives
 * the element centre, matching the constant L2 element. */
int orc_gll_nodes(int p, double *x)
{
    if (p < 0 || p > ORC_MAXP) return -1;
    if (p == 0) { x[0] = 0.5; return 0; }
    x[0] = 0.0; x[p] = 1.0;
    for (int i = 1; i < p; i++) {
        double t = -cos(M_PI * i / p);
        for (int it = 0; it < 100; it++) {
            /* f = P'_p(t); f' = P''_p(t) = (2t P'_p - p(p+1) P_p)/(1-t^2) */
            double P, dP;
            legendre(p, t, &P, &dP);
            double d2P = (2.0 * t * dP - p * (p + 1.0) * P) / (1.0 - t * t);
            double dt = dP / d2P;
            t -= dt;
            if (fabs(dt) < 1e-17) break;
        }
        x[i] = 0.5 * (1.0 + t);
    }
    return 0;
}

static void lagrange_1d(int p, const double *xn, double t, double *val, double *der)
{
    for (int k = 0; k <= p; k++) {
        double v = 1.0, d = 0.0;
        for (int m = 0; m <= p; m++) {
            if (m == k) continue;
            double r = 1.0 / (xn[k] - xn[m]);
            d = d * (t - xn[m]) * r + v * r;
            v *= (t - xn[m]) * r;
        }
        val[k] = v;
        der[k] = d;
    }
}

/* ------------------------------------------------------------------------- */
/* geometry helpers                                                          */
/* ------------------------------------------------------------------------- */

static void elem_index(const orc_problem *pr, int e, int *c)
{
    c[0] = e % pr->nel[0];
    c[1] = (e / pr->nel[0]) % pr->nel[1];
    c[2] = e / (pr->nel[0] * pr->nel[1]);
}

static const double *vertex(const orc_problem *pr, int i, int j, int k)
{
    return pr->vert + (size_t)pr->dim * (i + pr->nv[0] * (j + pr->nv[1] * k));
}

/* d-linear map of element e at reference point xi: X, J (J[m*3+d] = dX_m/dxi_d) */
static void elem_map(const orc_problem *pr, int e, const double *xi, double *X, double *J)
{
    int c[3];
    elem_index(pr, e, c);
    const int dim = pr->dim;
    for (int m = 0; m < 3; m++) { X[m] = 0.0; for (int d = 0; d < 3; d++) J[m * 3 + d] = 0.0; }
    const int ncorner = 1 << dim;
    for (int cc = 0; cc < ncorner; cc++) {
        int b[3] = {cc & 1, (cc >> 1) & 1, (cc >> 2) & 1};
        double N = 1.0, dN[3];
        for (int d = 0; d < dim; d++) N *= b[d] ? xi[d] : 1.0 - xi[d];
        for (int d = 0; d < dim; d++) {
            double v = b[d] ? 1.0 : -1.0;
            for (int d2 = 0; d2 < dim; d2++)
                if (d2 != d) v *= b[d2] ? xi[d2] : 1.0 - xi[d2];
            dN[d] = v;
        }
        const double *xv = vertex(pr, c[0] + b[0], dim > 1 ? c[1] + b[1] : 0, dim > 2 ? c[2] + b[2] : 0);
        for (int m = 0; m < dim; m++) {
            X[m] += N * xv[m];
            for (int d = 0; d < dim; d++) J[m * 3 + d] += dN[d] * xv[m];
        }
    }
}

static double det_adj(int dim, const double *J, double *adj)
{
    /* adj = det(J) * J^{-1}, stored adj[d*3+m] */
    if (dim == 1) { adj[0] = 1.0; return J[0]; }
    if (dim == 2) {
        double a = J[0], b = J[1], c = J[3], d = J[4];
        adj[0] = d; adj[1] = -b; adj[3] = -c; adj[4] = a;
        return a * d - b * c;
    }
    double a = J[0], b = J[1], c = J[2], d = J[3], e = J[4], f = J[5], g = J[6], h = J[7], i = J[8];
    adj[0] = e * i - f * h; adj[1] = c * h - b * i; adj[2] = b * f - c * e;
    adj[3] = f * g - d * i; adj[4] = a * i - c * g; adj[5] = c * d - a * f;
    adj[6] = d * h - e * g; adj[7] = b * g - a * h; adj[8] = a * e - b * d;
    return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
}

/* tensor basis values and reference derivatives at xi */
static void basis_eval(const orc_problem *pr, const double *gll, const double *xi, double *phi,
                       double *dphi /* n x 3 */)
{
    const int p = pr->p, q1 = pr->q1, dim = pr->dim;
    double v[3][ORC_MAXP + 1], d[3][ORC_MAXP + 1];
    for (int a = 0; a < 3; a++) {
        if (a < dim) lagrange_1d(p, gll, xi[a], v[a], d[a]);
        else { for (int k = 0; k <= p; k++) { v[a][k] = (k == 0); d[a][k] = 0.0; } }
    }
    const int qz = dim > 2 ? q1 : 1, qy = dim > 1 ? q1 : 1;
    int i = 0;
    for (int k = 0; k < qz; k++)
        for (int j = 0; j < qy; j++)
            for (int l = 0; l < q1; l++, i++) {
                phi[i] = v[0][l] * v[1][j] * v[2][k];
                dphi[i * 3 + 0] = d[0][l] * v[1][j] * v[2][k];
                dphi[i * 3 + 1] = v[0][l] * d[1][j] * v[2][k];
                dphi[i * 3 + 2] = v[0][l] * v[1][j] * d[2][k];
            }
}

/* a tensor rule over dim axes built from a 1D rule; lumped => corner rule
 * (LumpedIntegrationRule, src/lumping.cpp:3-12) */
typedef struct { int np; double x[512][3]; double w[512]; } rule_t;

static void make_rule(int dim, int npts1d, int lumped, rule_t *r)
{
    double x1[16], w1[16];
    if (lumped) {
        npts1d = 2; x1[0] = 0.0; x1[1] = 1.0; w1[0] = w1[1] = 0.5;
    } else {
        orc_gauss_rule(npts1d, x1, w1);
    }
    const int nz = dim > 2 ? npts1d : 1, ny = dim > 1 ? npts1d : 1;
    r->np = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < npts1d; i++) {
                r->x[r->np][0] = x1[i];
                r->x[r->np][1] = dim > 1 ? x1[j] : 0.0;
                r->x[r->np][2] = dim > 2 ? x1[k] : 0.0;
                r->w[r->np] = w1[i] * (dim > 1 ? w1[j] : 1.0) * (dim > 2 ? w1[k] : 1.0);
                r->np++;
            }
}

/* npts for mfem::IntRules.Get(geom, order): order/2 + 1 */
static int npts_for_order(int order) { return order / 2 + 1; }

/* ------------------------------------------------------------------------- */
/* assembly (src/sweep.cpp:924-1022, src/mg_form.cpp:4-65)                  */
/* ------------------------------------------------------------------------- */

static void assemble_element(orc_problem *pr, int e, const double *gll)
{
    const int n = pr->n, dim = pr->dim, G = pr->G;
    double phi[ORC_MAXN], dphi[ORC_MAXN * 3], X[3], J[9], adj[9];
    rule_t *r = (rule_t *)malloc(sizeof(rule_t));

    /* MGMassIntegrator: order 2p, coefficient sigma_g evaluated per point */
    make_rule(dim, npts_for_order(2 * pr->p), pr->lump & 1, r);
    for (int g = 0; g < G; g++) memset(pr->Mg + ((size_t)g * pr->Ne + e) * n * n, 0, sizeof(double) * n * n);
    for (int q = 0; q < r->np; q++) {
        elem_map(pr, e, r->x[q], X, J);
        double det = det_adj(dim, J, adj);
        basis_eval(pr, gll, r->x[q], phi, dphi);
        double wq = det * r->w[q];
        for (int g = 0; g < G; g++) {
            double *M = pr->Mg + ((size_t)g * pr->Ne + e) * n * n;
            double c = wq * pr->sigma[(size_t)g * pr->Ne + e];
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) M[i + j * n] += c * (phi[i] * phi[j]);
        }
    }

    /* ConservativeConvectionIntegrator(e_d, 1): G_d(i,j) = -int phi_j d_d phi_i;
     * ConvectionIntegrator rule order = OrderGrad + Order + p = dim + 2p - 1 */
    make_rule(dim, npts_for_order(dim + 2 * pr->p - 1), pr->lump & 2, r);
    for (int d = 0; d < dim; d++) memset(pr->Gd + ((size_t)d * pr->Ne + e) * n * n, 0, sizeof(double) * n * n);
    for (int q = 0; q < r->np; q++) {
        elem_map(pr, e, r->x[q], X, J);
        det_adj(dim, J, adj);
        basis_eval(pr, gll, r->x[q], phi, dphi);
        for (int d = 0; d < dim; d++) {
            double *Gm = pr->Gd + ((size_t)d * pr->Ne + e) * n * n;
            /* vec2 = w * adj(J) e_d ; grad_i = dshape_i . vec2 (alpha = -1) */
            double v2[3];
            for (int m = 0; m < dim; m++) v2[m] = -r->w[q] * adj[m * 3 + d];
            for (int i = 0; i < n; i++) {
                double bi = 0.0;
                for (int m = 0; m < dim; m++) bi += dphi[i * 3 + m] * v2[m];
                for (int j = 0; j < n; j++) Gm[i + j * n] += phi[j] * bi;
            }
        }
    }

    /* MassIntegrator rule: order 2p + OrderW, OrderW = dim-1 for d-linear maps */
    make_rule(dim, npts_for_order(2 * pr->p + dim - 1), pr->lump & 1, r);
    double *Mm = pr->Mm + (size_t)e * n * n;
    memset(Mm, 0, sizeof(double) * n * n);
    for (int q = 0; q < r->np; q++) {
        elem_map(pr, e, r->x[q], X, J);
        double det = det_adj(dim, J, adj);
        basis_eval(pr, gll, r->x[q], phi, dphi);
        double wq = r->w[q] * det;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) Mm[i + j * n] += wq * phi[i] * phi[j];
    }
    free(r);
}

/* reference point of element side (axis, side) at face parameters u */
static void face_to_elem(int dim, int axis, int side, const double *u, double *xi)
{
    int k = 0;
    for (int d = 0; d < dim; d++) xi[d] = (d == axis) ? (double)side : u[k++];
    for (int d = dim; d < 3; d++) xi[d] = 0.0;
}

/* surface measure and un-normalised outward normal (CalcOrtho) of element
 * side at reference point xi */
static double face_ortho(const orc_problem *pr, int e, int axis, int side, const double *xi,
                         double *N)
{
    double X[3], J[9];
    elem_map(pr, e, xi, X, J);
    const int dim = pr->dim;
    if (dim == 1) { N[0] = side ? 1.0 : -1.0; return 1.0; }
    int t[2], k = 0;
    for (int d = 0; d < dim; d++) if (d != axis) t[k++] = d;
    if (dim == 2) {
        double tx = J[0 * 3 + t[0]], ty = J[1 * 3 + t[0]];
        N[0] = ty; N[1] = -tx;
    } else {
        double a[3], b[3];
        for (int m = 0; m < 3; m++) { a[m] = J[m * 3 + t[0]]; b[m] = J[m * 3 + t[1]]; }
        N[0] = a[1] * b[2] - a[2] * b[1];
        N[1] = a[2] * b[0] - a[0] * b[2];
        N[2] = a[0] * b[1] - a[1] * b[0];
    }
    /* orient outward: along +dX/dxi_axis on the high side, against it on the low */
    double s = 0.0;
    for (int m = 0; m < dim; m++) s += N[m] * J[m * 3 + axis];
    if ((side == 1 && s < 0) || (side == 0 && s > 0))
        for (int m = 0; m < dim; m++) N[m] = -N[m];
    double len = 0.0;
    for (int m = 0; m < dim; m++) len += N[m] * N[m];
    return sqrt(len);
}

/* FaceMassMatricesIntegrator (src/sweep.cpp:1187-1244), order 2*max(p) */
static void assemble_face(orc_problem *pr, int f, const double *gll)
{
    const int n = pr->n, dim = pr->dim;
    const int e1 = pr->f_e1[f], e2 = pr->f_e2[f], ax = pr->f_axis[f], s1 = pr->f_side1[f];
    double *B = pr->F + (size_t)f * 4 * n * n;
    memset(B, 0, sizeof(double) * 4 * n * n);
    rule_t *r = (rule_t *)malloc(sizeof(rule_t));
    if (dim == 1) { r->np = 1; r->x[0][0] = 0.0; r->w[0] = 1.0; }
    else make_rule(dim - 1, npts_for_order(2 * pr->p), pr->lump & 4, r);
    double phi1[ORC_MAXN], phi2[ORC_MAXN], dphi[ORC_MAXN * 3], xi1[3], xi2[3], N[3];
    for (int q = 0; q < r->np; q++) {
        face_to_elem(dim, ax, s1, r->x[q], xi1);
        double ws = r->w[q] * face_ortho(pr, e1, ax, s1, xi1, N);
        basis_eval(pr, gll, xi1, phi1, dphi);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) B[i + j * n] += phi1[i] * phi1[j] * ws;
        if (e2 >= 0) {
            face_to_elem(dim, ax, 0, r->x[q], xi2);
            basis_eval(pr, gll, xi2, phi2, dphi);
            double *B01 = B + n * n, *B10 = B + 2 * n * n, *B11 = B + 3 * n * n;
            for (int i = 0; i < n; i++)
                for (int j = 0; j < n; j++) {
                    double val = phi1[i] * phi2[j] * ws;
                    B01[i + j * n] += val;
                    B10[j + i * n] += val;
                }
            for (int i = 0; i < n; i++)
                for (int j = 0; j < n; j++) B11[i + j * n] += phi2[i] * phi2[j] * ws;
        }
    }
    /* centre normal, normalised (src/sweep.cpp:142-153) */
    double uc[3] = {0.5, 0.5, 0.5}, xic[3];
    face_to_elem(dim, ax, s1, uc, xic);
    double len = face_ortho(pr, e1, ax, s1, xic, N);
    for (int m = 0; m < dim; m++) pr->f_nor[(size_t)f * dim + m] = N[m] / len;
    free(r);
}

static void assemble_all(orc_problem *pr)
{
    double gll[ORC_MAXP + 1];
    orc_gll_nodes(pr->p, gll);
#pragma omp parallel for schedule(dynamic, 256)
    for (int e = 0; e < pr->Ne; e++) assemble_element(pr, e, gll);
#pragma omp parallel for schedule(dynamic, 256)
    for (int f = 0; f < pr->Nf; f++) assemble_face(pr, f, gll);
}

/* ------------------------------------------------------------------------- */
/* topology, graph and frontiers (src/sweep.cpp:123-208, 308-322, 477-512)    */
/* ------------------------------------------------------------------------- */

static void build_faces(orc_problem *pr)
{
    const int dim = pr->dim;
    long nf = 0;
    for (int ax = 0; ax < dim; ax++) {
        long c = 1;
        for (int d = 0; d < dim; d++) c *= pr->nel[d] + (d == ax);
        nf += c;
    }
    pr->Nf = (int)nf;
    pr->f_e1 = malloc(sizeof(int) * nf); pr->f_e2 = malloc(sizeof(int) * nf);
    pr->f_axis = malloc(sizeof(int) * nf); pr->f_side1 = malloc(sizeof(int) * nf);
    pr->f_bside = malloc(sizeof(int) * nf);
    pr->f_nor = malloc(sizeof(double) * nf * dim);
    pr->e2f = malloc(sizeof(int) * (size_t)pr->Ne * 2 * dim);
    int f = 0;
    for (int ax = 0; ax < dim; ax++) {
        int ext[3] = {1, 1, 1};
        for (int d = 0; d < dim; d++) ext[d] = pr->nel[d] + (d == ax);
        for (int k = 0; k < ext[2]; k++)
            for (int j = 0; j < ext[1]; j++)
                for (int i = 0; i < ext[0]; i++, f++) {
                    int c[3] = {i, j, k};
                    int lo = -1, hi = -1;
                    int cl[3] = {i, j, k};
                    if (c[ax] > 0) { cl[ax] = c[ax] - 1; lo = cl[0] + pr->nel[0] * (cl[1] + pr->nel[1] * cl[2]); }
                    if (c[ax] < pr->nel[ax]) hi = c[0] + pr->nel[0] * (c[1] + pr->nel[1] * c[2]);
                    pr->f_axis[f] = ax;
                    if (lo >= 0 && hi >= 0) {
                        pr->f_e1[f] = lo; pr->f_e2[f] = hi; pr->f_side1[f] = 1; pr->f_bside[f] = -1;
                    } else if (hi >= 0) {
                        pr->f_e1[f] = hi; pr->f_e2[f] = -1; pr->f_side1[f] = 0; pr->f_bside[f] = 2 * ax;
                    } else {
                        pr->f_e1[f] = lo; pr->f_e2[f] = -1; pr->f_side1[f] = 1; pr->f_bside[f] = 2 * ax + 1;
                    }
                    if (lo >= 0) pr->e2f[(size_t)lo * 2 * dim + 2 * ax + 1] = f;
                    if (hi >= 0) pr->e2f[(size_t)hi * 2 * dim + 2 * ax + 0] = f;
                }
    }
}

static int reflect_index(const orc_problem *pr, int a, const double *nor)
{
    const int dim = pr->dim;
    double Om[3], dot = 0.0;
    for (int d = 0; d < dim; d++) dot += pr->omega[a * dim + d] * nor[d];
    for (int d = 0; d < dim; d++) Om[d] = pr->omega[a * dim + d] - 2 * dot * nor[d];
    for (int b = 0; b < pr->Nomega; b++) {
        double diff = 0.0;
        for (int d = 0; d < dim; d++) diff += pow(pr->omega[b * dim + d] - Om[d], 2);
        if (sqrt(diff) < 1e-14) return b;
    }
    return -1;
}

static int build_graph(orc_problem *pr)
{
    const int dim = pr->dim, No = pr->Nomega;
    const long nn = (long)pr->Ne * No;
    pr->nnodes = nn;
    /* count edges */
    long ne = 0;
    for (int f = 0; f < pr->Nf; f++) {
        const double *nor = pr->f_nor + (size_t)f * dim;
        if (pr->f_e2[f] >= 0) ne += No;
        else if (pr->bc[pr->f_bside[f]] == 1) {
            for (int a = 0; a < No; a++) {
                double dot = 0.0;
                for (int d = 0; d < dim; d++) dot += nor[d] * pr->omega[a * dim + d];
                if (dot < 0) ne++;
            }
        }
    }
    pr->nedges = ne;
    int *src = malloc(sizeof(int) * ne), *dst = malloc(sizeof(int) * ne), *fac = malloc(sizeof(int) * ne);
    long k = 0;
    for (int f = 0; f < pr->Nf; f++) {
        const double *nor = pr->f_nor + (size_t)f * dim;
        if (pr->f_e2[f] >= 0) {
            for (int a = 0; a < No; a++) {
                int e1 = pr->f_e1[f], e2 = pr->f_e2[f];
                double dot = 0.0;
                for (int d = 0; d < dim; d++) dot += nor[d] * pr->omega[a * dim + d];
                if (dot < 0) { int t = e1; e1 = e2; e2 = t; }
                src[k] = e1 * No + a; dst[k] = e2 * No + a; fac[k] = f; k++;
            }
        } else if (pr->bc[pr->f_bside[f]] == 1) {
            for (int a = 0; a < No; a++) {
                double dot = 0.0;
                for (int d = 0; d < dim; d++) dot += nor[d] * pr->omega[a * dim + d];
                if (dot < 0) {
                    int ra = reflect_index(pr, a, nor);
                    if (ra < 0) { free(src); free(dst); free(fac); return -3; }
                    src[k] = pr->f_e1[f] * No + ra; dst[k] = pr->f_e1[f] * No + a; fac[k] = f; k++;
                }
            }
        }
    }
    /* in-edge CSR sorted by source id (igraph incident-edge order) and out CSR */
    pr->in_ptr = calloc(nn + 1, sizeof(long));
    long *out_ptr = calloc(nn + 1, sizeof(long));
    for (long i = 0; i < ne; i++) { pr->in_ptr[dst[i] + 1]++; out_ptr[src[i] + 1]++; }
    for (long i = 0; i < nn; i++) { pr->in_ptr[i + 1] += pr->in_ptr[i]; out_ptr[i + 1] += out_ptr[i]; }
    pr->in_src = malloc(sizeof(int) * (ne ? ne : 1));
    pr->in_face = malloc(sizeof(int) * (ne ? ne : 1));
    int *out_dst = malloc(sizeof(int) * (ne ? ne : 1));
    long *fi = malloc(sizeof(long) * nn), *fo = malloc(sizeof(long) * nn);
    for (long i = 0; i < nn; i++) { fi[i] = pr->in_ptr[i]; fo[i] = out_ptr[i]; }
    for (long i = 0; i < ne; i++) {
        long s = fi[dst[i]]++;
        pr->in_src[s] = src[i]; pr->in_face[s] = fac[i];
        out_dst[fo[src[i]]++] = dst[i];
    }
    free(fi); free(fo); free(src); free(dst); free(fac);
    for (long v = 0; v < nn; v++) {
        for (long i = pr->in_ptr[v] + 1; i < pr->in_ptr[v + 1]; i++) {
            int s = pr->in_src[i], fc = pr->in_face[i];
            long j = i - 1;
            while (j >= pr->in_ptr[v] && pr->in_src[j] > s) {
                pr->in_src[j + 1] = pr->in_src[j]; pr->in_face[j + 1] = pr->in_face[j]; j--;
            }
            pr->in_src[j + 1] = s; pr->in_face[j + 1] = fc;
        }
    }
    /* Kahn by whole frontiers; roots ascending */
    int *deg = malloc(sizeof(int) * nn);
    for (long v = 0; v < nn; v++) deg[v] = (int)(pr->in_ptr[v + 1] - pr->in_ptr[v]);
    pr->order = malloc(sizeof(int) * nn);
    pr->lvl = malloc(sizeof(int) * nn);
    long cap = 1024;
    pr->lvl_ptr = malloc(sizeof(long) * cap);
    long head = 0, tail = 0;
    for (long v = 0; v < nn; v++) if (deg[v] == 0) pr->order[tail++] = (int)v;
    int L = 0;
    pr->lvl_ptr[0] = 0;
    while (head < tail) {
        long end = tail;
        for (long i = head; i < end; i++) {
            int v = pr->order[i];
            pr->lvl[v] = L;
            deg[v] = -1;
            for (long o = out_ptr[v]; o < out_ptr[v + 1]; o++) {
                int w = out_dst[o];
                if (--deg[w] == 0) pr->order[tail++] = w;
            }
        }
        head = end;
        L++;
        if (L + 1 >= cap) { cap *= 2; pr->lvl_ptr = realloc(pr->lvl_ptr, sizeof(long) * cap); }
        pr->lvl_ptr[L] = end;
    }
    pr->nlev = L;
    free(deg); free(out_ptr); free(out_dst);
    if (tail != nn) return -2; /* graph is not a dag (src/sweep.cpp:205-208) */
    return 0;
}

/* ------------------------------------------------------------------------- */
/* construction                                                              */
/* ------------------------------------------------------------------------- */

orc_problem *orc_create(int dim, const int *nel, const double *vertices, int p, int nomega,
                        const double *omega, const double *weights, int G,
                        const double *sigma_t, const int *bc, int lump, int *status)
{
    *status = 0;
    if (dim < 1 || dim > 3 || p < 0 || p > ORC_MAXP || nomega < 1 || G < 1) { *status = -1; return NULL; }
    orc_problem *pr = calloc(1, sizeof(orc_problem));
    pr->dim = dim; pr->p = p; pr->q1 = p + 1; pr->lump = lump;
    pr->n = 1; pr->nf = 1;
    for (int d = 0; d < dim; d++) pr->n *= p + 1;
    for (int d = 0; d < dim - 1; d++) pr->nf *= p + 1;
    if (pr->n > ORC_MAXN) { free(pr); *status = -1; return NULL; }
    pr->Ne = 1;
    long nvt = 1;
    for (int d = 0; d < 3; d++) {
        pr->nel[d] = d < dim ? nel[d] : 1;
        pr->nv[d] = d < dim ? nel[d] + 1 : 1;
        pr->Ne *= pr->nel[d];
        nvt *= pr->nv[d];
    }
    pr->Ndof = pr->Ne * pr->n;
    pr->Nomega = nomega; pr->G = G;
    pr->vert = malloc(sizeof(double) * nvt * dim);
    memcpy(pr->vert, vertices, sizeof(double) * nvt * dim);
    pr->omega = malloc(sizeof(double) * nomega * dim);
    pr->w = malloc(sizeof(double) * nomega);
    memcpy(pr->omega, omega, sizeof(double) * nomega * dim);
    memcpy(pr->w, weights, sizeof(double) * nomega);
    pr->wsum = 0.0;
    for (int a = 0; a < nomega; a++) pr->wsum += weights[a];
    for (int s = 0; s < 2 * dim; s++) pr->bc[s] = bc ? bc[s] : 0;
    pr->sigma = malloc(sizeof(double) * (size_t)G * pr->Ne);
    memcpy(pr->sigma, sigma_t, sizeof(double) * (size_t)G * pr->Ne);
    build_faces(pr);
    pr->Mg = malloc(sizeof(double) * (size_t)G * pr->Ne * pr->n * pr->n);
    pr->Gd = malloc(sizeof(double) * (size_t)dim * pr->Ne * pr->n * pr->n);
    pr->Mm = malloc(sizeof(double) * (size_t)pr->Ne * pr->n * pr->n);
    pr->F = malloc(sizeof(double) * (size_t)pr->Nf * 4 * pr->n * pr->n);
    assemble_all(pr);
    int st = build_graph(pr);
    if (st) { *status = st; orc_destroy(pr); return NULL; }
    return pr;
}

void orc_destroy(orc_problem *pr)
{
    if (!pr) return;
    free(pr->vert); free(pr->omega); free(pr->w); free(pr->sigma);
    free(pr->f_e1); free(pr->f_e2); free(pr->f_axis); free(pr->f_side1); free(pr->f_bside);
    free(pr->f_nor); free(pr->F); free(pr->e2f);
    free(pr->Mg); free(pr->Gd); free(pr->Mt); free(pr->Mm);
    free(pr->in_ptr); free(pr->in_src); free(pr->in_face);
    free(pr->order); free(pr->lvl_ptr); free(pr->lvl);
    free(pr);
}

void orc_sizes(const orc_problem *pr, long *out)
{
    out[0] = pr->Ne; out[1] = pr->n; out[2] = pr->Ndof; out[3] = pr->nnodes;
    out[4] = pr->nedges; out[5] = pr->nlev; out[6] = pr->Nf;
    out[7] = (long)pr->G * pr->Nomega * pr->Ndof;
}

int orc_levels(const orc_problem *pr, int *level_of_node)
{
    memcpy(level_of_node, pr->lvl, sizeof(int) * pr->nnodes);
    return pr->nlev;
}

int orc_set_sigma(orc_problem *pr, const double *sigma_t)
{
    memcpy(pr->sigma, sigma_t, sizeof(double) * (size_t)pr->G * pr->Ne);
    double gll[ORC_MAXP + 1];
    orc_gll_nodes(pr->p, gll);
#pragma omp parallel for schedule(dynamic, 256)
    for (int e = 0; e < pr->Ne; e++) assemble_element(pr, e, gll);
    return 0;
}

int orc_set_time_absorption(orc_problem *pr, double tau)
{
    pr->tau = tau;
    pr->timedep = 1;
    if (pr->Mt) return 0;
    /* MassIntegrator without coefficient (src/sweep.cpp:1009-1021) */
    pr->Mt = malloc(sizeof(double) * (size_t)pr->Ne * pr->n * pr->n);
    memcpy(pr->Mt, pr->Mm, sizeof(double) * (size_t)pr->Ne * pr->n * pr->n);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* mfem::LinearSolve (LU with partial pivoting, DenseMatrix column-major)    */
/* ------------------------------------------------------------------------- */

static int linear_solve(int m, double *A, double *X)
{
    if (m == 1) {
        if (fabs(A[0]) <= ORC_LU_TOL) return 0;
        X[0] /= A[0];
        return 1;
    }
    if (m == 2) {
        double det = A[0] * A[3] - A[1] * A[2];
        if (fabs(det) <= ORC_LU_TOL) return 0;
        double invdet = 1.0 / det, b0 = X[0], b1 = X[1];
        X[0] = (A[3] * b0 - A[2] * b1) * invdet;
        X[1] = (-A[1] * b0 + A[0] * b1) * invdet;
        return 1;
    }
    int ipiv[ORC_MAXN];
    for (int i = 0; i < m; i++) {
        int piv = i;
        double a = fabs(A[piv + i * m]);
        for (int j = i + 1; j < m; j++) {
            double b = fabs(A[j + i * m]);
            if (b > a) { a = b; piv = j; }
        }
        ipiv[i] = piv;
        if (piv != i)
            for (int j = 0; j < m; j++) { double t = A[i + j * m]; A[i + j * m] = A[piv + j * m]; A[piv + j * m] = t; }
        if (fabs(A[i + i * m]) <= ORC_LU_TOL) return 0;
        const double inv = 1.0 / A[i + i * m];
        for (int j = i + 1; j < m; j++) A[j + i * m] *= inv;
        for (int k = i + 1; k < m; k++) {
            const double aik = A[i + k * m];
            for (int j = i + 1; j < m; j++) A[j + k * m] -= aik * A[j + i * m];
        }
    }
    for (int i = 0; i < m; i++) { double t = X[i]; X[i] = X[ipiv[i]]; X[ipiv[i]] = t; }
    for (int j = 0; j < m; j++) {
        const double xj = X[j];
        for (int i = j + 1; i < m; i++) X[i] -= A[i + j * m] * xj;
    }
    for (int j = m - 1; j >= 0; j--) {
        const double xj = (X[j] /= A[j + j * m]);
        for (int i = 0; i < j; i++) X[i] -= A[i + j * m] * xj;
    }
    return 1;
}

/* y += a * M x  (DenseMatrix::AddMult_a, column-oriented) */
static void add_mult(int n, const double *M, const double *x, double *y, double a)
{
    for (int c = 0; c < n; c++) {
        const double xc = a * x[c];
        for (int r = 0; r < n; r++) y[r] += xc * M[r + c * n];
    }
}

/* ------------------------------------------------------------------------- */
/* negative-flux fixups (src/fixup.cpp), applied after each local solve      */
/* (src/sweep.cpp:465-469).  A is the assembled local matrix (column-major,  */
/* before the LU), rhs its right-hand side, x the solution.  Returns 1       */
/* applied, -1 failed, 0 not applied; *err is set where the reference aborts */
/* ------------------------------------------------------------------------- */

/* weights(i) = sum_j A(j, i): column sums */
static void column_sums(int n, const double *A, double *wcol)
{
    for (int i = 0; i < n; i++) {
        wcol[i] = 0.0;
        for (int j = 0; j < n; j++) wcol[i] += A[j + (size_t)i * n];
    }
}

/* ZeroFixupOperator::Apply (src/fixup.cpp:3-13) */
static int fixup_zero(int n, double mn, double *x)
{
    int applied = 0;
    for (int i = 0; i < n; i++)
        if (x[i] <= mn) { x[i] = mn; applied = 1; }
    return applied;
}

/* ZeroAndScaleFixupOperator::Apply (src/fixup.cpp:15-57) */
static int fixup_zero_scale(int n, double mn, const double *A, const double *rhs, double *x)
{
    double xmin = x[0];
    for (int i = 1; i < n; i++) xmin = x[i] < xmin ? x[i] : xmin;
    if (xmin > mn) return 0;
    double wcol[ORC_MAXN];
    column_sums(n, A, wcol);
    for (int i = 0; i < n; i++)
        if (x[i] <= mn) x[i] = mn;
    double orig = 0.0, now = 0.0;
    for (int i = 0; i < n; i++) orig += rhs[i];
    for (int i = 0; i < n; i++) now += wcol[i] * x[i];
    if (orig < 0.0 || now < 0.0) return -1;
    if (orig == 0.0 || now == 0.0) {
        for (int i = 0; i < n; i++) x[i] = 0.0;
    } else {
        const double sc = orig / now;
        for (int i = 0; i < n; i++) x[i] *= sc;
    }
    return 1;
}

/* RyosukeFixupOperator::perform_nff (src/fixup.cpp:90-170): clip, then solve
 * for the positive unknowns with their ratios kept and the balance
 * sum_i colsum_i x_i = sum(rhs) restored; at most 3 positive unknowns */
static int fixup_ryosuke(int n, double mn, const double *A, const double *rhs, double *x, int *err)
{
    int cnt[ORC_MAXN], pos = 0;
    for (int i = 0; i < n; i++) {
        if (x[i] <= mn) x[i] = mn;
        else cnt[pos++] = i;
    }
    if (pos == 0 || pos == n) return 0;
    if (pos > 3) { *err = 1; return -1; }
    double wcol[ORC_MAXN];
    column_sums(n, A, wcol);
    double d0 = 0.0;
    for (int j = 0; j < n; j++) d0 += rhs[j];
    const double w0 = wcol[cnt[0]];
    if (pos == 1) {
        x[cnt[0]] = d0 / w0;
    } else if (pos == 2) {
        const double c2 = -x[cnt[1]] / x[cnt[0]];
        const double det = 1.0 / (w0 - wcol[cnt[1]] * c2);
        x[cnt[0]] = d0 * det;
        x[cnt[1]] = -d0 * c2 * det;
    } else {
        const double c3 = -x[cnt[1]] / x[cnt[0]], c6 = -x[cnt[2]] / x[cnt[0]];
        const double det = 1.0 / (w0 - wcol[cnt[1]] * c3 - wcol[cnt[2]] * c6);
        x[cnt[0]] = d0 * det;
        x[cnt[1]] = -d0 * c3 * det;
        x[cnt[2]] = -d0 * c6 * det;
    }
    for (int i = 0; i < pos; i++)
        if (x[cnt[i]] < mn) {
            for (int j = 0; j < pos; j++) x[cnt[j]] = mn;
            return -1;
        }
    return 1;
}

int orc_set_fixup(orc_problem *pr, int type, double psi_min, double *monitor)
{
    if (type < 0 || type > 3) return -1;
    pr->fixup = type;
    pr->fix_min = psi_min;
    pr->fix_monitor = monitor;
    pr->fix_applied = pr->fix_failed = 0;
    pr->fix_error = 0;
    return 0;
}

int orc_fixup_counts(const orc_problem *pr, long *out)
{
    out[0] = pr->fix_applied;
    out[1] = pr->fix_failed;
    out[2] = pr->fix_error;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Mult_Upwind (src/sweep.cpp:277-669)                                       */
/* ------------------------------------------------------------------------- */

long orc_sweep(orc_problem *pr, const double *source, double *psi, int nthreads)
{
    const int n = pr->n, dim = pr->dim, No = pr->Nomega, G = pr->G;
    const size_t Ndof = pr->Ndof;
    long failures = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    long fa = 0, ff = 0;
    int ferr = 0;
    for (int L = 0; L < pr->nlev; L++) {
        const long b = pr->lvl_ptr[L], eidx = pr->lvl_ptr[L + 1];
#pragma omp parallel reduction(+ : failures, fa, ff) reduction(| : ferr)
        {
            double grad[ORC_MAXN * ORC_MAXN], A[ORC_MAXN * ORC_MAXN], rhs[ORC_MAXN], psi2[ORC_MAXN], nor[3];
            double A0[ORC_MAXN * ORC_MAXN], b0[ORC_MAXN];
#pragma omp for schedule(dynamic, 32)
            for (long idx = b; idx < eidx; idx++) {
                const int node = pr->order[idx];
                const int a = node % No, e = node / No;
                const double *Om = pr->omega + (size_t)a * dim;
                const size_t nn = (size_t)n * n;
                memset(grad, 0, sizeof(double) * nn);
                for (int d = 0; d < dim; d++) {
                    const double *Gm = pr->Gd + ((size_t)d * pr->Ne + e) * nn;
                    for (size_t i = 0; i < nn; i++) grad[i] += Om[d] * Gm[i];
                }
                for (int s = 0; s < 2 * dim; s++) {
                    const int f = pr->e2f[(size_t)e * 2 * dim + s];
                    const int keep = (e == pr->f_e1[f]);
                    double dot = 0.0;
                    for (int d = 0; d < dim; d++) {
                        nor[d] = pr->f_nor[(size_t)f * dim + d];
                        if (!keep) nor[d] = -nor[d];
                        dot += Om[d] * nor[d];
                    }
                    const int id = keep ? 0 : 1;
                    if (dot > 0) {
                        const double *Fm = pr->F + ((size_t)f * 4 + id * 2 + id) * nn;
                        for (size_t i = 0; i < nn; i++) grad[i] += dot * Fm[i];
                    }
                }
                for (int g = 0; g < G; g++) {
                    const double *Mg = pr->Mg + ((size_t)g * pr->Ne + e) * nn;
                    for (size_t i = 0; i < nn; i++) A[i] = grad[i];
                    for (size_t i = 0; i < nn; i++) A[i] += 1.0 * Mg[i];
                    const size_t off = Ndof * ((size_t)a + (size_t)No * g) + (size_t)e * n;
                    for (int i = 0; i < n; i++) rhs[i] = source[off + i];
                    for (long ie = pr->in_ptr[node]; ie < pr->in_ptr[node + 1]; ie++) {
                        const int nb = pr->in_src[ie], fid = pr->in_face[ie];
                        const int nbe = nb / No, nba = nb % No;
                        const int keep = (e == pr->f_e1[fid]);
                        double dot = 0.0;
                        for (int d = 0; d < dim; d++) {
                            double v = pr->f_nor[(size_t)fid * dim + d];
                            dot += Om[d] * (keep ? v : -v);
                        }
                        const int id = keep ? 0 : 1, nid = (id + 1) % 2;
                        if (nbe == e) {
                            const double *Fm = pr->F + ((size_t)fid * 4 + id * 2 + id) * nn;
                            const size_t o2 = Ndof * ((size_t)nba + (size_t)No * g) + (size_t)e * n;
                            for (int i = 0; i < n; i++) psi2[i] = psi[o2 + i];
                            add_mult(n, Fm, psi2, rhs, -dot);
                        } else {
                            const double *Fm = pr->F + ((size_t)fid * 4 + id * 2 + nid) * nn;
                            const size_t o2 = Ndof * ((size_t)a + (size_t)No * g) + (size_t)nbe * n;
                            for (int i = 0; i < n; i++) psi2[i] = psi[o2 + i];
                            add_mult(n, Fm, psi2, rhs, -dot);
                        }
                    }
                    if (pr->timedep) {
                        const double *Mt = pr->Mt + (size_t)e * nn;
                        for (size_t i = 0; i < nn; i++) A[i] += pr->tau * Mt[i];
                    }
                    if (pr->fixup) {
                        memcpy(A0, A, sizeof(double) * nn);
                        memcpy(b0, rhs, sizeof(double) * n);
                    }
                    if (!linear_solve(n, A, rhs)) failures++;
                    if (pr->fixup) {
                        int applied = 0;
                        if (pr->fixup == 1) applied = fixup_zero(n, pr->fix_min, rhs);
                        else if (pr->fixup == 2) applied = fixup_zero_scale(n, pr->fix_min, A0, b0, rhs);
                        else applied = fixup_ryosuke(n, pr->fix_min, A0, b0, rhs, &ferr);
                        if (applied == 1) fa++;
                        else if (applied == -1) ff++;
                        if (pr->fix_monitor && applied) {
#pragma omp atomic
                            pr->fix_monitor[e + (size_t)g * pr->Ne] += 1.0;
                        }
                    }
                    for (int i = 0; i < n; i++) psi[off + i] = rhs[i];
                }
            }
        }
    }
    pr->fix_applied += fa;
    pr->fix_failed += ff;
    pr->fix_error |= ferr;
    return failures;
}

/* ------------------------------------------------------------------------- */
/* AdvectionOperator::Mult, ComputeBalance (src/sweep.cpp:1246-1462)         */
/* ------------------------------------------------------------------------- */

int orc_apply_L(const orc_problem *pr, const double *x, double *y)
{
    const int n = pr->n, dim = pr->dim, No = pr->Nomega, G = pr->G;
    const size_t Ndof = pr->Ndof, nn = (size_t)n * n;
    memset(y, 0, sizeof(double) * Ndof * No * G);
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (int g = 0; g < G; g++)
        for (int a = 0; a < No; a++) {
            const double *Om = pr->omega + (size_t)a * dim;
            const size_t base = Ndof * ((size_t)a + (size_t)No * g);
            double grad[ORC_MAXN * ORC_MAXN], yl[ORC_MAXN], yl2[ORC_MAXN];
            for (int e = 0; e < pr->Ne; e++) {
                const double *xl = x + base + (size_t)e * n;
                memset(yl, 0, sizeof(double) * n);
                /* mass_mat->Mult(xlocal, ylocal) */
                const double *Mg = pr->Mg + ((size_t)g * pr->Ne + e) * nn;
                add_mult(n, Mg, xl, yl, 1.0);
                if (pr->timedep) add_mult(n, pr->Mt + (size_t)e * nn, xl, yl, pr->tau);
                memset(grad, 0, sizeof(double) * nn);
                for (int d = 0; d < dim; d++) {
                    const double *Gm = pr->Gd + ((size_t)d * pr->Ne + e) * nn;
                    for (size_t i = 0; i < nn; i++) grad[i] += Om[d] * Gm[i];
                }
                add_mult(n, grad, xl, yl, 1.0);
                for (int i = 0; i < n; i++) y[base + (size_t)e * n + i] += yl[i];
            }
            for (int f = 0; f < pr->Nf; f++) {
                double dot = 0.0;
                for (int d = 0; d < dim; d++) dot += Om[d] * pr->f_nor[(size_t)f * dim + d];
                const int e1 = pr->f_e1[f], e2 = pr->f_e2[f];
                const double *x1 = x + base + (size_t)e1 * n;
                const double *B = pr->F + (size_t)f * 4 * nn;
                memset(yl, 0, sizeof(double) * n);
                if (e2 >= 0) {
                    const double *x2 = x + base + (size_t)e2 * n;
                    memset(yl2, 0, sizeof(double) * n);
                    if (dot >= 0) {
                        add_mult(n, B, x1, yl, dot);
                        add_mult(n, B + 2 * nn, x1, yl2, -dot);
                    } else {
                        add_mult(n, B + nn, x2, yl, dot);
                        add_mult(n, B + 3 * nn, x2, yl2, -dot);
                    }
                    for (int i = 0; i < n; i++) y[base + (size_t)e2 * n + i] += yl2[i];
                } else if (dot >= 0) {
                    add_mult(n, B, x1, yl, dot);
                }
                for (int i = 0; i < n; i++) y[base + (size_t)e1 * n + i] += yl[i];
            }
        }
    return 0;
}

double orc_balance(const orc_problem *pr, const double *psi)
{
    const int n = pr->n, dim = pr->dim, No = pr->Nomega, G = pr->G;
    const size_t Ndof = pr->Ndof, nn = (size_t)n * n;
    double absorption = 0.0, leakage = 0.0;
    double M[ORC_MAXN * ORC_MAXN], loc[ORC_MAXN];
    for (int e = 0; e < pr->Ne; e++)
        for (int g = 0; g < G; g++) {
            memcpy(M, pr->Mg + ((size_t)g * pr->Ne + e) * nn, sizeof(double) * nn);
            if (pr->timedep)
                for (size_t i = 0; i < nn; i++) M[i] += pr->tau * pr->Mt[(size_t)e * nn + i];
            for (int a = 0; a < No; a++) {
                const double *pl = psi + Ndof * ((size_t)a + (size_t)No * g) + (size_t)e * n;
                memset(loc, 0, sizeof(double) * n);
                add_mult(n, M, pl, loc, 1.0);
                double s = 0.0;
                for (int i = 0; i < n; i++) s += loc[i];
                absorption += s * pr->w[a];
            }
        }
    for (int f = 0; f < pr->Nf; f++) {
        if (pr->f_e2[f] >= 0) continue;
        const int e = pr->f_e1[f];
        for (int a = 0; a < No; a++) {
            double dot = 0.0;
            for (int d = 0; d < dim; d++) dot += pr->omega[(size_t)a * dim + d] * pr->f_nor[(size_t)f * dim + d];
            if (dot >= 0)
                for (int g = 0; g < G; g++) {
                    const double *pl = psi + Ndof * ((size_t)a + (size_t)No * g) + (size_t)e * n;
                    memset(loc, 0, sizeof(double) * n);
                    add_mult(n, pr->F + (size_t)f * 4 * nn, pl, loc, 1.0);
                    double s = 0.0;
                    for (int i = 0; i < n; i++) s += loc[i];
                    leakage += pr->w[a] * dot * s;
                }
        }
    }
    return leakage + absorption;
}

/* ------------------------------------------------------------------------- */
/* FormTransportSource (src/sweep.cpp:1160-1185)                             */
/* ------------------------------------------------------------------------- */

int orc_form_source(const orc_problem *pr, const double *Q, const double *inflow, double *source)
{
    const int n = pr->n, dim = pr->dim, No = pr->Nomega, G = pr->G;
    const size_t Ndof = pr->Ndof;
    double gll[ORC_MAXP + 1];
    orc_gll_nodes(pr->p, gll);
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (int g = 0; g < G; g++)
        for (int a = 0; a < No; a++) {
            const double *Om = pr->omega + (size_t)a * dim;
            double *q = source + Ndof * ((size_t)a + (size_t)No * g);
            memset(q, 0, sizeof(double) * Ndof);
            rule_t *r = malloc(sizeof(rule_t));
            double phi[ORC_MAXN], dphi[ORC_MAXN * 3], X[3], J[9], adj[9], xi[3], N[3];
            /* DomainLFIntegrator, order 2p */
            make_rule(dim, npts_for_order(2 * pr->p), 0, r);
            for (int e = 0; e < pr->Ne; e++) {
                const double Qv = Q ? Q[(size_t)g * pr->Ne + e] : 0.0;
                if (Qv == 0.0) continue;
                for (int k = 0; k < r->np; k++) {
                    elem_map(pr, e, r->x[k], X, J);
                    double det = det_adj(dim, J, adj);
                    basis_eval(pr, gll, r->x[k], phi, dphi);
                    double c = r->w[k] * det * Qv;
                    for (int i = 0; i < n; i++) q[(size_t)e * n + i] += c * phi[i];
                }
            }
            /* BoundaryFlowIntegrator(inflow, Omega, -1, -0.5), order 2p */
            if (inflow) {
                if (dim == 1) { r->np = 1; r->x[0][0] = 0.0; r->w[0] = 1.0; }
                else make_rule(dim - 1, npts_for_order(2 * pr->p), 0, r);
                for (int f = 0; f < pr->Nf; f++) {
                    if (pr->f_e2[f] >= 0) continue;
                    const double fv = inflow[(size_t)g * 2 * dim + pr->f_bside[f]];
                    if (fv == 0.0) continue;
                    const int e = pr->f_e1[f];
                    for (int k = 0; k < r->np; k++) {
                        face_to_elem(dim, pr->f_axis[f], pr->f_side1[f], r->x[k], xi);
                        face_ortho(pr, e, pr->f_axis[f], pr->f_side1[f], xi, N);
                        basis_eval(pr, gll, xi, phi, dphi);
                        double un = 0.0;
                        for (int d = 0; d < dim; d++) un += Om[d] * N[d];
                        double w = 0.5 * (-1.0) * un - (-0.5) * fabs(un);
                        w *= r->w[k] * fv;
                        for (int i = 0; i < n; i++) q[(size_t)e * n + i] += w * phi[i];
                    }
                }
            }
            free(r);
        }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* DiscreteToMoment (src/transport_op.cpp:3-46), scattering mass             */
/* ------------------------------------------------------------------------- */

int orc_D(const orc_problem *pr, const double *psi, double *phi)
{
    const size_t Ndof = pr->Ndof;
    memset(phi, 0, sizeof(double) * Ndof * pr->G);
    for (int g = 0; g < pr->G; g++)
        for (int a = 0; a < pr->Nomega; a++) {
            const double *ps = psi + Ndof * ((size_t)a + (size_t)pr->Nomega * g);
            double *ph = phi + Ndof * g;
            const double wa = pr->w[a];
#pragma omp parallel for schedule(static)
            for (size_t s = 0; s < Ndof; s++) ph[s] += wa * ps[s];
        }
    return 0;
}

int orc_DT(const orc_problem *pr, const double *phi, double *psi)
{
    const size_t Ndof = pr->Ndof;
    for (int g = 0; g < pr->G; g++)
        for (int a = 0; a < pr->Nomega; a++) {
            double *ps = psi + Ndof * ((size_t)a + (size_t)pr->Nomega * g);
            const double *ph = phi + Ndof * g;
#pragma omp parallel for schedule(static)
            for (size_t s = 0; s < Ndof; s++) ps[s] = ph[s] / pr->wsum;
        }
    return 0;
}

int orc_mass(const orc_problem *pr, const double *coef, const double *x, double *y)
{
    const int n = pr->n;
    const size_t nn = (size_t)n * n, Ndof = pr->Ndof;
    for (int g = 0; g < pr->G; g++) {
#pragma omp parallel for schedule(static)
        for (int e = 0; e < pr->Ne; e++) {
            double yl[ORC_MAXN];
            memset(yl, 0, sizeof(double) * n);
            add_mult(n, pr->Mm + (size_t)e * nn, x + Ndof * g + (size_t)e * n, yl, coef[e]);
            for (int i = 0; i < n; i++) y[Ndof * g + (size_t)e * n + i] = yl[i];
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* source iteration: chive SLISolver on TransportOperator                    */
/* (exe/chive.cpp:700-735, src/transport_op.hpp:23-33, src/linalg.cpp:103-153)*/
/* ------------------------------------------------------------------------- */

int orc_source_iteration(orc_problem *pr, const double *sigma_s, const double *source,
                         double abstol, double reltol, int maxit, double *phi, double *psi,
                         double *norms, int nthreads)
{
    const size_t Np = (size_t)pr->Ndof * pr->G;
    double *b = malloc(sizeof(double) * Np), *y = malloc(sizeof(double) * Np);
    orc_sweep(pr, source, psi, nthreads);
    orc_D(pr, psi, b);
    memset(phi, 0, sizeof(double) * Np);
    double r0 = 0.0;
    int it;
    for (it = 1;; it++) {
        orc_mass(pr, sigma_s, phi, y);
        orc_DT(pr, y, psi);
        orc_sweep(pr, psi, psi, nthreads);
        orc_D(pr, psi, y);
        double nrm = 0.0;
        for (size_t i = 0; i < Np; i++) {
            /* T x = x - D Linv D^T Ms x ; r = b - T x */
            double r = b[i] - (phi[i] - y[i]);
            phi[i] += r;
            nrm += r * r;
        }
        nrm = sqrt(nrm);
        norms[it - 1] = nrm;
        if (it == 1) r0 = fmax(nrm * reltol, abstol);
        int converged = nrm < r0;
        if (it >= maxit || converged) break;
    }
    orc_mass(pr, sigma_s, phi, y);
    orc_DT(pr, y, psi);
    const size_t Npsi = Np * pr->Nomega;
    for (size_t i = 0; i < Npsi; i++) psi[i] += source[i];
    orc_sweep(pr, psi, psi, nthreads);
    free(b); free(y);
    return it;
}

/* ------------------------------------------------------------------------- */
/* accessors for known-answer tests                                          */
/* ------------------------------------------------------------------------- */

int orc_element_matrices(const orc_problem *pr, int e, double *mass_g0, double *grad, double *tmass)
{
    const size_t nn = (size_t)pr->n * pr->n;
    if (e < 0 || e >= pr->Ne) return -1;
    if (mass_g0) memcpy(mass_g0, pr->Mg + (size_t)e * nn, sizeof(double) * nn);
    if (grad)
        for (int d = 0; d < pr->dim; d++)
            memcpy(grad + d * nn, pr->Gd + ((size_t)d * pr->Ne + e) * nn, sizeof(double) * nn);
    if (tmass) memcpy(tmass, pr->Mm + (size_t)e * nn, sizeof(double) * nn);
    return 0;
}

static int find_face(const orc_problem *pr, int e1, int e2)
{
    for (int s = 0; s < 2 * pr->dim; s++) {
        int f = pr->e2f[(size_t)e1 * 2 * pr->dim + s];
        if ((pr->f_e1[f] == e1 && pr->f_e2[f] == e2) || (pr->f_e1[f] == e2 && pr->f_e2[f] == e1)) return f;
    }
    return -1;
}

int orc_face_blocks(const orc_problem *pr, int e1, int e2, double *blocks)
{
    int f = find_face(pr, e1, e2);
    if (f < 0) return -1;
    memcpy(blocks, pr->F + (size_t)f * 4 * pr->n * pr->n, sizeof(double) * 4 * pr->n * pr->n);
    return 0;
}

int orc_face_normal(const orc_problem *pr, int e1, int e2, double *nor)
{
    int f = find_face(pr, e1, e2);
    if (f < 0) return -1;
    memcpy(nor, pr->f_nor + (size_t)f * pr->dim, sizeof(double) * pr->dim);
    return 0;
}
