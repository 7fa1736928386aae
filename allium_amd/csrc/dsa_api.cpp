// C ABI of diffusion synthetic acceleration (include/allium_dsa.h): host
// setup of the MIP matrix and of its multigrid hierarchy, the preconditioned
// conjugate-gradient driver and the hook the source iteration calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/allium_dsa.h"
#include "cnn_internal.hpp"
#include "dsa.hpp"
#include "dsa_internal.hpp"
#include "krylov.hpp"
#include "sweep_handle.hpp"

namespace {

int fail(int code, const std::string &msg)
{
    allium_cnn_internal::set_last_error(msg);
    return code;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return fail(ALLIUM_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

// P1Discretization's alpha (BlockMomentDiscretization default; chive sets no other for p1sa)
constexpr double kP1Alpha = 0.5;

// K in CSR, rows and columns in its element-block layout; row_ptr == NULL
// returns the entry count only
int export_csr(const allium::MipMatrix &K, int Ne, long long *nnz, long long *row_ptr, int *col, double *val)
{
    const int n = K.n, ns = K.nside;
    long long cnt = 0;
    for (int e = 0; e < Ne; e++) {
        int blocks = 1;
        for (int s = 0; s < ns; s++) blocks += K.nb[(size_t)e * ns + s] >= 0;
        cnt += (long long)blocks * n * n;
    }
    if (!row_ptr) { *nnz = cnt; return ALLIUM_OK; }
    if (*nnz < cnt || !col || !val) return fail(ALLIUM_ERR_ARG, "CSR capacity too small");
    long long k = 0;
    for (int e = 0; e < Ne; e++)
        for (int i = 0; i < n; i++) {
            row_ptr[(long long)e * n + i] = k;
            // columns ascending: blocks ordered by element id
            std::vector<std::pair<int, const double *>> blk;
            blk.push_back({e, &K.Kd[((size_t)e * n + i) * n]});
            for (int s = 0; s < ns; s++) {
                const int o = K.nb[(size_t)e * ns + s];
                if (o >= 0) blk.push_back({o, &K.Ko[(((size_t)e * ns + s) * n + i) * n]});
            }
            std::sort(blk.begin(), blk.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
            for (auto &bb : blk)
                for (int j = 0; j < n; j++) { col[k] = bb.first * n + j; val[k] = bb.second[j]; k++; }
        }
    row_ptr[(long long)Ne * n] = k;
    *nnz = k;
    return ALLIUM_OK;
}


// the coarsest vertex grid is inverted densely up to this size; coarsening
// stops once a grid is at most kCoarseTarget vertices
constexpr long long kDenseMax = 1331;
constexpr long long kCoarseTarget = 729;
// Jacobi sweeps on a coarsest grid too large for the dense inverse
constexpr int kCoarseSweeps = 40;
constexpr int kPowerIts = 30;

struct HostStencil {
    int dim = 0;
    int nv[3] = {1, 1, 1};
    long long N = 0;
    std::vector<double> A;  // S x N, entry (k, v) at k * N + v
    std::vector<double> dinv;
    double omega = 0.0;
};

int nstencil(int dim) { return dim > 2 ? 27 : 9; }
int sidx(int dim, int dx, int dy, int dz) { return (dx + 1) + 3 * (dy + 1) + (dim > 2 ? 9 * (dz + 1) : 0); }

allium::MipParams mip_params(const allium_dsa_desc &d, int p, const allium::Quadrature &q)
{
    allium::MipParams m;
    m.kappa = d.kappa < 0 ? -d.kappa * (p + 1) * (p + 1) : d.kappa;
    double alpha = d.alpha;
    if (!(alpha > 0)) {
        // ComputeAlpha(quad, e_x) (src/angular_quadrature.cpp:266-276)
        double a = 0.0;
        for (int i = 0; i < q.size(); i++) a += std::fabs(q.omega[3 * i]) * q.w[i];
        alpha = a / q.wsum;
    }
    m.alpha = alpha;
    m.lower = d.bound_below ? alpha / 2 : 0.0;
    m.sym = d.sym;
    m.lump = 0;
    return m;
}

// chive's "mip" penalty options (exe/chive.cpp:635-650) from the descriptor
allium::BlockIpParams block_ip_params(const allium_dsa_desc &d, int p, double alpha)
{
    allium::BlockIpParams ip;
    ip.kappa = d.kappa < 0 ? -d.kappa * (p + 1) * (p + 1) : d.kappa;
    ip.limit = d.bound_below ? alpha / 2 : 0.0;
    ip.scale_D = d.scale_penalty != 0;
    return ip;
}

// in-place Gauss-Jordan inverse with partial pivoting of a row-major n x n matrix
bool invert(int n, double *A)
{
    std::vector<double> I((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) I[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++)
            if (std::fabs(A[(size_t)r * n + c]) > std::fabs(A[(size_t)piv * n + c])) piv = r;
        if (!(std::fabs(A[(size_t)piv * n + c]) > 0)) return false;
        if (piv != c)
            for (int k = 0; k < n; k++) {
                std::swap(A[(size_t)c * n + k], A[(size_t)piv * n + k]);
                std::swap(I[(size_t)c * n + k], I[(size_t)piv * n + k]);
            }
        const double inv = 1.0 / A[(size_t)c * n + c];
        for (int k = 0; k < n; k++) { A[(size_t)c * n + k] *= inv; I[(size_t)c * n + k] *= inv; }
        for (int r = 0; r < n; r++) {
            if (r == c) continue;
            const double f = A[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; k++) { A[(size_t)r * n + k] -= f * A[(size_t)c * n + k]; I[(size_t)r * n + k] -= f * I[(size_t)c * n + k]; }
        }
    }
    std::memcpy(A, I.data(), sizeof(double) * I.size());
    return true;
}

// y = K x with the row-major host blocks
void host_mv(const allium::MipMatrix &K, int Ne, const double *x, double *y)
{
    const int n = K.n, ns = K.nside;
    allium::parallel_for(Ne, [&](long b, long eend) {
        for (long e = b; e < eend; e++)
            for (int i = 0; i < n; i++) {
                double acc = 0.0;
                const double *Kd = &K.Kd[((size_t)e * n + i) * n];
                for (int j = 0; j < n; j++) acc += Kd[j] * x[e * n + j];
                for (int s = 0; s < ns; s++) {
                    const int o = K.nb[(size_t)e * ns + s];
                    if (o < 0) continue;
                    const double *Ko = &K.Ko[(((size_t)e * ns + s) * n + i) * n];
                    for (int j = 0; j < n; j++) acc += Ko[j] * x[(size_t)o * n + j];
                }
                y[e * n + i] = acc;
            }
    });
}

void host_st_mv(const HostStencil &L, const double *x, double *y)
{
    const int S = nstencil(L.dim);
    allium::parallel_for(L.N, [&](long b, long vend) {
        for (long v = b; v < vend; v++) {
            const int c[3] = {(int)(v % L.nv[0]), (int)((v / L.nv[0]) % L.nv[1]), (int)(v / ((long)L.nv[0] * L.nv[1]))};
            double acc = 0.0;
            for (int k = 0; k < S; k++) {
                const int xx = c[0] + k % 3 - 1, yy = c[1] + (k / 3) % 3 - 1, zz = c[2] + (L.dim > 2 ? k / 9 - 1 : 0);
                if (xx < 0 || xx >= L.nv[0] || yy < 0 || yy >= L.nv[1] || zz < 0 || zz >= L.nv[2]) continue;
                acc += L.A[(size_t)k * L.N + v] * x[xx + (long)L.nv[0] * (yy + (long)L.nv[1] * zz)];
            }
            y[v] = acc;
        }
    });
}

// largest eigenvalue of D^{-1} A by power iteration (D^{-1} A is similar to a
// symmetric positive definite matrix, so its spectrum is real and positive)
template <typename Apply>
double power_max(long long N, Apply apply)
{
    std::vector<double> v(N), u(N);
    for (long long i = 0; i < N; i++) v[i] = 1.0 + 0.5 * std::sin(0.7 * (double)i + 0.3);
    double lam = 0.0;
    for (int it = 0; it < kPowerIts; it++) {
        double nv = 0.0;
        for (double a : v) nv += a * a;
        nv = std::sqrt(nv);
        for (double &a : v) a /= nv;
        apply(v.data(), u.data());
        double nu = 0.0;
        for (double a : u) nu += a * a;
        lam = std::sqrt(nu);
        std::swap(u, v);
    }
    return lam;
}

// Galerkin P^T K P on the vertex grid of the element grid
bool galerkin_dg(const allium::Mesh &m, const allium::MipMatrix &K, HostStencil &L, std::string &err)
{
    const int dim = m.dim, n = K.n, nc = K.nc, ns = K.nside, S = nstencil(dim);
    L.dim = dim;
    L.N = 1;
    for (int d = 0; d < 3; d++) { L.nv[d] = d < dim ? m.nel[d] + 1 : 1; L.N *= L.nv[d]; }
    L.A.assign((size_t)S * L.N, 0.0);
    double scale = 0.0;
    for (long e = 0; e < m.Ne; e++)
        for (int i = 0; i < n; i++) scale = std::max(scale, std::fabs(K.Kd[((size_t)e * n + i) * n + i]));
    std::atomic<bool> bad{false};
    const int ncol = 1 << dim;
    for (int color = 0; color < ncol; color++) {
        allium::parallel_for(m.Ne, [&](long b, long eend) {
            std::vector<double> KP((size_t)n * nc);
            for (long e = b; e < eend; e++) {
                int ec[3];
                m.elem_coords((int)e, ec);
                if (((ec[0] & 1) | ((ec[1] & 1) << 1) | ((ec[2] & 1) << 2)) != color) continue;
                for (int blk = -1; blk < ns; blk++) {
                    const double *B;
                    int oc[3] = {ec[0], ec[1], ec[2]};
                    if (blk < 0) {
                        B = &K.Kd[(size_t)e * n * n];
                    } else {
                        const int o = K.nb[(size_t)e * ns + blk];
                        if (o < 0) continue;
                        B = &K.Ko[((size_t)e * ns + blk) * n * n];
                        m.elem_coords(o, oc);
                    }
                    // KP = B P (n x nc), then P^T KP
                    for (int i = 0; i < n; i++)
                        for (int c = 0; c < nc; c++) {
                            double acc = 0.0;
                            for (int j = 0; j < n; j++) acc += B[(size_t)i * n + j] * K.P[(size_t)j * nc + c];
                            KP[(size_t)i * nc + c] = acc;
                        }
                    for (int c = 0; c < nc; c++) {
                        const int vr[3] = {ec[0] + (c & 1), ec[1] + ((c >> 1) & 1), ec[2] + ((c >> 2) & 1)};
                        const long v = vr[0] + (long)L.nv[0] * (vr[1] + (long)L.nv[1] * vr[2]);
                        for (int c2 = 0; c2 < nc; c2++) {
                            double acc = 0.0;
                            for (int i = 0; i < n; i++) acc += K.P[(size_t)i * nc + c] * KP[(size_t)i * nc + c2];
                            const int off[3] = {oc[0] + (c2 & 1) - vr[0], oc[1] + ((c2 >> 1) & 1) - vr[1],
                                                oc[2] + ((c2 >> 2) & 1) - vr[2]};
                            if (std::abs(off[0]) > 1 || std::abs(off[1]) > 1 || std::abs(off[2]) > 1) {
                                if (std::fabs(acc) > 1e-12 * scale) bad = true;
                                continue;
                            }
                            L.A[(size_t)sidx(dim, off[0], off[1], off[2]) * L.N + v] += acc;
                        }
                    }
                }
            }
        });
    }
    if (bad) { err = "vertex-space Galerkin operator is not a 3^d-point stencil"; return false; }
    return true;
}

// Galerkin coarsening by two of a vertex grid (d-linear interpolation)
bool coarsen(const HostStencil &F, HostStencil &C)
{
    const int dim = F.dim, S = nstencil(dim);
    C.dim = dim;
    C.N = 1;
    for (int d = 0; d < 3; d++) {
        if (d < dim) {
            if (F.nv[d] < 3 || (F.nv[d] - 1) % 2) return false;
            C.nv[d] = (F.nv[d] - 1) / 2 + 1;
        } else {
            C.nv[d] = 1;
        }
        C.N *= C.nv[d];
    }
    C.A.assign((size_t)S * C.N, 0.0);
    const int rz = dim > 2 ? 1 : 0;
    allium::parallel_for(C.N, [&](long b, long Iend) {
        for (long I = b; I < Iend; I++) {
            const int ic[3] = {(int)(I % C.nv[0]), (int)((I / C.nv[0]) % C.nv[1]), (int)(I / ((long)C.nv[0] * C.nv[1]))};
            for (int dz = -rz; dz <= rz; dz++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) {
                        const int fi[3] = {2 * ic[0] + dx, 2 * ic[1] + dy, 2 * ic[2] + dz};
                        if (fi[0] < 0 || fi[0] >= F.nv[0] || fi[1] < 0 || fi[1] >= F.nv[1] || fi[2] < 0 || fi[2] >= F.nv[2])
                            continue;
                        const double wI = (dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0) * (dz ? 0.5 : 1.0);
                        const long i = fi[0] + (long)F.nv[0] * (fi[1] + (long)F.nv[1] * fi[2]);
                        for (int k = 0; k < S; k++) {
                            const double a = F.A[(size_t)k * F.N + i];
                            if (a == 0.0) continue;
                            const int fj[3] = {fi[0] + k % 3 - 1, fi[1] + (k / 3) % 3 - 1, fi[2] + (dim > 2 ? k / 9 - 1 : 0)};
                            int lo[3], cnt[3];
                            for (int d = 0; d < 3; d++) { lo[d] = fj[d] >> 1; cnt[d] = (fj[d] & 1) ? 2 : 1; }
                            for (int kz = 0; kz < cnt[2]; kz++)
                                for (int ky = 0; ky < cnt[1]; ky++)
                                    for (int kx = 0; kx < cnt[0]; kx++) {
                                        const int J[3] = {lo[0] + kx, lo[1] + ky, lo[2] + kz};
                                        const double wJ = (cnt[0] == 2 ? 0.5 : 1.0) * (cnt[1] == 2 ? 0.5 : 1.0) *
                                                          (cnt[2] == 2 ? 0.5 : 1.0);
                                        const int o[3] = {J[0] - ic[0], J[1] - ic[1], J[2] - ic[2]};
                                        C.A[(size_t)sidx(dim, o[0], o[1], o[2]) * C.N + I] += wI * a * wJ;
                                    }
                        }
                    }
        }
    });
    return true;
}

void set_diag(HostStencil &L)
{
    const int c = sidx(L.dim, 0, 0, 0);
    L.dinv.resize(L.N);
    for (long long v = 0; v < L.N; v++) L.dinv[v] = 1.0 / L.A[(size_t)c * L.N + v];
}

template <typename T>
hipError_t upload(std::vector<void *> &owned, T **dst, const T *src, size_t count)
{
    *dst = nullptr;
    if (count == 0) return hipSuccess;
    hipError_t e = hipMalloc((void **)dst, sizeof(T) * count);
    if (e != hipSuccess) return e;
    owned.push_back(*dst);
    if (src) return hipMemcpy(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice);
    return hipSuccess;
}

}  // namespace

struct allium_dsa {
    allium_sweep *h = nullptr;
    allium_dsa_desc desc{};
    allium::DgLevel dg{};
    std::vector<allium::StLevel> st;
    std::vector<double *> sb, sx, st_t, sr;   // per vertex level: rhs, solution, ping-pong, residual
    double *Ainv = nullptr;                   // dense coarsest inverse (row-major) or null
    double *sigs = nullptr;
    // CG and V-cycle work vectors (phi-sized)
    double *bv = nullptr, *xv = nullptr, *rv = nullptr, *zv = nullptr, *pv = nullptr, *qv = nullptr;
    double *zt = nullptr, *res0 = nullptr;
    double *partial = nullptr, *sc = nullptr;
    double *hsc = nullptr;                    // pinned host copy of sc
    int npartial = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // P1SA: p1 holds the 2x2 P1 blocks ((dim + 1) n unknowns per element); the
    // P1 system is solved by GMRES in the vectors below, preconditioned with
    // the help of the MIP hierarchy above
    int method = ALLIUM_DSA_MIP;
    long long pn = 0;                         // P1 unknowns
    allium::DgLevel p1{};
    double *pb = nullptr, *px = nullptr, *pw = nullptr, *pt = nullptr, *pr = nullptr, *pz = nullptr;
    double *pe = nullptr, *pV = nullptr;
    double *ppart = nullptr, *pH = nullptr, *ppart_spmv = nullptr;
    int pnb = 0;
    std::vector<void *> owned;
    allium_dsa_info info{};
    ~allium_dsa()
    {
        for (void *p : owned) hipFree(p);
        if (hsc) hipHostFree(hsc);
        if (ev0) hipEventDestroy(ev0);
        if (ev1) hipEventDestroy(ev1);
    }
};

namespace {

hipError_t vcycle_st(allium_dsa *d, size_t l, hipStream_t s)
{
    const allium::StLevel &L = d->st[l];
    hipError_t e;
    const bool last = l + 1 == d->st.size();
    if (last && d->Ainv) return allium::dense_apply(L.N, d->Ainv, d->sb[l], d->sx[l], s);
    const int nu = last ? kCoarseSweeps : std::max(1, d->desc.smooth);
    const int total = last ? nu : 2 * nu;
    // ping-pong so that the last sweep writes sx[l]
    double *cur = (total % 2) ? d->sx[l] : d->st_t[l];
    double *oth = (cur == d->sx[l]) ? d->st_t[l] : d->sx[l];
    if ((e = allium::st_jacobi(L, nullptr, d->sb[l], cur, s)) != hipSuccess) return e;
    for (int k = 1; k < nu; k++) {
        if ((e = allium::st_jacobi(L, cur, d->sb[l], oth, s)) != hipSuccess) return e;
        std::swap(cur, oth);
    }
    if (last) return hipSuccess;
    if ((e = allium::st_resid(L, cur, d->sb[l], d->sr[l], s)) != hipSuccess) return e;
    if ((e = allium::st_restrict(L, d->st[l + 1], d->sr[l], d->sb[l + 1], s)) != hipSuccess) return e;
    if ((e = vcycle_st(d, l + 1, s)) != hipSuccess) return e;
    if ((e = allium::st_prolong_add(L, d->st[l + 1], d->sx[l + 1], cur, s)) != hipSuccess) return e;
    for (int k = 0; k < nu; k++) {
        if ((e = allium::st_jacobi(L, cur, d->sb[l], oth, s)) != hipSuccess) return e;
        std::swap(cur, oth);
    }
    return hipSuccess;
}

// z = M^{-1} r: one V-cycle, symmetric (equal damped sweeps before and after)
hipError_t vcycle(allium_dsa *d, const double *r, double *z, hipStream_t s)
{
    const allium::DgLevel &L = d->dg;
    const int nu = std::max(1, d->desc.smooth);
    double *cur = d->zt, *oth = z;   // 2 nu sweeps end in z
    hipError_t e;
    if ((e = allium::dg_jacobi(L, nullptr, r, cur, s)) != hipSuccess) return e;
    for (int k = 1; k < nu; k++) {
        if ((e = allium::dg_jacobi(L, cur, r, oth, s)) != hipSuccess) return e;
        std::swap(cur, oth);
    }
    if (!d->st.empty()) {
        if ((e = allium::dg_resid(L, cur, r, d->res0, s)) != hipSuccess) return e;
        if ((e = allium::dg_restrict(L, d->res0, d->sb[0], s)) != hipSuccess) return e;
        if ((e = vcycle_st(d, 0, s)) != hipSuccess) return e;
        if ((e = allium::dg_prolong_add(L, d->sx[0], cur, s)) != hipSuccess) return e;
    }
    for (int k = 0; k < nu; k++) {
        if ((e = allium::dg_jacobi(L, cur, r, oth, s)) != hipSuccess) return e;
        std::swap(cur, oth);
    }
    return hipSuccess;
}

int read_sc(allium_dsa *d, hipStream_t s)
{
    HIP_TRY(hipMemcpyAsync(d->hsc, d->sc, sizeof(double) * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return ALLIUM_OK;
}

// x = K^{-1} b by preconditioned CG from x = 0 (mfem::CGSolver::Mult semantics)
int pcg(allium_dsa *d, const double *b, double *x, hipStream_t s)
{
    const long long N = (long long)d->dg.Ne * d->dg.n;
    const size_t bytes = sizeof(double) * N;
    const int nbd = allium::dot_blocks(N), nbk = allium::dg_blocks(d->dg);
    HIP_TRY(hipMemsetAsync(x, 0, bytes, s));
    HIP_TRY(hipMemcpyAsync(d->rv, b, bytes, hipMemcpyDeviceToDevice, s));
    HIP_TRY(vcycle(d, d->rv, d->zv, s));
    HIP_TRY(hipMemcpyAsync(d->pv, d->zv, bytes, hipMemcpyDeviceToDevice, s));
    HIP_TRY(allium::dot_partial(N, d->zv, d->rv, d->partial, s));
    HIP_TRY(allium::reduce_to(d->partial, nbd, d->sc, 0, s));
    int rc = read_sc(d, s);
    if (rc) return rc;
    const double nom0 = d->hsc[0];
    d->info.last_cg_iterations = 0;
    d->info.last_cg_norm = std::sqrt(std::max(nom0, 0.0));
    if (!std::isfinite(nom0) || nom0 < 0) return fail(ALLIUM_ERR_NUMERIC, "DSA: preconditioned residual is not positive");
    const double r0 = std::max(nom0 * d->desc.cg_reltol * d->desc.cg_reltol, d->desc.cg_abstol * d->desc.cg_abstol);
    if (nom0 <= r0) return ALLIUM_OK;
    HIP_TRY(allium::dg_spmv_dot(d->dg, d->pv, d->qv, d->partial, s));
    HIP_TRY(allium::reduce_to(d->partial, nbk, d->sc, 1, s));
    int inom = 0, inew = 2;
    for (int i = 1;; i++) {
        HIP_TRY(allium::cg_update_xr(N, d->sc, inom, 1, d->pv, d->qv, x, d->rv, s));
        HIP_TRY(vcycle(d, d->rv, d->zv, s));
        HIP_TRY(allium::dot_partial(N, d->rv, d->zv, d->partial, s));
        HIP_TRY(allium::reduce_to(d->partial, nbd, d->sc, inew, s));
        rc = read_sc(d, s);
        if (rc) return rc;
        const double den = d->hsc[1], betanom = d->hsc[inew];
        if (!(den > 0) || !std::isfinite(betanom))
            return fail(ALLIUM_ERR_NUMERIC, "DSA: CG breakdown (operator or preconditioner not positive definite)");
        d->info.last_cg_iterations = i;
        d->info.last_cg_norm = std::sqrt(std::max(betanom, 0.0));
        if (betanom <= r0 || i >= d->desc.cg_maxit) break;
        HIP_TRY(allium::cg_update_p(N, d->sc, inew, inom, d->zv, d->pv, s));
        HIP_TRY(allium::dg_spmv_dot(d->dg, d->pv, d->qv, d->partial, s));
        HIP_TRY(allium::reduce_to(d->partial, nbk, d->sc, 1, s));
        std::swap(inom, inew);
    }
    d->info.cg_iterations += d->info.last_cg_iterations;
    return ALLIUM_OK;
}

// GMRES restart length of the P1 solve
constexpr int kP1Kdim = 40;
constexpr int kP1Chunk = 64;

// DG blocks go to the device column-major
void to_colmajor(int n, const std::vector<double> &src, std::vector<double> &dst)
{
    dst.resize(src.size());
    const size_t nb = src.size() / ((size_t)n * n);
    allium::parallel_for((long)nb, [&](long b, long bend) {
        for (long k = b; k < bend; k++)
            for (int i = 0; i < n; i++)
                for (int j = 0; j < n; j++) dst[(size_t)k * n * n + (size_t)j * n + i] = src[(size_t)k * n * n + (size_t)i * n + j];
    });
}

// P1SA: the P1 level and its GMRES vectors, added to a complete MIP object
// (whose V-cycle preconditions the P1 solve)
int add_p1(allium_dsa *d, const double *siga)
{
    const allium::Mesh &m = d->h->mesh;
    const int Ne = m.Ne, mk = kP1Kdim;
    allium::MipMatrix K;
    std::string why;
    bool ok_asm;
    if (d->method == ALLIUM_DSA_BLOCK_MIP || d->method == ALLIUM_DSA_LDG) {
        const double alpha = mip_params(d->desc, d->h->ed.p, d->h->quad).alpha;
        allium::BlockIpParams ip = block_ip_params(d->desc, d->h->ed.p, alpha);
        ip.ldg = d->method == ALLIUM_DSA_LDG;
        ok_asm = allium::assemble_p1(m, d->h->ed.p, d->h->sigma.data(), siga, alpha, K, why, &ip);
    } else {
        ok_asm = allium::assemble_p1(m, d->h->ed.p, d->h->sigma.data(), siga, kP1Alpha, K, why);
    }
    if (!ok_asm) return fail(ALLIUM_ERR_ARG, why);
    const int n = K.n;
    std::vector<double> Dinv(K.Kd);
    std::atomic<bool> ok{true};
    allium::parallel_for(Ne, [&](long b, long eend) {
        for (long e = b; e < eend; e++)
            if (!invert(n, &Dinv[(size_t)e * n * n])) ok = false;
    });
    if (!ok) return fail(ALLIUM_ERR_NUMERIC, "DSA: singular diagonal block of the P1 matrix");
    {
        std::vector<double> cm;
        double *dp;
        int *ip;
        to_colmajor(n, K.Kd, cm);
        HIP_TRY(upload(d->owned, &dp, cm.data(), cm.size()));
        d->p1.Kd = dp;
        to_colmajor(n, K.Ko, cm);
        HIP_TRY(upload(d->owned, &dp, cm.data(), cm.size()));
        d->p1.Ko = dp;
        to_colmajor(n, Dinv, cm);
        HIP_TRY(upload(d->owned, &dp, cm.data(), cm.size()));
        d->p1.Dinv = dp;
        HIP_TRY(upload(d->owned, &ip, K.nb.data(), K.nb.size()));
        d->p1.nb = ip;
    }
    d->p1.dim = m.dim;
    d->p1.n = n;
    d->p1.ns = K.nside;
    d->p1.nc = 0;
    d->p1.P = nullptr;
    d->p1.Ne = Ne;
    for (int k = 0; k < 3; k++) d->p1.nel[k] = m.nel[k];
    d->p1.omega = 1.0;
    d->pn = (long long)Ne * n;
    // the P1 level goes first in the info, the MIP levels follow
    if (d->info.levels >= 16) return fail(ALLIUM_ERR_UNSUPPORTED, "DSA: too many levels");
    for (int l = d->info.levels; l > 0; l--) {
        d->info.level_size[l] = d->info.level_size[l - 1];
        d->info.omega[l] = d->info.omega[l - 1];
    }
    d->info.levels++;
    d->info.level_size[0] = d->pn;
    d->info.omega[0] = 1.0;
    for (double **q : {&d->pb, &d->px, &d->pw, &d->pt, &d->pr, &d->pz})
        HIP_TRY(upload<double>(d->owned, q, nullptr, (size_t)d->pn));
    // pe: phi-part embedding, zero outside the phi slots
    HIP_TRY(upload<double>(d->owned, &d->pe, nullptr, (size_t)d->pn));
    HIP_TRY(hipMemset(d->pe, 0, sizeof(double) * d->pn));
    HIP_TRY(upload<double>(d->owned, &d->pV, nullptr, (size_t)(mk + 1) * d->pn));
    d->pnb = allium::krylov_blocks(d->pn);
    HIP_TRY(upload<double>(d->owned, &d->ppart, nullptr, (size_t)2 * d->pnb));
    HIP_TRY(upload<double>(d->owned, &d->pH, nullptr, (size_t)(mk + 1) * (mk + 1) + mk + 1));
    // dg_spmv_dot on the P1 level writes its own partial sums
    HIP_TRY(upload<double>(d->owned, &d->ppart_spmv, nullptr, (size_t)allium::dg_blocks(d->p1)));
    return ALLIUM_OK;
}

// z = B r, the P1 preconditioner: block-Jacobi, a MIP V-cycle on the phi-part
// of the remaining residual (the MIP operator stands in for the Schur
// complement of the P1 system on phi), block-Jacobi again
int p1_prec(allium_dsa *d, const double *r, double *z, hipStream_t s)
{
    const allium::DgLevel &L = d->p1;
    const int n = L.n / (L.dim + 1);
    const size_t rowb = sizeof(double) * n, pitch = sizeof(double) * L.n;
    const size_t off = (size_t)L.dim * n;
    HIP_TRY(allium::dg_jacobi(L, nullptr, r, d->pz, s));
    HIP_TRY(allium::dg_resid(L, d->pz, r, d->pr, s));
    // rv, zv: the MIP CG vectors, free in this mode
    HIP_TRY(hipMemcpy2DAsync(d->rv, rowb, d->pr + off, pitch, rowb, L.Ne, hipMemcpyDeviceToDevice, s));
    HIP_TRY(vcycle(d, d->rv, d->zv, s));
    HIP_TRY(hipMemcpy2DAsync(d->pe + off, pitch, d->zv, rowb, rowb, L.Ne, hipMemcpyDeviceToDevice, s));
    HIP_TRY(allium::launch_axpby(d->pn, 1.0, d->pz, 1.0, d->pe, d->pz, s));
    HIP_TRY(allium::dg_jacobi(L, d->pz, r, z, s));
    return ALLIUM_OK;
}

// phi-part of P^{-1} [0; b] (exe/chive.cpp:608-617: the TripleProductOperator
// of ComponentExtractionOperator around the P1 solve).  The reference solves P
// with a direct solver; here GMRES(kP1Kdim), preconditioned from the right by
// p1_prec, from x = 0 until the true residual
// ||[0; b] - P x|| <= max(cg_reltol ||b||, cg_abstol) or cg_maxit steps.
int p1_solve(allium_dsa *d, const double *b, double *x, hipStream_t s)
{
    const allium::DgLevel &L = d->p1;
    const long long np = d->pn;
    const int m = kP1Kdim, nb = d->pnb, n = L.n / (L.dim + 1), Ne = L.Ne;
    const size_t rowb = sizeof(double) * n, pitch = sizeof(double) * L.n;
    double *part[2] = {d->ppart, d->ppart + nb};
    double *Hd = d->pH, *yd = d->pH + (size_t)(m + 1) * (m + 1);
    auto V = [&](int i) { return d->pV + (size_t)i * np; };
    // pb = [0; b] element by element
    HIP_TRY(hipMemsetAsync(d->pb, 0, sizeof(double) * np, s));
    HIP_TRY(hipMemcpy2DAsync(d->pb + (size_t)L.dim * n, pitch, b, rowb, rowb, Ne, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemsetAsync(d->px, 0, sizeof(double) * np, s));
    HIP_TRY(hipMemcpyAsync(d->pw, d->pb, sizeof(double) * np, hipMemcpyDeviceToDevice, s));
    double beta = 0.0;
    auto start_cycle = [&]() -> int {
        HIP_TRY(allium::launch_mgs(np, nullptr, nb, nullptr, nullptr, nullptr, d->pw, part[0], s));
        HIP_TRY(allium::launch_normalize(np, part[0], nb, Hd + (size_t)m * (m + 1), d->pw, V(0), s));
        HIP_TRY(hipMemcpyAsync(d->hsc, Hd + (size_t)m * (m + 1), sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        beta = d->hsc[0];
        return std::isfinite(beta) ? ALLIUM_OK : fail(ALLIUM_ERR_NUMERIC, "P1SA: non-finite residual");
    };
    std::vector<double> Hh((size_t)(m + 1) * m, 0.0), sv(m + 1), cs(m), sn(m), y;
    // px += Dinv (sum_{j<k} y_j V_j)
    auto update = [&](int k) -> int {
        if (k < 1) return ALLIUM_OK;
        y.assign(sv.begin(), sv.begin() + k);
        for (int i = k - 1; i >= 0; i--) {
            y[i] /= Hh[(size_t)i * (m + 1) + i];
            for (int j = i - 1; j >= 0; j--) y[j] -= Hh[(size_t)i * (m + 1) + j] * y[i];
        }
        HIP_TRY(hipMemcpyAsync(yd, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemsetAsync(d->pt, 0, sizeof(double) * np, s));
        for (int j0 = 0; j0 < k; j0 += kP1Chunk)
            HIP_TRY(allium::launch_basis_update(np, std::min(kP1Chunk, k - j0), V(j0), np, yd + j0, d->pt, s));
        int rc = p1_prec(d, d->pt, d->pw, s);
        if (rc) return rc;
        HIP_TRY(allium::launch_axpby(np, 1.0, d->px, 1.0, d->pw, d->px, s));
        // y lives in pageable host memory
        HIP_TRY(hipStreamSynchronize(s));
        return ALLIUM_OK;
    };
    int rc = start_cycle();
    if (rc) return rc;
    const double stop = std::max(d->desc.cg_reltol * beta, d->desc.cg_abstol);
    int its = 0;
    double res = beta;
    bool done = beta <= stop;
    while (!done && its < d->desc.cg_maxit) {
        std::fill(sv.begin(), sv.end(), 0.0);
        sv[0] = beta;
        int i = 0;
        for (; i < m && its < d->desc.cg_maxit; i++) {
            its++;
            // w = P B v_i
            if ((rc = p1_prec(d, V(i), d->pt, s))) return rc;
            HIP_TRY(allium::dg_spmv_dot(L, d->pt, d->pw, d->ppart_spmv, s));
            double *Hcol = Hd + (size_t)i * (m + 1);
            for (int k = 0; k <= i + 1; k++)
                HIP_TRY(allium::launch_mgs(np, part[(k + 1) & 1], nb, k > 0 ? Hcol + (k - 1) : nullptr,
                                           k > 0 ? V(k - 1) : nullptr, k <= i ? V(k) : nullptr, d->pw,
                                           part[k & 1], s));
            HIP_TRY(allium::launch_normalize(np, part[(i + 1) & 1], nb, Hcol + i + 1, d->pw, V(i + 1), s));
            double *hc = Hh.data() + (size_t)i * (m + 1);
            HIP_TRY(hipMemcpyAsync(hc, Hcol, sizeof(double) * (i + 2), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            for (int k = 0; k < i; k++) allium::apply_givens(hc[k], hc[k + 1], cs[k], sn[k]);
            allium::givens(hc[i], hc[i + 1], cs[i], sn[i]);
            allium::apply_givens(hc[i], hc[i + 1], cs[i], sn[i]);
            allium::apply_givens(sv[i], sv[i + 1], cs[i], sn[i]);
            res = std::fabs(sv[i + 1]);
            if (!std::isfinite(res)) return fail(ALLIUM_ERR_NUMERIC, "P1SA: GMRES breakdown");
            if (res <= stop) { i++; done = true; break; }
        }
        if ((rc = update(i))) return rc;
        if (done || its >= d->desc.cg_maxit) break;
        // restart from the true residual
        HIP_TRY(allium::dg_resid(L, d->px, d->pb, d->pw, s));
        if ((rc = start_cycle())) return rc;
        res = beta;
        done = beta <= stop;
    }
    d->info.last_cg_iterations = its;
    d->info.last_cg_norm = res;
    d->info.cg_iterations += its;
    HIP_TRY(hipMemcpy2DAsync(x, rowb, d->px + (size_t)L.dim * n, pitch, rowb, Ne, hipMemcpyDeviceToDevice, s));
    return ALLIUM_OK;
}

// x = K^{-1} b (MIP) or the phi-part of P^{-1} [0; b] (P1SA)
int inner_solve(allium_dsa *d, const double *b, double *x, hipStream_t s)
{
    return d->method == ALLIUM_DSA_MIP ? pcg(d, b, x, s) : p1_solve(d, b, x, s);
}

}  // namespace

namespace allium_dsa_internal {

int correct(allium_dsa *d, const double *d_r, double *d_x, hipStream_t s, double *nrm)
{
    allium_sweep *h = d->h;
    const long long N = h->phi_size;
    HIP_TRY(hipEventRecord(d->ev0, s));
    HIP_TRY(allium::launch_mass(h->ds.prm, h->ds.prm.Mt, d->sigs, 1.0, d_r, d->bv, s));
    int rc = inner_solve(d, d->bv, d->xv, s);
    if (rc) return rc;
    HIP_TRY(allium::dsa_finish(N, d_r, d->xv, d_x, d->partial, s));
    HIP_TRY(allium::reduce_to(d->partial, allium::dot_blocks(N), d->sc, 4, s));
    HIP_TRY(hipEventRecord(d->ev1, s));
    rc = read_sc(d, s);
    if (rc) return rc;
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, d->ev0, d->ev1));
    d->info.seconds += 1e-3 * ms;
    d->info.applications++;
    *nrm = std::sqrt(d->hsc[4]);
    return ALLIUM_OK;
}

const allium_sweep *handle_of(const allium_dsa *d) { return d->h; }

int last_inner(const allium_dsa *d) { return d->info.last_cg_iterations; }

}  // namespace allium_dsa_internal

extern "C" {

void allium_dsa_default_desc(allium_dsa_desc *desc)
{
    if (!desc) return;
    desc->kappa = -1.0;
    desc->bound_below = 1;
    desc->alpha = 0.0;
    desc->sym = -1.0;
    desc->cg_reltol = 1e-6;
    desc->cg_abstol = 0.0;
    desc->cg_maxit = 100;
    desc->smooth = 2;
    desc->method = ALLIUM_DSA_MIP;
    desc->scale_penalty = 1;
}

int allium_dsa_create(allium_sweep_t h, const double *sigma_s, const allium_dsa_desc *desc, allium_dsa_t *out)
{
    if (!h || !sigma_s || !out) return fail(ALLIUM_ERR_ARG, "null argument");
    *out = nullptr;
    if (h->G != 1) return fail(ALLIUM_ERR_UNSUPPORTED, "DSA is implemented for one group");
    // the vertex coarse space needs nodal traces (p >= 1): with p = 0 the
    // Galerkin operator couples vertices two elements apart
    if (h->ed.p < 1) return fail(ALLIUM_ERR_UNSUPPORTED, "DSA needs fe_order >= 1");
    // chive assembles its DSA operators with the transport's lumping
    // (exe/chive.cpp:610-666); these are assembled unlumped
    if (h->ed.lump) return fail(ALLIUM_ERR_UNSUPPORTED, "DSA is implemented for unlumped sweep handles (lump = 0)");
    auto d = std::make_unique<allium_dsa>();
    d->h = h;
    if (desc) d->desc = *desc; else allium_dsa_default_desc(&d->desc);
    if (d->desc.cg_maxit < 1 || d->desc.smooth < 1) return fail(ALLIUM_ERR_ARG, "cg_maxit and smooth must be >= 1");
    if (d->desc.method < ALLIUM_DSA_MIP || d->desc.method > ALLIUM_DSA_LDG)
        return fail(ALLIUM_ERR_ARG, "method must be ALLIUM_DSA_MIP, _P1, _BLOCK_MIP or _LDG");
    d->method = d->desc.method;
    const bool p1 = d->method != ALLIUM_DSA_MIP;
    const allium::Mesh &m = h->mesh;
    const int Ne = m.Ne, p = h->ed.p;
    std::vector<double> siga(Ne);
    for (int e = 0; e < Ne; e++) {
        if (!std::isfinite(sigma_s[e])) return fail(ALLIUM_ERR_NUMERIC, "non-finite sigma_s");
        siga[e] = h->sigma[e] - sigma_s[e];
    }
    allium::MipMatrix K;
    std::string why;
    if (!allium::assemble_mip(m, p, h->sigma.data(), siga.data(), mip_params(d->desc, p, h->quad), K, why))
        return fail(ALLIUM_ERR_ARG, why);
    const int n = K.n, ns = K.nside;
    // block-Jacobi inverses and the smoother damping of the DG level
    std::vector<double> Dinv(K.Kd);
    std::atomic<bool> ok{true};
    allium::parallel_for(Ne, [&](long b, long eend) {
        for (long e = b; e < eend; e++)
            if (!invert(n, &Dinv[(size_t)e * n * n])) ok = false;
    });
    if (!ok) return fail(ALLIUM_ERR_NUMERIC, "DSA: singular diagonal block of the MIP matrix");
    const long long N0 = (long long)Ne * n;
    std::vector<double> tmp(N0);
    const double lam0 = power_max(N0, [&](const double *v, double *u) {
        host_mv(K, Ne, v, tmp.data());
        allium::parallel_for(Ne, [&](long b, long eend) {
            for (long e = b; e < eend; e++)
                for (int i = 0; i < n; i++) {
                    double acc = 0.0;
                    for (int j = 0; j < n; j++) acc += Dinv[((size_t)e * n + i) * n + j] * tmp[e * n + j];
                    u[e * n + i] = acc;
                }
        });
    });
    // vertex levels
    std::vector<HostStencil> lev(1);
    if (!galerkin_dg(m, K, lev[0], why)) return fail(ALLIUM_ERR_NUMERIC, why);
    while (lev.back().N > kCoarseTarget && lev.size() < 15) {
        HostStencil c;
        if (!coarsen(lev.back(), c)) break;
        lev.push_back(std::move(c));
    }
    for (auto &L : lev) {
        set_diag(L);
        for (double v : L.dinv)
            if (!(v > 0) || !std::isfinite(v)) return fail(ALLIUM_ERR_NUMERIC, "DSA: non-positive vertex-grid diagonal");
        const double lam = power_max(L.N, [&](const double *v, double *u) {
            host_st_mv(L, v, u);
            for (long long i = 0; i < L.N; i++) u[i] *= L.dinv[i];
        });
        L.omega = 4.0 / (3.0 * 1.05 * lam);
    }
    std::vector<double> Ainv;
    const HostStencil &Lc = lev.back();
    if (Lc.N <= kDenseMax) {
        const int Nc = (int)Lc.N, S = nstencil(Lc.dim);
        Ainv.assign((size_t)Nc * Nc, 0.0);
        for (int v = 0; v < Nc; v++) {
            const int c[3] = {v % Lc.nv[0], (v / Lc.nv[0]) % Lc.nv[1], v / (Lc.nv[0] * Lc.nv[1])};
            for (int k = 0; k < S; k++) {
                const int xx = c[0] + k % 3 - 1, yy = c[1] + (k / 3) % 3 - 1, zz = c[2] + (Lc.dim > 2 ? k / 9 - 1 : 0);
                if (xx < 0 || xx >= Lc.nv[0] || yy < 0 || yy >= Lc.nv[1] || zz < 0 || zz >= Lc.nv[2]) continue;
                Ainv[(size_t)v * Nc + xx + Lc.nv[0] * (yy + Lc.nv[1] * zz)] += Lc.A[(size_t)k * Lc.N + v];
            }
        }
        // row-major: each device wave reads one row
        if (!invert(Nc, Ainv.data())) return fail(ALLIUM_ERR_NUMERIC, "DSA: singular coarsest operator");
    }
    // upload: DG blocks column-major
    auto colmajor = [&](const std::vector<double> &src, std::vector<double> &dst) { to_colmajor(n, src, dst); };
    {
        std::vector<double> cm;
        double *dp;
        colmajor(K.Kd, cm);
        HIP_TRY(upload(d->owned, &dp, cm.data(), cm.size()));
        d->dg.Kd = dp;
        colmajor(K.Ko, cm);
        HIP_TRY(upload(d->owned, &dp, cm.data(), cm.size()));
        d->dg.Ko = dp;
        colmajor(Dinv, cm);
        HIP_TRY(upload(d->owned, &dp, cm.data(), cm.size()));
        d->dg.Dinv = dp;
        HIP_TRY(upload(d->owned, &dp, K.P.data(), K.P.size()));
        d->dg.P = dp;
        int *ip;
        HIP_TRY(upload(d->owned, &ip, K.nb.data(), K.nb.size()));
        d->dg.nb = ip;
    }
    d->dg.dim = m.dim;
    d->dg.n = n;
    d->dg.ns = ns;
    d->dg.nc = K.nc;
    d->dg.Ne = Ne;
    for (int k = 0; k < 3; k++) d->dg.nel[k] = m.nel[k];
    d->dg.omega = 4.0 / (3.0 * 1.05 * lam0);
    d->info.levels = 1 + (int)lev.size();
    d->info.level_size[0] = N0;
    d->info.omega[0] = d->dg.omega;
    for (size_t l = 0; l < lev.size(); l++) {
        allium::StLevel L{};
        L.dim = lev[l].dim;
        for (int k = 0; k < 3; k++) L.nv[k] = lev[l].nv[k];
        L.N = lev[l].N;
        L.omega = lev[l].omega;
        double *dp;
        HIP_TRY(upload(d->owned, &dp, lev[l].A.data(), lev[l].A.size()));
        L.A = dp;
        HIP_TRY(upload(d->owned, &dp, lev[l].dinv.data(), lev[l].dinv.size()));
        L.dinv = dp;
        d->st.push_back(L);
        double *b4[4];
        for (auto &q : b4) HIP_TRY(upload<double>(d->owned, &q, nullptr, (size_t)L.N));
        d->sb.push_back(b4[0]); d->sx.push_back(b4[1]); d->st_t.push_back(b4[2]); d->sr.push_back(b4[3]);
        d->info.level_size[l + 1] = L.N;
        d->info.omega[l + 1] = L.omega;
    }
    if (!Ainv.empty()) HIP_TRY(upload(d->owned, &d->Ainv, Ainv.data(), Ainv.size()));
    d->info.dense_coarsest = Ainv.empty() ? 0 : 1;
    HIP_TRY(upload(d->owned, &d->sigs, sigma_s, (size_t)Ne));
    for (double **q : {&d->bv, &d->xv, &d->rv, &d->zv, &d->pv, &d->qv, &d->zt, &d->res0})
        HIP_TRY(upload<double>(d->owned, q, nullptr, (size_t)N0));
    d->npartial = std::max(allium::dot_blocks(N0), allium::dg_blocks(d->dg));
    HIP_TRY(upload<double>(d->owned, &d->partial, nullptr, (size_t)d->npartial));
    HIP_TRY(upload<double>(d->owned, &d->sc, nullptr, 8));
    HIP_TRY(hipMemset(d->sc, 0, sizeof(double) * 8));
    HIP_TRY(hipHostMalloc((void **)&d->hsc, sizeof(double) * 8, hipHostMallocDefault));
    HIP_TRY(hipEventCreate(&d->ev0));
    HIP_TRY(hipEventCreate(&d->ev1));
    if (p1) {
        int rc = add_p1(d.get(), siga.data());
        if (rc) return rc;
    }
    *out = d.release();
    return ALLIUM_OK;
}

void allium_dsa_destroy(allium_dsa_t d) { delete d; }

int allium_dsa_mult_device(allium_dsa_t d, const double *d_x, double *d_y, void *stream)
{
    if (!d || !d_x || !d_y) return fail(ALLIUM_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    allium_sweep *h = d->h;
    HIP_TRY(allium::launch_mass(h->ds.prm, h->ds.prm.Mt, d->sigs, 1.0, d_x, d->bv, s));
    int rc = inner_solve(d, d->bv, d->xv, s);
    if (rc) return rc;
    HIP_TRY(allium::launch_axpby(h->phi_size, 1.0, d_x, 1.0, d->xv, d_y, s));
    d->info.applications++;
    return ALLIUM_OK;
}

int allium_dsa_solve_device(allium_dsa_t d, const double *d_b, double *d_x, void *stream)
{
    if (!d || !d_b || !d_x) return fail(ALLIUM_ERR_ARG, "null argument");
    return inner_solve(d, d_b, d_x, (hipStream_t)stream);
}

int allium_dsa_solve(allium_dsa_t d, const double *b, double *x)
{
    if (!d || !b || !x) return fail(ALLIUM_ERR_ARG, "null argument");
    hipStream_t s = d->h->stream;
    const size_t bytes = sizeof(double) * d->h->phi_size;
    HIP_TRY(hipMemcpyAsync(d->bv, b, bytes, hipMemcpyHostToDevice, s));
    int rc = inner_solve(d, d->bv, d->xv, s);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(x, d->xv, bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return ALLIUM_OK;
}

int allium_dsa_get_info(allium_dsa_t d, allium_dsa_info *info)
{
    if (!d || !info) return fail(ALLIUM_ERR_ARG, "null argument");
    *info = d->info;
    return ALLIUM_OK;
}

int allium_host_mip_matrix(const allium_mesh_desc *mesh, int fe_order, const allium_quad_desc *quad,
                           const double *sigma_t, const double *sigma_s, const allium_dsa_desc *desc, long long *nnz,
                           long long *row_ptr, int *col, double *val)
{
    if (!mesh || !quad || !sigma_t || !sigma_s || !nnz) return fail(ALLIUM_ERR_ARG, "null argument");
    allium::Mesh m;
    allium::Quadrature q;
    std::string why;
    if (!allium::build_mesh(mesh->dim, mesh->nel, mesh->vertices, mesh->extents, mesh->bc, m, why))
        return fail(ALLIUM_ERR_ARG, why);
    int rc = allium_internal_quad(quad, mesh->dim, q);
    if (rc) return rc;
    allium_dsa_desc dd;
    if (desc) dd = *desc; else allium_dsa_default_desc(&dd);
    std::vector<double> siga(m.Ne);
    for (int e = 0; e < m.Ne; e++) siga[e] = sigma_t[e] - sigma_s[e];
    allium::MipMatrix K;
    if (!allium::assemble_mip(m, fe_order, sigma_t, siga.data(), mip_params(dd, fe_order, q), K, why))
        return fail(ALLIUM_ERR_ARG, why);
    return export_csr(K, m.Ne, nnz, row_ptr, col, val);
}

static int host_block_matrix(bool ldg, const allium_mesh_desc *mesh, int fe_order, const allium_quad_desc *quad,
                             const double *sigma_t, const double *sigma_s, const allium_dsa_desc *desc,
                             long long *nnz, long long *row_ptr, int *col, double *val)
{
    if (!mesh || !quad || !sigma_t || !sigma_s || !nnz) return fail(ALLIUM_ERR_ARG, "null argument");
    allium::Mesh m;
    allium::Quadrature q;
    std::string why;
    if (!allium::build_mesh(mesh->dim, mesh->nel, mesh->vertices, mesh->extents, mesh->bc, m, why))
        return fail(ALLIUM_ERR_ARG, why);
    int rc = allium_internal_quad(quad, mesh->dim, q);
    if (rc) return rc;
    allium_dsa_desc dd;
    if (desc) dd = *desc; else allium_dsa_default_desc(&dd);
    std::vector<double> siga(m.Ne);
    for (int e = 0; e < m.Ne; e++) siga[e] = sigma_t[e] - sigma_s[e];
    const double alpha = mip_params(dd, fe_order, q).alpha;
    allium::BlockIpParams ip = block_ip_params(dd, fe_order, alpha);
    ip.ldg = ldg;
    allium::MipMatrix K;
    if (!allium::assemble_p1(m, fe_order, sigma_t, siga.data(), alpha, K, why, &ip)) return fail(ALLIUM_ERR_ARG, why);
    return export_csr(K, m.Ne, nnz, row_ptr, col, val);
}

int allium_host_block_mip_matrix(const allium_mesh_desc *mesh, int fe_order, const allium_quad_desc *quad,
                                 const double *sigma_t, const double *sigma_s, const allium_dsa_desc *desc,
                                 long long *nnz, long long *row_ptr, int *col, double *val)
{
    return host_block_matrix(false, mesh, fe_order, quad, sigma_t, sigma_s, desc, nnz, row_ptr, col, val);
}

int allium_host_block_ldg_matrix(const allium_mesh_desc *mesh, int fe_order, const allium_quad_desc *quad,
                                 const double *sigma_t, const double *sigma_s, const allium_dsa_desc *desc,
                                 long long *nnz, long long *row_ptr, int *col, double *val)
{
    return host_block_matrix(true, mesh, fe_order, quad, sigma_t, sigma_s, desc, nnz, row_ptr, col, val);
}

int allium_host_p1_matrix(const allium_mesh_desc *mesh, int fe_order, const double *sigma_t, const double *sigma_s,
                          long long *nnz, long long *row_ptr, int *col, double *val)
{
    if (!mesh || !sigma_t || !sigma_s || !nnz) return fail(ALLIUM_ERR_ARG, "null argument");
    allium::Mesh m;
    std::string why;
    if (!allium::build_mesh(mesh->dim, mesh->nel, mesh->vertices, mesh->extents, mesh->bc, m, why))
        return fail(ALLIUM_ERR_ARG, why);
    std::vector<double> siga(m.Ne);
    for (int e = 0; e < m.Ne; e++) siga[e] = sigma_t[e] - sigma_s[e];
    allium::MipMatrix K;
    if (!allium::assemble_p1(m, fe_order, sigma_t, siga.data(), kP1Alpha, K, why)) return fail(ALLIUM_ERR_ARG, why);
    return export_csr(K, m.Ne, nnz, row_ptr, col, val);
}

}  // extern "C"
