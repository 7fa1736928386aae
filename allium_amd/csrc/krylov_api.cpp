// GMRES(m) and FGMRES(m) on the transport operator, the Krylov outer solvers
// chive offers besides source iteration (src/io.cpp:179-230: "gmres" is
// mfem::GMRESSolver, "fgmres" mfem::FGMRESSolver, kdim 50 by default;
// exe/chive.cpp:702-721 applies them to T = I - D L^{-1} D^T Ms with
// b = D L^{-1} q and the DSA operator as preconditioner).
//
// The algorithms are mfem's (linalg/solvers.cpp, releases 4.x; mfem is not
// vendored by the reference): Arnoldi with modified Gram-Schmidt, Givens
// rotations on each Hessenberg column, the estimate |s_{i+1}| tested against
// max(reltol * beta0, abstol) after every step, restart after kdim steps with
// a recomputed residual.  GMRES preconditions from the left (w = M T v, norms of
// M r); FGMRES from the right (z = M v, w = T z, true residual norms) and lets
// the preconditioner change between steps, which the learned correction
// needs: it is not linear.
//
// Device work per step: one operator application (mass, a FP64 sweep, the
// moments), i + 2 fused Gram-Schmidt launches over phi-sized vectors
// (krylov.hip), and one host read of the new Hessenberg column.  The rotations
// and the triangular solve run on the host (kdim <= 200 numbers).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/allium_cnn.h"
#include "../../include/allium_dsa.h"
#include "../../include/allium_sweep.h"
#include "cnn_internal.hpp"
#include "device.hpp"
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

constexpr int kMaxKdim = 200;
constexpr int kUpdateChunk = 64;

struct Krylov {
    allium_sweep *h;
    hipStream_t s;
    long np;
    int m;
    bool flexible;
    double inv_w;
    int nb;
    double *V, *Z, *w, *t, *part[2], *Hd;
    allium_cnn *cnn = nullptr;
    allium_dsa *dsa = nullptr;       // preconditioner in use (DSA or the fallback)
    bool learned_on = false;
    int prolong = 1;
    int inner = 0;                   // CG iterations of the last preconditioner application

    double *v(int i) const { return V + (long)i * np; }
    double *z(int i) const { return Z + (long)i * np; }

    // out = T x = x - D L^{-1} D^T Ms x
    int apply_T(const double *x, double *out)
    {
        HIP_TRY(allium::launch_mass(h->ds.prm, h->ds.prm.Mt, h->coef.d(), inv_w, x, h->iso.d(), s));
        int rc = allium_sweep_internal::sweep(h, nullptr, h->iso.d(), h->a.d(), s, h->y.d());
        if (rc) return rc;
        HIP_TRY(allium::launch_axpby(np, 1.0, x, -1.0, h->y.d(), out, s));
        return ALLIUM_OK;
    }

    bool has_prec() const { return learned_on || dsa; }

    // out = M x (out != x)
    int apply_M(const double *x, double *out)
    {
        inner = 0;
        if (learned_on) {
            HIP_TRY(hipMemcpyAsync(out, x, sizeof(double) * np, hipMemcpyDeviceToDevice, s));
            return allium_cnn_internal::correct(cnn, h->mesh.nel[2], h->mesh.nel[1], h->mesh.nel[0], h->ed.n,
                                                h->wavg.d(), x, h->d_sigma, h->coef.d(), out, s, prolong);
        }
        if (dsa) {
            int rc = allium_dsa_mult_device(dsa, x, out, s);
            if (rc) return rc;
            inner = allium_dsa_internal::last_inner(dsa);
            return ALLIUM_OK;
        }
        HIP_TRY(hipMemcpyAsync(out, x, sizeof(double) * np, hipMemcpyDeviceToDevice, s));
        return ALLIUM_OK;
    }

    // v_0 = r / ||r|| with r in w; returns beta
    int start_cycle(double &beta)
    {
        HIP_TRY(allium::launch_mgs(np, nullptr, nb, nullptr, nullptr, nullptr, w, part[0], s));
        HIP_TRY(allium::launch_normalize(np, part[0], nb, Hd + (long)m * (m + 1), w, v(0), s));
        HIP_TRY(hipMemcpyAsync(&beta, Hd + (long)m * (m + 1), sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        return ALLIUM_OK;
    }

    // residual r = b - T x into w, then (GMRES with M) w = M r
    int residual(const double *b, const double *x)
    {
        int rc = apply_T(x, t);
        if (rc) return rc;
        double *r = (!flexible && has_prec()) ? t : w;
        HIP_TRY(allium::launch_axpby(np, 1.0, b, -1.0, t, r, s));
        if (!flexible && has_prec()) return apply_M(t, w);
        return ALLIUM_OK;
    }

    // x += sum_{j<k} y_j basis_j
    int update(const std::vector<double> &Hh, const std::vector<double> &sv, int k, double *x)
    {
        if (k < 1) return ALLIUM_OK;
        std::vector<double> y(sv.begin(), sv.begin() + k);
        for (int i = k - 1; i >= 0; i--) {
            y[i] /= Hh[(long)i * (m + 1) + i];
            for (int j = i - 1; j >= 0; j--) y[j] -= Hh[(long)i * (m + 1) + j] * y[i];
        }
        double *yd = Hd + (long)(m + 1) * (m + 1);
        HIP_TRY(hipMemcpyAsync(yd, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
        const double *B = flexible ? Z : V;
        for (int j0 = 0; j0 < k; j0 += kUpdateChunk) {
            const int kk = std::min(kUpdateChunk, k - j0);
            HIP_TRY(allium::launch_basis_update(np, kk, B + (long)j0 * np, np, yd + j0, x, s));
        }
        // y lives in pageable host memory: finish the copy before it goes out of scope
        HIP_TRY(hipStreamSynchronize(s));
        return ALLIUM_OK;
    }
};

}  // namespace

extern "C" {

int allium_krylov_default_options(allium_krylov_options *opt)
{
    if (!opt) return fail(ALLIUM_ERR_ARG, "null argument");
    *opt = allium_krylov_options{};
    opt->solver = ALLIUM_SOLVER_FGMRES;
    opt->kdim = 50;
    opt->abstol = 1e-8;
    opt->reltol = 0.0;
    opt->maxit = 100;
    opt->accel_prolong = 1;
    opt->accel_patience = 3;
    return ALLIUM_OK;
}

}  // extern "C"

namespace {

// source, phi and psi are host pointers, or device pointers when `device`
int krylov_solve(allium_sweep_t h, const double *sigma_s, const double *source, const allium_krylov_options *opt,
                 double *phi, double *psi, double *norms, int *iterations, bool device)
{
    if (!h || !sigma_s || !source || !opt || !phi) return fail(ALLIUM_ERR_ARG, "null argument");
    if (opt->solver != ALLIUM_SOLVER_GMRES && opt->solver != ALLIUM_SOLVER_FGMRES)
        return fail(ALLIUM_ERR_ARG, "solver must be ALLIUM_SOLVER_GMRES or ALLIUM_SOLVER_FGMRES");
    if (opt->kdim < 1 || opt->kdim > kMaxKdim) return fail(ALLIUM_ERR_ARG, "kdim must be in 1..200");
    if (opt->maxit < 1) return fail(ALLIUM_ERR_ARG, "maxit < 1");
    if (!(opt->abstol >= 0.0) || !(opt->reltol >= 0.0)) return fail(ALLIUM_ERR_ARG, "negative tolerance");
    for (long e = 0; e < h->mesh.Ne; e++)
        if (!std::isfinite(sigma_s[e])) return fail(ALLIUM_ERR_NUMERIC, "non-finite sigma_s");
    const bool flexible = opt->solver == ALLIUM_SOLVER_FGMRES;
    Krylov K{};
    K.h = h;
    K.s = h->stream;
    K.np = h->phi_size;
    K.m = opt->kdim;
    K.flexible = flexible;
    K.inv_w = 1.0 / h->quad.wsum;
    K.nb = allium::krylov_blocks(K.np);
    allium_dsa *fallback = nullptr;
    if (opt->accel == 1) {
        if (!opt->accel_handle) return fail(ALLIUM_ERR_ARG, "accel 1 needs an allium_cnn_t in accel_handle");
        if (!flexible)
            return fail(ALLIUM_ERR_UNSUPPORTED, "the learned preconditioner is not linear: use FGMRES");
        if (h->mesh.dim != 3 || !h->mesh.uniform || h->G != 1)
            return fail(ALLIUM_ERR_UNSUPPORTED, "the learned correction needs a uniform 3D mesh and one group");
        if (opt->accel_prolong < 0 || opt->accel_prolong > 1 || (opt->accel_prolong == 1 && h->ed.n != 8))
            return fail(ALLIUM_ERR_ARG, "accel_prolong must be 0, or 1 with p = 1");
        if (opt->accel_patience < 1) return fail(ALLIUM_ERR_ARG, "accel_patience must be >= 1");
        K.cnn = static_cast<allium_cnn *>(opt->accel_handle);
        K.learned_on = true;
        K.prolong = opt->accel_prolong;
        fallback = static_cast<allium_dsa *>(opt->accel_fallback);
    } else if (opt->accel == 2) {
        if (!opt->accel_handle) return fail(ALLIUM_ERR_ARG, "accel 2 needs an allium_dsa_t in accel_handle");
        K.dsa = static_cast<allium_dsa *>(opt->accel_handle);
    } else if (opt->accel != 0) {
        return fail(ALLIUM_ERR_ARG, "unknown preconditioner");
    }
    for (allium_dsa *d : {K.dsa, fallback})
        if (d && allium_dsa_internal::handle_of(d) != h)
            return fail(ALLIUM_ERR_ARG, "the DSA object was built on another sweep handle");

    const long np = K.np, nps = h->psi_size;
    const int m = K.m;
    hipStream_t s = K.s;
    if (!device) HIP_TRY(h->q.ensure(sizeof(double) * nps));
    HIP_TRY(h->a.ensure(sizeof(double) * nps));
    HIP_TRY(h->b.ensure(sizeof(double) * np));
    HIP_TRY(h->x.ensure(sizeof(double) * np));
    HIP_TRY(h->y.ensure(sizeof(double) * np));
    HIP_TRY(h->iso.ensure(sizeof(double) * np));
    HIP_TRY(h->coef.ensure(sizeof(double) * h->mesh.Ne));
    HIP_TRY(h->kv.ensure(sizeof(double) * np * (m + 1)));
    if (flexible) HIP_TRY(h->kz.ensure(sizeof(double) * np * m));
    HIP_TRY(h->kw.ensure(sizeof(double) * np));
    HIP_TRY(h->kt.ensure(sizeof(double) * np));
    HIP_TRY(h->kpart.ensure(sizeof(double) * 2 * K.nb));
    HIP_TRY(h->kh.ensure(sizeof(double) * ((size_t)(m + 1) * (m + 1) + m + 1)));
    K.V = h->kv.d();
    K.Z = flexible ? h->kz.d() : nullptr;
    K.w = h->kw.d();
    K.t = h->kt.d();
    K.part[0] = h->kpart.d();
    K.part[1] = h->kpart.d() + K.nb;
    K.Hd = h->kh.d();
    if (K.cnn) {
        // element average = sum_i w_i phi_i, w_i = (row sum of the unit mass) / volume
        const int n = h->ed.n;
        std::vector<double> wv(n, 0.0);
        double vol = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                wv[i] += h->ed.Mu[(size_t)i * n + j];
                vol += h->ed.Mu[(size_t)i * n + j];
            }
        for (double &v : wv) v /= vol;
        HIP_TRY(h->wavg.ensure(sizeof(double) * n));
        HIP_TRY(hipMemcpy(h->wavg.p, wv.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    }
    if (!device) HIP_TRY(hipMemcpyAsync(h->q.p, source, sizeof(double) * nps, hipMemcpyHostToDevice, s));
    const double *dq = device ? source : h->q.d();
    const hipMemcpyKind out_kind = device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    HIP_TRY(hipMemcpyAsync(h->coef.p, sigma_s, sizeof(double) * h->mesh.Ne, hipMemcpyHostToDevice, s));
    // FP64 sweeps inside; the handle's precision setting comes back on every return
    struct PrecisionGuard {
        allium_sweep *h;
        int saved;
        ~PrecisionGuard() { h->ds.precision = saved; }
    } guard{h, h->ds.precision};
    h->ds.precision = 0;
    // every sweep here is followed by the moment kernel's read of psi
    struct StreamGuard {
        allium_sweep *h;
        int saved;
        ~StreamGuard() { h->ds.stream_psi = saved; }
    } sguard{h, h->ds.stream_psi};
    h->ds.stream_psi = 1;

    // b = D L^{-1} q ; x = 0
    int rc = allium_sweep_internal::sweep(h, dq, nullptr, h->a.d(), s, h->b.d());
    if (rc) return rc;
    HIP_TRY(hipMemsetAsync(h->x.p, 0, sizeof(double) * np, s));
    double *b = h->b.d(), *x = h->x.d();
    if (opt->inner_iterations) std::fill(opt->inner_iterations, opt->inner_iterations + opt->maxit, 0);
    if (!flexible && K.has_prec()) {
        rc = K.apply_M(b, K.w);
        if (rc) return rc;
    } else {
        HIP_TRY(hipMemcpyAsync(K.w, b, sizeof(double) * np, hipMemcpyDeviceToDevice, s));
    }
    double beta = 0.0;
    rc = K.start_cycle(beta);
    if (rc) return rc;
    if (!std::isfinite(beta)) return fail(ALLIUM_ERR_NUMERIC, "non-finite initial residual");
    const double stop = std::max(opt->reltol * beta, opt->abstol);
    if (norms) norms[0] = beta;
    int final_iter = 0, learned_steps = 0, slow = 0;
    bool converged = beta <= stop;
    double prev = beta;
    std::vector<double> Hh((size_t)(m + 1) * m, 0.0), sv(m + 1), cs(m), sn(m);
    for (int j = 1; !converged && j <= opt->maxit;) {
        std::fill(sv.begin(), sv.end(), 0.0);
        sv[0] = beta;
        int i = 0;
        for (; i < m && j <= opt->maxit; i++, j++) {
            // w = T z_i (FGMRES) or M T v_i (GMRES)
            if (K.learned_on && opt->accel_max_steps > 0 && learned_steps >= opt->accel_max_steps) {
                K.learned_on = false;
                K.dsa = fallback;
            }
            if (flexible) {
                const double *zi = K.v(i);
                if (K.has_prec()) {
                    if (K.learned_on) learned_steps++;
                    rc = K.apply_M(K.v(i), K.z(i));
                    if (rc) return rc;
                    zi = K.z(i);
                } else {
                    HIP_TRY(hipMemcpyAsync(K.z(i), K.v(i), sizeof(double) * np, hipMemcpyDeviceToDevice, s));
                }
                rc = K.apply_T(zi, K.w);
            } else if (K.has_prec()) {
                rc = K.apply_T(K.v(i), K.t);
                if (!rc) rc = K.apply_M(K.t, K.w);
            } else {
                rc = K.apply_T(K.v(i), K.w);
            }
            if (rc) return rc;
            if (opt->inner_iterations) opt->inner_iterations[j - 1] = K.inner;
            // modified Gram-Schmidt; column i of H at Hd + i (m + 1)
            double *Hcol = K.Hd + (long)i * (m + 1);
            for (int k = 0; k <= i + 1; k++) {
                const double *vprev = k > 0 ? K.v(k - 1) : nullptr;
                const double *vcur = k <= i ? K.v(k) : nullptr;
                HIP_TRY(allium::launch_mgs(np, K.part[(k + 1) & 1], K.nb, k > 0 ? Hcol + (k - 1) : nullptr, vprev,
                                           vcur, K.w, K.part[k & 1], s));
            }
            HIP_TRY(allium::launch_normalize(np, K.part[(i + 1) & 1], K.nb, Hcol + i + 1, K.w, K.v(i + 1), s));
            double *hc = Hh.data() + (size_t)i * (m + 1);
            HIP_TRY(hipMemcpyAsync(hc, Hcol, sizeof(double) * (i + 2), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            for (int k = 0; k < i; k++) allium::apply_givens(hc[k], hc[k + 1], cs[k], sn[k]);
            allium::givens(hc[i], hc[i + 1], cs[i], sn[i]);
            allium::apply_givens(hc[i], hc[i + 1], cs[i], sn[i]);
            allium::apply_givens(sv[i], sv[i + 1], cs[i], sn[i]);
            const double res = std::fabs(sv[i + 1]);
            if (norms) norms[j] = res;
            if (!std::isfinite(res)) return fail(ALLIUM_ERR_NUMERIC, "non-finite residual estimate");
            if (res <= stop) {
                rc = K.update(Hh, sv, i + 1, x);
                if (rc) return rc;
                final_iter = j;
                converged = true;
                break;
            }
            // the learned preconditioner gives way once it stops paying
            if (K.learned_on) {
                slow = res > 0.9 * prev ? slow + 1 : 0;
                if (slow >= opt->accel_patience) {
                    K.learned_on = false;
                    K.dsa = fallback;
                }
            }
            prev = res;
        }
        if (converged) break;
        // end of a cycle: x += basis y, recompute the residual, restart
        rc = K.update(Hh, sv, i, x);
        if (rc) return rc;
        rc = K.residual(b, x);
        if (rc) return rc;
        rc = K.start_cycle(beta);
        if (rc) return rc;
        if (!std::isfinite(beta)) return fail(ALLIUM_ERR_NUMERIC, "non-finite residual");
        if (beta <= stop) {
            // mfem reports the step counter, one past the cycle's last step
            final_iter = j;
            converged = true;
            if (norms) norms[j] = beta;
        }
    }
    if (!converged) final_iter = opt->maxit;
    if (opt->accel_steps) *opt->accel_steps = learned_steps;
    // final sweep: psi = L^{-1}(D^T Ms phi + q)
    if (psi) {
        HIP_TRY(allium::launch_mass(h->ds.prm, h->ds.prm.Mt, h->coef.d(), K.inv_w, x, h->iso.d(), s));
        rc = allium_sweep_internal::sweep(h, dq, h->iso.d(), h->a.d(), s);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(psi, h->a.p, sizeof(double) * nps, out_kind, s));
    }
    HIP_TRY(hipMemcpyAsync(phi, x, sizeof(double) * np, out_kind, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (iterations) *iterations = final_iter;
    if (!converged) return fail(ALLIUM_ERR_NOT_CONVERGED, "Krylov solver reached maxit");
    return ALLIUM_OK;
}

}  // namespace

extern "C" {

int allium_krylov_solve(allium_sweep_t h, const double *sigma_s, const double *source,
                        const allium_krylov_options *opt, double *phi, double *psi, double *norms,
                        int *iterations)
{
    return krylov_solve(h, sigma_s, source, opt, phi, psi, norms, iterations, false);
}

int allium_krylov_solve_device(allium_sweep_t h, const double *sigma_s, const double *d_source,
                               const allium_krylov_options *opt, double *d_phi, double *d_psi, double *norms,
                               int *iterations)
{
    return krylov_solve(h, sigma_s, d_source, opt, d_phi, d_psi, norms, iterations, true);
}

}  // extern "C"
