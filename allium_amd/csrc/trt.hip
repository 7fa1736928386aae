// Multigroup linearized backward-Euler thermal radiative transfer on the
// MI355X sweep: element-local emission/energy-balance kernels and the
// time-step driver behind include/allium_trt.h.
//
// One step restates LinearizedTRTOperator::Mult (src/trt_linearized.cpp:230-312):
//   resid = (c_v/dt) M T^n - Bt(T);  Bt' = (c_v/dt) M + dGray(T);  t1 = Bt'^{-1} resid
//   em = B(T) + dB t1;  phi_src = D Linv(D^T em + q_psi)
//   (I - D Linv D^T Ms) phi = phi_src,  Ms = dB Bt'^{-1} M_sigma   (source iteration)
//   em += Ms phi;  psi = Linv(D^T em + q_psi)
//   T += Bt'^{-1}(M_sigma phi + resid), floored at 1e-8
// with q_psi = inflow + (1/dt) Mt psi^n and dimensionless black-body emission T^4.
// Every element-local operator is one thread per element over the tensor
// Gauss rules the reference's integrators use (order 2p+1 for the group
// emission, 2p+2 for the gray energy balance).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "../../include/allium_device.h"
#include "../../include/allium_trt.h"
#include "cnn_internal.hpp"
#include "device.hpp"
#include "discretization.hpp"
#include "planck.hpp"

namespace {

constexpr int N = 8;       // p = 1 hexahedra
constexpr int GMAX = 64;   // groups a thread keeps spectra for

struct TrtParams {
    int Ne, G, nq3, nq4;
    double cvdt;               // c_v / dt
    const double *phi3, *w3;   // emission rule: basis values nq3 x N, weights
    const double *phi4, *w4;   // energy-balance rule
    const double *Mm;          // MassIntegrator-rule element mass (c_v, time mass)
    const double *Mu;          // MG mass (absorption)
    const double *bounds;      // G + 1
    const double *sigma;       // G x Ne
};

__device__ inline double dot_basis(const double *phi_q, const double *x)
{
    double v = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) v = fma(phi_q[j], x[j], v);
    return v;
}

// Gauss-Jordan inverse with partial pivoting of an N x N matrix (in place into inv)
__device__ inline bool invert(double (&A)[N][N], double (&inv)[N][N])
{
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) inv[i][j] = i == j ? 1.0 : 0.0;
    for (int k = 0; k < N; ++k) {
        int piv = k;
        double best = fabs(A[k][k]);
        for (int i = k + 1; i < N; ++i)
            if (fabs(A[i][k]) > best) { best = fabs(A[i][k]); piv = i; }
        if (!(best > 0.0)) return false;
        if (piv != k)
            for (int j = 0; j < N; ++j) {
                double t = A[k][j]; A[k][j] = A[piv][j]; A[piv][j] = t;
                t = inv[k][j]; inv[k][j] = inv[piv][j]; inv[piv][j] = t;
            }
        const double r = 1.0 / A[k][k];
        for (int j = 0; j < N; ++j) { A[k][j] *= r; inv[k][j] *= r; }
        for (int i = 0; i < N; ++i) {
            if (i == k) continue;
            const double f = A[i][k];
            if (f == 0.0) continue;
            for (int j = 0; j < N; ++j) { A[i][j] -= f * A[k][j]; inv[i][j] -= f * inv[k][j]; }
        }
    }
    return true;
}

// sigma[g][e] = coef / (mid_g^3 sqrt(T at the element centre))
__global__ void k_opacity(int Ne, int G, const double *__restrict__ T, const double *__restrict__ mid, double coef,
                          double *__restrict__ sigma)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Ne) return;
    double tc = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) tc += T[(long long)e * N + i];
    tc *= 1.0 / N;   // trilinear value at the centre
    const double rs = 1.0 / sqrt(tc);
    for (int g = 0; g < G; ++g) sigma[(long long)g * Ne + e] = coef * rs / (mid[g] * mid[g] * mid[g]);
}

// per (element, quadrature point): the Planck sums, which dominate the
// linearization.  A rule-4 point stores its gray emission and derivative
// weights W4[e][q][2] (Rosseland spectrum); a rule-3 point stores the group
// emission weights E3[e][q][g] and the emission-gradient weights
// dpw[e][q][g] = w_q sigma_g 4 T_q^3 r_g(T_q)
__global__ void __launch_bounds__(256) k_planck_points(TrtParams P, const double *__restrict__ T,
                                                       double *__restrict__ W4, double *__restrict__ E3,
                                                       double *__restrict__ dpw)
{
    const int nq = P.nq4 + P.nq3;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)P.Ne * nq) return;
    const int e = (int)(t / nq), q = (int)(t - (long long)e * nq);
    const double *Te = T + (long long)e * N;
    double tv[N];
#pragma unroll
    for (int i = 0; i < N; ++i) tv[i] = Te[i];
    const double *sig = P.sigma + e;
    const long long gs = P.Ne;
    if (q < P.nq4) {
        const double *ph = P.phi4 + q * N;
        const double Tq = dot_basis(ph, tv);
        double sp = 0.0, sr = 0.0;
        allium::planck_spectrum_each<0>(P.bounds, P.G, Tq, [&](int g, double v) { sp = fma(sig[g * gs], v, sp); });
        allium::planck_spectrum_each<1>(P.bounds, P.G, Tq, [&](int g, double v) { sr = fma(sig[g * gs], v, sr); });
        const double T3 = Tq * Tq * Tq;
        W4[((long long)e * P.nq4 + q) * 2] = P.w4[q] * sp * T3 * Tq;     // emission
        W4[((long long)e * P.nq4 + q) * 2 + 1] = P.w4[q] * 4.0 * sr * T3;  // its derivative
    } else {
        const int r = q - P.nq4;
        const double *ph = P.phi3 + r * N;
        const double Tq = dot_basis(ph, tv);
        const double T3 = Tq * Tq * Tq;
        double *e3 = E3 + ((long long)e * P.nq3 + r) * P.G;
        double *dp = dpw + ((long long)e * P.nq3 + r) * P.G;
        const double w = P.w3[r];
        allium::planck_spectrum_each<0>(P.bounds, P.G, Tq,
                                        [&](int g, double v) { e3[g] = w * sig[g * gs] * v * T3 * Tq; });
        allium::planck_spectrum_each<1>(P.bounds, P.G, Tq,
                                        [&](int g, double v) { dp[g] = w * sig[g * gs] * 4.0 * v * T3; });
    }
}

// per element, from the point weights of k_planck_points: resid, Bt'^{-1}
// and the group emission em (phi-sized)
__global__ void __launch_bounds__(64) k_linearize(TrtParams P, const double *__restrict__ T,
                                                  const double *__restrict__ Tn, const double *__restrict__ W4,
                                                  const double *__restrict__ E3, double *__restrict__ resid,
                                                  double *__restrict__ Binv, double *__restrict__ em,
                                                  int *__restrict__ bad)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= P.Ne) return;
    const double *Te = T + (long long)e * N;
    double tv[N], A[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i) tv[i] = Te[i];
    // (c_v/dt) M and resid = (c_v/dt) M (T^n - T)
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double r = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            A[i][j] = P.cvdt * P.Mm[i * N + j];
            r = fma(A[i][j], Tn[(long long)e * N + j] - tv[j], r);
        }
        resid[(long long)e * N + i] = r;
    }
    // gray emission and its gradient (Rosseland spectrum), rule 4
    for (int q = 0; q < P.nq4; ++q) {
        const double *ph = P.phi4 + q * N;
        const double we = W4[((long long)e * P.nq4 + q) * 2];
        const double wg = W4[((long long)e * P.nq4 + q) * 2 + 1];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            resid[(long long)e * N + i] -= we * ph[i];
#pragma unroll
            for (int j = 0; j < N; ++j) A[i][j] = fma(wg * ph[i], ph[j], A[i][j]);
        }
    }
    double inv[N][N];
    if (!invert(A, inv)) atomicAdd(bad, 1);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) Binv[((long long)e * N + i) * N + j] = inv[i][j];
    // group emission, rule 3
    for (int i = 0; i < N; ++i)
        for (int g = 0; g < P.G; ++g) em[((long long)g * P.Ne + e) * N + i] = 0.0;
    for (int q = 0; q < P.nq3; ++q) {
        const double *ph = P.phi3 + q * N;
        for (int g = 0; g < P.G; ++g) {
            const double s = E3[((long long)e * P.nq3 + q) * P.G + g];
            double *eg = em + ((long long)g * P.Ne + e) * N;
            for (int i = 0; i < N; ++i) eg[i] = fma(s, ph[i], eg[i]);
        }
    }
}

// out[g][e] (+)= scale * dB_g u_e with u_e = Binv_e (src_e), src in T space,
// in two kernels.  Per element: u_e and its values at the rule-3 points,
// UQ[e][q] = phi3_q . u_e; mode 0: src = x directly; mode 1: src = sum_g
// sigma_g Mu phi_g (Ms phi)
__global__ void __launch_bounds__(64) k_dB_points(TrtParams P, int mode, const double *__restrict__ x,
                                                  const double *__restrict__ Binv, double *__restrict__ UQ)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= P.Ne) return;
    double t[N], u[N];
    if (mode == 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) t[i] = x[(long long)e * N + i];
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) t[i] = 0.0;
        for (int g = 0; g < P.G; ++g) {
            const double sg = P.sigma[(long long)g * P.Ne + e];
            const double *xg = x + ((long long)g * P.Ne + e) * N;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                double a = 0.0;
#pragma unroll
                for (int j = 0; j < N; ++j) a = fma(P.Mu[i * N + j], xg[j], a);
                t[i] = fma(sg, a, t[i]);
            }
        }
    }
    const double *B = Binv + (long long)e * N * N;
#pragma unroll
    for (int i = 0; i < N; ++i) u[i] = dot_basis(B + i * N, t);
    for (int q = 0; q < P.nq3; ++q) UQ[(long long)e * P.nq3 + q] = dot_basis(P.phi3 + q * N, u);
}

// per (element, group), group fastest: out[g][e][i] (+)= scale sum_q dpw[e][q][g] UQ[e][q] phi3_q[i]
__global__ void __launch_bounds__(256) k_dB_groups(TrtParams P, const double *__restrict__ UQ,
                                                   const double *__restrict__ dpw, double scale, int accumulate,
                                                   double *__restrict__ out)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)P.Ne * P.G) return;
    const int e = (int)(t / P.G), g = (int)(t - (long long)e * P.G);
    double v[N];
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = 0.0;
    for (int q = 0; q < P.nq3; ++q) {
        const double sq = dpw[((long long)e * P.nq3 + q) * P.G + g] * UQ[(long long)e * P.nq3 + q];
        const double *ph = P.phi3 + q * N;
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = fma(sq, ph[i], v[i]);
    }
    double *og = out + ((long long)g * P.Ne + e) * N;
#pragma unroll
    for (int i = 0; i < N; ++i) og[i] = accumulate ? fma(scale, v[i], og[i]) : scale * v[i];
}

// T += Binv (sum_g sigma_g Mu phi_g + resid), floored at 1e-8; Traw (may be
// null) receives the update before the floor
__global__ void k_update_T(TrtParams P, const double *__restrict__ phi, const double *__restrict__ resid,
                           const double *__restrict__ Binv, double *__restrict__ T, int *__restrict__ floors,
                           double *__restrict__ Traw)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= P.Ne) return;
    double t[N];
#pragma unroll
    for (int i = 0; i < N; ++i) t[i] = resid[(long long)e * N + i];
    for (int g = 0; g < P.G; ++g) {
        const double sg = P.sigma[(long long)g * P.Ne + e];
        const double *xg = phi + ((long long)g * P.Ne + e) * N;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double a = 0.0;
#pragma unroll
            for (int j = 0; j < N; ++j) a = fma(P.Mu[i * N + j], xg[j], a);
            t[i] = fma(sg, a, t[i]);
        }
    }
    const double *B = Binv + (long long)e * N * N;
    int fl = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double v = T[(long long)e * N + i] + dot_basis(B + i * N, t);
        if (Traw) Traw[(long long)e * N + i] = v;
        if (!(v >= 1e-8)) { v = 1e-8; fl++; }
        T[(long long)e * N + i] = v;
    }
    if (fl) atomicAdd(floors, fl);
}

// q_psi = (1/dt) Mt psi^n, one thread per cell (8 consecutive doubles in, 8 out)
__global__ void k_time_source(long long cells, double tau, const double *__restrict__ Mm,
                              const double *__restrict__ psin, double *__restrict__ out)
{
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    double x[N];
#pragma unroll
    for (int j = 0; j < N; ++j) x[j] = psin[c * N + j];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double a = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) a = fma(Mm[i * N + j], x[j], a);
        out[c * N + i] = tau * a;
    }
}

// the inflow source, nonzero on boundary cells only: q_psi[k] = q_in[k] + (1/dt) (Mt psi^n)[k]
// for the listed entries, with the same fused rounding as a dense pass
__global__ void k_time_source_inflow(long long nnz, double tau, const double *__restrict__ Mm,
                                     const long long *__restrict__ idx, const double *__restrict__ val,
                                     const double *__restrict__ psin, double *__restrict__ out)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const long long t = idx[k];
    const int i = (int)(t % N);
    const double *x = psin + (t - i);
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) a = fma(Mm[i * N + j], x[j], a);
    out[t] = fma(tau, a, val[k]);
}

// part[blk] = sum (a - b)^2 over a grid-stride share (summed on the host in block order)
__global__ void k_diff2(long long n, const double *__restrict__ a, const double *__restrict__ b,
                        double *__restrict__ part)
{
    __shared__ double sh[256];
    double acc = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double d = a[i] - b[i];
        acc = fma(d, d, acc);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_minmax(long long n, const double *__restrict__ x, double *__restrict__ part)
{
    __shared__ double lo[256], hi[256];
    double a = 1e300, b = -1e300;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        a = fmin(a, x[i]);
        b = fmax(b, x[i]);
    }
    lo[threadIdx.x] = a;
    hi[threadIdx.x] = b;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) {
            lo[threadIdx.x] = fmin(lo[threadIdx.x], lo[threadIdx.x + k]);
            hi[threadIdx.x] = fmax(hi[threadIdx.x], hi[threadIdx.x + k]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = lo[0]; part[2 * blockIdx.x + 1] = hi[0]; }
}

// part[blk] = sum of x[i] w[i % N] over a grid-stride share (summed on the host in block order)
__global__ void k_wsum(long long n, const double *__restrict__ x, const double *__restrict__ w,
                       double *__restrict__ part)
{
    __shared__ double sh[256];
    double acc = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        acc = fma(x[i], w[i % N], acc);
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Two block sums written to part[2 blk], part[2 blk + 1] (summed on the host in
// block order, so the totals are deterministic)
__device__ inline void block_sum2(double a, double b, double *__restrict__ part)
{
    __shared__ double sa[256], sb[256];
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) { sa[threadIdx.x] += sa[threadIdx.x + k]; sb[threadIdx.x] += sb[threadIdx.x + k]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = sa[0]; part[2 * blockIdx.x + 1] = sb[0]; }
}

// Energy bookkeeping per (element, point), 256 threads per block: a rule-4
// point adds its linearized gray emission W4_0 + W4_1 (phi4_q . (Traw - T*)) (the
// Jacobian weights of the pass), a rule-3 point the emission form at the new
// temperature, w_q sum_g sigma_g b_g(T_q) T_q^4 (the basis sums to one)
__global__ void __launch_bounds__(256) k_energy_points(TrtParams P, const double *__restrict__ T,
                                                       const double *__restrict__ Traw,
                                                       const double *__restrict__ Tstar,
                                                       const double *__restrict__ W4, double *__restrict__ part)
{
    const int nq = P.nq4 + P.nq3;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double gray = 0.0, emis = 0.0;
    if (t < (long long)P.Ne * nq) {
        const int e = (int)(t / nq), q = (int)(t - (long long)e * nq);
        double tv[N];
#pragma unroll
        for (int i = 0; i < N; ++i) tv[i] = T[(long long)e * N + i];
        if (q < P.nq4) {
            double dv[N];
#pragma unroll
            for (int i = 0; i < N; ++i) dv[i] = Traw[(long long)e * N + i] - Tstar[(long long)e * N + i];
            const double *w = W4 + ((long long)e * P.nq4 + q) * 2;
            gray = fma(w[1], dot_basis(P.phi4 + q * N, dv), w[0]);
        } else {
            const int r = q - P.nq4;
            const double Tq = dot_basis(P.phi3 + r * N, tv);
            const double T2 = Tq * Tq;
            double sp = 0.0;
            const double *sig = P.sigma + e;
            allium::planck_spectrum_each<0>(P.bounds, P.G, Tq, [&](int g, double v) { sp = fma(sig[(long long)g * P.Ne], v, sp); });
            emis = P.w3[r] * sp * T2 * T2;
        }
    }
    block_sum2(gray, emis, part);
}

// sum_g sigma_g int phi_g (MG mass row sums rMu) and int sum_g phi_g (time mass
// row sums rMm) over a grid-stride share, 256 threads per block
__global__ void __launch_bounds__(256) k_phi_sums(TrtParams P, const double *__restrict__ rMu,
                                                  const double *__restrict__ rMm, const double *__restrict__ phi,
                                                  double *__restrict__ part)
{
    double a = 0.0, b = 0.0;
    const long long n = (long long)P.G * P.Ne * N;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(i % N);
        const long long ge = i / N;   // g * Ne + e
        a = fma(P.sigma[ge] * rMu[k], phi[i], a);
        b = fma(rMm[k], phi[i], b);
    }
    block_sum2(a, b, part);
}

// part[blk] = sum of (a[i] - b[i]) w[i % N] over a grid-stride share
__global__ void k_wdiff(long long n, const double *__restrict__ a, const double *__restrict__ b,
                        const double *__restrict__ w, double *__restrict__ part)
{
    double acc = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        acc = fma(a[i] - b[i], w[i % N], acc);
    block_sum2(acc, 0.0, part);
}

thread_local std::string g_err;
int tfail(int code, const std::string &m)
{
    allium_cnn_internal::set_last_error(m);
    return code;
}
#define TRT_TRY(expr)                                                                               \
    do {                                                                                            \
        hipError_t _e = (expr);                                                                     \
        if (_e != hipSuccess) return tfail(ALLIUM_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)
#define ABI_TRY(expr)                \
    do {                             \
        int _rc = (expr);            \
        if (_rc) return _rc;         \
    } while (0)

}  // namespace

struct allium_trt {
    allium_sweep_t L = nullptr;
    allium_trt_desc d{};
    int G = 0, No = 0, Ne = 0, nel[3] = {0, 0, 0};
    long Ndof = 0, psi_size = 0, phi_size = 0;
    double wsum = 0.0, time = 0.0;
    int step = 0;
    std::vector<double> bounds, mid, sigma_h;
    TrtParams P{};
    std::vector<void *> owned;
    double *Tk = nullptr;   // temperature before the current outer pass
    double *T = nullptr, *Tn = nullptr, *psi = nullptr, *qpsi = nullptr;
    long long qin_nnz = 0;                  // inflow source as its nonzero entries
    long long *qin_idx = nullptr;
    double *qin_val = nullptr;
    double *volw = nullptr;   // element-mass row sums: int phi = sum_k phi[k] volw[k % N]
    double *sigma = nullptr, *em = nullptr, *phi = nullptr, *phisrc = nullptr, *phit = nullptr, *iso = nullptr;
    double *resid = nullptr, *Binv = nullptr, *dpw = nullptr, *part = nullptr, *dmid = nullptr;
    double *W4 = nullptr, *E3 = nullptr;   // Planck weights per quadrature point (k_planck_points)
    double *UQ = nullptr;                  // Binv-weighted source at the rule-3 points (k_dB_points)
    int *flags = nullptr;
    // energy bookkeeping (allium_trt_track_energy)
    bool track = false;
    double *rMu = nullptr, *rMm = nullptr, *ones = nullptr, *epart = nullptr, *Traw = nullptr;
    double inflow_rate = 0.0, En = 0.0;   // sum D q_in; int sum_g phi^n_g
    allium_trt_energy energy{};
    hipStream_t s = nullptr;
    ~allium_trt()
    {
        if (L) allium_sweep_destroy(L);
        for (void *p : owned) hipFree(p);
        if (s) hipStreamDestroy(s);
    }
    template <typename T>
    hipError_t alloc(T **p, size_t count, const T *src = nullptr)
    {
        hipError_t e = hipMalloc((void **)p, sizeof(T) * std::max<size_t>(count, 1));
        if (e != hipSuccess) return e;
        owned.push_back(*p);
        if (src) return hipMemcpy(*p, src, sizeof(T) * count, hipMemcpyHostToDevice);
        return hipSuccess;
    }
};

namespace {

constexpr int kPart = 1024;

int update_opacity(allium_trt *h)
{
    hipLaunchKernelGGL(k_opacity, dim3((h->Ne + 255) / 256), dim3(256), 0, h->s, h->Ne, h->G, h->T, h->dmid,
                       h->d.opacity_coef, h->sigma);
    TRT_TRY(hipGetLastError());
    TRT_TRY(hipMemcpyAsync(h->sigma_h.data(), h->sigma, sizeof(double) * h->sigma_h.size(), hipMemcpyDeviceToHost,
                           h->s));
    TRT_TRY(hipStreamSynchronize(h->s));
    for (double v : h->sigma_h)
        if (!std::isfinite(v)) return tfail(ALLIUM_ERR_NUMERIC, "non-finite opacity");
    ABI_TRY(allium_sweep_set_sigma(h->L, h->sigma_h.data()));
    return ALLIUM_OK;
}

}  // namespace

extern "C" {

int allium_planck_integrals(double x, double *planck, double *rosseland)
{
    if (!(x >= 0.0)) return tfail(ALLIUM_ERR_ARG, "Planck integral needs x >= 0");
    if (planck) *planck = allium::planck_integral(x);
    if (rosseland) *rosseland = allium::rosseland_integral(x);
    return ALLIUM_OK;
}

int allium_log_groups(double emin, double emax, int G, double *bounds, double *midpoints)
{
    if (!(emin > 0.0) || !(emax > emin) || G < 1) return tfail(ALLIUM_ERR_ARG, "bad energy grid");
    // MultiGroupEnergyGrid::MakeLogSpaced (src/multigroup.cpp:40-59)
    const double lo = std::log(emin), d = (std::log(emax) - lo) / G;
    std::vector<double> b(G + 1);
    b[0] = emin;
    for (int g = 1; g <= G; g++) b[g] = std::exp(lo + g * d);
    if (bounds)
        for (int g = 0; g <= G; g++) bounds[g] = b[g];
    if (midpoints)
        for (int g = 0; g < G; g++) midpoints[g] = std::sqrt(b[g] * b[g + 1]);
    return ALLIUM_OK;
}

int allium_trt_create(const allium_mesh_desc *mesh, int fe_order, const allium_quad_desc *quad,
                      const allium_trt_desc *desc, allium_trt_t *out)
{
    if (!mesh || !quad || !desc || !out) return tfail(ALLIUM_ERR_ARG, "null argument");
    *out = nullptr;
    if (mesh->dim != 3 || fe_order != 1)
        return tfail(ALLIUM_ERR_UNSUPPORTED, "the TRT solver implements 3D meshes with p = 1");
    const allium_trt_desc &d = *desc;
    if (d.groups < 1 || d.groups > GMAX || !(d.dt > 0) || !(d.cv > 0) || !(d.T0 > 0) || !(d.T_inflow >= 0) ||
        !(d.opacity_coef > 0) || d.si_maxit < 1 || !(d.si_reltol >= 0) || !(d.si_abstol >= 0) ||
        !(d.si_reltol > 0 || d.si_abstol > 0) || d.outer_maxit < 0 || !(d.outer_reltol >= 0) ||
        !(d.outer_abstol >= 0) || (d.outer_maxit > 1 && !(d.outer_reltol > 0 || d.outer_abstol > 0)))
        return tfail(ALLIUM_ERR_ARG, "bad TRT description");
    if (d.implicit_opacity && d.outer_maxit <= 1)
        return tfail(ALLIUM_ERR_ARG, "implicit opacity needs the outer fixed-point iteration (outer_maxit > 1), "
                                     "as in exe/scallion.cpp:986");
    auto h = std::make_unique<allium_trt>();
    h->d = d;
    h->G = d.groups;
    std::string why;
    allium::Mesh m;
    if (!allium::build_mesh(3, mesh->nel, mesh->vertices, mesh->extents, mesh->bc, m, why))
        return tfail(ALLIUM_ERR_ARG, why);
    if (!m.uniform) return tfail(ALLIUM_ERR_UNSUPPORTED, "the TRT solver needs uniform geometry");
    for (int s = 0; s < 6; s++)
        if (mesh->bc[s] != ALLIUM_BC_INFLOW)
            return tfail(ALLIUM_ERR_UNSUPPORTED, "the TRT solver takes inflow/vacuum boundaries");
    h->Ne = m.Ne;
    for (int k = 0; k < 3; k++) h->nel[k] = m.nel[k];
    h->bounds.resize(h->G + 1);
    h->mid.resize(h->G);
    ABI_TRY(allium_log_groups(d.emin, d.emax, h->G, h->bounds.data(), h->mid.data()));
    // sweep handle with the initial opacities (element-centre T0)
    std::vector<double> sig0((size_t)h->G * h->Ne);
    for (int g = 0; g < h->G; g++)
        for (int e = 0; e < h->Ne; e++)
            sig0[(size_t)g * h->Ne + e] = d.opacity_coef / std::sqrt(d.T0) / (h->mid[g] * h->mid[g] * h->mid[g]);
    ABI_TRY(allium_sweep_create(mesh, 1, quad, h->G, sig0.data(), 0, ALLIUM_FP64, &h->L));
    ABI_TRY(allium_sweep_set_time_absorption(h->L, 1.0 / d.dt));
    allium_sweep_info_t info;
    ABI_TRY(allium_sweep_info(h->L, &info));
    h->No = (int)info.Nomega;
    h->Ndof = info.Ndof;
    h->psi_size = info.psi_size;
    h->phi_size = info.phi_size;
    std::vector<double> om((size_t)h->No * 3), w(h->No);
    ABI_TRY(allium_sweep_quadrature(h->L, om.data(), w.data()));
    for (double v : w) h->wsum += v;
    // element matrices and rules
    std::vector<double> Mu(N * N), Gd(3 * N * N), Mm(N * N);
    ABI_TRY(allium_host_element_matrices(mesh, 1, 0, Mu.data(), Gd.data(), Mm.data()));
    std::vector<double> phi3, w3, phi4, w4;
    allium::uniform_element_rule(m, 1, 2 * 1 + 1, phi3, w3);   // PlanckEmissionNFI default order 2p + 1
    allium::uniform_element_rule(m, 1, 2 * 1 + 2, phi4, w4);   // GrayPlanckEmissionNFI(.., 2, 2 + 0)
    TRT_TRY(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
    TrtParams &P = h->P;
    P.Ne = h->Ne;
    P.G = h->G;
    P.nq3 = (int)w3.size();
    P.nq4 = (int)w4.size();
    P.cvdt = d.cv / d.dt;
    double *dp;
    TRT_TRY(h->alloc(&dp, phi3.size(), phi3.data())); P.phi3 = dp;
    TRT_TRY(h->alloc(&dp, w3.size(), w3.data())); P.w3 = dp;
    TRT_TRY(h->alloc(&dp, phi4.size(), phi4.data())); P.phi4 = dp;
    TRT_TRY(h->alloc(&dp, w4.size(), w4.data())); P.w4 = dp;
    TRT_TRY(h->alloc(&dp, Mm.size(), Mm.data())); P.Mm = dp;
    TRT_TRY(h->alloc(&dp, Mu.size(), Mu.data())); P.Mu = dp;
    {
        std::vector<double> vw(N, 0.0);
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) vw[i] += Mu[(size_t)i * N + j];
        TRT_TRY(h->alloc(&h->volw, vw.size(), vw.data()));
    }
    {
        std::vector<double> ru(N, 0.0), rm(N, 0.0), one(N, 1.0);
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) { ru[i] += Mu[(size_t)i * N + j]; rm[i] += Mm[(size_t)i * N + j]; }
        TRT_TRY(h->alloc(&h->rMu, ru.size(), ru.data()));
        TRT_TRY(h->alloc(&h->rMm, rm.size(), rm.data()));
        TRT_TRY(h->alloc(&h->ones, one.size(), one.data()));
    }
    TRT_TRY(h->alloc(&dp, h->bounds.size(), h->bounds.data())); P.bounds = dp;
    TRT_TRY(h->alloc(&h->dmid, h->mid.size(), h->mid.data()));
    TRT_TRY(h->alloc(&h->sigma, sig0.size(), sig0.data()));
    P.sigma = h->sigma;
    h->sigma_h = sig0;
    TRT_TRY(h->alloc(&h->T, h->Ndof));
    TRT_TRY(h->alloc(&h->Tn, h->Ndof));
    TRT_TRY(h->alloc(&h->Tk, h->Ndof));
    TRT_TRY(h->alloc(&h->Traw, h->Ndof));
    TRT_TRY(h->alloc(&h->resid, h->Ndof));
    TRT_TRY(h->alloc(&h->Binv, (size_t)h->Ne * N * N));
    TRT_TRY(h->alloc(&h->dpw, (size_t)h->Ne * P.nq3 * h->G));
    TRT_TRY(h->alloc(&h->E3, (size_t)h->Ne * P.nq3 * h->G));
    TRT_TRY(h->alloc(&h->W4, (size_t)h->Ne * P.nq4 * 2));
    TRT_TRY(h->alloc(&h->UQ, (size_t)h->Ne * P.nq3));
    TRT_TRY(h->alloc(&h->psi, h->psi_size));
    TRT_TRY(h->alloc(&h->qpsi, h->psi_size));
    TRT_TRY(h->alloc(&h->em, h->phi_size));
    TRT_TRY(h->alloc(&h->phi, h->phi_size));
    TRT_TRY(h->alloc(&h->phisrc, h->phi_size));
    TRT_TRY(h->alloc(&h->phit, h->phi_size));
    TRT_TRY(h->alloc(&h->iso, h->phi_size));
    TRT_TRY(h->alloc(&h->part, 3 * kPart));
    TRT_TRY(h->alloc(&h->epart, 2 * (((size_t)h->Ne * (P.nq4 + P.nq3) + 255) / 256)));
    TRT_TRY(h->alloc(&h->flags, 4));
    // initial state: T = T0, psi in equilibrium with T0 (PlanckEmissionPSCoefficient,
    // src/phase_coefficient.cpp:72-76: the Planck fraction between the group bounds)
    std::vector<double> Th(h->Ndof, d.T0);
    TRT_TRY(hipMemcpy(h->T, Th.data(), sizeof(double) * h->Ndof, hipMemcpyHostToDevice));
    std::vector<double> psi0(h->psi_size);
    const double fourpi = 4.0 * M_PI;
    for (int g = 0; g < h->G; g++) {
        const double b = allium::planck_integral(h->bounds[g + 1] / d.T0) - allium::planck_integral(h->bounds[g] / d.T0);
        const double v = b * std::pow(d.T0, 4) / fourpi;
        std::fill(psi0.begin() + (size_t)g * h->No * h->Ndof, psi0.begin() + (size_t)(g + 1) * h->No * h->Ndof, v);
    }
    TRT_TRY(hipMemcpy(h->psi, psi0.data(), sizeof(double) * h->psi_size, hipMemcpyHostToDevice));
    // inflow source: Planckian at T_inflow on the flagged sides, vacuum elsewhere
    std::vector<double> inflow((size_t)h->G * 6, 0.0);
    for (int g = 0; g < h->G; g++) {
        const double Ti = d.T_inflow;
        const double b = Ti > 0 ? allium::planck_integral(h->bounds[g + 1] / Ti) - allium::planck_integral(h->bounds[g] / Ti)
                                : 0.0;
        for (int s = 0; s < 6; s++)
            if ((d.inflow_sides >> s) & 1) inflow[(size_t)g * 6 + s] = b * std::pow(Ti, 4) / fourpi;
    }
    ABI_TRY(allium_form_source(h->L, nullptr, inflow.data(), psi0.data()));
    // sum D q_in: the inflow rate of the energy bookkeeping
    for (long long k = 0; k < h->psi_size; k++)
        if (psi0[k] != 0.0) h->inflow_rate += w[(k / h->Ndof) % h->No] * psi0[k];
    {
        std::vector<long long> idx;
        std::vector<double> val;
        for (long long k = 0; k < h->psi_size; k++)
            if (psi0[k] != 0.0) { idx.push_back(k); val.push_back(psi0[k]); }
        h->qin_nnz = (long long)idx.size();
        if (h->qin_nnz) {
            TRT_TRY(h->alloc(&h->qin_idx, idx.size(), idx.data()));
            TRT_TRY(h->alloc(&h->qin_val, val.size(), val.data()));
        }
    }
    ABI_TRY(update_opacity(h.get()));
    *out = h.release();
    return ALLIUM_OK;
}

void allium_trt_destroy(allium_trt_t h) { delete h; }

allium_sweep_t allium_trt_sweep(allium_trt_t h) { return h ? h->L : nullptr; }

namespace {
// out (+)= scale dB Binv (src): k_dB_points then k_dB_groups on the handle's stream
void apply_dB(allium_trt *h, int mode, const double *x, double scale, int accumulate, double *out)
{
    const TrtParams &P = h->P;
    hipLaunchKernelGGL(k_dB_points, dim3((h->Ne + 63) / 64), dim3(64), 0, h->s, P, mode, x, h->Binv, h->UQ);
    const long long items = (long long)h->Ne * P.G;
    hipLaunchKernelGGL(k_dB_groups, dim3((unsigned)((items + 255) / 256)), dim3(256), 0, h->s, P, h->UQ, h->dpw,
                       scale, accumulate, out);
}
// both block sums of a part buffer (nb blocks), in block order
int sum_parts(allium_trt *h, int nb, double &a, double &b)
{
    std::vector<double> hp(2 * (size_t)nb);
    TRT_TRY(hipMemcpyAsync(hp.data(), h->part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, h->s));
    TRT_TRY(hipStreamSynchronize(h->s));
    a = b = 0.0;
    for (int k = 0; k < nb; k++) { a += hp[2 * k]; b += hp[2 * k + 1]; }
    return ALLIUM_OK;
}

// (sum_g sigma_g int phi_g, int sum_g phi_g) of x (phi-sized)
int phi_sums(allium_trt *h, const double *x, double &absorbed, double &energy)
{
    hipLaunchKernelGGL(k_phi_sums, dim3(kPart / 2), dim3(256), 0, h->s, h->P, h->rMu, h->rMm, x, h->part);
    TRT_TRY(hipGetLastError());
    return sum_parts(h, kPart / 2, absorbed, energy);
}

// the energy bookkeeping of the pass that just ended (allium_trt_energy):
// T is the new temperature, Tstar the linearization point, the Jacobian
// weights W4 and the emission em those of the pass; phit is free
int energy_terms(allium_trt *h, const double *Tstar)
{
    allium_trt_energy &E = h->energy;
    const double tau = 1.0 / h->d.dt;
    double En1 = 0.0, unused = 0.0;
    ABI_TRY(allium_moments_device(h->L, h->psi, h->phit, h->s));
    ABI_TRY(phi_sums(h, h->phit, E.absorption, En1));
    E.time_absorption = tau * En1;
    ABI_TRY(phi_sums(h, h->phi, E.absorption_si, unused));
    double lb = 0.0;
    ABI_TRY(allium_balance_device(h->L, h->psi, &lb, h->s));
    E.leakage = lb - E.absorption - E.time_absorption;
    E.inflow = h->inflow_rate;
    E.source = h->inflow_rate + tau * h->En;
    hipLaunchKernelGGL(k_wdiff, dim3(kPart / 2), dim3(256), 0, h->s, (long long)h->Ndof, h->T, h->Tn, h->rMm, h->part);
    TRT_TRY(hipGetLastError());
    ABI_TRY(sum_parts(h, kPart / 2, E.material, unused));
    E.material *= h->P.cvdt;
    hipLaunchKernelGGL(k_wdiff, dim3(kPart / 2), dim3(256), 0, h->s, (long long)h->Ndof, h->T, h->Traw, h->rMm, h->part);
    TRT_TRY(hipGetLastError());
    ABI_TRY(sum_parts(h, kPart / 2, E.floor, unused));
    E.floor *= h->P.cvdt;
    hipLaunchKernelGGL(k_wsum, dim3(kPart), dim3(256), 0, h->s, (long long)h->phi_size, h->em, h->ones, h->part);
    TRT_TRY(hipGetLastError());
    {
        std::vector<double> hp(kPart);
        TRT_TRY(hipMemcpyAsync(hp.data(), h->part, sizeof(double) * kPart, hipMemcpyDeviceToHost, h->s));
        TRT_TRY(hipStreamSynchronize(h->s));
        E.emission_linear = 0.0;
        for (double v : hp) E.emission_linear += v;
    }
    const long long pts = (long long)h->Ne * (h->P.nq4 + h->P.nq3);
    const int nb = (int)((pts + 255) / 256);
    hipLaunchKernelGGL(k_energy_points, dim3(nb), dim3(256), 0, h->s, h->P, h->T, h->Traw, Tstar, h->W4, h->epart);
    TRT_TRY(hipGetLastError());
    {
        std::vector<double> hp(2 * (size_t)nb);
        TRT_TRY(hipMemcpyAsync(hp.data(), h->epart, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, h->s));
        TRT_TRY(hipStreamSynchronize(h->s));
        E.gray_linear = E.emission = 0.0;
        for (int k = 0; k < nb; k++) { E.gray_linear += hp[2 * k]; E.emission += hp[2 * k + 1]; }
    }
    const double rhs = E.emission + E.source;
    E.balance = std::fabs(lb - rhs) / rhs;
    return ALLIUM_OK;
}

}  // namespace

int allium_trt_step(allium_trt_t h, allium_trt_info *info)
{
    if (!h) return tfail(ALLIUM_ERR_ARG, "null handle");
    hipStream_t s = h->s;
    const TrtParams &P = h->P;
    const int eb = (h->Ne + 63) / 64;
    // previous time level
    TRT_TRY(hipMemcpyAsync(h->Tn, h->T, sizeof(double) * h->Ndof, hipMemcpyDeviceToDevice, s));
    if (h->track) {
        // int sum_g phi^n_g of D psi^n, the radiation energy the time source carries
        double unused = 0.0;
        ABI_TRY(allium_moments_device(h->L, h->psi, h->phit, s));
        ABI_TRY(phi_sums(h, h->phit, unused, h->En));
    }
    // the time source reads psi^n before the first sweep of the step overwrites it
    const long long cells = h->psi_size / N;
    hipLaunchKernelGGL(k_time_source, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, cells, 1.0 / h->d.dt,
                       P.Mm, h->psi, h->qpsi);
    if (h->qin_nnz)
        hipLaunchKernelGGL(k_time_source_inflow, dim3((unsigned)((h->qin_nnz + 255) / 256)), dim3(256), 0, s,
                           h->qin_nnz, 1.0 / h->d.dt, P.Mm, h->qin_idx, h->qin_val, h->psi, h->qpsi);
    TRT_TRY(hipMemsetAsync(h->flags, 0, sizeof(int) * 4, s));
    // LinearizedTRTOperator::Mult turns a fixup off for the transport solve and
    // back on for the final sweep (src/trt_linearized.cpp:236-238,284-289)
    int fix_type = 0, fix_on = 0;
    double fix_min = 0.0;
    ABI_TRY(allium_sweep_get_fixup(h->L, &fix_type, &fix_min, &fix_on));
    const double inv_w = 1.0 / h->wsum;
    const int nb = allium::si_update_blocks(h->phi_size);
    std::vector<double> hp(nb);
    const int outer_max = std::max(1, h->d.outer_maxit);
    double nrm = 0.0, onrm = 0.0, or0 = 0.0;
    int it = 0, k = 0;
    // FixedPointIterationSolver on the temperature (src/fixed_point.cpp:134-172,
    // ComponentReductionOperator on T, exe/scallion.cpp:830): each pass is one
    // LinearizedTRTOperator::Mult linearized at the current T with the sources
    // of the previous time level; one pass without an outer solver
    for (k = 1;; k++) {
        if (outer_max > 1 || h->track)
            TRT_TRY(hipMemcpyAsync(h->Tk, h->T, sizeof(double) * h->Ndof, hipMemcpyDeviceToDevice, s));
        if (fix_on) ABI_TRY(allium_sweep_use_fixup(h->L, 0));
        // linearize at the current T
        {
            const long long pts = (long long)h->Ne * (P.nq4 + P.nq3);
            hipLaunchKernelGGL(k_planck_points, dim3((unsigned)((pts + 255) / 256)), dim3(256), 0, s, P, h->T, h->W4,
                               h->E3, h->dpw);
        }
        hipLaunchKernelGGL(k_linearize, dim3(eb), dim3(64), 0, s, P, h->T, h->Tn, h->W4, h->E3, h->resid, h->Binv,
                           h->em, h->flags);
        // em += dB Bt'^{-1} resid
        apply_dB(h, 0, h->resid, 1.0, 1, h->em);
        TRT_TRY(hipGetLastError());
        int bad = 0;
        TRT_TRY(hipMemcpyAsync(&bad, h->flags, sizeof(int), hipMemcpyDeviceToHost, s));
        TRT_TRY(hipStreamSynchronize(s));
        if (bad) return tfail(ALLIUM_ERR_NUMERIC, "singular energy-balance Jacobian");
        // phi_src = D Linv (D^T em + q_psi)
        TRT_TRY(allium::launch_axpby(h->phi_size, inv_w, h->em, 0.0, h->em, h->iso, s));
        ABI_TRY(allium_sweep_mult_moments_device(h->L, h->qpsi, h->iso, h->psi, h->phisrc, s));
        // source iteration on (I - D Linv D^T Ms) phi = phi_src (SLISolver, src/linalg.cpp:103-153);
        // a warm start begins from D psi of the current iterate
        // (in the first pass psi is still psi^n)
        if (h->d.si_warm_start) ABI_TRY(allium_moments_device(h->L, h->psi, h->phi, s));
        else TRT_TRY(hipMemsetAsync(h->phi, 0, sizeof(double) * h->phi_size, s));
        double r0 = 0.0;
        for (it = 1;; it++) {
            apply_dB(h, 1, h->phi, inv_w, 0, h->iso);
            TRT_TRY(hipGetLastError());
            ABI_TRY(allium_sweep_mult_moments_device(h->L, nullptr, h->iso, h->psi, h->phit, s));
            TRT_TRY(allium::launch_si_update(h->phi_size, h->phisrc, h->phit, h->phi, h->part, s));
            TRT_TRY(hipMemcpyAsync(hp.data(), h->part, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
            TRT_TRY(hipStreamSynchronize(s));
            nrm = 0.0;
            for (double v : hp) nrm += v;
            nrm = std::sqrt(nrm);
            if (!std::isfinite(nrm)) return tfail(ALLIUM_ERR_NUMERIC, "non-finite transport residual");
            if (it == 1) r0 = std::max(nrm * h->d.si_reltol, h->d.si_abstol);
            if (nrm < r0 || it >= h->d.si_maxit) break;
        }
        // em += Ms phi ; psi = Linv(D^T em + q_psi), with the fixup back on
        apply_dB(h, 1, h->phi, 1.0, 1, h->em);
        TRT_TRY(allium::launch_axpby(h->phi_size, inv_w, h->em, 0.0, h->em, h->iso, s));
        if (fix_on) ABI_TRY(allium_sweep_use_fixup(h->L, 1));
        ABI_TRY(allium_sweep_mult_iso_device(h->L, h->qpsi, h->iso, h->psi, s));
        // temperature update
        hipLaunchKernelGGL(k_update_T, dim3((h->Ne + 255) / 256), dim3(256), 0, s, P, h->phi, h->resid, h->Binv,
                           h->T, h->flags + 1, h->track ? h->Traw : nullptr);
        TRT_TRY(hipGetLastError());
        // bookkeeping with the operator of this pass (before an implicit opacity update)
        if (h->track) ABI_TRY(energy_terms(h, h->Tk));
        // implicit opacity: OpacityUpdate::Update after every pass (src/trt_linearized.cpp:311)
        if (h->d.implicit_opacity) ABI_TRY(update_opacity(h));
        if (outer_max == 1) break;
        hipLaunchKernelGGL(k_diff2, dim3(kPart), dim3(256), 0, s, (long long)h->Ndof, h->T, h->Tk, h->part);
        TRT_TRY(hipGetLastError());
        std::vector<double> dp(kPart);
        TRT_TRY(hipMemcpyAsync(dp.data(), h->part, sizeof(double) * kPart, hipMemcpyDeviceToHost, s));
        TRT_TRY(hipStreamSynchronize(s));
        onrm = 0.0;
        for (double v : dp) onrm += v;
        onrm = std::sqrt(onrm);
        if (!std::isfinite(onrm)) return tfail(ALLIUM_ERR_NUMERIC, "non-finite temperature change");
        if (k == 1) or0 = std::max(h->d.outer_abstol, onrm * h->d.outer_reltol);
        if (onrm < or0 || k >= outer_max) break;
    }
    h->step++;
    h->time += h->d.dt;
    // explicit opacity for the next step
    if (!h->d.implicit_opacity) ABI_TRY(update_opacity(h));
    if (info) {
        int fl = 0;
        TRT_TRY(hipMemcpyAsync(&fl, h->flags + 1, sizeof(int), hipMemcpyDeviceToHost, s));
        hipLaunchKernelGGL(k_minmax, dim3(kPart), dim3(256), 0, s, (long long)h->Ndof, h->T, h->part);
        // sum_g int phi_g = sum over dofs of phi times the element-mass row sums
        hipLaunchKernelGGL(k_wsum, dim3(kPart), dim3(256), 0, s, (long long)h->phi_size, h->phi, h->volw,
                           h->part + 2 * kPart);
        TRT_TRY(hipGetLastError());
        std::vector<double> mm(3 * kPart);
        TRT_TRY(hipMemcpyAsync(mm.data(), h->part, sizeof(double) * mm.size(), hipMemcpyDeviceToHost, s));
        TRT_TRY(hipStreamSynchronize(s));
        info->step = h->step;
        info->time = h->time;
        info->transport_iterations = it;
        info->transport_residual = nrm;
        info->outer_iterations = k;
        info->outer_residual = onrm;
        info->floored = fl;
        info->T_min = 1e300;
        info->T_max = -1e300;
        for (int k = 0; k < kPart; k++) {
            info->T_min = std::min(info->T_min, mm[2 * k]);
            info->T_max = std::max(info->T_max, mm[2 * k + 1]);
        }
        double E = 0.0;
        for (int k = 0; k < kPart; k++) E += mm[2 * kPart + k];
        info->radiation_energy = E;
        ABI_TRY(allium_sweep_timing(h->L, &info->sweep_seconds, &info->sweeps));
    }
    return ALLIUM_OK;
}

int allium_trt_track_energy(allium_trt_t h, int on)
{
    if (!h) return tfail(ALLIUM_ERR_ARG, "null handle");
    h->track = on != 0;
    h->energy = allium_trt_energy{};
    return ALLIUM_OK;
}

int allium_trt_set_time_step(allium_trt_t h, double dt)
{
    if (!h) return tfail(ALLIUM_ERR_ARG, "null handle");
    if (!(dt > 0) || !std::isfinite(dt)) return tfail(ALLIUM_ERR_ARG, "the time step must be positive");
    // the time source, c_v/dt and the sweep's time absorption read it from the next step on
    ABI_TRY(allium_sweep_set_time_absorption(h->L, 1.0 / dt));
    h->d.dt = dt;
    h->P.cvdt = h->d.cv / dt;
    return ALLIUM_OK;
}

int allium_trt_energy_terms(allium_trt_t h, allium_trt_energy *out)
{
    if (!h || !out) return tfail(ALLIUM_ERR_ARG, "null argument");
    if (!h->track) return tfail(ALLIUM_ERR_ARG, "energy bookkeeping is off (allium_trt_track_energy)");
    *out = h->energy;
    return ALLIUM_OK;
}

int allium_trt_state_size(allium_trt_t h, long *size)
{
    if (!h || !size) return tfail(ALLIUM_ERR_ARG, "null argument");
    *size = h->psi_size + h->Ndof;
    return ALLIUM_OK;
}

int allium_trt_get_state(allium_trt_t h, double *x, int *cycle, double *time)
{
    if (!h || !x) return tfail(ALLIUM_ERR_ARG, "null argument");
    TRT_TRY(hipMemcpyAsync(x, h->psi, sizeof(double) * h->psi_size, hipMemcpyDeviceToHost, h->s));
    TRT_TRY(hipMemcpyAsync(x + h->psi_size, h->T, sizeof(double) * h->Ndof, hipMemcpyDeviceToHost, h->s));
    TRT_TRY(hipStreamSynchronize(h->s));
    if (cycle) *cycle = h->step;
    if (time) *time = h->time;
    return ALLIUM_OK;
}

int allium_trt_set_state(allium_trt_t h, const double *x, int cycle, double time)
{
    if (!h || !x) return tfail(ALLIUM_ERR_ARG, "null argument");
    if (cycle < 0 || !std::isfinite(time)) return tfail(ALLIUM_ERR_ARG, "bad cycle or time");
    for (long k = 0; k < h->psi_size + h->Ndof; k++)
        if (!std::isfinite(x[k])) return tfail(ALLIUM_ERR_NUMERIC, "non-finite state");
    for (long k = 0; k < h->Ndof; k++)
        if (!(x[h->psi_size + k] > 0.0)) return tfail(ALLIUM_ERR_NUMERIC, "temperature must be positive");
    TRT_TRY(hipMemcpyAsync(h->psi, x, sizeof(double) * h->psi_size, hipMemcpyHostToDevice, h->s));
    TRT_TRY(hipMemcpyAsync(h->T, x + h->psi_size, sizeof(double) * h->Ndof, hipMemcpyHostToDevice, h->s));
    // what the state determines: the scalar flux and the opacities of the next step
    ABI_TRY(allium_moments_device(h->L, h->psi, h->phi, h->s));
    ABI_TRY(update_opacity(h));
    h->step = cycle;
    h->time = time;
    return ALLIUM_OK;
}

int allium_trt_temperature(allium_trt_t h, double *T)
{
    if (!h || !T) return tfail(ALLIUM_ERR_ARG, "null argument");
    TRT_TRY(hipMemcpy(T, h->T, sizeof(double) * h->Ndof, hipMemcpyDeviceToHost));
    return ALLIUM_OK;
}

int allium_trt_scalar_flux(allium_trt_t h, double *phi)
{
    if (!h || !phi) return tfail(ALLIUM_ERR_ARG, "null argument");
    TRT_TRY(hipMemcpy(phi, h->phi, sizeof(double) * h->phi_size, hipMemcpyDeviceToHost));
    return ALLIUM_OK;
}

}  // extern "C"
