// HIP kernels of the DG upwind sweep for CDNA4 (gfx950).
//
// One work item is one cell (element e, direction a, group g): it assembles
// the local DG system  A = sum_d Omega_d G_d + sum_{outflow} (Omega.n) F + sigma_g M [+ tau Mt],
// gathers the upwind face traces from HBM into the right-hand side and solves
// with an unrolled register LU (src/sweep.cpp:366-475).  Cells of one
// topological level are independent; the launchers below run the levels as
// separate launches, as one captured hipGraph, or inside one persistent kernel
// whose workgroups take tiles in level order from an atomic ticket and wait on
// per-(level, direction) completion counters.
#include "device.hpp"
#include "kernel_common.hpp"

#include <algorithm>
#include <cstdio>
#include <mutex>

namespace allium {

// One cell (e, a, g).  Every global load (source, isotropic source, upwind
// traces, general-geometry matrices) is issued before any of them is used and
// without data-dependent branches, so a wave makes one memory round trip
// before its register LU.  OCT >= 0 fixes the direction's octant at compile
// time (uniform geometry, direction-uniform waves): only the dim inflow sides
// are gathered.  OCT < 0 decides inflow per lane with selects.
// Negative-flux fixup of one solved cell (src/fixup.cpp), given the column
// sums wcol of the assembled matrix and the sum of its right-hand side:
// 1 ZeroFixupOperator (clip at fix_min), 2 ZeroAndScaleFixupOperator (clip,
// then scale to the balance sum(rhs) = wcol . x; failed if either sum is
// negative), 3 RyosukeFixupOperator (clip, then solve for the positive
// unknowns with their ratios kept and the balance restored; the reference
// aborts beyond 3 positive unknowns, here a flag is raised).  Applications
// and failures are counted (EventLog "fixup applied" / "fixup failed") and,
// with a monitor, per element (SetFixupMonitorData).
template <int N>
__device__ __forceinline__ void fixup_cell(const SweepParams &p, const double (&wcol)[N], double bsum, double (&x)[N],
                                           int e, int g)
{
    const double mn = p.fix_min;
    int applied = 0;
    if (p.fixup == 1) {
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (x[i] <= mn) { x[i] = mn; applied = 1; }
    } else if (p.fixup == 2) {
        double xmin = x[0];
#pragma unroll
        for (int i = 1; i < N; ++i) xmin = x[i] < xmin ? x[i] : xmin;
        if (xmin <= mn) {
            double now = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                if (x[i] <= mn) x[i] = mn;
                now += wcol[i] * x[i];
            }
            if (bsum < 0.0 || now < 0.0) {
                applied = -1;
            } else {
                const double sc = (bsum == 0.0 || now == 0.0) ? 0.0 : bsum / now;
#pragma unroll
                for (int i = 0; i < N; ++i) x[i] = sc == 0.0 ? 0.0 : x[i] * sc;
                applied = 1;
            }
        }
    } else {
        int pos = 0, c0 = 0, c1 = 0, c2 = 0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            if (x[i] <= mn) {
                x[i] = mn;
            } else {
                if (pos == 0) c0 = i;
                else if (pos == 1) c1 = i;
                else if (pos == 2) c2 = i;
                ++pos;
            }
        }
        if (pos > 0 && pos < N) {
            if (pos > 3) {
                atomicAdd(p.fix_counts + 2, 1ull);
                applied = -1;
            } else {
                double x0 = 0.0, x1 = 0.0, x2 = 0.0, w0 = 0.0, w1 = 0.0, w2 = 0.0;
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    x0 = i == c0 ? x[i] : x0; w0 = i == c0 ? wcol[i] : w0;
                    x1 = i == c1 ? x[i] : x1; w1 = i == c1 ? wcol[i] : w1;
                    x2 = i == c2 ? x[i] : x2; w2 = i == c2 ? wcol[i] : w2;
                }
                double y0, y1 = 0.0, y2 = 0.0;
                if (pos == 1) {
                    y0 = bsum / w0;
                } else if (pos == 2) {
                    const double r1 = -x1 / x0;
                    const double det = 1.0 / (w0 - w1 * r1);
                    y0 = bsum * det;
                    y1 = -bsum * r1 * det;
                } else {
                    const double r1 = -x1 / x0, r2 = -x2 / x0;
                    const double det = 1.0 / (w0 - w1 * r1 - w2 * r2);
                    y0 = bsum * det;
                    y1 = -bsum * r1 * det;
                    y2 = -bsum * r2 * det;
                }
                const bool fail = y0 < mn || (pos > 1 && y1 < mn) || (pos > 2 && y2 < mn);
                if (fail) { y0 = y1 = y2 = mn; }
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    x[i] = i == c0 ? y0 : x[i];
                    if (pos > 1) x[i] = i == c1 ? y1 : x[i];
                    if (pos > 2) x[i] = i == c2 ? y2 : x[i];
                }
                applied = fail ? -1 : 1;
            }
        }
    }
    if (applied == 1) atomicAdd(p.fix_counts, 1ull);
    else if (applied == -1) atomicAdd(p.fix_counts + 1, 1ull);
    if (applied != 0 && p.fix_monitor) atomicAdd(p.fix_monitor + (long long)g * p.Ne + e, 1.0);
}

template <int DIM, int P, bool UNI, typename T, int OCT, bool FIX = false>
__device__ __forceinline__ void cell_solve(const SweepParams &p, const int e, const int a, const int g)
{
    using S = Sh<DIM, P>;
    constexpr int N = S::N, NF = S::NF, NS = S::NS;
    int c[3];
    {
        const int q01 = (DIM == 3) ? fast_div(e, p.stride[2], p.inv_nel01) : 0;
        const int rem = e - q01 * p.stride[2];
        const int q0 = fast_div(rem, p.nel[0], p.inv_nel0);
        c[0] = rem - q0 * p.nel[0];
        c[1] = q0;
        c[2] = q01;
    }
    const cdptr omg = as_const(p.omega);
    const double om[3] = {omg[4 * a + 0], omg[4 * a + 1], omg[4 * a + 2]};
    const long long base = ((long long)g * p.Nomega + a) * p.Ndof;
    const long long off = base + (long long)e * N;
    const double *self = p.psi + off;

    // ---- phase 1: issue loads -------------------------------------------
    const double sig = p.sigma[(long long)g * p.Ne + e];
    double b[N];
    if (p.src) {
        if constexpr (N % 2 == 0) {
            const d2v *q2 = reinterpret_cast<const d2v *>(p.src + off);
#pragma unroll
            for (int i = 0; i < N / 2; ++i) { const d2v v = q2[i]; b[2 * i] = v.x; b[2 * i + 1] = v.y; }
        } else {
#pragma unroll
            for (int i = 0; i < N; ++i) b[i] = p.src[off + i];
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) b[i] = 0.0;
    }
    double isv[N];
    if (p.iso) {
        const double *is = p.iso + (long long)g * p.Ndof + (long long)e * N;
#pragma unroll
        for (int i = 0; i < N; ++i) isv[i] = is[i];
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) isv[i] = 0.0;
    }
    // upwind traces: u[s][m] and weight -dot (0 when side s is not an inflow side)
    double u[NS][NF], wgt[NS], dots[NS];
    int fid[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int ax = s >> 1, hi = s & 1;
        const bool interior = hi ? (c[ax] < p.nel[ax] - 1) : (c[ax] > 0);
        double dot;
        if constexpr (UNI) {
            dot = hi ? om[ax] : -om[ax];
            fid[s] = 0;
        } else {
            int cf[3] = {c[0], c[1], c[2]};
            cf[ax] += hi;
            const int f = p.foff[ax] + cf[0] + p.fext[ax][0] * (cf[1] + p.fext[ax][1] * cf[2]);
            fid[s] = f;
            dot = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) dot = fma(om[d], p.fnor[3 * f + d], dot);
            if (!hi && interior) dot = -dot;
        }
        dots[s] = dot;
        // compile-time octant: inflow sides are the low side of each positive axis
        bool inflow_known = true, inflow = false;
        if constexpr (OCT >= 0) inflow = ((OCT >> ax) & 1) ? (hi == 1) : (hi == 0);
        else inflow_known = false;
        if (inflow_known && !inflow) { wgt[s] = 0.0; continue; }
        const bool in = inflow_known ? true : (dot < 0);
        const bool refl = !interior && ((p.refl_mask >> s) & 1);
        const bool use_nb = in && interior, use_rf = in && refl;
        wgt[s] = (use_nb || use_rf) ? -dot : 0.0;
        const double *nbp = p.psi + base + (long long)(hi ? e + p.stride[ax] : e - p.stride[ax]) * N;
        const double *rfp = self;
        if (refl) rfp = p.psi + ((long long)g * p.Nomega + p.refl[a * NS + s]) * p.Ndof + (long long)e * N;
#pragma unroll
        for (int m = 0; m < NF; ++m) {
            const double *addr = use_nb ? nbp + side_node<DIM, P>(s ^ 1, m) : rfp + side_node<DIM, P>(s, m);
            u[s][m] = *addr;
        }
    }

    // ---- phase 2: right-hand side (the traces die here) ----------------------------------------
#pragma unroll
    for (int i = 0; i < N; ++i) b[i] += isv[i];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if constexpr (OCT >= 0) {
            const int ax = s >> 1, hi = s & 1;
            const bool inflow = ((OCT >> ax) & 1) ? (hi == 1) : (hi == 0);
            if (!inflow) continue;
        }
        const cdptr F = UNI ? as_const(p.Fface + s * NF * NF) : as_const(p.Fface + (size_t)fid[s] * NF * NF);
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            double acc = 0.0;
#pragma unroll
            for (int l = 0; l < NF; ++l) acc = fma(F[k * NF + l], u[s][l], acc);
            // select, not multiply by zero: a side that is not an inflow side read
            // an arbitrary (possibly non-finite) value
            const double nb = fma(wgt[s], acc, b[side_node<DIM, P>(s, k)]);
            b[side_node<DIM, P>(s, k)] = wgt[s] != 0.0 ? nb : b[side_node<DIM, P>(s, k)];
        }
    }

    // ---- phase 3: local matrix -------------------------------------------
    T A[N][N];
    if constexpr (UNI) {
        const cdptr K = as_const(p.K + (size_t)a * N * N);   // includes tau * Mt
        const cdptr Mu = as_const(p.Mu);
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < N; ++j) A[i][j] = (T)fma(sig, Mu[i * N + j], K[i * N + j]);
    } else {
        const double *Gd = p.Gd + (size_t)e * DIM * N * N;
        const double *Mu = p.Mu + (size_t)e * N * N;
        const double *Mt = p.Mt + (size_t)e * N * N;
        const double tau = p.timedep ? p.tau : 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < N; ++j) {
                double v = om[0] * Gd[i * N + j];
#pragma unroll
                for (int d = 1; d < DIM; ++d) v = fma(om[d], Gd[(d * N + i) * N + j], v);
                v = fma(sig, Mu[i * N + j], v);
                v = fma(tau, Mt[i * N + j], v);
                A[i][j] = (T)v;
            }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const double w = dots[s] > 0 ? dots[s] : 0.0;   // outflow faces
            const double *F = p.Fface + (size_t)fid[s] * NF * NF;
#pragma unroll
            for (int k = 0; k < NF; ++k)
#pragma unroll
                for (int l = 0; l < NF; ++l)
                    A[side_node<DIM, P>(s, k)][side_node<DIM, P>(s, l)] += (T)(w * F[k * NF + l]);
        }
    }

    // ---- phase 4: solve, fix up, store -----------------------------------
    [[maybe_unused]] double wcol[N];
    [[maybe_unused]] double bsum = 0.0;
    if constexpr (FIX) {
        // the fixups weigh the solution with the column sums of the matrix
#pragma unroll
        for (int j = 0; j < N; ++j) {
            double sj = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) sj += (double)A[i][j];
            wcol[j] = sj;
        }
#pragma unroll
        for (int i = 0; i < N; ++i) bsum += b[i];
    }
    T x[N];
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = (T)b[i];
    lu_factor_solve<N, T>(A, x);
    if constexpr (FIX) {
        static_assert(sizeof(T) == sizeof(double), "fixups run on FP64 cell solves");
        fixup_cell<N>(p, wcol, bsum, x, e, g);
    }
    if constexpr (N % 2 == 0) {
        d2v *o2 = reinterpret_cast<d2v *>(p.psi + off);
#pragma unroll
        for (int i = 0; i < N / 2; ++i) { d2v v; v.x = (double)x[2 * i]; v.y = (double)x[2 * i + 1]; o2[i] = v; }
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) p.psi[off + i] = (double)x[i];
    }
}

// octant sign bits of direction a (bit d set when Omega_d < 0)
__device__ __forceinline__ int octant_of(const SweepParams &p, int a)
{
    return (p.omega[4 * a] < 0 ? 1 : 0) | (p.omega[4 * a + 1] < 0 ? 2 : 0) | (p.omega[4 * a + 2] < 0 ? 4 : 0);
}

// dispatch a direction-uniform cell solve on the (wave-uniform) octant
template <int DIM, int P, typename T>
__device__ __forceinline__ void cell_solve_octant(const SweepParams &p, int e, int a, int g, int oct)
{
    switch (oct) {
    case 0: cell_solve<DIM, P, true, T, 0>(p, e, a, g); break;
    case 1: cell_solve<DIM, P, true, T, 1>(p, e, a, g); break;
    case 2: cell_solve<DIM, P, true, T, 2>(p, e, a, g); break;
    case 3: cell_solve<DIM, P, true, T, 3>(p, e, a, g); break;
    case 4: if constexpr (DIM == 3) cell_solve<DIM, P, true, T, 4>(p, e, a, g); break;
    case 5: if constexpr (DIM == 3) cell_solve<DIM, P, true, T, 5>(p, e, a, g); break;
    case 6: if constexpr (DIM == 3) cell_solve<DIM, P, true, T, 6>(p, e, a, g); break;
    default: if constexpr (DIM == 3) cell_solve<DIM, P, true, T, 7>(p, e, a, g); break;
    }
}

// ---------------------------------------------------------------------------
// Multi-lane cell solve: LPC lanes (a pair or a quad) share one cell; lane l
// owns rows [l*RPL, (l+1)*RPL) of the local system.  Loads and stores of q and
// psi are 8*RPL contiguous bytes per lane, so one instruction of a wave covers
// 64/LPC whole cells; pivot rows and the back-substituted unknowns move between
// the lanes of a group with DPP quad_perm broadcasts (no LDS traffic).
// ---------------------------------------------------------------------------
template <int LPC, int SRC>
__device__ __forceinline__ constexpr int bcast_ctrl()
{
    // quad_perm(s0,s1,s2,s3) = s0 | s1<<2 | s2<<4 | s3<<6
    return LPC == 4 ? SRC * 0x55 : (SRC | (SRC << 2) | ((2 + SRC) << 4) | ((2 + SRC) << 6));
}

template <int LPC, int SRC>
__device__ __forceinline__ double group_bcast(double v)
{
    if constexpr (LPC == 1) {
        return v;
    } else {
        constexpr int ctrl = bcast_ctrl<LPC, SRC>();
        const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), ctrl, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), ctrl, 0xF, 0xF, false);
        return __hiloint2double(hi, lo);
    }
}

template <int LPC, int SRC>
__device__ __forceinline__ float group_bcast(float v)
{
    if constexpr (LPC == 1) {
        return v;
    } else {
        constexpr int ctrl = bcast_ctrl<LPC, SRC>();
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xF, 0xF, false));
    }
}

template <int DIM, int P, bool UNI, typename T, int LPC>
__device__ __forceinline__ void group_cell_solve(const SweepParams &p, const int e, const int a, const int g,
                                                 const int l)
{
    using S = Sh<DIM, P>;
    constexpr int N = S::N, NF = S::NF, NS = S::NS, RPL = N / LPC;
    static_assert(N % LPC == 0, "lanes per cell must divide the dofs per cell");
    int c[3];
    {
        const int q01 = (DIM == 3) ? fast_div(e, p.stride[2], p.inv_nel01) : 0;
        const int rem = e - q01 * p.stride[2];
        const int q0 = fast_div(rem, p.nel[0], p.inv_nel0);
        c[0] = rem - q0 * p.nel[0];
        c[1] = q0;
        c[2] = q01;
    }
    const double om[3] = {p.omega[4 * a + 0], p.omega[4 * a + 1], p.omega[4 * a + 2]};
    const long long base = ((long long)g * p.Nomega + a) * p.Ndof;
    const long long off = base + (long long)e * N;
    const double sig = p.sigma[(long long)g * p.Ne + e];
    const int r0 = l * RPL;

    T A[RPL][N];
    if constexpr (UNI) {
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
            const double *Kr = p.K + ((size_t)a * N + r0 + j) * N;   // includes tau * Mt
            const double *Mr = p.Mu + (size_t)(r0 + j) * N;
#pragma unroll
            for (int cc = 0; cc < N; ++cc) A[j][cc] = (T)fma(sig, Mr[cc], Kr[cc]);
        }
    } else {
        const double *Gd = p.Gd + (size_t)e * DIM * N * N;
        const double *Mu = p.Mu + (size_t)e * N * N;
        const double *Mt = p.Mt + (size_t)e * N * N;
#pragma unroll
        for (int j = 0; j < RPL; ++j)
#pragma unroll
            for (int cc = 0; cc < N; ++cc) {
                const int ij = (r0 + j) * N + cc;
                double v = om[0] * Gd[ij];
#pragma unroll
                for (int d = 1; d < DIM; ++d) v = fma(om[d], Gd[d * N * N + ij], v);
                v = fma(sig, Mu[ij], v);
                if (p.timedep) v = fma(p.tau, Mt[ij], v);
                A[j][cc] = (T)v;
            }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int ax = s >> 1, hi = s & 1;
            int cf[3] = {c[0], c[1], c[2]};
            cf[ax] += hi;
            const int f = p.foff[ax] + cf[0] + p.fext[ax][0] * (cf[1] + p.fext[ax][1] * cf[2]);
            const bool interior = hi ? (c[ax] < p.nel[ax] - 1) : (c[ax] > 0);
            double dot = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) dot = fma(om[d], p.fnor[3 * f + d], dot);
            if (!hi && interior) dot = -dot;
            if (dot > 0) {
                const double *F = p.Fface + (size_t)f * NF * NF;
#pragma unroll
                for (int j = 0; j < RPL; ++j) {
                    int k;
                    if (on_side<DIM, P>(r0 + j, s, k)) {
#pragma unroll
                        for (int m = 0; m < NF; ++m)
                            A[j][side_node<DIM, P>(s, m)] += (T)(dot * F[k * NF + m]);
                    }
                }
            }
        }
    }

    double b[RPL];
#pragma unroll
    for (int j = 0; j < RPL; ++j) b[j] = 0.0;
    if (p.src) {
        if constexpr (RPL % 2 == 0) {
            const d2v *q2 = reinterpret_cast<const d2v *>(p.src + off + r0);
#pragma unroll
            for (int j = 0; j < RPL / 2; ++j) { const d2v v = q2[j]; b[2 * j] = v.x; b[2 * j + 1] = v.y; }
        } else {
#pragma unroll
            for (int j = 0; j < RPL; ++j) b[j] = p.src[off + r0 + j];
        }
    }
    if (p.iso) {
        const double *is = p.iso + (long long)g * p.Ndof + (long long)e * N + r0;
#pragma unroll
        for (int j = 0; j < RPL; ++j) b[j] += is[j];
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int ax = s >> 1, hi = s & 1;
        const bool interior = hi ? (c[ax] < p.nel[ax] - 1) : (c[ax] > 0);
        double dot;
        const double *F;
        if constexpr (UNI) {
            dot = hi ? om[ax] : -om[ax];
            F = p.Fface + s * NF * NF;
        } else {
            int cf[3] = {c[0], c[1], c[2]};
            cf[ax] += hi;
            const int f = p.foff[ax] + cf[0] + p.fext[ax][0] * (cf[1] + p.fext[ax][1] * cf[2]);
            dot = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) dot = fma(om[d], p.fnor[3 * f + d], dot);
            if (!hi && interior) dot = -dot;
            F = p.Fface + (size_t)f * NF * NF;
        }
        if (dot < 0) {
            const double *up = nullptr;
            bool same = false;
            if (interior) {
                const int nb = hi ? e + p.stride[ax] : e - p.stride[ax];
                up = p.psi + base + (long long)nb * N;
            } else if ((p.refl_mask >> s) & 1) {
                const int ra = p.refl[a * NS + s];
                up = p.psi + ((long long)g * p.Nomega + ra) * p.Ndof + (long long)e * N;
                same = true;
            }
            if (up) {
                double u[NF];
                if (same) {
#pragma unroll
                    for (int m = 0; m < NF; ++m) u[m] = up[side_node<DIM, P>(s, m)];
                } else {
#pragma unroll
                    for (int m = 0; m < NF; ++m) u[m] = up[side_node<DIM, P>(s ^ 1, m)];
                }
#pragma unroll
                for (int j = 0; j < RPL; ++j) {
                    int k;
                    if (on_side<DIM, P>(r0 + j, s, k)) {
                        double acc = 0.0;
#pragma unroll
                        for (int m = 0; m < NF; ++m) acc = fma(F[k * NF + m], u[m], acc);
                        b[j] = fma(-dot, acc, b[j]);
                    }
                }
            }
        }
    }

    // LU without pivoting across the lanes of the group
    T x[RPL];
#pragma unroll
    for (int j = 0; j < RPL; ++j) x[j] = (T)b[j];
    T inv[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        constexpr int dummy = 0;
        (void)dummy;
        const int kl = k / RPL, kj = k % RPL;
        T piv[N];
        T pb;
        // the owner lane kl broadcasts row kj of its block (compile-time lane)
#define ALLIUM_BCAST_ROW(LANE)                                                        \
    if (kl == LANE) {                                                                 \
        _Pragma("unroll") for (int cc = 0; cc < N; ++cc) if (cc >= k) piv[cc] =        \
            group_bcast<LPC, LANE>(A[kj][cc]);                                        \
        pb = group_bcast<LPC, LANE>(x[kj]);                                           \
    }
        ALLIUM_BCAST_ROW(0)
        if constexpr (LPC > 1) { ALLIUM_BCAST_ROW(1) }
        if constexpr (LPC > 2) { ALLIUM_BCAST_ROW(2) ALLIUM_BCAST_ROW(3) }
#undef ALLIUM_BCAST_ROW
        inv[k] = fast_rcp(piv[k]);
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
            const T lm = (r0 + j > k) ? A[j][k] * inv[k] : T(0);
#pragma unroll
            for (int cc = 0; cc < N; ++cc)
                if (cc > k) A[j][cc] = fma(-lm, piv[cc], A[j][cc]);
            x[j] = fma(-lm, pb, x[j]);
        }
    }
#pragma unroll
    for (int k = N - 1; k >= 0; --k) {
        const int kl = k / RPL, kj = k % RPL;
        T xk = T(0);
#define ALLIUM_BCAST_X(LANE) if (kl == LANE) xk = group_bcast<LPC, LANE>(x[kj]);
        ALLIUM_BCAST_X(0)
        if constexpr (LPC > 1) { ALLIUM_BCAST_X(1) }
        if constexpr (LPC > 2) { ALLIUM_BCAST_X(2) ALLIUM_BCAST_X(3) }
#undef ALLIUM_BCAST_X
        xk *= inv[k];
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
            if (r0 + j < k) x[j] = fma(-A[j][k], xk, x[j]);
            else if (r0 + j == k) x[j] = xk;
        }
    }
    if constexpr (RPL % 2 == 0) {
        d2v *o2 = reinterpret_cast<d2v *>(p.psi + off + r0);
#pragma unroll
        for (int j = 0; j < RPL / 2; ++j) { d2v v; v.x = (double)x[2 * j]; v.y = (double)x[2 * j + 1]; o2[j] = v; }
    } else {
#pragma unroll
        for (int j = 0; j < RPL; ++j) p.psi[off + r0 + j] = (double)x[j];
    }
}

template <int DIM, int P, bool UNI, typename T, int LPC>
__global__ void __launch_bounds__(256) k_sweep_level_group(SweepParams p, long long cbeg, int ncells)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int item = t / LPC, l = t % LPC;
    if (item >= ncells * p.G) return;
    int g = 0, c = item;
    if (p.G > 1) { g = item / ncells; c = item - g * ncells; }
    const uint32_t cell = p.cells[cbeg + c];
    group_cell_solve<DIM, P, UNI, T, LPC>(p, (int)(cell >> 10), (int)(cell & 1023u), g, l);
}

// Direction-uniform waves over XCD bins: every wave holds 64 cells of one
// (level, direction) class, so the direction index is wave-uniform and the
// direction's streaming matrix K[a] is read with scalar loads; workgroup
// blockIdx.x % 8 serves spatial bin blockIdx.x % 8, so on MI355X (workgroups
// dispatched round-robin over the 8 XCDs) the upwind traces a cell reads were
// written through the same XCD's L2.
template <int DIM, int P, typename T>
__global__ void __launch_bounds__(256) k_sweep_wave(SweepParams p, const uint32_t *__restrict__ wcells, WaveLevel wl)
{
    const int bin = blockIdx.x & 7;
    const int rest = blockIdx.x >> 3;
    const int part = rest % wl.nparts, chunk = rest / wl.nparts;
    const int pb = part * 8 + bin;
    const int w = chunk * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6);
    if (w >= wl.pb[pb + 1] - wl.pb[pb]) return;
    const long long idx = wl.base + (long long)(wl.pb[pb] + w) * 64 + (threadIdx.x & 63);
    const uint32_t cell = wcells[idx];
    const int a = __builtin_amdgcn_readfirstlane((int)(cell & 1023u));
    if (cell == 0xFFFFFFFFu) return;
    cell_solve_octant<DIM, P, T>(p, (int)(cell >> 10), a, (int)blockIdx.y, octant_of(p, a));
}

// Direction sweeps: workgroup (a, g) sweeps direction a of group g through
// all of its levels, separated by workgroup barriers.  Directions never
// couple without reflective sides, so workgroups share no data: every upwind
// trace is written and read through the same CU, the direction is uniform for
// the whole workgroup (its streaming matrix comes through scalar loads) and a
// sweep is one launch.
template <int DIM, int P, bool UNI, typename T>
__global__ void __launch_bounds__(512) k_sweep_direction(SweepParams p, const long long *__restrict__ cls_ptr,
                                                         int nlevels)
{
    const int a = blockIdx.x, g = blockIdx.y, No = p.Nomega;
    for (int L = 0; L < nlevels; ++L) {
        const long long b = cls_ptr[(long long)L * No + a], e = cls_ptr[(long long)L * No + a + 1];
        for (long long i = b + threadIdx.x; i < e; i += blockDim.x) {
            if constexpr (UNI) cell_solve_octant<DIM, P, T>(p, (int)(p.cells[i] >> 10), a, g, octant_of(p, a));
            else cell_solve<DIM, P, false, T, -1>(p, (int)(p.cells[i] >> 10), a, g);
        }
        __syncthreads();
    }
}

template <int DIM, int P, bool UNI, typename T>
__global__ void __launch_bounds__(256) k_sweep_level(SweepParams p, long long cbeg, int ncells)
{
    const int item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= ncells * p.G) return;
    int g = 0, c = item;
    if (p.G > 1) { g = item / ncells; c = item - g * ncells; }
    const uint32_t cell = p.cells[cbeg + c];
    cell_solve<DIM, P, UNI, T, -1>(p, (int)(cell >> 10), (int)(cell & 1023u), g);
}

// level kernel with the negative-flux fixup: one lane per cell, FP64
template <int DIM, int P, bool UNI>
__global__ void __launch_bounds__(256) k_sweep_level_fixup(SweepParams p, long long cbeg, int ncells)
{
    const int item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= ncells * p.G) return;
    int g = 0, c = item;
    if (p.G > 1) { g = item / ncells; c = item - g * ncells; }
    const uint32_t cell = p.cells[cbeg + c];
    cell_solve<DIM, P, UNI, double, -1, true>(p, (int)(cell >> 10), (int)(cell & 1023u), g);
}

// Persistent sweep.  Tiles are handed out in level order by an atomic ticket,
// so any tile a workgroup waits on was taken earlier by a running workgroup:
// the wait always ends.  Each wait is also bounded by spin_limit; on overrun the
// error flag is raised and the workgroup proceeds, so the grid always drains.
template <int DIM, int P, bool UNI, typename T>
__global__ void __launch_bounds__(256) k_sweep_persistent(SweepParams p, const Tile *__restrict__ tiles,
                                                          int ntiles, int *counters, const int *need,
                                                          int nkeys, long long spin_limit)
{
    __shared__ int s_t;
    int *ticket = counters + nkeys;
    int *err = counters + nkeys + 1;
    for (;;) {
        if (threadIdx.x == 0) s_t = atomicAdd(ticket, 1);
        __syncthreads();
        const int t = s_t;
        __syncthreads();
        if (t >= ntiles) return;
        const Tile tl = tiles[t];
        if (threadIdx.x == 0 && tl.wait_key >= 0) {
            const int target = need[tl.wait_key];
            long long spins = 0;
            while (__hip_atomic_load(&counters[tl.wait_key], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                   target) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > spin_limit) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                // once any wait has overrun, stop waiting so the grid drains at once
                if ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    break;
            }
        }
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        for (int i = threadIdx.x; i < tl.nitems; i += blockDim.x) {
            const int item = tl.item_begin + i;
            int g = 0, c = item;
            if (p.G > 1) { g = item / tl.ncells; c = item - g * tl.ncells; }
            const uint32_t cell = p.cells[tl.cell_begin + c];
            cell_solve<DIM, P, UNI, T, -1>(p, (int)(cell >> 10), (int)(cell & 1023u), g);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (threadIdx.x == 0)
            __hip_atomic_fetch_add(&counters[tl.key], tl.nitems, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void k_fill_int(int *p, int n, int v)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// dispatch over (dim, p, uniform, precision)
// ---------------------------------------------------------------------------
bool sweep_supported(int dim, int p, std::string &why)
{
    if (dim == 2 && p >= 0 && p <= 2) return true;
    if (dim == 3 && p >= 0 && p <= 2) return true;
    why = "GPU sweep implements 2D p=0..2 and 3D p=0..2";
    return false;
}

// levels of one instance through the cooperative kernel (3D p = 2)
static hipError_t run_coop(DeviceSweep &ds, const SweepParams &prm, hipStream_t s, int mode)
{
    if (mode != 0) return hipErrorInvalidValue;
    for (int L = 0; L < ds.nlevels; ++L) {
        CoopLevel cl{};
        cl.ninst = 1;
        cl.cbeg[0] = ds.level_ptr[L];
        cl.ncells[0] = (int)(ds.level_ptr[L + 1] - ds.level_ptr[L]);
        cl.start[0] = 0;
        cl.start[1] = cl.ncells[0] * prm.G;
        if (cl.start[1] == 0) continue;
        hipError_t e = launch_coop_level(prm, nullptr, cl, prm.dim, prm.p, prm.uniform != 0, ds.precision, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <int DIM, int P, bool UNI, typename T, int LPC>
static void launch_level(const DeviceSweep &ds, const SweepParams &prm, hipStream_t stream, long long cb, int nc)
{
    const long threads = (long)nc * prm.G * LPC;
    const int grid = (int)((threads + ds.block - 1) / ds.block);
    if constexpr (LPC == 1)
        hipLaunchKernelGGL((k_sweep_level<DIM, P, UNI, T>), dim3(grid), dim3(ds.block), 0, stream, prm, cb, nc);
    else
        hipLaunchKernelGGL((k_sweep_level_group<DIM, P, UNI, T, LPC>), dim3(grid), dim3(ds.block), 0, stream, prm,
                           cb, nc);
}

template <int DIM, int P>
constexpr bool lanes_ok(int lpc)
{
    return lpc == 1 || (Sh<DIM, P>::N % lpc == 0 && Sh<DIM, P>::N / lpc >= 1);
}

template <int DIM, int P, bool UNI, typename T>
static hipError_t run_levels(DeviceSweep &ds, const SweepParams &prm, hipStream_t stream)
{
    if constexpr (UNI) {
        if (ds.lanes == 0 && ds.d_wcells) {
            const int wpb = ds.block / 64;
            for (int L = 0; L < ds.nlevels; ++L) {
                const int *pb = &ds.wpb[(size_t)L * 9];
                WaveLevel wl;
                wl.nparts = 1;
                wl.base = ds.wlevel_base[L];
                int maxw = 0;
                for (int k = 0; k <= 8; ++k) wl.pb[k] = pb[k];
                for (int k = 0; k < 8; ++k) maxw = std::max(maxw, wl.pb[k + 1] - wl.pb[k]);
                if (maxw == 0) continue;
                const int chunks = (maxw + wpb - 1) / wpb;
                hipLaunchKernelGGL((k_sweep_wave<DIM, P, T>), dim3(8 * chunks, prm.G), dim3(ds.block), 0, stream, prm,
                                   ds.d_wcells, wl);
            }
            return hipGetLastError();
        }
    }
    const int lpc = (ds.lanes == 4 && lanes_ok<DIM, P>(4)) ? 4 : (ds.lanes == 2 && lanes_ok<DIM, P>(2)) ? 2 : 1;
    for (int L = 0; L < ds.nlevels; ++L) {
        const long long cb = ds.level_ptr[L];
        const int nc = (int)(ds.level_ptr[L + 1] - cb);
        if ((long)nc * prm.G == 0) continue;
        if constexpr (Sh<DIM, P>::N % 4 == 0) {
            if (lpc == 4) { launch_level<DIM, P, UNI, T, 4>(ds, prm, stream, cb, nc); continue; }
        }
        if constexpr (Sh<DIM, P>::N % 2 == 0) {
            if (lpc == 2) { launch_level<DIM, P, UNI, T, 2>(ds, prm, stream, cb, nc); continue; }
        }
        launch_level<DIM, P, UNI, T, 1>(ds, prm, stream, cb, nc);
    }
    return hipGetLastError();
}

template <int DIM, int P, bool UNI, typename T>
static hipError_t run_persistent(DeviceSweep &ds, const SweepParams &prm, hipStream_t stream)
{
    hipLaunchKernelGGL(k_fill_int, dim3(std::max(1, std::min(1024, (ds.nkeys + 2 + 255) / 256))), dim3(256), 0,
                       stream, ds.d_counters, ds.nkeys + 2, 0);
    int dev = 0, ncu = 256;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const int grid = std::max(1, std::min(ds.ntiles, ncu * 4));
    hipLaunchKernelGGL((k_sweep_persistent<DIM, P, UNI, T>), dim3(grid), dim3(ds.persistent_tile), 0, stream, prm,
                       ds.d_tiles, ds.ntiles, ds.d_counters, ds.d_need, ds.nkeys, (long long)(1LL << 21));
    return hipGetLastError();
}

template <int DIM, int P, bool UNI, typename T>
static hipError_t run_directions(DeviceSweep &ds, const SweepParams &prm, hipStream_t stream)
{
    const int block = std::max(64, std::min(512, ds.dir_block));
    hipLaunchKernelGGL((k_sweep_direction<DIM, P, UNI, T>), dim3(prm.Nomega, prm.G), dim3(block), 0, stream, prm,
                       ds.d_cls_ptr, ds.nlevels);
    return hipGetLastError();
}

template <int DIM, int P, bool UNI, typename T>
static hipError_t run_mode(DeviceSweep &ds, const SweepParams &prm, hipStream_t stream, int mode)
{
    if (mode == 3) return run_directions<DIM, P, UNI, T>(ds, prm, stream);
    if (mode == 2) return run_persistent<DIM, P, UNI, T>(ds, prm, stream);
    return run_levels<DIM, P, UNI, T>(ds, prm, stream);
}

template <int DIM, int P>
static hipError_t dispatch_dp(DeviceSweep &ds, const SweepParams &prm, hipStream_t s, int mode)
{
    if (prm.uniform) {
        if (ds.precision == 1) return run_mode<DIM, P, true, float>(ds, prm, s, mode);
        return run_mode<DIM, P, true, double>(ds, prm, s, mode);
    }
    if (ds.precision == 1) return run_mode<DIM, P, false, float>(ds, prm, s, mode);
    return run_mode<DIM, P, false, double>(ds, prm, s, mode);
}

// sweeps with a fixup: the level kernel in every launch mode (the box and
// direction kernels have no fixup), FP64 cells
template <int DIM, int P>
static hipError_t run_fixup_levels(DeviceSweep &ds, const SweepParams &prm, hipStream_t stream)
{
    const int block = 256;
    for (int L = 0; L < ds.nlevels; ++L) {
        const long long cb = ds.level_ptr[L];
        const int nc = (int)(ds.level_ptr[L + 1] - cb);
        if ((long)nc * prm.G == 0) continue;
        const int grid = (int)(((long)nc * prm.G + block - 1) / block);
        if (prm.uniform)
            hipLaunchKernelGGL((k_sweep_level_fixup<DIM, P, true>), dim3(grid), dim3(block), 0, stream, prm, cb, nc);
        else
            hipLaunchKernelGGL((k_sweep_level_fixup<DIM, P, false>), dim3(grid), dim3(block), 0, stream, prm, cb, nc);
    }
    return hipGetLastError();
}

bool fixup_supported(int dim, int p) { return (dim == 2 && p >= 0 && p <= 2) || (dim == 3 && p >= 0 && p <= 1); }

static hipError_t dispatch_fixup(DeviceSweep &ds, const SweepParams &prm, hipStream_t s)
{
    if (prm.dim == 2) {
        switch (prm.p) {
        case 0: return run_fixup_levels<2, 0>(ds, prm, s);
        case 1: return run_fixup_levels<2, 1>(ds, prm, s);
        case 2: return run_fixup_levels<2, 2>(ds, prm, s);
        }
    } else if (prm.dim == 3) {
        switch (prm.p) {
        case 0: return run_fixup_levels<3, 0>(ds, prm, s);
        case 1: return run_fixup_levels<3, 1>(ds, prm, s);
        }
    }
    return hipErrorInvalidValue;
}

static hipError_t dispatch(DeviceSweep &ds, const SweepParams &prm, hipStream_t s, int mode)
{
    if (prm.dim == 2) {
        switch (prm.p) {
        case 0: return dispatch_dp<2, 0>(ds, prm, s, mode);
        case 1: return dispatch_dp<2, 1>(ds, prm, s, mode);
        case 2: return dispatch_dp<2, 2>(ds, prm, s, mode);
        }
    } else if (prm.dim == 3) {
        switch (prm.p) {
        case 0: return dispatch_dp<3, 0>(ds, prm, s, mode);
        case 1: return dispatch_dp<3, 1>(ds, prm, s, mode);
        case 2: return run_coop(ds, prm, s, mode);
        }
    }
    return hipErrorInvalidValue;
}

hipError_t reset_persistent(DeviceSweep &ds, hipStream_t stream)
{
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, stream, ds.d_counters, ds.nkeys + 2, 0);
    return hipGetLastError();
}

hipError_t launch_sweep(DeviceSweep &ds, const double *src, const double *iso, double *psi, hipStream_t stream,
                        double *phi)
{
    SweepParams prm = ds.prm;
    prm.src = src;
    prm.iso = iso;
    prm.psi = psi;
    prm.nt_psi = (ds.stream_psi || phi) ? 1 : 0;
    if (phi && (phi == iso || (const double *)phi == src || phi == psi)) return hipErrorInvalidValue;
    hipError_t err = hipSuccess;
    if (ds.timing) hipEventRecord(ds.ev0, stream);
    if (prm.fixup) {
        err = dispatch_fixup(ds, prm, stream);
    } else if (ds.launch_mode == 1 || ds.launch_mode == 5) {
        // hipGraph of the level (1) or box-level (5) launches, re-captured when
        // the vectors change
        const int variant = ds.precision + 2 * ds.refine + 4 * prm.nt_psi;
        if (!ds.graph_exec || ds.graph_src != src || ds.graph_iso != iso || ds.graph_psi != psi ||
            ds.graph_prec != variant) {
            if (ds.graph_exec) { hipGraphExecDestroy(ds.graph_exec); ds.graph_exec = nullptr; }
            if (!ds.graph_stream) hipStreamCreateWithFlags(&ds.graph_stream, hipStreamNonBlocking);
            hipGraph_t graph;
            err = hipStreamBeginCapture(ds.graph_stream, hipStreamCaptureModeThreadLocal);
            if (err != hipSuccess) return err;
            const hipError_t lerr = ds.launch_mode == 5 ? launch_box_sweep(ds, prm, ds.graph_stream)
                                                        : dispatch(ds, prm, ds.graph_stream, 0);
            err = hipStreamEndCapture(ds.graph_stream, &graph);
            if (lerr != hipSuccess) {
                if (err == hipSuccess) hipGraphDestroy(graph);
                return lerr;
            }
            if (err != hipSuccess) return err;
            err = hipGraphInstantiate(&ds.graph_exec, graph, nullptr, nullptr, 0);
            hipGraphDestroy(graph);
            if (err != hipSuccess) return err;
            ds.graph_src = src; ds.graph_iso = iso; ds.graph_psi = psi; ds.graph_prec = variant;
        }
        err = hipGraphLaunch(ds.graph_exec, stream);
    } else if (ds.launch_mode == 4) {
        err = launch_box_sweep(ds, prm, stream);
    } else {
        err = dispatch(ds, prm, stream, ds.launch_mode);
    }
    if (err == hipSuccess && phi) err = launch_moments(ds.prm, psi, phi, stream);
    if (ds.timing) hipEventRecord(ds.ev1, stream);
    ds.sweeps++;
    return err;
}

// ---------------------------------------------------------------------------
// streaming operator L (AdvectionOperator::Mult, src/sweep.cpp:1246-1387),
// one thread per (element, direction, group)
// ---------------------------------------------------------------------------
template <int DIM, int P, bool UNI>
__global__ void __launch_bounds__(256) k_apply_L(SweepParams p, const double *__restrict__ x, double *__restrict__ y,
                                                 const double *__restrict__ q)
{
    using S = Sh<DIM, P>;
    constexpr int N = S::N, NF = S::NF, NS = S::NS;
    const long long item = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)p.Ne * p.Nomega * p.G;
    if (item >= total) return;
    const int e = (int)(item % p.Ne);
    const int a = (int)((item / p.Ne) % p.Nomega);
    const int g = (int)(item / ((long long)p.Ne * p.Nomega));
    int c[3] = {e % p.nel[0], (e / p.nel[0]) % p.nel[1], e / (p.nel[0] * p.nel[1])};
    const double om[3] = {p.omega[4 * a], p.omega[4 * a + 1], p.omega[4 * a + 2]};
    const long long base = ((long long)g * p.Nomega + a) * p.Ndof;
    const double *xe = x + base + (long long)e * N;
    const double sig = p.sigma[(long long)g * p.Ne + e];
    double xl[N], yl[N];
#pragma unroll
    for (int i = 0; i < N; ++i) xl[i] = xe[i];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            double v;
            if constexpr (UNI) {
                v = fma(sig, p.Mu[i * N + j], p.K[((size_t)a * N + i) * N + j]);   // K includes tau * Mt
            } else {
                const double *Gd = p.Gd + (size_t)e * DIM * N * N;
                v = om[0] * Gd[i * N + j];
#pragma unroll
                for (int d = 1; d < DIM; ++d) v = fma(om[d], Gd[(d * N + i) * N + j], v);
                v = fma(sig, p.Mu[(size_t)e * N * N + i * N + j], v);
                if (p.timedep) v = fma(p.tau, p.Mt[(size_t)e * N * N + i * N + j], v);
            }
            acc = fma(v, xl[j], acc);
        }
        yl[i] = acc;
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int ax = s >> 1, hi = s & 1;
        const bool interior = hi ? (c[ax] < p.nel[ax] - 1) : (c[ax] > 0);
        double dot;
        const double *F;
        if constexpr (UNI) {
            dot = hi ? om[ax] : -om[ax];
            F = p.Fface + s * NF * NF;
        } else {
            int cf[3] = {c[0], c[1], c[2]};
            cf[ax] += hi;
            const int f = p.foff[ax] + cf[0] + p.fext[ax][0] * (cf[1] + p.fext[ax][1] * cf[2]);
            const double sgn = (!hi && interior) ? -1.0 : 1.0;
            dot = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) dot = fma(om[d], p.fnor[3 * f + d], dot);
            dot *= sgn;
            F = p.Fface + (size_t)f * NF * NF;
            if (dot > 0) {  // outflow term (uniform geometry has it inside K)
#pragma unroll
                for (int k = 0; k < NF; ++k) {
                    double acc = 0.0;
#pragma unroll
                    for (int l = 0; l < NF; ++l) acc = fma(F[k * NF + l], xl[side_node<DIM, P>(s, l)], acc);
                    yl[side_node<DIM, P>(s, k)] = fma(dot, acc, yl[side_node<DIM, P>(s, k)]);
                }
            }
        }
        if (dot < 0 && interior) {
            const int nb = hi ? e + p.stride[ax] : e - p.stride[ax];
            const double *xn = x + base + (long long)nb * N;
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                double acc = 0.0;
#pragma unroll
                for (int l = 0; l < NF; ++l) acc = fma(F[k * NF + l], xn[side_node<DIM, P>(s ^ 1, l)], acc);
                yl[side_node<DIM, P>(s, k)] = fma(dot, acc, yl[side_node<DIM, P>(s, k)]);
            }
        } else if (q && dot < 0 && ((p.refl_mask >> s) & 1)) {
            // the defect is that of the operator the sweep inverts, which couples a
            // reflective inflow side to the reflected direction on the element's own
            // nodes (AdvectionOperator::Mult leaves it out: y = L x stays as the reference's)
            const double *xr = x + ((long long)g * p.Nomega + p.refl[a * NS + s]) * p.Ndof + (long long)e * N;
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                double acc = 0.0;
#pragma unroll
                for (int l = 0; l < NF; ++l) acc = fma(F[k * NF + l], xr[side_node<DIM, P>(s, l)], acc);
                yl[side_node<DIM, P>(s, k)] = fma(dot, acc, yl[side_node<DIM, P>(s, k)]);
            }
        }
    }
    // y = L x, or the defect y = q - L x when q is given
    double *ye = y + base + (long long)e * N;
    if (q) {
        const double *qe = q + base + (long long)e * N;
#pragma unroll
        for (int i = 0; i < N; ++i) ye[i] = qe[i] - yl[i];
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) ye[i] = yl[i];
    }
}

template <int DIM, int P>
static hipError_t apply_L_dp(const SweepParams &prm, const double *x, double *y, const double *q, hipStream_t s)
{
    const long long total = (long long)prm.Ne * prm.Nomega * prm.G;
    const int grid = (int)((total + 255) / 256);
    if (prm.uniform) hipLaunchKernelGGL((k_apply_L<DIM, P, true>), dim3(grid), dim3(256), 0, s, prm, x, y, q);
    else hipLaunchKernelGGL((k_apply_L<DIM, P, false>), dim3(grid), dim3(256), 0, s, prm, x, y, q);
    return hipGetLastError();
}

static hipError_t apply_L_any(const SweepParams &prm, const double *x, double *y, const double *q, hipStream_t s)
{
    if (prm.dim == 2) {
        if (prm.p == 0) return apply_L_dp<2, 0>(prm, x, y, q, s);
        if (prm.p == 1) return apply_L_dp<2, 1>(prm, x, y, q, s);
        if (prm.p == 2) return apply_L_dp<2, 2>(prm, x, y, q, s);
    } else {
        if (prm.p == 0) return apply_L_dp<3, 0>(prm, x, y, q, s);
        if (prm.p == 1) return apply_L_dp<3, 1>(prm, x, y, q, s);
        if (prm.p == 2) return apply_L_dp<3, 2>(prm, x, y, q, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_apply_L(const SweepParams &prm, const double *x, double *y, hipStream_t s)
{
    return apply_L_any(prm, x, y, nullptr, s);
}

hipError_t launch_defect_L(const SweepParams &prm, const double *q, const double *x, double *y, hipStream_t s)
{
    return apply_L_any(prm, x, y, q, s);
}

// ---------------------------------------------------------------------------
// particle balance (ComputeBalance, src/sweep.cpp:1400-1462): per (g, e)
// absorption sum_a w_a 1^T (sigma M + tau Mt) psi and boundary leakage
// sum_{dot >= 0} w_a dot 1^T F psi, block partials in FP64
// ---------------------------------------------------------------------------
constexpr int kBalanceBlock = 256;

template <int DIM, int P>
__global__ void __launch_bounds__(kBalanceBlock) k_balance(SweepParams p, const double *__restrict__ psi,
                                                           double *__restrict__ partial)
{
    using S = Sh<DIM, P>;
    constexpr int N = S::N, NF = S::NF, NS = S::NS;
    __shared__ double sa[kBalanceBlock], sl[kBalanceBlock];
    const long long item = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double absorb = 0.0, leak = 0.0;
    if (item < (long long)p.Ne * p.G) {
        const int e = (int)(item % p.Ne), g = (int)(item / p.Ne);
        const double sig = p.sigma[(long long)g * p.Ne + e];
        const double *Mu = p.uniform ? p.Mu : p.Mu + (size_t)e * N * N;
        const double *Mt = p.uniform ? p.Mt : p.Mt + (size_t)e * N * N;
        double col[N];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                double v = sig * Mu[i * N + j];
                if (p.timedep) v = fma(p.tau, Mt[i * N + j], v);
                s += v;
            }
            col[j] = s;
        }
        int c[3] = {e % p.nel[0], (e / p.nel[0]) % p.nel[1], e / (p.nel[0] * p.nel[1])};
        for (int a = 0; a < p.Nomega; ++a) {
            const double *pe = psi + ((long long)g * p.Nomega + a) * p.Ndof + (long long)e * N;
            const double wa = p.omega[4 * a + 3];
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < N; ++j) s = fma(col[j], pe[j], s);
            absorb = fma(s, wa, absorb);
#pragma unroll
            for (int sd = 0; sd < NS; ++sd) {
                const int ax = sd >> 1, hi = sd & 1;
                const bool bdr = hi ? (c[ax] == p.nel[ax] - 1) : (c[ax] == 0);
                if (!bdr) continue;
                double dot;
                const double *F;
                if (p.uniform) {
                    dot = hi ? p.omega[4 * a + ax] : -p.omega[4 * a + ax];
                    F = p.Fface + sd * NF * NF;
                } else {
                    int cf[3] = {c[0], c[1], c[2]};
                    cf[ax] += hi;
                    const int f = p.foff[ax] + cf[0] + p.fext[ax][0] * (cf[1] + p.fext[ax][1] * cf[2]);
                    dot = 0.0;
                    for (int d = 0; d < DIM; ++d) dot = fma(p.omega[4 * a + d], p.fnor[3 * f + d], dot);
                    F = p.Fface + (size_t)f * NF * NF;
                }
                if (dot >= 0) {
                    double t = 0.0;
#pragma unroll
                    for (int k = 0; k < NF; ++k)
#pragma unroll
                        for (int l = 0; l < NF; ++l) t = fma(F[k * NF + l], pe[side_node<DIM, P>(sd, l)], t);
                    leak = fma(wa * dot, t, leak);
                }
            }
        }
    }
    sa[threadIdx.x] = absorb;
    sl[threadIdx.x] = leak;
    __syncthreads();
    for (int w = kBalanceBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) { sa[threadIdx.x] += sa[threadIdx.x + w]; sl[threadIdx.x] += sl[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { partial[2 * blockIdx.x] = sa[0]; partial[2 * blockIdx.x + 1] = sl[0]; }
}

int balance_blocks(const SweepParams &prm)
{
    return (int)(((long long)prm.Ne * prm.G + kBalanceBlock - 1) / kBalanceBlock);
}

hipError_t launch_balance(const SweepParams &prm, const double *psi, double *partial, hipStream_t s)
{
    const int grid = balance_blocks(prm);
#define BAL(D, PP) hipLaunchKernelGGL((k_balance<D, PP>), dim3(grid), dim3(kBalanceBlock), 0, s, prm, psi, partial)
    if (prm.dim == 2) {
        if (prm.p == 0) BAL(2, 0);
        else if (prm.p == 1) BAL(2, 1);
        else if (prm.p == 2) BAL(2, 2);
    } else {
        if (prm.p == 0) BAL(3, 0);
        else if (prm.p == 1) BAL(3, 1);
        else if (prm.p == 2) BAL(3, 2);
    }
#undef BAL
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The balance with a static workspace: the block partials and the result live
// in `scratch`, a file-scope device array (one copy per GPU, held until the
// process ends), so the call allocates nothing.  Its device address comes
// from hipGetSymbolAddress once per device; the array's name is a host-side
// stand-in and is never passed to a kernel or a copy.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[2 * kStaticBalanceBlocks + 2];

namespace {
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// partial[0 .. 2 nb) -> out[0] = sum of absorptions, out[1] = sum of leakages,
// one wave: lane l sums blocks l, l + 64, ... in order, then a fixed tree
__global__ void __launch_bounds__(64) k_balance_finish(const double *__restrict__ partial, int nb,
                                                       double *__restrict__ out)
{
    __shared__ double sa[64], sl[64];
    double a = 0.0, l = 0.0;
    for (int i = threadIdx.x; i < nb; i += 64) {
        a += partial[2 * i];
        l += partial[2 * i + 1];
    }
    sa[threadIdx.x] = a;
    sl[threadIdx.x] = l;
    __syncthreads();
    for (int w = 32; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) { sa[threadIdx.x] += sa[threadIdx.x + w]; sl[threadIdx.x] += sl[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = sa[0]; out[1] = sl[0]; }
}

hipError_t static_scratch(double **ptr)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *ptr = g_scratch_ptr[dev];
    return hipSuccess;
}

hipError_t balance_static(const SweepParams &prm, const double *psi, double *absorption, double *leakage,
                          hipStream_t s)
{
    const int nb = balance_blocks(prm);
    if (nb > kStaticBalanceBlocks) return hipErrorInvalidValue;
    double *ws = nullptr;
    hipError_t e = static_scratch(&ws);
    if (e != hipSuccess) return e;
    // every partial the finish kernel reads is written by launch_balance, and
    // the result slots by the finish kernel: nothing needs clearing
    e = launch_balance(prm, psi, ws, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_balance_finish, dim3(1), dim3(64), 0, s, ws, nb, ws + 2 * kStaticBalanceBlocks);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    double res[2] = {0.0, 0.0};
    e = hipMemcpyAsync(res, ws + 2 * kStaticBalanceBlocks, sizeof(res), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    *absorption = res[0];
    *leakage = res[1];
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// moments (DiscreteToMoment, src/transport_op.cpp:3-46) and block mass
// ---------------------------------------------------------------------------
// Four entries per thread, 256 apart: a workgroup reads 8 KB runs of each
// direction and keeps 16 loads in flight per thread (two and eight entries
// per thread, and one, were slower on the 64^3 S8 psi).  Each entry is summed
// over the directions in order.
__global__ void __launch_bounds__(256) k_moments(const double *__restrict__ psi, const double *__restrict__ omega,
                                                 int Nomega, long long Ndof, double *__restrict__ phi)
{
    constexpr int E = 4;
    const long long b = (long long)blockIdx.x * (256 * E) + threadIdx.x;
    const long long g = blockIdx.y;
    const double *ps = psi + g * Nomega * Ndof;
    bool ok[E];
    long long i[E];
    double acc[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
        i[k] = b + 256 * k;
        ok[k] = i[k] < Ndof;
        acc[k] = 0.0;
    }
#pragma unroll 4
    for (int a = 0; a < Nomega; ++a) {
        const double w = omega[4 * a + 3];
        const double *pa = ps + (long long)a * Ndof;
#pragma unroll
        for (int k = 0; k < E; ++k) {
            // psi is streamed once: nontemporal loads keep it from evicting the
            // L2 and MALL lines the next sweep reuses (281 -> 251 us at 64^3 S8)
            const double v = __builtin_nontemporal_load(pa + (ok[k] ? i[k] : 0));
            acc[k] = fma(w, v, acc[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < E; ++k)
        if (ok[k]) phi[g * Ndof + i[k]] = acc[k];
}

hipError_t launch_moments(const SweepParams &prm, const double *psi, double *phi, hipStream_t s)
{
    hipLaunchKernelGGL(k_moments, dim3((unsigned)((prm.Ndof + 1023) / 1024), (unsigned)prm.G), dim3(256), 0, s, psi,
                       prm.omega, prm.Nomega, prm.Ndof, phi);
    return hipGetLastError();
}

// [phi, J_0, .., J_{dim-1}] with the moment index slowest: mom[(m G + g) Ndof + s]
// (MomentVectorLayout); J_m = sum_a w_a Omega_m psi_a
__global__ void k_moments_aniso(const double *__restrict__ psi, const double *__restrict__ omega, int Nomega,
                                long long Ndof, int G, int dim, double *__restrict__ mom)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Ndof * G) return;
    const long long g = i / Ndof, s = i - g * Ndof;
    const double *ps = psi + g * Nomega * Ndof + s;
    double p0 = 0.0, j[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < Nomega; ++a) {
        const double w = omega[4 * a + 3], v = ps[(long long)a * Ndof];
        p0 = fma(w, v, p0);
        for (int m = 0; m < dim; ++m) j[m] = fma(omega[4 * a + m] * w, v, j[m]);
    }
    mom[g * Ndof + s] = p0;
    for (int m = 0; m < dim; ++m) mom[((long long)(1 + m) * G + g) * Ndof + s] = j[m];
}

// psi = phi / sum w + 3 Omega . J / sum w (DiscreteToMoment::MultTranspose with dim + 1 moments)
__global__ void k_moments_aniso_T(const double *__restrict__ mom, const double *__restrict__ omega, double wsum,
                                  int Nomega, long long Ndof, int G, int dim, double *__restrict__ psi)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Ndof * Nomega * G) return;
    const long long s = i % Ndof, a = (i / Ndof) % Nomega, g = i / (Ndof * Nomega);
    double v = mom[g * Ndof + s] / wsum;
    for (int m = 0; m < dim; ++m) v += 3 * omega[4 * a + m] * mom[((long long)(1 + m) * G + g) * Ndof + s] / wsum;
    psi[i] = v;
}

hipError_t launch_moments_aniso(const SweepParams &prm, int dim, const double *psi, double *mom, hipStream_t s)
{
    const long long n = prm.Ndof * prm.G;
    hipLaunchKernelGGL(k_moments_aniso, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, psi, prm.omega,
                       prm.Nomega, prm.Ndof, prm.G, dim, mom);
    return hipGetLastError();
}

hipError_t launch_moments_aniso_T(const SweepParams &prm, int dim, double wsum, const double *mom, double *psi,
                                  hipStream_t s)
{
    const long long n = prm.Ndof * prm.Nomega * prm.G;
    hipLaunchKernelGGL(k_moments_aniso_T, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, mom, prm.omega, wsum,
                       prm.Nomega, prm.Ndof, prm.G, dim, psi);
    return hipGetLastError();
}

__global__ void k_moments_T(const double *__restrict__ phi, double inv_wsum, int Nomega, long long Ndof, int G,
                            double *__restrict__ psi)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = Ndof * Nomega * G;
    if (i >= tot) return;
    const long long s = i % Ndof, g = i / (Ndof * Nomega);
    psi[i] = phi[g * Ndof + s] * inv_wsum;
}

hipError_t launch_moments_T(const SweepParams &prm, double inv_wsum, const double *phi, double *psi, hipStream_t s)
{
    const long long n = prm.Ndof * prm.Nomega * prm.G;
    hipLaunchKernelGGL(k_moments_T, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, phi, inv_wsum, prm.Nomega,
                       prm.Ndof, prm.G, psi);
    return hipGetLastError();
}

__global__ void k_mass(const double *__restrict__ Mm, int uniform, int n, int Ne, int G, const double *__restrict__ coef,
                       double scale, const double *__restrict__ x, double *__restrict__ y)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)Ne * G * n) return;
    const int i = (int)(t % n);
    const long long ge = t / n;
    const int e = (int)(ge % Ne);
    const double *M = uniform ? Mm : Mm + (size_t)e * n * n;
    const double *xe = x + ge * n;
    double acc = 0.0;
    for (int j = 0; j < n; ++j) acc = fma(M[i * n + j], xe[j], acc);
    y[t] = (coef[e] * acc) * scale;
}

hipError_t launch_mass(const SweepParams &prm, const double *Mm, const double *coef, double scale, const double *x,
                       double *y, hipStream_t s)
{
    const long long n = (long long)prm.Ne * prm.G * prm.n;
    hipLaunchKernelGGL(k_mass, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, Mm, prm.uniform, prm.n, prm.Ne,
                       prm.G, coef, scale, x, y);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// SI update and vector helpers
// ---------------------------------------------------------------------------
constexpr int kUpdBlock = 256;
int si_update_blocks(long n) { return (int)std::min<long>(2048, (n + kUpdBlock - 1) / kUpdBlock); }

__global__ void __launch_bounds__(kUpdBlock) k_si_update(long long n, const double *__restrict__ b,
                                                         const double *__restrict__ y, double *__restrict__ x,
                                                         double *__restrict__ partial, double *__restrict__ rout)
{
    __shared__ double sh[kUpdBlock];
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double r = y ? b[i] - (x[i] - y[i]) : b[i] - x[i];
        x[i] += r;
        if (rout) rout[i] = r;
        acc = fma(r, r, acc);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kUpdBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

hipError_t launch_si_update(long n, const double *b, const double *y, double *x, double *partial, hipStream_t s,
                            double *rout)
{
    hipLaunchKernelGGL(k_si_update, dim3(si_update_blocks(n)), dim3(kUpdBlock), 0, s, (long long)n, b, y, x, partial,
                       rout);
    return hipGetLastError();
}

__global__ void k_axpby(long long n, double a, const double *__restrict__ x, double b, const double *__restrict__ y,
                        double *__restrict__ z)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = a * x[i] + b * y[i];
}

hipError_t launch_axpby(long n, double a, const double *x, double b, const double *y, double *z, hipStream_t s)
{
    hipLaunchKernelGGL(k_axpby, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (long long)n, a, x, b, y, z);
    return hipGetLastError();
}

}  // namespace allium
