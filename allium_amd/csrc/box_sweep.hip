// Box sweeps for uniform geometry (src/sweep.cpp:366-475 restructured for a
// CDNA4 compute unit).
//
// The mesh is cut into boxes.  One workgroup sweeps one box for nd directions
// of one octant and one group: the box's cells are visited level by level
// (li + lj + lk in the octant's orientation) with a workgroup barrier between
// levels, and the face traces between the box's cells never leave LDS.  Each
// axis has one LDS plane per direction: slot (lj, lk) of the x plane holds the
// x-outflow face of the last cell swept in that row, so a cell reads its dim
// inflow traces from its three slots and overwrites them with its outflow
// traces (no two cells of one level share a slot).  Only the traces that enter
// the box through its inflow faces are read from HBM, once per box, from the
// psi written by the previous box level.
//
// Boxes on one anti-diagonal of the box grid (in the octant's orientation) do
// not depend on each other; one launch runs one box level for every octant.
// blockIdx.x = task * noct + octant, so in 3D the workgroups of octant o are
// dispatched to XCD o (MI355X hands workgroups to its 8 XCDs round-robin) and
// every trace a box reads from HBM was written through the same XCD's L2.
#include "device.hpp"
#include "kernel_common.hpp"

#include <algorithm>
#include <type_traits>

namespace allium {

bool box_sweep_supported(int dim, int p) { return (dim == 2 && p >= 0 && p <= 2) || (dim == 3 && p >= 0 && p <= 1); }

double host_unit_mass(int dim, int p, int i, int j)
{
    double v = 1.0;
    const int q = p + 1;
    for (int d = 0; d < dim; ++d) {
        const int a = i % q, b = j % q;
        v *= p == 0 ? unit_mass_1d<0>(a, b) : p == 1 ? unit_mass_1d<1>(a, b) : unit_mass_1d<2>(a, b);
        i /= q;
        j /= q;
    }
    return v;
}

// Workgroup tables in LDS: the trace planes, the directions' streaming
// matrices K[a], |Omega_ax| of each direction, the unit mass, the box's cells
// by level and the level offsets.
struct BoxLds {
    double *planes;
    const double *K;    // nd x kstride
    const double *om;   // nd x 3: |Omega_x| A_x, |Omega_y| A_y, |Omega_z| A_z (face areas A)
    const int *dirs;    // nd direction ids (-1: none)
    const uint32_t *lcl;
    const int *lpl;
    // p = 1 lane-split path: per direction the psi offset of the group and
    // direction (-1: none); per box cell in level order its element (-1:
    // outside the mesh)
    const long long *dbase;
    const int *cell_e;
};

// Inputs of one cell that do not depend on the sweep: the element, its total
// cross section and source + isotropic source.
template <int N>
struct BoxInputs {
    double q[N];
    double is[N];
    double sig;
    int e, d, a;
    int lc[3];
    bool ok;
};

template <int DIM, int P, int OCT>
__device__ __forceinline__ void box_fetch(const SweepParams &p, const BoxPlan &bp, const BoxLds &L, int ndir,
                                          const int (&lo)[3], const int (&ext)[3], int g, int l, int it,
                                          BoxInputs<Sh<DIM, P>::N> &in)
{
    constexpr int N = Sh<DIM, P>::N;
    in.ok = false;
    in.d = 0;
    in.a = 0;
    in.lc[0] = in.lc[1] = in.lc[2] = 0;
    if (l < bp.nl) {
        const int c0 = L.lpl[l], cnt = L.lpl[l + 1] - c0;
        if (it < cnt * ndir) {
            const int d = it / cnt, k = it - d * cnt;   // direction-major: few directions per wave
            const uint32_t cw = L.lcl[c0 + k];
            in.lc[0] = (int)(cw & 1023u);
            in.lc[1] = (int)((cw >> 10) & 1023u);
            in.lc[2] = (int)(cw >> 20);
            in.d = d;
            in.a = L.dirs[d];
            in.ok = in.a >= 0 && in.lc[0] < ext[0] && in.lc[1] < ext[1] && in.lc[2] < ext[2];
        }
    }
    if (!in.ok) {
#pragma unroll
        for (int i = 0; i < N; ++i) { in.q[i] = 0.0; in.is[i] = 0.0; }
        in.sig = 0.0;
        in.e = 0;
        return;
    }
    int c[3] = {0, 0, 0};
#pragma unroll
    for (int ax = 0; ax < DIM; ++ax)
        c[ax] = ((OCT >> ax) & 1) ? lo[ax] + ext[ax] - 1 - in.lc[ax] : lo[ax] + in.lc[ax];
    in.e = c[0] + p.nel[0] * (c[1] + p.nel[1] * c[2]);
    const long long off = ((long long)g * p.Nomega + in.a) * p.Ndof + (long long)in.e * N;
    in.sig = p.sigma[(long long)g * p.Ne + in.e];
    if (p.src) {
        if constexpr (N % 2 == 0) {
            const d2v *q2 = reinterpret_cast<const d2v *>(p.src + off);
#pragma unroll
            for (int i = 0; i < N / 2; ++i) { const d2v v = q2[i]; in.q[2 * i] = v.x; in.q[2 * i + 1] = v.y; }
        } else {
#pragma unroll
            for (int i = 0; i < N; ++i) in.q[i] = p.src[off + i];
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) in.q[i] = 0.0;
    }
    if (p.iso) {
        const double *is = p.iso + (long long)g * p.Ndof + (long long)in.e * N;
#pragma unroll
        for (int i = 0; i < N; ++i) in.is[i] = is[i];
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) in.is[i] = 0.0;
    }
}

// Uniform p = 1 cells by fast diagonalisation.  On an axis-aligned box every
// operator of the cell system is a Kronecker product of 1D Gauss-Lobatto
// matrices: with the exact 1D mass m = [[1/3,1/6],[1/6,1/3]] and upwind
// convection C+ = [[1/2,1/2],[-1/2,1/2]] (C- its mirror image),
//   M^{-1} A = sum_d t_d (B_d on axis d) + sigma I,  t_d = |Omega_d| / h_d,
// with B+ = m^{-1} C+ = [[3,1],[-3,1]] = W R W^{-1}, R = [[2,sqrt2],[-sqrt2,2]]
// = 2 I + sqrt2 J, W+ = [[1,0],[-1,sqrt2]], W- = P W+ (P the swap).  So
//   A psi = b  <=>  (alpha I + sum_d beta_d J_d) y = W^{-1} M^{-1} b,
//   alpha = sigma + 2 sum_d t_d,  beta_d = sqrt2 t_d,  psi = W y,
// and the commuting J_d (J^2 = -I) are diagonal after complexifying along x
// (z = y0 - i y1 makes J_x multiplication by i) and splitting y and z into
// (z0 -+ i z1): 2^(dim-1) complex divisions replace the N x N LU, and
// M^{-1} maps an inflow face term to t_d (m^{-1} e_in) (x) u_d.
template <int DIM, typename T, int OCT>
__device__ __forceinline__ void fd_solve_p1(const double (&b)[1 << DIM], const double (&u)[DIM][1 << (DIM - 1)],
                                            const double (&t)[3], double sig, double inv_vol, T (&x)[1 << DIM])
{
    constexpr int N = 1 << DIM;
    constexpr T r2 = (T)1.4142135623730951, ir2 = (T)0.70710678118654752;
    T r[N];
#pragma unroll
    for (int i = 0; i < N; ++i) r[i] = (T)(b[i] * inv_vol);
    // pairs along axis ax: (i, i + (1 << ax)) for i with bit ax clear
    auto pass = [&](int ax, T a00, T a01, T a10, T a11) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            if (i & (1 << ax)) continue;
            const T v0 = r[i], v1 = r[i + (1 << ax)];
            r[i] = a00 * v0 + a01 * v1;
            r[i + (1 << ax)] = a10 * v0 + a11 * v1;
        }
    };
#pragma unroll
    for (int ax = 0; ax < DIM; ++ax) pass(ax, (T)4, (T)-2, (T)-2, (T)4);
    // inflow faces: r += t_ax (m^{-1} e_in) (x) u_ax; face index m runs over
    // the other axes, the lower one fastest
#pragma unroll
    for (int ax = 0; ax < DIM; ++ax) {
        const bool neg = (OCT >> ax) & 1;
        const T c0 = neg ? (T)-2 : (T)4, c1 = neg ? (T)4 : (T)-2;
        const T ta = (T)t[ax];
#pragma unroll
        for (int m = 0; m < (1 << (DIM - 1)); ++m) {
            int i = 0, k = 0;
#pragma unroll
            for (int o = 0; o < DIM; ++o)
                if (o != ax) { i |= ((m >> k) & 1) << o; ++k; }
            const T v = ta * (T)u[ax][m];
            r[i] += c0 * v;
            r[i | (1 << ax)] += c1 * v;
        }
    }
    // y = W^{-1} r: W+^{-1} = [[1,0],[1/sqrt2,1/sqrt2]], W-^{-1} = [[0,1],[1/sqrt2,1/sqrt2]]
#pragma unroll
    for (int ax = 0; ax < DIM; ++ax) {
        if ((OCT >> ax) & 1) pass(ax, (T)0, (T)1, ir2, ir2);
        else pass(ax, (T)1, (T)0, ir2, ir2);
    }
    T alpha = (T)sig, beta[3];
#pragma unroll
    for (int ax = 0; ax < DIM; ++ax) {
        alpha += (T)2 * (T)t[ax];
        beta[ax] = r2 * (T)t[ax];
    }
    // complex z[j] = y[2j] - i y[2j+1] over the remaining axes (j = i1 + 2 i2)
    constexpr int NZ = N / 2;
    T zr[NZ], zi[NZ];
#pragma unroll
    for (int j = 0; j < NZ; ++j) { zr[j] = r[2 * j]; zi[j] = -r[2 * j + 1]; }
    // split axis k (bit k of j): (z0, z1) -> (p, m) = (z0 - i z1, z0 + i z1)
    auto split = [&](int k) {
#pragma unroll
        for (int j = 0; j < NZ; ++j) {
            if (j & (1 << k)) continue;
            const int j1 = j | (1 << k);
            const T ar = zr[j], ai = zi[j], br = zr[j1], bi = zi[j1];
            zr[j] = ar + bi; zi[j] = ai - br;
            zr[j1] = ar - bi; zi[j1] = ai + br;
        }
    };
    // inverse: z0 = (p + m) / 2, z1 = i (p - m) / 2
    auto merge = [&](int k) {
#pragma unroll
        for (int j = 0; j < NZ; ++j) {
            if (j & (1 << k)) continue;
            const int j1 = j | (1 << k);
            const T pr = zr[j], pi = zi[j], mr = zr[j1], mi = zi[j1];
            zr[j] = (T)0.5 * (pr + mr); zi[j] = (T)0.5 * (pi + mi);
            zr[j1] = (T)0.5 * (mi - pi); zi[j1] = (T)0.5 * (pr - mr);
        }
    };
#pragma unroll
    for (int k = 0; k < DIM - 1; ++k) split(k);
    // eigenvalues alpha + i (beta_x +- beta_y +- beta_z): bit k of j set -> minus on axis k+1
#pragma unroll
    for (int j = 0; j < NZ; ++j) {
        T li = beta[0];
#pragma unroll
        for (int k = 0; k < DIM - 1; ++k) li += ((j >> k) & 1) ? -beta[k + 1] : beta[k + 1];
        const T inv = (T)1 / (alpha * alpha + li * li);
        const T ar = zr[j], ai = zi[j];
        zr[j] = (ar * alpha + ai * li) * inv;
        zi[j] = (ai * alpha - ar * li) * inv;
    }
#pragma unroll
    for (int k = DIM - 2; k >= 0; --k) merge(k);
#pragma unroll
    for (int j = 0; j < NZ; ++j) { r[2 * j] = zr[j]; r[2 * j + 1] = -zi[j]; }
    // psi = W y: W+ = [[1,0],[-1,sqrt2]], W- = [[-1,sqrt2],[1,0]]
#pragma unroll
    for (int ax = 0; ax < DIM; ++ax) {
        if ((OCT >> ax) & 1) pass(ax, (T)-1, r2, (T)1, (T)0);
        else pass(ax, (T)1, (T)0, (T)-1, r2);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = r[i];
}

// psi of a solved cell whose store is deferred.  gfx950 counts vector memory
// loads and stores in one in-order counter, so a store issued before the next
// cell's loads would make waiting for those loads wait for the store too;
// storing after the next loads are in flight keeps stores off the critical path.
template <int N>
struct PendingPsi {
    double v[N];
    double *dst = nullptr;
    __device__ __forceinline__ void flush()
    {
        if (!dst) return;
        if constexpr (N % 2 == 0) {
            d2v *o2 = reinterpret_cast<d2v *>(dst);
#pragma unroll
            for (int i = 0; i < N / 2; ++i) { d2v w; w.x = v[2 * i]; w.y = v[2 * i + 1]; o2[i] = w; }
        } else {
#pragma unroll
            for (int i = 0; i < N; ++i) dst[i] = v[i];
        }
        dst = nullptr;
    }
};

// Solve one cell whose inputs are loaded: p = 1 by fast diagonalisation;
// otherwise traces and K[a] from LDS, the mass and face-mass matrices as
// compile-time tensor products (uniform axis-aligned boxes: Mu = vol *
// unit_mass, F_side = area * unit_face_mass), and an LU; store
// the outflow traces to the planes and leave psi in `pend` (only when in.ok);
// the previous cell's pending psi is stored once this cell's inputs are in.
// mid() runs between the assembly of the local system and its solve (the next
// level's loads go there: issued after this cell's inputs were consumed and
// before the pending psi store, so their latency overlaps the solve).
template <int DIM, int P, typename T, int OCT, typename Mid>
__device__ __forceinline__ void box_solve(const SweepParams &p, const BoxPlan &bp, const BoxLds &L,
                                          const BoxInputs<Sh<DIM, P>::N> &in, int g, Mid &&mid,
                                          PendingPsi<Sh<DIM, P>::N> &pend)
{
    using S = Sh<DIM, P>;
    constexpr int N = S::N, NF = S::NF, PS = box_slot_stride(S::NF);
    const int a = in.a, d = in.d;
    double *slot[DIM];
    double u[DIM][NF];
    {
        double *pl = L.planes;
#pragma unroll
        for (int ax = 0; ax < DIM; ++ax) {
            // slot of this cell in the axis-ax plane: the other axes' local coordinates
            int s = 0, mul = 1;
#pragma unroll
            for (int o = 0; o < DIM; ++o)
                if (o != ax) { s += mul * in.lc[o]; mul *= bp.bs[o]; }
            slot[ax] = pl + ((long long)d * bp.psz[ax] + s) * PS;
            pl += (long long)bp.nd * bp.psz[ax] * PS;
#pragma unroll
            for (int m = 0; m < NF; ++m) u[ax][m] = slot[ax][m];
        }
    }

    T x[N];
    if constexpr (P == 1) {
        double b[N];
#pragma unroll
        for (int i = 0; i < N; ++i) b[i] = in.q[i] + in.is[i];
        const double inv_vol = 1.0 / bp.vol;
        const double t[3] = {L.om[3 * d] * inv_vol, L.om[3 * d + 1] * inv_vol, L.om[3 * d + 2] * inv_vol};
        mid();
        pend.flush();
        fd_solve_p1<DIM, T, OCT>(b, u, t, in.sig + (p.timedep ? p.tau : 0.0), inv_vol, x);
    } else {
    // ---- right-hand side: q + iso + |Omega_ax| F_in u -----------------------
    double b[N];
#pragma unroll
    for (int i = 0; i < N; ++i) b[i] = in.q[i] + in.is[i];
#pragma unroll
    for (int ax = 0; ax < DIM; ++ax) {
        const int s_in = ((OCT >> ax) & 1) ? 2 * ax + 1 : 2 * ax;
        const double w = L.om[3 * d + ax];
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            double acc = 0.0;
#pragma unroll
            for (int l = 0; l < NF; ++l) acc = fma(unit_face_mass<DIM, P>(k, l), u[ax][l], acc);
            b[side_node<DIM, P>(s_in, k)] = fma(w, acc, b[side_node<DIM, P>(s_in, k)]);
        }
    }

    // ---- matrix K[a] (includes tau Mt) + sigma Mu --------------------------
    const double *Kl = L.K + d * bp.kstride;
    const double sv = in.sig * bp.vol;
    T A[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) A[i][j] = (T)fma(sv, unit_mass<DIM, P>(i, j), Kl[i * N + j]);
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = (T)b[i];
    mid();
    pend.flush();
    lu_factor_solve<N, T>(A, x);
    }

    // ---- store psi and the outflow traces ----------------------------------
    const long long off = ((long long)g * p.Nomega + a) * p.Ndof + (long long)in.e * N;
#pragma unroll
    for (int i = 0; i < N; ++i) pend.v[i] = (double)x[i];
    pend.dst = p.psi + off;
#pragma unroll
    for (int ax = 0; ax < DIM; ++ax) {
        const int s_out = ((OCT >> ax) & 1) ? 2 * ax : 2 * ax + 1;
#pragma unroll
        for (int m = 0; m < NF; ++m) slot[ax][m] = (double)x[side_node<DIM, P>(s_out, m)];
    }
}

// Workgroup barrier that orders LDS only.  Inside a box the cells exchange
// traces through LDS alone: psi goes to HBM for the next box level (a later
// launch) and q, the isotropic source and sigma are read-only, so the barrier
// need not wait for outstanding global loads (the prefetched next level) or
// stores (this level's psi), which __syncthreads() would.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// PF: 0 loads each cell's inputs when it is solved; 2 loads the next level's
// first-pass inputs during the current solve (between assembly and LU).
// Everything a box workgroup sets up before its levels: the box's extent,
// its directions, the LDS tables and the inflow trace planes.
struct BoxCtx {
    BoxLds L;
    int lo[3], ext[3];
    int g, ndir;
};

template <int DIM, int P, int OCT>
__device__ __forceinline__ BoxCtx box_stage(const SweepParams &p, const BoxPlan &bp, int box_base, double *lds)
{
    using S = Sh<DIM, P>;
    constexpr int N = S::N, NF = S::NF, PS = box_slot_stride(S::NF);
    const int task = blockIdx.x / bp.noct;
    const int g = blockIdx.y / bp.nchunk, chunk = blockIdx.y - g * bp.nchunk;
    const int tid = threadIdx.x, nt = blockDim.x;
    const uint32_t bc = bp.boxes[box_base + task];
    const int ob[3] = {(int)(bc & 1023u), (int)((bc >> 10) & 1023u), (int)(bc >> 20)};
    int lo[3] = {0, 0, 0}, ext[3] = {1, 1, 1};
#pragma unroll
    for (int ax = 0; ax < DIM; ++ax) {
        const int bi = ((OCT >> ax) & 1) ? bp.nb[ax] - 1 - ob[ax] : ob[ax];
        lo[ax] = bi * bp.bs[ax];
        ext[ax] = min(bp.bs[ax], p.nel[ax] - lo[ax]);
    }
    const int *gdirs = bp.oct_dirs + OCT * bp.maxoct + chunk * bp.nd;
    const int ndir = min(bp.nd, bp.maxoct - chunk * bp.nd);
    int nplane = 0;
#pragma unroll
    for (int ax = 0; ax < DIM; ++ax) nplane += bp.psz[ax];
    // LDS layout (doubles first, then 32-bit tables)
    double *Kw = lds + (long long)bp.nd * nplane * PS;
    double *omw = Kw + (long long)bp.nd * bp.kstride;
    long long *dbw = reinterpret_cast<long long *>(omw + 3 * bp.nd);
    int *dirw = reinterpret_cast<int *>(dbw + (P == 1 ? bp.nd : 0));
    uint32_t *lcw = reinterpret_cast<uint32_t *>(dirw + bp.nd);
    const int vol = bp.bs[0] * bp.bs[1] * bp.bs[2];
    int *lpw = reinterpret_cast<int *>(lcw + vol);
    int *cew = lpw + bp.nl + 1;
    const BoxLds L{lds, Kw, omw, dirw, lcw, lpw, dbw, cew};

    // ---- stage the workgroup tables ----------------------------------------
    if constexpr (P != 1)
        for (int i = tid; i < ndir * N * N; i += nt) {
            const int d = i / (N * N), r = i - d * N * N;
            const int a = gdirs[d];
            Kw[d * bp.kstride + r] = a >= 0 ? p.K[(long long)a * N * N + r] : 0.0;
        }
    for (int i = tid; i < bp.nd; i += nt) {
        const int a = i < ndir ? gdirs[i] : -1;
        dirw[i] = a;
        for (int ax = 0; ax < 3; ++ax)
            omw[3 * i + ax] = (a >= 0 && ax < DIM) ? fabs(p.omega[4 * a + ax]) * bp.farea[ax] : 0.0;
    }
    if constexpr (P != 1)
        for (int i = tid; i < vol; i += nt) lcw[i] = bp.lcells[i];
    for (int i = tid; i <= bp.nl; i += nt) lpw[i] = bp.lptr[i];
    if constexpr (P == 1) {
        for (int i = tid; i < bp.nd; i += nt) {
            const int a = i < ndir ? gdirs[i] : -1;
            dbw[i] = a >= 0 ? ((long long)g * p.Nomega + a) * p.Ndof : -1;
        }
        // per box cell: its element, and its slot in each axis's trace plane
        // (the other axes' local coordinates), packed 10 bits per axis, so the
        // cell loop computes no plane index
        for (int i = tid; i < vol; i += nt) {
            const uint32_t cw = bp.lcells[i];
            const int lc[3] = {(int)(cw & 1023u), (int)((cw >> 10) & 1023u), (int)(cw >> 20)};
            int c[3] = {0, 0, 0};
            bool in_box = true;
            uint32_t sw = 0;
#pragma unroll
            for (int ax = 0; ax < DIM; ++ax) {
                in_box = in_box && lc[ax] < ext[ax];
                c[ax] = ((OCT >> ax) & 1) ? lo[ax] + ext[ax] - 1 - lc[ax] : lo[ax] + lc[ax];
                int sidx = 0, mul = 1;
#pragma unroll
                for (int o = 0; o < DIM; ++o)
                    if (o != ax) { sidx += mul * lc[o]; mul *= bp.bs[o]; }
                sw |= (uint32_t)sidx << (10 * ax);
            }
            cew[i] = in_box ? c[0] + p.nel[0] * (c[1] + p.nel[1] * c[2]) : -1;
            lcw[i] = sw;
        }
    }
    // ---- inflow planes from the upwind neighbours (zero on the boundary) ---
    {
        double *pl = lds;
#pragma unroll
        for (int ax = 0; ax < DIM; ++ax) {
            const bool neg = (OCT >> ax) & 1;
            const int s_out = neg ? 2 * ax : 2 * ax + 1;   // neighbour's side facing this box
            const int cax = neg ? lo[ax] + ext[ax] : lo[ax] - 1;
            const bool in_dom = cax >= 0 && cax < p.nel[ax];
            const int np = bp.psz[ax];
            for (int i = tid; i < ndir * np; i += nt) {
                const int d = i / np, s = i - d * np;
                double *dst = pl + ((long long)d * np + s) * PS;
                int c[3] = {0, 0, 0}, rem = s;
                const int a = gdirs[d];
                bool valid = in_dom && a >= 0;
#pragma unroll
                for (int o = 0; o < DIM; ++o) {
                    if (o == ax) { c[o] = cax; continue; }
                    const int l = rem % bp.bs[o];
                    rem /= bp.bs[o];
                    valid = valid && l < ext[o];
                    c[o] = ((OCT >> o) & 1) ? lo[o] + ext[o] - 1 - l : lo[o] + l;
                }
                if (valid) {
                    const int e = c[0] + p.nel[0] * (c[1] + p.nel[1] * c[2]);
                    const double *src = p.psi + ((long long)g * p.Nomega + a) * p.Ndof + (long long)e * N;
#pragma unroll
                    for (int m = 0; m < NF; ++m) dst[m] = src[side_node<DIM, P>(s_out, m)];
                } else {
#pragma unroll
                    for (int m = 0; m < NF; ++m) dst[m] = 0.0;
                }
            }
            pl += (long long)bp.nd * np * PS;
        }
    }
    lds_barrier();
    return BoxCtx{L, {lo[0], lo[1], lo[2]}, {ext[0], ext[1], ext[2]}, g, ndir};
}

template <int DIM, int P, typename T, int OCT, int PF>
__device__ __forceinline__ void box_task(const SweepParams &p, const BoxPlan &bp, int box_base, double *lds)
{
    using S = Sh<DIM, P>;
    constexpr int N = S::N;
    const BoxCtx cx = box_stage<DIM, P, OCT>(p, bp, box_base, lds);
    const BoxLds &L = cx.L;
    const int g = cx.g, ndir = cx.ndir;
    const int(&lo)[3] = cx.lo;
    const int(&ext)[3] = cx.ext;
    const int tid = threadIdx.x, nt = blockDim.x;

    // ---- the box's levels ---------------------------------------------------
    auto nothing = [] {};
    PendingPsi<N> pend;
    if constexpr (PF == 0) {
        for (int l = 0; l < bp.nl; ++l) {
            const int items = (L.lpl[l + 1] - L.lpl[l]) * ndir;
            for (int it = tid; it < items; it += nt) {
                BoxInputs<N> in;
                box_fetch<DIM, P, OCT>(p, bp, L, ndir, lo, ext, g, l, it, in);
                if (in.ok) box_solve<DIM, P, T, OCT>(p, bp, L, in, g, nothing, pend);
            }
            lds_barrier();
        }
    } else {
        // two input buffers with alternating roles (no register copy between
        // levels, so the prefetched loads stay in flight across the barrier)
        auto level = [&](int l, BoxInputs<N> &cur, BoxInputs<N> &nxt) {
            const int items = (L.lpl[l + 1] - L.lpl[l]) * ndir;
            auto pre = [&] { box_fetch<DIM, P, OCT>(p, bp, L, ndir, lo, ext, g, l + 1, tid, nxt); };
            if (cur.ok) box_solve<DIM, P, T, OCT>(p, bp, L, cur, g, pre, pend);
            else pre();
            for (int it = tid + nt; it < items; it += nt) {
                BoxInputs<N> in;
                box_fetch<DIM, P, OCT>(p, bp, L, ndir, lo, ext, g, l, it, in);
                if (in.ok) box_solve<DIM, P, T, OCT>(p, bp, L, in, g, nothing, pend);
            }
            lds_barrier();
        };
        BoxInputs<N> b0, b1;
        box_fetch<DIM, P, OCT>(p, bp, L, ndir, lo, ext, g, 0, tid, b0);
        for (int l = 0; l < bp.nl; l += 2) {
            level(l, b0, b1);
            if (l + 1 < bp.nl) level(l + 1, b1, b0);
        }
    }
    pend.flush();
}

// ---------------------------------------------------------------------------
// p = 1 with 2^(dim-1) lanes per cell.  Lane r of a cell holds its two nodes
// along x, (0, r) and (1, r) with r = i_y + 2 i_z, which are adjacent in
// memory: a cell's q, isotropic source and psi are one contiguous 64-byte
// (3D) access shared by its lanes instead of 64 cache lines per wave-wide
// access.  The fast diagonalisation above runs on the lanes: the x passes in
// registers, the y and z passes and the complex splits as swaps inside
// quads (DPP quad_perm), the division lane by lane.
// ---------------------------------------------------------------------------
template <typename T, int CTRL>
__device__ __forceinline__ T quad_swap(T x)
{
    if constexpr (sizeof(T) == 8) {
        const long long v = __builtin_bit_cast(long long, x);
        const int lo = __builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_mov_dpp((int)(v >> 32), CTRL, 0xF, 0xF, false);
        return __builtin_bit_cast(T, ((long long)hi << 32) | (unsigned int)lo);
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), CTRL, 0xF, 0xF, false));
    }
}
constexpr int kSwapY = 0xB1;    // quad_perm [1,0,3,2]: partner along y (lane bit 0)
constexpr int kSwapZ = 0x4E;    // quad_perm [2,3,0,1]: partner along z (lane bit 1)
constexpr int kSwapYZ = 0x1B;   // quad_perm [3,2,1,0]: partner along y and z

// A direction's constants of the lane-split solve (t_d = |Omega_d| area_d / vol
// on the unit cell): computed once per workgroup when it sweeps one direction.
template <typename T>
struct DirConst {
    T t[3];
    T s2t;        // 2 (t_x + t_y + t_z)
    T bx, by, bz; // sqrt(2) t_d
    T q;          // by^2 + bz^2 - bx^2
    T b2;         // 2 bx
    T c, c2;      // 2 by bz and its square
    // signed by this lane's node bits (constant per lane): +by on a bit-0 lane
    // along y, -by on a bit-1 lane; the same along z; +c where the bits agree
    T syb, szb, cs;
    T wiy[2], wiz[2];   // the lane's W_y^{-1} and W_z^{-1} rows: v <- [0] v + [1] partner
};
template <int DIM, typename T>
__device__ __forceinline__ DirConst<T> dir_const(const T (&t)[3])
{
    constexpr T r2 = (T)1.4142135623730951;
    DirConst<T> k;
    k.t[0] = t[0];
    k.t[1] = t[1];
    k.t[2] = DIM == 3 ? t[2] : (T)0;
    k.s2t = (T)2 * (k.t[0] + k.t[1] + k.t[2]);
    k.bx = r2 * k.t[0];
    k.by = r2 * k.t[1];
    k.bz = r2 * k.t[2];
    k.q = k.by * k.by + k.bz * k.bz - k.bx * k.bx;
    k.b2 = (T)2 * k.bx;
    k.c = (T)2 * k.by * k.bz;
    k.c2 = k.c * k.c;
    return k;
}

// Operators along y and z act between the lanes of a cell: lane bit b is the
// node index on that axis, and the partner lane (quad_perm swap) holds the
// other node.  J = [[0,1],[-1,0]] is the partner's value with sign + on the
// bit-0 lane and - on the bit-1 lane.

// psi on this lane's nodes (0, r), (1, r) from (v0, v1) = b / vol there:
// r = M^{-1} b + inflow terms, y = W^{-1} r, then
//   S^{-1} = (B^2 - c^2)^{-1} (B + c K) (A_x - beta_y J_y - beta_z J_z),
// with A_x = alpha + beta_x J_x, B = A_x^2 + beta_y^2 + beta_z^2,
// c = 2 beta_y beta_z and K = J_y J_z (commuting, J^2 = -I, K^2 = I), so B and
// B^2 - c^2 act inside the lane's x pair; psi = W y.
// The right-hand side of the lane's x pair, v = M^{-1} b + sum_d t_d (m^{-1} e_in) (x) u_d,
// from (v0, v1) = b / vol there.
template <int DIM, typename T, int OCT>
__device__ __forceinline__ void fd_split_rhs(T &v0, T &v1, const T (&t)[3], T ux, const T (&uy)[2], const T (&uz)[2],
                                             bool hy, bool hz)
{
    constexpr bool nx = OCT & 1, ny = (OCT >> 1) & 1, nz = (OCT >> 2) & 1;
    // M^{-1}: m^{-1} = [[4,-2],[-2,4]] on every axis
    {
        const T w0 = v0, w1 = v1;
        v0 = (T)4 * w0 - (T)2 * w1;
        v1 = (T)4 * w1 - (T)2 * w0;
    }
    {
        const T p0 = quad_swap<T, kSwapY>(v0), p1 = quad_swap<T, kSwapY>(v1);
        v0 = (T)4 * v0 - (T)2 * p0;
        v1 = (T)4 * v1 - (T)2 * p1;
    }
    if constexpr (DIM == 3) {
        const T p0 = quad_swap<T, kSwapZ>(v0), p1 = quad_swap<T, kSwapZ>(v1);
        v0 = (T)4 * v0 - (T)2 * p0;
        v1 = (T)4 * v1 - (T)2 * p1;
    }
    // inflow faces: t_ax (m^{-1} e_in) (x) u_ax; (m^{-1} e_in) = 4 on the inflow node, -2 on the other
    v0 += t[0] * (nx ? (T)-2 : (T)4) * ux;
    v1 += t[0] * (nx ? (T)4 : (T)-2) * ux;
    {
        const T wy = t[1] * ((hy == ny) ? (T)4 : (T)-2);
        v0 += wy * uy[0];
        v1 += wy * uy[1];
    }
    if constexpr (DIM == 3) {
        const T wz = t[2] * ((hz == nz) ? (T)4 : (T)-2);
        v0 += wz * uz[0];
        v1 += wz * uz[1];
    }
}

// (sigma I + sum_d t_d B_d) psi = v for the lane's x pair: W^{-1}, S^{-1}, W (above)
template <int DIM, typename T, int OCT>
__device__ __forceinline__ void fd_split_solve(T &v0, T &v1, const T (&t)[3], T sig, bool hy, bool hz)
{
    constexpr T r2 = (T)1.4142135623730951, ir2 = (T)0.70710678118654752;
    constexpr bool nx = OCT & 1, ny = (OCT >> 1) & 1, nz = (OCT >> 2) & 1;
    // W^{-1}: W+^{-1} = [[1,0],[1/r2,1/r2]], W-^{-1} = [[0,1],[1/r2,1/r2]]
    {
        const T w0 = v0, w1 = v1;
        v0 = nx ? w1 : w0;
        v1 = ir2 * (w0 + w1);
    }
    auto winv = [&](auto swp, bool hi, bool neg) {
        const T p0 = swp(v0), p1 = swp(v1);
        const T a0 = ir2 * (v0 + p0), a1 = ir2 * (v1 + p1);
        v0 = hi ? a0 : (neg ? p0 : v0);
        v1 = hi ? a1 : (neg ? p1 : v1);
    };
    winv([](T x) { return quad_swap<T, kSwapY>(x); }, hy, ny);
    if constexpr (DIM == 3) winv([](T x) { return quad_swap<T, kSwapZ>(x); }, hz, nz);
    // S^{-1}
    T alpha = sig + (T)2 * (t[0] + t[1]);
    if constexpr (DIM == 3) alpha += (T)2 * t[2];
    const T bx = r2 * t[0], by = r2 * t[1], bz = DIM == 3 ? r2 * t[2] : (T)0;
    T u0 = alpha * v0 + bx * v1, u1 = alpha * v1 - bx * v0;   // A_x v
    {
        // - beta_y J_y v: J_y v = +partner on the bit-0 lane, -partner on the bit-1 lane
        const T gy0 = by * quad_swap<T, kSwapY>(v0), gy1 = by * quad_swap<T, kSwapY>(v1);
        u0 = hy ? u0 + gy0 : u0 - gy0;
        u1 = hy ? u1 + gy1 : u1 - gy1;
        if constexpr (DIM == 3) {
            const T gz0 = bz * quad_swap<T, kSwapZ>(v0), gz1 = bz * quad_swap<T, kSwapZ>(v1);
            u0 = hz ? u0 + gz0 : u0 - gz0;
            u1 = hz ? u1 + gz1 : u1 - gz1;
        }
    }
    const T a = alpha * alpha - bx * bx + by * by + bz * bz, b = (T)2 * alpha * bx;
    if constexpr (DIM == 3) {
        // K u = s_y s_z u(partner along y and z): + when the bits agree
        const T c = (T)2 * by * bz;
        const T k0 = c * quad_swap<T, kSwapYZ>(u0), k1 = c * quad_swap<T, kSwapYZ>(u1);
        const bool same = hy == hz;
        const T w0 = a * u0 + b * u1 + (same ? k0 : -k0);
        const T w1 = a * u1 - b * u0 + (same ? k1 : -k1);
        const T e = a * a - b * b - c * c, f = (T)2 * a * b;
        const T inv = fast_rcp(e * e + f * f);
        v0 = (e * w0 - f * w1) * inv;
        v1 = (e * w1 + f * w0) * inv;
    } else {
        const T inv = fast_rcp(a * a + b * b);
        v0 = (a * u0 - b * u1) * inv;
        v1 = (a * u1 + b * u0) * inv;
    }
    // W: W+ = [[1,0],[-1,r2]], W- = [[-1,r2],[1,0]]
    {
        const T w0 = v0, w1 = v1;
        v0 = nx ? -w0 + r2 * w1 : w0;
        v1 = nx ? w0 : -w0 + r2 * w1;
    }
    auto wfwd = [&](auto swp, bool hi, bool neg) {
        const T p0 = swp(v0), p1 = swp(v1);
        if (neg) {
            v0 = hi ? p0 : r2 * p0 - v0;
            v1 = hi ? p1 : r2 * p1 - v1;
        } else {
            v0 = hi ? r2 * v0 - p0 : v0;
            v1 = hi ? r2 * v1 - p1 : v1;
        }
    };
    wfwd([](T x) { return quad_swap<T, kSwapY>(x); }, hy, ny);
    if constexpr (DIM == 3) wfwd([](T x) { return quad_swap<T, kSwapZ>(x); }, hz, nz);
}

// fd_split_solve with the direction's constants precomputed (one direction per
// workgroup): the same arithmetic less the per-cell products of t_d
template <int DIM, typename T, int OCT>
__device__ __forceinline__ void fd_split_solve_k(T &v0, T &v1, const DirConst<T> &K, T sig, bool hy, bool hz)
{
    constexpr T r2 = (T)1.4142135623730951, ir2 = (T)0.70710678118654752;
    constexpr bool nx = OCT & 1, ny = (OCT >> 1) & 1, nz = (OCT >> 2) & 1;
    // W^{-1}: W+^{-1} = [[1,0],[1/r2,1/r2]], W-^{-1} = [[0,1],[1/r2,1/r2]]
    {
        const T w0 = v0, w1 = v1;
        v0 = nx ? w1 : w0;
        v1 = ir2 * (w0 + w1);
    }
    // W_y^{-1}, W_z^{-1} between the lanes: v <- cA v + cB partner, with the
    // lane's (cA, cB) = (1/r2, 1/r2) on a bit-1 lane, else (0, 1) or (1, 0)
    auto winv = [&](auto swp, T cA, T cB) {
        const T p0 = swp(v0), p1 = swp(v1);
        v0 = fma(cB, p0, cA * v0);
        v1 = fma(cB, p1, cA * v1);
    };
    winv([](T x) { return quad_swap<T, kSwapY>(x); }, K.wiy[0], K.wiy[1]);
    if constexpr (DIM == 3) winv([](T x) { return quad_swap<T, kSwapZ>(x); }, K.wiz[0], K.wiz[1]);
    // S^{-1}
    const T alpha = sig + K.s2t;
    const T bx = K.bx;
    T u0 = alpha * v0 + bx * v1, u1 = alpha * v1 - bx * v0;   // A_x v
    // - beta_y J_y v - beta_z J_z v, the signs carried by the lane's constants
    u0 = fma(K.syb, quad_swap<T, kSwapY>(v0), u0);
    u1 = fma(K.syb, quad_swap<T, kSwapY>(v1), u1);
    if constexpr (DIM == 3) {
        u0 = fma(K.szb, quad_swap<T, kSwapZ>(v0), u0);
        u1 = fma(K.szb, quad_swap<T, kSwapZ>(v1), u1);
    }
    const T a = fma(alpha, alpha, K.q), b = alpha * K.b2;
    if constexpr (DIM == 3) {
        // K u = s_y s_z u(partner along y and z): + when the bits agree
        const T w0 = fma(K.cs, quad_swap<T, kSwapYZ>(u0), a * u0 + b * u1);
        const T w1 = fma(K.cs, quad_swap<T, kSwapYZ>(u1), a * u1 - b * u0);
        const T e = a * a - b * b - K.c2, f = (T)2 * a * b;
        const T inv = fast_rcp(e * e + f * f);
        v0 = (e * w0 - f * w1) * inv;
        v1 = (e * w1 + f * w0) * inv;
    } else {
        const T inv = fast_rcp(a * a + b * b);
        v0 = (a * u0 - b * u1) * inv;
        v1 = (a * u1 + b * u0) * inv;
    }
    // W: W+ = [[1,0],[-1,r2]], W- = [[-1,r2],[1,0]]
    {
        const T w0 = v0, w1 = v1;
        v0 = nx ? -w0 + r2 * w1 : w0;
        v1 = nx ? w0 : -w0 + r2 * w1;
    }
    auto wfwd = [&](auto swp, bool hi, bool neg) {
        const T p0 = swp(v0), p1 = swp(v1);
        if (neg) {
            v0 = hi ? p0 : r2 * p0 - v0;
            v1 = hi ? p1 : r2 * p1 - v1;
        } else {
            v0 = hi ? r2 * v0 - p0 : v0;
            v1 = hi ? r2 * v1 - p1 : v1;
        }
    };
    wfwd([](T x) { return quad_swap<T, kSwapY>(x); }, hy, ny);
    if constexpr (DIM == 3) wfwd([](T x) { return quad_swap<T, kSwapZ>(x); }, hz, nz);
}

// fd_split_rhs for a source already scaled by 2^dim / vol: m^{-1} = 2 m~ with
// m~ = [[2,-1],[-1,2]] on every axis, the 2s folded into that scale, so each
// axis is one FMA per value
template <int DIM, typename T, int OCT>
__device__ __forceinline__ void fd_split_rhs_h(T &v0, T &v1, const T (&t)[3], T ux, const T (&uy)[2],
                                               const T (&uz)[2], bool hy, bool hz)
{
    constexpr bool nx = OCT & 1, ny = (OCT >> 1) & 1, nz = (OCT >> 2) & 1;
    {
        const T w0 = v0, w1 = v1;
        v0 = fma((T)2, w0, -w1);
        v1 = fma((T)2, w1, -w0);
    }
    {
        const T p0 = quad_swap<T, kSwapY>(v0), p1 = quad_swap<T, kSwapY>(v1);
        v0 = fma((T)2, v0, -p0);
        v1 = fma((T)2, v1, -p1);
    }
    if constexpr (DIM == 3) {
        const T p0 = quad_swap<T, kSwapZ>(v0), p1 = quad_swap<T, kSwapZ>(v1);
        v0 = fma((T)2, v0, -p0);
        v1 = fma((T)2, v1, -p1);
    }
    // inflow faces as in fd_split_rhs
    v0 += t[0] * (nx ? (T)-2 : (T)4) * ux;
    v1 += t[0] * (nx ? (T)4 : (T)-2) * ux;
    {
        const T wy = t[1] * ((hy == ny) ? (T)4 : (T)-2);
        v0 += wy * uy[0];
        v1 += wy * uy[1];
    }
    if constexpr (DIM == 3) {
        const T wz = t[2] * ((hz == nz) ? (T)4 : (T)-2);
        v0 += wz * uz[0];
        v1 += wz * uz[1];
    }
}

template <int DIM, typename T, int OCT>
__device__ __forceinline__ void fd_split_p1_k(T &v0, T &v1, const DirConst<T> &K, T sig, T ux, const T (&uy)[2],
                                              const T (&uz)[2], bool hy, bool hz)
{
    fd_split_rhs_h<DIM, T, OCT>(v0, v1, K.t, ux, uy, uz, hy, hz);
    fd_split_solve_k<DIM, T, OCT>(v0, v1, K, sig, hy, hz);
}

template <int DIM, typename T, int OCT>
__device__ __forceinline__ void fd_split_p1(T &v0, T &v1, const T (&t)[3], T sig, T ux, const T (&uy)[2],
                                            const T (&uz)[2], bool hy, bool hz)
{
    fd_split_rhs<DIM, T, OCT>(v0, v1, t, ux, uy, uz, hy, hz);
    fd_split_solve<DIM, T, OCT>(v0, v1, t, sig, hy, hz);
}

// (sigma I + sum_d t_d B_d) x for the lane's x pair (x0, x1), with the 1D upwind
// operators B+ = m^{-1} C+ = [[3, 1], [-3, 1]] and B- = [[1, -3], [1, 3]]: on
// the upwind node 3 own + other, on the downwind node own - 3 other
template <int DIM, int OCT>
__device__ __forceinline__ void fd_split_apply(double x0, double x1, const double (&t)[3], double sig, bool hy,
                                               bool hz, double &y0, double &y1)
{
    constexpr bool nx = OCT & 1, ny = (OCT >> 1) & 1, nz = (OCT >> 2) & 1;
    y0 = sig * x0 + t[0] * (nx ? x0 - 3.0 * x1 : 3.0 * x0 + x1);
    y1 = sig * x1 + t[0] * (nx ? 3.0 * x1 + x0 : x1 - 3.0 * x0);
    {
        const double p0 = quad_swap<double, kSwapY>(x0), p1 = quad_swap<double, kSwapY>(x1);
        const bool up = hy == ny;
        y0 = fma(t[1], up ? fma(3.0, x0, p0) : fma(-3.0, p0, x0), y0);
        y1 = fma(t[1], up ? fma(3.0, x1, p1) : fma(-3.0, p1, x1), y1);
    }
    if constexpr (DIM == 3) {
        const double p0 = quad_swap<double, kSwapZ>(x0), p1 = quad_swap<double, kSwapZ>(x1);
        const bool up = hz == nz;
        y0 = fma(t[2], up ? fma(3.0, x0, p0) : fma(-3.0, p0, x0), y0);
        y1 = fma(t[2], up ? fma(3.0, x1, p1) : fma(-3.0, p1, x1), y1);
    }
}

// one lane's inputs of one cell, loaded ahead of its solve
struct SplitIn {
    double q0, q1, i0, i1, sig;
    double o0, o1;   // refinement: the cell's previous psi
    long long off;
    int e, d, ci;   // element, direction slot, box cell in level order
    bool ok;
};

template <int DIM, int OCT, bool REF = false, bool ONE = false>
__device__ __forceinline__ void split_fetch(const SweepParams &p, const BoxPlan &bp, const BoxCtx &cx, int l, int it,
                                            int r, SplitIn &in)
{
    constexpr int N = 2 << (DIM - 1);
    // The three loads are issued on every path: a lane without a cell (or a
    // level past the box) reads element 0 and never uses it, and a missing
    // source reads zeros.  Loads behind branches would make the compiler's
    // wait counts pessimistic at the join, and the wait for this level's
    // operands would also drain the prefetch of the levels ahead.
    const BoxLds &L = cx.L;
    const bool lv = l < bp.nl;
    const int lc = lv ? l : 0;
    const int c0 = L.lpl[lc], cnt = L.lpl[lc + 1] - c0;
    // ONE: one direction per workgroup, so the item is the cell (no division)
    const bool inr = lv && it < (ONE ? cnt : cnt * cx.ndir);
    const int itc = inr ? it : 0;
    const int d = ONE ? 0 : fast_div(itc, cnt, 1.0f / (float)cnt), ci = c0 + (itc - d * cnt);
    const int e = L.cell_e[ci];
    const long long base = L.dbase[d];
    const bool ok = inr && e >= 0 && base >= 0;
    const int es = ok ? e : 0;
    const long long off = ok ? base + (long long)es * N + 2 * r : 0;
    const double *srcp = p.src ? p.src + off : p.zeros;
    const double *isop = p.iso ? p.iso + (long long)cx.g * p.Ndof + (long long)es * N + 2 * r : p.zeros;
    const d2v vq = *reinterpret_cast<const d2v *>(srcp);
    const d2v vi = *reinterpret_cast<const d2v *>(isop);
    const double sg = p.sigma[(long long)cx.g * p.Ne + es];
    in.ok = ok;
    in.e = es;
    in.d = ok ? d : 0;
    in.ci = ok ? ci : 0;
    in.off = off;
    // no selects here: split_solve runs only where ok (all lanes of a cell
    // agree), and a select would wait for the loads now instead of at use
    in.q0 = vq.x;
    in.q1 = vq.y;
    in.i0 = vi.x;
    in.i1 = vi.y;
    in.sig = sg;
    if constexpr (REF) {
        // the cell's psi before this sweep overwrites it (no cell of a later level is written yet)
        const d2v vo = *reinterpret_cast<const d2v *>(p.psi + off);
        in.o0 = vo.x;
        in.o1 = vo.y;
    }
}

template <int DIM, typename T, int OCT, bool REF = false, bool ONE = false, bool NT = false>
__device__ __forceinline__ void split_solve(const SweepParams &p, const BoxPlan &bp, const BoxCtx &cx, int r,
                                            const SplitIn &in, const DirConst<T> &K)
{
    constexpr int LPC = 1 << (DIM - 1), NF = LPC, PS = box_slot_stride(NF);
    const BoxLds &L = cx.L;
    const int by = r & 1, bz = (r >> 1) & 1, d = in.d;
    const double inv_vol = 1.0 / bp.vol;
    // trace slots of this cell (precomputed per box cell by box_stage)
    double *slot[DIM];
    {
        const uint32_t sw = L.lcl[in.ci];
        double *pl = L.planes;
#pragma unroll
        for (int ax = 0; ax < DIM; ++ax) {
            const int sidx = (int)((sw >> (10 * ax)) & 1023u);
            slot[ax] = pl + (ONE ? sidx : d * bp.psz[ax] + sidx) * PS;
            pl += bp.nd * bp.psz[ax] * PS;
        }
    }
    // x face: node (., r); y face: (i0, i_z) -> i0 + 2 i_z; z face: (i0, i_y) -> i0 + 2 i_y
    const int my0 = DIM == 3 ? 2 * bz : 0;
    d2v o;
    if constexpr (REF) {
        // defect correction: the right-hand side with the new upwind traces minus the
        // operator on the cell's previous psi, in FP64; the correction solved in T;
        // psi = previous + correction
        const double ux = slot[0][r];
        const double uy[2] = {slot[1][my0], slot[1][my0 + 1]};
        double uz[2] = {0.0, 0.0};
        if constexpr (DIM == 3) {
            uz[0] = slot[2][2 * by];
            uz[1] = slot[2][2 * by + 1];
        }
        const double t[3] = {L.om[3 * d] * inv_vol, L.om[3 * d + 1] * inv_vol, L.om[3 * d + 2] * inv_vol};
        const double sg = in.sig + (p.timedep ? p.tau : 0.0);
        double v0 = (in.q0 + in.i0) * inv_vol, v1 = (in.q1 + in.i1) * inv_vol;
        fd_split_rhs<DIM, double, OCT>(v0, v1, t, ux, uy, uz, by != 0, bz != 0);
        double a0, a1;
        fd_split_apply<DIM, OCT>(in.o0, in.o1, t, sg, by != 0, bz != 0, a0, a1);
        T r0 = (T)(v0 - a0), r1 = (T)(v1 - a1);
        const T tt[3] = {(T)t[0], (T)t[1], (T)t[2]};
        fd_split_solve<DIM, T, OCT>(r0, r1, tt, (T)sg, by != 0, bz != 0);
        o.x = in.o0 + (double)r0;
        o.y = in.o1 + (double)r1;
    } else {
        const T ux = (T)slot[0][r];
        const T uy[2] = {(T)slot[1][my0], (T)slot[1][my0 + 1]};
        T uz[2] = {(T)0, (T)0};
        if constexpr (DIM == 3) {
            uz[0] = (T)slot[2][2 * by];
            uz[1] = (T)slot[2][2 * by + 1];
        }
        if constexpr (ONE) {
            // one direction per workgroup: its constants were formed once, in registers;
            // the source carries M^{-1}'s factor 2^dim with its 1 / vol
            constexpr double sc = DIM == 3 ? 8.0 : 4.0;
            T v0 = (T)((in.q0 + in.i0) * (sc * inv_vol)), v1 = (T)((in.q1 + in.i1) * (sc * inv_vol));
            fd_split_p1_k<DIM, T, OCT>(v0, v1, K, (T)(in.sig + (p.timedep ? p.tau : 0.0)), ux, uy, uz, by != 0,
                                       bz != 0);
            o.x = (double)v0;
            o.y = (double)v1;
        } else {
            T v0 = (T)((in.q0 + in.i0) * inv_vol), v1 = (T)((in.q1 + in.i1) * inv_vol);
            const T t[3] = {(T)(L.om[3 * d] * inv_vol), (T)(L.om[3 * d + 1] * inv_vol),
                            (T)(L.om[3 * d + 2] * inv_vol)};
            fd_split_p1<DIM, T, OCT>(v0, v1, t, (T)(in.sig + (p.timedep ? p.tau : 0.0)), ux, uy, uz, by != 0, bz != 0);
            o.x = (double)v0;
            o.y = (double)v1;
        }
    }
    // NT: the moment kernel reads psi once after the sweep (source iteration,
    // the Krylov solvers, sweeps with moments), and psi leaves through
    // nontemporal stores.  At 64^3 S8 source iteration the box levels took 90.8
    // instead of 96.2 us and the moment kernel 203 instead of 251 us.
    // Back-to-back standalone sweeps into one psi buffer were slower with them
    // (1.346 against 1.280 ms), so other sweeps keep plain stores.
    if constexpr (NT)
        __builtin_nontemporal_store(o, reinterpret_cast<d2v *>(p.psi + in.off));
    else
        *reinterpret_cast<d2v *>(p.psi + in.off) = o;
    // outflow traces
    slot[0][r] = (OCT & 1) ? o.x : o.y;
    if (by == (((OCT >> 1) & 1) ? 0 : 1)) {
        slot[1][my0] = o.x;
        slot[1][my0 + 1] = o.y;
    }
    if constexpr (DIM == 3) {
        if (bz == (((OCT >> 2) & 1) ? 0 : 1)) {
            slot[2][2 * by] = o.x;
            slot[2][2 * by + 1] = o.y;
        }
    }
}

// PF 0: each cell's inputs are loaded when it is solved.  PF 2: the first
// pass of every level is prefetched during the level before (two input
// buffers with alternating roles, so no register copy waits for loads in
// flight; the LDS-only barrier lets them cross it), at the price of registers.
template <int DIM, typename T, int OCT, int PF, bool REF, bool ONE, bool NT>
__device__ __forceinline__ void box_task_split(const SweepParams &p, const BoxPlan &bp, int box_base, double *lds)
{
    constexpr int LPC = 1 << (DIM - 1);
    const BoxCtx cx = box_stage<DIM, 1, OCT>(p, bp, box_base, lds);
    const int tid = threadIdx.x, nt = blockDim.x;
    // One direction per workgroup: its solve constants once, in registers.
    // (Kept in LDS instead, the reads sat in each cell's dependent chain and
    // cost more than the products they saved.)
    DirConst<T> K{};
    if constexpr (ONE && !REF) {
        const double inv_vol = 1.0 / bp.vol;
        const T t[3] = {(T)(cx.L.om[0] * inv_vol), (T)(cx.L.om[1] * inv_vol), (T)(cx.L.om[2] * inv_vol)};
        K = dir_const<DIM, T>(t);
        // this lane's node bits (the lane's r = tid mod 2^(DIM-1) in the level loop)
        const bool ly = (tid & 1) != 0, lz = DIM == 3 && ((tid >> 1) & 1) != 0;
        K.syb = ly ? K.by : -K.by;
        K.szb = lz ? K.bz : -K.bz;
        K.cs = ly == lz ? K.c : -K.c;
        constexpr T ir2 = (T)0.70710678118654752;
        constexpr bool ny = (OCT >> 1) & 1, nz = (OCT >> 2) & 1;
        K.wiy[0] = ly ? ir2 : (ny ? (T)0 : (T)1);
        K.wiy[1] = ly ? ir2 : (ny ? (T)1 : (T)0);
        K.wiz[0] = lz ? ir2 : (nz ? (T)0 : (T)1);
        K.wiz[1] = lz ? ir2 : (nz ? (T)1 : (T)0);
    }
    const int r = tid & (LPC - 1), qd = tid / LPC, nq = nt / LPC;
    if constexpr (PF == 0) {
        for (int l = 0; l < bp.nl; ++l) {
            const int items = (cx.L.lpl[l + 1] - cx.L.lpl[l]) * (ONE ? 1 : cx.ndir);
            for (int it = qd; it < items; it += nq) {   // uniform over the lanes of a cell
                SplitIn in;
                split_fetch<DIM, OCT, REF, ONE>(p, bp, cx, l, it, r, in);
                if (in.ok) split_solve<DIM, T, OCT, REF, ONE, NT>(p, bp, cx, r, in, K);
            }
            lds_barrier();
        }
        return;
    }
    // AHEAD levels in flight: the first pass of level l + AHEAD is fetched
    // while level l solves (PF 2: one level ahead, PF 3: two)
    constexpr int AHEAD = PF == 3 ? 2 : 1;
    auto level = [&](int l, SplitIn &cur, SplitIn &nxt) {
        const int items = (cx.L.lpl[l + 1] - cx.L.lpl[l]) * (ONE ? 1 : cx.ndir);
        split_fetch<DIM, OCT, REF, ONE>(p, bp, cx, l + AHEAD, qd, r, nxt);
        if (cur.ok) split_solve<DIM, T, OCT, REF, ONE, NT>(p, bp, cx, r, cur, K);
        for (int it = qd + nq; it < items; it += nq) {   // uniform over the lanes of a cell
            SplitIn in;
            split_fetch<DIM, OCT, REF, ONE>(p, bp, cx, l, it, r, in);
            if (in.ok) split_solve<DIM, T, OCT, REF, ONE, NT>(p, bp, cx, r, in, K);
        }
        lds_barrier();
    };
    if constexpr (AHEAD == 1) {
        SplitIn b0, b1;
        split_fetch<DIM, OCT, REF, ONE>(p, bp, cx, 0, qd, r, b0);
        for (int l = 0; l < bp.nl; l += 2) {
            level(l, b0, b1);
            if (l + 1 < bp.nl) level(l + 1, b1, b0);
        }
    } else {
        SplitIn b0, b1, b2;
        split_fetch<DIM, OCT, REF, ONE>(p, bp, cx, 0, qd, r, b0);
        split_fetch<DIM, OCT, REF, ONE>(p, bp, cx, 1, qd, r, b1);
        for (int l = 0; l < bp.nl; l += 3) {
            level(l, b0, b2);
            if (l + 1 < bp.nl) level(l + 1, b1, b0);
            if (l + 2 < bp.nl) level(l + 2, b2, b1);
        }
    }
}

template <int DIM, typename T, int PF, bool REF, bool ONE, bool NT = false>
__global__ void __launch_bounds__(1024) k_sweep_box_split(SweepParams p, BoxPlan bp, int box_base)
{
    extern __shared__ double lds[];
    const int oct = blockIdx.x % bp.noct;   // workgroup-uniform
    switch (oct) {
    case 0: box_task_split<DIM, T, 0, PF, REF, ONE, NT>(p, bp, box_base, lds); break;
    case 1: box_task_split<DIM, T, 1, PF, REF, ONE, NT>(p, bp, box_base, lds); break;
    case 2: box_task_split<DIM, T, 2, PF, REF, ONE, NT>(p, bp, box_base, lds); break;
    case 3: box_task_split<DIM, T, 3, PF, REF, ONE, NT>(p, bp, box_base, lds); break;
    case 4: if constexpr (DIM == 3) box_task_split<DIM, T, 4, PF, REF, ONE, NT>(p, bp, box_base, lds); break;
    case 5: if constexpr (DIM == 3) box_task_split<DIM, T, 5, PF, REF, ONE, NT>(p, bp, box_base, lds); break;
    case 6: if constexpr (DIM == 3) box_task_split<DIM, T, 6, PF, REF, ONE, NT>(p, bp, box_base, lds); break;
    default: if constexpr (DIM == 3) box_task_split<DIM, T, 7, PF, REF, ONE, NT>(p, bp, box_base, lds); break;
    }
}

template <int DIM, typename T, int PF, bool REF = false>
static hipError_t run_boxes_split(const DeviceSweep &ds, const SweepParams &prm, hipStream_t s)
{
    const BoxPlan &bp = ds.box;
    const int nlev = (int)ds.box_ptr.size() - 1;
    for (int k = 0; k < nlev; ++k) {
        const int nbox = ds.box_ptr[k + 1] - ds.box_ptr[k];
        if (nbox == 0) continue;
        // A box level with at most one workgroup per CU (the levels near the
        // mesh corners) runs twice the threads per workgroup, up to 1024: the
        // CU has the room, and a box level then takes fewer passes over its cells.
        const long long wgs = (long long)nbox * bp.noct * prm.G * bp.nchunk;
        const int th = (wgs <= ds.box_ncu && ds.box_threads <= 512) ? 2 * ds.box_threads : ds.box_threads;
        // one direction per workgroup (the tiling the autotuner picks at 64^3) has its own
        // instantiation without the item -> (direction, cell) division
        // nontemporal psi stores (prm.nt_psi) have their own instantiation, for
        // the FP64 one-direction tiling the solvers run
        if constexpr (std::is_same<T, double>::value && !REF) {
            if (bp.nd == 1 && prm.nt_psi) {
                hipLaunchKernelGGL((k_sweep_box_split<DIM, T, PF, REF, true, true>),
                                   dim3(nbox * bp.noct, prm.G * bp.nchunk), dim3(th), bp.lds_bytes, s, prm, bp,
                                   ds.box_ptr[k]);
                continue;
            }
        }
        if (bp.nd == 1)
            hipLaunchKernelGGL((k_sweep_box_split<DIM, T, PF, REF, true>), dim3(nbox * bp.noct, prm.G * bp.nchunk),
                               dim3(th), bp.lds_bytes, s, prm, bp, ds.box_ptr[k]);
        else
            hipLaunchKernelGGL((k_sweep_box_split<DIM, T, PF, REF, false>), dim3(nbox * bp.noct, prm.G * bp.nchunk),
                               dim3(th), bp.lds_bytes, s, prm, bp, ds.box_ptr[k]);
    }
    return hipGetLastError();
}

template <int DIM, int P, typename T, int PF>
__global__ void __launch_bounds__(256) k_sweep_box(SweepParams p, BoxPlan bp, int box_base)
{
    extern __shared__ double lds[];
    const int oct = blockIdx.x % bp.noct;   // workgroup-uniform
    switch (oct) {
    case 0: box_task<DIM, P, T, 0, PF>(p, bp, box_base, lds); break;
    case 1: box_task<DIM, P, T, 1, PF>(p, bp, box_base, lds); break;
    case 2: box_task<DIM, P, T, 2, PF>(p, bp, box_base, lds); break;
    case 3: box_task<DIM, P, T, 3, PF>(p, bp, box_base, lds); break;
    case 4: if constexpr (DIM == 3) box_task<DIM, P, T, 4, PF>(p, bp, box_base, lds); break;
    case 5: if constexpr (DIM == 3) box_task<DIM, P, T, 5, PF>(p, bp, box_base, lds); break;
    case 6: if constexpr (DIM == 3) box_task<DIM, P, T, 6, PF>(p, bp, box_base, lds); break;
    default: if constexpr (DIM == 3) box_task<DIM, P, T, 7, PF>(p, bp, box_base, lds); break;
    }
}

template <int DIM, int P, typename T, int PF>
static hipError_t run_boxes(const DeviceSweep &ds, const SweepParams &prm, hipStream_t s)
{
    const BoxPlan &bp = ds.box;
    const int nlev = (int)ds.box_ptr.size() - 1;
    for (int k = 0; k < nlev; ++k) {
        const int nbox = ds.box_ptr[k + 1] - ds.box_ptr[k];
        if (nbox == 0) continue;
        hipLaunchKernelGGL((k_sweep_box<DIM, P, T, PF>), dim3(nbox * bp.noct, prm.G * bp.nchunk),
                           dim3(ds.box_threads), bp.lds_bytes, s, prm, bp, ds.box_ptr[k]);
    }
    return hipGetLastError();
}

template <int DIM, int P, typename T>
static hipError_t run_boxes_pf(const DeviceSweep &ds, const SweepParams &prm, hipStream_t s)
{
    if (ds.box_prefetch == 0) return run_boxes<DIM, P, T, 0>(ds, prm, s);
    return run_boxes<DIM, P, T, 2>(ds, prm, s);
}

template <int DIM, int P>
static hipError_t run_boxes_p(const DeviceSweep &ds, const SweepParams &prm, hipStream_t s)
{
    if constexpr (P == 1) {
        // lane-split cells; the workgroup must hold whole cells
        if (ds.box_threads % (1 << (DIM - 1))) return hipErrorInvalidValue;
        if (ds.refine) {
            // in-place defect correction with FP32 cell solves (psi holds the previous iterate)
            if (ds.box_prefetch == 3) return run_boxes_split<DIM, float, 3, true>(ds, prm, s);
            if (ds.box_prefetch == 2) return run_boxes_split<DIM, float, 2, true>(ds, prm, s);
            return run_boxes_split<DIM, float, 0, true>(ds, prm, s);
        }
        if (ds.precision == 1) {
            if (ds.box_prefetch == 3) return run_boxes_split<DIM, float, 3>(ds, prm, s);
            if (ds.box_prefetch == 2) return run_boxes_split<DIM, float, 2>(ds, prm, s);
            return run_boxes_split<DIM, float, 0>(ds, prm, s);
        }
        if (ds.box_prefetch == 3) return run_boxes_split<DIM, double, 3>(ds, prm, s);
        if (ds.box_prefetch == 2) return run_boxes_split<DIM, double, 2>(ds, prm, s);
        return run_boxes_split<DIM, double, 0>(ds, prm, s);
    }
    if (ds.box_threads > 256) return hipErrorInvalidValue;
    if (ds.precision == 1) return run_boxes_pf<DIM, P, float>(ds, prm, s);
    return run_boxes_pf<DIM, P, double>(ds, prm, s);
}

hipError_t launch_box_sweep(const DeviceSweep &ds, const SweepParams &prm, hipStream_t s)
{
    if (!prm.uniform || !ds.box_buf || ds.box_threads < 64 || ds.box_threads > 1024) return hipErrorInvalidValue;
    if (prm.dim == 2) {
        switch (prm.p) {
        case 0: return run_boxes_p<2, 0>(ds, prm, s);
        case 1: return run_boxes_p<2, 1>(ds, prm, s);
        case 2: return run_boxes_p<2, 2>(ds, prm, s);
        }
    } else if (prm.dim == 3) {
        switch (prm.p) {
        case 0: return run_boxes_p<3, 0>(ds, prm, s);
        case 1: return run_boxes_p<3, 1>(ds, prm, s);
        }
    }
    return hipErrorInvalidValue;
}

}  // namespace allium
