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

// Solve one cell whose inputs are loaded: traces and K[a] from LDS, the mass
// and face-mass matrices as compile-time tensor products (uniform axis-aligned
// boxes: Mu = vol * unit_mass, F_side = area * unit_face_mass); store
// psi to HBM and the outflow traces to the planes (only when in.ok).  mid()
// runs between the assembly of the local system and its LU and may overwrite
// `in` (the next level's loads go there: issued after this cell's inputs were
// consumed and before its psi stores, so their latency overlaps the LU).
template <int DIM, int P, typename T, int OCT, typename Mid>
__device__ __forceinline__ void box_solve(const SweepParams &p, const BoxPlan &bp, const BoxLds &L,
                                          const BoxInputs<Sh<DIM, P>::N> &in, int g, Mid &&mid)
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
    T x[N];
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = (T)b[i];
    mid();
    lu_factor_solve<N, T>(A, x);

    // ---- store psi and the outflow traces ----------------------------------
    const long long off = ((long long)g * p.Nomega + a) * p.Ndof + (long long)in.e * N;
    if constexpr (N % 2 == 0) {
        d2v *o2 = reinterpret_cast<d2v *>(p.psi + off);
#pragma unroll
        for (int i = 0; i < N / 2; ++i) { d2v v; v.x = (double)x[2 * i]; v.y = (double)x[2 * i + 1]; o2[i] = v; }
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) p.psi[off + i] = (double)x[i];
    }
#pragma unroll
    for (int ax = 0; ax < DIM; ++ax) {
        const int s_out = ((OCT >> ax) & 1) ? 2 * ax : 2 * ax + 1;
#pragma unroll
        for (int m = 0; m < NF; ++m) slot[ax][m] = (double)x[side_node<DIM, P>(s_out, m)];
    }
}

// PF: 0 loads each cell's inputs when it is solved; 2 loads the next level's
// first-pass inputs during the current solve (between assembly and LU).
template <int DIM, int P, typename T, int OCT, int PF>
__device__ __forceinline__ void box_task(const SweepParams &p, const BoxPlan &bp, int box_base, double *lds)
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
    int *dirw = reinterpret_cast<int *>(omw + 3 * bp.nd);
    uint32_t *lcw = reinterpret_cast<uint32_t *>(dirw + bp.nd);
    const int vol = bp.bs[0] * bp.bs[1] * bp.bs[2];
    int *lpw = reinterpret_cast<int *>(lcw + vol);
    const BoxLds L{lds, Kw, omw, dirw, lcw, lpw};

    // ---- stage the workgroup tables ----------------------------------------
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
    for (int i = tid; i < vol; i += nt) lcw[i] = bp.lcells[i];
    for (int i = tid; i <= bp.nl; i += nt) lpw[i] = bp.lptr[i];
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
    __syncthreads();

    // ---- the box's levels ---------------------------------------------------
    auto nothing = [] {};
    if constexpr (PF == 0) {
        for (int l = 0; l < bp.nl; ++l) {
            const int items = (L.lpl[l + 1] - L.lpl[l]) * ndir;
            for (int it = tid; it < items; it += nt) {
                BoxInputs<N> in;
                box_fetch<DIM, P, OCT>(p, bp, L, ndir, lo, ext, g, l, it, in);
                if (in.ok) box_solve<DIM, P, T, OCT>(p, bp, L, in, g, nothing);
            }
            __syncthreads();
        }
    } else {
        BoxInputs<N> cur, nxt;
        box_fetch<DIM, P, OCT>(p, bp, L, ndir, lo, ext, g, 0, tid, cur);
        for (int l = 0; l < bp.nl; ++l) {
            const int items = (L.lpl[l + 1] - L.lpl[l]) * ndir;
            auto pre = [&] { box_fetch<DIM, P, OCT>(p, bp, L, ndir, lo, ext, g, l + 1, tid, nxt); };
            if (cur.ok) box_solve<DIM, P, T, OCT>(p, bp, L, cur, g, pre);
            else pre();
            for (int it = tid + nt; it < items; it += nt) {
                BoxInputs<N> in;
                box_fetch<DIM, P, OCT>(p, bp, L, ndir, lo, ext, g, l, it, in);
                if (in.ok) box_solve<DIM, P, T, OCT>(p, bp, L, in, g, nothing);
            }
            cur = nxt;
            __syncthreads();
        }
    }
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
    if (ds.precision == 1) return run_boxes_pf<DIM, P, float>(ds, prm, s);
    return run_boxes_pf<DIM, P, double>(ds, prm, s);
}

hipError_t launch_box_sweep(const DeviceSweep &ds, const SweepParams &prm, hipStream_t s)
{
    if (!prm.uniform || !ds.box_buf || ds.box_threads < 64 || ds.box_threads > 256) return hipErrorInvalidValue;
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
