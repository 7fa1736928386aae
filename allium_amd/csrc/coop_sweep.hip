// Cooperative cell solves for local DG systems that do not fit one lane's
// registers (3D p = 2: 27 x 27), restating src/sweep.cpp:366-475.
//
// 16 lanes solve one cell (e, a, g), one row of a wave's DPP row each: lane t
// assembles rows t and t + 16 (t + 16 < n) of
//   A = sum_d Omega_d G_d + sigma M + tau Mt + sum_{outflow} (Omega.n) F
// and of the right-hand side (source, isotropic source, upwind traces), then
// the cell's lanes run an LU without pivoting (A + A^T is positive definite)
// and a column-oriented back substitution.  The pivot row and each solved
// unknown reach the other lanes of the cell by DPP row broadcast
// (row_newbcast) from the owner lane's registers: no LDS, no synchronisation.
// A 256-thread workgroup holds 16 cells, four per wave.
// The launch covers one level of up to kMaxBatch independent instances, so
// a batch of uneven meshes sweeps in max(levels) launches instead of
// sum(levels).
#include <utility>

#include "device.hpp"
#include "kernel_common.hpp"

namespace allium {

bool coop_supported(int dim, int p) { return dim == 3 && p == 2; }

// every entry of r is computed before any code after this point that the
// compiler would otherwise hoist (empty asm statements keep their order)
template <int N>
__device__ __forceinline__ void fence_row(double (&r)[N])
{
#pragma unroll
    for (int j = 0; j < N; ++j) asm volatile("" : "+v"(r[j]));
}

template <int First, int Count>
struct AxisRange {
    static constexpr int first = First, count = Count;
};

// lane L of each 16-lane row of the wave, to every lane of that row.  Every
// lane is active where this is called, so the source lane is always valid and
// bound_ctrl (zero for an invalid source) only spares the copy of an old value.
template <int L, typename T>
__device__ __forceinline__ T row_bcast(T x)
{
    static_assert(L >= 0 && L < 16, "row_newbcast takes a lane of the row");
    return __builtin_amdgcn_update_dpp((T)0, x, 0x150 + L, 0xF, 0xF, true);
}

template <typename F, int... K>
__device__ __forceinline__ void for_each_k(F &&f, std::integer_sequence<int, K...>)
{
    (f(std::integral_constant<int, K>{}), ...);
}

// one cell (e, a, g) of a level and what every row of it shares
template <int DIM, int P>
struct CellRef {
    int e, a, g;
    double om[3], sig;
    long long base, off;   // psi offsets of direction (g, a) and of the cell
    int c[3];              // element coordinates
};

// the cell of workgroup-local slot `slot`; valid is false past the level's end
template <int DIM, int P>
__device__ __forceinline__ CellRef<DIM, P> load_cell(const SweepParams &p, const CoopLevel &cl, int inst, long long wg0,
                                                     int slot, bool &valid)
{
    constexpr int N = Sh<DIM, P>::N;
    CellRef<DIM, P> cr;
    const int local = (int)(wg0 - cl.start[inst]) + slot;
    valid = local < cl.ncells[inst] * p.G;
    cr.e = cr.a = cr.g = 0;
    if (valid) {
        const int nc = cl.ncells[inst];
        cr.g = local / nc;
        const uint32_t cell = as_global(p.cells)[cl.cbeg[inst] + (local - cr.g * nc)];
        cr.e = (int)(cell >> 10);
        cr.a = (int)(cell & 1023u);
    }
    const auto g_omega = as_global(p.omega);
    for (int d = 0; d < 3; ++d) cr.om[d] = g_omega[4 * cr.a + d];
    cr.sig = as_global(p.sigma)[(long long)cr.g * p.Ne + cr.e];
    cr.base = ((long long)cr.g * p.Nomega + cr.a) * p.Ndof;
    cr.off = cr.base + (long long)cr.e * N;
    cr.c[0] = cr.e % p.nel[0];
    cr.c[1] = (cr.e / p.nel[0]) % p.nel[1];
    cr.c[2] = cr.e / (p.nel[0] * p.nel[1]);
    return cr;
}

// row rr of the cell's local system into (Ar, b)
template <int DIM, int P, bool UNI>
__device__ __forceinline__ void assemble_row(const SweepParams &p, const CellRef<DIM, P> &cr, const int rr,
                                             double (&Ar)[Sh<DIM, P>::N], double &b)
{
    using S = Sh<DIM, P>;
    constexpr int N = S::N, NF = S::NF, NS = S::NS, Q = P + 1;
    // every table read through p is global memory
    const auto g_K = as_global(p.K), g_Mu = as_global(p.Mu), g_Mt = as_global(p.Mt), g_Gd = as_global(p.Gd);
    const auto g_src = as_global(p.src), g_iso = as_global(p.iso), g_F = as_global(p.Fface);
    const auto g_fnor = as_global(p.fnor);
    const auto g_refl = as_global(p.refl);
    const auto g_psi = as_global(p.psi);
    const int e = cr.e, a = cr.a, g = cr.g;
    const double(&om)[3] = cr.om;
    const double sig = cr.sig;
    const long long base = cr.base, off = cr.off;
    const int(&c)[3] = cr.c;
    b = 0.0;
    if (p.src) b = g_src[off + rr];
    if (p.iso) b += g_iso[(long long)g * p.Ndof + (long long)e * N + rr];

    // Faces, one axis at a time: a node lies on at most one side per axis (the
    // low side where its coordinate on the axis is 0, the high side where it
    // is P).  Every address follows from the cell alone, so the face normal,
    // the face-mass row and the upwind trace are loaded without branches, and
    // the sign of Omega.n only selects what they are added to.  The inflow
    // traces go into b before the element matrices are read (the row sum is
    // not live yet); on non-uniform geometry the outflow face blocks join the
    // row after them (a uniform mesh has them in K).  Both passes take the
    // sides in the order 2 axis + {low, high}, as one loop over the sides would.
    const int ic[3] = {rr % Q, (rr / Q) % Q, rr / (Q * Q)};
    // the node's side on axis ax (-1: none) and its face-lexicographic index there
    auto side_of = [&](int ax, int &k) {
        k = ax == 0 ? ic[1] + Q * ic[2] : (ax == 1 ? ic[0] + Q * ic[2] : ic[0] + Q * ic[1]);
        return ic[ax] == 0 ? 2 * ax : (ic[ax] == P ? 2 * ax + 1 : -1);
    };
    // node l of a side of axis ax, with coordinate 0 on the axis
    auto side_base = [](int ax, int l) {
        return ax == 0 ? Q * (l % Q) + Q * Q * (l / Q) : (ax == 1 ? l % Q + Q * Q * (l / Q) : l);
    };
    // face of side s (non-uniform geometry)
    auto face_of = [&](int s) {
        const int ax = s >> 1;
        int cf[3] = {c[0], c[1], c[2]};
        cf[ax] += s & 1;
        return p.foff[ax] + cf[0] + p.fext[ax][0] * (cf[1] + p.fext[ax][1] * cf[2]);
    };
    double dots[DIM];   // Omega.n on the node's side of each axis, 0 where it has none
    auto gF = g_F;
    auto gN = g_fnor;
    auto gP = g_psi;
#pragma unroll
    for (int ax = 0; ax < DIM; ++ax) {
        // one axis's loads in flight at a time
        if (ax > 0) asm volatile("" : "+v"(b), "+s"(gF), "+s"(gN), "+s"(gP));
        int k;
        const int sd = side_of(ax, k);
        const bool on = sd >= 0;
        const int hi = sd & 1, s = 2 * ax + hi;
        const bool interior = hi ? (c[ax] < p.nel[ax] - 1) : (c[ax] > 0);
        double dot;
        gptr<const double> F;
        if constexpr (UNI) {
            dot = hi ? om[ax] : -om[ax];
            F = gF + s * NF * NF + k * NF;
        } else {
            const int f = face_of(s);
            dot = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) dot = fma(om[d], gN[3 * f + d], dot);
            if (!hi && interior) dot = -dot;
            F = gF + (size_t)f * NF * NF + k * NF;
        }
        // upwind trace: the neighbour's nodes on the opposite side, or on a
        // reflective side the reflected direction's own nodes; elsewhere the
        // cell's own entries stand in as an in-bounds address and are not used
        const bool refl = !interior && ((p.refl_mask >> s) & 1);
        const int sax = ax == 0 ? 1 : (ax == 1 ? Q : Q * Q);   // node stride along the axis
        gptr<const double> u = gP + off;
        int ub = 0;
        if (interior) {
            u = gP + base + (long long)(hi ? e + p.stride[ax] : e - p.stride[ax]) * N;
            ub = hi ? 0 : P * sax;
        } else if (refl) {
            u = gP + ((long long)g * p.Nomega + g_refl[a * NS + s]) * p.Ndof + (long long)e * N;
            ub = hi ? P * sax : 0;
        }
        dots[ax] = on ? dot : 0.0;
        double acc = 0.0;
#pragma unroll
        for (int l = 0; l < NF; ++l) acc = fma(F[l], u[ub + side_base(ax, l)], acc);
        const bool inflow = on && dot < 0 && (interior || refl);
        b = inflow ? fma(-dot, acc, b) : b;
    }

    // Each matrix row is read in two halves.  The half's row pointers pass
    // through an empty asm after fences on every entry of the running sum, so
    // the compiler cannot hoist all the row loads ahead of the sums (all 5 N of
    // the jittered case took 256 VGPRs + AGPRs, one wave per SIMD); halves keep
    // the kernel at 128 VGPRs, four waves per SIMD.
    constexpr int H = (N + 1) / 2;
    if constexpr (UNI) {
        auto K = g_K + ((size_t)a * N + rr) * N;   // includes tau * Mt and the outflow faces
        auto Mu = g_Mu + (size_t)rr * N;
        asm volatile("" : "+v"(K), "+v"(Mu), "+v"(b));
#pragma unroll
        for (int j = 0; j < H; ++j) Ar[j] = fma(sig, Mu[j], K[j]);
        fence_row<H>(reinterpret_cast<double(&)[H]>(Ar));
        asm volatile("" : "+v"(K), "+v"(Mu));
#pragma unroll
        for (int j = H; j < N; ++j) Ar[j] = fma(sig, Mu[j], K[j]);
    } else {
        auto Gd = g_Gd + (size_t)e * DIM * N * N + (size_t)rr * N;
        auto Mu = g_Mu + (size_t)e * N * N + (size_t)rr * N;
        auto Mt = g_Mt + (size_t)e * N * N + (size_t)rr * N;
        // A steady problem (tau = 0) skips the time-mass row, whose
        // fma(0, Mt, .) would leave the sum unchanged.
        // (the first phase starts after the face pass: its pointer passes through
        // the same empty asm as b)
        auto phase = [&](auto q, double c, bool first) {
            if (!first) fence_row<N>(Ar);
            asm volatile("" : "+v"(q), "+v"(b));
#pragma unroll
            for (int j = 0; j < H; ++j) Ar[j] = first ? c * q[j] : fma(c, q[j], Ar[j]);
            fence_row<H>(reinterpret_cast<double(&)[H]>(Ar));
            asm volatile("" : "+v"(q));
#pragma unroll
            for (int j = H; j < N; ++j) Ar[j] = first ? c * q[j] : fma(c, q[j], Ar[j]);
        };
        phase(Gd, om[0], true);
#pragma unroll
        for (int d = 1; d < DIM; ++d) phase(Gd + (size_t)d * N * N, om[d], false);
        phase(Mu, sig, false);
        if (p.timedep) phase(Mt, p.tau, false);
        // outflow: the own face block joins the row (the face-mass rows were
        // read by the face pass, so these loads are cache hits)
        // Each row is loaded before any of it is used (a row loaded only where
        // it is added would become one branch and one round trip per entry):
        // the first axis's row, then the other two in one batch.
        auto outflow_rows = [&](auto axes) {
            fence_row<N>(Ar);
            asm volatile("" : "+s"(gF));
            constexpr int A0 = decltype(axes)::first, NA = decltype(axes)::count;
            double Fk[NA][NF];
#pragma unroll
            for (int t = 0; t < NA; ++t) {
                const int ax = A0 + t;
                int k;
                const int hi = side_of(ax, k) & 1;
                const auto F = gF + (size_t)face_of(2 * ax + hi) * NF * NF + k * NF;
#pragma unroll
                for (int l = 0; l < NF; ++l) Fk[t][l] = F[l];
            }
#pragma unroll
            for (int t = 0; t < NA; ++t) fence_row<NF>(Fk[t]);
#pragma unroll
            for (int t = 0; t < NA; ++t) {
                const int ax = A0 + t;
                int k;
                const int hi = side_of(ax, k) & 1;
                const int sax = ax == 0 ? 1 : (ax == 1 ? Q : Q * Q);
                const bool outflow = dots[ax] > 0;
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int l = 0; l < NF; ++l) {
                        const int j = h * P * sax + side_base(ax, l);
                        Ar[j] = (outflow && hi == h) ? fma(dots[ax], Fk[t][l], Ar[j]) : Ar[j];
                    }
            }
        };
        outflow_rows(AxisRange<0, 1>{});
        outflow_rows(AxisRange<1, DIM - 1>{});
    }
}

// 16 lanes per cell, two rows per lane, LU by DPP row broadcast (no LDS
// exchange): the FP64 solves and the uniform-geometry FP32 ones
template <int DIM, int P, bool UNI, typename T>
__device__ __forceinline__ void coop_cell_dpp(const SweepParams &p, const CoopLevel &cl, int inst, long long wg0)
{
    constexpr int N = Sh<DIM, P>::N;
    static_assert(N > 16 && N <= 32, "two rows per lane of a 16-lane row");
    const int slot = threadIdx.x >> 4, t = threadIdx.x & 15;
    bool valid;
    const CellRef<DIM, P> cr = load_cell<DIM, P>(p, cl, inst, wg0, slot, valid);
    const auto g_psi = as_global(p.psi);
    const long long off = cr.off;
    auto assemble = [&](const int rr, double (&Ar)[N], double &b) { assemble_row<DIM, P, UNI>(p, cr, rr, Ar, b); };

    // rows t and t + 16 of the cell (a lane without a second row assembles its
    // first one again; that copy takes no part in the solve)
    constexpr int R1 = N - 16;   // lanes with a second row
    const int r1 = t < R1 ? t + 16 : t;
    T A0[N], A1[N], b0, b1;
    {
        double Ar[N], b;
        assemble(t, Ar, b);
        if constexpr (sizeof(T) == 4) {
            // FP32 rows wait in LDS (one column per thread, 27.6 KB per
            // workgroup) while the second row is assembled: registers for
            // more waves.  The empty asm keeps the compiler from carrying
            // the stored values across in registers.
            __shared__ float stash[N][256];
#pragma unroll
            for (int j = 0; j < N; ++j) stash[j][threadIdx.x] = (float)Ar[j];
            b0 = (T)b;
            asm volatile("" ::: "memory");
            fence_row<N>(Ar);
            assemble(r1, Ar, b);
#pragma unroll
            for (int j = 0; j < N; ++j) A1[j] = (T)Ar[j];
            b1 = (T)b;
            asm volatile("" ::: "memory");
#pragma unroll
            for (int j = 0; j < N; ++j) A0[j] = stash[j][threadIdx.x];
        } else {
#pragma unroll
            for (int j = 0; j < N; ++j) A0[j] = (T)Ar[j];
            b0 = (T)b;
            fence_row<N>(Ar);
            assemble(r1, Ar, b);
#pragma unroll
            for (int j = 0; j < N; ++j) A1[j] = (T)Ar[j];
            b1 = (T)b;
        }
    }

    // ---- LU without pivoting ------------------------------------------------
    // Pivot row k is row (k & 15) of slot k >> 4 on lane k & 15.  Every lane
    // runs every update; a row at or above the pivot takes multiplier 0.
    for_each_k(
        [&](auto kc) {
            constexpr int k = decltype(kc)::value, kl = k & 15;
            const T piv = k < 16 ? row_bcast<kl>(A0[k]) : row_bcast<kl>(A1[k]);
            const T rp = fast_rcp(piv);
            T l0 = (T)0, l1 = (T)0;
            if constexpr (k < 15) l0 = t > k ? A0[k] * rp : (T)0;
            l1 = (t < R1 && t + 16 > k) ? A1[k] * rp : (T)0;
#pragma unroll
            for (int j = k + 1; j < N; ++j) {
                const T pj = k < 16 ? row_bcast<kl>(A0[j]) : row_bcast<kl>(A1[j]);
                if constexpr (k < 15) A0[j] = fma(-l0, pj, A0[j]);
                A1[j] = fma(-l1, pj, A1[j]);
            }
            const T pb = k < 16 ? row_bcast<kl>(b0) : row_bcast<kl>(b1);
            if constexpr (k < 15) b0 = fma(-l0, pb, b0);
            b1 = fma(-l1, pb, b1);
        },
        std::make_integer_sequence<int, N>{});

    // ---- back substitution, one unknown per step ----------------------------
    T x0 = (T)0, x1 = (T)0;
    for_each_k(
        [&](auto kc) {
            constexpr int k = N - 1 - decltype(kc)::value, kl = k & 15;
            const T own = k < 16 ? b0 * fast_rcp(A0[k]) : b1 * fast_rcp(A1[k]);
            const T xk = row_bcast<kl>(own);
            if constexpr (k < 16) x0 = t == kl ? xk : x0;
            else x1 = t == kl ? xk : x1;
            b0 = t < k ? fma(-A0[k], xk, b0) : b0;
            b1 = (t < R1 && t + 16 < k) ? fma(-A1[k], xk, b1) : b1;
        },
        std::make_integer_sequence<int, N>{});
    if (valid) {
        g_psi[off + t] = (double)x0;
        if (t < R1) g_psi[off + t + 16] = (double)x1;
    }
}

// 32 lanes per cell, one row per lane, the pivot row published through LDS:
// the non-uniform FP32 solves, whose limit is the memory pipeline rather than
// the LDS, and which keep four waves per SIMD this way
template <int DIM, int P, bool UNI, typename T>
__device__ __forceinline__ void coop_cell_lds(const SweepParams &p, const CoopLevel &cl, int inst, long long wg0)
{
    constexpr int N = Sh<DIM, P>::N;
    static_assert(N <= 32, "one row per lane of a 32-lane group");
    __shared__ T piv[8][N + 1];
    const int slot = threadIdx.x >> 5, r = threadIdx.x & 31;
    bool valid;
    const CellRef<DIM, P> cr = load_cell<DIM, P>(p, cl, inst, wg0, slot, valid);
    const bool row = valid && r < N;
    const int rr = r < N ? r : 0;
    double Ar[N], b;
    assemble_row<DIM, P, UNI>(p, cr, rr, Ar, b);
    // lanes without a row carry an identity row so that nothing is singular
    T A[N], bb = (T)b;
#pragma unroll
    for (int j = 0; j < N; ++j) A[j] = row ? (T)Ar[j] : (T)(j == r ? 1.0 : 0.0);
    if (!row) bb = (T)0;
    // LU without pivoting across the cell's lanes: the owner of pivot row k
    // publishes it through LDS; the 32 lanes of a cell are half of one wave
    // and its pivot row is LDS no other cell touches, so the exchange needs
    // no workgroup barrier (LDS executes a wave's instructions in order, and
    // the wavefront-scope fences keep the compiler from moving the reads
    // above the owner lane's write)
    auto cell_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (r == k) {
#pragma unroll
            for (int j = k; j < N; ++j) piv[slot][j] = A[j];
            piv[slot][N] = bb;
        }
        cell_sync();
        if (r > k && r < N) {
            const T l = A[k] * fast_rcp(piv[slot][k]);
#pragma unroll
            for (int j = k + 1; j < N; ++j) A[j] = fma(-l, piv[slot][j], A[j]);
            bb = fma(-l, piv[slot][N], bb);
        }
        cell_sync();
    }
    // back substitution, one unknown per step
    T x = (T)0;
#pragma unroll
    for (int k = N - 1; k >= 0; --k) {
        if (r == k) {
            x = bb * fast_rcp(A[k]);
            piv[slot][0] = x;
        }
        cell_sync();
        const T xk = piv[slot][0];
        if (r < k) bb = fma(-A[k], xk, bb);
        cell_sync();
    }
    if (row) as_global(p.psi)[cr.off + rr] = (double)x;
}

// BATCH: the instances' parameters are the device array prms (a level of
// several instances); otherwise p0, one instance.  The two are separate
// instantiations so that the parameters' address space is known (a pointer
// selected between the two would be generic and its loads flat).
// Layout per instantiation: the DPP cells (16 per workgroup) except for
// non-uniform FP32 solves, which keep the LDS exchange (8 cells per workgroup).
template <bool UNI, typename T>
constexpr bool kCoopDpp = UNI || sizeof(T) == 8;
template <bool UNI, typename T>
constexpr int kCoopCellsPerGroup = kCoopDpp<UNI, T> ? 16 : 8;
static_assert(kCoopCells % 16 == 0, "instances start on a workgroup of either layout");

template <int DIM, int P, bool UNI, typename T, bool BATCH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 3)))
k_sweep_coop(SweepParams p0, const SweepParams *__restrict__ prms, CoopLevel cl)
{
    // Workgroups go to the 8 XCDs round robin (blockIdx.x % 8).  Each XCD takes
    // a contiguous run of the level's cells instead, so the cells of one element
    // (adjacent in the element-ordered level) share that XCD's L2.
    const int nwg = (int)gridDim.x, x = (int)(blockIdx.x & 7u), q = (int)(blockIdx.x >> 3);
    const int per = nwg >> 3, rem = nwg & 7;
    const long long wg = (long long)x * per + (x < rem ? x : rem) + q;
    // each instance's items start on a multiple of kCoopCells, so the instance (and
    // with it every parameter read through p) is uniform over the workgroup
    const long long wg0 = wg * kCoopCellsPerGroup<UNI, T>;
    int inst = 0;
    while (inst + 1 < cl.ninst && wg0 >= cl.start[inst + 1]) ++inst;
    if constexpr (kCoopDpp<UNI, T>) {
        if constexpr (BATCH)
            coop_cell_dpp<DIM, P, UNI, T>(*(const SweepParams *)as_global(prms + inst), cl, inst, wg0);
        else
            coop_cell_dpp<DIM, P, UNI, T>(p0, cl, inst, wg0);
    } else {
        if constexpr (BATCH)
            coop_cell_lds<DIM, P, UNI, T>(*(const SweepParams *)as_global(prms + inst), cl, inst, wg0);
        else
            coop_cell_lds<DIM, P, UNI, T>(p0, cl, inst, wg0);
    }
}

template <int DIM, int P, bool UNI>
static hipError_t launch_dp(const SweepParams &p0, const SweepParams *d_prms, const CoopLevel &cl, int precision,
                            hipStream_t s)
{
    const long long items = cl.start[cl.ninst];
    if (items <= 0) return hipSuccess;
    auto grid_of = [items](int cells) { return (int)((items + cells - 1) / cells); };
    if (precision == 1) {
        const int grid = grid_of(kCoopCellsPerGroup<UNI, float>);
        if (d_prms)
            hipLaunchKernelGGL((k_sweep_coop<DIM, P, UNI, float, true>), dim3(grid), dim3(256), 0, s, p0, d_prms, cl);
        else
            hipLaunchKernelGGL((k_sweep_coop<DIM, P, UNI, float, false>), dim3(grid), dim3(256), 0, s, p0, d_prms, cl);
    } else {
        const int grid = grid_of(kCoopCellsPerGroup<UNI, double>);
        if (d_prms)
            hipLaunchKernelGGL((k_sweep_coop<DIM, P, UNI, double, true>), dim3(grid), dim3(256), 0, s, p0, d_prms, cl);
        else
            hipLaunchKernelGGL((k_sweep_coop<DIM, P, UNI, double, false>), dim3(grid), dim3(256), 0, s, p0, d_prms, cl);
    }
    return hipGetLastError();
}

hipError_t launch_coop_level(const SweepParams &p0, const SweepParams *d_prms, const CoopLevel &cl, int dim, int p,
                             bool uniform, int precision, hipStream_t s)
{
    if (dim == 3 && p == 2)
        return uniform ? launch_dp<3, 2, true>(p0, d_prms, cl, precision, s)
                       : launch_dp<3, 2, false>(p0, d_prms, cl, precision, s);
    return hipErrorInvalidValue;
}

}  // namespace allium
