// HIP kernels of diffusion synthetic acceleration on gfx950: the block-sparse
// MIP matrix (products, residuals, damped block-Jacobi sweeps), the transfers
// between DG nodes and the vertex grid, the stencil levels of the multigrid
// V-cycle and the conjugate-gradient vector updates.  All of it is
// bandwidth-bound FP64 streaming: each kernel reads its operands once, with
// consecutive lanes on consecutive rows.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dsa.hpp"

namespace allium {

namespace {
constexpr int kBlock = 256;

__device__ __forceinline__ double block_sum(double v, double *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    return sh[0];
}

// MODE 0: x_out = x + omega Dinv (b - K x); 1: out = b - K x; 2: out = K x and x.(K x) partials
template <int N, int MODE>
__global__ void __launch_bounds__(kBlock) k_dg(DgLevel L, const double *__restrict__ x, const double *__restrict__ b,
                                               double *__restrict__ out, double *__restrict__ partial)
{
    constexpr int EPB = kBlock / N;
    __shared__ double sh[kBlock];
    const int t = threadIdx.x;
    const int le = t / N, i = t - le * N;
    const long long e = (long long)blockIdx.x * EPB + le;
    const bool act = (le < EPB) && e < L.Ne;
    double acc = 0.0;
    if (act && x) {
        const double *Kd = L.Kd + e * N * N;
        const double *xe = x + e * N;
#pragma unroll
        for (int j = 0; j < N; j++) acc = fma(Kd[j * N + i], xe[j], acc);
        for (int s = 0; s < L.ns; s++) {
            const int o = L.nb[e * L.ns + s];
            if (o < 0) continue;
            const double *Ko = L.Ko + (e * L.ns + s) * N * N;
            const double *xo = x + (long long)o * N;
#pragma unroll
            for (int j = 0; j < N; j++) acc = fma(Ko[j * N + i], xo[j], acc);
        }
    }
    if constexpr (MODE == 0) {
        sh[t] = act ? b[e * N + i] - acc : 0.0;
        __syncthreads();
        if (act) {
            const double *Di = L.Dinv + e * N * N;
            double v = 0.0;
#pragma unroll
            for (int j = 0; j < N; j++) v = fma(Di[j * N + i], sh[le * N + j], v);
            out[e * N + i] = (x ? x[e * N + i] : 0.0) + L.omega * v;
        }
    } else if constexpr (MODE == 1) {
        if (act) out[e * N + i] = b[e * N + i] - acc;
    } else {
        if (act) out[e * N + i] = acc;
        const double tot = block_sum(act ? acc * x[e * N + i] : 0.0, sh);
        if (t == 0) partial[blockIdx.x] = tot;
    }
}

template <int MODE>
hipError_t launch_dg(const DgLevel &L, const double *x, const double *b, double *out, double *partial, hipStream_t s)
{
    const int epb = kBlock / L.n;
    const unsigned grid = (unsigned)((L.Ne + epb - 1) / epb);
    switch (L.n) {
    case 1: hipLaunchKernelGGL((k_dg<1, MODE>), dim3(grid), dim3(kBlock), 0, s, L, x, b, out, partial); break;
    case 4: hipLaunchKernelGGL((k_dg<4, MODE>), dim3(grid), dim3(kBlock), 0, s, L, x, b, out, partial); break;
    case 8: hipLaunchKernelGGL((k_dg<8, MODE>), dim3(grid), dim3(kBlock), 0, s, L, x, b, out, partial); break;
    case 9: hipLaunchKernelGGL((k_dg<9, MODE>), dim3(grid), dim3(kBlock), 0, s, L, x, b, out, partial); break;
    case 12: hipLaunchKernelGGL((k_dg<12, MODE>), dim3(grid), dim3(kBlock), 0, s, L, x, b, out, partial); break;
    case 32: hipLaunchKernelGGL((k_dg<32, MODE>), dim3(grid), dim3(kBlock), 0, s, L, x, b, out, partial); break;
    case 27: hipLaunchKernelGGL((k_dg<27, MODE>), dim3(grid), dim3(kBlock), 0, s, L, x, b, out, partial); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

__device__ __forceinline__ void coords3(long long v, const int *nv, int *c)
{
    c[0] = (int)(v % nv[0]);
    c[1] = (int)((v / nv[0]) % nv[1]);
    c[2] = (int)(v / ((long long)nv[0] * nv[1]));
}

__global__ void __launch_bounds__(kBlock) k_dg_restrict(DgLevel L, const double *__restrict__ r, double *__restrict__ rv)
{
    const int nv[3] = {L.nel[0] + 1, L.dim > 1 ? L.nel[1] + 1 : 1, L.dim > 2 ? L.nel[2] + 1 : 1};
    const long long Nv = (long long)nv[0] * nv[1] * nv[2];
    const long long v = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (v >= Nv) return;
    int c[3];
    coords3(v, nv, c);
    double acc = 0.0;
    const int bz = L.dim > 2 ? 2 : 1, by = L.dim > 1 ? 2 : 1;
    for (int k = 0; k < bz; k++)
        for (int j = 0; j < by; j++)
            for (int i = 0; i < 2; i++) {
                const int ex = c[0] - i, ey = c[1] - j, ez = c[2] - k;
                if (ex < 0 || ex >= L.nel[0] || ey < 0 || ey >= L.nel[1] || ez < 0 || ez >= L.nel[2]) continue;
                const long long e = ex + (long long)L.nel[0] * (ey + (long long)L.nel[1] * ez);
                const int corner = i | (j << 1) | (k << 2);
                const double *re = r + e * L.n;
                for (int d = 0; d < L.n; d++) acc = fma(L.P[d * L.nc + corner], re[d], acc);
            }
    rv[v] = acc;
}

__global__ void __launch_bounds__(kBlock) k_dg_prolong(DgLevel L, const double *__restrict__ xv, double *__restrict__ x)
{
    const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (long long)L.Ne * L.n) return;
    const long long e = t / L.n;
    const int i = (int)(t - e * L.n);
    const int ec[3] = {(int)(e % L.nel[0]), (int)((e / L.nel[0]) % L.nel[1]), (int)(e / ((long long)L.nel[0] * L.nel[1]))};
    const int nv0 = L.nel[0] + 1, nv1 = L.dim > 1 ? L.nel[1] + 1 : 1;
    double acc = 0.0;
    for (int corner = 0; corner < L.nc; corner++) {
        const long long v = (ec[0] + (corner & 1)) + (long long)nv0 * ((ec[1] + ((corner >> 1) & 1)) +
                                                                       (long long)nv1 * (ec[2] + ((corner >> 2) & 1)));
        acc = fma(L.P[i * L.nc + corner], xv[v], acc);
    }
    x[t] += acc;
}

// MODE 0: x_out = x + omega dinv (b - A x); 1: out = b - A x
template <int MODE>
__global__ void __launch_bounds__(kBlock) k_st(StLevel L, const double *__restrict__ x, const double *__restrict__ b,
                                               double *__restrict__ out)
{
    const long long v = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (v >= L.N) return;
    int c[3];
    coords3(v, L.nv, c);
    double acc = 0.0;
    if (x) {
        const int S = L.dim > 2 ? 27 : 9;
        for (int k = 0; k < S; k++) {
            const int dx = k % 3 - 1, dy = (k / 3) % 3 - 1, dz = L.dim > 2 ? k / 9 - 1 : 0;
            const int xx = c[0] + dx, yy = c[1] + dy, zz = c[2] + dz;
            if (xx < 0 || xx >= L.nv[0] || yy < 0 || yy >= L.nv[1] || zz < 0 || zz >= L.nv[2]) continue;
            acc = fma(L.A[(long long)k * L.N + v], x[xx + (long long)L.nv[0] * (yy + (long long)L.nv[1] * zz)], acc);
        }
    }
    if constexpr (MODE == 0)
        out[v] = (x ? x[v] : 0.0) + L.omega * L.dinv[v] * (b[v] - acc);
    else
        out[v] = b[v] - acc;
}

__global__ void __launch_bounds__(kBlock) k_st_restrict(StLevel F, StLevel C, const double *__restrict__ rf,
                                                        double *__restrict__ rc)
{
    const long long I = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (I >= C.N) return;
    int c[3];
    coords3(I, C.nv, c);
    const int rz = F.dim > 2 ? 1 : 0, ry = F.dim > 1 ? 1 : 0;
    double acc = 0.0;
    for (int dz = -rz; dz <= rz; dz++)
        for (int dy = -ry; dy <= ry; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int xx = 2 * c[0] + dx, yy = 2 * c[1] + dy, zz = 2 * c[2] + dz;
                if (xx < 0 || xx >= F.nv[0] || yy < 0 || yy >= F.nv[1] || zz < 0 || zz >= F.nv[2]) continue;
                const double w = (dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0) * (dz ? 0.5 : 1.0);
                acc = fma(w, rf[xx + (long long)F.nv[0] * (yy + (long long)F.nv[1] * zz)], acc);
            }
    rc[I] = acc;
}

__global__ void __launch_bounds__(kBlock) k_st_prolong(StLevel F, StLevel C, const double *__restrict__ xc,
                                                       double *__restrict__ xf)
{
    const long long v = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (v >= F.N) return;
    int c[3];
    coords3(v, F.nv, c);
    // parents per axis: even index -> one (weight 1), odd -> two (weight 1/2)
    int lo[3], cnt[3];
    for (int d = 0; d < 3; d++) {
        lo[d] = c[d] >> 1;
        cnt[d] = (c[d] & 1) ? 2 : 1;
    }
    double acc = 0.0;
    for (int k = 0; k < cnt[2]; k++)
        for (int j = 0; j < cnt[1]; j++)
            for (int i = 0; i < cnt[0]; i++) {
                const double w = (cnt[0] == 2 ? 0.5 : 1.0) * (cnt[1] == 2 ? 0.5 : 1.0) * (cnt[2] == 2 ? 0.5 : 1.0);
                acc = fma(w, xc[(lo[0] + i) + (long long)C.nv[0] * ((lo[1] + j) + (long long)C.nv[1] * (lo[2] + k))], acc);
            }
    xf[v] += acc;
}

// one wave per row of the row-major inverse: lanes stride over the columns,
// then a wave reduction
__global__ void __launch_bounds__(kBlock) k_dense(long long N, const double *__restrict__ Ainv,
                                                  const double *__restrict__ b, double *__restrict__ x)
{
    const long long r = (long long)blockIdx.x * (kBlock / 64) + threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    if (r >= N) return;   // wave-uniform
    double acc = 0.0;
    for (long long c = lane; c < N; c += 64) acc = fma(Ainv[r * N + c], b[c], acc);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) x[r] = acc;
}

__global__ void __launch_bounds__(kBlock) k_dot(long long n, const double *__restrict__ a, const double *__restrict__ b,
                                                double *__restrict__ partial)
{
    __shared__ double sh[kBlock];
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        acc = fma(a[i], b[i], acc);
    const double tot = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kBlock) k_reduce(const double *__restrict__ partial, int nblk, double *sc, int slot)
{
    __shared__ double sh[kBlock];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nblk; i += kBlock) acc += partial[i];
    const double tot = block_sum(acc, sh);
    if (threadIdx.x == 0) sc[slot] = tot;
}

__global__ void __launch_bounds__(kBlock) k_cg_xr(long long n, const double *__restrict__ sc, int inom, int iden,
                                                  const double *__restrict__ p, const double *__restrict__ q,
                                                  double *__restrict__ x, double *__restrict__ r)
{
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double alpha = sc[inom] / sc[iden];
    x[i] = fma(alpha, p[i], x[i]);
    r[i] = fma(-alpha, q[i], r[i]);
}

__global__ void __launch_bounds__(kBlock) k_cg_p(long long n, const double *__restrict__ sc, int inew, int iold,
                                                 const double *__restrict__ z, double *__restrict__ p)
{
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double beta = sc[inew] / sc[iold];
    p[i] = fma(beta, p[i], z[i]);
}

__global__ void __launch_bounds__(kBlock) k_dsa_finish(long long n, const double *__restrict__ r,
                                                       const double *__restrict__ c, double *__restrict__ x,
                                                       double *__restrict__ partial)
{
    __shared__ double sh[kBlock];
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock) {
        x[i] += c[i];
        const double z = r[i] + c[i];
        acc = fma(z, z, acc);
    }
    const double tot = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

unsigned grid_of(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }
}  // namespace

int dg_blocks(const DgLevel &L) { return (L.Ne + kBlock / L.n - 1) / (kBlock / L.n); }

hipError_t dg_jacobi(const DgLevel &L, const double *x_in, const double *b, double *x_out, hipStream_t s)
{
    return launch_dg<0>(L, x_in, b, x_out, nullptr, s);
}

hipError_t dg_resid(const DgLevel &L, const double *x, const double *b, double *r, hipStream_t s)
{
    return launch_dg<1>(L, x, b, r, nullptr, s);
}

hipError_t dg_spmv_dot(const DgLevel &L, const double *x, double *y, double *partial, hipStream_t s)
{
    return launch_dg<2>(L, x, nullptr, y, partial, s);
}

hipError_t dg_restrict(const DgLevel &L, const double *r, double *rv, hipStream_t s)
{
    const long long Nv = (long long)(L.nel[0] + 1) * (L.dim > 1 ? L.nel[1] + 1 : 1) * (L.dim > 2 ? L.nel[2] + 1 : 1);
    hipLaunchKernelGGL(k_dg_restrict, dim3(grid_of(Nv)), dim3(kBlock), 0, s, L, r, rv);
    return hipGetLastError();
}

hipError_t dg_prolong_add(const DgLevel &L, const double *xv, double *x, hipStream_t s)
{
    hipLaunchKernelGGL(k_dg_prolong, dim3(grid_of((long long)L.Ne * L.n)), dim3(kBlock), 0, s, L, xv, x);
    return hipGetLastError();
}

hipError_t st_jacobi(const StLevel &L, const double *x_in, const double *b, double *x_out, hipStream_t s)
{
    hipLaunchKernelGGL(k_st<0>, dim3(grid_of(L.N)), dim3(kBlock), 0, s, L, x_in, b, x_out);
    return hipGetLastError();
}

hipError_t st_resid(const StLevel &L, const double *x, const double *b, double *r, hipStream_t s)
{
    hipLaunchKernelGGL(k_st<1>, dim3(grid_of(L.N)), dim3(kBlock), 0, s, L, x, b, r);
    return hipGetLastError();
}

hipError_t st_restrict(const StLevel &F, const StLevel &C, const double *rf, double *rc, hipStream_t s)
{
    hipLaunchKernelGGL(k_st_restrict, dim3(grid_of(C.N)), dim3(kBlock), 0, s, F, C, rf, rc);
    return hipGetLastError();
}

hipError_t st_prolong_add(const StLevel &F, const StLevel &C, const double *xc, double *xf, hipStream_t s)
{
    hipLaunchKernelGGL(k_st_prolong, dim3(grid_of(F.N)), dim3(kBlock), 0, s, F, C, xc, xf);
    return hipGetLastError();
}

hipError_t dense_apply(long long N, const double *Ainv, const double *b, double *x, hipStream_t s)
{
    hipLaunchKernelGGL(k_dense, dim3((unsigned)((N + kBlock / 64 - 1) / (kBlock / 64))), dim3(kBlock), 0, s, N, Ainv,
                       b, x);
    return hipGetLastError();
}

int dot_blocks(long long n) { return (int)std::min<long long>(2048, (n + kBlock - 1) / kBlock); }

hipError_t dot_partial(long long n, const double *a, const double *b, double *partial, hipStream_t s)
{
    hipLaunchKernelGGL(k_dot, dim3(dot_blocks(n)), dim3(kBlock), 0, s, n, a, b, partial);
    return hipGetLastError();
}

hipError_t reduce_to(const double *partial, int nblk, double *sc, int slot, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kBlock), 0, s, partial, nblk, sc, slot);
    return hipGetLastError();
}

hipError_t cg_update_xr(long long n, const double *sc, int inom, int iden, const double *p, const double *q, double *x,
                        double *r, hipStream_t s)
{
    hipLaunchKernelGGL(k_cg_xr, dim3(grid_of(n)), dim3(kBlock), 0, s, n, sc, inom, iden, p, q, x, r);
    return hipGetLastError();
}

hipError_t cg_update_p(long long n, const double *sc, int inew, int iold, const double *z, double *p, hipStream_t s)
{
    hipLaunchKernelGGL(k_cg_p, dim3(grid_of(n)), dim3(kBlock), 0, s, n, sc, inew, iold, z, p);
    return hipGetLastError();
}

hipError_t dsa_finish(long long n, const double *r, const double *c, double *x, double *partial, hipStream_t s)
{
    hipLaunchKernelGGL(k_dsa_finish, dim3(dot_blocks(n)), dim3(kBlock), 0, s, n, r, c, x, partial);
    return hipGetLastError();
}

}  // namespace allium
