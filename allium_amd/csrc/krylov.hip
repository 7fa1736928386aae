// Vector kernels of the GMRES / FGMRES outer solvers.
//
// Each Arnoldi step of the outer solver orthogonalises w = T v_i (or T z_i)
// against v_0..v_i by modified Gram-Schmidt.  Every projection coefficient is
// a global dot product; instead of returning it to the host, the kernel that
// applies coefficient k also forms the partial sums of coefficient k + 1, and
// every block of the next launch re-sums those partials itself (in the same
// fixed order, so all blocks agree bit for bit).  One launch per projection,
// no host round trip until the Hessenberg column is complete.
//
// The vectors are phi-sized (2.1e6 doubles for 64^3 p = 1): each launch
// streams w, v_{k-1} and v_k once, HBM-bound.

#include "krylov.hpp"

namespace allium {

namespace {

constexpr int kKB = 256;           // threads per block
constexpr int kMaxBlocks = 1024;   // partial sums per reduction
constexpr int kMaxBasisUpdate = 64;

__device__ double block_sum(double v, double *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int st = kKB / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__device__ double sum_partials(const double *part, int nb, double *sh)
{
    double a = 0.0;
    for (int i = threadIdx.x; i < nb; i += kKB) a += part[i];
    return block_sum(a, sh);
}

__global__ __launch_bounds__(kKB) void k_mgs(long n, const double *__restrict__ part_in, int nb,
                                             double *__restrict__ h_out, const double *__restrict__ vprev,
                                             const double *__restrict__ vcur, double *__restrict__ w,
                                             double *__restrict__ part_out)
{
    __shared__ double sh[kKB];
    double h = 0.0;
    if (vprev) {
        h = sum_partials(part_in, nb, sh);
        if (blockIdx.x == 0 && threadIdx.x == 0) h_out[0] = h;
    }
    double acc = 0.0;
    const long stride = (long)gridDim.x * kKB;
    for (long i = (long)blockIdx.x * kKB + threadIdx.x; i < n; i += stride) {
        double wi = w[i];
        if (vprev) {
            wi = fma(-h, vprev[i], wi);
            w[i] = wi;
        }
        acc = fma(wi, vcur ? vcur[i] : wi, acc);
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part_out[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kKB) void k_normalize(long n, const double *__restrict__ part_in, int nb,
                                                   double *__restrict__ h_out, const double *__restrict__ w,
                                                   double *__restrict__ vout)
{
    __shared__ double sh[kKB];
    const double h = sqrt(sum_partials(part_in, nb, sh));
    if (blockIdx.x == 0 && threadIdx.x == 0) h_out[0] = h;
    const double inv = 1.0 / h;
    const long stride = (long)gridDim.x * kKB;
    for (long i = (long)blockIdx.x * kKB + threadIdx.x; i < n; i += stride) vout[i] = inv * w[i];
}

__global__ __launch_bounds__(kKB) void k_basis_update(long n, int k, const double *__restrict__ V, long ld,
                                                      const double *__restrict__ y, double *__restrict__ x)
{
    __shared__ double ys[kMaxBasisUpdate];
    if ((int)threadIdx.x < k) ys[threadIdx.x] = y[threadIdx.x];
    __syncthreads();
    const long stride = (long)gridDim.x * kKB;
    for (long i = (long)blockIdx.x * kKB + threadIdx.x; i < n; i += stride) {
        double a = x[i];
        for (int j = 0; j < k; j++) a = fma(ys[j], V[(long)j * ld + i], a);
        x[i] = a;
    }
}

}  // namespace

int krylov_blocks(long n)
{
    long nb = (n + 4L * kKB - 1) / (4L * kKB);
    if (nb < 1) nb = 1;
    if (nb > kMaxBlocks) nb = kMaxBlocks;
    return (int)nb;
}

hipError_t launch_mgs(long n, const double *part_in, int nb, double *h_out, const double *vprev,
                      const double *vcur, double *w, double *part_out, hipStream_t s)
{
    const int grid = krylov_blocks(n);
    hipLaunchKernelGGL(k_mgs, dim3(grid), dim3(kKB), 0, s, n, part_in, nb, h_out, vprev, vcur, w, part_out);
    return hipGetLastError();
}

hipError_t launch_normalize(long n, const double *part_in, int nb, double *h_out, const double *w, double *vout,
                            hipStream_t s)
{
    const int grid = krylov_blocks(n);
    hipLaunchKernelGGL(k_normalize, dim3(grid), dim3(kKB), 0, s, n, part_in, nb, h_out, w, vout);
    return hipGetLastError();
}

hipError_t launch_basis_update(long n, int k, const double *V, long ld, const double *y, double *x, hipStream_t s)
{
    if (k < 1) return hipSuccess;
    if (k > kMaxBasisUpdate) return hipErrorInvalidValue;
    const int grid = krylov_blocks(n);
    hipLaunchKernelGGL(k_basis_update, dim3(grid), dim3(kKB), 0, s, n, k, V, ld, y, x);
    return hipGetLastError();
}

}  // namespace allium
