// Helpers shared by the sweep kernels: element shapes, face-node maps and the
// unrolled register LU used for every cell solve.
#pragma once

#include <hip/hip_runtime.h>

namespace allium {

typedef double d2v __attribute__((ext_vector_type(2)));
// read-only tables seen through the constant address space: uniform
// addresses become scalar loads, non-uniform ones stay vector loads
typedef const __attribute__((address_space(4))) double *cdptr;
__device__ __forceinline__ cdptr as_const(const double *p) { return (cdptr)p; }
// Global-address-space views of pointers the compiler cannot prove global
// (pointers read from device memory are generic): loads through them are
// global_load instructions instead of flat ones, which also count against the
// LDS wait counter and so make every LDS wait drain them.
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gptr<T> as_global(T *p) { return (gptr<T>)p; }

template <int DIM, int P>
struct Sh {
    static constexpr int Q = P + 1;
    static constexpr int N = DIM == 2 ? Q * Q : Q * Q * Q;
    static constexpr int NF = DIM == 2 ? Q : Q * Q;
    static constexpr int NS = 2 * DIM;
};

// k-th node (face-lexicographic) of element side s = 2*axis + {0 low, 1 high}
template <int DIM, int P>
__host__ __device__ constexpr int side_node(int s, int k)
{
    const int ax = s >> 1, hi = s & 1, Q = P + 1;
    int idx[3] = {0, 0, 0};
    int r = k;
    for (int d = 0; d < DIM; ++d) {
        if (d == ax) idx[d] = hi ? P : 0;
        else { idx[d] = r % Q; r /= Q; }
    }
    return idx[0] + Q * (idx[1] + Q * idx[2]);
}

__device__ __forceinline__ int fast_div(int a, int d, float inv)
{
    int q = (int)((float)a * inv);
    const int r = a - q * d;
    if (r < 0) q--;
    else if (r >= d) q++;
    return q;
}

__device__ __forceinline__ double fast_rcp(double a)
{
    double r = __builtin_amdgcn_rcp(a);
    double e = fma(-a, r, 1.0);
    r = fma(r, e, r);
    e = fma(-a, r, 1.0);
    return fma(r, e, r);
}

__device__ __forceinline__ float fast_rcp(float a)
{
    const float r = __builtin_amdgcn_rcpf(a);
    return fmaf(r, fmaf(-a, r, 1.0f), r);
}

template <int N, typename T>
__device__ __forceinline__ void lu_factor_solve(T (&A)[N][N], T (&x)[N])
{
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const T inv = fast_rcp(A[k][k]);
        A[k][k] = inv;
#pragma unroll
        for (int i = k + 1; i < N; ++i) {
            const T l = A[i][k] * inv;
            A[i][k] = l;
#pragma unroll
            for (int j = k + 1; j < N; ++j) A[i][j] = fma(-l, A[k][j], A[i][j]);
            x[i] = fma(-l, x[k], x[i]);
        }
    }
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
        T s = x[i];
#pragma unroll
        for (int j = i + 1; j < N; ++j) s = fma(-A[i][j], x[j], s);
        x[i] = s * A[i][i];
    }
}

template <int N, typename T>
__device__ __forceinline__ void lu_resolve(const T (&A)[N][N], T (&x)[N])
{
#pragma unroll
    for (int k = 0; k < N; ++k)
#pragma unroll
        for (int i = k + 1; i < N; ++i) x[i] = fma(-A[i][k], x[k], x[i]);
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
        T s = x[i];
#pragma unroll
        for (int j = i + 1; j < N; ++j) s = fma(-A[i][j], x[j], s);
        x[i] = s * A[i][i];
    }
}

// is local node r on side s, and its face-lexicographic index there
template <int DIM, int P>
__device__ __forceinline__ bool on_side(int r, int s, int &k)
{
    constexpr int Q = P + 1;
    const int i0 = r % Q, i1 = (r / Q) % Q, i2 = r / (Q * Q);
    const int ax = s >> 1, want = (s & 1) ? P : 0;
    const int ia = ax == 0 ? i0 : (ax == 1 ? i1 : i2);
    if (DIM == 2) k = ax == 0 ? i1 : i0;
    else k = ax == 0 ? i1 + Q * i2 : (ax == 1 ? i0 + Q * i2 : i0 + Q * i1);
    return ia == want;
}

// 1D mass matrix of the Gauss-Lobatto Lagrange basis of degree P on [0, 1]
// (exact integrals: P=1 h/6 [[2, 1], [1, 2]]; P=2 nodes 0, 1/2, 1:
// h/30 [[4, 2, -1], [2, 16, 2], [-1, 2, 4]]).
template <int P>
__host__ __device__ constexpr double unit_mass_1d(int i, int j)
{
    if constexpr (P == 0) {
        return 1.0;
    } else if constexpr (P == 1) {
        return i == j ? 1.0 / 3.0 : 1.0 / 6.0;
    } else {
        constexpr double m[3][3] = {{4.0 / 30.0, 2.0 / 30.0, -1.0 / 30.0},
                                    {2.0 / 30.0, 16.0 / 30.0, 2.0 / 30.0},
                                    {-1.0 / 30.0, 2.0 / 30.0, 4.0 / 30.0}};
        return m[i][j];
    }
}

// Element mass of an axis-aligned box of unit volume: the tensor product of
// the 1D masses over the DIM lexicographic node coordinates.
template <int DIM, int P>
__host__ __device__ constexpr double unit_mass(int i, int j)
{
    constexpr int Q = P + 1;
    double v = 1.0;
    for (int d = 0; d < DIM; ++d) {
        v *= unit_mass_1d<P>(i % Q, j % Q);
        i /= Q;
        j /= Q;
    }
    return v;
}

// Face mass of a face of unit area in face-lexicographic numbering.
template <int DIM, int P>
__host__ __device__ constexpr double unit_face_mass(int k, int l)
{
    return DIM == 1 ? 1.0 : unit_mass<DIM - 1, P>(k, l);
}

}  // namespace allium
