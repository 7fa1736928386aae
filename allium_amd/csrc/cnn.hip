// Learned low-order flux correction: a 5-layer 3x3x3 convolutional network
// on the element-centred grid, with BF16 MFMA forward and backward kernels
// (FP32 accumulation) for gfx950 and an FP32 Adam optimiser.  No counterpart
// in the reference; it stands in for the diffusion synthetic acceleration of
// src/transport_op.hpp:38-59 inside the source iteration.
//
// Convolution as an implicit GEMM on v_mfma_f32_16x16x32_bf16: a wave owns 4
// tiles of 16 consecutive voxels (rows of the A tiles) and all output channels
// (N tiles of 16); the K dimension runs over (kernel offset, input channel) in
// chunks of 8 channels of one offset, so every A operand is one 16-byte load
// of a channels-last voxel row.  Lane l holds A[l % 16][8 (l / 16) + t],
// B[8 (l / 16) + t][l % 16] and D[4 (l / 16) + r][l % 16].
//
// Weight gradients: a workgroup owns one (sample, z) plane; for each row and
// 32-voxel chunk it stages the output gradient (transposed, channel-major)
// and the 9 neighbouring input rows with a one-voxel x halo in LDS, then
// accumulates all 27 offsets of its (output-tile, input-tile) pair with
// voxels as the K dimension.  Per-plane partials are reduced in a second pass
// in a fixed order, so gradients are deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/allium_cnn.h"
#include "cnn_internal.hpp"

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NL = 5;
constexpr int CI_[NL] = {3, 32, 32, 32, 32};
constexpr int CO_[NL] = {32, 32, 32, 32, 1};
constexpr int CIP_[NL] = {8, 32, 32, 32, 32};
constexpr int COP_[NL] = {32, 32, 32, 32, 16};
constexpr int HID = 32;

struct Geo {
    int B, nz, ny, nx;
    long long T;   // voxels
    int dil;       // dilation of the 3x3x3 taps: offsets are dil * (-1, 0, 1) per axis
};
constexpr int kMaxDil = 8;

// ---------------------------------------------------------------------------
// convolution (forward, and backward-data with flipped/transposed weights)
// MODE 0: + bias, ReLU, bf16 out [T][COP]
// MODE 1: + bias, FP32 out [T] (channel 0)
// MODE 2: masked by (mask[T][COP] > 0), bf16 out [T][COP]
// ---------------------------------------------------------------------------
// A workgroup computes one 8 x 8 x 8 tile of voxels with eight waves; a wave
// owns MT = 4 16-voxel row tiles (two y rows of one z plane each), so every B
// operand it loads (a 16-byte weight row) feeds MT MFMAs.  A dilated 3x3x3
// stencil couples only voxels of one sub-grid (x mod d, y mod d, z mod d),
// where it is an undilated stencil: the tile is taken in one sub-grid and its
// 10 x 10 x 10 halo is loaded into LDS once, so each input voxel feeds its 27
// taps from LDS instead of 27 cache reads.  Each output element accumulates
// its K steps (tap-major, then 8-channel chunks) in order.
constexpr int kConvMT = 4;
constexpr int kTX = 8, kTY = 8, kTZ = 8, kHX = kTX + 2, kHY = kTY + 2, kHZ = kTZ + 2;
constexpr int kHalo = kHX * kHY * kHZ;

struct TileGeo {
    int nsx, nsy, nsz;   // sub-grid extents (the largest, at offset 0)
    int ntx, nty, ntz;   // tiles per sub-grid axis
};

__host__ __device__ inline TileGeo tile_geo(const Geo &g)
{
    TileGeo t;
    t.nsx = (g.nx + g.dil - 1) / g.dil;
    t.nsy = (g.ny + g.dil - 1) / g.dil;
    t.nsz = (g.nz + g.dil - 1) / g.dil;
    t.ntx = (t.nsx + kTX - 1) / kTX;
    t.nty = (t.nsy + kTY - 1) / kTY;
    t.ntz = (t.nsz + kTZ - 1) / kTZ;
    return t;
}

template <int CIP, int COP, int MODE>
__global__ void __launch_bounds__(512) k_conv_t(Geo g, const bf16 *__restrict__ in, const bf16 *__restrict__ W,
                                                const float *__restrict__ bias, bf16 *__restrict__ out_bf,
                                                float *__restrict__ out_f, const bf16 *__restrict__ mask)
{
    constexpr int CH = CIP / 8, NCH = 27 * CH, KS = (NCH + 3) / 4, NT = COP / 16, MT = kConvMT;
    static_assert(CIP % 8 == 0 && COP % 16 == 0, "padded channel counts");
    // voxel stride in LDS: the channels, unpadded (timed against 16, 32 and 48
    // bytes of padding per voxel, which were 28 % slower in a 64^3 inference)
    constexpr int VS = CIP * 2;
    static_assert(MT * 16 * 8 == kTX * kTY * kTZ, "a workgroup's voxels are one tile");
    extern __shared__ __align__(16) unsigned char halo[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, grp = lane >> 4;
    const TileGeo tg = tile_geo(g);
    // workgroup -> (batch, sub-grid offset, tile)
    long long q = blockIdx.x;
    const int tx = (int)(q % tg.ntx); q /= tg.ntx;
    const int ty = (int)(q % tg.nty); q /= tg.nty;
    const int tz = (int)(q % tg.ntz); q /= tg.ntz;
    const int d = g.dil;
    const int sx = (int)(q % d); q /= d;
    const int sy = (int)(q % d); q /= d;
    const int sz = (int)(q % d); q /= d;
    const long long b = q;
    const long long plane = (long long)g.nx * g.ny, vol = plane * g.nz;
    // halo: CH 16-byte chunks per voxel, zeros outside the grid
    for (int c = threadIdx.x; c < kHalo * CH; c += 512) {
        const int hv = c / CH, part = c - hv * CH;
        const int hx = hv % kHX, hy = (hv / kHX) % kHY, hz = hv / (kHX * kHY);
        const int x = sx + d * (tx * kTX + hx - 1), y = sy + d * (ty * kTY + hy - 1), z = sz + d * (tz * kTZ + hz - 1);
        const bool ok = x >= 0 && x < g.nx && y >= 0 && y < g.ny && z >= 0 && z < g.nz;
        const long long m = ok ? b * vol + (long long)z * plane + (long long)y * g.nx + x : 0;
        const bf16x8 v = *reinterpret_cast<const bf16x8 *>(in + m * CIP + part * 8);
        const bf16x8 zero8 = {};
        *reinterpret_cast<bf16x8 *>(halo + hv * VS + part * 16) = ok ? v : zero8;
    }
    __syncthreads();
    // this lane's operand rows: voxel (i, j, k) = (r & 7, 2 t + (r >> 3), w) of the tile
    int hbase[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) {
        const int r = lane & 15;
        const int i = r & 7, j = 2 * t + (r >> 3), k = w;
        hbase[t] = ((k + 1) * kHY + (j + 1)) * kHX + (i + 1);
    }
    f32x4 acc[MT][NT];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[i][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bf16x8 zero8 = {};
    for (int s = 0; s < KS; ++s) {
        const int c = 4 * s + grp;
        const bool cv = c < NCH;
        const int cc = cv ? c : 0;
        const int o = cc / CH;
        const int ci0 = (cc - o * CH) * 8;
        const int dz = o / 9 - 1, dy = (o / 3) % 3 - 1, dx = o % 3 - 1;
        const int doff = (dz * kHY + dy) * kHX + dx;
        bf16x8 a[MT];
#pragma unroll
        for (int i = 0; i < MT; ++i) {
            const bf16x8 v = *reinterpret_cast<const bf16x8 *>(halo + (hbase[i] + doff) * VS + ci0 * 2);
            a[i] = cv ? v : zero8;
        }
        bf16x8 bw[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const bf16x8 v = *reinterpret_cast<const bf16x8 *>(W + ((long long)o * COP + nt * 16 + (lane & 15)) * CIP + ci0);
            bw[nt] = cv ? v : zero8;
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int i = 0; i < MT; ++i) acc[i][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], bw[nt], acc[i][nt], 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int jc = nt * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                // output row 4 grp + r of operand tile t
                const int v = 4 * grp + r;
                const int i = v & 7, j = 2 * t + (v >> 3), k = w;
                const int x = sx + d * (tx * kTX + i), y = sy + d * (ty * kTY + j), z = sz + d * (tz * kTZ + k);
                if (x >= g.nx || y >= g.ny || z >= g.nz) continue;
                const long long mo = b * vol + (long long)z * plane + (long long)y * g.nx + x;
                float val = acc[t][nt][r];
                if constexpr (MODE == 0) {
                    val = fmaxf(val + bias[jc], 0.f);
                    out_bf[mo * COP + jc] = (bf16)val;
                } else if constexpr (MODE == 1) {
                    if (jc == 0) out_f[mo] = val + bias[0];
                } else {
                    val = (float)mask[mo * COP + jc] > 0.f ? val : 0.f;
                    out_bf[mo * COP + jc] = (bf16)val;
                }
            }
        }
}

template <int CIP, int COP, int MODE>
void launch_conv(const Geo &g, const bf16 *in, const bf16 *W, const float *bias, bf16 *out_bf, float *out_f,
                 const bf16 *mask, hipStream_t s)
{
    const TileGeo tg = tile_geo(g);
    const long long grid = (long long)g.B * g.dil * g.dil * g.dil * tg.ntx * tg.nty * tg.ntz;
    const size_t lds = (size_t)kHalo * CIP * 2;
    hipLaunchKernelGGL((k_conv_t<CIP, COP, MODE>), dim3((unsigned)grid), dim3(512), lds, s, g, in, W, bias, out_bf,
                       out_f, mask);
}

// ---------------------------------------------------------------------------
// weight gradient partials: part[plane][27][COP][CIT]
// ---------------------------------------------------------------------------
// One workgroup per (batch, z) plane; K steps of 32 voxels along x, row by row.
// Each step stages the step's gradient rows transposed (gT[co][x]) and the
// nine (dz, dy) input rows with a dil-voxel halo each side (aT[r9][ci][x]) in
// LDS.  The global reads are 16-byte channel chunks; those of the next step are
// issued into registers before this step's MFMAs, so their latency hides
// behind them.  A thread stages two neighbouring input voxels, so each channel
// pair goes to LDS as one 4-byte store.  Row strides are padded: gT's keeps
// 16-byte operand reads aligned and spreads its 16 rows over the banks, aT's is
// an odd number of dwords.
template <int CIP, int COP>
__global__ void __launch_bounds__(256) k_wgrad(Geo g, const bf16 *__restrict__ gz, const bf16 *__restrict__ a,
                                               float *__restrict__ part)
{
    constexpr int NTO = COP / 16, NTI = (CIP + 15) / 16, CIT = NTI * 16, NPAIR = NTO * NTI;
    constexpr int XW = 32;                        // voxels per K step
    constexpr int XSMAX = XW + 2 * kMaxDil;       // staged x range (dil-voxel halo each side)
    constexpr int GS = XW + 8, AS = XSMAX + 2;    // LDS row strides (bf16)
    constexpr int CHI = CIP / 8, CHO = COP / 8;   // 16-byte chunks per voxel
    constexpr int NA = (9 * (XSMAX / 2) * CHI + 255) / 256, NG = (XW * CHO + 255) / 256;
    static_assert(NPAIR == 1 || NPAIR == 2 || NPAIR == 4, "tile pairs divide the four waves");
    // The waves take the NPAIR (16 co x 16 ci) tile pairs.  With 8 input
    // channels, the waves left over share a pair and split its 27 taps (input
    // layer: 126 -> 86 us on 8 x 32^3); with 32 they stay idle, which timed
    // faster than splitting.
    constexpr int WPP = CIP <= 8 ? 4 / NPAIR : 1;
    __shared__ __attribute__((aligned(16))) bf16 gT[COP][GS];
    __shared__ __attribute__((aligned(16))) bf16 aT[9][CIP][AS];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, grp = lane >> 4;
    const int plane = blockIdx.x;
    const int b = plane / g.nz, z = plane - b * g.nz;
    const bool active = w < NPAIR * WPP;
    const int pw = WPP == 1 ? w : w % NPAIR, tap0 = WPP == 1 ? 0 : w / NPAIR;   // taps o = tap0 mod WPP
    const int cot = active ? pw / NTI : 0, cit = active ? pw % NTI : 0;
    const int dl = g.dil, xr = XW + 2 * dl;
    const int nxs = (g.nx + XW - 1) / XW, nsteps = g.ny * nxs;
    const long long plane_base = ((long long)b * g.nz + z) * g.ny * g.nx;
    const bf16x8 zero8 = {};
    bf16x8 ra[NA][2], rg[NG];
    const int xh = xr / 2;
    // the step's global reads into registers; out-of-grid and padding chunks read
    // voxel 0 and are zeroed by a select
    auto load = [&](int st) {
        const int y = st / nxs, x0 = (st - y * nxs) * XW;
#pragma unroll
        for (int k = 0; k < NG; ++k) {
            const int i = threadIdx.x + 256 * k;
            const int xi = i / CHO, c = i - xi * CHO;
            const bool ok = i < XW * CHO && x0 + xi < g.nx;
            const long long v = ok ? plane_base + (long long)y * g.nx + x0 + xi : 0;
            const bf16x8 r = *reinterpret_cast<const bf16x8 *>(gz + v * COP + 8 * c);
            rg[k] = ok ? r : zero8;
        }
#pragma unroll
        for (int k = 0; k < NA; ++k) {
            const int i = threadIdx.x + 256 * k;
            const int c = i % CHI, rest = i / CHI;
            const int xp = rest % xh, r9 = rest / xh;
            const int yy = y + (r9 % 3 - 1) * dl, zz = z + (r9 / 3 - 1) * dl;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int xx = x0 - dl + 2 * xp + h;
                const bool ok = r9 < 9 && xx >= 0 && xx < g.nx && yy >= 0 && yy < g.ny && zz >= 0 && zz < g.nz;
                const long long v = ok ? (((long long)b * g.nz + zz) * g.ny + yy) * g.nx + xx : 0;
                const bf16x8 r = *reinterpret_cast<const bf16x8 *>(a + v * CIP + 8 * c);
                ra[k][h] = ok ? r : zero8;
            }
        }
    };
    f32x4 acc[27];
#pragma unroll
    for (int o = 0; o < 27; ++o) acc[o] = f32x4{0.f, 0.f, 0.f, 0.f};
    if (nsteps > 0) load(0);
    for (int st = 0; st < nsteps; ++st) {
        __syncthreads();   // the previous step's operand reads are done
#pragma unroll
        for (int k = 0; k < NG; ++k) {
            const int i = threadIdx.x + 256 * k;
            if (i < XW * CHO) {
                const int xi = i / CHO, c = i - xi * CHO;
#pragma unroll
                for (int t = 0; t < 8; ++t) gT[8 * c + t][xi] = rg[k][t];
            }
        }
#pragma unroll
        for (int k = 0; k < NA; ++k) {
            const int i = threadIdx.x + 256 * k;
            const int c = i % CHI, rest = i / CHI;
            const int xp = rest % xh, r9 = rest / xh;
            if (r9 < 9) {
#pragma unroll
                for (int t = 0; t < 8; ++t)
                    *reinterpret_cast<bf16x2 *>(&aT[r9][8 * c + t][2 * xp]) = bf16x2{ra[k][0][t], ra[k][1][t]};
            }
        }
        __syncthreads();
        if (st + 1 < nsteps) load(st + 1);
        if (active) {
            const bf16x8 av = *reinterpret_cast<const bf16x8 *>(&gT[cot * 16 + (lane & 15)][8 * grp]);
            const int ci = cit * 16 + (lane & 15);
#pragma unroll
            for (int o = 0; o < 27; ++o) {
                if (WPP > 1 && o % WPP != tap0) continue;
                const int dz = o / 9 - 1, dy = (o / 3) % 3 - 1, dx = o % 3 - 1;
                const int r9 = (dz + 1) * 3 + (dy + 1);
                bf16x8 bv;
#pragma unroll
                for (int t = 0; t < 8; ++t) bv[t] = ci < CIP ? aT[r9][ci][dl * (1 + dx) + 8 * grp + t] : (bf16)0.f;
                acc[o] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[o], 0, 0, 0);
            }
        }
    }
    if (!active) return;
#pragma unroll
    for (int o = 0; o < 27; ++o) {
        if (WPP > 1 && o % WPP != tap0) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int co = cot * 16 + 4 * grp + r, ci = cit * 16 + (lane & 15);
            part[(((long long)plane * 27 + o) * COP + co) * CIT + ci] = acc[o][r];
        }
    }
}

// grad[o][co][ci] (packed, real sizes) = sum over planes of the partials.  A
// block sums 64 outputs: each of its four waves adds every fourth plane (loads
// unrolled eight deep) and the four sums are added in wave order, so the result
// does not depend on timing.
__global__ void __launch_bounds__(256) k_wgrad_reduce(const float *__restrict__ part, int nplane, int COP, int CIT,
                                                      int CO, int CI, float *__restrict__ grad)
{
    __shared__ float red[4][64];
    const int n = 27 * CO * CI;
    const int jl = threadIdx.x & 63, sl = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + jl;
    float s = 0.f;
    if (i < n) {
        const int o = i / (CO * CI), rem = i - o * CO * CI, co = rem / CI, ci = rem - co * CI;
        const long long stride = 27LL * COP * CIT;
        const float *p0 = part + ((long long)o * COP + co) * CIT + ci;
#pragma unroll 8
        for (int p = sl; p < nplane; p += 4) s += p0[p * stride];
    }
    red[sl][jl] = s;
    __syncthreads();
    if (sl == 0 && i < n) grad[i] = ((red[0][jl] + red[1][jl]) + red[2][jl]) + red[3][jl];
}

// bias gradient partials: bpart[block][COP] = sum over the block's voxels.  A
// thread reads one 16-byte chunk (8 channels) of a voxel per step; the block's
// per-thread sums are added in a fixed order.
template <int COP>
__global__ void __launch_bounds__(256) k_bias_grad(long long T, const bf16 *__restrict__ gz, float *__restrict__ bpart)
{
    constexpr int CHO = COP / 8, VPB = 256 / CHO;
    __shared__ float red[256][9];
    const int c = threadIdx.x % CHO, vi = threadIdx.x / CHO;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
    for (long long v = (long long)blockIdx.x * VPB + vi; v < T; v += (long long)gridDim.x * VPB) {
        const bf16x8 r = *reinterpret_cast<const bf16x8 *>(gz + v * COP + 8 * c);
#pragma unroll
        for (int t = 0; t < 8; ++t) s[t] += (float)r[t];
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) red[threadIdx.x][t] = s[t];
    __syncthreads();
    if (threadIdx.x < COP) {
        const int cc = threadIdx.x >> 3, t = threadIdx.x & 7;
        float a = 0.f;
        for (int k = 0; k < VPB; ++k) a += red[k * CHO + cc][t];
        bpart[(long long)blockIdx.x * COP + threadIdx.x] = a;
    }
}

// grad[co] = sum over the partials of k_bias_grad: eight slices of 32 threads
// (one per channel) add every eighth partial, then the slices in order
__global__ void __launch_bounds__(256) k_bias_reduce(const float *__restrict__ bpart, int nblk, int COP, int CO,
                                                     float *__restrict__ grad)
{
    static_assert(HID <= 32, "one thread per channel in a 32-thread slice");
    __shared__ float red[8][32];
    const int co = threadIdx.x & 31, sl = threadIdx.x >> 5;
    float s = 0.f;
    if (co < CO) {
#pragma unroll 8
        for (int b = sl; b < nblk; b += 8) s += bpart[(long long)b * COP + co];
    }
    red[sl][co] = s;
    __syncthreads();
    if (sl == 0 && co < CO) {
        float a = 0.f;
#pragma unroll
        for (int k = 0; k < 8; ++k) a += red[k][co];
        grad[co] = a;
    }
}

// x [T][3] float -> [T][8] bf16
__global__ void k_pack_input(long long T, const float *__restrict__ x, bf16 *__restrict__ xb)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < T * 8; i += (long long)gridDim.x * blockDim.x) {
        const long long v = i >> 3;
        const int c = (int)(i & 7);
        xb[i] = c < 3 ? (bf16)x[v * 3 + c] : (bf16)0.f;
    }
}

// master FP32 layer (W[27][CO][CI], b[CO]) -> bf16 forward [27][COP][CIP],
// bf16 backward-data [27][CIP][COP] (offset flipped), FP32 bias [COP]
__global__ void k_pack_layer(const float *__restrict__ Wm, const float *__restrict__ bm, int CO, int CI, int COP,
                             int CIP, bf16 *__restrict__ Wf, bf16 *__restrict__ Wb, float *__restrict__ bias)
{
    const int n = 27 * COP * CIP;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int o = i / (COP * CIP), rem = i - o * COP * CIP, co = rem / CIP, ci = rem - co * CIP;
        const float v = (co < CO && ci < CI) ? Wm[(o * CO + co) * CI + ci] : 0.f;
        Wf[i] = (bf16)v;
        Wb[((26 - o) * CIP + ci) * COP + co] = (bf16)v;
        if (i < COP) bias[i] = i < CO ? bm[i] : 0.f;
    }
}

// loss: g[T][16] bf16 (channel 0 = 2 (out - target) / T), lpart[block] = sum diff^2
__global__ void __launch_bounds__(256) k_loss(long long T, const float *__restrict__ out, const float *__restrict__ tgt,
                                              bf16 *__restrict__ g, double *__restrict__ lpart)
{
    __shared__ double red[256];
    double s = 0.0;
    const float scale = 2.0f / (float)T;
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < T; v += (long long)gridDim.x * blockDim.x) {
        const float d = out[v] - tgt[v];
        s += (double)d * d;
        g[v * 16] = (bf16)(scale * d);
#pragma unroll
        for (int c = 1; c < 16; ++c) g[v * 16 + c] = (bf16)0.f;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
}

__global__ void k_adam(long n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// ---- source-iteration correction -------------------------------------------
// ravg[e] = sum_i w_i r[e n + i] ; part[block] = sum ravg^2
__global__ void __launch_bounds__(256) k_elem_avg(long long Ne, int n, const double *__restrict__ w,
                                                  const double *__restrict__ r, double *__restrict__ ravg,
                                                  double *__restrict__ part)
{
    __shared__ double red[256];
    double s = 0.0;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < Ne; e += (long long)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int i = 0; i < n; ++i) v = fma(w[i], r[e * n + i], v);
        ravg[e] = v;
        s += v * v;
    }
    if (!part) return;   // block-uniform
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// scale = sqrt(mean ravg^2) (one block)
__global__ void k_scale(const double *__restrict__ part, int nb, long long Ne, double *__restrict__ scale)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) scale[0] = sqrt(red[0] / (double)Ne);
}

// network input [Ne][8] bf16: (ravg / scale, log sigma_t, sigma_s / sigma_t)
__global__ void k_make_input(long long Ne, const double *__restrict__ ravg, const double *__restrict__ scale,
                             const double *__restrict__ sigt, const double *__restrict__ sigs, bf16 *__restrict__ xb)
{
    const double s = scale[0];
    const double inv = s > 0.0 ? 1.0 / s : 0.0;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < Ne; e += (long long)gridDim.x * blockDim.x) {
        const double st = sigt[e];
        xb[e * 8 + 0] = (bf16)(float)(ravg[e] * inv);
        xb[e * 8 + 1] = (bf16)(float)log(st);
        xb[e * 8 + 2] = (bf16)(float)(sigs[e] / st);
#pragma unroll
        for (int c = 3; c < 8; ++c) xb[e * 8 + c] = (bf16)0.f;
    }
}

// phi[e n + i] += scale * out[e]
__global__ void k_apply_corr(long long Ne, int n, const float *__restrict__ out, const double *__restrict__ scale,
                             double *__restrict__ phi)
{
    const double s = scale[0];
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < Ne * n; i += (long long)gridDim.x * blockDim.x)
        phi[i] += s * (double)out[i / n];
}

// p = 1 (nodes at the vertices): phi[e, i] += scale * (mean of out over the
// elements that share node i's vertex) -- a continuous trilinear correction
__global__ void k_apply_corr_vertex(int nz, int ny, int nx, const float *__restrict__ out,
                                    const double *__restrict__ scale, double *__restrict__ phi)
{
    const double s = scale[0];
    const long long Ne = (long long)nz * ny * nx;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < Ne * 8; i += (long long)gridDim.x * blockDim.x) {
        const long long e = i >> 3;
        const int loc = (int)(i & 7);
        const int x = (int)(e % nx), y = (int)((e / nx) % ny), z = (int)(e / ((long long)nx * ny));
        const int vx = x + (loc & 1), vy = y + ((loc >> 1) & 1), vz = z + (loc >> 2);
        double sum = 0.0;
        int cnt = 0;
        for (int jz = vz - 1; jz <= vz; ++jz)
            for (int jy = vy - 1; jy <= vy; ++jy)
                for (int jx = vx - 1; jx <= vx; ++jx)
                    if (jx >= 0 && jx < nx && jy >= 0 && jy < ny && jz >= 0 && jz < nz) {
                        sum += (double)out[((long long)jz * ny + jy) * nx + jx];
                        ++cnt;
                    }
        phi[i] += s * sum / (double)cnt;
    }
}

inline int grid_for(long long n, int block = 256, int cap = 4096)
{
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    return (int)(g > cap ? cap : g);
}

thread_local std::string g_cnn_err = "no error";
int cfail(int code, const std::string &m)
{
    g_cnn_err = m;
    allium_cnn_internal::set_last_error(m);
    return code;
}
#define CNN_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess) return cfail(-4, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

long layer_offset(int l)
{
    long off = 0;
    for (int k = 0; k < l; ++k) off += 27L * CO_[k] * CI_[k] + CO_[k];
    return off;
}
const long NPARAM = layer_offset(NL);

}  // namespace

struct allium_cnn {
    float *params = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    bf16 *Wf[NL] = {}, *Wb[NL] = {};
    float *bias[NL] = {};
    bool packed = false;
    long long cap = 0;
    int cap_planes = 0;
    bf16 *xin = nullptr, *act[NL - 1] = {}, *ga = nullptr, *gb = nullptr;
    float *out = nullptr, *part = nullptr, *bpart = nullptr, *xf = nullptr;
    double *lpart = nullptr, *ravg = nullptr, *rpart = nullptr, *scale = nullptr;
    long step = 0;
    int dil[NL] = {1, 1, 1, 1, 1};
    hipStream_t stream = nullptr;
    std::vector<void *> owned;
    ~allium_cnn()
    {
        for (void *p : owned) hipFree(p);
        if (stream) hipStreamDestroy(stream);
    }
    template <typename T>
    hipError_t alloc(T **p, size_t count)
    {
        hipError_t e = hipMalloc((void **)p, sizeof(T) * (count ? count : 1));
        if (e == hipSuccess) owned.push_back(*p);
        return e;
    }
};

namespace {

constexpr int kBiasBlocks = 256;
constexpr int kLossBlocks = 512;

int ensure(allium_cnn *h, const Geo &g)
{
    const int planes = g.B * g.nz;
    if (g.T <= h->cap && planes <= h->cap_planes) return 0;
    // grow: free the old work buffers by re-creating the list
    std::vector<void *> keep;
    for (void *p : h->owned) {
        bool work = p == h->xin || p == h->ga || p == h->gb || p == h->out || p == h->part || p == h->xf ||
                    p == h->ravg;
        for (int l = 0; l < NL - 1; ++l) work = work || p == h->act[l];
        if (work) hipFree(p);
        else keep.push_back(p);
    }
    h->owned = keep;
    const long long T = g.T;
    CNN_TRY(h->alloc(&h->xin, (size_t)T * 8));
    for (int l = 0; l < NL - 1; ++l) CNN_TRY(h->alloc(&h->act[l], (size_t)T * HID));
    CNN_TRY(h->alloc(&h->ga, (size_t)T * HID));
    CNN_TRY(h->alloc(&h->gb, (size_t)T * HID));
    CNN_TRY(h->alloc(&h->out, (size_t)T));
    CNN_TRY(h->alloc(&h->xf, (size_t)T * 3));
    CNN_TRY(h->alloc(&h->ravg, (size_t)T));
    CNN_TRY(h->alloc(&h->part, (size_t)planes * 27 * 32 * 32));
    h->cap = T;
    h->cap_planes = planes;
    return 0;
}

int pack_weights(allium_cnn *h, hipStream_t s)
{
    if (h->packed) return 0;
    for (int l = 0; l < NL; ++l) {
        const float *Wm = h->params + layer_offset(l);
        const float *bm = Wm + 27L * CO_[l] * CI_[l];
        hipLaunchKernelGGL(k_pack_layer, dim3(grid_for(27L * COP_[l] * CIP_[l])), dim3(256), 0, s, Wm, bm, CO_[l],
                           CI_[l], COP_[l], CIP_[l], h->Wf[l], h->Wb[l], h->bias[l]);
    }
    CNN_TRY(hipGetLastError());
    h->packed = true;
    return 0;
}

// forward from the packed bf16 input h->xin; output FP32 h->out
Geo layer_geo(const allium_cnn *h, const Geo &g, int l)
{
    Geo gl = g;
    gl.dil = h->dil[l];
    return gl;
}

int forward_packed(allium_cnn *h, const Geo &g, hipStream_t s)
{
    launch_conv<8, 32, 0>(layer_geo(h, g, 0), h->xin, h->Wf[0], h->bias[0], h->act[0], nullptr, nullptr, s);
    launch_conv<32, 32, 0>(layer_geo(h, g, 1), h->act[0], h->Wf[1], h->bias[1], h->act[1], nullptr, nullptr, s);
    launch_conv<32, 32, 0>(layer_geo(h, g, 2), h->act[1], h->Wf[2], h->bias[2], h->act[2], nullptr, nullptr, s);
    launch_conv<32, 32, 0>(layer_geo(h, g, 3), h->act[2], h->Wf[3], h->bias[3], h->act[3], nullptr, nullptr, s);
    launch_conv<32, 16, 1>(layer_geo(h, g, 4), h->act[3], h->Wf[4], h->bias[4], nullptr, h->out, nullptr, s);
    CNN_TRY(hipGetLastError());
    return 0;
}

int check_geo(int B, int nz, int ny, int nx, Geo &g)
{
    if (B < 1 || nz < 1 || ny < 1 || nx < 1 || nx > (1 << 20)) return cfail(-1, "bad grid size");
    g.B = B;
    g.nz = nz;
    g.ny = ny;
    g.nx = nx;
    g.T = (long long)B * nz * ny * nx;
    g.dil = 1;
    return 0;
}

// forward + backward into h->grad (device, packed); *loss on the host
int gradient(allium_cnn *h, const Geo &g, const float *d_x, const float *d_t, double *loss, hipStream_t s)
{
    int rc = ensure(h, g);
    if (rc) return rc;
    rc = pack_weights(h, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pack_input, dim3(grid_for(g.T * 8)), dim3(256), 0, s, g.T, d_x, h->xin);
    rc = forward_packed(h, g, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_loss, dim3(kLossBlocks), dim3(256), 0, s, g.T, h->out, d_t, h->ga, h->lpart);
    const int planes = g.B * g.nz;
    // layer 5: gz = ga [T][16]
    const bf16 *gz = h->ga;
    bf16 *gnext = h->gb;
    for (int l = NL - 1; l >= 0; --l) {
        const bf16 *ain = l == 0 ? h->xin : h->act[l - 1];
        float *gW = h->grad + layer_offset(l);
        float *gB = gW + 27L * CO_[l] * CI_[l];
        const int CIT = ((CIP_[l] + 15) / 16) * 16;
        const Geo gl = layer_geo(h, g, l);
        if (l == NL - 1) {
            hipLaunchKernelGGL((k_wgrad<32, 16>), dim3(planes), dim3(256), 0, s, gl, gz, ain, h->part);
            hipLaunchKernelGGL(k_bias_grad<16>, dim3(kBiasBlocks), dim3(256), 0, s, g.T, gz, h->bpart);
        } else if (l == 0) {
            hipLaunchKernelGGL((k_wgrad<8, 32>), dim3(planes), dim3(256), 0, s, gl, gz, ain, h->part);
            hipLaunchKernelGGL(k_bias_grad<32>, dim3(kBiasBlocks), dim3(256), 0, s, g.T, gz, h->bpart);
        } else {
            hipLaunchKernelGGL((k_wgrad<32, 32>), dim3(planes), dim3(256), 0, s, gl, gz, ain, h->part);
            hipLaunchKernelGGL(k_bias_grad<32>, dim3(kBiasBlocks), dim3(256), 0, s, g.T, gz, h->bpart);
        }
        hipLaunchKernelGGL(k_wgrad_reduce, dim3((unsigned)((27L * CO_[l] * CI_[l] + 63) / 64)), dim3(256), 0, s,
                           h->part, planes, COP_[l], CIT, CO_[l], CI_[l], gW);
        hipLaunchKernelGGL(k_bias_reduce, dim3(1), dim3(256), 0, s, h->bpart, kBiasBlocks, COP_[l], CO_[l], gB);
        if (l > 0) {
            // gradient of the previous pre-activation: conv with flipped weights, ReLU mask of its output
            if (l == NL - 1)
                launch_conv<16, 32, 2>(gl, gz, h->Wb[l], nullptr, gnext, nullptr, h->act[l - 1], s);
            else
                launch_conv<32, 32, 2>(gl, gz, h->Wb[l], nullptr, gnext, nullptr, h->act[l - 1], s);
            const bf16 *t = gnext;
            gnext = const_cast<bf16 *>(gz);
            gz = t;
        }
    }
    CNN_TRY(hipGetLastError());
    std::vector<double> lp(kLossBlocks);
    CNN_TRY(hipMemcpyAsync(lp.data(), h->lpart, sizeof(double) * kLossBlocks, hipMemcpyDeviceToHost, s));
    CNN_TRY(hipStreamSynchronize(s));
    double sum = 0.0;
    for (double v : lp) sum += v;
    if (loss) *loss = sum / (double)g.T;
    return 0;
}

}  // namespace

namespace allium_cnn_internal {

hipError_t elem_avg(long long Ne, int n, const double *d_w, const double *d_x, double *d_out, hipStream_t s)
{
    hipLaunchKernelGGL(k_elem_avg, dim3(grid_for(Ne)), dim3(256), 0, s, Ne, n, d_w, d_x, d_out, nullptr);
    return hipGetLastError();
}

int correct(allium_cnn *h, int nz, int ny, int nx, int n, const double *d_w, const double *d_r,
            const double *d_sigt, const double *d_sigs, double *d_phi, hipStream_t s, int prolong)
{
    Geo g;
    int rc = check_geo(1, nz, ny, nx, g);
    if (rc) return rc;
    rc = ensure(h, g);
    if (rc) return rc;
    rc = pack_weights(h, s);
    if (rc) return rc;
    const int nb = 256;
    hipLaunchKernelGGL(k_elem_avg, dim3(nb), dim3(256), 0, s, g.T, n, d_w, d_r, h->ravg, h->rpart);
    hipLaunchKernelGGL(k_scale, dim3(1), dim3(256), 0, s, h->rpart, nb, g.T, h->scale);
    hipLaunchKernelGGL(k_make_input, dim3(grid_for(g.T)), dim3(256), 0, s, g.T, h->ravg, h->scale, d_sigt, d_sigs,
                       h->xin);
    rc = forward_packed(h, g, s);
    if (rc) return rc;
    if (prolong == 1 && n == 8)
        hipLaunchKernelGGL(k_apply_corr_vertex, dim3(grid_for(g.T * 8)), dim3(256), 0, s, nz, ny, nx, h->out, h->scale,
                           d_phi);
    else
        hipLaunchKernelGGL(k_apply_corr, dim3(grid_for(g.T * n)), dim3(256), 0, s, g.T, n, h->out, h->scale, d_phi);
    CNN_TRY(hipGetLastError());
    return 0;
}

}  // namespace allium_cnn_internal

extern "C" {

long allium_cnn_num_params(void) { return NPARAM; }

int allium_cnn_set_dilation(allium_cnn_t h, const int *dilation)
{
    if (!h || !dilation) return cfail(-1, "null argument");
    for (int l = 0; l < NL; ++l)
        if (dilation[l] < 1 || dilation[l] > kMaxDil) return cfail(-1, "dilation must be 1..8 per layer");
    for (int l = 0; l < NL; ++l) h->dil[l] = dilation[l];
    return 0;
}

int allium_cnn_get_dilation(allium_cnn_t h, int *dilation)
{
    if (!h || !dilation) return cfail(-1, "null argument");
    for (int l = 0; l < NL; ++l) dilation[l] = h->dil[l];
    return 0;
}

int allium_cnn_create(unsigned long long seed, allium_cnn_t *out)
{
    if (!out) return cfail(-1, "null argument");
    *out = nullptr;
    auto *h = new allium_cnn();
    auto bad = [&](hipError_t e) {
        delete h;
        return cfail(-4, std::string("allium_cnn_create: ") + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = h->alloc(&h->params, NPARAM)) != hipSuccess) return bad(e);
    if ((e = h->alloc(&h->m, NPARAM)) != hipSuccess) return bad(e);
    if ((e = h->alloc(&h->v, NPARAM)) != hipSuccess) return bad(e);
    if ((e = h->alloc(&h->grad, NPARAM)) != hipSuccess) return bad(e);
    for (int l = 0; l < NL; ++l) {
        if ((e = h->alloc(&h->Wf[l], 27L * COP_[l] * CIP_[l])) != hipSuccess) return bad(e);
        if ((e = h->alloc(&h->Wb[l], 27L * COP_[l] * CIP_[l])) != hipSuccess) return bad(e);
        if ((e = h->alloc(&h->bias[l], COP_[l])) != hipSuccess) return bad(e);
    }
    if ((e = h->alloc(&h->bpart, (size_t)kBiasBlocks * 32)) != hipSuccess) return bad(e);
    if ((e = h->alloc(&h->lpart, kLossBlocks)) != hipSuccess) return bad(e);
    if ((e = h->alloc(&h->rpart, 256)) != hipSuccess) return bad(e);
    if ((e = h->alloc(&h->scale, 1)) != hipSuccess) return bad(e);
    if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess) return bad(e);
    // He-normal weights, zero biases
    std::vector<float> p(NPARAM, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < NL; ++l) {
        std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / (27.0 * CI_[l])));
        float *W = p.data() + layer_offset(l);
        for (long i = 0; i < 27L * CO_[l] * CI_[l]; ++i) W[i] = (float)nd(rng);
    }
    if ((e = hipMemcpy(h->params, p.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice)) != hipSuccess) return bad(e);
    if ((e = hipMemset(h->m, 0, sizeof(float) * NPARAM)) != hipSuccess) return bad(e);
    if ((e = hipMemset(h->v, 0, sizeof(float) * NPARAM)) != hipSuccess) return bad(e);
    *out = h;
    return 0;
}

void allium_cnn_destroy(allium_cnn_t h) { delete h; }

int allium_cnn_get_params(allium_cnn_t h, float *params)
{
    if (!h || !params) return cfail(-1, "null argument");
    CNN_TRY(hipMemcpy(params, h->params, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
    return 0;
}

int allium_cnn_set_params(allium_cnn_t h, const float *params)
{
    if (!h || !params) return cfail(-1, "null argument");
    for (long i = 0; i < NPARAM; ++i)
        if (!std::isfinite(params[i])) return cfail(-6, "non-finite network parameter");
    CNN_TRY(hipMemcpy(h->params, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    CNN_TRY(hipMemset(h->m, 0, sizeof(float) * NPARAM));
    CNN_TRY(hipMemset(h->v, 0, sizeof(float) * NPARAM));
    h->step = 0;
    h->packed = false;
    return 0;
}

int allium_cnn_forward_device(allium_cnn_t h, int batch, int nz, int ny, int nx, const float *d_x, float *d_out,
                              void *stream)
{
    if (!h || !d_x || !d_out) return cfail(-1, "null argument");
    Geo g;
    int rc = check_geo(batch, nz, ny, nx, g);
    if (rc) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    rc = ensure(h, g);
    if (rc) return rc;
    rc = pack_weights(h, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pack_input, dim3(grid_for(g.T * 8)), dim3(256), 0, s, g.T, d_x, h->xin);
    rc = forward_packed(h, g, s);
    if (rc) return rc;
    CNN_TRY(hipMemcpyAsync(d_out, h->out, sizeof(float) * g.T, hipMemcpyDeviceToDevice, s));
    CNN_TRY(hipStreamSynchronize(s));
    return 0;
}

int allium_cnn_gradient_device(allium_cnn_t h, int batch, int nz, int ny, int nx, const float *d_x,
                               const float *d_target, float *grad, double *loss)
{
    if (!h || !d_x || !d_target || !grad) return cfail(-1, "null argument");
    Geo g;
    int rc = check_geo(batch, nz, ny, nx, g);
    if (rc) return rc;
    rc = gradient(h, g, d_x, d_target, loss, h->stream);
    if (rc) return rc;
    CNN_TRY(hipMemcpy(grad, h->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
    return 0;
}

int allium_cnn_train_step_device(allium_cnn_t h, int batch, int nz, int ny, int nx, const float *d_x,
                                 const float *d_target, float lr, double *loss)
{
    if (!h || !d_x || !d_target) return cfail(-1, "null argument");
    if (!(lr > 0.f) || !std::isfinite(lr)) return cfail(-1, "learning rate must be positive");
    Geo g;
    int rc = check_geo(batch, nz, ny, nx, g);
    if (rc) return rc;
    double l = 0.0;
    rc = gradient(h, g, d_x, d_target, &l, h->stream);
    if (rc) return rc;
    if (!std::isfinite(l)) return cfail(-6, "non-finite training loss");
    h->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)h->step), c2 = 1.f - std::pow(b2, (float)h->step);
    hipLaunchKernelGGL(k_adam, dim3(grid_for(NPARAM)), dim3(256), 0, h->stream, NPARAM, h->params, h->grad, h->m,
                       h->v, lr, b1, b2, 1e-8f, c1, c2);
    CNN_TRY(hipGetLastError());
    h->packed = false;
    CNN_TRY(hipStreamSynchronize(h->stream));
    if (loss) *loss = l;
    return 0;
}

}  // extern "C"
