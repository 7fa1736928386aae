// C ABI of the MI355X sweep (include/allium_sweep.h): handle construction,
// device upload, the Mult entry points, the streaming operator, moments,
// balance and the source iteration driver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/allium_device.h"
#include "../../include/allium_sweep.h"
#include "cnn_internal.hpp"
#include "dsa_internal.hpp"
#include "device.hpp"
#include "discretization.hpp"
#include "sweep_handle.hpp"

namespace {
thread_local std::string g_last_error = "no error";

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return fail(ALLIUM_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

template <typename T>
hipError_t upload(std::vector<void *> &owned, T **dst, const T *src, size_t count)
{
    *dst = nullptr;
    if (count == 0) return hipSuccess;
    hipError_t e = hipMalloc((void **)dst, sizeof(T) * count);
    if (e != hipSuccess) return e;
    owned.push_back(*dst);
    if (src) return hipMemcpy(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice);
    return hipSuccess;
}

}  // namespace

void allium_cnn_internal::set_last_error(const std::string &msg) { g_last_error = msg; }


namespace {

// Uniform geometry: K_a = sum_d Omega_d G_d + sum_{sides with Omega.n > 0}
// (Omega.n) F_side [+ tau Mt], the same for every element
void build_streaming(allium_sweep *h)
{
    const auto &ed = h->ed;
    const int n = ed.n, nf = ed.nf, dim = ed.dim, No = h->quad.size();
    const double tau = h->ds.prm.timedep ? h->ds.prm.tau : 0.0;
    h->Kst.assign((size_t)No * n * n, 0.0);
    for (int a = 0; a < No; a++) {
        double *K = &h->Kst[(size_t)a * n * n];
        if (tau != 0.0)
            for (int i = 0; i < n * n; i++) K[i] += tau * ed.Mm[i];
        for (int d = 0; d < dim; d++) {
            const double od = h->quad.omega[3 * a + d];
            for (int i = 0; i < n * n; i++) K[i] += od * ed.Gd[(size_t)d * n * n + i];
        }
        for (int s = 0; s < 2 * dim; s++) {
            const double dot = (s & 1) ? h->quad.omega[3 * a + (s >> 1)] : -h->quad.omega[3 * a + (s >> 1)];
            if (dot > 0)
                for (int k = 0; k < nf; k++)
                    for (int l = 0; l < nf; l++)
                        K[ed.side_nodes[s * nf + k] * n + ed.side_nodes[s * nf + l]] +=
                            dot * ed.Fface[(size_t)s * nf * nf + k * nf + l];
        }
    }
}

void build_tiles(allium_sweep *h, std::vector<allium::Tile> &tiles, std::vector<int> &need, int &nkeys)
{
    const auto &plan = h->plan;
    const int No = h->quad.size(), G = h->G, T = h->ds.persistent_tile;
    tiles.clear();
    need.clear();
    if (plan.angle_independent) {
        nkeys = plan.nlevels * No;
        need.assign(nkeys, 0);
        long i = 0;
        const long ncell = (long)plan.cells.size();
        while (i < ncell) {
            const int a = (int)(plan.cells[i] & 1023u);
            const int e = (int)(plan.cells[i] >> 10);
            const int L = plan.level[(size_t)e * No + a];
            long j = i;
            while (j < ncell && (int)(plan.cells[j] & 1023u) == a &&
                   plan.level[(size_t)(plan.cells[j] >> 10) * No + a] == L)
                j++;
            const int key = L * No + a;
            const int nc = (int)(j - i);
            need[key] = nc * G;
            for (int it = 0; it < nc * G; it += T)
                tiles.push_back({key, L > 0 ? key - No : -1, (int)i, nc, it, std::min(T, nc * G - it)});
            i = j;
        }
    } else {
        nkeys = plan.nlevels;
        need.assign(nkeys, 0);
        for (int L = 0; L < plan.nlevels; L++) {
            const long b = plan.level_ptr[L];
            const int nc = (int)(plan.level_ptr[L + 1] - b);
            need[L] = nc * G;
            for (int it = 0; it < nc * G; it += T)
                tiles.push_back({L, L > 0 ? L - 1 : -1, (int)b, nc, it, std::min(T, nc * G - it)});
        }
    }
}

int upload_all(allium_sweep *h)
{
    auto &ds = h->ds;
    auto &prm = ds.prm;
    const auto &m = h->mesh;
    const auto &ed = h->ed;
    const int No = h->quad.size(), n = ed.n, dim = m.dim;
    std::memset(&prm, 0, sizeof(prm));
    prm.dim = dim; prm.p = ed.p; prm.n = n; prm.nf = ed.nf;
    prm.Ne = m.Ne; prm.Nomega = No; prm.G = h->G;
    for (int d = 0; d < 3; d++) prm.nel[d] = m.nel[d];
    prm.stride[0] = 1; prm.stride[1] = m.nel[0]; prm.stride[2] = m.nel[0] * m.nel[1];
    prm.inv_nel0 = 1.0f / (float)m.nel[0];
    prm.inv_nel01 = 1.0f / (float)(m.nel[0] * m.nel[1]);
    prm.Ndof = (long long)m.Ne * n;
    prm.uniform = m.uniform ? 1 : 0;
    for (int i = 0; i < 4; i++) prm.foff[i] = m.foff[i];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) prm.fext[i][j] = m.fext[i][j];
    std::vector<double> om4((size_t)No * 4);
    for (int a = 0; a < No; a++) {
        for (int d = 0; d < 3; d++) om4[4 * a + d] = h->quad.omega[3 * a + d];
        om4[4 * a + 3] = h->quad.w[a];
    }
    HIP_TRY(upload(ds.owned, &ds.d_omega, om4.data(), om4.size()));
    prm.omega = ds.d_omega;
    auto tofloat = [](const std::vector<double> &v) {
        std::vector<float> f(v.size());
        for (size_t i = 0; i < v.size(); i++) f[i] = (float)v[i];
        return f;
    };
    double *dK = nullptr, *dMu = nullptr, *dMm = nullptr, *dGd = nullptr, *dF = nullptr, *dN = nullptr;
    float *fK = nullptr, *fMu = nullptr, *fMm = nullptr, *fF = nullptr;
    if (m.uniform) {
        build_streaming(h);
        HIP_TRY(upload(ds.owned, &dK, h->Kst.data(), h->Kst.size()));
        auto kf = tofloat(h->Kst);
        HIP_TRY(upload(ds.owned, &fK, kf.data(), kf.size()));
    } else {
        HIP_TRY(upload(ds.owned, &dGd, ed.Gd.data(), ed.Gd.size()));
        HIP_TRY(upload(ds.owned, &dN, m.f_nor.data(), m.f_nor.size()));
    }
    HIP_TRY(upload(ds.owned, &dMu, ed.Mu.data(), ed.Mu.size()));
    HIP_TRY(upload(ds.owned, &dMm, ed.Mm.data(), ed.Mm.size()));
    HIP_TRY(upload(ds.owned, &dF, ed.Fface.data(), ed.Fface.size()));
    {
        auto a = tofloat(ed.Mu), b = tofloat(ed.Mm), c = tofloat(ed.Fface);
        HIP_TRY(upload(ds.owned, &fMu, a.data(), a.size()));
        HIP_TRY(upload(ds.owned, &fMm, b.data(), b.size()));
        HIP_TRY(upload(ds.owned, &fF, c.data(), c.size()));
    }
    prm.K = dK; prm.Kf = fK; prm.Mu = dMu; prm.Muf = fMu; prm.Mt = dMm; prm.Mtf = fMm;
    prm.Gd = dGd; prm.Fface = dF; prm.Ffacef = fF; prm.fnor = dN;
    HIP_TRY(upload(ds.owned, &h->d_sigma, h->sigma.data(), h->sigma.size()));
    prm.sigma = h->d_sigma;
    {
        const std::vector<double> z(64, 0.0);
        double *dz = nullptr;
        HIP_TRY(upload(ds.owned, &dz, z.data(), z.size()));
        prm.zeros = dz;
    }
    if (allium::coop_supported(dim, ed.p)) {
        // cooperative solves read their element's matrices row by row: order
        // each level by element, then direction, so that the cells one
        // workgroup solves mostly share an element (cache reuse)
        std::vector<uint32_t> cells = h->plan.cells;
        for (int L = 0; L < h->plan.nlevels; L++)
            std::sort(cells.begin() + h->plan.level_ptr[L], cells.begin() + h->plan.level_ptr[L + 1]);
        HIP_TRY(upload(ds.owned, &ds.d_cells, cells.data(), cells.size()));
    } else {
        HIP_TRY(upload(ds.owned, &ds.d_cells, h->plan.cells.data(), h->plan.cells.size()));
    }
    prm.cells = ds.d_cells;
    prm.refl_mask = 0;
    for (int s = 0; s < 2 * dim; s++)
        if (m.bc[s] == 1) prm.refl_mask |= 1 << s;
    if (prm.refl_mask) {
        HIP_TRY(upload(ds.owned, &ds.d_refl, h->plan.refl.data(), h->plan.refl.size()));
        prm.refl = ds.d_refl;
    }
    if (m.uniform) {
        int nmax = std::max(m.nel[0], std::max(m.nel[1], m.nel[2]));
        allium::build_wave_plan(m, h->quad, h->plan, std::max(2, (nmax + 7) / 8), 1);
        HIP_TRY(upload(ds.owned, &ds.d_wcells, h->plan.wcells.data(), h->plan.wcells.size()));
        ds.wlevel_base = h->plan.wlevel_base;
        ds.wpb = h->plan.wpb;
        ds.nparts = 1;
        std::vector<uint32_t>().swap(h->plan.wcells);
    }
    {
        // start of every (level, direction) class in the cells array
        const int nl = h->plan.nlevels;
        std::vector<long long> cls((size_t)nl * No + 1, 0);
        std::vector<long long> cnt((size_t)nl * No, 0);
        for (uint32_t c : h->plan.cells) {
            const int e = (int)(c >> 10), a = (int)(c & 1023u);
            cnt[(size_t)h->plan.level[(size_t)e * No + a] * No + a]++;
        }
        for (size_t k = 0; k < cnt.size(); k++) cls[k + 1] = cls[k] + cnt[k];
        HIP_TRY(upload(ds.owned, &ds.d_cls_ptr, cls.data(), cls.size()));
        ds.angle_independent = h->plan.angle_independent;
    }
    ds.level_ptr = h->plan.level_ptr;
    ds.nlevels = h->plan.nlevels;
    std::vector<allium::Tile> tiles;
    std::vector<int> need;
    build_tiles(h, tiles, need, ds.nkeys);
    ds.ntiles = (int)tiles.size();
    HIP_TRY(upload(ds.owned, &ds.d_tiles, tiles.data(), tiles.size()));
    HIP_TRY(upload(ds.owned, &ds.d_need, need.data(), need.size()));
    HIP_TRY(upload(ds.owned, &ds.d_counters, (const int *)nullptr, (size_t)ds.nkeys + 2));
    HIP_TRY(hipMemset(ds.d_counters, 0, sizeof(int) * (ds.nkeys + 2)));
    HIP_TRY(hipEventCreate(&ds.ev0));
    HIP_TRY(hipEventCreate(&ds.ev1));
    HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    return ALLIUM_OK;
}

// Box-sweep plan (launch modes 4/5): oriented box order, the level order of the
// cells inside a full box and the directions of each octant.
int build_box_plan(allium_sweep *h, int bx, int by, int bz, int nd, int threads, int prefetch)
{
    auto &ds = h->ds;
    const int dim = h->mesh.dim, n = h->ed.n, nf = h->ed.nf;
    if (!h->mesh.uniform) return fail(ALLIUM_ERR_UNSUPPORTED, "box sweeps need uniform geometry");
    if (!ds.angle_independent)
        return fail(ALLIUM_ERR_UNSUPPORTED, "box sweeps need directions that do not couple (no reflective side)");
    if (!allium::box_sweep_supported(dim, h->ed.p))
        return fail(ALLIUM_ERR_UNSUPPORTED, "box sweeps implement 2D p=0..2 and 3D p=0..1");
    // the box kernels build the cell operator from the exact 1D masses and derivatives
    if (h->ed.lump)
        return fail(ALLIUM_ERR_UNSUPPORTED, "box sweeps take the unlumped element matrices (lump = 0)");
    int bs[3] = {bx, by, dim == 3 ? bz : 1};
    for (int d = 0; d < dim; d++) {
        if (bs[d] < 1 || bs[d] > 1023) return fail(ALLIUM_ERR_ARG, "box extents must be 1..1023");
        bs[d] = std::min(bs[d], h->mesh.nel[d]);
    }
    if (nd < 1 || nd > 64) return fail(ALLIUM_ERR_ARG, "directions per workgroup must be 1..64");
    // p = 1 cells are split over 2^(dim-1) lanes and take workgroups up to 1024
    const int max_threads = h->ed.p == 1 ? 1024 : 256;
    if (threads < 64 || threads > max_threads || threads % 64)
        return fail(ALLIUM_ERR_ARG, "box sweep workgroup size must be a multiple of 64 in 64.." +
                                        std::to_string(max_threads));
    if (prefetch != 0 && prefetch != 2 && !(prefetch == 3 && h->ed.p == 1))
        return fail(ALLIUM_ERR_ARG, "box sweep prefetch must be 0 or 2 (or 3 at p = 1)");
    const int No = h->quad.size(), noct = 1 << dim;
    std::vector<std::vector<int>> oct(noct);
    for (int a = 0; a < No; a++) {
        int o = 0;
        for (int d = 0; d < dim; d++)
            if (h->quad.omega[3 * a + d] < 0) o |= 1 << d;
        oct[o].push_back(a);
    }
    int maxoct = 1;
    for (auto &v : oct) maxoct = std::max(maxoct, (int)v.size());
    nd = std::min(nd, maxoct);
    allium::BoxPlan bp{};
    for (int d = 0; d < 3; d++) {
        bp.bs[d] = bs[d];
        bp.nb[d] = d < dim ? (h->mesh.nel[d] + bs[d] - 1) / bs[d] : 1;
    }
    bp.nd = nd;
    bp.nchunk = (maxoct + nd - 1) / nd;
    bp.maxoct = bp.nchunk * nd;
    bp.noct = noct;
    bp.nl = bs[0] + bs[1] + bs[2] - 2;
    // odd stride: the directions' matrices start in different LDS banks; p = 1
    // cells are solved by fast diagonalisation and need no matrix in LDS
    bp.kstride = h->ed.p == 1 ? 0 : n * n + 1;
    long plane = 0;
    for (int ax = 0; ax < 3; ax++) {
        int v = 1;
        for (int o = 0; o < dim; o++)
            if (o != ax) v *= bs[o];
        bp.psz[ax] = ax < dim ? v : 0;
        plane += bp.psz[ax];
    }
    const long vol = (long)bs[0] * bs[1] * bs[2];
    // p = 1 adds per-direction psi offsets and a per-cell element table
    const long split = h->ed.p == 1 ? 1 : 0;
    const long lds = 8L * ((long)nd * allium::box_slot_stride(nf) * plane + (long)nd * bp.kstride + 3L * nd) +
                     4L * (nd + vol + bp.nl + 1) + split * (8L * nd + 4L * vol);
    // the box kernel builds Mu and the face masses as tensor products of the
    // exact 1D masses: check that the assembled matrices are those
    {
        const auto &m = h->mesh;
        bp.vol = 1.0;
        for (int d = 0; d < dim; d++) bp.vol *= m.h[d];
        for (int ax = 0; ax < 3; ax++) {
            bp.farea[ax] = 1.0;
            for (int d = 0; d < dim; d++)
                if (d != ax) bp.farea[ax] *= m.h[d];
        }
        double err = 0.0, ref = 0.0, errf = 0.0, reff = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                const double c = bp.vol * allium::host_unit_mass(dim, h->ed.p, i, j);
                err = std::max(err, std::fabs(h->ed.Mu[(size_t)i * n + j] - c));
                ref = std::max(ref, std::fabs(c));
            }
        for (int sd = 0; sd < 2 * dim; sd++)
            for (int k = 0; k < nf; k++)
                for (int l = 0; l < nf; l++) {
                    const double c = bp.farea[sd >> 1] * allium::host_unit_mass(dim - 1, h->ed.p, k, l);
                    errf = std::max(errf, std::fabs(h->ed.Fface[((size_t)sd * nf + k) * nf + l] - c));
                    reff = std::max(reff, std::fabs(c));
                }
        // p = 1 also takes the time mass (sigma + tau shift) as the same tensor product
        if (h->ed.p == 1)
            for (int i = 0; i < n; i++)
                for (int j = 0; j < n; j++)
                    err = std::max(err, std::fabs(h->ed.Mm[(size_t)i * n + j] -
                                                  bp.vol * allium::host_unit_mass(dim, 1, i, j)));
        if (!(err <= 1e-13 * ref) || !(errf <= 1e-13 * reff))
            return fail(ALLIUM_ERR_UNSUPPORTED, "box sweeps: element matrices are not the tensor-product masses "
                                                "of an axis-aligned box");
    }
    int dev = 0, maxlds = 65536, ncu = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (lds > std::min(maxlds, 65536))
        return fail(ALLIUM_ERR_ARG, "box tiling needs " + std::to_string(lds) +
                                        " bytes of LDS per workgroup (limit 65536): use smaller boxes or fewer "
                                        "directions per workgroup");
    bp.lds_bytes = (int)lds;
    std::vector<uint32_t> lcells, boxes;
    std::vector<int> lptr(1, 0), bptr(1, 0), odirs((size_t)noct * bp.maxoct, -1);
    for (int l = 0; l < bp.nl; l++) {
        for (int k = 0; k < bs[2]; k++)
            for (int j = 0; j < bs[1]; j++) {
                const int i = l - j - k;
                if (i >= 0 && i < bs[0]) lcells.push_back((uint32_t)i | (uint32_t)j << 10 | (uint32_t)k << 20);
            }
        lptr.push_back((int)lcells.size());
    }
    const int nbl = bp.nb[0] + bp.nb[1] + bp.nb[2] - 2;
    for (int l = 0; l < nbl; l++) {
        for (int k = 0; k < bp.nb[2]; k++)
            for (int j = 0; j < bp.nb[1]; j++) {
                const int i = l - j - k;
                if (i >= 0 && i < bp.nb[0]) boxes.push_back((uint32_t)i | (uint32_t)j << 10 | (uint32_t)k << 20);
            }
        bptr.push_back((int)boxes.size());
    }
    for (int o = 0; o < noct; o++)
        for (size_t i = 0; i < oct[o].size(); i++) odirs[(size_t)o * bp.maxoct + i] = oct[o][i];
    // one device buffer: oct_dirs | lptr | lcells | boxes (all 32-bit)
    std::vector<uint32_t> buf;
    buf.insert(buf.end(), odirs.begin(), odirs.end());
    const size_t o_lptr = buf.size();
    buf.insert(buf.end(), lptr.begin(), lptr.end());
    const size_t o_lcells = buf.size();
    buf.insert(buf.end(), lcells.begin(), lcells.end());
    const size_t o_boxes = buf.size();
    buf.insert(buf.end(), boxes.begin(), boxes.end());
    void *d = nullptr;
    HIP_TRY(hipMalloc(&d, sizeof(uint32_t) * buf.size()));
    hipError_t e = hipMemcpy(d, buf.data(), sizeof(uint32_t) * buf.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(d);
        return fail(ALLIUM_ERR_DEVICE, std::string("box plan upload: ") + hipGetErrorString(e));
    }
    if (ds.box_buf) HIP_TRY(hipFree(ds.box_buf));
    ds.box_buf = d;
    const uint32_t *db = (const uint32_t *)d;
    bp.oct_dirs = (const int *)db;
    bp.lptr = (const int *)(db + o_lptr);
    bp.lcells = db + o_lcells;
    bp.boxes = db + o_boxes;
    ds.box = bp;
    ds.box_ptr = bptr;
    ds.box_threads = threads;
    ds.box_ncu = ncu;
    ds.box_prefetch = prefetch;
    if (ds.graph_exec) { hipGraphExecDestroy(ds.graph_exec); ds.graph_exec = nullptr; }
    return ALLIUM_OK;
}

// default box tiling: 8-element boxes, directions per workgroup as LDS allows
int default_box_plan(allium_sweep *h)
{
    const int dim = h->mesh.dim;
    int rc = ALLIUM_ERR_ARG;
    for (int nd : {5, 4, 3, 2, 1}) {
        rc = build_box_plan(h, 8, 8, dim == 3 ? 8 : 1, nd, 256, 2);
        if (rc != ALLIUM_ERR_ARG) return rc;
    }
    return rc;
}

bool finite_all(const double *v, long n)
{
    for (long i = 0; i < n; i++)
        if (!std::isfinite(v[i])) return false;
    return true;
}

int check_sweep_error(allium_sweep *h, hipStream_t stream)
{
    if (h->ds.prm.fixup == 3) {
        // RyosukeFixupOperator handles at most 3 positive unknowns; the reference aborts beyond
        unsigned long long over = 0;
        HIP_TRY(hipMemcpyAsync(&over, (unsigned long long *)h->fixbuf.p + 2, sizeof(over), hipMemcpyDeviceToHost,
                               stream));
        HIP_TRY(hipStreamSynchronize(stream));
        if (over) {
            HIP_TRY(hipMemsetAsync((unsigned long long *)h->fixbuf.p + 2, 0, sizeof(over), stream));
            return fail(ALLIUM_ERR_NUMERIC, "ryosuke fixup: a cell has more than 3 positive unknowns after clipping "
                                            "(src/fixup.cpp aborts here)");
        }
    }
    if (h->ds.launch_mode != 2) return ALLIUM_OK;
    int flag = 0;
    HIP_TRY(hipMemcpyAsync(&flag, h->ds.d_counters + h->ds.nkeys + 1, sizeof(int), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (flag) return fail(ALLIUM_ERR_DEVICE, "persistent sweep: a dependency wait exceeded its spin limit");
    return ALLIUM_OK;
}

// add the time of a finished ring slot to the total
int harvest_timing(allium::DeviceSweep &ds, int slot)
{
    if (!ds.ring_busy[slot]) return ALLIUM_OK;
    ds.ring_busy[slot] = false;
    HIP_TRY(hipEventSynchronize(ds.ring1[slot]));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ds.ring0[slot], ds.ring1[slot]));
    ds.sweep_seconds += ms * 1e-3;
    return ALLIUM_OK;
}

int harvest_all(allium::DeviceSweep &ds)
{
    for (int i = 0; i < allium::DeviceSweep::kTimingRing; i++) {
        int rc = harvest_timing(ds, i);
        if (rc) return rc;
    }
    return ALLIUM_OK;
}

// psi = L^{-1}(src + iso); phi (may be null) = D psi
int do_sweep(allium_sweep *h, const double *src, const double *iso, double *psi, hipStream_t stream,
             double *phi = nullptr)
{
    allium::DeviceSweep &ds = h->ds;
    if (!ds.timing) {
        HIP_TRY(allium::launch_sweep(ds, src, iso, psi, stream, phi));
        return check_sweep_error(h, stream);
    }
    // the events of this sweep come from the ring; a slot still in use holds a
    // sweep 64 launches old, long finished in practice
    const int slot = ds.ring_next;
    ds.ring_next = (slot + 1) % allium::DeviceSweep::kTimingRing;
    int rc = harvest_timing(ds, slot);
    if (rc) return rc;
    if (!ds.ring0[slot]) HIP_TRY(hipEventCreate(&ds.ring0[slot]));
    if (!ds.ring1[slot]) HIP_TRY(hipEventCreate(&ds.ring1[slot]));
    hipEvent_t e0 = ds.ev0, e1 = ds.ev1;
    ds.ev0 = ds.ring0[slot];
    ds.ev1 = ds.ring1[slot];
    const hipError_t err = allium::launch_sweep(ds, src, iso, psi, stream, phi);
    ds.ev0 = e0;
    ds.ev1 = e1;
    HIP_TRY(err);
    ds.ring_busy[slot] = true;
    return check_sweep_error(h, stream);
}

}  // namespace

int allium_sweep_internal::sweep(allium_sweep *h, const double *src, const double *iso, double *psi,
                                 hipStream_t stream, double *phi)
{
    return do_sweep(h, src, iso, psi, stream, phi);
}

extern "C" {

const char *allium_last_error(void) { return g_last_error.c_str(); }

int allium_sweep_create(const allium_mesh_desc *mesh, int fe_order, const allium_quad_desc *quad, int G,
                        const double *sigma_t, int lump, int precision, allium_sweep_t *out)
{
    if (!mesh || !quad || !sigma_t || !out || G < 1) return fail(ALLIUM_ERR_ARG, "null argument or G < 1");
    *out = nullptr;
    if (lump < 0 || lump > 7)
        return fail(ALLIUM_ERR_ARG, "lump is a bit set: 1 mass, 2 gradient, 4 face terms on the vertex rule");
    // the vertex rule leaves a p = 2 mass matrix singular (8 points, 27 or 9
    // dofs); the GPU cell solves factor without pivoting
    if (lump != 0 && fe_order > 1)
        return fail(ALLIUM_ERR_UNSUPPORTED, "lumped matrices are implemented for fe_order 0 and 1");
    std::string why;
    if (!allium::sweep_supported(mesh->dim, fe_order, why)) return fail(ALLIUM_ERR_UNSUPPORTED, why);
    auto h = std::make_unique<allium_sweep>();
    if (!allium::build_mesh(mesh->dim, mesh->nel, mesh->vertices, mesh->extents, mesh->bc, h->mesh, why))
        return fail(ALLIUM_ERR_ARG, why);
    if (quad->nomega > 0) {
        if (!quad->omega || !quad->weights) return fail(ALLIUM_ERR_ARG, "explicit quadrature needs omega and weights");
        h->quad.dim = mesh->dim;
        for (int a = 0; a < quad->nomega; a++) {
            for (int d = 0; d < 3; d++) h->quad.omega.push_back(d < mesh->dim ? quad->omega[a * mesh->dim + d] : 0.0);
            h->quad.w.push_back(quad->weights[a]);
            h->quad.wsum += quad->weights[a];
        }
    } else if (!allium::level_symmetric(quad->order, mesh->dim, h->quad, why)) {
        return fail(ALLIUM_ERR_ARG, why);
    }
    h->G = G;
    const long Ne = h->mesh.Ne;
    h->sigma.assign(sigma_t, sigma_t + (size_t)G * Ne);
    if (!finite_all(sigma_t, (long)G * Ne)) return fail(ALLIUM_ERR_NUMERIC, "non-finite sigma_t");
    if (!allium::assemble(h->mesh, fe_order, lump, h->ed, why)) return fail(ALLIUM_ERR_ARG, why);
    int st = allium::build_plan(h->mesh, h->quad, h->plan, why);
    if (st == -2) return fail(ALLIUM_ERR_NOT_DAG, why);
    if (st == -3) return fail(ALLIUM_ERR_REFLECT, why);
    if (st) return fail(ALLIUM_ERR_ARG, why);
    h->psi_size = (long)G * h->quad.size() * Ne * h->ed.n;
    h->phi_size = (long)G * Ne * h->ed.n;
    h->ds.precision = precision == ALLIUM_FP32_REFINED ? 1 : 0;
    int rc = upload_all(h.get());
    if (rc) return rc;
    // launch strategy: persistent kernel when the direction classes are independent
    h->ds.launch_mode = 0;
    h->ds.lanes = h->ds.d_wcells ? 0 : 1;
    *out = h.release();
    return ALLIUM_OK;
}

void allium_sweep_destroy(allium_sweep_t h) { delete h; }

int allium_sweep_info(allium_sweep_t h, allium_sweep_info_t *info)
{
    if (!h || !info) return fail(ALLIUM_ERR_ARG, "null argument");
    info->Ne = h->mesh.Ne;
    info->ndof_per_elem = h->ed.n;
    info->Ndof = (long)h->mesh.Ne * h->ed.n;
    info->Nomega = h->quad.size();
    info->G = h->G;
    info->psi_size = h->psi_size;
    info->phi_size = h->phi_size;
    info->levels = h->plan.nlevels;
    info->cells = (long)h->plan.cells.size();
    info->edges = h->plan.nedges;
    info->uniform_geometry = h->mesh.uniform ? 1 : 0;
    info->launch_mode = h->ds.launch_mode;
    return ALLIUM_OK;
}

int allium_sweep_quadrature(allium_sweep_t h, double *omega, double *weights)
{
    if (!h) return fail(ALLIUM_ERR_ARG, "null handle");
    const int dim = h->mesh.dim;
    for (int a = 0; a < h->quad.size(); a++) {
        if (omega)
            for (int d = 0; d < dim; d++) omega[a * dim + d] = h->quad.omega[3 * a + d];
        if (weights) weights[a] = h->quad.w[a];
    }
    return ALLIUM_OK;
}

int allium_sweep_set_sigma(allium_sweep_t h, const double *sigma_t)
{
    if (!h || !sigma_t) return fail(ALLIUM_ERR_ARG, "null argument");
    const long n = (long)h->G * h->mesh.Ne;
    if (!finite_all(sigma_t, n)) return fail(ALLIUM_ERR_NUMERIC, "non-finite sigma_t");
    h->sigma.assign(sigma_t, sigma_t + n);
    // sweeps still queued on the handle's stream read the old values
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(h->d_sigma, sigma_t, sizeof(double) * n, hipMemcpyHostToDevice));
    return ALLIUM_OK;
}

int allium_sweep_set_time_absorption(allium_sweep_t h, double tau)
{
    if (!h) return fail(ALLIUM_ERR_ARG, "null handle");
    if (!std::isfinite(tau)) return fail(ALLIUM_ERR_NUMERIC, "non-finite time absorption");
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->ds.prm.tau = tau;
    h->ds.prm.timedep = 1;
    if (h->mesh.uniform) {
        // the uniform streaming tables carry tau * Mt
        build_streaming(h);
        std::vector<float> kf(h->Kst.size());
        for (size_t i = 0; i < kf.size(); i++) kf[i] = (float)h->Kst[i];
        HIP_TRY(hipMemcpy((void *)h->ds.prm.K, h->Kst.data(), sizeof(double) * h->Kst.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy((void *)h->ds.prm.Kf, kf.data(), sizeof(float) * kf.size(), hipMemcpyHostToDevice));
    }
    if (h->ds.graph_exec) { hipGraphExecDestroy(h->ds.graph_exec); h->ds.graph_exec = nullptr; }
    return ALLIUM_OK;
}

int allium_sweep_get_time_absorption(allium_sweep_t h, double *tau)
{
    if (!h || !tau) return fail(ALLIUM_ERR_ARG, "null argument");
    *tau = h->ds.prm.timedep ? h->ds.prm.tau : 0.0;
    return ALLIUM_OK;
}

int allium_sweep_get_lumping(allium_sweep_t h, int *lump, int *parallel_block_jacobi)
{
    if (!h) return fail(ALLIUM_ERR_ARG, "null handle");
    if (lump) *lump = h->ed.lump;
    if (parallel_block_jacobi) *parallel_block_jacobi = 0;
    return ALLIUM_OK;
}

int allium_sweep_set_precision(allium_sweep_t h, int precision)
{
    if (!h || (precision != ALLIUM_FP64 && precision != ALLIUM_FP32_REFINED))
        return fail(ALLIUM_ERR_ARG, "bad precision");
    h->ds.precision = precision == ALLIUM_FP32_REFINED ? 1 : 0;
    return ALLIUM_OK;
}

int allium_sweep_levels(allium_sweep_t h, int *level_of_node)
{
    if (!h || !level_of_node) return fail(ALLIUM_ERR_ARG, "null argument");
    std::memcpy(level_of_node, h->plan.level.data(), sizeof(int) * h->plan.level.size());
    return ALLIUM_OK;
}

int allium_sweep_mult_device(allium_sweep_t h, const double *d_source, double *d_psi, void *stream)
{
    if (!h || !d_source || !d_psi) return fail(ALLIUM_ERR_ARG, "null argument");
    return do_sweep(h, d_source, nullptr, d_psi, (hipStream_t)stream);
}

int allium_sweep_mult_iso_device(allium_sweep_t h, const double *d_source, const double *d_iso, double *d_psi,
                                 void *stream)
{
    if (!h || !d_iso || !d_psi) return fail(ALLIUM_ERR_ARG, "null argument");
    return do_sweep(h, d_source, d_iso, d_psi, (hipStream_t)stream);
}

int allium_sweep_mult_moments_device(allium_sweep_t h, const double *d_source, const double *d_iso, double *d_psi,
                                     double *d_phi, void *stream)
{
    if (!h || !d_psi || !d_phi || (!d_source && !d_iso)) return fail(ALLIUM_ERR_ARG, "null argument");
    if (d_phi == d_iso || d_phi == d_source || d_phi == d_psi)
        return fail(ALLIUM_ERR_ARG, "phi must not alias the source, iso or psi");
    return do_sweep(h, d_source, d_iso, d_psi, (hipStream_t)stream, d_phi);
}

int allium_sweep_set_fixup(allium_sweep_t h, int type, double psi_min)
{
    if (!h) return fail(ALLIUM_ERR_ARG, "null handle");
    if (type < ALLIUM_FIXUP_NONE || type > ALLIUM_FIXUP_RYOSUKE) return fail(ALLIUM_ERR_ARG, "unknown fixup type");
    if (!std::isfinite(psi_min)) return fail(ALLIUM_ERR_NUMERIC, "non-finite psi_min");
    if (type != ALLIUM_FIXUP_NONE && !allium::fixup_supported(h->mesh.dim, h->ed.p))
        return fail(ALLIUM_ERR_UNSUPPORTED, "fixups are implemented for 2D p=0..2 and 3D p=0..1");
    HIP_TRY(h->fixbuf.ensure(3 * sizeof(unsigned long long)));
    HIP_TRY(hipMemsetAsync(h->fixbuf.p, 0, 3 * sizeof(unsigned long long), h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    auto &ds = h->ds;
    ds.fix_type = type;
    ds.fix_on = type != ALLIUM_FIXUP_NONE;
    ds.prm.fix_min = psi_min;
    ds.prm.fix_counts = (unsigned long long *)h->fixbuf.p;
    ds.prm.fixup = ds.fix_on ? type : 0;
    return ALLIUM_OK;
}

int allium_sweep_use_fixup(allium_sweep_t h, int use)
{
    if (!h) return fail(ALLIUM_ERR_ARG, "null handle");
    auto &ds = h->ds;
    ds.fix_on = use != 0 && ds.fix_type != ALLIUM_FIXUP_NONE;
    ds.prm.fixup = ds.fix_on ? ds.fix_type : 0;
    return ALLIUM_OK;
}

int allium_sweep_get_fixup(allium_sweep_t h, int *type, double *psi_min, int *on)
{
    if (!h) return fail(ALLIUM_ERR_ARG, "null handle");
    if (type) *type = h->ds.fix_type;
    if (psi_min) *psi_min = h->ds.prm.fix_min;
    if (on) *on = h->ds.fix_on ? 1 : 0;
    return ALLIUM_OK;
}

int allium_sweep_set_fixup_monitor(allium_sweep_t h, double *d_monitor)
{
    if (!h) return fail(ALLIUM_ERR_ARG, "null handle");
    h->ds.prm.fix_monitor = d_monitor;
    return ALLIUM_OK;
}

int allium_sweep_fixup_counts(allium_sweep_t h, long long *applied, long long *failed)
{
    if (!h) return fail(ALLIUM_ERR_ARG, "null handle");
    unsigned long long c[3] = {0, 0, 0};
    if (h->fixbuf.p) {
        HIP_TRY(hipMemcpyAsync(c, h->fixbuf.p, sizeof(c), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    if (applied) *applied = (long long)c[0];
    if (failed) *failed = (long long)c[1];
    return ALLIUM_OK;
}

int allium_sweep_mult(allium_sweep_t h, const double *source, double *psi)
{
    if (!h || !source || !psi) return fail(ALLIUM_ERR_ARG, "null argument");
    const size_t bytes = sizeof(double) * h->psi_size;
    HIP_TRY(h->a.ensure(bytes));
    HIP_TRY(hipMemcpyAsync(h->a.p, source, bytes, hipMemcpyHostToDevice, h->stream));
    int rc = do_sweep(h, h->a.d(), nullptr, h->a.d(), h->stream);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(psi, h->a.p, bytes, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return ALLIUM_OK;
}

int allium_sweep_set_launch(allium_sweep_t h, int mode)
{
    if (!h || mode < -1 || mode > 5) return fail(ALLIUM_ERR_ARG, "bad launch mode");
    if (mode == 3 && !h->ds.angle_independent)
        return fail(ALLIUM_ERR_UNSUPPORTED, "direction sweeps need directions that do not couple (no reflective side)");
    if (allium::coop_supported(h->mesh.dim, h->ed.p) && mode != 0 && mode != 1 && mode != -1)
        return fail(ALLIUM_ERR_UNSUPPORTED, "3D p=2 sweeps run cooperative level launches (modes 0 and 1)");
    if ((mode == 4 || mode == 5) && !h->ds.box_buf) {
        int rc = default_box_plan(h);
        if (rc) return rc;
    }
    if (mode == 2)
        return fail(ALLIUM_ERR_UNSUPPORTED,
                    "persistent sweep kernel is disabled: a run on MI355X did not terminate and the cause "
                    "is not identified");
    if (mode >= 0) { h->ds.launch_mode = mode; return ALLIUM_OK; }
    // autotune: time each strategy and block size on scratch vectors
    const size_t bytes = sizeof(double) * h->psi_size;
    HIP_TRY(h->b.ensure(bytes));
    HIP_TRY(h->c.ensure(bytes));
    HIP_TRY(hipMemsetAsync(h->b.p, 0, bytes, h->stream));
    struct Cand { int mode, block, lanes, bx, by, bz, nd, pf; };
    std::vector<Cand> cands;
    const bool coop = allium::coop_supported(h->mesh.dim, h->ed.p);
    if (coop) {
        cands.push_back({0, 256, 1, 0, 0, 0, 0, 0});
        cands.push_back({1, 256, 1, 0, 0, 0, 0, 0});
    }
    if (coop) {
    } else if (h->ds.angle_independent)
        for (int bl : {256, 512}) cands.push_back({3, bl, 1, 0, 0, 0, 0, 0});
    for (int md : {0, 1})
        for (int bl : {64, 128, 256})
            for (int ln : {0, 1, 2, 4}) {
                if (coop) continue;
                if (ln == 0 && !h->ds.d_wcells) continue;
                if (ln > 0 && h->ed.n % ln) continue;
                cands.push_back({md, bl, ln, 0, 0, 0, 0, 0});
            }
    const bool boxes = !coop && h->mesh.uniform && h->ds.angle_independent && h->ed.lump == 0 &&
                       allium::box_sweep_supported(h->mesh.dim, h->ed.p);
    if (boxes) {
        const bool d3 = h->mesh.dim == 3;
        // p = 1: lane-split cells, workgroups up to 1024
        const bool split = h->ed.p == 1;
        const std::vector<int> ths = split ? std::vector<int>{256, 512, 1024} : std::vector<int>{128, 256};
        const std::vector<int> pfs = split ? std::vector<int>{0, 2, 3} : std::vector<int>{0, 2};
        // cubic boxes, and boxes halved along one axis (shorter in-box level
        // chains, smaller LDS planes)
        struct Shape { int x, y, z; };
        const std::vector<Shape> shapes =
            d3 ? std::vector<Shape>{{4, 4, 4}, {8, 8, 8}, {16, 16, 16}, {16, 16, 8}, {16, 8, 16}, {8, 16, 16}}
               : std::vector<Shape>{{8, 8, 1}, {16, 16, 1}, {32, 32, 1}, {32, 16, 1}, {16, 32, 1}};
        for (const Shape &b : shapes)
            for (int nd : {1, 2, 3, 5, 10})
                for (int th : ths)
                    for (int pf : pfs) cands.push_back({5, th, 1, b.x, b.y, b.z, nd, pf});
    }
    // keep the caller's box tiling to restore it when a non-box mode wins
    const allium::BoxPlan keep_box = h->ds.box;
    const int keep_threads = h->ds.box_threads, keep_pf = h->ds.box_prefetch;
    double best = 1e30;
    Cand pick = {0, 128, 1, 0, 0, 0, 0, 0};
    const bool timing = h->ds.timing;
    h->ds.timing = false;
    for (const auto &cd : cands) {
        if (cd.mode >= 4) {
            int rc = build_box_plan(h, cd.bx, cd.by, cd.bz, cd.nd, cd.block, cd.pf);
            if (rc == ALLIUM_ERR_ARG) continue;   // LDS does not fit this tiling
            if (rc) { h->ds.timing = timing; return rc; }
        }
        h->ds.launch_mode = cd.mode;
        if (cd.mode == 3) h->ds.dir_block = cd.block;
        else if (cd.mode < 3) h->ds.block = cd.block;
        h->ds.lanes = cd.lanes;
        if (h->ds.graph_exec) { hipGraphExecDestroy(h->ds.graph_exec); h->ds.graph_exec = nullptr; }
        int rc = do_sweep(h, h->b.d(), nullptr, h->c.d(), h->stream);
        if (rc) { h->ds.timing = timing; return rc; }
        HIP_TRY(hipStreamSynchronize(h->stream));
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        HIP_TRY(hipEventRecord(e0, h->stream));
        const int reps = 3;
        for (int r = 0; r < reps; r++) HIP_TRY(allium::launch_sweep(h->ds, h->b.d(), nullptr, h->c.d(), h->stream));
        HIP_TRY(hipEventRecord(e1, h->stream));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        int rc2 = check_sweep_error(h, h->stream);
        if (rc2) { h->ds.timing = timing; return rc2; }
        if (ms / reps < best) { best = ms / reps; pick = cd; }
    }
    h->ds.timing = timing;
    h->ds.launch_mode = pick.mode;
    if (pick.mode == 3) h->ds.dir_block = pick.block;
    else if (pick.mode < 3) h->ds.block = pick.block;
    h->ds.lanes = pick.lanes;
    if (pick.mode >= 4) {
        int rc = build_box_plan(h, pick.bx, pick.by, pick.bz, pick.nd, pick.block, pick.pf);
        if (rc) return rc;
    } else if (boxes && keep_box.nd > 0) {
        int rc = build_box_plan(h, keep_box.bs[0], keep_box.bs[1], keep_box.bs[2], keep_box.nd, keep_threads, keep_pf);
        if (rc) return rc;
    }
    if (h->ds.graph_exec) { hipGraphExecDestroy(h->ds.graph_exec); h->ds.graph_exec = nullptr; }
    int rc = harvest_all(h->ds);
    if (rc) return rc;
    h->ds.sweeps = 0;
    h->ds.sweep_seconds = 0;
    return ALLIUM_OK;
}

int allium_sweep_set_tuning(allium_sweep_t h, int mode, int block, int lanes)
{
    // the level and box kernels are compiled with __launch_bounds__(256), the
    // direction-sweep kernel (mode 3) with __launch_bounds__(512) and the
    // lane-split p = 1 box kernel with __launch_bounds__(1024)
    const int maxb = !h ? 0 : mode == 3 ? 512 : (mode >= 4 && h->ed.p == 1) ? 1024 : 256;
    if (!h || mode < 0 || mode > 5 || block < 64 || block > maxb || block % 64 ||
        (lanes != 0 && lanes != 1 && lanes != 2 && lanes != 4))
        return fail(ALLIUM_ERR_ARG, "bad tuning (mode 0..5, block 64..256 step 64 (512 for mode 3, 1024 for p = 1 "
                                    "box sweeps), lanes 0/1/2/4)");
    if (mode >= 3 && lanes != 1)
        return fail(ALLIUM_ERR_ARG, "bad tuning: direction and box sweeps use one lane per cell");
    if (mode >= 4) {
        int rc = allium_sweep_set_launch(h, mode);
        if (rc) return rc;
        const allium::BoxPlan &bp = h->ds.box;
        return build_box_plan(h, bp.bs[0], bp.bs[1], bp.bs[2], bp.nd, block, h->ds.box_prefetch);
    }
    if (lanes == 0 && !h->ds.d_wcells)
        return fail(ALLIUM_ERR_ARG, "direction-uniform waves need uniform geometry");
    if (lanes > 0 && h->ed.n % lanes) return fail(ALLIUM_ERR_ARG, "lanes per cell must divide the dofs per element");
    int rc = allium_sweep_set_launch(h, mode);
    if (rc) return rc;
    if (mode == 3) h->ds.dir_block = block;
    else h->ds.block = block;
    h->ds.lanes = lanes;
    if (h->ds.graph_exec) { hipGraphExecDestroy(h->ds.graph_exec); h->ds.graph_exec = nullptr; }
    return ALLIUM_OK;
}

int allium_sweep_get_tuning(allium_sweep_t h, int *mode, int *block, int *lanes)
{
    if (!h) return fail(ALLIUM_ERR_ARG, "null handle");
    if (mode) *mode = h->ds.launch_mode;
    if (block)
        *block = h->ds.launch_mode == 3 ? h->ds.dir_block : h->ds.launch_mode >= 4 ? h->ds.box_threads : h->ds.block;
    if (lanes) *lanes = h->ds.launch_mode >= 3 ? 1 : h->ds.lanes;
    return ALLIUM_OK;
}

int allium_sweep_set_box_tiling(allium_sweep_t h, int bx, int by, int bz, int dirs_per_group, int threads,
                                int prefetch)
{
    if (!h) return fail(ALLIUM_ERR_ARG, "null handle");
    return build_box_plan(h, bx, by, bz, dirs_per_group, threads, prefetch);
}

int allium_sweep_get_box_tiling(allium_sweep_t h, int *bx, int *by, int *bz, int *dirs_per_group, int *threads,
                                int *prefetch)
{
    if (!h) return fail(ALLIUM_ERR_ARG, "null handle");
    if (!h->ds.box_buf) return fail(ALLIUM_ERR_UNSUPPORTED, "no box tiling (box sweeps not set up)");
    const allium::BoxPlan &bp = h->ds.box;
    if (bx) *bx = bp.bs[0];
    if (by) *by = bp.bs[1];
    if (bz) *bz = bp.bs[2];
    if (dirs_per_group) *dirs_per_group = bp.nd;
    if (threads) *threads = h->ds.box_threads;
    if (prefetch) *prefetch = h->ds.box_prefetch;
    return ALLIUM_OK;
}

int allium_apply_L_device(allium_sweep_t h, const double *d_psi, double *d_Lpsi, void *stream)
{
    if (!h || !d_psi || !d_Lpsi || d_psi == d_Lpsi) return fail(ALLIUM_ERR_ARG, "bad argument (in-place not allowed)");
    HIP_TRY(allium::launch_apply_L(h->ds.prm, d_psi, d_Lpsi, (hipStream_t)stream));
    return ALLIUM_OK;
}

int allium_apply_L(allium_sweep_t h, const double *psi, double *Lpsi)
{
    if (!h || !psi || !Lpsi) return fail(ALLIUM_ERR_ARG, "null argument");
    const size_t bytes = sizeof(double) * h->psi_size;
    HIP_TRY(h->a.ensure(bytes));
    HIP_TRY(h->b.ensure(bytes));
    HIP_TRY(hipMemcpyAsync(h->a.p, psi, bytes, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(allium::launch_apply_L(h->ds.prm, h->a.d(), h->b.d(), h->stream));
    HIP_TRY(hipMemcpyAsync(Lpsi, h->b.p, bytes, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return ALLIUM_OK;
}

int allium_balance_device(allium_sweep_t h, const double *d_psi, double *out, void *stream)
{
    if (!h || !d_psi || !out) return fail(ALLIUM_ERR_ARG, "null argument");
    const int nb = allium::balance_blocks(h->ds.prm);
    HIP_TRY(h->part.ensure(sizeof(double) * 2 * nb));
    HIP_TRY(allium::launch_balance(h->ds.prm, d_psi, h->part.d(), (hipStream_t)stream));
    std::vector<double> hp(2 * nb);
    HIP_TRY(hipMemcpyAsync(hp.data(), h->part.p, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    double ab = 0.0, lk = 0.0;
    for (int i = 0; i < nb; i++) { ab += hp[2 * i]; lk += hp[2 * i + 1]; }
    *out = lk + ab;
    return ALLIUM_OK;
}

int allium_balance_device_noalloc(allium_sweep_t h, const double *d_psi, double *out, void *stream)
{
    if (!h || !d_psi || !out) return fail(ALLIUM_ERR_ARG, "null argument");
    if (allium::balance_blocks(h->ds.prm) > allium::kStaticBalanceBlocks)
        return fail(ALLIUM_ERR_ARG, "Ne * G exceeds the static balance workspace (at most 8388608 element-groups)");
    double ab = 0.0, lk = 0.0;
    HIP_TRY(allium::balance_static(h->ds.prm, d_psi, &ab, &lk, (hipStream_t)stream));
    *out = lk + ab;
    return ALLIUM_OK;
}

int allium_balance(allium_sweep_t h, const double *psi, double *out)
{
    if (!h || !psi || !out) return fail(ALLIUM_ERR_ARG, "null argument");
    const size_t bytes = sizeof(double) * h->psi_size;
    HIP_TRY(h->a.ensure(bytes));
    HIP_TRY(hipMemcpyAsync(h->a.p, psi, bytes, hipMemcpyHostToDevice, h->stream));
    return allium_balance_device(h, h->a.d(), out, h->stream);
}

int allium_form_source(allium_sweep_t h, const double *Q, const double *inflow, double *source)
{
    if (!h || !source) return fail(ALLIUM_ERR_ARG, "null argument");
    allium::form_source(h->mesh, h->ed, h->quad, h->G, Q, inflow, source);
    return ALLIUM_OK;
}

int allium_moments_device(allium_sweep_t h, const double *d_psi, double *d_phi, void *stream)
{
    if (!h || !d_psi || !d_phi) return fail(ALLIUM_ERR_ARG, "null argument");
    HIP_TRY(allium::launch_moments(h->ds.prm, d_psi, d_phi, (hipStream_t)stream));
    return ALLIUM_OK;
}

int allium_moments_transpose_device(allium_sweep_t h, const double *d_phi, double *d_psi, void *stream)
{
    if (!h || !d_psi || !d_phi) return fail(ALLIUM_ERR_ARG, "null argument");
    HIP_TRY(allium::launch_moments_T(h->ds.prm, 1.0 / h->quad.wsum, d_phi, d_psi, (hipStream_t)stream));
    return ALLIUM_OK;
}

int allium_moments_aniso_device(allium_sweep_t h, const double *d_psi, double *d_mom, void *stream)
{
    if (!h || !d_psi || !d_mom) return fail(ALLIUM_ERR_ARG, "null argument");
    HIP_TRY(allium::launch_moments_aniso(h->ds.prm, h->mesh.dim, d_psi, d_mom, (hipStream_t)stream));
    return ALLIUM_OK;
}

int allium_moments_aniso_transpose_device(allium_sweep_t h, const double *d_mom, double *d_psi, void *stream)
{
    if (!h || !d_psi || !d_mom) return fail(ALLIUM_ERR_ARG, "null argument");
    HIP_TRY(allium::launch_moments_aniso_T(h->ds.prm, h->mesh.dim, h->quad.wsum, d_mom, d_psi, (hipStream_t)stream));
    return ALLIUM_OK;
}

int allium_mass_device(allium_sweep_t h, const double *d_coef, const double *d_x, double *d_y, void *stream)
{
    if (!h || !d_coef || !d_x || !d_y) return fail(ALLIUM_ERR_ARG, "null argument");
    HIP_TRY(allium::launch_mass(h->ds.prm, h->ds.prm.Mt, d_coef, 1.0, d_x, d_y, (hipStream_t)stream));
    return ALLIUM_OK;
}

// source, phi and psi are host pointers, or device pointers when `device`
static int source_iteration(allium_sweep_t h, const double *sigma_s, const double *source,
                            const allium_si_options *opt, double *phi, double *psi, double *norms, int *iterations,
                            bool device)
{
    if (!h || !sigma_s || !source || !opt || !phi) return fail(ALLIUM_ERR_ARG, "null argument");
    if (opt->maxit < 1) return fail(ALLIUM_ERR_ARG, "maxit < 1");
    if (!finite_all(sigma_s, h->mesh.Ne)) return fail(ALLIUM_ERR_NUMERIC, "non-finite sigma_s");
    hipStream_t s = h->stream;
    const long np = h->phi_size, nps = h->psi_size;
    if (!device) HIP_TRY(h->q.ensure(sizeof(double) * nps));
    HIP_TRY(h->a.ensure(sizeof(double) * nps));
    HIP_TRY(h->b.ensure(sizeof(double) * np));
    HIP_TRY(h->x.ensure(sizeof(double) * np));
    HIP_TRY(h->y.ensure(sizeof(double) * np));
    HIP_TRY(h->iso.ensure(sizeof(double) * np));
    HIP_TRY(h->coef.ensure(sizeof(double) * h->mesh.Ne));
    const int nb = allium::si_update_blocks(np);
    HIP_TRY(h->part.ensure(sizeof(double) * nb));
    if (!device) HIP_TRY(hipMemcpyAsync(h->q.p, source, sizeof(double) * nps, hipMemcpyHostToDevice, s));
    const double *dq = device ? source : h->q.d();
    const hipMemcpyKind out_kind = device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    HIP_TRY(hipMemcpyAsync(h->coef.p, sigma_s, sizeof(double) * h->mesh.Ne, hipMemcpyHostToDevice, s));
    const double inv_w = 1.0 / h->quad.wsum;
    const int prec_save = h->ds.precision;
    // every sweep here is followed by the moment kernel's read of psi
    struct StreamGuard {
        allium_sweep *h;
        int saved;
        ~StreamGuard() { h->ds.stream_psi = saved; }
    } sguard{h, h->ds.stream_psi};
    h->ds.stream_psi = 1;
    int rc = ALLIUM_OK;
    std::vector<double> hp(nb);
    double r0 = 0.0;
    int it = 0;
    // learned correction and the per-iteration history both use element averages
    allium_cnn *cnn = nullptr;
    allium_dsa *dsa = nullptr;
    if (opt->accel == 1) {
        if (!opt->accel_handle) return fail(ALLIUM_ERR_ARG, "accel 1 needs an allium_cnn_t in accel_handle");
        if (h->mesh.dim != 3 || !h->mesh.uniform || h->G != 1)
            return fail(ALLIUM_ERR_UNSUPPORTED, "the learned correction needs a uniform 3D mesh and one group");
        cnn = static_cast<allium_cnn *>(opt->accel_handle);
        if (opt->accel_fallback) dsa = static_cast<allium_dsa *>(opt->accel_fallback);
    } else if (opt->accel == 2) {
        if (!opt->accel_handle) return fail(ALLIUM_ERR_ARG, "accel 2 needs an allium_dsa_t in accel_handle");
        dsa = static_cast<allium_dsa *>(opt->accel_handle);
    } else if (opt->accel != 0) {
        return fail(ALLIUM_ERR_ARG, "unknown acceleration");
    }
    if (dsa && allium_dsa_internal::handle_of(dsa) != h)
        return fail(ALLIUM_ERR_ARG, "the DSA object was built on another sweep handle");
    // DSA runs from the first iteration (accel 2) or once the learned correction is off
    bool dsa_on = dsa && !cnn;
    if (opt->history && h->G != 1) return fail(ALLIUM_ERR_UNSUPPORTED, "iteration history needs one group");
    const int n = h->ed.n;
    const long Ne = h->mesh.Ne;
    if (cnn || opt->history) {
        // element average = sum_i w_i phi_i with w_i = (row sum of the element mass) / volume
        std::vector<double> w(n, 0.0);
        double vol = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) { w[i] += h->ed.Mu[(size_t)i * n + j]; vol += h->ed.Mu[(size_t)i * n + j]; }
        for (double &v : w) v /= vol;
        HIP_TRY(h->wavg.ensure(sizeof(double) * n));
        HIP_TRY(hipMemcpy(h->wavg.p, w.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    }
    const bool want_r = cnn || dsa;
    if (want_r) HIP_TRY(h->r.ensure(sizeof(double) * np));
    if (want_r) HIP_TRY(hipMemsetAsync(h->r.p, 0, sizeof(double) * np, s));
    if (opt->history) HIP_TRY(h->havg.ensure(sizeof(double) * Ne));
    if (cnn && (opt->accel_prolong < 0 || opt->accel_prolong > 1 || (opt->accel_prolong == 1 && n != 8)))
        return fail(ALLIUM_ERR_ARG, "accel_prolong must be 0, or 1 with p = 1");
    bool cnn_on = cnn != nullptr;
    const int patience = std::max(1, opt->accel_patience);
    double best = 0.0;
    int best_it = 0;
    if (cnn) HIP_TRY(h->xb.ensure(sizeof(double) * np));
    // after x += r: apply the learned correction until `patience` iterations
    // have passed without a new lowest residual; then the iterate of the
    // lowest residual is restored and plain SI (or the DSA fallback) goes on
    // from it.  Record the history.
    auto after_update = [&](double nrm) -> int {
        if (cnn_on) {
            if (it == 1 || nrm < best) {
                best = nrm;
                best_it = it;
                HIP_TRY(hipMemcpyAsync(h->xb.p, h->x.p, sizeof(double) * np, hipMemcpyDeviceToDevice, s));
            } else if (it - best_it >= patience) {
                cnn_on = false;
                dsa_on = dsa != nullptr;
                HIP_TRY(hipMemcpyAsync(h->x.p, h->xb.p, sizeof(double) * np, hipMemcpyDeviceToDevice, s));
            }
            if (cnn_on) {
                int rc2 = allium_cnn_internal::correct(cnn, h->mesh.nel[2], h->mesh.nel[1], h->mesh.nel[0], n,
                                                       h->wavg.d(), h->r.d(), h->d_sigma, h->coef.d(), h->x.d(), s,
                                                       opt->accel_prolong);
                if (rc2) return rc2;
            }
        }
        if (opt->history) {
            double *hrow = opt->history + (size_t)(it - 1) * 2 * Ne;
            HIP_TRY(allium_cnn_internal::elem_avg(Ne, n, h->wavg.d(), h->x.d(), h->havg.d(), s));
            HIP_TRY(hipMemcpyAsync(hrow, h->havg.p, sizeof(double) * Ne, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            if (cnn) {
                HIP_TRY(allium_cnn_internal::elem_avg(Ne, n, h->wavg.d(), h->r.d(), h->havg.d(), s));
                HIP_TRY(hipMemcpyAsync(hrow + Ne, h->havg.p, sizeof(double) * Ne, hipMemcpyDeviceToHost, s));
                HIP_TRY(hipStreamSynchronize(s));
            } else {
                const double *prev = it > 1 ? hrow - 2 * Ne : nullptr;
                for (long e = 0; e < Ne; e++) hrow[Ne + e] = hrow[e] - (prev ? prev[e] : 0.0);
            }
        }
        return ALLIUM_OK;
    };
    // ||phi_{k+1} - phi_k|| from the update's partial sums, or with DSA on the
    // preconditioned change z = r + K^{-1} Ms r (SLISolver with a preconditioner)
    auto residual_norm = [&](double &nrm) -> int {
        if (!dsa_on || opt->raw_norms) {
            HIP_TRY(hipMemcpyAsync(hp.data(), h->part.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            nrm = 0.0;
            for (int i = 0; i < nb; i++) nrm += hp[i];
            nrm = std::sqrt(nrm);
            if (opt->raw_norms) opt->raw_norms[it - 1] = nrm;
        }
        if (opt->inner_iterations) opt->inner_iterations[it - 1] = 0;
        if (dsa_on) {
            int rc2 = allium_dsa_internal::correct(dsa, h->r.d(), h->x.d(), s, &nrm);
            if (rc2) return rc2;
            if (opt->inner_iterations) opt->inner_iterations[it - 1] = allium_dsa_internal::last_inner(dsa);
        }
        return ALLIUM_OK;
    };
    // FP32-refined iteration inside the box sweep (uniform p = 1, box launches): the
    // scattered part psi_s = L^{-1} D^T Ms phi is corrected in place each iteration,
    // psi_s += L32^{-1}(D^T Ms phi - L psi_s), the defect formed by the sweep itself in
    // FP64 from the new upwind traces and the cell's previous psi_s; phi = D L^{-1} q + D psi_s
    const bool fused_refine = opt->precision == ALLIUM_FP32_REFINED && !h->ds.fix_on && h->mesh.uniform &&
                              h->ed.p == 1 && (h->ds.launch_mode == 4 || h->ds.launch_mode == 5) && h->ds.box_buf;
    if (fused_refine) {
        h->ds.precision = 0;
        rc = do_sweep(h, dq, nullptr, h->a.d(), s, h->b.d());
        if (rc) { h->ds.precision = prec_save; return rc; }
        HIP_TRY(h->c.ensure(sizeof(double) * nps));
        HIP_TRY(hipMemsetAsync(h->c.p, 0, sizeof(double) * nps, s));
        HIP_TRY(hipMemsetAsync(h->x.p, 0, sizeof(double) * np, s));
        for (it = 1;; it++) {
            HIP_TRY(allium::launch_mass(h->ds.prm, h->ds.prm.Mt, h->coef.d(), inv_w, h->x.d(), h->iso.d(), s));
            h->ds.refine = 1;
            rc = do_sweep(h, nullptr, h->iso.d(), h->c.d(), s, h->y.d());
            h->ds.refine = 0;
            if (rc) break;
            HIP_TRY(allium::launch_si_update(np, h->b.d(), h->y.d(), h->x.d(), h->part.d(), s,
                                             want_r ? h->r.d() : nullptr));
            double nrm = 0.0;
            rc = residual_norm(nrm);
            if (rc) break;
            if (norms) norms[it - 1] = nrm;
            if (!std::isfinite(nrm)) { rc = fail(ALLIUM_ERR_NUMERIC, "non-finite residual"); break; }
            if (it == 1) r0 = std::max(nrm * opt->reltol, opt->abstol);
            if (nrm < r0 || it >= opt->maxit) break;
            rc = after_update(nrm);
            if (rc) break;
        }
        if (rc) { h->ds.precision = prec_save; return rc; }
        // final FP64 sweep: psi = L^{-1}(D^T Ms phi + q)
        HIP_TRY(allium::launch_mass(h->ds.prm, h->ds.prm.Mt, h->coef.d(), inv_w, h->x.d(), h->iso.d(), s));
        rc = do_sweep(h, dq, h->iso.d(), h->a.d(), s);
        h->ds.precision = prec_save;
        if (rc) return rc;
        if (psi) HIP_TRY(hipMemcpyAsync(psi, h->a.p, sizeof(double) * nps, out_kind, s));
    } else if (opt->precision == ALLIUM_FP32_REFINED) {
        // defect correction: psi += L32^{-1} (q + D^T Ms phi - L psi), residuals in FP64.
        // A fixup would clip the correction, which may legitimately be negative.
        if (h->ds.fix_on)
            return fail(ALLIUM_ERR_UNSUPPORTED, "FP32-refined source iteration with a negative-flux fixup");
        HIP_TRY(h->c.ensure(sizeof(double) * nps));
        HIP_TRY(hipMemsetAsync(h->a.p, 0, sizeof(double) * nps, s));
        HIP_TRY(hipMemsetAsync(h->x.p, 0, sizeof(double) * np, s));
        h->ds.precision = 1;
        for (it = 1;; it++) {
            HIP_TRY(allium::launch_mass(h->ds.prm, h->ds.prm.Mt, h->coef.d(), inv_w, h->x.d(), h->iso.d(), s));
            HIP_TRY(allium::launch_defect_L(h->ds.prm, dq, h->a.d(), h->c.d(), s));
            rc = do_sweep(h, h->c.d(), h->iso.d(), h->c.d(), s);
            if (rc) break;
            HIP_TRY(allium::launch_axpby(nps, 1.0, h->a.d(), 1.0, h->c.d(), h->a.d(), s));
            HIP_TRY(allium::launch_moments(h->ds.prm, h->a.d(), h->y.d(), s));
            // r = D psi_{k+1} - phi_k ; phi_{k+1} = D psi_{k+1}
            HIP_TRY(allium::launch_si_update(np, h->y.d(), nullptr, h->x.d(), h->part.d(), s,
                                             want_r ? h->r.d() : nullptr));
            double nrm = 0.0;
            rc = residual_norm(nrm);
            if (rc) break;
            if (norms) norms[it - 1] = nrm;
            if (!std::isfinite(nrm)) { rc = fail(ALLIUM_ERR_NUMERIC, "non-finite residual"); break; }
            if (it == 1) r0 = std::max(nrm * opt->reltol, opt->abstol);
            if (nrm < r0 || it >= opt->maxit) break;
            rc = after_update(nrm);
            if (rc) break;
        }
        h->ds.precision = prec_save;
        if (rc) return rc;
        // psi holds the FP64-corrected angular flux of the last iterate
        if (psi) HIP_TRY(hipMemcpyAsync(psi, h->a.p, sizeof(double) * nps, out_kind, s));
    } else {
        h->ds.precision = 0;
        rc = do_sweep(h, dq, nullptr, h->a.d(), s, h->b.d());
        if (rc) { h->ds.precision = prec_save; return rc; }
        HIP_TRY(hipMemsetAsync(h->x.p, 0, sizeof(double) * np, s));
        for (it = 1;; it++) {
            HIP_TRY(allium::launch_mass(h->ds.prm, h->ds.prm.Mt, h->coef.d(), inv_w, h->x.d(), h->iso.d(), s));
            rc = do_sweep(h, nullptr, h->iso.d(), h->a.d(), s, h->y.d());
            if (rc) break;
            HIP_TRY(allium::launch_si_update(np, h->b.d(), h->y.d(), h->x.d(), h->part.d(), s,
                                             want_r ? h->r.d() : nullptr));
            double nrm = 0.0;
            rc = residual_norm(nrm);
            if (rc) break;
            if (norms) norms[it - 1] = nrm;
            if (!std::isfinite(nrm)) { rc = fail(ALLIUM_ERR_NUMERIC, "non-finite residual"); break; }
            if (it == 1) r0 = std::max(nrm * opt->reltol, opt->abstol);
            if (nrm < r0 || it >= opt->maxit) break;
            rc = after_update(nrm);
            if (rc) break;
        }
        if (rc) { h->ds.precision = prec_save; return rc; }
        // final sweep: psi = L^{-1}(D^T Ms phi + q)
        HIP_TRY(allium::launch_mass(h->ds.prm, h->ds.prm.Mt, h->coef.d(), inv_w, h->x.d(), h->iso.d(), s));
        rc = do_sweep(h, dq, h->iso.d(), h->a.d(), s);
        h->ds.precision = prec_save;
        if (rc) return rc;
        if (psi) HIP_TRY(hipMemcpyAsync(psi, h->a.p, sizeof(double) * nps, out_kind, s));
    }
    HIP_TRY(hipMemcpyAsync(phi, h->x.p, sizeof(double) * np, out_kind, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (iterations) *iterations = it;
    const bool converged = norms ? (norms[it - 1] < r0) : true;
    if (!converged) return fail(ALLIUM_ERR_NOT_CONVERGED, "source iteration reached maxit");
    return ALLIUM_OK;
}

int allium_source_iteration(allium_sweep_t h, const double *sigma_s, const double *source,
                            const allium_si_options *opt, double *phi, double *psi, double *norms, int *iterations)
{
    return source_iteration(h, sigma_s, source, opt, phi, psi, norms, iterations, false);
}

int allium_source_iteration_device(allium_sweep_t h, const double *sigma_s, const double *d_source,
                                   const allium_si_options *opt, double *d_phi, double *d_psi, double *norms,
                                   int *iterations)
{
    return source_iteration(h, sigma_s, d_source, opt, d_phi, d_psi, norms, iterations, true);
}

}  // extern "C"

struct allium_sweep_batch {
    std::vector<allium_sweep *> h;
    std::vector<allium::CoopLevel> levels;
    std::vector<allium::SweepParams> prms;
    allium::SweepParams *d_prms = nullptr;
    bool uniform = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double seconds = 0.0;
    long sweeps = 0;
    ~allium_sweep_batch()
    {
        if (d_prms) hipFree(d_prms);
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
    }
};

extern "C" {

int allium_sweep_batch_create(const allium_sweep_t *handles, int count, allium_sweep_batch_t *out)
{
    if (!handles || !out || count < 1 || count > allium::kMaxBatch)
        return fail(ALLIUM_ERR_ARG, "batch needs 1..16 handles");
    *out = nullptr;
    auto b = std::make_unique<allium_sweep_batch>();
    for (int i = 0; i < count; i++) {
        allium_sweep *h = handles[i];
        if (!h) return fail(ALLIUM_ERR_ARG, "null handle in batch");
        if (!allium::coop_supported(h->mesh.dim, h->ed.p))
            return fail(ALLIUM_ERR_UNSUPPORTED, "batched sweeps are implemented for 3D p=2 handles");
        if (i > 0 && h->mesh.uniform != handles[0]->mesh.uniform)
            return fail(ALLIUM_ERR_ARG, "batched handles must all be uniform or all general geometry");
        b->h.push_back(h);
    }
    b->uniform = handles[0]->mesh.uniform;
    int maxl = 0;
    for (auto *h : b->h) maxl = std::max(maxl, h->ds.nlevels);
    for (int L = 0; L < maxl; L++) {
        allium::CoopLevel cl{};
        cl.ninst = count;
        cl.start[0] = 0;
        for (int i = 0; i < count; i++) {
            const auto &ds = b->h[i]->ds;
            const bool has = L < ds.nlevels;
            cl.cbeg[i] = has ? ds.level_ptr[L] : 0;
            cl.ncells[i] = has ? (int)(ds.level_ptr[L + 1] - ds.level_ptr[L]) : 0;
            const int items = cl.ncells[i] * b->h[i]->G;   // workgroup-aligned
            cl.start[i + 1] = cl.start[i] + (items + allium::kCoopCells - 1) / allium::kCoopCells * allium::kCoopCells;
        }
        b->levels.push_back(cl);
    }
    b->prms.resize(count);
    HIP_TRY(hipMalloc(&b->d_prms, sizeof(allium::SweepParams) * count));
    HIP_TRY(hipEventCreate(&b->e0));
    HIP_TRY(hipEventCreate(&b->e1));
    *out = b.release();
    return ALLIUM_OK;
}

void allium_sweep_batch_destroy(allium_sweep_batch_t b) { delete b; }

int allium_sweep_batch_mult_device(allium_sweep_batch_t b, const double *const *d_source, const double *const *d_iso,
                                   double *const *d_psi, void *stream)
{
    if (!b || !d_psi) return fail(ALLIUM_ERR_ARG, "null argument");
    const int n = (int)b->h.size();
    const int prec = b->h[0]->ds.precision;
    for (int i = 0; i < n; i++) {
        if (!d_psi[i]) return fail(ALLIUM_ERR_ARG, "null psi in batch");
        if (b->h[i]->ds.precision != prec) return fail(ALLIUM_ERR_ARG, "batched handles must share precision");
        b->prms[i] = b->h[i]->ds.prm;
        b->prms[i].src = d_source ? d_source[i] : nullptr;
        b->prms[i].iso = d_iso ? d_iso[i] : nullptr;
        b->prms[i].psi = d_psi[i];
    }
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipMemcpy(b->d_prms, b->prms.data(), sizeof(allium::SweepParams) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipEventRecord(b->e0, s));
    for (const auto &cl : b->levels)
        HIP_TRY(allium::launch_coop_level(b->prms[0], b->d_prms, cl, 3, 2, b->uniform, prec, s));
    HIP_TRY(hipEventRecord(b->e1, s));
    HIP_TRY(hipEventSynchronize(b->e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, b->e0, b->e1));
    b->seconds += ms * 1e-3;
    b->sweeps++;
    return ALLIUM_OK;
}

int allium_sweep_batch_timing(allium_sweep_batch_t b, double *seconds, long *sweeps)
{
    if (!b) return fail(ALLIUM_ERR_ARG, "null batch");
    if (seconds) *seconds = b->seconds;
    if (sweeps) *sweeps = b->sweeps;
    return ALLIUM_OK;
}

int allium_sweep_timing(allium_sweep_t h, double *sweep_seconds, long *sweeps)
{
    if (!h) return fail(ALLIUM_ERR_ARG, "null handle");
    int rc = harvest_all(h->ds);
    if (rc) return rc;
    if (sweep_seconds) *sweep_seconds = h->ds.sweep_seconds;
    if (sweeps) *sweeps = h->ds.sweeps;
    return ALLIUM_OK;
}

// ---- host-side setup queries ------------------------------------------------
}  // extern "C"

int allium_internal_quad(const allium_quad_desc *quad, int dim, allium::Quadrature &q)
{
    std::string why;
    if (quad->nomega > 0) {
        q.dim = dim;
        for (int a = 0; a < quad->nomega; a++) {
            for (int d = 0; d < 3; d++) q.omega.push_back(d < dim ? quad->omega[a * dim + d] : 0.0);
            q.w.push_back(quad->weights[a]);
            q.wsum += quad->weights[a];
        }
        return ALLIUM_OK;
    }
    if (!allium::level_symmetric(quad->order, dim, q, why)) return fail(ALLIUM_ERR_ARG, why);
    return ALLIUM_OK;
}

extern "C" {

int allium_host_quadrature(int order, int dim, double *omega, double *weights, int *nomega)
{
    if (!nomega) return fail(ALLIUM_ERR_ARG, "null argument");
    allium::Quadrature q;
    std::string why;
    if (!allium::level_symmetric(order, dim, q, why)) return fail(ALLIUM_ERR_ARG, why);
    if (*nomega < q.size()) { *nomega = q.size(); return fail(ALLIUM_ERR_ARG, "capacity too small"); }
    *nomega = q.size();
    for (int a = 0; a < q.size(); a++) {
        for (int d = 0; d < dim; d++) omega[a * dim + d] = q.omega[3 * a + d];
        weights[a] = q.w[a];
    }
    return ALLIUM_OK;
}

int allium_host_levels(const allium_mesh_desc *mesh, const allium_quad_desc *quad, int *level_of_node, long *nlevels)
{
    if (!mesh || !quad || !level_of_node) return fail(ALLIUM_ERR_ARG, "null argument");
    allium::Mesh m;
    allium::Quadrature q;
    allium::SweepPlan plan;
    std::string why;
    if (!allium::build_mesh(mesh->dim, mesh->nel, mesh->vertices, mesh->extents, mesh->bc, m, why))
        return fail(ALLIUM_ERR_ARG, why);
    int rc = allium_internal_quad(quad, mesh->dim, q);
    if (rc) return rc;
    int st = allium::build_plan(m, q, plan, why);
    if (st == -2) return fail(ALLIUM_ERR_NOT_DAG, why);
    if (st == -3) return fail(ALLIUM_ERR_REFLECT, why);
    if (st) return fail(ALLIUM_ERR_ARG, why);
    std::memcpy(level_of_node, plan.level.data(), sizeof(int) * plan.level.size());
    if (nlevels) *nlevels = plan.nlevels;
    return ALLIUM_OK;
}

int allium_host_form_source(const allium_mesh_desc *mesh, int fe_order, const allium_quad_desc *quad, int G,
                            const double *Q, const double *inflow, double *source)
{
    if (!mesh || !quad || !source || G < 1) return fail(ALLIUM_ERR_ARG, "bad argument");
    allium::Mesh m;
    allium::Quadrature q;
    allium::ElementData ed;
    std::string why;
    if (!allium::build_mesh(mesh->dim, mesh->nel, mesh->vertices, mesh->extents, mesh->bc, m, why))
        return fail(ALLIUM_ERR_ARG, why);
    int rc = allium_internal_quad(quad, mesh->dim, q);
    if (rc) return rc;
    if (!allium::assemble(m, fe_order, 0, ed, why)) return fail(ALLIUM_ERR_ARG, why);
    allium::form_source(m, ed, q, G, Q, inflow, source);
    return ALLIUM_OK;
}

int allium_host_element_matrices(const allium_mesh_desc *mesh, int fe_order, int e, double *mass, double *grad,
                                 double *tmass)
{
    if (!mesh) return fail(ALLIUM_ERR_ARG, "null argument");
    allium::Mesh m;
    allium::ElementData ed;
    std::string why;
    if (!allium::build_mesh(mesh->dim, mesh->nel, mesh->vertices, mesh->extents, mesh->bc, m, why))
        return fail(ALLIUM_ERR_ARG, why);
    if (!allium::assemble(m, fe_order, 0, ed, why)) return fail(ALLIUM_ERR_ARG, why);
    if (e < 0 || e >= m.Ne) return fail(ALLIUM_ERR_ARG, "element out of range");
    const size_t nn = (size_t)ed.n * ed.n, k = m.uniform ? 0 : (size_t)e;
    if (mass) std::memcpy(mass, &ed.Mu[k * nn], sizeof(double) * nn);
    if (tmass) std::memcpy(tmass, &ed.Mm[k * nn], sizeof(double) * nn);
    if (grad) std::memcpy(grad, &ed.Gd[k * m.dim * nn], sizeof(double) * m.dim * nn);
    return ALLIUM_OK;
}

// ---- device plumbing (include/allium_device.h) ------------------------------
int allium_device_count(int *count)
{
    HIP_TRY(hipGetDeviceCount(count));
    return ALLIUM_OK;
}

int allium_device_name(char *buf, int len)
{
    hipDeviceProp_t prop;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    std::snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return ALLIUM_OK;
}

int allium_device_malloc(void **ptr, size_t bytes)
{
    if (!ptr) return fail(ALLIUM_ERR_ARG, "null argument");
    HIP_TRY(hipMalloc(ptr, bytes ? bytes : 1));
    return ALLIUM_OK;
}

int allium_device_free(void *ptr)
{
    if (ptr) HIP_TRY(hipFree(ptr));
    return ALLIUM_OK;
}

int allium_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return ALLIUM_OK;
}

int allium_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return ALLIUM_OK;
}

int allium_memcpy_d2d(void *dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return ALLIUM_OK;
}

int allium_device_synchronize(void)
{
    HIP_TRY(hipDeviceSynchronize());
    return ALLIUM_OK;
}

int allium_event_timer(int op, double *ms)
{
    // op 0: record start on the null stream, op 1: record stop and return ms
    static hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!e0) { HIP_TRY(hipEventCreate(&e0)); HIP_TRY(hipEventCreate(&e1)); }
    if (op == 0) { HIP_TRY(hipEventRecord(e0, nullptr)); return ALLIUM_OK; }
    HIP_TRY(hipEventRecord(e1, nullptr));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    return ALLIUM_OK;
}

}  // extern "C"
