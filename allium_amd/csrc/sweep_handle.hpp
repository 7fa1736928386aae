// The sweep handle behind allium_sweep_t, shared by the C-ABI sources.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/allium_sweep.h"
#include "device.hpp"
#include "discretization.hpp"

// device buffer that grows on demand and is freed with its owner
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { if (p) hipFree(p); }
    hipError_t ensure(size_t b)
    {
        if (bytes >= b) return hipSuccess;
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, b);
        if (e == hipSuccess) bytes = b;
        return e;
    }
    double *d() const { return (double *)p; }
};

struct allium_sweep {
    allium::Mesh mesh;
    allium::Quadrature quad;
    allium::ElementData ed;
    allium::SweepPlan plan;
    allium::DeviceSweep ds;
    int G = 1;
    long psi_size = 0, phi_size = 0;
    std::vector<double> sigma;
    double *d_sigma = nullptr;
    // matrices kept for host-side rebuilds
    std::vector<double> Kst;
    // scratch for host-pointer entry points and the source iteration
    DevBuf a, b, c, x, y, iso, coef, part, q, r, wavg, havg, xb;
    // Krylov solver: basis V (and Z for FGMRES), work vectors, partial sums, Hessenberg
    DevBuf kv, kz, kw, kt, kpart, kh;
    // negative-flux fixup counters: applied, failed, ryosuke beyond 3 positive unknowns
    DevBuf fixbuf;
    hipStream_t stream = nullptr;
    ~allium_sweep()
    {
        for (void *p : ds.owned) hipFree(p);
        if (ds.box_buf) hipFree(ds.box_buf);
        if (ds.graph_exec) hipGraphExecDestroy(ds.graph_exec);
        if (ds.graph_stream) hipStreamDestroy(ds.graph_stream);
        if (ds.ev0) hipEventDestroy(ds.ev0);
        if (ds.ev1) hipEventDestroy(ds.ev1);
        for (int i = 0; i < allium::DeviceSweep::kTimingRing; i++) {
            if (ds.ring0[i]) hipEventDestroy(ds.ring0[i]);
            if (ds.ring1[i]) hipEventDestroy(ds.ring1[i]);
        }
        if (stream) hipStreamDestroy(stream);
    }
};

namespace allium_sweep_internal {
// psi = L^{-1}(src + iso), iso phi-sized and added in every direction (src or iso may be
// NULL), with the handle's timing and error checks; phi (may be NULL) then receives D psi
int sweep(allium_sweep *h, const double *src, const double *iso, double *psi, hipStream_t stream,
          double *phi = nullptr);
}  // namespace allium_sweep_internal

// quadrature of a descriptor: explicit directions or the level-symmetric set
int allium_internal_quad(const allium_quad_desc *quad, int dim, allium::Quadrature &q);
