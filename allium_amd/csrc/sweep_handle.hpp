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
    hipStream_t stream = nullptr;
    ~allium_sweep()
    {
        for (void *p : ds.owned) hipFree(p);
        if (ds.box_buf) hipFree(ds.box_buf);
        if (ds.graph_exec) hipGraphExecDestroy(ds.graph_exec);
        if (ds.graph_stream) hipStreamDestroy(ds.graph_stream);
        if (ds.ev0) hipEventDestroy(ds.ev0);
        if (ds.ev1) hipEventDestroy(ds.ev1);
        if (stream) hipStreamDestroy(stream);
    }
};

// quadrature of a descriptor: explicit directions or the level-symmetric set
int allium_internal_quad(const allium_quad_desc *quad, int dim, allium::Quadrature &q);
