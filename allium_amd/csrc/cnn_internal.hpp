// Hooks between the source iteration (sweep_api.cpp) and the learned flux
// correction (cnn.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

struct allium_cnn;

namespace allium_cnn_internal {

// phi += scale * net(ravg / scale, log sigma_t, sigma_s / sigma_t) on the
// element grid nz x ny x nx, where ravg is the w-weighted element average of
// the scalar-flux change r (n dofs per element) and scale its RMS.
// prolong 0: element-constant correction; 1 (p = 1): each node gets the mean
// of the corrections of the elements around its vertex (continuous)
int correct(allium_cnn *h, int nz, int ny, int nx, int n, const double *d_w, const double *d_r,
            const double *d_sigt, const double *d_sigs, double *d_phi, hipStream_t s, int prolong);
// out[e] = sum_i w_i x[e n + i]
hipError_t elem_avg(long long Ne, int n, const double *d_w, const double *d_x, double *d_out, hipStream_t s);
// record a failure for allium_last_error() (defined in sweep_api.cpp)
void set_last_error(const std::string &msg);

}  // namespace allium_cnn_internal
