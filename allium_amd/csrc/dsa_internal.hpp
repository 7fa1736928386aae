// Hook between the source iteration (sweep_api.cpp) and diffusion synthetic
// acceleration (dsa_api.cpp).
#pragma once

#include <hip/hip_runtime.h>

struct allium_dsa;
struct allium_sweep;

namespace allium_dsa_internal {

// SLISolver step with the DSA preconditioner: given the source-iteration
// change r (phi-sized, device), c = K^{-1} Ms r, x += c and *nrm = ||r + c||_2.
int correct(allium_dsa *d, const double *d_r, double *d_x, hipStream_t s, double *nrm);
// the sweep handle the DSA object was built on
const allium_sweep *handle_of(const allium_dsa *d);
// CG iterations of the last solve
int last_inner(const allium_dsa *d);

}  // namespace allium_dsa_internal
