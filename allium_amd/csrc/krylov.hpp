// Vector kernels of the GMRES / FGMRES outer solvers (krylov.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace allium {

// Givens rotation (cs, sn) that zeroes dy against dx, and its application
// (mfem's GeneratePlaneRotation / ApplyPlaneRotation, linalg/solvers.cpp)
inline void givens(double dx, double dy, double &cs, double &sn)
{
    if (dy == 0.0) {
        cs = 1.0;
        sn = 0.0;
    } else if (std::fabs(dy) > std::fabs(dx)) {
        const double t = dx / dy;
        sn = 1.0 / std::sqrt(1.0 + t * t);
        cs = t * sn;
    } else {
        const double t = dy / dx;
        cs = 1.0 / std::sqrt(1.0 + t * t);
        sn = t * cs;
    }
}

inline void apply_givens(double &dx, double &dy, double cs, double sn)
{
    const double t = cs * dx + sn * dy;
    dy = -sn * dx + cs * dy;
    dx = t;
}

// Reductions write one partial sum per block; the consumer of a partial array
// re-sums it in a fixed order, so results do not depend on scheduling.
int krylov_blocks(long n);

// Modified Gram-Schmidt step on w.  If vprev is given: h = sum(part_in) is
// stored to *h_out and w -= h vprev.  Then part_out[blk] = partial sum of
// w . vcur (vcur given) or of w . w (vcur NULL).  part_in and part_out must
// differ.
hipError_t launch_mgs(long n, const double *part_in, int nb, double *h_out, const double *vprev,
                      const double *vcur, double *w, double *part_out, hipStream_t s);
// h = sqrt(sum(part_in)) stored to *h_out; vout = (1 / h) w
hipError_t launch_normalize(long n, const double *part_in, int nb, double *h_out, const double *w, double *vout,
                            hipStream_t s);
// x += sum_{j<k} y[j] V_j with V_j = V + j * ld (y on the device, k <= 64)
hipError_t launch_basis_update(long n, int k, const double *V, long ld, const double *y, double *x, hipStream_t s);

}  // namespace allium
