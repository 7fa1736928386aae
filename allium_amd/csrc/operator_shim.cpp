// A C entry over the C++ adapter include/allium_operator.hpp, used the way the
// reference's drivers use InverseAdvectionOperator and AdvectionOperator
// (exe/chive.cpp:709-735: Linv.Mult(source, psi) into a std/mfem vector, in
// place, then L psi and the balance).  Host-only code: built with g++ against
// liballium_amd.so, so the adapter's templates, exceptions and vector
// plumbing run on the GPU path.
#include <cstring>
#include <vector>

#include "allium_operator.hpp"

extern "C" {

// out_psi = L^{-1} source through Mult(const V&, V&); out_inplace = the same
// through Mult(psi, psi) on a copy of the source; out_Lpsi = ApplyL(out_psi);
// *balance = ComputeBalance(out_psi).  Returns 0 or the allium_status of the
// allium::Error thrown (its message stays in allium_last_error()).
int allium_operator_shim_run(const allium_mesh_desc *mesh, int fe_order, const allium_quad_desc *quad, int G,
                             const double *sigma_t, const double *source, long n, double *out_psi,
                             double *out_inplace, double *out_Lpsi, double *balance)
{
    try {
        allium::InverseAdvectionOperator Linv(*mesh, fe_order, *quad, G, sigma_t);
        if (Linv.Height() != n || Linv.Width() != n) return ALLIUM_ERR_ARG;
        std::vector<double> src(source, source + n), psi(n);
        Linv.Mult(src, psi);
        std::memcpy(out_psi, psi.data(), sizeof(double) * n);
        std::vector<double> buf(src);
        Linv.Mult(buf, buf);
        std::memcpy(out_inplace, buf.data(), sizeof(double) * n);
        std::vector<double> Lpsi(n);
        Linv.ApplyL(psi, Lpsi);
        std::memcpy(out_Lpsi, Lpsi.data(), sizeof(double) * n);
        *balance = Linv.ComputeBalance(psi);
        // a size mismatch is an allium::Error, not an abort
        std::vector<double> shorter(n - 1);
        try {
            Linv.Mult(src, shorter);
            return ALLIUM_ERR_ARG;
        } catch (const allium::Error &e) {
            if (e.code() != ALLIUM_ERR_ARG) return e.code();
        }
        return ALLIUM_OK;
    } catch (const allium::Error &e) {
        return e.code();
    }
}

}  // extern "C"
