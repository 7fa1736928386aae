/*
 * allium_operator.hpp -- header-only C++ adapter over the C ABI in allium_sweep.h.
 *
 * Gives the sweep the shape of allium's InverseAdvectionOperator
 * (src/sweep.hpp:13-130): Height()/Width() = G * Nomega * Ndof and a const
 * Mult(source, psi) that may alias (src/transport_op.hpp:28).  Failures throw
 * allium::Error carrying the allium_status and allium_last_error()'s text
 * instead of MFEM_ABORT.
 *
 * Mult is templated on the vector type: anything with GetData() and Size()
 * (mfem::Vector) or data() and size() (std::vector<double>) works, so the
 * header does not depend on MFEM.
 */
#ifndef ALLIUM_OPERATOR_HPP
#define ALLIUM_OPERATOR_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "allium_sweep.h"

namespace allium {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc) {
    if (rc != ALLIUM_OK)
        throw Error(rc, std::string("allium error ") + std::to_string(rc) + ": " + allium_last_error());
}

namespace detail {
template <class V> auto data_of(V &v, int) -> decltype(v.GetData()) { return v.GetData(); }
template <class V> auto data_of(V &v, long) -> decltype(v.data()) { return v.data(); }
template <class V> auto size_of(const V &v, int) -> decltype(static_cast<long>(v.Size())) {
    return static_cast<long>(v.Size());
}
template <class V> auto size_of(const V &v, long) -> decltype(static_cast<long>(v.size())) {
    return static_cast<long>(v.size());
}
}  // namespace detail

/* ≙ InverseAdvectionOperator(fes, quad, total, bc_map, lump): the sweep
 * L^{-1} on a structured quad/hex mesh, with sigma_t[G][Ne]. */
class InverseAdvectionOperator {
public:
    InverseAdvectionOperator(const allium_mesh_desc &mesh, int fe_order, const allium_quad_desc &quad, int G,
                             const double *sigma_t, int lump = 0) {
        check(allium_sweep_create(&mesh, fe_order, &quad, G, sigma_t, lump, ALLIUM_FP64, &h_));
        check(allium_sweep_info(h_, &info_));
    }
    ~InverseAdvectionOperator() {
        if (h_) allium_sweep_destroy(h_);
    }
    InverseAdvectionOperator(const InverseAdvectionOperator &) = delete;
    InverseAdvectionOperator &operator=(const InverseAdvectionOperator &) = delete;
    InverseAdvectionOperator(InverseAdvectionOperator &&o) noexcept : h_(o.h_), info_(o.info_) { o.h_ = nullptr; }

    long Height() const { return info_.psi_size; }
    long Width() const { return info_.psi_size; }
    const allium_sweep_info_t &Info() const { return info_; }
    allium_sweep_t Handle() const { return h_; }

    /* psi = L^{-1} source (host vectors, RIGHT layout; source may alias psi) */
    template <class VIn, class VOut> void Mult(const VIn &source, VOut &psi) const {
        if (detail::size_of(source, 0) != Width() || detail::size_of(psi, 0) != Height())
            throw Error(ALLIUM_ERR_ARG, "InverseAdvectionOperator::Mult: vector size does not match Width()");
        check(allium_sweep_mult(h_, detail::data_of(const_cast<VIn &>(source), 0), detail::data_of(psi, 0)));
    }
    /* device pointers in HBM; stream is a hipStream_t (NULL = the library's stream) */
    void MultDevice(const double *d_source, double *d_psi, void *stream = nullptr) const {
        check(allium_sweep_mult_device(h_, d_source, d_psi, stream));
    }
    /* ≙ AssembleLocalMatrices() after the caller changed sigma_t[G][Ne] */
    void AssembleLocalMatrices(const double *sigma_t) { check(allium_sweep_set_sigma(h_, sigma_t)); }
    /* ≙ SetTimeAbsorption(tau) */
    void SetTimeAbsorption(double tau) {
        check(allium_sweep_set_time_absorption(h_, tau));
    }
    double GetTimeAbsorption() const {
        double tau = 0.0;
        check(allium_sweep_get_time_absorption(h_, &tau));
        return tau;
    }
    /* ≙ GetLumpingType (always 0) and IsParallelBlockJacobi (always false) */
    int GetLumpingType() const {
        int lump = 0, pbj = 0;
        check(allium_sweep_get_lumping(h_, &lump, &pbj));
        return lump;
    }
    bool IsParallelBlockJacobi() const {
        int lump = 0, pbj = 0;
        check(allium_sweep_get_lumping(h_, &lump, &pbj));
        return pbj != 0;
    }
    /* ≙ SetFixupOperator / UseFixup / IsFixupOn / SetFixupMonitorData (src/sweep.hpp:116-122):
     * type ALLIUM_FIXUP_ZERO, _ZERO_AND_SCALE or _RYOSUKE; the monitor is a device array of
     * G Ne doubles */
    void SetFixupOperator(int type, double psi_min = 0.0) { check(allium_sweep_set_fixup(h_, type, psi_min)); }
    void UseFixup(bool use) { check(allium_sweep_use_fixup(h_, use ? 1 : 0)); }
    bool IsFixupOn() const {
        int type = 0, on = 0;
        double mn = 0.0;
        check(allium_sweep_get_fixup(h_, &type, &mn, &on));
        return on != 0;
    }
    void SetFixupMonitorData(double *d_monitor) { check(allium_sweep_set_fixup_monitor(h_, d_monitor)); }
    void SetPrecision(int precision) { check(allium_sweep_set_precision(h_, precision)); }
    void SetLaunch(int mode) { check(allium_sweep_set_launch(h_, mode)); }

    /* ≙ AdvectionOperator::Mult (L psi) and ComputeBalance (src/sweep.cpp:1246-1462) */
    template <class VIn, class VOut> void ApplyL(const VIn &psi, VOut &Lpsi) const {
        check(allium_apply_L(h_, detail::data_of(const_cast<VIn &>(psi), 0), detail::data_of(Lpsi, 0)));
    }
    /* absorption + leakage, summed over angles and groups */
    template <class VIn> double ComputeBalance(const VIn &psi) const {
        double out = 0.0;
        check(allium_balance(h_, detail::data_of(const_cast<VIn &>(psi), 0), &out));
        return out;
    }

private:
    allium_sweep_t h_ = nullptr;
    allium_sweep_info_t info_{};
};

}  // namespace allium

#endif
