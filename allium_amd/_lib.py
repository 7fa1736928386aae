"""ctypes declarations of liballium_amd.so (include/allium_*.h).

The library is built in-tree by ``allium_amd/build.py``.  There is no fallback:
if the shared library is missing every call raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liballium_amd.so")

c_dp = ctypes.POINTER(ctypes.c_double)
c_ip = ctypes.POINTER(ctypes.c_int)
c_lp = ctypes.POINTER(ctypes.c_long)
c_vp = ctypes.c_void_p


class MeshDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("nel", ctypes.c_int * 3), ("vertices", c_dp),
                ("extents", ctypes.c_double * 3), ("bc", ctypes.c_int * 6)]


class QuadDesc(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int), ("nomega", ctypes.c_int), ("omega", c_dp), ("weights", c_dp)]


class SweepInfo(ctypes.Structure):
    _fields_ = [("Ne", ctypes.c_long), ("ndof_per_elem", ctypes.c_long), ("Ndof", ctypes.c_long),
                ("Nomega", ctypes.c_long), ("G", ctypes.c_long), ("psi_size", ctypes.c_long),
                ("phi_size", ctypes.c_long), ("levels", ctypes.c_long), ("cells", ctypes.c_long),
                ("edges", ctypes.c_long), ("uniform_geometry", ctypes.c_int), ("launch_mode", ctypes.c_int)]


class SIOptions(ctypes.Structure):
    _fields_ = [("abstol", ctypes.c_double), ("reltol", ctypes.c_double), ("maxit", ctypes.c_int),
                ("precision", ctypes.c_int), ("accel", ctypes.c_int), ("accel_handle", c_vp),
                ("accel_fallback", c_vp), ("accel_prolong", ctypes.c_int), ("accel_patience", ctypes.c_int),
                ("history", c_dp), ("raw_norms", c_dp), ("inner_iterations", c_ip)]


class KrylovOptions(ctypes.Structure):
    _fields_ = [("solver", ctypes.c_int), ("kdim", ctypes.c_int), ("abstol", ctypes.c_double),
                ("reltol", ctypes.c_double), ("maxit", ctypes.c_int), ("accel", ctypes.c_int),
                ("accel_handle", c_vp), ("accel_fallback", c_vp), ("accel_prolong", ctypes.c_int),
                ("accel_patience", ctypes.c_int), ("inner_iterations", c_ip), ("accel_steps", c_ip),
                ("accel_max_steps", ctypes.c_int)]


class DSADesc(ctypes.Structure):
    _fields_ = [("kappa", ctypes.c_double), ("bound_below", ctypes.c_int), ("alpha", ctypes.c_double),
                ("sym", ctypes.c_double), ("cg_reltol", ctypes.c_double), ("cg_abstol", ctypes.c_double),
                ("cg_maxit", ctypes.c_int), ("smooth", ctypes.c_int), ("method", ctypes.c_int),
                ("scale_penalty", ctypes.c_int)]


class DSAInfo(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int), ("level_size", ctypes.c_longlong * 16), ("omega", ctypes.c_double * 16),
                ("dense_coarsest", ctypes.c_int), ("last_cg_iterations", ctypes.c_int),
                ("last_cg_norm", ctypes.c_double), ("applications", ctypes.c_long), ("cg_iterations", ctypes.c_long),
                ("seconds", ctypes.c_double)]


# name -> (restype, argtypes)
SIGNATURES = {
    # allium_sweep.h
    "allium_sweep_create": (ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.c_int, ctypes.POINTER(QuadDesc),
                                           ctypes.c_int, c_dp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(c_vp)]),
    "allium_sweep_destroy": (None, [c_vp]),
    "allium_sweep_info": (ctypes.c_int, [c_vp, ctypes.POINTER(SweepInfo)]),
    "allium_sweep_quadrature": (ctypes.c_int, [c_vp, c_dp, c_dp]),
    "allium_sweep_set_sigma": (ctypes.c_int, [c_vp, c_dp]),
    "allium_sweep_set_time_absorption": (ctypes.c_int, [c_vp, ctypes.c_double]),
    "allium_sweep_get_time_absorption": (ctypes.c_int, [c_vp, c_dp]),
    "allium_sweep_get_lumping": (ctypes.c_int, [c_vp, c_ip, c_ip]),
    "allium_sweep_set_precision": (ctypes.c_int, [c_vp, ctypes.c_int]),
    "allium_sweep_set_launch": (ctypes.c_int, [c_vp, ctypes.c_int]),
    "allium_sweep_set_tuning": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "allium_sweep_get_tuning": (ctypes.c_int, [c_vp, c_ip, c_ip, c_ip]),
    "allium_sweep_set_box_tiling": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                   ctypes.c_int, ctypes.c_int]),
    "allium_sweep_get_box_tiling": (ctypes.c_int, [c_vp, c_ip, c_ip, c_ip, c_ip, c_ip, c_ip]),
    "allium_sweep_mult": (ctypes.c_int, [c_vp, c_dp, c_dp]),
    "allium_sweep_mult_device": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "allium_sweep_mult_iso_device": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    "allium_sweep_levels": (ctypes.c_int, [c_vp, c_ip]),
    "allium_apply_L_device": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "allium_apply_L": (ctypes.c_int, [c_vp, c_dp, c_dp]),
    "allium_balance": (ctypes.c_int, [c_vp, c_dp, c_dp]),
    "allium_balance_device": (ctypes.c_int, [c_vp, c_vp, c_dp, c_vp]),
    "allium_balance_device_noalloc": (ctypes.c_int, [c_vp, c_vp, c_dp, c_vp]),
    "allium_form_source": (ctypes.c_int, [c_vp, c_dp, c_dp, c_dp]),
    "allium_moments_device": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "allium_moments_transpose_device": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "allium_moments_aniso_device": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "allium_moments_aniso_transpose_device": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "allium_mass_device": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    "allium_source_iteration": (ctypes.c_int, [c_vp, c_dp, c_dp, ctypes.POINTER(SIOptions), c_dp, c_dp, c_dp,
                                               c_ip]),
    "allium_sweep_set_fixup": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.c_double]),
    "allium_sweep_use_fixup": (ctypes.c_int, [c_vp, ctypes.c_int]),
    "allium_sweep_get_fixup": (ctypes.c_int, [c_vp, c_ip, c_dp, c_ip]),
    "allium_sweep_set_fixup_monitor": (ctypes.c_int, [c_vp, c_vp]),
    "allium_sweep_fixup_counts": (ctypes.c_int, [c_vp, ctypes.POINTER(ctypes.c_longlong),
                                                 ctypes.POINTER(ctypes.c_longlong)]),
    "allium_sweep_mult_moments_device": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "allium_krylov_default_options": (ctypes.c_int, [ctypes.POINTER(KrylovOptions)]),
    "allium_krylov_solve": (ctypes.c_int, [c_vp, c_dp, c_dp, ctypes.POINTER(KrylovOptions), c_dp, c_dp, c_dp,
                                           c_ip]),
    "allium_krylov_solve_device": (ctypes.c_int, [c_vp, c_dp, c_vp, ctypes.POINTER(KrylovOptions), c_vp, c_vp,
                                                  c_dp, c_ip]),
    "allium_source_iteration_device": (ctypes.c_int, [c_vp, c_dp, c_vp, ctypes.POINTER(SIOptions), c_vp, c_vp,
                                                      c_dp, c_ip]),
    "allium_sweep_timing": (ctypes.c_int, [c_vp, c_dp, c_lp]),
    "allium_sweep_batch_create": (ctypes.c_int, [ctypes.POINTER(c_vp), ctypes.c_int, ctypes.POINTER(c_vp)]),
    "allium_sweep_batch_destroy": (None, [c_vp]),
    "allium_sweep_batch_mult_device": (ctypes.c_int, [c_vp, ctypes.POINTER(c_vp), ctypes.POINTER(c_vp),
                                                      ctypes.POINTER(c_vp), c_vp]),
    "allium_sweep_batch_timing": (ctypes.c_int, [c_vp, c_dp, c_lp]),
    "allium_host_quadrature": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, c_dp, c_dp, c_ip]),
    "allium_host_levels": (ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.POINTER(QuadDesc), c_ip, c_lp]),
    "allium_host_form_source": (ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.c_int, ctypes.POINTER(QuadDesc),
                                               ctypes.c_int, c_dp, c_dp, c_dp]),
    "allium_host_element_matrices": (ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.c_int, ctypes.c_int, c_dp,
                                                    c_dp, c_dp]),
    "allium_last_error": (ctypes.c_char_p, []),
    # allium_cnn.h
    "allium_cnn_create": (ctypes.c_int, [ctypes.c_ulonglong, ctypes.POINTER(c_vp)]),
    "allium_cnn_destroy": (None, [c_vp]),
    "allium_cnn_num_params": (ctypes.c_long, []),
    "allium_cnn_set_dilation": (ctypes.c_int, [c_vp, c_ip]),
    "allium_cnn_get_dilation": (ctypes.c_int, [c_vp, c_ip]),
    "allium_cnn_get_params": (ctypes.c_int, [c_vp, ctypes.POINTER(ctypes.c_float)]),
    "allium_cnn_set_params": (ctypes.c_int, [c_vp, ctypes.POINTER(ctypes.c_float)]),
    "allium_cnn_forward_device": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_vp,
                                                 c_vp, c_vp]),
    "allium_cnn_gradient_device": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_vp,
                                                  c_vp, ctypes.POINTER(ctypes.c_float), c_dp]),
    "allium_cnn_train_step_device": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                    c_vp, c_vp, ctypes.c_float, c_dp]),
    # allium_dsa.h
    "allium_dsa_default_desc": (None, [ctypes.POINTER(DSADesc)]),
    "allium_dsa_create": (ctypes.c_int, [c_vp, c_dp, ctypes.POINTER(DSADesc), ctypes.POINTER(c_vp)]),
    "allium_dsa_destroy": (None, [c_vp]),
    "allium_dsa_mult_device": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "allium_dsa_solve_device": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "allium_dsa_solve": (ctypes.c_int, [c_vp, c_dp, c_dp]),
    "allium_dsa_get_info": (ctypes.c_int, [c_vp, ctypes.POINTER(DSAInfo)]),
    "allium_host_mip_matrix": (ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.c_int, ctypes.POINTER(QuadDesc), c_dp,
                                              c_dp, ctypes.POINTER(DSADesc), ctypes.POINTER(ctypes.c_longlong),
                                              ctypes.POINTER(ctypes.c_longlong), c_ip, c_dp]),
    "allium_host_block_mip_matrix": (ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.c_int, ctypes.POINTER(QuadDesc),
                                                    c_dp, c_dp, ctypes.POINTER(DSADesc),
                                                    ctypes.POINTER(ctypes.c_longlong),
                                                    ctypes.POINTER(ctypes.c_longlong), c_ip, c_dp]),
    "allium_host_block_ldg_matrix": (ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.c_int, ctypes.POINTER(QuadDesc),
                                                    c_dp, c_dp, ctypes.POINTER(DSADesc),
                                                    ctypes.POINTER(ctypes.c_longlong),
                                                    ctypes.POINTER(ctypes.c_longlong), c_ip, c_dp]),
    "allium_host_p1_matrix": (ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.c_int, c_dp, c_dp,
                                             ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                                             c_ip, c_dp]),
    # allium_trt.h (the descriptor/info structs are declared in allium_amd/trt.py)
    "allium_trt_create": (ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.c_int, ctypes.POINTER(QuadDesc), c_vp,
                                         ctypes.POINTER(c_vp)]),
    "allium_trt_destroy": (None, [c_vp]),
    "allium_trt_step": (ctypes.c_int, [c_vp, c_vp]),
    "allium_trt_temperature": (ctypes.c_int, [c_vp, c_dp]),
    "allium_trt_scalar_flux": (ctypes.c_int, [c_vp, c_dp]),
    "allium_trt_sweep": (c_vp, [c_vp]),
    "allium_trt_state_size": (ctypes.c_int, [c_vp, c_lp]),
    "allium_trt_get_state": (ctypes.c_int, [c_vp, c_dp, c_ip, c_dp]),
    "allium_trt_set_state": (ctypes.c_int, [c_vp, c_dp, ctypes.c_int, ctypes.c_double]),
    "allium_trt_track_energy": (ctypes.c_int, [c_vp, ctypes.c_int]),
    "allium_trt_energy_terms": (ctypes.c_int, [c_vp, c_vp]),
    "allium_trt_set_time_step": (ctypes.c_int, [c_vp, ctypes.c_double]),
    "allium_planck_integrals": (ctypes.c_int, [ctypes.c_double, c_dp, c_dp]),
    "allium_log_groups": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_int, c_dp, c_dp]),
    # allium_device.h
    "allium_device_count": (ctypes.c_int, [c_ip]),
    "allium_device_name": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "allium_device_malloc": (ctypes.c_int, [ctypes.POINTER(c_vp), ctypes.c_size_t]),
    "allium_device_free": (ctypes.c_int, [c_vp]),
    "allium_memcpy_h2d": (ctypes.c_int, [c_vp, c_vp, ctypes.c_size_t]),
    "allium_memcpy_d2h": (ctypes.c_int, [c_vp, c_vp, ctypes.c_size_t]),
    "allium_memcpy_d2d": (ctypes.c_int, [c_vp, c_vp, ctypes.c_size_t]),
    "allium_device_synchronize": (ctypes.c_int, []),
    "allium_event_timer": (ctypes.c_int, [ctypes.c_int, c_dp]),
}

_lib = None


class AlliumError(RuntimeError):
    """A non-zero allium_status, with allium_last_error()'s text."""

    def __init__(self, code, msg):
        super().__init__(f"allium error {code}: {msg}")
        self.code = code


def load():
    """Load liballium_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m allium_amd.build` "
                           "(the HIP library is required; there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        raise AlliumError(rc, load().allium_last_error().decode())
    return rc
