"""Multigroup linearized backward-Euler TRT on the GPU (include/allium_trt.h).

Mirrors scallion's "linearized" solver: one LinearizedTRTOperator::Mult per
time step (src/trt_linearized.cpp:230-312) in the time loop of
exe/scallion.cpp:1354-1555, in dimensionless units (c = 1, black-body
emission T^4)."""
import ctypes

import numpy as np

from . import _lib
from ._lib import check


class TRTDesc(ctypes.Structure):
    _fields_ = [("groups", ctypes.c_int), ("emin", ctypes.c_double), ("emax", ctypes.c_double),
                ("opacity_coef", ctypes.c_double), ("cv", ctypes.c_double), ("dt", ctypes.c_double),
                ("T0", ctypes.c_double), ("T_inflow", ctypes.c_double), ("inflow_sides", ctypes.c_int),
                ("si_reltol", ctypes.c_double), ("si_abstol", ctypes.c_double), ("si_maxit", ctypes.c_int),
                ("si_warm_start", ctypes.c_int), ("outer_maxit", ctypes.c_int), ("outer_abstol", ctypes.c_double),
                ("outer_reltol", ctypes.c_double), ("implicit_opacity", ctypes.c_int)]


class TRTInfo(ctypes.Structure):
    _fields_ = [("step", ctypes.c_int), ("time", ctypes.c_double), ("transport_iterations", ctypes.c_int),
                ("transport_residual", ctypes.c_double), ("floored", ctypes.c_long), ("T_min", ctypes.c_double),
                ("T_max", ctypes.c_double), ("radiation_energy", ctypes.c_double),
                ("sweep_seconds", ctypes.c_double), ("sweeps", ctypes.c_long), ("outer_iterations", ctypes.c_int),
                ("outer_residual", ctypes.c_double)]


class TRTEnergy(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("absorption", "time_absorption", "leakage", "source", "inflow",
                                                "emission_linear", "absorption_si", "gray_linear", "material",
                                                "floor", "emission", "balance")]


def planck_integrals(x):
    """(normalized Planck, Rosseland) integrals over [0, x] from the library."""
    p, r = ctypes.c_double(), ctypes.c_double()
    check(_lib.load().allium_planck_integrals(float(x), ctypes.byref(p), ctypes.byref(r)))
    return p.value, r.value


def log_groups(emin, emax, G):
    b, m = np.zeros(G + 1), np.zeros(G)
    check(_lib.load().allium_log_groups(emin, emax, G, b.ctypes.data_as(_lib.c_dp), m.ctypes.data_as(_lib.c_dp)))
    return b, m


class TRT:
    """Linearized TRT solver on a uniform 3D hex mesh (p = 1)."""

    def __init__(self, nel, sn=6, groups=16, emin=1e-2, emax=1e1, opacity_coef=10.0, cv=1.0, dt=1e-3, T0=0.01,
                 T_inflow=1.0, inflow_sides=1, si_reltol=1e-8, si_abstol=0.0, si_maxit=1000, warm_start=False,
                 quad=None, outer_maxit=1, outer_abstol=0.0, outer_reltol=0.0, implicit_opacity=False):
        """outer_*: scallion's fixed-point iteration on T around the linearized
        step (one pass by default); implicit_opacity re-evaluates the opacities
        after every pass."""
        from . import Mesh, _quad_desc
        self._lib = _lib.load()
        self.mesh = Mesh(nel)
        q, self._keep = _quad_desc(sn if quad is None else quad, 3)
        d = TRTDesc(groups, emin, emax, opacity_coef, cv, dt, T0, T_inflow, inflow_sides, si_reltol, si_abstol,
                    si_maxit, int(warm_start), int(outer_maxit), float(outer_abstol), float(outer_reltol),
                    int(bool(implicit_opacity)))
        h = ctypes.c_void_p()
        m = self.mesh.desc()
        check(self._lib.allium_trt_create(ctypes.byref(m), 1, ctypes.byref(q), ctypes.addressof(d), ctypes.byref(h)))
        self._h = h
        self.G = groups
        self.dt = dt
        self.Ndof = self.mesh.Ne * 8

    def step(self):
        info = TRTInfo()
        check(self._lib.allium_trt_step(self._h, ctypes.addressof(info)))
        return {k: getattr(info, k) for k, _ in TRTInfo._fields_}

    def track_energy(self, on=True):
        """Energy bookkeeping of every step (allium_trt_track_energy)."""
        check(self._lib.allium_trt_track_energy(self._h, int(bool(on))))

    def energy_terms(self):
        """The last step's energy terms (allium_trt_energy), as a dict."""
        e = TRTEnergy()
        check(self._lib.allium_trt_energy_terms(self._h, ctypes.byref(e)))
        return {k: getattr(e, k) for k, _ in TRTEnergy._fields_}

    def temperature(self):
        T = np.empty(self.Ndof)
        check(self._lib.allium_trt_temperature(self._h, T.ctypes.data_as(_lib.c_dp)))
        return T

    def scalar_flux(self):
        phi = np.empty(self.G * self.Ndof)
        check(self._lib.allium_trt_scalar_flux(self._h, phi.ctypes.data_as(_lib.c_dp)))
        return phi

    def state(self):
        """(x = [psi, T], step count, time): the state scallion restarts from."""
        n = ctypes.c_long()
        check(self._lib.allium_trt_state_size(self._h, ctypes.byref(n)))
        x = np.empty(n.value)
        cyc, t = ctypes.c_int(), ctypes.c_double()
        check(self._lib.allium_trt_get_state(self._h, x.ctypes.data_as(_lib.c_dp), ctypes.byref(cyc), ctypes.byref(t)))
        return x, cyc.value, t.value

    def set_state(self, x, cycle, time):
        x = np.ascontiguousarray(x, dtype=np.float64)
        check(self._lib.allium_trt_set_state(self._h, x.ctypes.data_as(_lib.c_dp), int(cycle), float(time)))

    def write_restart(self, root, output_cycle=0, keep=2):
        """RestartWriter::Write of [psi, T] with this solver's cycle, time and time step."""
        from .restart import write_restart
        x, cyc, t = self.state()
        write_restart(root, x, output_cycle=output_cycle, simulation_cycle=cyc, time=t, time_step=self.dt, keep=keep)

    def set_time_step(self, dt):
        """Time step of the following steps (allium_trt_set_time_step)."""
        check(self._lib.allium_trt_set_time_step(self._h, float(dt)))
        self.dt = float(dt)

    def load_restart(self, base, restart_id=0, use_restart_dt=True):
        """LoadFromRestart into this solver; returns the file's metadata.  With
        use_restart_dt (scallion's default) the solver continues with the file's
        time step, as scallion does (exe/scallion.cpp:616-674)."""
        from .restart import load_restart
        n = ctypes.c_long()
        check(self._lib.allium_trt_state_size(self._h, ctypes.byref(n)))
        r = load_restart(base, restart_id, size=n.value)
        self.set_state(r["x"], r["simulation_cycle"], r["time"])
        if use_restart_dt and r["time_step"] != self.dt:
            self.set_time_step(r["time_step"])
        return r

    def set_launch(self, mode):
        check(self._lib.allium_sweep_set_launch(self._lib.allium_trt_sweep(self._h), mode))

    def __del__(self):
        try:
            if self._h:
                self._lib.allium_trt_destroy(self._h)
        except Exception:
            pass
