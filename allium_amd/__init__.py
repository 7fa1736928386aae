"""allium_amd -- MI355X-native DG upwind transport sweep and source iteration.

Python mirror of the reference's operator interface for the sweep path:

* ``InverseAdvectionOperator``  (src/sweep.hpp:13-130)   ``Mult(source, psi)``
* ``AdvectionOperator``         (src/sweep.hpp:158-167)  ``Mult``, ``ComputeBalance``
* ``DiscreteToMoment``          (src/transport_op.hpp:99-123)
* ``source_iteration``          (SLISolver on TransportOperator, src/linalg.cpp:103-153)
* ``krylov_solve``              (chive's "gmres" / "fgmres" outer solvers, src/io.cpp:179-230)

Every call goes through the C ABI of ``lib/liballium_amd.so`` (HIP kernels for
gfx950).  Host numpy arrays and ``DeviceVector`` (HBM) are both accepted.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import AlliumError, check

FP64 = 0
FP32_REFINED = 1
BC_INFLOW = 0
BC_REFLECTIVE = 1


def _d(a):
    return a.ctypes.data_as(_lib.c_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class DeviceVector:
    """A float64 vector in device memory (HBM)."""

    def __init__(self, n, host=None):
        lib = _lib.load()
        self.n = int(n)
        p = ctypes.c_void_p()
        check(lib.allium_device_malloc(ctypes.byref(p), 8 * max(1, self.n)))
        self.ptr = p
        if host is not None:
            self.upload(host)

    @classmethod
    def from_host(cls, a):
        a = _f64(a).reshape(-1)
        return cls(a.size, a)

    def upload(self, a):
        a = _f64(a).reshape(-1)
        if a.size != self.n:
            raise ValueError("size mismatch")
        check(_lib.load().allium_memcpy_h2d(self.ptr, a.ctypes.data_as(ctypes.c_void_p), 8 * self.n))

    def to_host(self, out=None):
        out = np.empty(self.n) if out is None else out
        check(_lib.load().allium_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), self.ptr, 8 * self.n))
        return out

    def copy_from(self, other):
        check(_lib.load().allium_memcpy_d2d(self.ptr, other.ptr, 8 * self.n))

    def free(self):
        if self.ptr is not None and self.ptr.value:
            _lib.load().allium_device_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(_lib.load().allium_device_synchronize())


def device_name():
    buf = ctypes.create_string_buffer(256)
    check(_lib.load().allium_device_name(buf, 256))
    return buf.value.decode()


class Mesh:
    """Structured quad/hex mesh: elements per axis, optional vertex array
    ((n0+1)(n1+1)[(n2+1)] x dim, x fastest), per-side boundary conditions."""

    def __init__(self, nel, vertices=None, extents=None, bc=None):
        self.nel = tuple(int(v) for v in nel)
        self.dim = len(self.nel)
        self.vertices = None if vertices is None else _f64(vertices).reshape(-1, self.dim)
        self.extents = tuple(extents) if extents is not None else (1.0,) * self.dim
        self.bc = tuple(bc) if bc is not None else (BC_INFLOW,) * (2 * self.dim)
        self.Ne = int(np.prod(self.nel))

    def desc(self):
        m = _lib.MeshDesc()
        m.dim = self.dim
        for d in range(3):
            m.nel[d] = self.nel[d] if d < self.dim else 1
            m.extents[d] = self.extents[d] if d < self.dim else 1.0
        for s in range(6):
            m.bc[s] = self.bc[s] if s < 2 * self.dim else 0
        m.vertices = _d(self.vertices) if self.vertices is not None else None
        return m


def _quad_desc(quad, dim):
    q = _lib.QuadDesc()
    keep = None
    if isinstance(quad, int):
        q.order, q.nomega = quad, 0
    else:
        om, w = _f64(quad[0]).reshape(-1, dim), _f64(quad[1]).reshape(-1)
        keep = (om, w)
        q.order, q.nomega, q.omega, q.weights = 0, w.size, _d(om), _d(w)
    return q, keep


def level_symmetric(order, dim):
    """Directions (N x dim) and weights of the library's level-symmetric S_order set."""
    lib = _lib.load()
    cap = ctypes.c_int(1024)
    om, w = np.zeros(1024 * dim), np.zeros(1024)
    check(lib.allium_host_quadrature(order, dim, _d(om), _d(w), ctypes.byref(cap)))
    n = cap.value
    return om[: n * dim].reshape(n, dim).copy(), w[:n].copy()


def host_levels(mesh, quad):
    """Frontier index of every (element, angle) node, shape (Ne, Nomega)."""
    lib = _lib.load()
    qd, keep = _quad_desc(quad, mesh.dim)
    nom = qd.nomega if qd.nomega else level_symmetric(quad, mesh.dim)[1].size
    out = np.zeros(mesh.Ne * nom, dtype=np.int32)
    nl = ctypes.c_long()
    check(lib.allium_host_levels(ctypes.byref(mesh.desc()), ctypes.byref(qd), out.ctypes.data_as(_lib.c_ip),
                                 ctypes.byref(nl)))
    return out.reshape(mesh.Ne, nom), nl.value


def host_form_source(mesh, fe_order, quad, G, Q=None, inflow=None):
    lib = _lib.load()
    qd, keep = _quad_desc(quad, mesh.dim)
    nom = qd.nomega if qd.nomega else level_symmetric(quad, mesh.dim)[1].size
    n = (fe_order + 1) ** mesh.dim
    out = np.zeros(G * nom * mesh.Ne * n)
    Qa = None if Q is None else _f64(np.broadcast_to(Q, (G, mesh.Ne)))
    Ia = None if inflow is None else _f64(np.broadcast_to(inflow, (G, 2 * mesh.dim)))
    check(lib.allium_host_form_source(ctypes.byref(mesh.desc()), fe_order, ctypes.byref(qd), G,
                                      None if Qa is None else _d(Qa), None if Ia is None else _d(Ia), _d(out)))
    return out


def host_element_matrices(mesh, fe_order, e):
    """(unit mass, gradient blocks (dim, n, n), MassIntegrator-rule mass) of element e."""
    lib = _lib.load()
    n = (fe_order + 1) ** mesh.dim
    m, g, t = np.zeros(n * n), np.zeros(mesh.dim * n * n), np.zeros(n * n)
    check(lib.allium_host_element_matrices(ctypes.byref(mesh.desc()), fe_order, e, _d(m), _d(g), _d(t)))
    return m.reshape(n, n), g.reshape(mesh.dim, n, n), t.reshape(n, n)


class InverseAdvectionOperator:
    """psi = L^{-1} source on the GPU (InverseAdvectionOperator, src/sweep.hpp:13-130).

    ``quad`` is an S_N order (level-symmetric set generated by the library) or
    an ``(omega, weights)`` pair.  ``sigma_t`` is (G, Ne) element-constant."""

    def __init__(self, mesh, fe_order, quad, sigma_t, lump=0, precision=FP64):
        lib = _lib.load()
        self.mesh = mesh
        sig = _f64(sigma_t)
        self.G = sig.shape[0] if sig.ndim == 2 else 1
        sig = sig.reshape(-1)
        if sig.size != self.G * mesh.Ne:
            raise ValueError("sigma_t must be (G, Ne)")
        qd, self._qkeep = _quad_desc(quad, mesh.dim)
        md = mesh.desc()
        h = ctypes.c_void_p()
        check(lib.allium_sweep_create(ctypes.byref(md), fe_order, ctypes.byref(qd), self.G, _d(sig), lump,
                                      precision, ctypes.byref(h)))
        self._h = h
        info = _lib.SweepInfo()
        check(lib.allium_sweep_info(h, ctypes.byref(info)))
        self.info = {k: getattr(info, k) for k, _ in _lib.SweepInfo._fields_}
        self.Nomega = info.Nomega
        self.n = info.ndof_per_elem
        self.Ndof = info.Ndof
        self.size = info.psi_size
        self.phi_size = info.phi_size
        self.nlevels = info.levels
        self.fe_order = fe_order
        self._lib = lib

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.allium_sweep_destroy(h)
            self._h = None

    # mfem::Operator shape
    def Height(self):
        return self.size

    def Width(self):
        return self.size

    def quadrature(self):
        dim = self.mesh.dim
        om, w = np.zeros(self.Nomega * dim), np.zeros(self.Nomega)
        check(self._lib.allium_sweep_quadrature(self._h, _d(om), _d(w)))
        return om.reshape(-1, dim), w

    def Mult(self, source, psi):
        """psi = L^{-1} source; host arrays (may be the same array) or DeviceVectors."""
        if isinstance(source, DeviceVector):
            check(self._lib.allium_sweep_mult_device(self._h, source.ptr, psi.ptr, None))
            return psi
        if source.dtype != np.float64 or not source.flags.c_contiguous or source.size != self.size:
            raise ValueError("source must be a contiguous float64 array of Height() entries")
        if psi.dtype != np.float64 or not psi.flags.c_contiguous or psi.size != self.size:
            raise ValueError("psi must be a contiguous float64 array of Height() entries")
        check(self._lib.allium_sweep_mult(self._h, _d(source), _d(psi)))
        return psi

    FIXUPS = {"none": 0, "zero": 1, "zero and scale": 2, "ryosuke": 3}

    def SetFixupOperator(self, kind, psi_min=0.0):
        """Negative-flux fixup after every cell solve (src/fixup.cpp): "zero",
        "zero and scale" or "ryosuke" (the deck's fixup type), "none" to turn
        it off; resets the counts."""
        check(self._lib.allium_sweep_set_fixup(self._h, self.FIXUPS[kind], float(psi_min)))

    def UseFixup(self, use=True):
        check(self._lib.allium_sweep_use_fixup(self._h, int(bool(use))))

    def IsFixupOn(self):
        t, m, on = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
        check(self._lib.allium_sweep_get_fixup(self._h, ctypes.byref(t), ctypes.byref(m), ctypes.byref(on)))
        return bool(on.value)

    def SetFixupMonitorData(self, monitor):
        """monitor: a DeviceVector of G x Ne counts (None stops monitoring)."""
        self._fix_monitor = monitor
        check(self._lib.allium_sweep_set_fixup_monitor(self._h, None if monitor is None else monitor.ptr))

    def fixup_counts(self):
        """(applied, failed): EventLog "fixup applied" / "fixup failed"."""
        a, f = ctypes.c_longlong(), ctypes.c_longlong()
        check(self._lib.allium_sweep_fixup_counts(self._h, ctypes.byref(a), ctypes.byref(f)))
        return a.value, f.value

    def MultMoments(self, source, psi, phi, iso=None):
        """psi = L^{-1}(source + iso) and phi = D psi in one call (DeviceVectors;
        source or iso may be None): the sweep-and-moment pair of
        TransportOperator::Mult."""
        check(self._lib.allium_sweep_mult_moments_device(self._h, None if source is None else source.ptr,
                                                         None if iso is None else iso.ptr, psi.ptr, phi.ptr, None))
        return psi, phi

    def AssembleLocalMatrices(self, sigma_t):
        check(self._lib.allium_sweep_set_sigma(self._h, _d(_f64(sigma_t).reshape(-1))))

    def SetTimeAbsorption(self, tau):
        check(self._lib.allium_sweep_set_time_absorption(self._h, float(tau)))

    def GetTimeAbsorption(self):
        tau = ctypes.c_double()
        check(self._lib.allium_sweep_get_time_absorption(self._h, ctypes.byref(tau)))
        return tau.value

    def GetLumpingType(self):
        lump, pbj = ctypes.c_int(), ctypes.c_int()
        check(self._lib.allium_sweep_get_lumping(self._h, ctypes.byref(lump), ctypes.byref(pbj)))
        return lump.value

    def IsParallelBlockJacobi(self):
        lump, pbj = ctypes.c_int(), ctypes.c_int()
        check(self._lib.allium_sweep_get_lumping(self._h, ctypes.byref(lump), ctypes.byref(pbj)))
        return bool(pbj.value)

    def set_precision(self, precision):
        check(self._lib.allium_sweep_set_precision(self._h, precision))

    def set_launch(self, mode):
        """0 per-level launches, 1 hipGraph, 2 persistent kernel (disabled),
        3 direction sweeps, 4 box sweeps, 5 box sweeps in a hipGraph, -1 autotune."""
        check(self._lib.allium_sweep_set_launch(self._h, mode))
        info = _lib.SweepInfo()
        check(self._lib.allium_sweep_info(self._h, ctypes.byref(info)))
        return info.launch_mode

    def set_tuning(self, mode, block, lanes):
        """Explicit launch mode, workgroup size and lanes per cell."""
        check(self._lib.allium_sweep_set_tuning(self._h, mode, block, lanes))

    def tuning(self):
        m, b, l = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(self._lib.allium_sweep_get_tuning(self._h, ctypes.byref(m), ctypes.byref(b), ctypes.byref(l)))
        return {"mode": m.value, "block": b.value, "lanes": l.value}

    def set_box_tiling(self, bx, by, bz=1, dirs_per_group=5, threads=256, prefetch=2):
        """Box extents, directions of one octant per workgroup, workgroup size and
        input prefetch (0/1/2) of box sweeps."""
        check(self._lib.allium_sweep_set_box_tiling(self._h, bx, by, bz, dirs_per_group, threads, prefetch))

    def box_tiling(self):
        v = [ctypes.c_int() for _ in range(6)]
        check(self._lib.allium_sweep_get_box_tiling(self._h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("bx", "by", "bz", "dirs_per_group", "threads", "prefetch"), (x.value for x in v)))

    def levels(self):
        out = np.zeros(self.mesh.Ne * self.Nomega, dtype=np.int32)
        check(self._lib.allium_sweep_levels(self._h, out.ctypes.data_as(_lib.c_ip)))
        return out.reshape(self.mesh.Ne, self.Nomega)

    def form_source(self, Q=None, inflow=None):
        out = np.empty(self.size)
        Qa = None if Q is None else _f64(np.broadcast_to(Q, (self.G, self.mesh.Ne)))
        Ia = None if inflow is None else _f64(np.broadcast_to(inflow, (self.G, 2 * self.mesh.dim)))
        check(self._lib.allium_form_source(self._h, None if Qa is None else _d(Qa),
                                           None if Ia is None else _d(Ia), _d(out)))
        return out

    def timing(self):
        """(seconds, sweeps): the "sweep" TimingLog and "sweeps" EventLog entries."""
        t, n = ctypes.c_double(), ctypes.c_long()
        check(self._lib.allium_sweep_timing(self._h, ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value


class SweepBatch:
    """Independent instances swept together, one launch per level of all of
    them (3D p = 2 handles): psi_i = L_i^{-1}(source_i + iso_i)."""

    def __init__(self, operators):
        self._lib = _lib.load()
        self.ops = list(operators)
        arr = (ctypes.c_void_p * len(self.ops))(*[op._h.value for op in self.ops])
        h = ctypes.c_void_p()
        check(self._lib.allium_sweep_batch_create(arr, len(self.ops), ctypes.byref(h)))
        self._h = h

    def Mult(self, sources, psis, isos=None):
        """Device vectors: sources (entries may be None), psis, optional isos."""
        n = len(self.ops)
        ptr = lambda v: None if v is None else v.ptr.value
        src = (ctypes.c_void_p * n)(*[ptr(v) for v in sources]) if sources is not None else None
        iso = (ctypes.c_void_p * n)(*[ptr(v) for v in isos]) if isos is not None else None
        out = (ctypes.c_void_p * n)(*[ptr(v) for v in psis])
        check(self._lib.allium_sweep_batch_mult_device(self._h, src, iso, out, None))
        return psis

    def timing(self):
        t, n = ctypes.c_double(), ctypes.c_long()
        check(self._lib.allium_sweep_batch_timing(self._h, ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value

    def __del__(self):
        try:
            if self._h:
                self._lib.allium_sweep_batch_destroy(self._h)
        except Exception:
            pass


class AdvectionOperator:
    """L psi and particle balance from the sweep's matrices (src/sweep.cpp:1246-1462)."""

    def __init__(self, Linv):
        self.Linv = Linv

    def Mult(self, x, y):
        L = self.Linv
        if isinstance(x, DeviceVector):
            check(L._lib.allium_apply_L_device(L._h, x.ptr, y.ptr, None))
            return y
        check(L._lib.allium_apply_L(L._h, _d(_f64(x)), _d(y)))
        return y

    def ComputeBalance(self, psi):
        L = self.Linv
        out = ctypes.c_double()
        if isinstance(psi, DeviceVector):
            check(L._lib.allium_balance_device(L._h, psi.ptr, ctypes.byref(out), None))
        else:
            check(L._lib.allium_balance(L._h, _d(_f64(psi)), ctypes.byref(out)))
        return out.value


class DiscreteToMoment:
    """D and D^T on device vectors (src/transport_op.cpp:3-89): the scalar flux
    (num_moments 1) or [phi, J] (num_moments dim + 1, moment index slowest)."""

    def __init__(self, Linv, num_moments=1):
        if num_moments not in (1, Linv.mesh.dim + 1):
            raise ValueError("num_moments must be 1 or dim + 1")
        self.Linv = Linv
        self.num_moments = num_moments

    def Mult(self, psi, phi):
        fn = self.Linv._lib.allium_moments_device if self.num_moments == 1 else \
            self.Linv._lib.allium_moments_aniso_device
        check(fn(self.Linv._h, psi.ptr, phi.ptr, None))
        return phi

    def MultTranspose(self, phi, psi):
        fn = self.Linv._lib.allium_moments_transpose_device if self.num_moments == 1 else \
            self.Linv._lib.allium_moments_aniso_transpose_device
        check(fn(self.Linv._h, phi.ptr, psi.ptr, None))
        return psi


# "mip": the scalar MIP operator (chive's "scalar mip"); "block mip": chive's "mip"
# (block interior-penalty system, Schur complement solved); "p1sa": chive's "p1sa";
# "ldgsa": chive's "ldgsa" (block LDG system)
DSA_METHODS = {"mip": 0, "p1sa": 1, "block mip": 2, "ldgsa": 3}


def _dsa_desc(kappa=-1.0, bound_below=True, alpha=0.0, sym=-1.0, cg_reltol=1e-6, cg_abstol=0.0, cg_maxit=100,
              smooth=2, method="mip", scale_penalty=True):
    return _lib.DSADesc(float(kappa), int(bool(bound_below)), float(alpha), float(sym), float(cg_reltol),
                        float(cg_abstol), int(cg_maxit), int(smooth), DSA_METHODS[method], int(bool(scale_penalty)))


def host_mip_matrix(mesh, fe_order, quad, sigma_t, sigma_s, **desc):
    """The scalar MIP diffusion matrix of DSA as CSR arrays (row_ptr, col, val), host only."""
    lib = _lib.load()
    qd, keep = _quad_desc(quad, mesh.dim)
    dd = _dsa_desc(**desc)
    st = _f64(np.broadcast_to(sigma_t, (mesh.Ne,)))
    ss = _f64(np.broadcast_to(sigma_s, (mesh.Ne,)))
    nnz = ctypes.c_longlong(0)
    md = mesh.desc()
    check(lib.allium_host_mip_matrix(ctypes.byref(md), fe_order, ctypes.byref(qd), _d(st), _d(ss), ctypes.byref(dd),
                                     ctypes.byref(nnz), None, None, None))
    n = (fe_order + 1) ** mesh.dim
    rp = np.zeros(mesh.Ne * n + 1, dtype=np.int64)
    col = np.zeros(nnz.value, dtype=np.int32)
    val = np.zeros(nnz.value)
    check(lib.allium_host_mip_matrix(ctypes.byref(md), fe_order, ctypes.byref(qd), _d(st), _d(ss), ctypes.byref(dd),
                                     ctypes.byref(nnz), rp.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                                     col.ctypes.data_as(_lib.c_ip), _d(val)))
    return rp, col, val


def host_block_ldg_matrix(mesh, fe_order, quad, sigma_t, sigma_s, **desc):
    """chive's "ldgsa" 2x2 block operator as CSR arrays in the layout of
    host_p1_matrix, host only (option: alpha)."""
    return host_block_mip_matrix(mesh, fe_order, quad, sigma_t, sigma_s, _entry="allium_host_block_ldg_matrix",
                                 **desc)


def host_block_mip_matrix(mesh, fe_order, quad, sigma_t, sigma_s, _entry="allium_host_block_mip_matrix", **desc):
    """chive's "mip" 2x2 block operator as CSR arrays (row_ptr, col, val) in the
    layout of host_p1_matrix, host only (options: kappa, bound_below, alpha,
    scale_penalty)."""
    lib = _lib.load()
    fn = getattr(lib, _entry)
    qd, keep = _quad_desc(quad, mesh.dim)
    dd = _dsa_desc(**desc)
    st = _f64(np.broadcast_to(sigma_t, (mesh.Ne,)))
    ss = _f64(np.broadcast_to(sigma_s, (mesh.Ne,)))
    nnz = ctypes.c_longlong(0)
    md = mesh.desc()
    check(fn(ctypes.byref(md), fe_order, ctypes.byref(qd), _d(st), _d(ss),
                                           ctypes.byref(dd), ctypes.byref(nnz), None, None, None))
    N = (mesh.dim + 1) * (fe_order + 1) ** mesh.dim
    rp = np.zeros(mesh.Ne * N + 1, dtype=np.int64)
    col = np.zeros(nnz.value, dtype=np.int32)
    val = np.zeros(nnz.value)
    check(fn(ctypes.byref(md), fe_order, ctypes.byref(qd), _d(st), _d(ss),
                                           ctypes.byref(dd), ctypes.byref(nnz),
                                           rp.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                                           col.ctypes.data_as(_lib.c_ip), _d(val)))
    return rp, col, val


def host_p1_matrix(mesh, fe_order, sigma_t, sigma_s):
    """chive's "p1sa" 2x2 block operator as CSR arrays (row_ptr, col, val) over
    the element-block layout [J_0, .., J_{dim-1}, phi] x n per element, host only."""
    lib = _lib.load()
    st = _f64(np.broadcast_to(sigma_t, (mesh.Ne,)))
    ss = _f64(np.broadcast_to(sigma_s, (mesh.Ne,)))
    nnz = ctypes.c_longlong(0)
    md = mesh.desc()
    check(lib.allium_host_p1_matrix(ctypes.byref(md), fe_order, _d(st), _d(ss), ctypes.byref(nnz), None, None, None))
    N = (mesh.dim + 1) * (fe_order + 1) ** mesh.dim
    rp = np.zeros(mesh.Ne * N + 1, dtype=np.int64)
    col = np.zeros(nnz.value, dtype=np.int32)
    val = np.zeros(nnz.value)
    check(lib.allium_host_p1_matrix(ctypes.byref(md), fe_order, _d(st), _d(ss), ctypes.byref(nnz),
                                    rp.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                                    col.ctypes.data_as(_lib.c_ip), _d(val)))
    return rp, col, val


class DSA:
    """Diffusion synthetic acceleration on the GPU
    (DiffusionSyntheticAccelerationOperator, src/transport_op.hpp:38-59) with
    chive's "scalar mip" operator (method "mip", exe/chive.cpp:663-682), its
    block "mip" one (method "block mip", exe/chive.cpp:635-662), its "p1sa"
    one (method "p1sa", exe/chive.cpp:608-619) or its "ldgsa" one (method
    "ldgsa", exe/chive.cpp:620-634).

    ``Mult(x, y)``: y = x + K^{-1} Ms x on device vectors; ``solve(b)``: K^{-1} b
    for host arrays (for "p1sa": the phi-part of P^{-1} [0; b]).  Options of
    "mip": kappa (< 0: |kappa| (p+1)^2), bound_below, alpha (<= 0: from the
    quadrature), sym, smooth (damped Jacobi sweeps per multigrid level and
    direction); cg_reltol, cg_abstol, cg_maxit are the CG ("mip") or GMRES
    ("p1sa") stopping rule."""

    def __init__(self, Linv, sigma_s, **desc):
        self.Linv = Linv
        self._lib = Linv._lib
        self._desc = _dsa_desc(**desc)
        h = ctypes.c_void_p()
        check(self._lib.allium_dsa_create(Linv._h, _d(_f64(np.broadcast_to(sigma_s, (Linv.mesh.Ne,)))),
                                          ctypes.byref(self._desc), ctypes.byref(h)))
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self._lib.allium_dsa_destroy(h)
            self.handle = None

    def Mult(self, x, y):
        check(self._lib.allium_dsa_mult_device(self.handle, x.ptr, y.ptr, None))
        return y

    def solve(self, b):
        x = np.empty(self.Linv.phi_size)
        check(self._lib.allium_dsa_solve(self.handle, _d(_f64(b)), _d(x)))
        return x

    def info(self):
        inf = _lib.DSAInfo()
        check(self._lib.allium_dsa_get_info(self.handle, ctypes.byref(inf)))
        nl = inf.levels
        return dict(levels=nl, level_size=list(inf.level_size[:nl]), omega=list(inf.omega[:nl]),
                    dense_coarsest=bool(inf.dense_coarsest), last_cg_iterations=inf.last_cg_iterations,
                    last_cg_norm=inf.last_cg_norm, applications=inf.applications, cg_iterations=inf.cg_iterations,
                    seconds=inf.seconds)


def source_iteration(Linv, sigma_s, source, abstol=1e-8, reltol=0.0, maxit=1000, precision=FP64,
                     want_psi=True, accel=None, history=None, prolong=1, patience=3, fallback=None,
                     raw_norms=None, inner_iterations=None):
    """chive's SI: returns (phi, psi, iterations, norms) (exe/chive.cpp:700-735).

    ``accel``: a DSA object (the SLISolver preconditioner: every iteration is
    corrected and the norm is that of the corrected change), or a
    cnn.FluxCorrector applied after every iteration (switched off once
    ``patience`` iterations have passed without a new lowest residual; the
    iterate of the lowest residual is restored and ``fallback``, a DSA object
    or None, takes over); ``prolong`` 1 spreads the
    network's element corrections continuously over the vertices (p = 1), 0
    keeps them element-constant.  ``history``: a (maxit, 2, Ne) float64 array
    that receives, for every iteration but the last, the element-averaged
    scalar flux after it and the element-averaged change it made before any
    correction.  ``raw_norms``: a float64 array of maxit that receives the
    unpreconditioned change of every iteration (before any DSA or learned
    correction); ``inner_iterations``: an int32 array of maxit that receives
    the DSA CG iterations of every iteration.  A DeviceVector ``source``
    (HBM-resident) runs the device entry point: phi comes back to the host,
    psi stays a DeviceVector."""
    lib = Linv._lib
    hist_p = None
    if history is not None:
        if history.dtype != np.float64 or not history.flags.c_contiguous or history.size < 2 * maxit * Linv.mesh.Ne:
            raise ValueError("history must be a contiguous float64 array of maxit x 2 x Ne")
        hist_p = _d(history)
    kind = 0 if accel is None else (2 if isinstance(accel, DSA) else 1)
    raw_p = None
    if raw_norms is not None:
        if raw_norms.dtype != np.float64 or not raw_norms.flags.c_contiguous or raw_norms.size < maxit:
            raise ValueError("raw_norms must be a contiguous float64 array of maxit")
        raw_p = _d(raw_norms)
    inner_p = None
    if inner_iterations is not None:
        if inner_iterations.dtype != np.int32 or not inner_iterations.flags.c_contiguous or \
                inner_iterations.size < maxit:
            raise ValueError("inner_iterations must be a contiguous int32 array of maxit")
        inner_p = inner_iterations.ctypes.data_as(_lib.c_ip)
    opt = _lib.SIOptions(abstol, reltol, int(maxit), int(precision), kind,
                         None if accel is None else accel.handle, None if fallback is None else fallback.handle,
                         int(prolong), int(patience), hist_p, raw_p, inner_p)
    norms = np.zeros(maxit + 1)
    it = ctypes.c_int(0)
    sig = _d(_f64(np.broadcast_to(sigma_s, (Linv.mesh.Ne,))))
    if isinstance(source, DeviceVector):
        dphi = DeviceVector(Linv.phi_size)
        psi = DeviceVector(Linv.size) if want_psi else None
        rc = lib.allium_source_iteration_device(Linv._h, sig, source.ptr, ctypes.byref(opt), dphi.ptr,
                                                None if psi is None else psi.ptr, _d(norms), ctypes.byref(it))
        if rc not in (0, -7):
            check(rc)
        phi = dphi.to_host()
    else:
        phi = np.empty(Linv.phi_size)
        psi = np.empty(Linv.size) if want_psi else None
        rc = lib.allium_source_iteration(Linv._h, sig, _d(_f64(source)), ctypes.byref(opt), _d(phi),
                                         None if psi is None else _d(psi), _d(norms), ctypes.byref(it))
        if rc not in (0, -7):
            check(rc)
    return phi, psi, it.value, norms[: it.value].copy()


SOLVERS = {"gmres": 1, "fgmres": 2}


def krylov_solve(Linv, sigma_s, source, solver="fgmres", kdim=50, abstol=1e-8, reltol=0.0, maxit=100,
                 want_psi=True, accel=None, fallback=None, prolong=1, patience=3, max_steps=0,
                 inner_iterations=None):
    """chive's "gmres" / "fgmres" outer solver on T = I - D L^{-1} D^T Ms
    (src/io.cpp:179-230, exe/chive.cpp:700-735): returns (phi, psi,
    iterations, norms) with norms[0] = beta0 and norms[j] the residual
    estimate after step j (mfem's monitor values).

    ``accel``: a DSA object (M = I + K^{-1} Ms; left preconditioner of GMRES,
    right of FGMRES) or, with FGMRES only, a cnn.FluxCorrector (M v = v +
    net(v)), replaced by ``fallback`` (a DSA object or None) once ``patience``
    consecutive steps each cut the residual by less than 10 %, or after
    ``max_steps`` steps (0: no limit); the number of steps that used the
    network is left in ``krylov_solve.last_accel_steps``.  A DeviceVector
    ``source`` (HBM-resident) runs the device entry point: phi comes back to
    the host, psi stays a DeviceVector.
    ``inner_iterations`` (int32, maxit) receives the DSA CG iterations of each
    step.  FP64 sweeps; zero initial guess."""
    lib = Linv._lib
    if solver not in SOLVERS:
        raise ValueError(f"solver must be one of {sorted(SOLVERS)}")
    kind = 0 if accel is None else (2 if isinstance(accel, DSA) else 1)
    inner_p = None
    if inner_iterations is not None:
        if inner_iterations.dtype != np.int32 or not inner_iterations.flags.c_contiguous or \
                inner_iterations.size < maxit:
            raise ValueError("inner_iterations must be a contiguous int32 array of maxit")
        inner_p = inner_iterations.ctypes.data_as(_lib.c_ip)
    steps = ctypes.c_int(0)
    opt = _lib.KrylovOptions()
    check(lib.allium_krylov_default_options(ctypes.byref(opt)))
    opt.solver, opt.kdim, opt.abstol, opt.reltol, opt.maxit = SOLVERS[solver], int(kdim), abstol, reltol, int(maxit)
    opt.accel = kind
    opt.accel_handle = None if accel is None else accel.handle
    opt.accel_fallback = None if fallback is None else fallback.handle
    opt.accel_prolong, opt.accel_patience = int(prolong), int(patience)
    opt.inner_iterations = inner_p
    opt.accel_steps = ctypes.cast(ctypes.byref(steps), _lib.c_ip)
    opt.accel_max_steps = int(max_steps)
    norms = np.zeros(maxit + 2)
    it = ctypes.c_int(0)
    sig = _d(_f64(np.broadcast_to(sigma_s, (Linv.mesh.Ne,))))
    if isinstance(source, DeviceVector):
        dphi = DeviceVector(Linv.phi_size)
        psi = DeviceVector(Linv.size) if want_psi else None
        rc = lib.allium_krylov_solve_device(Linv._h, sig, source.ptr, ctypes.byref(opt), dphi.ptr,
                                            None if psi is None else psi.ptr, _d(norms), ctypes.byref(it))
        if rc not in (0, -7):
            check(rc)
        phi = dphi.to_host()
    else:
        phi = np.empty(Linv.phi_size)
        psi = np.empty(Linv.size) if want_psi else None
        rc = lib.allium_krylov_solve(Linv._h, sig, _d(_f64(source)), ctypes.byref(opt), _d(phi),
                                     None if psi is None else _d(psi), _d(norms), ctypes.byref(it))
        if rc not in (0, -7):
            check(rc)
    krylov_solve.last_accel_steps = steps.value
    return phi, psi, it.value, norms[: it.value + 1].copy()
