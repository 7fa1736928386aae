"""CPU FP64 oracle for allium's DG upwind sweep -- TEST INFRASTRUCTURE ONLY.

ctypes binding of ``liballium_oracle.so`` (``allium_oracle.c``), a plain-C
restatement of the reference's ``InverseAdvectionOperator`` /
``AdvectionOperator`` / ``DiscreteToMoment`` / ``SLISolver`` path
(src/sweep.cpp, src/transport_op.cpp, src/linalg.cpp:103-153).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module; the product package never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liballium_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


def build(verbose=False):
    """Compile the oracle with gcc (-O3 -march=native -fopenmp)."""
    cmd = ["gcc", "-O3", "-march=native", "-fopenmp", "-fPIC", "-shared", "-Wall",
           "-o", LIB_PATH, os.path.join(_HERE, "allium_oracle.c"), "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    return LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    lib.orc_create.restype = ctypes.c_void_p
    lib.orc_create.argtypes = [ctypes.c_int, _ip, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp,
                               ctypes.c_int, _dp, _ip, ctypes.c_int, _ip]
    lib.orc_destroy.argtypes = [ctypes.c_void_p]
    lib.orc_sizes.argtypes = [ctypes.c_void_p, _lp]
    lib.orc_levels.argtypes = [ctypes.c_void_p, _ip]
    lib.orc_levels.restype = ctypes.c_int
    lib.orc_set_sigma.argtypes = [ctypes.c_void_p, _dp]
    lib.orc_set_time_absorption.argtypes = [ctypes.c_void_p, ctypes.c_double]
    lib.orc_sweep.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
    lib.orc_sweep.restype = ctypes.c_long
    lib.orc_apply_L.argtypes = [ctypes.c_void_p, _dp, _dp]
    lib.orc_balance.argtypes = [ctypes.c_void_p, _dp]
    lib.orc_balance.restype = ctypes.c_double
    lib.orc_form_source.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    lib.orc_D.argtypes = [ctypes.c_void_p, _dp, _dp]
    lib.orc_DT.argtypes = [ctypes.c_void_p, _dp, _dp]
    lib.orc_mass.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    lib.orc_source_iteration.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, _dp, _dp, _dp,
                                         ctypes.c_int]
    lib.orc_source_iteration.restype = ctypes.c_int
    lib.orc_element_matrices.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
    lib.orc_face_blocks.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
    lib.orc_face_normal.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
    lib.orc_set_fixup.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _dp]
    lib.orc_fixup_counts.argtypes = [ctypes.c_void_p, _lp]
    lib.orc_gll_nodes.argtypes = [ctypes.c_int, _dp]
    lib.orc_gauss_rule.argtypes = [ctypes.c_int, _dp, _dp]
    _lib = lib
    return lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def gll_nodes(p):
    x = np.zeros(p + 1)
    _load().orc_gll_nodes(p, _d(x))
    return x


def gauss_rule(n):
    x, w = np.zeros(n), np.zeros(n)
    _load().orc_gauss_rule(n, _d(x), _d(w))
    return x, w


class Problem:
    """One sweep problem (mesh + FE order + quadrature + total cross sections)."""

    def __init__(self, dim, nel, vertices, p, omega, weights, sigma_t, bc=None, lump=0):
        lib = _load()
        self.dim, self.p = int(dim), int(p)
        self.nel = [int(v) for v in nel][:dim]
        self._vert = _c64(vertices).reshape(-1)
        self.omega = _c64(omega).reshape(-1, dim)
        self.weights = _c64(weights).reshape(-1)
        self.Nomega = self.omega.shape[0]
        sig = _c64(sigma_t)
        self.G = sig.shape[0] if sig.ndim == 2 else 1
        sig = sig.reshape(-1)
        nel_arr = (ctypes.c_int * 3)(*(self.nel + [1] * (3 - dim)))
        bc_arr = (ctypes.c_int * 6)(*((list(bc) if bc is not None else [0] * (2 * dim)) + [0] * 6)[:6])
        st = ctypes.c_int(0)
        h = lib.orc_create(dim, nel_arr, _d(self._vert), p, self.Nomega, _d(self.omega),
                           _d(self.weights), self.G, _d(sig), bc_arr, lump, ctypes.byref(st))
        if not h:
            raise RuntimeError(f"orc_create failed with status {st.value}"
                               + (" (graph is not a dag)" if st.value == -2 else ""))
        self._h = ctypes.c_void_p(h)
        s = (ctypes.c_long * 8)()
        lib.orc_sizes(self._h, s)
        (self.Ne, self.n, self.Ndof, self.nnodes, self.nedges, self.nlevels, self.nfaces,
         self.size) = [int(v) for v in s]
        self._lib = lib

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and self._lib is not None:
            self._lib.orc_destroy(h)
            self._h = None

    def levels(self):
        out = np.zeros(self.nnodes, dtype=np.int32)
        self._lib.orc_levels(self._h, out.ctypes.data_as(_ip))
        return out.reshape(self.Ne, self.Nomega)

    def set_sigma(self, sigma_t):
        self._lib.orc_set_sigma(self._h, _d(_c64(sigma_t).reshape(-1)))

    def set_time_absorption(self, tau):
        self._lib.orc_set_time_absorption(self._h, float(tau))

    FIXUPS = {"none": 0, "zero": 1, "zero and scale": 2, "ryosuke": 3}

    def set_fixup(self, kind, psi_min=0.0, monitor=None):
        """Negative-flux fixup after every local solve (src/fixup.cpp): "none",
        "zero", "zero and scale" or "ryosuke"; monitor: float64 array of G x Ne
        that counts the applications per element.  Resets the counts."""
        self._fix_monitor = monitor
        self._lib.orc_set_fixup(self._h, self.FIXUPS[kind], float(psi_min),
                                None if monitor is None else _d(monitor))

    def fixup_counts(self):
        """(applied, failed, ryosuke met more than 3 positive unknowns)"""
        out = (ctypes.c_long * 3)()
        self._lib.orc_fixup_counts(self._h, out)
        return int(out[0]), int(out[1]), bool(out[2])

    def sweep(self, source, psi=None, threads=0):
        """psi = L^{-1} source (Mult_Upwind). psi may be source (in place)."""
        src = source if (isinstance(source, np.ndarray) and source.dtype == np.float64
                         and source.flags.c_contiguous) else _c64(source)
        if psi is None:
            psi = np.empty(self.size)
        fails = self._lib.orc_sweep(self._h, _d(src), _d(psi), int(threads))
        if fails:
            raise RuntimeError(f"{fails} local solves hit the LinearSolve tolerance")
        return psi

    def apply_L(self, psi):
        y = np.empty(self.size)
        self._lib.orc_apply_L(self._h, _d(_c64(psi)), _d(y))
        return y

    def balance(self, psi):
        return self._lib.orc_balance(self._h, _d(_c64(psi)))

    def form_source(self, Q=None, inflow=None):
        q = np.empty(self.size)
        Qa = None if Q is None else _c64(np.broadcast_to(np.asarray(Q, float).reshape(self.G, -1) if np.ndim(Q) else Q,
                                                          (self.G, self.Ne)))
        Ia = None if inflow is None else _c64(np.broadcast_to(inflow, (self.G, 2 * self.dim)))
        self._lib.orc_form_source(self._h, None if Qa is None else _d(Qa),
                                  None if Ia is None else _d(Ia), _d(q))
        return q

    def D(self, psi):
        phi = np.empty(self.G * self.Ndof)
        self._lib.orc_D(self._h, _d(_c64(psi)), _d(phi))
        return phi

    def DT(self, phi):
        psi = np.empty(self.size)
        self._lib.orc_DT(self._h, _d(_c64(phi)), _d(psi))
        return psi

    def mass(self, coef, x):
        y = np.empty(self.G * self.Ndof)
        self._lib.orc_mass(self._h, _d(_c64(np.broadcast_to(coef, (self.Ne,)))), _d(_c64(x)), _d(y))
        return y

    def source_iteration(self, sigma_s, source, abstol=1e-8, reltol=0.0, maxit=1000, threads=0):
        phi = np.empty(self.G * self.Ndof)
        psi = np.empty(self.size)
        norms = np.zeros(maxit + 1)
        it = self._lib.orc_source_iteration(self._h, _d(_c64(np.broadcast_to(sigma_s, (self.Ne,)))),
                                            _d(_c64(source)), abstol, reltol, maxit, _d(phi),
                                            _d(psi), _d(norms), int(threads))
        return phi, psi, it, norms[:it].copy()

    def element_matrices(self, e):
        nn = self.n * self.n
        m, g, t = np.zeros(nn), np.zeros(self.dim * nn), np.zeros(nn)
        self._lib.orc_element_matrices(self._h, e, _d(m), _d(g), _d(t))
        f = lambda a: a.reshape(-1, self.n, self.n).transpose(0, 2, 1)  # column-major -> [i,j]
        return f(m)[0], f(g), f(t)[0]

    def face_blocks(self, e1, e2):
        b = np.zeros(4 * self.n * self.n)
        if self._lib.orc_face_blocks(self._h, e1, e2, _d(b)):
            raise ValueError("no face between the elements")
        return b.reshape(4, self.n, self.n).transpose(0, 2, 1)

    def face_normal(self, e1, e2):
        nor = np.zeros(self.dim)
        if self._lib.orc_face_normal(self._h, e1, e2, _d(nor)):
            raise ValueError("no face between the elements")
        return nor
