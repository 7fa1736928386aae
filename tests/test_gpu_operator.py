"""The C++ adapter include/allium_operator.hpp on the GPU: Mult (into a
std::vector and in place), ApplyL and ComputeBalance of
allium::InverseAdvectionOperator, called from C++ the way the reference's
drivers call InverseAdvectionOperator and AdvectionOperator
(exe/chive.cpp:709-735), through lib/liballium_operator.so (built with the
library), against the oracle."""
import ctypes
import os

import numpy as np
import pytest

import allium_amd as A
import oracle
from allium_amd import _lib
from allium_amd import problems as P
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

SHIM = os.path.join(os.path.dirname(_lib.LIB_PATH), "liballium_operator.so")


@pytest.mark.parametrize("nel,p,sn,jitter", [((12, 10), 2, 6, 0.1), ((8, 8, 8), 1, 8, 0.0), ((6, 5, 4), 2, 4, 0.05)])
def test_cpp_adapter_matches_oracle(nel, p, sn, jitter):
    print(f"C++ adapter {nel} p={p} S{sn}", flush=True)
    _lib.load()
    shim = ctypes.CDLL(SHIM)
    dim = len(nel)
    ne = int(np.prod(nel))
    rng = np.random.default_rng(3)
    sig = P.lognormal_sigma(ne, 1.0, rng).reshape(1, ne)
    v = P.jittered_vertices(nel, jitter, rng) if jitter else P.cartesian_vertices(nel)
    mesh = A.Mesh(nel, vertices=v)
    om, w = A.level_symmetric(sn, dim)
    pr = oracle.Problem(dim, nel, v, p, om, w, sig)
    q = pr.form_source(Q=1.0) + rng.uniform(0.0, 0.1, pr.size)
    n = pr.size
    psi, inplace, Lpsi = np.empty(n), np.empty(n), np.empty(n)
    bal = ctypes.c_double()
    md = mesh.desc()
    qd, keep = A._quad_desc(sn, dim)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rc = shim.allium_operator_shim_run(ctypes.byref(md), p, ctypes.byref(qd), 1, dp(np.ascontiguousarray(sig)),
                                       dp(q), ctypes.c_long(n), dp(psi), dp(inplace), dp(Lpsi), ctypes.byref(bal))
    assert rc == 0, _lib.load().allium_last_error().decode()
    ref = pr.sweep(q)
    e_psi = rel_err(psi, ref)
    e_L = rel_err(Lpsi, pr.apply_L(psi))
    e_q = rel_err(Lpsi, q)
    e_b = abs(bal.value - pr.balance(psi)) / abs(pr.balance(psi))
    print(f"  Mult {e_psi:.1e}, ApplyL {e_L:.1e}, L L^-1 q - q {e_q:.1e}, balance {e_b:.1e}")
    assert e_psi < 1e-12 and np.array_equal(psi, inplace)
    assert e_L < 1e-12 and e_q < 1e-11 and e_b < 1e-12
