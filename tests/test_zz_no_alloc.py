"""allium_balance_device_noalloc (the balance with a file-scope device
workspace, no allocation) against allium_balance_device on the same inputs,
through the C ABI: two calls in a row, the second on a smaller problem."""
import ctypes

import numpy as np
import pytest

import allium_amd as A
from allium_amd import _lib
from allium_amd import problems as P

pytestmark = pytest.mark.gpu


def _balance_pair(nel, sn, G, seed):
    rng = np.random.default_rng(seed)
    ne = int(np.prod(nel))
    sig = np.stack([P.lognormal_sigma(ne, 1.0, rng) for _ in range(G)])
    Linv = A.InverseAdvectionOperator(A.Mesh(nel), 1, sn, sig)
    psi = A.DeviceVector(Linv.size, host=rng.uniform(0.0, 1.0, Linv.size))
    lib = _lib.load()
    ref, got = ctypes.c_double(), ctypes.c_double()
    _lib.check(lib.allium_balance_device(Linv._h, psi.ptr, ctypes.byref(ref), None))
    _lib.check(lib.allium_balance_device_noalloc(Linv._h, psi.ptr, ctypes.byref(got), None))
    return ref.value, got.value


def test_noalloc_balance_matches_allocating_balance():
    cases = [((16, 12, 10), 8, 2, 1), ((6, 5, 4), 4, 1, 2)]   # the second call is smaller
    for nel, sn, G, seed in cases:
        print(f"noalloc balance {nel} S{sn} G={G}", flush=True)
        ref, got = _balance_pair(nel, sn, G, seed)
        print(f"  allocating {ref!r}, static workspace {got!r}")
        assert ref > 0 and abs(got - ref) <= 1e-13 * abs(ref)
