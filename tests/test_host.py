"""CPU tests of the product's host side: the C ABI exports, quadrature
generation, mesh/assembly and the sweep levels, each against the reference's
fixtures or the oracle.  No GPU compute is called."""
import ctypes
import os
import re

import numpy as np
import pytest

import allium_amd as A
import oracle
from allium_amd import _lib
from allium_amd import problems as P
from tests.helpers import ls_fixture

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names.update(re.findall(r"\b(allium_\w+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the binding declares every one of them
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


@pytest.mark.parametrize("order", [2, 4, 6, 8, 10, 12, 14, 16])
@pytest.mark.parametrize("dim", [2, 3])
def test_level_symmetric_matches_reference_tables(order, dim):
    om, w = A.level_symmetric(order, dim)
    ref_om, ref_w = ls_fixture(order, dim)
    assert om.shape == ref_om.shape
    assert np.abs(om - ref_om).max() <= 2.3e-16
    assert np.abs(w - ref_w).max() <= 2.3e-16


def test_bad_quadrature_order_reports_error():
    with pytest.raises(A.AlliumError, match="not available"):
        A.level_symmetric(7, 3)


@pytest.mark.parametrize("nel,jitter,sn,bc", [((5, 4, 3), 0.0, 4, None), ((6, 6, 6), 0.2, 6, None),
                                              ((8, 7), 0.25, 8, None), ((6, 5), 0.0, 4, (1, 0, 0, 0)),
                                              ((4, 4, 4), 0.0, 4, (0, 0, 1, 0, 0, 0))])
def test_levels_match_oracle(nel, jitter, sn, bc):
    dim = len(nel)
    v = P.jittered_vertices(nel, jitter, np.random.default_rng(1)) if jitter else P.cartesian_vertices(nel)
    m = A.Mesh(nel, vertices=v, bc=bc)
    lv, nl = A.host_levels(m, sn)
    om, w = ls_fixture(sn, dim)
    ne = int(np.prod(nel))
    pr = oracle.Problem(dim, nel, v, 1, om, w, np.ones((1, ne)), bc=bc)
    assert nl == pr.nlevels
    np.testing.assert_array_equal(lv, pr.levels())


def test_cyclic_graph_is_rejected():
    m = A.Mesh((4, 4), bc=(1, 1, 0, 0))
    with pytest.raises(A.AlliumError, match="DAG"):
        A.host_levels(m, 4)


@pytest.mark.parametrize("nel,jitter,p", [((4, 3), 0.0, 1), ((4, 3), 0.2, 2), ((3, 3, 3), 0.0, 1),
                                          ((3, 3, 3), 0.2, 1), ((3, 2, 2), 0.15, 2)])
def test_element_matrices_and_source_match_oracle(nel, jitter, p):
    dim = len(nel)
    ne = int(np.prod(nel))
    v = P.jittered_vertices(nel, jitter, np.random.default_rng(4)) if jitter else P.cartesian_vertices(nel)
    m = A.Mesh(nel, vertices=v)
    om, w = ls_fixture(4, dim)
    sig = np.vstack([np.ones(ne), 2 * np.ones(ne)])
    pr = oracle.Problem(dim, nel, v, p, om, w, sig)
    for e in (0, ne // 2, ne - 1):
        Mu, G, T = A.host_element_matrices(m, p, e)
        Mo, Go, To = pr.element_matrices(e)
        np.testing.assert_allclose(Mu, Mo, rtol=0, atol=4e-15 * np.abs(Mo).max())
        np.testing.assert_allclose(G, Go, rtol=0, atol=1e-14 * np.abs(Go).max())
        np.testing.assert_allclose(T, To, rtol=0, atol=4e-15 * np.abs(To).max())
    Q = np.vstack([np.linspace(0.5, 1.5, ne), np.ones(ne)])
    inflow = np.arange(4 * dim, dtype=float).reshape(2, 2 * dim)
    q1 = A.host_form_source(m, p, 4, 2, Q=Q, inflow=inflow)
    q2 = pr.form_source(Q=Q, inflow=inflow)
    assert np.abs(q1 - q2).max() <= 1e-14 * np.abs(q2).max()


def test_problem_generators_are_seeded():
    a = P.problem_3d_hetero(n=8, seed=11)
    b = P.problem_3d_hetero(n=8, seed=11)
    np.testing.assert_array_equal(a["sigma_t"], b["sigma_t"])
    s = P.lognormal_sigma(200000, 1.0, np.random.default_rng(0))
    assert abs(s.mean() - 1.0) < 0.02
    v = P.jittered_vertices((6, 6, 6), 0.2, np.random.default_rng(0))
    c = P.cartesian_vertices((6, 6, 6))
    assert np.abs(v - c).max() <= 0.2 / 6 + 1e-15
    # boundary nodes stay on their planes
    on_x0 = np.isclose(c[:, 0], 0.0)
    assert np.all(v[on_x0, 0] == 0.0)


def test_cpp_operator_adapter_compiles_and_links(tmp_path):
    """include/allium_operator.hpp builds against the library with g++ and its
    host-only calls work without a GPU (level-symmetric S8 set in 3D)."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "adapter.cpp"
    src.write_text(r'''
#include "allium_operator.hpp"
#include <cstdio>
int main() {
    double om[80 * 3], w[80];
    int n = 80;
    allium::check(allium_host_quadrature(8, 3, om, w, &n));
    double s = 0;
    for (int i = 0; i < n; ++i) s += w[i];
    try {
        allium::check(allium_host_quadrature(7, 3, om, w, &n));
        return 2;
    } catch (const allium::Error &e) {
        if (e.code() != ALLIUM_ERR_ARG) return 3;
    }
    std::printf("%d %.15f\n", n, s);
    return 0;
}
''')
    exe = tmp_path / "adapter"
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(root, "include"), str(src),
                    "-L", libdir, "-lallium_amd", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert int(out[0]) == 80
    assert abs(float(out[1]) - 4 * np.pi) < 1e-12
