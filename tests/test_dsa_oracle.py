"""CPU checks of diffusion synthetic acceleration: the MIP oracle against closed
forms and the reference's thick-diffusion-limit deck, and the library's host
assembly of the MIP matrix against the oracle.  No GPU compute is called."""
import numpy as np
import pytest
import scipy.sparse as sp

import allium_amd as A
import oracle
from allium_amd import problems as P
from oracle import mip_oracle as M
from tests.helpers import ls_fixture


def _uniform(dim, n):
    nel = (n,) * dim
    return nel, P.cartesian_vertices(nel)


def test_alpha_is_half_for_level_symmetric_sets():
    # sum |mu| w / sum w -> 1/2 for a set that integrates |mu| exactly
    for order, dim in [(4, 2), (8, 3), (12, 3)]:
        om, w = ls_fixture(order, dim)
        a = M.compute_alpha(om, w)
        assert 0.45 < a < 0.55


@pytest.mark.parametrize("dim,p", [(2, 1), (3, 1), (2, 2)])
def test_row_sums_are_absorption_plus_marshak(dim, p):
    """K 1 = sigma_a int phi_i + alpha int_{inflow sides} phi_i (gradients and
    jumps of a constant vanish).  For p = 1 the closed forms are vol / 2^dim per
    node and area / 2^(dim-1) per boundary node."""
    n = 4
    nel, vert = _uniform(dim, n)
    rng = np.random.default_rng(1)
    st = rng.uniform(0.5, 2.0, n ** dim)
    ss = st * rng.uniform(0.1, 0.9, n ** dim)
    om, w = ls_fixture(4, dim)
    bc = [0, 1] + [0] * (2 * dim - 2)           # x-high side reflective: no Marshak term
    K = M.mip_matrix(dim, nel, vert, p, st, ss, om, w, bc=bc)
    alpha = M.compute_alpha(om, w)
    rs = np.asarray(K.sum(axis=1)).ravel()
    h = 1.0 / n
    nn = (p + 1) ** dim
    if p == 1:
        expect = np.zeros_like(rs)
        for e in range(n ** dim):
            c = [(e // n ** d) % n for d in range(dim)]
            for i in range(nn):
                b = [(i >> d) & 1 for d in range(dim)]
                v = (st[e] - ss[e]) * h ** dim / 2 ** dim
                for ax in range(dim):
                    for hi in (0, 1):
                        if bc[2 * ax + hi] == 0 and b[ax] == hi and c[ax] == (n - 1) * hi:
                            v += alpha * h ** (dim - 1) / 2 ** (dim - 1)
                expect[e * nn + i] = v
        np.testing.assert_allclose(rs, expect, rtol=1e-12, atol=1e-15)
    else:
        # totals: sigma_a-weighted volume plus alpha times the inflow boundary measure
        tot = np.sum(st - ss) * h ** dim + alpha * (2 * dim - 1)
        assert abs(rs.sum() - tot) < 1e-12 * tot


@pytest.mark.parametrize("dim,p", [(2, 1), (3, 1), (2, 2), (3, 2)])
def test_linear_function_gives_boundary_flux(dim, p):
    """With constant sigma_t, sigma_a = 0 and no Marshak term, K x_nodal is the
    boundary flux int_{dD} D (du/dn) phi_i of u = x: +D int phi_i on x = 1,
    -D int phi_i on x = 0 (the consistency terms cancel the interior flux)."""
    n = 3
    nel, vert = _uniform(dim, n)
    st = np.full(n ** dim, 2.0)
    om, w = ls_fixture(4, dim)
    K = M.mip_matrix(dim, nel, vert, p, st, st, om, w, alpha=1e-300)   # alpha ~ 0
    nodes = oracle.gll_nodes(p)
    nn = (p + 1) ** dim
    x = np.zeros(n ** dim * nn)
    for e in range(n ** dim):
        for i in range(nn):
            x[e * nn + i] = ((e % n) + nodes[i % (p + 1)]) / n
    y = K @ x
    D = 1.0 / 6.0
    # boundary integrals of the GLL basis on a face: tensor of 1D integrals
    xg, wg = oracle.gauss_rule(p + 2)
    lag = M._lagrange(nodes, xg)[0]
    w1 = (wg[:, None] * lag).sum(0) / n            # int over one face edge of each 1D node function
    expect = np.zeros_like(y)
    for e in range(n ** dim):
        cx = e % n
        for i in range(nn):
            ix = i % (p + 1)
            rest = [(i // (p + 1) ** d) % (p + 1) for d in range(1, dim)]
            face = np.prod([w1[r] for r in rest])
            if cx == n - 1 and ix == p:
                expect[e * nn + i] += D * face
            if cx == 0 and ix == 0:
                expect[e * nn + i] -= D * face
    np.testing.assert_allclose(y, expect, atol=1e-13)


@pytest.mark.parametrize("dim,sym", [(2, -1.0), (3, -1.0), (3, 1.0)])
def test_scalar_mip_on_elementwise_constants(dim, sym):
    """The scalar MIP operator (MIPDiffusionIntegrator, src/moment_integrators.cpp:3-195)
    on element-wise constant f, uniform mesh, p = 1, every side Marshak.  For
    node i of element e, vol_i = (h/2)^dim its volume integral, A = (h/2)^(dim-1)
    its face integral on a side through it, and each interior side s of e
    with neighbour value f_s:
      (K f)_i = sigma_a vol_i f_e
              + sum_s [i on s] pen_s A (f_e - f_s)                 (penalty)
              + sum_s sym (D_e / 2) (A / h) s_i (f_e - f_s)        (consistency)
              + sum_{Marshak s through i} alpha A f_e,
    with pen_s = max(kappa (D_e + D_s) / 2 / h, alpha / 2), D = 1/(3 sigma_t),
    kappa = (p+1)^2 and s_i = +1 for i on s, -1 for i on the opposite side
    (the normal derivative of phi_i through s is s_i / h on the face); the
    jump of f and the gradient of phi_i are the only nonzero factors.
    sigma_t spans both branches of the penalty's max."""
    n_el = 3 if dim == 3 else 5
    nel, vert = _uniform(dim, n_el)
    rng = np.random.default_rng(31)
    ne = n_el ** dim
    st = np.exp(rng.uniform(np.log(0.2), np.log(300.0), ne))
    ss = st * rng.uniform(0.1, 0.9, ne)
    om, w = ls_fixture(4, dim)
    alpha = M.compute_alpha(om, w)
    kappa = 4.0
    K = M.mip_matrix(dim, nel, vert, 1, st, ss, om, w, sym=sym)
    h = 1.0 / n_el
    nn = 2 ** dim
    vol, A = (h / 2) ** dim, (h / 2) ** (dim - 1)
    Dc = 1.0 / (3.0 * st)
    f = rng.normal(size=ne)
    y = K @ np.repeat(f, nn)
    expect = np.zeros_like(y)
    floors = 0
    for e in range(ne):
        c = [(e // n_el ** d) % n_el for d in range(dim)]
        for i in range(nn):
            b = [(i >> d) & 1 for d in range(dim)]
            v = (st[e] - ss[e]) * vol * f[e]
            for ax in range(dim):
                for hi in (0, 1):
                    on = b[ax] == hi
                    if c[ax] == (n_el - 1) * hi:            # boundary side: Marshak
                        v += alpha * A * f[e] if on else 0.0
                        continue
                    nb = e + (n_el ** ax) * (1 if hi else -1)
                    pen = kappa * 0.5 * (Dc[e] + Dc[nb]) / h
                    if pen < alpha / 2:
                        pen = alpha / 2
                        floors += 1
                    jump = f[e] - f[nb]
                    if on:
                        v += pen * A * jump
                    v += sym * 0.5 * Dc[e] * (A / h) * (1.0 if on else -1.0) * jump
            expect[e * nn + i] = v
    assert floors > 0 and floors < 4 * ne * dim * nn      # both branches of the max
    np.testing.assert_allclose(y, expect, rtol=1e-12, atol=1e-14 * np.abs(expect).max())


@pytest.mark.parametrize("dim,p,jitter", [(2, 1, 0.0), (3, 1, 0.0), (3, 1, 0.15), (2, 2, 0.2), (3, 2, 0.1)])
def test_oracle_matrix_is_spd(dim, p, jitter):
    n = 3
    nel = (n,) * dim
    rng = np.random.default_rng(3)
    vert = P.jittered_vertices(nel, jitter, rng) if jitter else P.cartesian_vertices(nel)
    st = P.lognormal_sigma(n ** dim, 1.0, rng)
    ss = st * 0.99
    om, w = ls_fixture(4, dim)
    K = M.mip_matrix(dim, nel, vert, p, st, ss, om, w).toarray()
    assert np.abs(K - K.T).max() < 1e-13 * np.abs(K).max()
    assert np.linalg.eigvalsh(K).min() > 0


@pytest.mark.parametrize("dim,p,jitter,sig_scale", [(2, 1, 0.0, 1.0), (3, 1, 0.0, 1.0), (3, 1, 0.15, 1.0),
                                                    (2, 2, 0.2, 1.0), (3, 2, 0.1, 1.0), (2, 1, 0.0, 1e4),
                                                    (3, 0, 0.0, 1.0)])
def test_host_assembly_matches_oracle(dim, p, jitter, sig_scale):
    """allium_host_mip_matrix (C++) against the numpy restatement; sigma_t =
    1e4 drives every face onto the alpha/2 lower bound of the penalty."""
    n = 4 if dim == 2 else 3
    nel = (n,) * dim
    rng = np.random.default_rng(7)
    vert = P.jittered_vertices(nel, jitter, rng) if jitter else P.cartesian_vertices(nel)
    st = P.lognormal_sigma(n ** dim, 1.0, rng) * sig_scale
    ss = st * rng.uniform(0.5, 0.999, n ** dim)
    bc = [0, 1] + [0] * (2 * dim - 2)
    om, w = ls_fixture(4, dim)
    mesh = A.Mesh(nel, vertices=vert, bc=bc)
    rp, col, val = A.host_mip_matrix(mesh, p, (om, w), st, ss)
    Kh = sp.csr_matrix((val, col, rp), shape=(rp.size - 1,) * 2).toarray()
    Ko = M.mip_matrix(dim, nel, vert, p, st, ss, om, w, bc=bc).toarray()
    assert np.abs(Kh - Ko).max() < 1e-12 * np.abs(Ko).max()


def test_host_assembly_options():
    """kappa, the lower bound, alpha and sym reach the assembly as the oracle reads them."""
    nel, vert = _uniform(2, 4)
    rng = np.random.default_rng(2)
    st = P.lognormal_sigma(16, 1.0, rng)
    om, w = ls_fixture(4, 2)
    mesh = A.Mesh(nel, vertices=vert)
    for opts, oopts in [(dict(kappa=9.0), dict(kappa=9.0)), (dict(kappa=-2.0), dict(kappa=8.0)),
                        (dict(bound_below=False), dict(bound_below=False)), (dict(alpha=0.3), dict(alpha=0.3)),
                        (dict(sym=1.0), dict(sym=1.0)), (dict(kappa=0.0), dict(kappa=0.0))]:
        rp, col, val = A.host_mip_matrix(mesh, 1, (om, w), st, st * 0.5, **opts)
        Kh = sp.csr_matrix((val, col, rp), shape=(rp.size - 1,) * 2).toarray()
        Ko = M.mip_matrix(2, nel, vert, 1, st, st * 0.5, om, w, **oopts).toarray()
        assert np.abs(Kh - Ko).max() < 1e-12 * np.abs(Ko).max(), opts


def _tdl(eps=1e-4):
    """test/chive/tdl.lua: 8x8 unit square, sigma_t = 1/eps, sigma_s = 1/eps - eps,
    source eps (per steradian), vacuum, S4 level-symmetric, p = 1."""
    nel = (8, 8)
    vert = P.cartesian_vertices(nel)
    om, w = ls_fixture(4, 2)
    st = np.full(64, 1.0 / eps)
    ss = np.full(64, 1.0 / eps - eps)
    return nel, vert, om, w, st, ss, eps


def test_tdl_deck_converges_in_under_12_iterations():
    """The reference's chive.tdl.mip acceptance test (test/chive/test_tdl.py:
    `assert it < 12` with abstol 1e-6, max_iter 50), with the scalar MIP DSA and
    a direct inner solve.  Plain SI does not converge in 50 iterations there."""
    nel, vert, om, w, st, ss, eps = _tdl()
    pr = oracle.Problem(2, nel, vert, 1, om, w, st)
    q = pr.form_source(Q=eps)
    K = M.mip_matrix(2, nel, vert, 1, st, ss, om, w)
    phi, it, norms = M.dsa_source_iteration(pr, ss, q, K, abstol=1e-6, maxit=50)
    assert norms[-1] < 1e-6 and it < 12
    _, _, it_plain, n_plain = pr.source_iteration(ss, q, abstol=1e-6, maxit=50)
    assert it_plain == 50 and n_plain[-1] > 1e-6
    # the DSA fixed point is the transport solution: plain SI from it stays put
    b = pr.D(pr.sweep(q))
    r = b - phi + pr.D(pr.sweep(pr.DT(pr.mass(ss, phi))))
    assert np.linalg.norm(r) < 1e-5 * np.linalg.norm(phi)


@pytest.mark.parametrize("dim", [2, 3])
def test_p1_blocks_against_closed_forms(dim):
    """chive's "p1sa" operator (oracle p1_matrix) on a uniform mesh with one
    reflective side, p = 1 (node integrals: vol / 2^dim, side area / 2^(dim-1)):
      D c      = -sum_{all sides} int psi c.n + 1/2 sum_{Marshak sides} int psi c.n
                 for a constant current c (integration by parts; interior jumps vanish);
      Ma 1     = sigma_a int psi + alpha/2 int_{Marshak} psi;
      Mt c     = 3 sigma_t int psi c + beta int_side psi n (c.n), beta = 1/(2 alpha) on
                 Marshak and 1/alpha on reflective sides;
      G        = -D^T."""
    n_el = 3
    nel = (n_el,) * dim
    vert = P.cartesian_vertices(nel)
    rng = np.random.default_rng(4)
    ne = n_el ** dim
    st = rng.uniform(0.5, 2.0, ne)
    ss = st * rng.uniform(0.1, 0.9, ne)
    bc = [0, 1] + [0] * (2 * dim - 2)        # x-high side reflective
    alpha = 0.5
    Pm = M.p1_matrix(dim, nel, vert, 1, st, ss, bc=bc, alpha=alpha).toarray()
    n, N = 2 ** dim, (dim + 1) * 2 ** dim
    h = 1.0 / n_el
    vol, area = h ** dim / n, h ** (dim - 1) / 2 ** (dim - 1)

    def blk(e, c):
        return slice((e * (dim + 1) + c) * n, (e * (dim + 1) + c + 1) * n)

    def side_nodes(e):
        """(axis, hi, node mask) of each boundary side the element touches"""
        co = [(e // n_el ** d) % n_el for d in range(dim)]
        out = []
        for ax in range(dim):
            for hi in (0, 1):
                if co[ax] == (n_el - 1) * hi:
                    out.append((ax, hi, np.array([((i >> ax) & 1) == hi for i in range(n)])))
        return out

    c = rng.normal(size=dim)
    x = np.zeros(ne * N)
    for e in range(ne):
        for d in range(dim):
            x[blk(e, d)] = c[d]
    y = Pm @ x
    for e in range(ne):
        Dc = np.zeros(n)
        Mtc = np.array([3 * st[e] * vol * c[d] * np.ones(n) for d in range(dim)])
        for ax, hi, mask in side_nodes(e):
            cn = c[ax] * (1 if hi else -1)
            marshak = bc[2 * ax + hi] == 0
            Dc[mask] += (-1.0 + (0.5 if marshak else 0.0)) * area * cn
            beta = 1 / (2 * alpha) if marshak else 1 / alpha
            Mtc[ax][mask] += beta * area * c[ax]
        np.testing.assert_allclose(y[blk(e, dim)], Dc, atol=1e-14)
        for d in range(dim):
            np.testing.assert_allclose(y[blk(e, d)], Mtc[d], atol=1e-14)
    # Ma on constants (the phi block only)
    x = np.zeros(ne * N)
    for e in range(ne):
        x[blk(e, dim)] = 1.0
    y = Pm @ x
    for e in range(ne):
        expect = (st[e] - ss[e]) * vol * np.ones(n)
        for ax, hi, mask in side_nodes(e):
            if bc[2 * ax + hi] == 0:
                expect[mask] += 0.5 * alpha * area
        phi_rows = y[blk(e, dim)]
        # the G columns act on phi: D^T's contribution shows in the current rows, not here
        np.testing.assert_allclose(phi_rows, expect, atol=1e-14)
    # G = -D^T
    jdx = np.concatenate([np.arange(blk(e, 0).start, blk(e, dim).start) for e in range(ne)])
    pdx = np.concatenate([np.arange(blk(e, dim).start, blk(e, dim).stop) for e in range(ne)])
    np.testing.assert_allclose(Pm[np.ix_(jdx, pdx)], -Pm[np.ix_(pdx, jdx)].T, atol=1e-16)


def test_p1sa_accelerates_the_tdl_deck():
    """SLISolver with the P1SA preconditioner (direct inner solve, as chive
    requires for p1sa, exe/chive.cpp:610) on test/chive/tdl.lua: converged, the
    transport fixed point, and far fewer iterations than plain SI's 50."""
    nel, vert, om, w, st, ss, eps = _tdl()
    pr = oracle.Problem(2, nel, vert, 1, om, w, st)
    q = pr.form_source(Q=eps)
    Pm = M.p1_matrix(2, nel, vert, 1, st, ss)
    phi, it, norms = M.dsa_source_iteration(pr, ss, q, None, abstol=1e-6, maxit=50,
                                            solve=M.p1_solver(Pm, 2, 4))
    assert norms[-1] < 1e-6 and it < 12
    b = pr.D(pr.sweep(q))
    r = b - phi + pr.D(pr.sweep(pr.DT(pr.mass(ss, phi))))
    assert np.linalg.norm(r) < 1e-5 * np.linalg.norm(phi)


@pytest.mark.parametrize("dim", [2, 3])
def test_p1_interior_face_terms_on_elementwise_constants(dim):
    """The interior face terms of p1_matrix, which global constants do not see:
    element-wise constant currents c_e and fluxes f_e on a uniform mesh, p = 1,
    all sides Marshak.  For node i of element e on a side with outward normal n
    (axis a) and node face integral A = side area / 2^(dim-1):
      Mt row (component a):  beta A (c_e - c_nbr)_a on interior sides, beta = 1/(2 alpha)
                             (DGVectorJumpJump on the normal component);
      D row:                 A (c_e + c_nbr)/2 . n - A c_e . n = A (c_nbr - c_e) . n / 2
                             (integration by parts plus the {J.n} average);
      Ma row:                (alpha/2) A (f_e - f_nbr) (unscaled penalty),
    plus the volume and boundary parts of the constant-field test."""
    n_el = 3
    nel = (n_el,) * dim
    vert = P.cartesian_vertices(nel)
    rng = np.random.default_rng(9)
    ne = n_el ** dim
    st = rng.uniform(0.5, 2.0, ne)
    ss = st * rng.uniform(0.1, 0.9, ne)
    alpha = 0.5
    beta = 1 / (2 * alpha)
    Pm = M.p1_matrix(dim, nel, vert, 1, st, ss, alpha=alpha).toarray()
    n, N = 2 ** dim, (dim + 1) * 2 ** dim
    h = 1.0 / n_el
    vol, area = h ** dim / n, h ** (dim - 1) / 2 ** (dim - 1)
    C = rng.normal(size=(ne, dim))
    F = rng.normal(size=ne)
    x = np.zeros((ne, dim + 1, n))
    x[:, :dim, :] = C[:, :, None]
    x[:, dim, :] = F[:, None]
    y = (Pm @ x.ravel()).reshape(ne, dim + 1, n)
    for e in range(ne):
        co = [(e // n_el ** d) % n_el for d in range(dim)]
        Mt = 3 * st[e] * vol * np.repeat(C[e][:, None], n, 1)
        Dr = np.zeros(n)
        Ma = (st[e] - ss[e]) * vol * F[e] * np.ones(n)
        for ax in range(dim):
            for hi in (0, 1):
                mask = np.array([((i >> ax) & 1) == hi for i in range(n)])
                sgn = 1.0 if hi else -1.0
                if co[ax] == (n_el - 1) * hi:
                    # Marshak boundary side
                    Mt[ax][mask] += beta * area * C[e][ax]
                    Dr[mask] += -0.5 * area * sgn * C[e][ax]
                    Ma[mask] += 0.5 * alpha * area * F[e]
                else:
                    o = e + (1 if hi else -1) * n_el ** ax
                    Mt[ax][mask] += beta * area * (C[e][ax] - C[o][ax])
                    Dr[mask] += 0.5 * area * sgn * (C[o][ax] - C[e][ax])
                    Ma[mask] += 0.5 * alpha * area * (F[e] - F[o])
        # the phi rows hold D c + Ma f; the current rows Mt c + G f
        np.testing.assert_allclose(y[e, dim], Dr + Ma, atol=1e-14)
        Gf = y[e, :dim] - Mt
        # G = -D^T: pinned separately; here only that the current rows differ from Mt c by G f
        xf = np.zeros((ne, dim + 1, n))
        xf[:, dim, :] = F[:, None]
        np.testing.assert_allclose(Gf, (Pm @ xf.ravel()).reshape(ne, dim + 1, n)[e, :dim], atol=1e-13)


@pytest.mark.parametrize("dim,p,jitter", [(2, 1, 0.0), (2, 1, 0.2), (2, 2, 0.1), (3, 1, 0.0), (3, 1, 0.15)])
def test_host_p1_assembly_matches_oracle(dim, p, jitter):
    """allium_host_p1_matrix (C++) against the numpy restatement, with one
    reflective side."""
    n = 4 if dim == 2 else 3
    nel = (n,) * dim
    rng = np.random.default_rng(11)
    vert = P.jittered_vertices(nel, jitter, rng) if jitter else P.cartesian_vertices(nel)
    st = P.lognormal_sigma(n ** dim, 1.0, rng)
    ss = st * rng.uniform(0.5, 0.999, n ** dim)
    bc = [0, 1] + [0] * (2 * dim - 2)
    mesh = A.Mesh(nel, vertices=vert, bc=bc)
    rp, col, val = A.host_p1_matrix(mesh, p, st, ss)
    Ph = sp.csr_matrix((val, col, rp), shape=(rp.size - 1,) * 2).toarray()
    Po = M.p1_matrix(dim, nel, vert, p, st, ss, bc=bc).toarray()
    assert np.abs(Ph - Po).max() < 1e-12 * np.abs(Po).max()


@pytest.mark.parametrize("dim", [2, 3])
def test_block_ip_blocks_on_elementwise_constants(dim):
    """chive's "mip" block operator (oracle block_ip_matrix) on a uniform mesh,
    p = 1, all sides Marshak, element-wise constant currents c_e and fluxes f_e.
    For node i of element e on a side with outward normal n (axis a), node face
    integral A and h the element width (area / det J = 1/h):
      Mt row:  3 sigma_t vol c_e only (no face terms in BlockIPDiscretization);
      D row:   A (c_nbr - c_e).n / 2 on interior sides, -A c_e.n on Marshak sides
               (integration by parts; no boundary term in D);
      Ma row:  sigma_a vol f_e + pen A (f_e - f_nbr) on interior sides with
               pen = max(kappa (D_e + D_nbr)/2 / h, alpha/2), D = 1/(3 sigma_t),
               + alpha A f_e on Marshak sides;
      G = -D^T.  sigma_t spans both branches of the penalty's max."""
    n_el = 3
    nel = (n_el,) * dim
    vert = P.cartesian_vertices(nel)
    rng = np.random.default_rng(21)
    ne = n_el ** dim
    st = np.exp(rng.uniform(np.log(0.5), np.log(300.0), ne))
    ss = st * rng.uniform(0.1, 0.9, ne)
    om, w = ls_fixture(4, dim)
    alpha = M.compute_alpha(om, w)
    kappa = 4.0
    Pm = M.block_ip_matrix(dim, nel, vert, 1, st, ss, om, w).toarray()
    n, N = 2 ** dim, (dim + 1) * 2 ** dim
    h = 1.0 / n_el
    vol, area = h ** dim / n, h ** (dim - 1) / 2 ** (dim - 1)
    Dc = 1.0 / (3.0 * st)
    C = rng.normal(size=(ne, dim))
    F = rng.normal(size=ne)
    x = np.zeros((ne, dim + 1, n))
    x[:, :dim, :] = C[:, :, None]
    x[:, dim, :] = F[:, None]
    y = (Pm @ x.ravel()).reshape(ne, dim + 1, n)
    xf = np.zeros((ne, dim + 1, n))
    xf[:, dim, :] = F[:, None]
    yf = (Pm @ xf.ravel()).reshape(ne, dim + 1, n)
    branches = set()
    for e in range(ne):
        co = [(e // n_el ** d) % n_el for d in range(dim)]
        Dr = np.zeros(n)
        Ma = (st[e] - ss[e]) * vol * F[e] * np.ones(n)
        for ax in range(dim):
            for hi in (0, 1):
                mask = np.array([((i >> ax) & 1) == hi for i in range(n)])
                sgn = 1.0 if hi else -1.0
                if co[ax] == (n_el - 1) * hi:
                    Dr[mask] += -area * sgn * C[e][ax]
                    Ma[mask] += alpha * area * F[e]
                else:
                    o = e + (1 if hi else -1) * n_el ** ax
                    Dr[mask] += 0.5 * area * sgn * (C[o][ax] - C[e][ax])
                    val = kappa * 0.5 * (Dc[e] + Dc[o]) / h
                    branches.add(val > alpha / 2)
                    Ma[mask] += max(val, alpha / 2) * area * (F[e] - F[o])
        np.testing.assert_allclose(y[e, dim], Dr + Ma, atol=1e-13)
        Mt = 3 * st[e] * vol * np.repeat(C[e][:, None], n, 1)
        np.testing.assert_allclose(y[e, :dim] - yf[e, :dim], Mt, rtol=1e-13, atol=1e-13)
    assert branches == {True, False}
    jdx = np.concatenate([np.arange(e * N, e * N + dim * n) for e in range(ne)])
    pdx = np.concatenate([np.arange(e * N + dim * n, (e + 1) * N) for e in range(ne)])
    np.testing.assert_allclose(Pm[np.ix_(jdx, pdx)], -Pm[np.ix_(pdx, jdx)].T, atol=1e-16)


def test_block_ip_schur_complement_is_the_block_solve():
    """FormSchurComplement (src/moment_discretization.cpp:248-264): the Schur
    complement Ma + D Mt^-1 D^T is symmetric positive definite and its solve
    equals the phi-part of the block solve with [0; b]; with scale_penalty off
    the penalty is kappa / h (times 1) instead."""
    nel = (5, 4)
    rng = np.random.default_rng(2)
    vert = P.jittered_vertices(nel, 0.15, rng)
    st = P.lognormal_sigma(20, 1.0, rng) * 3.0
    ss = st * 0.95
    om, w = ls_fixture(4, 2)
    Pm = M.block_ip_matrix(2, nel, vert, 1, st, ss, om, w)
    S = M.schur_complement(Pm, 2, 4).toarray()
    np.testing.assert_allclose(S, S.T, atol=1e-12 * np.abs(S).max())
    assert np.linalg.eigvalsh(S).min() > 0
    b = rng.normal(size=80)
    np.testing.assert_allclose(M.p1_solver(Pm, 2, 4)(b), np.linalg.solve(S, b), rtol=1e-9)
    P0 = M.block_ip_matrix(2, nel, vert, 1, st, ss, om, w, scale=False)
    assert abs((P0 - Pm).toarray()).max() > 1e-3


def test_block_ip_accelerates_the_tdl_deck():
    """chive.tdl.mip (test/chive/CMakeLists.txt:117-121, test/chive/test_tdl.py):
    under 12 iterations with the block "mip" preconditioner, at the transport
    fixed point."""
    nel, vert, om, w, st, ss, eps = _tdl()
    pr = oracle.Problem(2, nel, vert, 1, om, w, st)
    q = pr.form_source(Q=eps)
    Pm = M.block_ip_matrix(2, nel, vert, 1, st, ss, om, w)
    phi, it, norms = M.dsa_source_iteration(pr, ss, q, None, abstol=1e-6, maxit=50, solve=M.p1_solver(Pm, 2, 4))
    assert norms[-1] < 1e-6 and it < 12
    b = pr.D(pr.sweep(q))
    r = b - phi + pr.D(pr.sweep(pr.DT(pr.mass(ss, phi))))
    assert np.linalg.norm(r) < 1e-5 * np.linalg.norm(phi)


@pytest.mark.parametrize("dim,p,jitter,scale", [(2, 1, 0.0, True), (2, 1, 0.2, True), (2, 2, 0.1, False),
                                                (3, 1, 0.0, True), (3, 1, 0.15, True)])
def test_host_block_mip_assembly_matches_oracle(dim, p, jitter, scale):
    """allium_host_block_mip_matrix (C++) against the numpy restatement, with
    one reflective side."""
    n = 4 if dim == 2 else 3
    nel = (n,) * dim
    rng = np.random.default_rng(13)
    vert = P.jittered_vertices(nel, jitter, rng) if jitter else P.cartesian_vertices(nel)
    st = np.exp(rng.uniform(np.log(0.5), np.log(100.0), n ** dim))
    ss = st * rng.uniform(0.5, 0.999, n ** dim)
    bc = [0, 1] + [0] * (2 * dim - 2)
    om, w = ls_fixture(4, dim)
    mesh = A.Mesh(nel, vertices=vert, bc=bc)
    rp, col, val = A.host_block_mip_matrix(mesh, p, (om, w), st, ss, scale_penalty=scale)
    Ph = sp.csr_matrix((val, col, rp), shape=(rp.size - 1,) * 2).toarray()
    Po = M.block_ip_matrix(dim, nel, vert, p, st, ss, om, w, bc=bc, scale=scale).toarray()
    np.testing.assert_allclose(Ph, Po, rtol=1e-12, atol=1e-12 * np.abs(Po).max())


@pytest.mark.parametrize("dim", [2, 3])
def test_block_ldg_blocks_on_elementwise_constants(dim):
    """chive's "ldgsa" block operator (oracle block_ldg_matrix), uniform mesh,
    p = 1, all sides Marshak, element-wise constant c_e and f_e.  The LDG trace
    takes the lower element's current (beta . N > 0), so for node i of element e
    on a side with outward normal n and node face integral A:
      D row:  0 on the high interior sides (the trace is e's own current),
              A (c_low - c_e).n on the low interior sides, -A c_e.n on Marshak sides;
      Ma row: sigma_a vol f_e + (alpha/2) A (f_e - f_nbr) interior + alpha A f_e Marshak;
      Mt row: 3 sigma_t vol c_e;  G = -D^T."""
    n_el = 3
    nel = (n_el,) * dim
    vert = P.cartesian_vertices(nel)
    rng = np.random.default_rng(23)
    ne = n_el ** dim
    st = rng.uniform(0.5, 3.0, ne)
    ss = st * rng.uniform(0.1, 0.9, ne)
    om, w = ls_fixture(4, dim)
    alpha = M.compute_alpha(om, w)
    Pm = M.block_ldg_matrix(dim, nel, vert, 1, st, ss, om, w).toarray()
    n, N = 2 ** dim, (dim + 1) * 2 ** dim
    h = 1.0 / n_el
    vol, area = h ** dim / n, h ** (dim - 1) / 2 ** (dim - 1)
    C = rng.normal(size=(ne, dim))
    F = rng.normal(size=ne)
    x = np.zeros((ne, dim + 1, n))
    x[:, :dim, :] = C[:, :, None]
    x[:, dim, :] = F[:, None]
    y = (Pm @ x.ravel()).reshape(ne, dim + 1, n)
    xf = np.zeros((ne, dim + 1, n))
    xf[:, dim, :] = F[:, None]
    yf = (Pm @ xf.ravel()).reshape(ne, dim + 1, n)
    for e in range(ne):
        co = [(e // n_el ** d) % n_el for d in range(dim)]
        Dr = np.zeros(n)
        Ma = (st[e] - ss[e]) * vol * F[e] * np.ones(n)
        for ax in range(dim):
            for hi in (0, 1):
                mask = np.array([((i >> ax) & 1) == hi for i in range(n)])
                sgn = 1.0 if hi else -1.0
                if co[ax] == (n_el - 1) * hi:
                    Dr[mask] += -area * sgn * C[e][ax]
                    Ma[mask] += alpha * area * F[e]
                else:
                    o = e + (1 if hi else -1) * n_el ** ax
                    if not hi:
                        Dr[mask] += area * sgn * (C[o][ax] - C[e][ax])
                    Ma[mask] += 0.5 * alpha * area * (F[e] - F[o])
        np.testing.assert_allclose(y[e, dim], Dr + Ma, atol=1e-13)
        np.testing.assert_allclose(y[e, :dim] - yf[e, :dim], 3 * st[e] * vol * np.repeat(C[e][:, None], n, 1),
                                   rtol=1e-13, atol=1e-13)
    jdx = np.concatenate([np.arange(e * N, e * N + dim * n) for e in range(ne)])
    pdx = np.concatenate([np.arange(e * N + dim * n, (e + 1) * N) for e in range(ne)])
    np.testing.assert_allclose(Pm[np.ix_(jdx, pdx)], -Pm[np.ix_(pdx, jdx)].T, atol=1e-16)


def test_block_ldg_accelerates_the_tdl_deck():
    """chive.tdl.ldgsa (test/chive/CMakeLists.txt:117-121): under 12 iterations,
    at the transport fixed point; its Schur complement is SPD."""
    nel, vert, om, w, st, ss, eps = _tdl()
    pr = oracle.Problem(2, nel, vert, 1, om, w, st)
    q = pr.form_source(Q=eps)
    Pm = M.block_ldg_matrix(2, nel, vert, 1, st, ss, om, w)
    S = M.schur_complement(Pm, 2, 4).toarray()
    assert np.linalg.eigvalsh(0.5 * (S + S.T)).min() > 0
    phi, it, norms = M.dsa_source_iteration(pr, ss, q, None, abstol=1e-6, maxit=50, solve=M.p1_solver(Pm, 2, 4))
    assert norms[-1] < 1e-6 and it < 12
    b = pr.D(pr.sweep(q))
    r = b - phi + pr.D(pr.sweep(pr.DT(pr.mass(ss, phi))))
    assert np.linalg.norm(r) < 1e-5 * np.linalg.norm(phi)


@pytest.mark.parametrize("dim,p,jitter", [(2, 1, 0.0), (2, 2, 0.15), (3, 1, 0.0), (3, 1, 0.15)])
def test_host_block_ldg_assembly_matches_oracle(dim, p, jitter):
    n = 4 if dim == 2 else 3
    nel = (n,) * dim
    rng = np.random.default_rng(17)
    vert = P.jittered_vertices(nel, jitter, rng) if jitter else P.cartesian_vertices(nel)
    st = P.lognormal_sigma(n ** dim, 1.0, rng)
    ss = st * rng.uniform(0.5, 0.999, n ** dim)
    bc = [0, 1] + [0] * (2 * dim - 2)
    om, w = ls_fixture(4, dim)
    mesh = A.Mesh(nel, vertices=vert, bc=bc)
    rp, col, val = A.host_block_ldg_matrix(mesh, p, (om, w), st, ss)
    Ph = sp.csr_matrix((val, col, rp), shape=(rp.size - 1,) * 2).toarray()
    Po = M.block_ldg_matrix(dim, nel, vert, p, st, ss, om, w, bc=bc).toarray()
    np.testing.assert_allclose(Ph, Po, rtol=1e-12, atol=1e-12 * np.abs(Po).max())
