"""CPU oracle of diffusion synthetic acceleration (test infrastructure only).

Restates, in FP64 numpy/scipy:
  * the scalar MIP diffusion matrix of chive's "scalar mip" preconditioner
    (InteriorPenaltyDiscretization::GetOperator, src/moment_discretization.cpp:73-109;
    MIPDiffusionIntegrator::AssembleFaceMatrix, src/moment_integrators.cpp:3-195;
    penalty factor and lower bound as exe/chive.cpp:663-682 sets them;
    ComputeAlpha, src/angular_quadrature.cpp:266-276) on structured quad/hex
    meshes, returned as a scipy CSR matrix in the phi layout;
  * chive's "p1sa" 2x2 block operator (p1_matrix, P1Discretization,
    src/moment_discretization.cpp:400-490);
  * DiffusionSyntheticAccelerationOperator (src/transport_op.hpp:38-59) inside
    SLISolver (src/linalg.cpp:103-153): z = r + K^{-1} Ms r, x += z, norm ||z||,
    with K solved directly (scipy's sparse LU), as chive's "direct" inner solver.
Pinned in tests/test_dsa_oracle.py by closed forms (row sums: absorption plus
Marshak boundary mass; the product with a linear function: the boundary flux
D du/dn) and by the reference's thick-diffusion-limit test deck
(test/chive/tdl.lua, test/chive/test_tdl.py: fewer than 12 iterations).
p1_matrix is pinned there by closed forms on constant and element-wise
constant fields (volume, boundary and interior face terms of every block,
G = -D^T) and by the same deck with the p1sa preconditioner.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import gauss_rule, gll_nodes


def compute_alpha(omega, weights):
    """sum_a |Omega_a . e_x| w_a / sum_a w_a."""
    omega = np.asarray(omega, float)
    w = np.asarray(weights, float)
    return float(np.sum(np.abs(omega[:, 0]) * w) / np.sum(w))


def _lagrange(nodes, x):
    """1D Lagrange basis on `nodes` at points x: values and derivatives (len(x), len(nodes))."""
    k = len(nodes)
    v = np.ones((len(x), k))
    d = np.zeros((len(x), k))
    for i in range(k):
        for m in range(k):
            if m == i:
                continue
            r = 1.0 / (nodes[i] - nodes[m])
            d[:, i] = d[:, i] * (x - nodes[m]) * r + v[:, i] * r
            v[:, i] *= (x - nodes[m]) * r
    return v, d


def _basis(dim, p, xi):
    """Tensor GLL basis at reference points xi (nq, dim): phi (nq, n), grad (nq, n, dim)."""
    nodes = gll_nodes(p) if p > 0 else np.array([0.5])
    vals, ders = zip(*[_lagrange(nodes, xi[:, d]) for d in range(dim)])
    nq, k = xi.shape[0], p + 1
    idx = np.array(np.unravel_index(np.arange(k ** dim), (k,) * dim)[::-1]).T   # x fastest
    phi = np.ones((nq, k ** dim))
    grad = np.ones((nq, k ** dim, dim))
    for d in range(dim):
        phi *= vals[d][:, idx[:, d]]
        for g in range(dim):
            grad[:, :, g] *= (ders[d] if g == d else vals[d])[:, idx[:, d]]
    return phi, grad


def _rule(dim, order):
    x, w = gauss_rule(order // 2 + 1)
    if dim == 0:
        return np.zeros((1, 0)), np.ones(1)
    grids = np.meshgrid(*([x] * dim), indexing="ij")
    wg = np.meshgrid(*([w] * dim), indexing="ij")
    pts = np.stack([g.ravel() for g in grids[::-1]], -1)           # x fastest
    wts = np.prod(np.stack([g.ravel() for g in wg]), axis=0)
    return pts, wts


class _Geometry:
    def __init__(self, dim, nel, vertices):
        self.dim = dim
        self.nel = list(nel)[:dim]
        self.nv = [n + 1 for n in self.nel]
        self.vert = np.asarray(vertices, float).reshape(-1, dim)
        self.Ne = int(np.prod(self.nel))

    def coords(self, e):
        c, r = [], e
        for n in self.nel:
            c.append(r % n)
            r //= n
        return c

    def corners(self, e):
        c = self.coords(e)
        out = []
        for k in range(2 ** self.dim):
            b = [(k >> d) & 1 for d in range(self.dim)]
            vi, s = 0, 1
            for d in range(self.dim):
                vi += (c[d] + b[d]) * s
                s *= self.nv[d]
            out.append(self.vert[vi])
        return np.array(out)

    def jac(self, e, xi):
        """J[q, m, d] = dX_m / dxi_d of the d-linear map at points xi (nq, dim)."""
        X = self.corners(e)
        _, g = _basis(self.dim, 1, xi)                               # d-linear shape gradients
        return np.einsum("km,qkd->qmd", X, g)


def _face_normal(J, ax, hi):
    """Area-weighted outward normal of the side (ax, hi) from the element Jacobians (nq, dim, dim)."""
    dim = J.shape[1]
    t = [d for d in range(dim) if d != ax]
    if dim == 2:
        N = np.stack([J[:, 1, t[0]], -J[:, 0, t[0]]], -1)
    else:
        N = np.cross(J[:, :, t[0]], J[:, :, t[1]])
    s = np.einsum("qm,qm->q", N, J[:, :, ax])
    sign = np.where((s > 0) == bool(hi), 1.0, -1.0)
    return N * sign[:, None]


def mip_matrix(dim, nel, vertices, p, sigma_t, sigma_s, omega, weights, bc=None, kappa=None, bound_below=True,
               alpha=None, sym=-1.0):
    """Scalar MIP diffusion matrix K (CSR, Ndof x Ndof in the phi layout)."""
    geo = _Geometry(dim, nel, vertices)
    sigma_t = np.asarray(sigma_t, float).reshape(-1)
    sigma_a = sigma_t - np.asarray(sigma_s, float).reshape(-1)
    D = 1.0 / (3.0 * sigma_t)
    if kappa is None:
        kappa = float((p + 1) ** 2)
    if alpha is None:
        alpha = compute_alpha(omega, weights)
    lower = alpha / 2 if bound_below else 0.0
    bc = list(bc) if bc is not None else [0] * (2 * dim)
    n = (p + 1) ** dim
    rows, cols, vals = [], [], []

    def add(e1, e2, B):
        r = np.repeat(np.arange(n) + e1 * n, n)
        c = np.tile(np.arange(n) + e2 * n, n)
        rows.append(r)
        cols.append(c)
        vals.append(B.ravel())

    xv, wv = _rule(dim, 2 * p + dim - 1)
    phi_v, g_v = _basis(dim, p, xv)
    for e in range(geo.Ne):
        J = geo.jac(e, xv)
        det = np.linalg.det(J)
        Jinv = np.linalg.inv(J)
        grad = np.einsum("qnr,qrd->qnd", g_v, Jinv)                 # physical gradients
        w = wv * np.abs(det)
        B = D[e] * np.einsum("q,qid,qjd->ij", w, grad, grad) + sigma_a[e] * np.einsum("q,qi,qj->ij", w, phi_v, phi_v)
        add(e, e, B)
    uf, wf = _rule(dim - 1, 2 * p)
    stride = [1, geo.nel[0], geo.nel[0] * (geo.nel[1] if dim > 1 else 1)]
    for e1 in range(geo.Ne):
        c = geo.coords(e1)
        for ax in range(dim):
            def side_pts(hi):
                xi = np.zeros((len(wf), dim))
                others = [d for d in range(dim) if d != ax]
                for k, d in enumerate(others):
                    xi[:, d] = uf[:, k]
                xi[:, ax] = hi
                return xi
            for hi in (0, 1):
                at_bdr = (c[ax] == 0 and hi == 0) or (c[ax] == geo.nel[ax] - 1 and hi == 1)
                if at_bdr:
                    if bc[2 * ax + hi] != 0:
                        continue
                    xi = side_pts(hi)
                    N = _face_normal(geo.jac(e1, xi), ax, hi)
                    ph, _ = _basis(dim, p, xi)
                    add(e1, e1, alpha * np.einsum("q,qi,qj->ij", wf * np.linalg.norm(N, axis=1), ph, ph))
                    continue
                if hi == 0:
                    continue                                          # each interior face once, from e1 = lower
                e2 = e1 + stride[ax]
                xi1, xi2 = side_pts(1), side_pts(0)
                J1, J2 = geo.jac(e1, xi1), geo.jac(e2, xi2)
                N = _face_normal(J1, ax, 1)
                area = np.linalg.norm(N, axis=1)
                ph1, g1 = _basis(dim, p, xi1)
                ph2, g2 = _basis(dim, p, xi2)
                gr1 = np.einsum("qnr,qrd->qnd", g1, np.linalg.inv(J1))
                gr2 = np.einsum("qnr,qrd->qnd", g2, np.linalg.inv(J2))
                # {D grad u . N} with the 1/2 average, per face point
                dn1 = 0.5 * D[e1] * np.einsum("qnd,qd->qn", gr1, N) * wf[:, None]
                dn2 = 0.5 * D[e2] * np.einsum("qnd,qd->qn", gr2, N) * wf[:, None]
                v = np.concatenate([ph1, -ph2], axis=1)                  # [v] = v1 - v2
                du = np.concatenate([dn1, dn2], axis=1)
                E = np.einsum("qi,qj->ij", v, du)
                pen = kappa * wf * area ** 2 * 0.5 * (D[e1] / np.abs(np.linalg.det(J1)) + D[e2] / np.abs(np.linalg.det(J2)))
                if kappa != 0.0:
                    pen = np.maximum(pen, lower * wf * area)
                else:
                    pen = np.zeros_like(pen)
                Jm = np.einsum("q,qi,qj->ij", pen, v, v)
                F = sym * E.T - E + Jm
                add(e1, e1, F[:n, :n])
                add(e1, e2, F[:n, n:])
                add(e2, e1, F[n:, :n])
                add(e2, e2, F[n:, n:])
    N0 = geo.Ne * n
    K = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N0, N0))
    return K.tocsr()


def dsa_source_iteration(pr, sigma_s, source, K, abstol=1e-8, reltol=0.0, maxit=1000, solve=None):
    """SLISolver on T = I - D L^{-1} D^T Ms with the DSA preconditioner
    z = r + K^{-1} Ms r; returns (phi, iterations, norms).  `solve` replaces
    the direct solve of K (e.g. p1_solver)."""
    if solve is None:
        solve = spla.factorized(sp.csc_matrix(K))
    ms = lambda x: pr.mass(sigma_s, x)
    b = pr.D(pr.sweep(source))
    x = np.zeros_like(b)
    norms = []
    r0 = None
    for it in range(1, maxit + 1):
        r = b - x + pr.D(pr.sweep(pr.DT(ms(x))))
        z = r + solve(ms(r))
        x = x + z
        nrm = float(np.sqrt(z @ z))
        norms.append(nrm)
        if it == 1:
            r0 = max(nrm * reltol, abstol)
        if nrm < r0:
            break
    return x, it, np.array(norms)


def p1_matrix(dim, nel, vertices, p, sigma_t, sigma_s, bc=None, alpha=0.5):
    """chive's "p1sa" operator, the 2x2 block system of P1Discretization::GetOperator
    (src/moment_discretization.cpp:400-490) as a CSR matrix, unknowns per element
    [J_0 (n), ..., J_{dim-1} (n), phi (n)] (element e, component c, node i at
    (e (dim+1) + c) n + i):
      Mt = VectorMass(3 sigma_t) + DGVectorJumpJump(1/(2 alpha)) on interior and
           inflow (Marshak) sides, 1/alpha on reflective sides (src/moment_integrators.cpp:495-580);
      Ma = Mass(sigma_a) + Penalty(alpha/2, unscaled) on interior faces (:206-287)
           + BoundaryMass(alpha/2) on Marshak sides;
      D  = -(grad psi, J) + DGJumpAverage [psi]{J.n} on interior faces and 0.5 psi J.n
           on Marshak sides (:289-388);  G = -D^T.
    alpha is P1Discretization's 0.5 (chive does not set it for p1sa,
    exe/chive.cpp:608-619).  Volume rule order 2p + dim - 1, face rule 2p."""
    return _block_moment_matrix(dim, nel, vertices, p, sigma_t, sigma_s, bc, alpha, None)


def block_ip_matrix(dim, nel, vertices, p, sigma_t, sigma_s, omega, weights, bc=None, kappa=None,
                    bound_below=True, scale=True):
    """chive's "mip" operator, the 2x2 block system of BlockIPDiscretization::GetOperator
    (src/moment_discretization.cpp:342-398) as exe/chive.cpp:635-650 sets it up, in
    p1_matrix's layout:
      Mt = VectorMass(3 sigma_t) (no face terms);
      Ma = Mass(sigma_a) + PenaltyIntegrator on interior faces (src/moment_integrators.cpp:206-287):
           weight w_q |N| max({D} kappa |N| / det J_1, limit), {D} the mean of 1/(3 sigma_t)
           (scale) or 1, limit alpha/2 (bound_below) or 0, kappa (p+1)^2 by default;
           + BoundaryMass(alpha) on Marshak sides, alpha = ComputeAlpha;
      D  = -(grad psi, J) + DGJumpAverage on interior faces;  G = -D^T.
    The reference solves its Schur complement Ma + D Mt^-1 D^T (FormSchurComplement,
    :248-264), whose solution is the phi-part of the block solve with [0; b]."""
    alpha = compute_alpha(omega, weights)
    ip = {"kappa": float((p + 1) ** 2 if kappa is None else kappa),
          "limit": alpha / 2 if bound_below else 0.0, "scale": bool(scale)}
    return _block_moment_matrix(dim, nel, vertices, p, sigma_t, sigma_s, bc, alpha, ip)


def block_ldg_matrix(dim, nel, vertices, p, sigma_t, sigma_s, omega, weights, bc=None):
    """chive's "ldgsa" operator, BlockLDGDiscretization::GetOperator
    (src/moment_discretization.cpp:266-340) with SetAlpha(ComputeAlpha), in
    p1_matrix's layout:
      Mt = VectorMass(3 sigma_t);
      Ma = Mass(sigma_a) + Penalty(alpha/2, unscaled) on interior faces + BoundaryMass(alpha)
           on Marshak sides;
      D  = -(grad psi, J) + LDGTraceIntegrator on interior faces (src/moment_integrators.cpp:593-692);
      G = -D^T.
    The trace weighs element 1's current by 1 + sign(beta . N) and element 2's by
    1 - sign(beta . N), with beta = |random| >= 0 (mfem Vector::Randomize) and N out
    of element 1.  On these meshes element 1 is the lower element along the face's
    axis, so beta . N > 0 and the trace is element 1's current: [psi] J_1 . N."""
    alpha = compute_alpha(omega, weights)
    ip = {"ldg": True}
    return _block_moment_matrix(dim, nel, vertices, p, sigma_t, sigma_s, bc, alpha, ip)


def schur_complement(Pb, dim, n):
    """Ma + D Mt^-1 D^T of a block matrix whose Mt is block diagonal by element
    (BlockMomentDiscretization::FormSchurComplement), over the phi unknowns."""
    Pb = sp.csr_matrix(Pb)
    N = (dim + 1) * n
    Ne = Pb.shape[0] // N
    cur = np.concatenate([np.arange(e * N, e * N + dim * n) for e in range(Ne)])
    phi = np.concatenate([np.arange(e * N + dim * n, (e + 1) * N) for e in range(Ne)])
    Mt = Pb[cur][:, cur].toarray()
    G = Pb[cur][:, phi]
    D = Pb[phi][:, cur]
    Ma = Pb[phi][:, phi]
    return (Ma - D @ sp.csr_matrix(np.linalg.inv(Mt)) @ G).tocsr()


def _block_moment_matrix(dim, nel, vertices, p, sigma_t, sigma_s, bc, alpha, ip):
    """p1_matrix (ip None) or block_ip_matrix (ip the penalty parameters)."""
    geo = _Geometry(dim, nel, vertices)
    sigma_t = np.asarray(sigma_t, float).reshape(-1)
    sigma_a = sigma_t - np.asarray(sigma_s, float).reshape(-1)
    bc = list(bc) if bc is not None else [0] * (2 * dim)
    n = (p + 1) ** dim
    N = (dim + 1) * n
    rows, cols, vals = [], [], []

    def add(r, c, v):
        rows.append(np.repeat(r, len(c)))
        cols.append(np.tile(c, len(r)))
        vals.append(v.ravel())

    def idx(e, c):
        return np.arange(n) + (e * (dim + 1) + c) * n

    xv, wv = _rule(dim, 2 * p + dim - 1)
    phi_v, g_v = _basis(dim, p, xv)
    for e in range(geo.Ne):
        J = geo.jac(e, xv)
        det = np.linalg.det(J)
        grad = np.einsum("qnr,qrd->qnd", g_v, np.linalg.inv(J))
        w = wv * np.abs(det)
        M = np.einsum("q,qi,qj->ij", w, phi_v, phi_v)
        for c in range(dim):
            add(idx(e, c), idx(e, c), 3.0 * sigma_t[e] * M)
        add(idx(e, dim), idx(e, dim), sigma_a[e] * M)
        for c in range(dim):
            Dv = -np.einsum("q,qi,qj->ij", w, grad[:, :, c], phi_v)     # row psi_i, column J_c,j
            add(idx(e, dim), idx(e, c), Dv)
            add(idx(e, c), idx(e, dim), -Dv.T)
    uf, wf = _rule(dim - 1, 2 * p)
    stride = [1, geo.nel[0], geo.nel[0] * (geo.nel[1] if dim > 1 else 1)]
    for e1 in range(geo.Ne):
        cc = geo.coords(e1)
        for ax in range(dim):
            def side_pts(hi):
                xi = np.zeros((len(wf), dim))
                others = [d for d in range(dim) if d != ax]
                for k, d in enumerate(others):
                    xi[:, d] = uf[:, k]
                xi[:, ax] = hi
                return xi
            for hi in (0, 1):
                at_bdr = (cc[ax] == 0 and hi == 0) or (cc[ax] == geo.nel[ax] - 1 and hi == 1)
                if at_bdr:
                    xi = side_pts(hi)
                    Nn = _face_normal(geo.jac(e1, xi), ax, hi)
                    area = np.linalg.norm(Nn, axis=1)
                    nh = Nn / area[:, None]
                    ph, _ = _basis(dim, p, xi)
                    if ip is not None:
                        if bc[2 * ax + hi] == 0:
                            add(idx(e1, dim), idx(e1, dim), alpha * np.einsum("q,qi,qj->ij", wf * area, ph, ph))
                        continue
                    beta = 1.0 / alpha if bc[2 * ax + hi] != 0 else 1.0 / (2 * alpha)
                    for c in range(dim):
                        for d in range(dim):
                            add(idx(e1, c), idx(e1, d),
                                beta * np.einsum("q,qi,qj->ij", wf * area * nh[:, c] * nh[:, d], ph, ph))
                    if bc[2 * ax + hi] == 0:
                        add(idx(e1, dim), idx(e1, dim), 0.5 * alpha * np.einsum("q,qi,qj->ij", wf * area, ph, ph))
                        for d in range(dim):
                            Db = 0.5 * np.einsum("q,qi,qj->ij", wf * Nn[:, d], ph, ph)
                            add(idx(e1, dim), idx(e1, d), Db)
                            add(idx(e1, d), idx(e1, dim), -Db.T)
                    continue
                if hi == 0:
                    continue
                e2 = e1 + stride[ax]
                xi1, xi2 = side_pts(1), side_pts(0)
                Nn = _face_normal(geo.jac(e1, xi1), ax, 1)
                area = np.linalg.norm(Nn, axis=1)
                nh = Nn / area[:, None]
                ph1, _ = _basis(dim, p, xi1)
                ph2, _ = _basis(dim, p, xi2)
                jmp = [(e1, ph1, 1.0), (e2, ph2, -1.0)]
                if ip is None or ip.get("ldg"):
                    pen = np.full(len(wf), 0.5 * alpha)
                else:
                    det1 = np.linalg.det(geo.jac(e1, xi1))
                    Davg = 0.5 * (1 / (3 * sigma_t[e1]) + 1 / (3 * sigma_t[e2])) if ip["scale"] else 1.0
                    pen = np.maximum(Davg * ip["kappa"] * area / det1, ip["limit"])
                for ea, pa, sa in jmp:
                    for eb, pb, sb in jmp:
                        base = sa * sb * np.einsum("q,qi,qj->ij", wf * area * pen, pa, pb)
                        add(idx(ea, dim), idx(eb, dim), base)
                        if ip is None:
                            for c in range(dim):
                                for d in range(dim):
                                    add(idx(ea, c), idx(eb, d), (0.5 / alpha) * sa * sb *
                                        np.einsum("q,qi,qj->ij", wf * area * nh[:, c] * nh[:, d], pa, pb))
                        # [psi]{J.n}: test side sign sa, trial average 1/2; LDG: trial element 1 only
                        if ip is not None and ip.get("ldg"):
                            if eb != e1:
                                continue
                            tw = 1.0
                        else:
                            tw = 0.5
                        for d in range(dim):
                            Df = tw * sa * np.einsum("q,qi,qj->ij", wf * Nn[:, d], pa, pb)
                            add(idx(ea, dim), idx(eb, d), Df)
                            add(idx(eb, d), idx(ea, dim), -Df.T)
    Ntot = geo.Ne * N
    P = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(Ntot, Ntot))
    return P.tocsr()


def p1_solver(P, dim, n):
    """phi-part of P^{-1} [0; b] (the TripleProductOperator of
    ComponentExtractionOperator, exe/chive.cpp:612-617), P solved directly."""
    solve = spla.factorized(sp.csc_matrix(P))
    N = (dim + 1) * n

    def apply(b):
        Ne = b.size // n
        full = np.zeros(Ne * N)
        full.reshape(Ne, dim + 1, n)[:, dim, :] = b.reshape(Ne, n)
        return solve(full).reshape(Ne, dim + 1, n)[:, dim, :].ravel()
    return apply
