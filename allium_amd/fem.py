"""Host-side finite-element post-processing on uniform box meshes: the
Gauss-Lobatto nodal basis of the L2 space (MFEM's L2_FECollection with
BasisType::GaussLobatto, exe/chive.cpp:411), the projection of a source
function Q(x, Omega) onto it (the DomainLFIntegrator part of
FormTransportSource, src/sweep.cpp:1135-1185) and the L2 error of a discrete
field against a function (GridFunction::ComputeL2Error, as chive's "L2 error",
exe/chive.cpp:1036-1072).

These run on the host once per problem (set-up and reporting), never inside
a sweep or an iteration.  Vectors follow the library's layouts: phi is
(Ne x n) with elements x-fastest and lexicographic nodes, psi is
(Nomega x Ne x n) for one group.
"""
import numpy as np


def gll_nodes(p):
    """Gauss-Lobatto nodes on [0, 1] of the order-p L2 basis (p = 0: the midpoint)."""
    if p == 0:
        return np.array([0.5])
    if p == 1:
        return np.array([0.0, 1.0])
    if p == 2:
        return np.array([0.0, 0.5, 1.0])
    if p == 3:
        r = 1.0 / np.sqrt(5.0)
        return np.array([0.0, 0.5 * (1 - r), 0.5 * (1 + r), 1.0])
    raise ValueError("Gauss-Lobatto nodes are tabulated for p = 0..3")


def lagrange(nodes, x):
    """Values (len(x), len(nodes)) of the Lagrange polynomials of `nodes` at x."""
    x = np.asarray(x, float)
    out = np.ones((x.size, len(nodes)))
    for j, xj in enumerate(nodes):
        for k, xk in enumerate(nodes):
            if k != j:
                out[:, j] *= (x - xk) / (xj - xk)
    return out


def gauss(npts):
    """Gauss-Legendre points and weights on [0, 1]."""
    x, w = np.polynomial.legendre.leggauss(npts)
    return 0.5 * (x + 1.0), 0.5 * w


class BoxSpace:
    """The order-p GLL L2 space on a uniform box mesh nel (2D or 3D) of the
    given extents, with a tensor Gauss rule of `npts` points per axis
    (default p + 3, exact for the products MFEM's default orders integrate)."""

    def __init__(self, nel, p, extents=None, npts=None):
        self.nel = tuple(int(v) for v in nel)
        self.dim = len(self.nel)
        if self.dim not in (2, 3):
            raise ValueError("2D or 3D meshes")
        self.p = int(p)
        self.ext = tuple(extents) if extents is not None else (1.0,) * self.dim
        self.h = np.array([self.ext[d] / self.nel[d] for d in range(self.dim)])
        self.Ne = int(np.prod(self.nel))
        self.n = (self.p + 1) ** self.dim
        self._nq1 = npts or self.p + 3
        xg, wg = gauss(self._nq1)
        lag = lagrange(gll_nodes(self.p), xg)             # (nq1, p+1)
        nq1 = xg.size
        # tensor rule, x fastest: local point coordinates, weights and basis values
        idx = np.stack(np.meshgrid(*[np.arange(nq1)] * self.dim, indexing="ij"), -1).reshape(-1, self.dim)[:, ::-1]
        self.xi = xg[idx]                                   # (nq, dim) in [0,1]^dim
        self.w = np.prod(wg[idx], axis=1) * np.prod(self.h)
        nidx = np.stack(np.meshgrid(*[np.arange(self.p + 1)] * self.dim, indexing="ij"), -1).reshape(-1, self.dim)[:, ::-1]
        self.B = np.ones((idx.shape[0], self.n))
        for d in range(self.dim):
            self.B *= lag[idx[:, d]][:, nidx[:, d]]
        # element origins, elements x-fastest
        eidx = np.stack(np.meshgrid(*[np.arange(k) for k in self.nel[::-1]], indexing="ij"), -1).reshape(-1, self.dim)[:, ::-1]
        self.origin = eidx * self.h

    def points(self):
        """Physical quadrature points (Ne, nq, dim)."""
        return self.origin[:, None, :] + self.xi[None, :, :] * self.h

    def project_source(self, Q, omega):
        """int Q(x, Omega_a) phi_i dx for every direction: (Nomega, Ne * n),
        the volume part of FormTransportSource for one group.  Q takes points
        (..., dim) and one direction (dim,) and returns (...)."""
        X = self.points()
        out = np.empty((len(omega), self.Ne * self.n))
        for a, om in enumerate(np.asarray(omega, float)):
            v = Q(X, om[: self.dim]) * self.w[None, :]
            out[a] = (v @ self.B).ravel()
        return out

    def project_inflow(self, value, omega, sides):
        """int_{inflow side} |Omega . n| psi_inc(x, Omega_a) phi_i ds on the listed
        boundary sides (side = axis * 2 + {0 low, 1 high}): the boundary part of
        FormTransportSource (BoundaryFlowIntegrator with alpha = -1, beta = -1/2,
        which keeps only Omega . n < 0).  Returns (Nomega, Ne * n)."""
        nq1 = self._nq1
        xg, wg = gauss(nq1)
        lagq = lagrange(gll_nodes(self.p), xg)                 # (nq1, p+1)
        lag_end = lagrange(gll_nodes(self.p), np.array([0.0, 1.0]))
        out = np.zeros((len(omega), self.Ne, self.n))
        om_all = np.asarray(omega, float)
        for side in sides:
            ax, hi = side >> 1, side & 1
            others = [d for d in range(self.dim) if d != ax]
            fidx = np.stack(np.meshgrid(*[np.arange(nq1)] * (self.dim - 1), indexing="ij"), -1).reshape(-1, self.dim - 1)[:, ::-1]
            xi = np.empty((fidx.shape[0], self.dim))
            xi[:, ax] = float(hi)
            wf = np.ones(fidx.shape[0])
            for k, d in enumerate(others):
                xi[:, d] = xg[fidx[:, k]]
                wf *= wg[fidx[:, k]] * self.h[d]
            nidx = np.stack(np.meshgrid(*[np.arange(self.p + 1)] * self.dim, indexing="ij"), -1).reshape(-1, self.dim)[:, ::-1]
            B = np.ones((fidx.shape[0], self.n))
            for d in range(self.dim):
                if d == ax:
                    B *= lag_end[hi][nidx[:, d]][None, :]
                else:
                    k = others.index(d)
                    B *= lagq[fidx[:, k]][:, nidx[:, d]]
            cell = np.nonzero(np.isclose(self.origin[:, ax], (self.nel[ax] - 1) * self.h[ax] * hi))[0]
            X = self.origin[cell][:, None, :] + xi[None, :, :] * self.h
            for a, om in enumerate(om_all):
                dot = om[ax] if hi else -om[ax]
                if dot >= 0:
                    continue
                v = value(X, om[: self.dim]) * wf[None, :] * (-dot)
                out[a, cell] += v @ B
        return out.reshape(len(omega), -1)

    def l2_error(self, u, exact):
        """sqrt(int (u_h - exact)^2): u holds Ne * n nodal values, exact takes
        points (..., dim)."""
        uh = np.asarray(u, float).reshape(self.Ne, self.n) @ self.B.T
        d = uh - exact(self.points())
        return float(np.sqrt(np.sum(d * d * self.w[None, :])))
