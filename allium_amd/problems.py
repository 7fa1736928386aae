"""Seeded synthetic transport problems (meshes, cross sections, sources).

Every workload the benchmarks and tests use is generated here from a seed, so
the HIP path and the CPU checker always receive identical inputs.

Mesh vertices are structured: ``(n0+1) x (n1+1) [x (n2+1)]`` points, x-fastest,
matching mfem::Mesh::MakeCartesian{2,3}D with sfc_ordering=false (the
reference's default, src/io.cpp:343-363).  Elements are numbered x-fastest.
"""
import numpy as np


def cartesian_vertices(nel, extents=None):
    """Vertex coordinates of a uniform Cartesian grid on [0, L_d]."""
    dim = len(nel)
    extents = [1.0] * dim if extents is None else list(extents)
    axes = [np.linspace(0.0, extents[d], nel[d] + 1) for d in range(dim)]
    grids = np.meshgrid(*axes, indexing="ij")
    # x-fastest ordering: transpose so axis 0 varies fastest in C order
    pts = np.stack([g.transpose(*reversed(range(dim))).reshape(-1) for g in grids], axis=1)
    return np.ascontiguousarray(pts)


def jittered_vertices(nel, frac, rng, extents=None):
    """Cartesian vertices with interior nodes moved uniformly in +-frac*h per axis.

    Boundary nodes only move tangentially (they stay on their boundary plane);
    corner and edge nodes keep the coordinates that pin them to the boundary.
    """
    dim = len(nel)
    extents = [1.0] * dim if extents is None else list(extents)
    v = cartesian_vertices(nel, extents).reshape(*reversed([n + 1 for n in nel]), dim)
    for d in range(dim):
        h = extents[d] / nel[d]
        shift = rng.uniform(-frac * h, frac * h, size=v.shape[:-1])
        # index along axis d in the reversed-shape array
        ax = dim - 1 - d
        idx = np.arange(nel[d] + 1)
        on_bdr = (idx == 0) | (idx == nel[d])
        shape = [1] * dim
        shape[ax] = nel[d] + 1
        shift = np.where(on_bdr.reshape(shape), 0.0, shift)
        v[..., d] += shift
    return np.ascontiguousarray(v.reshape(-1, dim))


def lognormal_sigma(ne, log_sd, rng, mean=1.0):
    """Per-element lognormal total cross sections with E[sigma] = mean.

    log sigma ~ N(mu, s^2) with mu = log(mean) - s^2/2, so the arithmetic mean
    of the distribution is ``mean`` (not its median)."""
    mu = np.log(mean) - 0.5 * log_sd * log_sd
    return np.exp(rng.normal(mu, log_sd, size=ne))


def problem_2d_cfg1():
    """32x32 quads on the unit square, p=1, S8 (40 dirs), sigma_t=1, c=0.9, Q=1."""
    nel = (32, 32)
    ne = nel[0] * nel[1]
    return dict(dim=2, nel=nel, vertices=cartesian_vertices(nel), p=1, sn=8,
                sigma_t=np.ones((1, ne)), sigma_s=np.full(ne, 0.9), Q=1.0)


def problem_3d_hetero(n=64, seed=0, log_sd=1.0, c_range=(0.5, 0.99), c_const=None, p=1, sn=8):
    """Unit cube n^3 hexes, lognormal sigma_t (E=1), uniform c, Q=1, vacuum."""
    rng = np.random.default_rng(seed)
    nel = (n, n, n)
    ne = n ** 3
    sig = lognormal_sigma(ne, log_sd, rng)
    c = np.full(ne, c_const) if c_const is not None else rng.uniform(c_range[0], c_range[1], ne)
    return dict(dim=3, nel=nel, vertices=cartesian_vertices(nel), p=p, sn=sn,
                sigma_t=sig.reshape(1, ne), sigma_s=c * sig, c=c, Q=1.0, seed=seed)
