"""Manufactured and exact solutions of the reference's order-of-accuracy
tests, restated as formulas (test/test_sweep.cpp:112-287,
test/test_transport_mms.cpp:7-197), with one 3D analogue of the MMS."""
import numpy as np

from allium_amd import problems as P
from allium_amd.fem import BoxSpace

FOURPI = 4.0 * np.pi


def exp_1d(sigma, x, om):
    """exp(-x sigma / Omega_x): inflow 1 on x = 0, no source (exp_sol_1d)."""
    return np.exp(-x[..., 0] / om[0] * sigma)


def exp_diag(sigma, x, om):
    """exp_sol_2d for a direction with Omega_x = Omega_y: the characteristic
    reaches the nearer of the sides x = 0, y = 0."""
    return np.where(x[..., 0] > x[..., 1], np.exp(-x[..., 1] / om[0] * sigma), np.exp(-x[..., 0] / om[1] * sigma))


def mms_2d(sig_t=1.0, sig_s=0.25):
    """(Q, phi_exact) of LinearTransportError: psi = (sin pi x sin pi y
    + |Omega|^2 sin 2pi x sin 2pi y + 2) / 4pi with the projected 2D |Omega|^2."""
    pi = np.pi

    def Q(x, om):
        X, Y = x[..., 0], x[..., 1]
        quad = float(om @ om)
        dx = np.cos(pi * X) * np.sin(pi * Y) / 4 + quad * np.cos(2 * pi * X) * np.sin(2 * pi * Y) / 2
        dy = np.sin(pi * X) * np.cos(pi * Y) / 4 + quad * np.sin(2 * pi * X) * np.cos(2 * pi * Y) / 2
        s1, s2 = np.sin(pi * X) * np.sin(pi * Y), np.sin(2 * pi * X) * np.sin(2 * pi * Y)
        return om[0] * dx + om[1] * dy + sig_t * (s1 + quad * s2 + 2.0) / FOURPI - sig_s * (s1 + 2. / 3 * s2 + 2.0) / FOURPI

    def phi(x):
        X, Y = x[..., 0], x[..., 1]
        return np.sin(pi * X) * np.sin(pi * Y) + 2. / 3 * np.sin(2 * pi * X) * np.sin(2 * pi * Y) + 2.0

    return Q, phi


def mms_3d(sig_t=1.0, sig_s=0.25):
    """3D analogue: psi = (s1 + Omega_x^2 s2 + 2) / 4pi with s1 = prod sin pi x_d,
    s2 = prod sin 2pi x_d, so phi = s1 + s2 / 3 + 2 (sum w mu^2 = 4pi / 3)."""
    pi = np.pi

    def parts(x):
        X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
        s1 = np.sin(pi * X) * np.sin(pi * Y) * np.sin(pi * Z)
        s2 = np.sin(2 * pi * X) * np.sin(2 * pi * Y) * np.sin(2 * pi * Z)
        g1 = pi * np.stack([np.cos(pi * X) * np.sin(pi * Y) * np.sin(pi * Z),
                            np.sin(pi * X) * np.cos(pi * Y) * np.sin(pi * Z),
                            np.sin(pi * X) * np.sin(pi * Y) * np.cos(pi * Z)], -1)
        g2 = 2 * pi * np.stack([np.cos(2 * pi * X) * np.sin(2 * pi * Y) * np.sin(2 * pi * Z),
                                np.sin(2 * pi * X) * np.cos(2 * pi * Y) * np.sin(2 * pi * Z),
                                np.sin(2 * pi * X) * np.sin(2 * pi * Y) * np.cos(2 * pi * Z)], -1)
        return s1, s2, g1, g2

    def Q(x, om):
        s1, s2, g1, g2 = parts(x)
        mu2 = om[0] ** 2
        grad = (g1 + mu2 * g2) @ om / FOURPI
        return grad + sig_t * (s1 + mu2 * s2 + 2.0) / FOURPI - sig_s * (s1 + s2 / 3 + 2.0) / FOURPI

    def phi(x):
        s1, s2, _, _ = parts(x)
        return s1 + s2 / 3 + 2.0

    return Q, phi


def exponential_problem(nel, p, om, sigmas=(1.0, 0.5)):
    """ExponentialSolution's inputs on a box mesh: one direction of weight 1,
    two groups of total cross sections `sigmas`, no source, inflow 1 on every
    side.  Returns (space, vertices, omega (1 x dim), weights, sigma_t (G x Ne),
    inflow (G x 2dim))."""
    sp = BoxSpace(nel, p)
    dim = len(nel)
    sig = np.repeat(np.asarray(sigmas, float)[:, None], sp.Ne, axis=1)
    return sp, P.cartesian_vertices(nel), np.asarray(om, float).reshape(1, dim), np.ones(1), sig, \
        np.ones((len(sigmas), 2 * dim))


def exponential_error(sp, psi, om, sigmas, exact):
    """sum over groups of the L2 error of psi_g (one direction) against exact."""
    psi = np.asarray(psi).reshape(len(sigmas), sp.Ne * sp.n)
    return sum(sp.l2_error(psi[g], lambda x, s=s: exact(s, x, om)) for g, s in enumerate(sigmas))
