"""Shared test helpers: fixtures from tests/golden and seeded problems."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def ls_fixture(order, dim):
    """Directions and weights of the reference's data/lsquad/LS_<order>.txt
    (src/angular_quadrature.cpp:114-140: 2D keeps the first half, weights x2)."""
    with open(os.path.join(GOLDEN, "lsquad.json")) as fh:
        rows = np.array(json.load(fh)["sets"][str(order)])
    if dim < 3:
        rows = rows[: len(rows) // 2]
        return rows[:, :dim].copy(), rows[:, 3] * 2.0
    return rows[:, :3].copy(), rows[:, 3].copy()


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
