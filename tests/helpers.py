"""Shared workload builders for the tests (seeded, deterministic)."""
import numpy as np

import vpfft_amd as V

TENSION = np.diag([1.0, -0.5, -0.5])


def edot_tension():
    return V.EigenRep(TENSION)


def random_unit_edot(seed):
    """Random symmetric traceless strain rate of Frobenius norm 1 (the basis is
    orthonormal, so a unit 5-vector)."""
    rng = np.random.default_rng(seed)
    e = rng.standard_normal(5)
    return e / np.linalg.norm(e)


def polycrystal(n, grains, seed, dims=None):
    nx, ny, nz = dims or (n, n, n)
    R, grain, eul = V.generate_voronoi(nx, ny, nz, grains, seed)
    return R, grain


def random_rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])
