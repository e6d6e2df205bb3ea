"""Host-side checks that need no GPU: the C ABI's exported symbols, the
microstructure generator, and the Python mirror's pure helpers."""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"\b(vpfft_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol(vpfft):
    lib = vpfft.load_library()
    declared = declared_symbols()
    assert len(declared) >= 15
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert set(vpfft.EXPORTED) == declared


def test_error_path_without_grid(vpfft):
    lib = vpfft.load_library()
    rc = lib.vpfft_grid_set_voxel_length(None, 1.0, 1.0, 1.0)
    assert rc == -1
    assert b"null grid" in lib.vpfft_last_error()


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        M = 0xFFFFFFFFFFFFFFFF
        self.s = (self.s + 0x9E3779B97F4A7C15) & M
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    def u01(self):
        return (self.next() >> 11) * 2.0 ** -53


def active_euler(phi, theta, psi):
    """SMatrix3x3::BuildActiveEulerMatrix (LinearAlgebra.cpp:161-181), FP64."""
    cp, ct, cs = np.cos(phi), np.cos(theta), np.cos(psi)
    sp, st, ss = np.sin(phi), np.sin(theta), np.sin(psi)
    return np.array([[cp * cs - sp * ct * ss, -cp * ss - sp * ct * cs, sp * st],
                     [sp * cs + cp * ct * ss, -sp * ss + cp * ct * cs, -cp * st],
                     [st * ss, st * cs, ct]])


def test_voronoi_generator_matches_brute_force(vpfft):
    dims, ng, seed = (14, 9, 11), 23, 77
    R, grain, eul = vpfft.generate_voronoi(*dims, ng, seed)
    rng = SplitMix64(seed)
    centres = np.empty((ng, 3))
    ref_eul = np.empty((ng, 3))
    for g in range(ng):
        centres[g] = [rng.u01() * dims[0], rng.u01() * dims[1], rng.u01() * dims[2]]
        ref_eul[g] = [2 * np.pi * rng.u01(), np.arccos(2 * rng.u01() - 1), 2 * np.pi * rng.u01()]
    assert np.allclose(eul, ref_eul, rtol=0, atol=1e-15)
    L = np.array(dims, dtype=float)
    idx = np.indices(dims).reshape(3, -1).T + 0.5
    d = np.abs(idx[:, None, :] - centres[None, :, :])
    d = np.minimum(d, L - d)
    d2 = (d ** 2).sum(-1)
    ref_grain = np.argmin(d2, axis=1)  # first minimum = lowest id on ties
    assert np.array_equal(grain, ref_grain)
    for v in range(0, R.shape[0], 37):
        assert np.allclose(R[v].reshape(3, 3), active_euler(*eul[grain[v]]), atol=1e-15)


def test_voronoi_orientations_are_rotations_and_deterministic(vpfft):
    R1, g1, _ = vpfft.generate_voronoi(16, 16, 16, 40, 5)
    R2, g2, _ = vpfft.generate_voronoi(16, 16, 16, 40, 5)
    assert np.array_equal(R1, R2) and np.array_equal(g1, g2)
    R3, _, _ = vpfft.generate_voronoi(16, 16, 16, 40, 6)
    assert not np.array_equal(R1, R3)
    for r in R1[::97]:
        M = r.reshape(3, 3)
        assert np.allclose(M @ M.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(M) - 1) < 1e-14
    assert len(np.unique(g1)) > 30


def test_haar_orientations_uniform_cos_theta(vpfft):
    _, _, eul = vpfft.generate_voronoi(4, 4, 4, 20000, 9)
    c = np.cos(eul[:, 1])
    hist, _ = np.histogram(c, bins=10, range=(-1, 1))
    assert hist.min() > 1700 and hist.max() < 2300
    assert eul[:, 0].min() >= 0 and eul[:, 0].max() < 2 * np.pi


def test_eigenrep_helpers(vpfft, oracle):
    rng = np.random.default_rng(3)
    v = rng.standard_normal(5)
    assert np.allclose(vpfft.ToMatrixRep(v), oracle.matrix_rep(v), atol=1e-15)
    M = vpfft.ToMatrixRep(v)
    assert np.allclose(vpfft.EigenRep(M), v, atol=1e-15)
    E = vpfft.EigenRep(np.diag([1.0, -0.5, -0.5]))
    assert np.allclose(E, [-3 / (2 * np.sqrt(2)), -np.sqrt(6) / 4, 0, 0, 0], atol=1e-15)


def test_material_struct_roundtrip(vpfft):
    m = vpfft.Material.uniform(15.5, 10)
    c = m.to_c()
    assert list(c.rate_sensitivity) == [10] * 12
    assert list(c.gamma_dot_base) == [15.5] * 12
    assert c.theta0[3] == 430.0 and c.hardening[143] == 1.0


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus N > 1 exits with a message before any
    GPU work."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)
