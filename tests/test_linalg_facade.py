"""The C++ value types of include/vpfft_linalg.hpp (VPFFT::LinearAlgebra and
FCC_SchmidtBasis, Src/LinearAlgebra.h, Src/BurgerVectorBasis.h), compiled
with g++ and run on the CPU (tests/cpp/linalg_check.cpp): conversions
against the oracle's, bit for bit, and against what the mathematics fixes."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def run_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("linalg") / "linalg_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-Wall", "-Wextra", "-Werror",
                    "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "linalg_check.cpp"),
                    "-o", exe], check=True)

    def run(*angles):
        out = subprocess.run([exe] + [repr(a) for a in angles], check=True, capture_output=True, text=True).stdout
        res = {}
        for line in out.strip().splitlines():
            k, *v = line.split()
            res.setdefault(k, []).append(np.array([float(x) for x in v]))
        return {k: (v[0] if len(v) == 1 else np.stack(v)) for k, v in res.items()}
    return run


M0 = np.array([[0.3, -1.2, 0.7], [0.4, 2.0, -0.1], [1.1, 0.25, -0.9]])
W = np.array([0.5, -0.25, 1.5, 2.0, -1.0])


def _mul3(A, B):
    C = np.empty((3, 3))
    for i in range(3):
        for j in range(3):
            C[i, j] = A[i, 0] * B[0, j] + A[i, 1] * B[1, j] + A[i, 2] * B[2, j]
    return C


def _mat5(v):  # column-major storage -> (i, j)
    return np.asarray(v).reshape(5, 5).T


@pytest.mark.parametrize("angles", [(0.3, 1.1, 2.0), (-0.6981317007977318, 0.0, 0.0), (5.5, 3.0, 0.1)])
def test_euler_matrix_single_precision_trig(run_check, angles):
    """BuildActiveEulerMatrix is the active ZXZ rotation with float cosines
    and sines (LinearAlgebra.cpp:161-181)."""
    r = run_check(*angles)
    R = r["R"].reshape(3, 3)
    phi, theta, psi = angles

    def rz(a):
        return np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    rx = np.array([[1, 0, 0], [0, np.cos(theta), -np.sin(theta)], [0, np.sin(theta), np.cos(theta)]])
    exact = rz(phi) @ rx @ rz(psi)
    assert np.max(np.abs(R - exact)) < 3e-7  # float trig
    assert np.max(np.abs(R @ R.T - np.eye(3))) < 1e-6
    # R_zz is cos(theta) alone: a float value; R_zx = sin(theta) sin(psi) is
    # a product of two floats, exact in double
    assert float(np.float32(R[2, 2])) == R[2, 2]
    st, sp = np.float32(np.sin(np.float32(theta))), np.float32(np.sin(np.float32(psi)))
    assert abs(R[2, 0] - float(st) * float(sp)) <= 2 * abs(R[2, 0]) * 2.0 ** -23


def test_eigenrep_round_trip_matches_oracle(run_check, oracle):
    r = run_check(0.3, 1.1, 2.0)
    assert np.array_equal(r["eigenrep"], oracle.eigenrep(M0))
    back = r["matrix"].reshape(3, 3)
    assert np.array_equal(back, oracle.matrix_rep(r["eigenrep"]))
    # EigenRep keeps the deviatoric symmetric part: dev(sym(M))
    S = 0.5 * (M0 + M0.T)
    assert np.max(np.abs(back - (S - np.trace(S) / 3 * np.eye(3)))) < 1e-15
    assert np.max(np.abs(r["sym"] - r["eigenrep"])) < 1e-15
    assert np.all(r["antisym"] == 0.0)
    R = r["R"].reshape(3, 3)
    assert np.array_equal(r["rotated"].reshape(3, 3), _mul3(_mul3(R, M0), R.T.copy()))


def test_products_and_running_mean(run_check):
    r = run_check(0.3, 1.1, 2.0)
    v = r["eigenrep"]
    inner = 0.0
    for a in range(5):
        inner += v[a] * W[a]
    assert r["inner"][0] == inner
    assert r["norm"][0] == 2.75
    A = np.empty((5, 5))
    for i in range(5):
        for j in range(5):
            A[i, j] = v[i] * W[j]
    A[2, 2] += 4.0
    assert np.array_equal(_mat5(r["outer"]), A)
    Av, vA = np.zeros(5), np.zeros(5)
    for i in range(5):
        for j in range(5):
            Av[i] += A[i, j] * W[j]
            vA[j] += W[i] * A[i, j]
    assert np.array_equal(r["Av"], Av) and np.array_equal(r["vA"], vA)
    B = np.eye(5)
    B[0, 4], B[1, 3] = 1e-12, 3.0
    AB = np.zeros((5, 5))
    for i in range(5):
        for j in range(5):
            for k in range(5):
                AB[i, j] += A[i, k] * B[k, j]
    assert np.array_equal(_mat5(r["AB"]), AB)
    # running mean of two samples is their mean; |x| < 1e-10 becomes 0
    mean = np.zeros((5, 5))
    mean += (A - mean) / 1.0
    mean += (B - mean) / 2.0
    assert np.array_equal(_mat5(r["moving_average"]), mean)
    assert np.max(np.abs(_mat5(r["moving_average"]) - 0.5 * (A + B))) < 1e-15
    Bn = B.copy()
    Bn[np.abs(Bn) < 1e-10] = 0.0
    assert np.array_equal(_mat5(r["normalized"]), Bn) and Bn[0, 4] == 0.0


def test_fcc_schmid_basis_matches_oracle(run_check, oracle):
    r = run_check(0.0, 0.0, 0.0)
    sym, anti = oracle.schmid_basis()
    assert np.array_equal(r["schmid_basis"], sym.reshape(12, 9))
    assert np.array_equal(r["anti_basis"], anti.reshape(12, 9))
    S = r["schmid_basis"].reshape(12, 3, 3)
    assert np.max(np.abs(np.trace(S, axis1=1, axis2=2))) < 1e-15
    assert np.allclose(np.sum(S * S, axis=(1, 2)), 0.5, rtol=0, atol=1e-15)  # |m0|_F = 1/sqrt(2)
