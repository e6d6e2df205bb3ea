"""Pins for the CPU FP64 oracle (runs without a GPU).

The reference ships no golden vectors, known-answer tests or fixtures for this
path, so each oracle function is checked against what the mathematics fixes:
closed forms, invariants, an independent derivation, numpy's FFT, or brute
force on tiny inputs -- chosen so that a dropped term, a wrong sign or a
transposed operand fails one of them.
"""
import numpy as np
import pytest

from helpers import random_rotation


def colmajor(M):
    return np.asarray(M, dtype=np.float64).T.reshape(25)


def from_colmajor(v):
    return np.asarray(v).reshape(5, 5).T


def spd(rng, scale=4.0):
    A = rng.standard_normal((5, 5))
    return scale * (A @ A.T / 5 + np.eye(5))


# ---- bases -------------------------------------------------------------------
def test_eigen_basis_orthonormal_deviatoric(oracle):
    B = oracle.eigen_basis()
    G = np.einsum("aij,bij->ab", B, B)
    assert np.allclose(G, np.eye(6), atol=1e-15)
    for a in range(5):
        assert abs(np.trace(B[a])) < 1e-15
        assert np.array_equal(B[a], B[a].T)
    assert np.allclose(B[5], np.eye(3) / np.sqrt(3))


def test_schmid_basis_fcc(oracle):
    sym, anti = oracle.schmid_basis()
    b_dirs, n_dirs = [], []
    for s in range(12):
        m, a = sym[s], anti[s]
        assert abs(np.trace(m)) < 1e-15
        assert np.array_equal(m, m.T)
        assert np.array_equal(a, -a.T)
        assert abs(np.linalg.norm(m) - 1 / np.sqrt(2)) < 1e-15
        P = (m + a) * 2 * np.sqrt(6)  # = b n^T
        b = P[:, np.argmax(np.abs(P).sum(0))]
        n = P[np.argmax(np.abs(P).sum(1)), :]
        assert abs(b @ n) < 1e-12  # glide direction lies in the slip plane
        b_dirs.append(tuple(np.sign(b) * np.sign(b[np.nonzero(b)[0][0]])))
        n_dirs.append(tuple(np.sign(n) * np.sign(n[0])))
    # 4 {111} planes x 3 <110> directions, each system distinct
    assert len(set(n_dirs)) == 4
    assert len(set(zip(b_dirs, n_dirs))) == 12


def test_eigenrep_roundtrip_and_exact_symmetry(oracle):
    rng = np.random.default_rng(1)
    for _ in range(50):
        v = rng.standard_normal(5)
        M = oracle.matrix_rep(v)
        # exactly symmetric: the dead FFT term of the orientation update
        # (AntiSymmetrize of this matrix) is exactly zero
        assert np.array_equal(M, M.T)
        assert np.allclose(oracle.eigenrep(M), v, atol=1e-15)
        assert abs(np.trace(M)) < 1e-14


def test_update_schmid_identity_and_rotation_invariance(oracle):
    sym, _ = oracle.schmid_basis()
    m_id = oracle.update_schmid(np.eye(3))
    for s in range(12):
        assert np.allclose(m_id[s], oracle.eigenrep(sym[s]), atol=1e-15)
    rng = np.random.default_rng(2)
    G0 = m_id @ m_id.T
    for _ in range(10):
        R = random_rotation(rng)
        m = oracle.update_schmid(R)
        assert np.allclose(m @ m.T, G0, atol=1e-13)  # orthogonal R acts isometrically
        s = oracle.matrix_rep(rng.standard_normal(5))
        # m_k . sigma = <R m0 R^T, S> = <m0, R^T S R>
        lhs = m @ oracle.eigenrep(s)
        rhs = np.array([np.sum(sym[k] * (R.T @ s @ R)) for k in range(12)])
        assert np.allclose(lhs, rhs, atol=1e-13)


# ---- linear algebra ------------------------------------------------------------
def test_lu_full_pivot_solves(oracle):
    rng = np.random.default_rng(3)
    for _ in range(40):
        A = rng.standard_normal((5, 5))
        b = rng.standard_normal(5)
        x, rank = oracle.lu_solve(colmajor(A), b)
        assert rank == 5
        assert np.allclose(A @ x, b, atol=1e-11)
    # rank-deficient: unknowns past the rank are zero, consistent system solved
    A = np.zeros((5, 5))
    A[:3, :3] = rng.standard_normal((3, 3))
    b = np.zeros(5)
    b[:3] = rng.standard_normal(3)
    x, rank = oracle.lu_solve(colmajor(A), b)
    assert rank == 3
    assert np.allclose(A @ x, b, atol=1e-11)
    assert np.sum(x == 0) >= 2


def test_inverse(oracle):
    rng = np.random.default_rng(4)
    for n in (4, 5):
        A = rng.standard_normal((n, n)) + n * np.eye(n)
        inv, rc = oracle.inverse(A.T.reshape(-1), n)
        assert rc == 0
        assert np.allclose(inv.reshape(n, n).T, np.linalg.inv(A), atol=1e-13)


def test_ipow_correctly_rounded(oracle):
    import ctypes as C
    from fractions import Fraction
    lib = oracle.load()
    lib.vo_ipow.restype = C.c_double
    lib.vo_ipow.argtypes = [C.c_double, C.c_int]
    rng = np.random.default_rng(5)
    for n in (0, 1, 2, 3, 9, 10, 19, 20):
        for x in rng.uniform(0, 2, 60):
            exact = Fraction(float(x)) ** n
            assert lib.vo_ipow(float(x), n) == float(exact)  # float(Fraction) rounds once


# ---- constitutive law ---------------------------------------------------------
def test_apply_constitutive_homogeneous_and_odd(oracle, vpfft):
    rng = np.random.default_rng(6)
    m = oracle.update_schmid(random_rotation(rng))
    tau = rng.uniform(0.5, 2, 12)
    for n in (1, 5, 10, 20):
        mat = vpfft.Material.uniform(1.7, n)
        s = rng.standard_normal(5)
        e = oracle.apply_constitutive(s, m, tau, mat)
        assert np.allclose(oracle.apply_constitutive(1.3 * s, m, tau, mat), 1.3 ** n * e, rtol=1e-12)
        assert np.allclose(oracle.apply_constitutive(-s, m, tau, mat), -e, rtol=1e-14)
        # closed form: sum_k m_k gdot0 |r_k|^n sgn r_k
        r = (m @ s) / tau
        ref = (m * (1.7 * np.abs(r) ** n * np.sign(r))[:, None]).sum(0)
        assert np.allclose(e, ref, rtol=1e-12, atol=1e-14)


def test_taylor_cube_orientation_closed_form(oracle, vpfft):
    """Cube orientation under axisymmetric tension: by cubic symmetry the
    Taylor stress is parallel to E, sigma = c E with c^n |edot(E)|/|E| = 1."""
    m = oracle.update_schmid(np.eye(3))
    tau = np.ones(12)
    E = vpfft.EigenRep(np.diag([-0.5, -0.5, 1.0]))
    for n in (3, 10):
        mat = vpfft.Material.uniform(1.0, n)
        e1 = oracle.apply_constitutive(E, m, tau, mat)
        lam = np.dot(e1, E) / np.dot(E, E)
        assert np.allclose(e1, lam * E, atol=1e-14)  # edot(E) || E
        c = lam ** (-1.0 / n)
        guess = E / np.max(np.abs(m @ E))
        s = oracle.taylor_nr(guess, m, tau, mat, E, 5e-4, 2.0, 1000)
        assert np.allclose(s, c * E, rtol=1e-6), (s, c * E)


def test_taylor_residual_small_at_return(oracle, vpfft):
    rng = np.random.default_rng(7)
    mat = vpfft.Material.uniform(1.0, 10)
    E = rng.standard_normal(5)
    E /= np.linalg.norm(E)
    for _ in range(20):
        m = oracle.update_schmid(random_rotation(rng))
        tau = np.ones(12)
        g = E / np.max(np.abs(m @ E))
        s = oracle.taylor_nr(g, m, tau, mat, E, 5e-4, 2.0, 1000)
        assert np.linalg.norm(oracle.apply_constitutive(s, m, tau, mat) - E) < 1e-3


def test_constrained_nr_solves_its_residual(oracle, vpfft):
    """sigma - lambda + L0 (edot(sigma) - E - d) = 0 at return (Solvers.cpp:205-207)."""
    rng = np.random.default_rng(8)
    for n in (1, 10, 20):
        mat = vpfft.Material.uniform(1.0, n)
        L0 = spd(rng)
        E = rng.standard_normal(5)
        E /= np.linalg.norm(E)
        for _ in range(10):
            m = oracle.update_schmid(random_rotation(rng))
            tau = rng.uniform(0.8, 1.2, 12)
            s0 = rng.standard_normal(5)
            s0 /= oracle.max_rss(s0, m, tau)
            lam = s0 + 0.05 * rng.standard_normal(5)
            d = 0.05 * rng.standard_normal(5)
            s, err = oracle.constrained_nr(s0, m, tau, mat, E, lam, d, colmajor(L0))
            def residual(x):
                return x - lam + L0 @ (oracle.apply_constitutive(x, m, tau, mat) - E - d)
            scale = np.linalg.norm(s) + np.linalg.norm(lam)
            # the Newton loop stops on a relative step below 5e-4, so the
            # residual left is of that order at most (far smaller when the
            # last steps converge quadratically)
            assert np.linalg.norm(residual(s)) < 1e-3 * scale, (n, np.linalg.norm(residual(s)))
            assert np.linalg.norm(residual(s)) < 1e-2 * np.linalg.norm(residual(s0))
            assert err < 5e-4


def test_constrained_nr_linear_case_is_exact(oracle, vpfft):
    """n=1: edot = M sigma is linear, so Newton lands on the solution of
    (I + L0 M) sigma = lambda + L0 (E + d) after one step."""
    rng = np.random.default_rng(9)
    mat = vpfft.Material.uniform(2.0, 1)
    m = oracle.update_schmid(random_rotation(rng))
    tau = rng.uniform(0.8, 1.2, 12)
    M = sum(2.0 / tau[k] * np.outer(m[k], m[k]) for k in range(12))
    L0 = spd(rng)
    E, lam, d = rng.standard_normal(5), rng.standard_normal(5), rng.standard_normal(5)
    s0 = rng.standard_normal(5)
    s, _ = oracle.constrained_nr(s0, m, tau, mat, E, lam, d, colmajor(L0))
    exact = np.linalg.solve(np.eye(5) + L0 @ M, lam + L0 @ (E + d))
    assert np.allclose(s, exact, rtol=1e-12, atol=1e-12)


def test_shear_rate_skew_and_hardening(oracle, vpfft):
    rng = np.random.default_rng(10)
    R = random_rotation(rng)
    m = oracle.update_schmid(R)
    tau = np.ones(12)
    mat = vpfft.Material.uniform(1.0, 10)
    s = rng.standard_normal(5)
    W = oracle.shear_rate(s, m, R, tau, mat)
    assert np.allclose(W, -W.T, atol=1e-15)
    # closed form of the rotated anti-Schmid sum
    _, anti = oracle.schmid_basis()
    r = m @ s
    ref = sum(R @ anti[k] @ R.T * np.abs(r[k]) ** 10 * np.sign(r[k]) for k in range(12))
    assert np.allclose(W, ref, atol=1e-13)
    # Voce hardening, scalar restatement of Solvers.cpp:343-371 as written
    crss0 = rng.uniform(0.9, 1.1, 12)
    A0, dt = 0.3, 0.005
    new, A1 = oracle.hardening(A0, dt, s, m, crss0, mat)
    g = np.abs((m @ s) / crss0) ** 10
    local = np.sum(g) * dt
    fr = abs(430.0 / 11.0)
    ei, ed = np.exp(-A0 * fr), np.exp(-local * fr)
    voce = 110.0 * local - (fr * 11.0 * 110.0) / fr * ei * (ed - 1) - 110.0 / fr * ei * (
        ed * ((A0 + local) * fr + 1) - (A0 * fr + 1))
    ref = crss0 + (np.ones((12, 12)) @ (g * dt)) * voce / local
    assert np.allclose(new, ref, rtol=1e-12)
    assert abs(A1 - (A0 + local)) < 1e-15


# ---- Green operator -------------------------------------------------------------
def green_closed_form(L0, k):
    """G = V N V^T (the negative of SymGreensOperator), V_a = B_a k,
    N = K^-1 - K^-1 k k^T K^-1 / (k^T K^-1 k), K = V^T L0 V.  Derived from the
    bordered acoustic matrix of MaterialGrid.cpp:589-612 by block inversion."""
    import vpfft_amd
    B = vpfft_amd.eigen_basis()[:5]
    V = np.stack([B[a] @ k for a in range(5)])  # 5 x 3
    K = V.T @ L0 @ V
    Ki = np.linalg.inv(K)
    p = Ki @ k
    N = Ki - np.outer(p, p) / (k @ p)
    return V @ N @ V.T


def test_green_operator_projector_and_closed_form(oracle):
    rng = np.random.default_rng(11)
    for _ in range(20):
        L0 = spd(rng)
        k = rng.standard_normal(3)
        k /= np.linalg.norm(k)
        S5 = from_colmajor(oracle.green_sym5(colmajor(L0), k))
        G = -S5
        assert np.allclose(G, green_closed_form(L0, k), atol=1e-12)
        assert np.allclose(G @ L0 @ G, G, atol=1e-12)          # projector w.r.t. L0
        assert np.linalg.matrix_rank(G, tol=1e-9) == 2
        assert np.allclose(from_colmajor(oracle.green_sym5(colmajor(L0), -k)), S5, atol=1e-13)   # even
        assert np.allclose(from_colmajor(oracle.green_sym5(colmajor(L0), 3.7 * k)), S5, atol=1e-12)  # scale-free


def test_green_operator_isotropic_closed_form(oracle, vpfft):
    """L0 = 2 mu I: G_ab = (1/mu)[(B_a k).(B_b k) - (k.B_a k)(k.B_b k)] for |k| = 1."""
    mu = 1.7
    B = vpfft.eigen_basis()[:5]
    rng = np.random.default_rng(12)
    for _ in range(10):
        k = rng.standard_normal(3)
        k /= np.linalg.norm(k)
        Bk = np.stack([B[a] @ k for a in range(5)])
        kBk = np.array([k @ B[a] @ k for a in range(5)])
        G = (Bk @ Bk.T - np.outer(kBk, kBk)) / mu
        S5 = from_colmajor(oracle.green_sym5(colmajor(2 * mu * np.eye(5)), k))
        assert np.allclose(-S5, G, atol=1e-13)


def test_special_plane_operator_is_inverse(oracle):
    rng = np.random.default_rng(13)
    L0 = spd(rng)
    S = from_colmajor(oracle.special_sym5(colmajor(L0)))
    assert np.allclose(S, np.linalg.inv(-L0), atol=1e-13)


# ---- FFT -----------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(8, 8, 8), (6, 10, 12), (24, 5, 7), (16, 48, 9), (80, 3, 4), (112, 2, 2),
                                   (96, 4, 2), (11, 13, 2)])
def test_fft3d_matches_numpy(oracle, shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    f = oracle.fft3d(x, -1)
    assert np.allclose(f, np.fft.fftn(x), atol=1e-10 * np.sqrt(x.size))
    b = oracle.fft3d(x, +1)
    assert np.allclose(b, np.fft.ifftn(x) * x.size, atol=1e-10 * np.sqrt(x.size))


def test_fft1d_primes_and_impulse(oracle):
    for n in (2, 3, 5, 7, 11, 13, 17, 64, 81, 125, 343):
        x = np.zeros(n, dtype=complex)
        x[1 % n] = 1
        f = oracle.fft1d(x, -1)
        assert np.allclose(f, np.exp(-2j * np.pi * np.arange(n) / n * (1 % n)), atol=1e-13)


def _reference_srev(oracle, dims, L0, phi, lengths):
    """numpy C2C restatement with the reference's special planes, using the
    oracle's per-frequency operator (pins the FFT sign, normalisation, the
    special-plane rule and the real-part read, MaterialGrid.cpp:638-752)."""
    nx, ny, nz = dims
    P = np.fft.fftn(phi.reshape(5, nx, ny, nz), axes=(1, 2, 3))
    D = np.zeros_like(P)
    Ssp = from_colmajor(oracle.special_sym5(colmajor(L0)))
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                if ix == iy == iz == 0:
                    continue
                spec = any(N % 2 == 0 and i in (N // 2, N // 2 + 1) for i, N in ((ix, nx), (iy, ny), (iz, nz)))
                if spec:
                    S = Ssp
                else:
                    k = np.array([ix - nx if ix > nx // 2 else ix, iy - ny if iy > ny // 2 else iy,
                                  iz - nz if iz > nz // 2 else iz], dtype=float)
                    k /= np.array(dims) * np.array(lengths)
                    k /= np.linalg.norm(k)
                    S = from_colmajor(oracle.green_sym5(colmajor(L0), k))
                D[:, ix, iy, iz] = -S @ P[:, ix, iy, iz]
    return np.real(np.fft.ifftn(D, axes=(1, 2, 3))).reshape(5, -1)


@pytest.mark.parametrize("dims", [(4, 6, 8), (5, 4, 6), (6, 6, 3)])
def test_strain_rate_evolution_matches_numpy_restatement(oracle, vpfft, dims):
    rng = np.random.default_rng(sum(dims))
    L0 = spd(rng)
    phi = rng.standard_normal((5, int(np.prod(dims))))
    mat = vpfft.Material.uniform()
    lengths = (1.0, 1.3, 0.7)
    d = oracle.strain_rate_evolution(dims, mat, colmajor(L0), phi, lengths)
    ref = _reference_srev(oracle, dims, L0, phi, lengths)
    assert np.allclose(d, ref, atol=1e-12 * np.abs(ref).max())


def test_special_plane_r2c_equivalence(oracle, vpfft):
    """The half-spectrum formulation the GPU uses: C2R of D'(k) with
    G_eff(k) = (G(k) + G(-k))/2 equals the reference's Re(C2C^-1 D)."""
    dims = (6, 4, 8)
    nx, ny, nz = dims
    rng = np.random.default_rng(14)
    L0 = spd(rng)
    phi = rng.standard_normal((5, nx * ny * nz))
    ref = oracle.strain_rate_evolution(dims, vpfft.Material.uniform(), colmajor(L0), phi)
    Ssp = from_colmajor(oracle.special_sym5(colmajor(L0)))

    def op(ix, iy, iz):
        if ix == iy == iz == 0:
            return np.zeros((5, 5))
        if any(N % 2 == 0 and i in (N // 2, N // 2 + 1) for i, N in ((ix, nx), (iy, ny), (iz, nz))):
            return Ssp
        k = np.array([ix - nx if ix > nx // 2 else ix, iy - ny if iy > ny // 2 else iy,
                      iz - nz if iz > nz // 2 else iz], dtype=float) / np.array(dims)
        return from_colmajor(oracle.green_sym5(colmajor(L0), k / np.linalg.norm(k)))

    P = np.fft.rfftn(phi.reshape(5, nx, ny, nz), axes=(1, 2, 3))
    D = np.zeros_like(P)
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz // 2 + 1):
                Geff = 0.5 * (op(ix, iy, iz) + op((-ix) % nx, (-iy) % ny, (-iz) % nz))
                D[:, ix, iy, iz] = -Geff @ P[:, ix, iy, iz]
    d = np.fft.irfftn(D, s=dims, axes=(1, 2, 3)).reshape(5, -1)
    assert np.allclose(d, ref, atol=1e-12 * np.abs(ref).max())
    # and the naive half-spectrum version (G(k) instead of G_eff) is wrong
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz // 2 + 1):
                D[:, ix, iy, iz] = -op(ix, iy, iz) @ P[:, ix, iy, iz]
    naive = np.fft.irfftn(D, s=dims, axes=(1, 2, 3)).reshape(5, -1)
    assert np.abs(naive - ref).max() > 1e-3 * np.abs(ref).max()


# ---- whole strain step -----------------------------------------------------------
def test_single_crystal_step_is_taylor(oracle, vpfft):
    """A homogeneous single crystal has no fluctuation: d~ = 0, one iteration,
    <sigma> = the Taylor stress, L0 = its own secant stiffness."""
    dims = (4, 4, 4)
    rng = np.random.default_rng(15)
    R = np.tile(random_rotation(rng).reshape(9), (64, 1))
    mat = vpfft.Material.uniform(1.0, 10)
    E = vpfft.EigenRep(np.diag([1.0, -0.5, -0.5]))
    out = oracle.strain_step(dims, mat, R, np.ones((64, 12)), E, 10, 0.005, 1e-5, epilogue=False)
    assert out["iterations"] == 1
    assert np.abs(out["dtilde"]).max() < 1e-12
    m = oracle.update_schmid(R[0])
    edot = oracle.apply_constitutive(out["avg_stress"], m, np.ones(12), mat)
    assert np.linalg.norm(edot - E) < 1e-4
    assert np.allclose(out["sigma"], out["sigma"][0], atol=1e-14)


def test_strain_step_average_strain_rate_and_cap(oracle, vpfft):
    import vpfft_amd as V
    dims = (12, 12, 12)
    R, _, _ = V.generate_voronoi(*dims, 6, 3)
    mat = vpfft.Material.uniform(1.0, 5)
    E = vpfft.EigenRep(np.diag([1.0, -0.5, -0.5]))
    crss = np.ones((R.shape[0], 12))
    out = oracle.strain_step(dims, mat, R, crss, E, 200, 0.005, 1e-4, epilogue=False)
    assert out["iterations"] < 200
    assert np.allclose(out["avg_strain_rate"], E, atol=2e-3 * np.linalg.norm(E))
    one = oracle.strain_step(dims, mat, R, crss, E, 1, 0.005, 1e-12, epilogue=False)
    assert one["iterations"] == 1 and one["err_strain_rate"] > 1e-12


def test_degenerate_loading_reports_error(oracle, vpfft):
    dims = (4, 4, 4)
    R = np.tile(np.eye(3).reshape(9), (64, 1))
    out = oracle.strain_step(dims, vpfft.Material.uniform(), R, np.ones((64, 12)), np.zeros(5), 5, 0.005, 1e-5)
    assert out["rc"] == -1
