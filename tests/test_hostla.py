"""Host linear algebra of the block solvers (band Sturm counts, inverse
iteration, dense Hermitian eigenpairs, Cholesky) against numpy."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hostla") / "libhostla.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", out,
                    os.path.join(HERE, "native", "hostla_shim.cpp")], check=True)
    return ctypes.CDLL(out)


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def band_matrix(p, m, rng):
    A = rng.normal(size=(m, p, p)) + 1j * rng.normal(size=(m, p, p))
    A = A + np.conj(np.swapaxes(A, 1, 2))
    B = np.triu(rng.normal(size=(m, p, p)) + 1j * rng.normal(size=(m, p, p)))
    T = np.zeros((m * p, m * p), dtype=complex)
    for k in range(m):
        T[k * p:(k + 1) * p, k * p:(k + 1) * p] = A[k]
        if k + 1 < m:  # (k+1, k) block B_k, (k, k+1) block B_k^H
            T[(k + 1) * p:(k + 2) * p, k * p:(k + 1) * p] = B[k]
            T[k * p:(k + 1) * p, (k + 1) * p:(k + 2) * p] = np.conj(B[k].T)
    return A, B, T


def test_band_count_and_inverse_iteration(lib):
    lib.hl_band_count.restype = ctypes.c_int
    lib.hl_band_count.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_double]
    lib.hl_band_eigvec.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]
    rng = np.random.default_rng(0)
    for p, m in [(1, 30), (4, 9), (8, 6)]:
        A, B, T = band_matrix(p, m, rng)
        w, V = np.linalg.eigh(T)
        Ac = np.ascontiguousarray(A)
        Bc = np.ascontiguousarray(B)
        for s in np.linspace(w[0] - 1, w[-1] + 1, 17):
            assert lib.hl_band_count(p, m, P(Ac), P(Bc), s) == np.sum(w < s)
        for k in (0, len(w) // 2, len(w) - 1):
            x = np.zeros(2 * len(w))
            lib.hl_band_eigvec(p, m, P(Ac), P(Bc), w[k], P(x))
            v = x[0::2] + 1j * x[1::2]
            assert abs(np.linalg.norm(v) - 1) < 1e-12
            assert np.linalg.norm(T @ v - w[k] * v) < 1e-9 * np.abs(w).max()


def test_dense_eigh_and_cholesky(lib):
    lib.hl_dense_eigh.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p]
    lib.hl_cholesky_upper.argtypes = [ctypes.c_int, ctypes.c_void_p]
    rng = np.random.default_rng(1)
    s = 12
    H = rng.normal(size=(s, s)) + 1j * rng.normal(size=(s, s))
    H = H + np.conj(H.T)
    # a degenerate pair: the vectors must still be orthonormal
    Q = np.linalg.qr(rng.normal(size=(s, s)) + 1j * rng.normal(size=(s, s)))[0]
    d = np.sort(rng.normal(size=s))
    d[3] = d[4]
    H2 = Q @ np.diag(d) @ np.conj(Q.T)
    for M, want in ((H, np.linalg.eigvalsh(H)), (H2, d)):
        Hc = np.ascontiguousarray(M.astype(complex))
        w = np.zeros(s)
        V = np.zeros(2 * s * s)
        lib.hl_dense_eigh(s, P(Hc), P(w), P(V))
        Vc = (V[0::2] + 1j * V[1::2]).reshape(s, s)
        assert np.abs(w - want).max() < 1e-12 * np.abs(want).max()
        assert np.abs(Vc.conj() @ Vc.T - np.eye(s)).max() < 1e-10
        for k in range(s):
            assert np.linalg.norm(M @ Vc[k] - w[k] * Vc[k]) < 1e-10
    G = rng.normal(size=(6, 6)) + 1j * rng.normal(size=(6, 6))
    G = np.conj(G.T) @ G + 0.1 * np.eye(6)
    Gc = np.ascontiguousarray(G.copy())
    assert lib.hl_cholesky_upper(6, P(Gc))
    R = Gc
    assert np.abs(np.triu(R) - R).max() == 0
    assert np.abs(np.conj(R.T) @ R - G).max() < 1e-12 * np.abs(G).max()


def test_householder_qr_r(lib):
    lib.hl_qr_r.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    rng = np.random.default_rng(8)
    A = rng.normal(size=(9, 5)) + 1j * rng.normal(size=(9, 5))
    R = np.zeros((5, 5), dtype=complex)
    lib.hl_qr_r(9, 5, P(np.ascontiguousarray(A)), P(R))
    assert np.allclose(np.tril(R, -1), 0)
    assert np.allclose(R.conj().T @ R, A.conj().T @ A, atol=1e-12)


def test_refined_ritz_is_min_singular_vector(lib):
    """Refined Ritz vector from the block Arnoldi factors (T, Rw) equals the
    smallest right singular vector of (H - theta) B."""
    lib.hl_refined_ritz.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]
    rng = np.random.default_rng(9)
    n, q, d = 80, 2, 5
    H = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = H + H.conj().T
    wH, vH = np.linalg.eigh(H)
    # start block: an eigenvector with a small error, plus a random column
    S = rng.normal(size=(n, q)) + 1j * rng.normal(size=(n, q))
    S[:, 0] = vH[:, 30] + 1e-6 * S[:, 0]
    blocks = [np.linalg.qr(S)[0]]
    m = q * (d + 1)
    T = np.zeros((m, m), dtype=complex)
    for i in range(d + 1):
        W = H @ blocks[-1]
        B = np.concatenate(blocks, axis=1)
        for _ in range(2):
            c = B.conj().T @ W
            T[:c.shape[0], i * q:(i + 1) * q] += c
            W = W - B @ c
        if i == d:
            break
        Qn, R = np.linalg.qr(W)
        T[(i + 1) * q:(i + 2) * q, i * q:(i + 1) * q] = R
        blocks.append(Qn)
    B = np.concatenate(blocks, axis=1)
    # block Arnoldi relation H B = B T + W E^T
    resid = H @ B - B @ T
    assert np.abs(resid[:, :m - q]).max() < 1e-10
    assert np.abs(resid[:, m - q:] - W).max() < 1e-10
    Rw = np.linalg.cholesky(W.conj().T @ W).conj().T
    theta = wH[30] + 1e-7
    y = np.zeros(m, dtype=complex)
    y[0] = 1.0
    lib.hl_refined_ritz(m, q, P(np.ascontiguousarray(T)), P(np.ascontiguousarray(Rw)),
                        theta, P(y))
    _, s, vh = np.linalg.svd((H - theta * np.eye(n)) @ B)
    v = vh[-1].conj()
    assert abs(abs(np.vdot(v, y)) - 1.0) < 1e-10
    assert abs(np.linalg.norm((H - theta * np.eye(n)) @ B @ y) - s[-1]) < 1e-3 * s[-1]
    assert s[-1] < 1e-5 * s[0]
