"""The Lanczos start vectors built on the device (CyclicRandom, Const1,
Random: lanczos_initialize, NEMO_3D/nemo3d/src/base/eigsys3d_par.c:113-215)
against their closed forms: the first Lanczos coefficients of the device
run equal those of the oracle's recurrence started from the closed-form
vector (a different start vector changes alpha_0 = r0^H H r0 already)."""
import ctypes
import ctypes.util

import numpy as np
import pytest

import _models
from nemo3d_amd import native
from nemo3d_amd.hamiltonian import Hamiltonian
from nemo3d_amd.solvers import DEFAULT_CYCLIC, eigensolve_lanczos
from oracle import oracle

pytestmark = pytest.mark.gpu


def closed_form(kind, n, seed=7):
    j = np.arange(n)
    if kind == "CyclicRandom":  # (c[2j mod L], c[2j+1 mod L])
        c = np.asarray(DEFAULT_CYCLIC)
        v = c[(2 * j) % len(c)] + 1j * c[(2 * j + 1) % len(c)]
    elif kind == "Const1":  # (1, 1)
        v = np.full(n, 1.0 + 1.0j)
    else:  # Random: (rand() / RAND_MAX, 0) after srand(seed), glibc
        libc = ctypes.CDLL(ctypes.util.find_library("c"))
        libc.rand.restype = ctypes.c_int
        libc.srand(seed)
        v = np.array([1.0 * libc.rand() / 2147483647 for _ in range(n)]) + 0j
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("kind", ["CyclicRandom", "Const1", "Random"])
def test_device_start_vectors_match_closed_form(kind):
    M = _models.gaas_box(cells=(2, 2, 2), strain_nm=0.01, seed=51)
    ref = oracle.assemble(M.arrays())
    H = Hamiltonian(M, layout=native.LAYOUT_SK)
    k = 10
    r = eigensolve_lanczos(H, num_vb=0, num_cb=1, erange_cb=(1e3, 1e3 + 1),
                           max_iter=k, start_iter=10 ** 6, want_vectors=False,
                           return_tridiag=True, start=kind, random_seed=7)
    a, b = ref.lanczos_tridiag(closed_form(kind, M.n), k)
    n = min(len(r["alpha"]), k)
    assert n >= 5
    assert np.abs(r["alpha"][:n] - a[:n]).max() < 1e-12 * 40
    assert np.abs(r["beta"][:n - 1] - b[:n - 1]).max() < 1e-12 * 40
    # a wrong start shows at once
    other = closed_form("Const1" if kind != "Const1" else "CyclicRandom", M.n)
    a2, _ = ref.lanczos_tridiag(other, 2)
    assert abs(a2[0] - r["alpha"][0]) > 1e-6


def test_restart_from_saved_tridiagonal(tmp_path):
    """Restart (restoreState, lanczos.c:1905-1960): a first run's
    tridiagonal, written in the reference's checkpoint layout and read
    back, makes a second solve skip the eigenvalue pass (the analyses are
    repeated at the first run's check points, so the same converged values
    are picked); it finds the same eigenpairs, eigenvectors bit for bit
    (the same start and products), with only the eigenvector pass's
    products."""
    from nemo3d_amd import ndio
    M = _models.gaas_box(cells=(2, 2, 2), strain_nm=0.01, seed=52)
    ref = oracle.assemble(M.arrays())
    w = np.linalg.eigvalsh(ref.dense())
    sel = w[(w > -1) & (w < 3)]
    k = int(np.argmax(np.diff(sel)))
    win = dict(erange_vb=(sel[k] - 1.0, sel[k] + 1e-6),
               erange_cb=(sel[k + 1] - 1e-6, sel[k + 1] + 1.0))
    H = Hamiltonian(M, layout=native.LAYOUT_SK)
    opts = dict(num_vb=2, num_cb=2, max_iter=4000, tol=1e-10, start_iter=100,
                skip=25, return_vectors=True, capacity=4, **win)
    r1 = eigensolve_lanczos(H, return_tridiag=True, **opts)
    assert r1["converged"]
    path = tmp_path / "t.tridiag"
    ndio.write_tridiagonal(path, r1["alpha"], r1["beta"])
    a, b, procs = ndio.read_tridiagonal(path)
    assert procs == 1 and np.array_equal(a, r1["alpha"])
    r2 = eigensolve_lanczos(H, restore=(a, b), **opts)
    assert r2["converged"]
    assert r2["iterations"] == r1["iterations"]
    assert np.array_equal(r2["eigenvalues"], r1["eigenvalues"])
    assert np.array_equal(r2["vectors"], r1["vectors"])
    for e, v in zip(r2["eigenvalues"], r2["vectors"]):
        assert np.linalg.norm(ref.hv(v) - e * v) < 1e-8
    assert r2["matvecs"] < r1["matvecs"] - len(a) // 2
