"""Reference-signature operator callbacks (matvect_mult, matmat_mult,
Hmult_spds_complete, Hmatmult_spds_complete) against the oracle."""
import ctypes

import numpy as np
import pytest

import _models
from nemo3d_amd import native
from nemo3d_amd.hamiltonian import Hamiltonian
from oracle import oracle

pytestmark = pytest.mark.gpu


def _arglist(*objs):
    """argList: array of pointers to the given ctypes objects (None: null)."""
    return (ctypes.c_void_p * len(objs))(
        *[None if o is None else ctypes.addressof(o) for o in objs])


def test_shim_operators_match_oracle():
    M = _models.gaas_box(cells=(2, 2, 2), strain_nm=0.01, seed=51)
    ref = oracle.assemble(M.arrays())
    H = Hamiltonian(M)
    lib = native.load()
    native.check(lib.nh_shim_bind(H.ctx))
    rng = np.random.default_rng(5)
    n = H.n
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    shift = ctypes.c_double(0.7)
    nproj = ctypes.c_int(0)
    y = np.zeros(n, dtype=np.complex128)
    args = _arglist(None, shift, nproj)
    print("matvect_mult", n, flush=True)
    lib.nh_matvect_mult(args, native.ptr(y), native.ptr(x))
    want = ref.hv(x, 0.7)
    assert np.linalg.norm(y - want) <= 1e-12 * np.linalg.norm(want)

    # spectral projection through the registered wave-function table
    psi = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
    E = np.array([0.3, -1.1, 2.5])
    native.check(lib.nh_shim_set_wavefunctions(3, native.ptr(E), native.ptr(psi)))
    which = (ctypes.c_int * 2)(2, 0)
    which_p = ctypes.cast(which, ctypes.POINTER(ctypes.c_int))
    nproj2 = ctypes.c_int(2)
    args2 = _arglist(None, shift, nproj2, which_p)
    lib.nh_matvect_mult(args2, native.ptr(y), native.ptr(x))
    want2 = want.copy()
    for j in (2, 0):
        want2 -= E[j] * np.vdot(psi[j], x) * psi[j]
    assert np.linalg.norm(y - want2) <= 1e-12 * np.linalg.norm(want2)

    # block products on row-pointer matrices (one row per vector)
    X = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
    Y = np.zeros((3, n), dtype=np.complex128)
    rows = lambda A: (ctypes.c_void_p * A.shape[0])(
        *[A.ctypes.data + i * A.strides[0] for i in range(A.shape[0])])
    ncol = ctypes.c_int(3)
    args3 = _arglist(None, shift, ncol)
    lib.nh_matmat_mult(args3, rows(Y), rows(X))
    for j in range(3):
        w = ref.hv(X[j], 0.7)
        assert np.linalg.norm(Y[j] - w) <= 1e-12 * np.linalg.norm(w)
    Y[:] = 0
    lib.nh_Hmatmult_spds_complete(rows(Y), None, -0.2, rows(X), 3)
    for j in range(3):
        w = ref.hv(X[j], -0.2)
        assert np.linalg.norm(Y[j] - w) <= 1e-12 * np.linalg.norm(w)
    y[:] = 0
    lib.nh_Hmult_spds_complete(native.ptr(y), None, 0.0, native.ptr(x))
    w = ref.hv(x, 0.0)
    assert np.linalg.norm(y - w) <= 1e-12 * np.linalg.norm(w)
    native.check(lib.nh_shim_set_wavefunctions(0, None, None))
    native.check(lib.nh_shim_bind(None))
