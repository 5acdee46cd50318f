"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
use this module, always as the checker.  See nemo_oracle.h for what each
function restates and how it is pinned.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ONSITE_STRIDE = 17


class ORModel(ctypes.Structure):
    _fields_ = [
        ("nb", ctypes.c_int),
        ("n_atoms", ctypes.c_int),
        ("n_species", ctypes.c_int),
        ("parmat", ctypes.c_void_p),
        ("strnmat", ctypes.c_void_p),
        ("species", ctypes.c_void_p),
        ("is_cation", ctypes.c_void_p),
        ("nbr", ctypes.c_void_p),
        ("bond_q", ctypes.c_void_p),
        ("disp", ctypes.c_void_p),
        ("lattice", ctypes.c_double * 3),
        ("passivate", ctypes.c_int),
        ("db_shift", ctypes.c_double),
        ("phi", ctypes.c_void_p),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        L.or_assemble.argtypes = [ctypes.POINTER(ORModel), P, P]
        L.or_assemble.restype = ctypes.c_int
        L.or_hv.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P, P,
                            ctypes.c_double, P, P]
        L.or_sk_bond.argtypes = [P, P, ctypes.c_double, P, P, P, P, P]
        L.or_onsite.argtypes = [P, ctypes.c_int, P, P]
        L.or_lanczos_tridiag.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P,
                                         P, ctypes.c_int, P, P]
        L.or_tridiag_eigs.argtypes = [ctypes.c_int, P, P, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, P,
                                      ctypes.c_int]
        L.or_tridiag_eigs.restype = ctypes.c_int
        L.or_set_threads.argtypes = [ctypes.c_int]
        L.or_get_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def set_threads(n):
    """Threads of the oracle's OpenMP loops (n <= 0: every processor)."""
    lib().or_set_threads(int(n))


def get_threads():
    return int(lib().or_get_threads())


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleHamiltonian:
    """Assembled Hamiltonian in the oracle's row-owner block layout."""

    def __init__(self, nb, nbr, onsite, blocks, nsurf):
        self.nb = nb
        self.nbr = nbr
        self.onsite = onsite
        self.blocks = blocks
        self.nsurf = nsurf
        self.n_atoms = nbr.shape[0]
        self.n = self.n_atoms * nb

    def hv(self, x, shift=0.0, phase=None):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.empty_like(x)
        ph = None if phase is None else np.ascontiguousarray(phase, dtype=np.complex128)
        lib().or_hv(self.nb, self.n_atoms, _p(self.onsite), _p(self.blocks),
                    _p(self.nbr), None if ph is None else _p(ph),
                    float(shift), _p(x), _p(y))
        return y

    def dense(self, phase=None):
        n = self.n
        H = np.zeros((n, n), dtype=np.complex128)
        e = np.zeros(n, dtype=np.complex128)
        for k in range(n):
            e[:] = 0
            e[k] = 1
            H[:, k] = self.hv(e, phase=phase)
        return H

    def lanczos_tridiag(self, r0, iters):
        r0 = np.ascontiguousarray(r0, dtype=np.complex128)
        a = np.zeros(iters)
        b = np.zeros(iters)
        lib().or_lanczos_tridiag(self.nb, self.n_atoms, _p(self.onsite),
                                 _p(self.blocks), _p(self.nbr), _p(r0), iters,
                                 _p(a), _p(b))
        return a, b


def assemble(model):
    """Assemble with the oracle.  `model` is a dict of numpy arrays as made
    by nemo3d_amd.hamiltonian.model_arrays()."""
    A = int(model["n_atoms"])
    keep = {k: np.ascontiguousarray(model[k], dtype=dt) for k, dt in [
        ("parmat", np.float64), ("strnmat", np.float64), ("species", np.int32),
        ("is_cation", np.int32), ("nbr", np.int32), ("bond_q", np.int32),
        ("disp", np.float64)]}
    m = ORModel()
    m.nb = int(model["nb"])
    m.n_atoms = A
    m.n_species = int(model["n_species"])
    for k in keep:
        setattr(m, k, _p(keep[k]))
    for d in range(3):
        m.lattice[d] = float(model["lattice"][d])
    m.passivate = int(model["passivate"])
    m.db_shift = float(model["db_shift"])
    phi = model.get("phi")
    if phi is not None:
        keep["phi"] = np.ascontiguousarray(phi, dtype=np.float64)
        m.phi = _p(keep["phi"])
    onsite = np.zeros((A, ONSITE_STRIDE))
    blocks = np.zeros((A, 4, 10, 10))
    nsurf = lib().or_assemble(ctypes.byref(m), _p(onsite), _p(blocks))
    return OracleHamiltonian(m.nb, keep["nbr"], onsite, blocks, nsurf)


def sk_bond(par, strn, a0, nnv, nnv0):
    par = np.ascontiguousarray(par, dtype=np.float64)
    strn = None if strn is None else np.ascontiguousarray(strn, dtype=np.float64)
    nnv = np.ascontiguousarray(nnv, dtype=np.float64)
    nnv0 = np.ascontiguousarray(nnv0, dtype=np.float64)
    T = np.zeros((10, 10))
    ca = np.zeros(10)
    cc = np.zeros(10)
    lib().or_sk_bond(_p(par), None if strn is None else _p(strn), float(a0),
                     _p(nnv), _p(nnv0), _p(T), _p(ca), _p(cc))
    return T, ca, cc


def onsite(par, anion):
    """or_onsite: spatial diagonal energies (10) and the spin-orbit
    constant of an (averaged) parameter set; anion selects the anion
    energies."""
    par = np.ascontiguousarray(par, dtype=np.float64)
    d = np.zeros(10)
    lam = ctypes.c_double()
    lib().or_onsite(_p(par), int(anion), _p(d), ctypes.byref(lam))
    return d, lam.value


def onsite_block(nb, diag, lam, sp=None):
    """Dense nb x nb onsite block of one atom with the given onsite record
    (diagonal, surface s-p couplings, spin-orbit constant), applied through
    or_hv to unit vectors."""
    rec = np.zeros((1, ONSITE_STRIDE))
    rec[0, :10] = diag
    if sp is not None:
        rec[0, 10:16] = sp
    rec[0, 16] = lam
    H = OracleHamiltonian(nb, -np.ones((1, 4), dtype=np.int32), rec,
                          np.zeros((1, 4, 10, 10)), 0)
    return H.dense()


def tridiag_eigs(d, e, lo, hi, tol=1e-15, wmax=100000):
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    w = np.zeros(min(wmax, len(d)))
    n = lib().or_tridiag_eigs(len(d), _p(d), _p(e), lo, hi, tol, _p(w), len(w))
    return w[:n]
