"""Keating valence-force-field relaxation (runs on the GPU).

Produces the displacements the strained Hamiltonian consumes, like the
reference's ``min_estrain`` (nemo3d/src/base/strain3d.c:157-475) with
``StrainModel = VFF_keating`` and ``Keating = FixSurfaceAll``: atoms on the
open surface of the box keep their ideal positions.
"""
import ctypes

import numpy as np

from . import native


def keating_tables(tbmodel):
    """alpha, beta, a0 [S][S] tables from the material records."""
    tab = tbmodel.tables
    S = len(tab.species)
    al = np.zeros((S, S))
    be = np.zeros((S, S))
    a0 = np.zeros((S, S))
    for comp, rec in tab.records.items():
        from .materials import split_compound
        c, a = (tab.index(e) for e in split_compound(comp))
        for i, j in ((c, a), (a, c)):
            al[i, j] = rec["strain_alpha"]
            be[i, j] = rec["strain_beta"]
            a0[i, j] = rec["unstrnd_cubic_cell_length"]
    return al, be, a0


class _Ctx:
    """Context holder for VFF-only work (no Hamiltonian build)."""

    def __init__(self, tbmodel):
        from .hamiltonian import Hamiltonian
        self.h = Hamiltonian(tbmodel, layout=None)
        self.ctx = self.h.ctx


def energy_grad(tbmodel, disp=None):
    al, be, a0 = keating_tables(tbmodel)
    c = _Ctx(tbmodel)
    A = tbmodel.n_atoms
    u = np.ascontiguousarray(tbmodel.structure.disp if disp is None else disp,
                             dtype=np.float64)
    g = np.zeros((A, 3))
    e = ctypes.c_double()
    native.check(native.load().nh_vff_energy_grad(
        c.ctx, native.ptr(al), native.ptr(be), native.ptr(a0), native.ptr(u),
        native.ptr(g), ctypes.byref(e)))
    return e.value, g


def relax(tbmodel, max_iter=4000, gtol=1e-6):
    """Relax tbmodel.structure.disp in place; returns a small info dict."""
    al, be, a0 = keating_tables(tbmodel)
    c = _Ctx(tbmodel)
    A = tbmodel.n_atoms
    out = np.zeros((A, 3))
    info = np.zeros(5)
    native.check(native.load().nh_vff_relax(
        c.ctx, native.ptr(al), native.ptr(be), native.ptr(a0), max_iter,
        float(gtol), native.ptr(out), native.ptr(info)))
    tbmodel.structure.disp = out
    return {"iterations": int(info[0]), "energy0": info[1], "energy": info[2],
            "grad0": info[3], "grad": info[4]}
