"""Keating valence-force-field relaxation (runs on the GPU).

Produces the displacements the strained Hamiltonian consumes, like the
reference's ``min_estrain`` (nemo3d/src/base/strain3d.c:157-475).  Models
(``StrainModel``): ``VFF_keating`` (constant Keating alpha, beta) and
``VFF_keating_strained`` (alpha, beta scaled with the local strain by the
acoustic anharmonicity constants strain_VFF[0..2] of the material,
StrainParam.c:60-214).  Boundaries (``Keating``): ``FixSurfaceAll`` keeps
atoms with fewer than four neighbours at their ideal positions,
``AllFree`` moves every atom.
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


def anharmonic_tables(tbmodel):
    """aa, ba, ca [S][S] tables (strain_VFF[0..2] of each pair's record,
    StrainParam::SetMatParm, StrainParam.c:265-280)."""
    tab = tbmodel.tables
    S = len(tab.species)
    out = [np.zeros((S, S)) for _ in range(3)]
    for comp, rec in tab.records.items():
        from .materials import split_compound
        c, a = (tab.index(e) for e in split_compound(comp))
        v = rec.get("strain_VFF")
        if v is None or len(v) < 3:
            raise ValueError("record %s has no strain_VFF constants" % comp)
        for k in range(3):
            out[k][c, a] = out[k][a, c] = v[k]
    return out


MODELS = {"VFF_keating": native.VFF_KEATING,
          "VFF_keating_strained": native.VFF_KEATING_STRAINED}
BOUNDARIES = {"FixSurfaceAll": 1, "AllFree": 0}


def _params(tbmodel, model, boundary):
    if model not in MODELS:
        raise ValueError("unsupported strain model %s" % model)
    if boundary not in BOUNDARIES:
        raise ValueError("unsupported Keating boundary %s" % boundary)
    al, be, a0 = keating_tables(tbmodel)
    tabs = [al, be, a0]
    if model == "VFF_keating_strained":
        tabs += anharmonic_tables(tbmodel)
    tabs = [np.ascontiguousarray(t, dtype=np.float64) for t in tabs]
    p = native.NHVffParams()
    for name, t in zip(("alpha", "beta", "a0", "aa", "ba", "ca"), tabs):
        setattr(p, name, t.ctypes.data)
    p.model = MODELS[model]
    p.fix_surface = BOUNDARIES[boundary]
    return p, tabs  # the tables must outlive the call


class _Ctx:
    """Context holder for VFF-only work (no Hamiltonian build)."""

    def __init__(self, tbmodel):
        from .hamiltonian import Hamiltonian
        self.h = Hamiltonian(tbmodel, layout=None)
        self.ctx = self.h.ctx


def energy_grad(tbmodel, disp=None, model="VFF_keating", boundary="FixSurfaceAll"):
    """Energy and gradient (for VFF_keating_strained: the gradient with the
    scaled constants held fixed, as the reference forms it)."""
    p, keep = _params(tbmodel, model, boundary)
    c = _Ctx(tbmodel)
    A = tbmodel.n_atoms
    u = np.ascontiguousarray(tbmodel.structure.disp if disp is None else disp,
                             dtype=np.float64)
    g = np.zeros((A, 3))
    e = ctypes.c_double()
    native.check(native.load().nh_vff_energy_grad_model(
        c.ctx, ctypes.byref(p), native.ptr(u), native.ptr(g), ctypes.byref(e)))
    del keep
    return e.value, g


def relax(tbmodel, max_iter=4000, gtol=1e-6, model="VFF_keating",
          boundary="FixSurfaceAll"):
    """Relax tbmodel.structure.disp in place (from its current value);
    returns a small info dict."""
    p, keep = _params(tbmodel, model, boundary)
    c = _Ctx(tbmodel)
    A = tbmodel.n_atoms
    out = np.zeros((A, 3))
    info = np.zeros(5)
    native.check(native.load().nh_vff_relax_model(
        c.ctx, ctypes.byref(p), int(max_iter), float(gtol), native.ptr(out),
        native.ptr(info)))
    del keep
    tbmodel.structure.disp = out
    return {"iterations": int(info[0]), "energy0": info[1], "energy": info[2],
            "grad0": info[3], "grad": info[4]}
