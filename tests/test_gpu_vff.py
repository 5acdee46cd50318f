"""GPU Keating VFF against the CPU energy oracle."""
import numpy as np
import pytest

import _models
from nemo3d_amd import materials, model, native, structure, vff
from oracle import vff_oracle

pytestmark = pytest.mark.gpu


def dot_box(cells=(4, 4, 3), seed=0, jitter=0.0):
    a = 0.56532
    rng = np.random.default_rng(seed)
    elements = ["Ga", "As", "In"]
    c = np.asarray(cells)

    def occ(pos, m, cell):
        cat = m < 4
        core = (np.abs(cell - (1 + c / 2.0)) <= 1.0).all(axis=1)
        return np.where(cat, np.where(core, 2, 0), 1)

    s = structure.build((1, 1, 1), 1 + c, occ, elements, (a, a, a))
    if jitter:
        s.disp = rng.normal(0, jitter, (s.n_atoms, 3))
        s.disp[s.surface_mask()] = 0.0
    return model.TBModel(s, ["GaAs", "InAs"], materials.BANDS_20_SPIN)


def test_energy_and_gradient_match_oracle():
    """The energy as the reference reports it (each bond term from both of
    its atoms) and the reference's gradient (each atom differentiates its
    own copy of a bond term: the derivative of the energy with each bond
    term once) against the oracle and its finite differences."""
    M = dot_box(jitter=0.01)
    al, be, a0 = vff.keating_tables(M)
    e, g = vff.energy_grad(M)
    e_ref = vff_oracle.energy(M.structure, M.species, al, be, a0)
    assert abs(e - e_ref) <= 1e-12 * abs(e_ref)
    free = ~M.structure.surface_mask()
    assert (g[~free] == 0).all()
    rng = np.random.default_rng(1)
    u0 = M.structure.disp.copy()
    h = 1e-6
    for i in rng.choice(np.nonzero(free)[0], 12, replace=False):
        for k in range(3):
            up, dn = u0.copy(), u0.copy()
            up[i, k] += h
            dn[i, k] -= h
            fd = (vff_oracle.energy(M.structure, M.species, al, be, a0, up, bond_copies=1) -
                  vff_oracle.energy(M.structure, M.species, al, be, a0, dn, bond_copies=1)) / (2 * h)
            assert abs(g[i, k] - fd) <= 1e-6 * (abs(fd) + 1e-3 * np.abs(g).max()), (i, k, g[i, k], fd)


def test_ideal_crystal_is_stationary():
    a = 0.56532
    s = structure.binary_box((3, 3, 3), "Ga", "As", (a, a, a))
    M = model.TBModel(s, ["GaAs"], materials.BANDS_20_SPIN)
    e, g = vff.energy_grad(M)
    assert abs(e) < 1e-20 and np.abs(g).max() < 1e-12


def test_relaxation_of_inas_inclusion():
    M = dot_box(cells=(5, 5, 4))
    al, be, a0 = vff.keating_tables(M)
    info = vff.relax(M, max_iter=3000, gtol=1e-9)
    print(info, flush=True)
    assert info["energy"] < 0.8 * info["energy0"]
    assert info["grad"] <= 1e-9 * 1.0001 or info["grad"] < 1e-6 * info["grad0"]
    e_ref = vff_oracle.energy(M.structure, M.species, al, be, a0)
    assert abs(e_ref - info["energy"]) <= 1e-10 * info["energy0"]
    # In-As bonds stretch towards InAs, Ga-As bonds near the inclusion stay
    # shorter: mean In-As bond length exceeds the mean Ga-As one
    s = M.structure
    p = s.positions()
    ln, isin = [], []
    for i in np.nonzero(s.is_cation)[0]:
        for j in s.nbr[i]:
            if j >= 0:
                ln.append(np.linalg.norm(p[j] - p[i]))
                isin.append(s.species[i] == 2)
    ln, isin = np.array(ln), np.array(isin)
    assert ln[isin].mean() > ln[~isin].mean() + 0.003


def test_relax_after_nan_filled_device_memory():
    """Relaxation starts from fresh search directions whatever the recycled
    device memory holds: device buffers are filled with NaN and released
    first (the first conjugate-gradient direction once read memory it had
    not written, 0 * NaN), then the relaxation must stay finite and match a
    relaxation of the same structure in clean memory."""
    M1 = _models.gaas_box(cells=(3, 3, 3), strain_nm=0.02, seed=41)
    ref = vff.relax(M1)
    M2 = _models.gaas_box(cells=(3, 3, 3), strain_nm=0.02, seed=41)
    from nemo3d_amd.hamiltonian import Hamiltonian
    owner = Hamiltonian(M2, layout=None)
    junk = [native.DeviceArray.from_host(owner, np.full(M2.n_atoms * 3 * 8, np.nan))
            for _ in range(8)]
    del junk
    info = vff.relax(M2)
    assert np.isfinite(M2.structure.disp).all()
    assert info["iterations"] == ref["iterations"]
    assert np.abs(M2.structure.disp - M1.structure.disp).max() < 1e-12


def _strained_tables(M):
    al, be, a0 = vff.keating_tables(M)
    aa, ba, ca = vff.anharmonic_tables(M)
    return dict(alpha=al, beta=be, a0=a0, aa=aa, ba=ba, ca=ca)


@pytest.mark.parametrize("boundary", ["AllFree", "FixSurfaceAll"])
def test_strained_keating_energy_and_gradient_match_oracle(boundary):
    """VFF_keating_strained (strain-scaled alpha and beta): the GPU energy
    equals the oracle's, and the GPU gradient equals the central finite
    differences of the oracle energy with the scaled constants frozen at the
    current displacements (how calc_estrain forms its gradient).  Large
    random displacements reach the truncated branches of beta_eff.  With
    AllFree the surface atoms carry a gradient too."""
    M = dot_box(jitter=0.02, seed=5)
    t = _strained_tables(M)
    s = M.structure
    u0 = s.disp.copy()
    kw = dict(model="VFF_keating_strained", aa=t["aa"], ba=t["ba"], ca=t["ca"])
    e, g = vff.energy_grad(M, model="VFF_keating_strained", boundary=boundary)
    e_ref = vff_oracle.energy(s, M.species, t["alpha"], t["beta"], t["a0"], **kw)
    e_harm = vff_oracle.energy(s, M.species, t["alpha"], t["beta"], t["a0"])
    print("strained E %.6e oracle %.6e harmonic %.6e" % (e, e_ref, e_harm), flush=True)
    assert abs(e - e_ref) <= 1e-12 * abs(e_ref)
    assert abs(e - e_harm) > 1e-3 * e_harm  # the scaling is in effect
    surf = s.surface_mask()
    if boundary == "FixSurfaceAll":
        assert (g[surf] == 0).all()
    else:
        assert np.abs(g[surf]).max() > 0
    rng = np.random.default_rng(2)
    h = 1e-6
    movable = np.nonzero(~surf if boundary == "FixSurfaceAll" else np.ones(len(surf), bool))[0]
    picks = list(rng.choice(movable, 10, replace=False))
    if boundary == "AllFree":
        picks += list(rng.choice(np.nonzero(surf)[0], 4, replace=False))
    for i in picks:
        for k in range(3):
            up, dn = u0.copy(), u0.copy()
            up[i, k] += h
            dn[i, k] -= h
            fd = (vff_oracle.energy(s, M.species, t["alpha"], t["beta"], t["a0"], up,
                                    frozen=u0, bond_copies=1, **kw) -
                  vff_oracle.energy(s, M.species, t["alpha"], t["beta"], t["a0"], dn,
                                    frozen=u0, bond_copies=1, **kw)) / (2 * h)
            assert abs(g[i, k] - fd) <= 1e-6 * (abs(fd) + 1e-3 * np.abs(g).max()), (i, k, g[i, k], fd)


def test_strained_keating_relaxation_is_stationary():
    """A relaxation under VFF_keating_strained / AllFree ends where the
    reference's gradient (frozen constants) vanishes, checked by the
    oracle's finite differences at the result; its In-As bonds are longer
    than under the harmonic model's fixed surface only through the model
    (both relax the same inclusion)."""
    M = dot_box(cells=(5, 5, 4))
    t = _strained_tables(M)
    info = vff.relax(M, max_iter=6000, gtol=1e-10, model="VFF_keating_strained",
                     boundary="AllFree")
    print(info, flush=True)
    assert info["grad"] <= 1e-10 * 1.0001
    s = M.structure
    u = s.disp.copy()
    kw = dict(model="VFF_keating_strained", aa=t["aa"], ba=t["ba"], ca=t["ca"])
    rng = np.random.default_rng(3)
    h = 1e-6
    for i in rng.choice(s.n_atoms, 6, replace=False):
        for k in range(3):
            up, dn = u.copy(), u.copy()
            up[i, k] += h
            dn[i, k] -= h
            fd = (vff_oracle.energy(s, M.species, t["alpha"], t["beta"], t["a0"], up,
                                    frozen=u, bond_copies=1, **kw) -
                  vff_oracle.energy(s, M.species, t["alpha"], t["beta"], t["a0"], dn,
                                    frozen=u, bond_copies=1, **kw)) / (2 * h)
            assert abs(fd) < 1e-7, (i, k, fd)
