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
            fd = (vff_oracle.energy(M.structure, M.species, al, be, a0, up) -
                  vff_oracle.energy(M.structure, M.species, al, be, a0, dn)) / (2 * h)
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
