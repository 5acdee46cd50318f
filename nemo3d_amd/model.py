"""Tight-binding model description: structure + material tables + options.

`TBModel` gathers what the reference's ``setup_hamiltonian``
(nemo3d/src/base/ham3d_final.c:731-1230) hands to the Hamiltonian
assembly: band model (``basis_init``, ham3d_final.c:576-620), material
tables (``param_init``), the atom list in row order with its neighbour
table, bond vectors and displacements, and the surface-passivation options
of ``Phys`` (``SurfacePassivation``, ``DanglingBondEnergyShift``).
"""
import numpy as np

from . import materials as mat

BAND_NB = {mat.BANDS_10_NOSPIN: 10, mat.BANDS_20_SPIN: 20}


class TBModel:
    def __init__(self, structure, compounds, band_model,
                 passivate=True, db_shift=30.0, source=None, phi=None,
                 deck_records=None):
        if band_model not in BAND_NB:
            raise ValueError("unsupported band model %s" % band_model)
        self.structure = structure
        self.band_model = band_model
        self.nb = BAND_NB[band_model]
        self.tables = mat.param_init(compounds, band_model, source=source,
                                     deck_records=deck_records)
        self.passivate = bool(passivate)
        self.db_shift = float(db_shift)
        # per-atom electrostatic potential subtracted from the onsite
        # diagonal (Phys.PotentialFile, ham3d_final.c:913-927)
        self.phi = None if phi is None else np.asarray(phi, dtype=np.float64)
        # structure element index -> table species index
        used = set(np.unique(structure.species).tolist())
        remap = np.array([self.tables.index(e) if k in used else -1
                          for k, e in enumerate(structure.elements)],
                         dtype=np.int32)
        self.species = remap[structure.species]

    @property
    def n_atoms(self):
        return self.structure.n_atoms

    @property
    def n(self):
        return self.n_atoms * self.nb

    def arrays(self):
        s = self.structure
        return {
            "nb": self.nb,
            "n_atoms": s.n_atoms,
            "n_species": len(self.tables.species),
            "parmat": np.ascontiguousarray(self.tables.parmat),
            "strnmat": np.ascontiguousarray(self.tables.strnmat),
            "species": np.ascontiguousarray(self.species, dtype=np.int32),
            "is_cation": np.ascontiguousarray(s.is_cation, dtype=np.int32),
            "nbr": np.ascontiguousarray(s.nbr, dtype=np.int32),
            "bond_q": np.ascontiguousarray(s.bond_q, dtype=np.int32),
            "disp": np.ascontiguousarray(s.disp, dtype=np.float64),
            "lattice": np.asarray(s.lattice, dtype=np.float64),
            "passivate": int(self.passivate),
            "db_shift": self.db_shift,
            "phi": self.phi,
        }
