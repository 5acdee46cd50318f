"""Material parameter tables for the sp3d5s* Hamiltonian.

Host-side setup that mirrors the reference's material-parameter API:

* ``find_material`` follows ``find_Material_struct``
  (nemo3d/src/io/i_mat.c:184-224): first record in deck order whose name,
  band model and temperature match.
* ``param_init`` follows ``param_init``/``strnmat_init``
  (nemo3d/src/base/mat3d.c:105-307): ``parmat[cation][anion]`` holds the 35
  ``mb_pv`` values with ``Ev_offset`` added to, and the strain reference
  shift (``mb_strain[15]`` cation / ``[16]`` anion, 27 eV in the shipped
  records) subtracted from, the onsite energies; ``strnmat[cation][anion]``
  holds the 33 ``mb_strain`` values; a missing ``[anion][cation]`` entry is
  copied from ``[cation][anion]`` (mat3d.c:120-129, 293-306).

The records themselves live in ``data/materials.json`` (extracted from the
reference decks by ``data/extract_materials.py``).
"""
import json
import os
import re

import numpy as np

N_PARAM = 35
N_STRAIN = 33

BANDS_10_NOSPIN = "Bands_10_sp3d5ss_nospin"
BANDS_20_SPIN = "Bands_20_sp3d5ss_spin"

# Anion elements (geo_struct::isCation, nemo3d/src/base/geo_struct.h:598-610).
ANIONS = {"As", "Sb", "Se", "N", "O", "P", "S"}
# Group-IV elements take the cation role on atoms 0-3 of the cell.
GROUP_IV = {"Si", "Ge"}

_DB = None


def database():
    global _DB
    if _DB is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                            "data", "materials.json")
        with open(path) as f:
            _DB = json.load(f)["records"]
    return _DB


def find_material(name, band_model, source=None, spinless_view=False,
                  tkel=300.0, records=None):
    """First matching record in deck order (i_mat.c:184-224).

    ``spinless_view`` lets a ``Bands_10_sp3d5ss_nospin`` request use a
    ``Bands_20_sp3d5ss_spin`` record (same 35 sp3d5s* parameters; the
    spin-orbit entries are unused without spin) when the decks hold no
    spinless record for the material.  ``records``: search these records
    (a deck's own material list) instead of the shipped database.
    """
    models = [band_model]
    if spinless_view and band_model == BANDS_10_NOSPIN:
        models.append(BANDS_20_SPIN)
    for bm in models:
        for rec in (database() if records is None else records):
            if source is not None and rec.get("source") != source:
                continue
            if rec["material"] != name or rec["band_model"] != bm:
                continue
            if rec.get("comp1_mat", "NULL") != "NULL":
                continue
            if rec.get("tkel", 300.0) != tkel:
                continue
            return rec
    raise KeyError("no material record %s / %s" % (name, band_model))


def split_compound(name):
    """'GaAs' -> ('Ga', 'As'); 'Si' -> ('Si', 'Si')."""
    parts = re.findall(r"[A-Z][a-z]?", name)
    if len(parts) == 1:
        return parts[0], parts[0]
    if len(parts) != 2:
        raise ValueError("not a binary compound: %s" % name)
    return parts[0], parts[1]


class MaterialTables:
    """Dense per-species tables indexed [cation][anion]."""

    def __init__(self, species, parmat, strnmat, records):
        self.species = list(species)
        self.parmat = parmat
        self.strnmat = strnmat
        self.records = records

    def index(self, element):
        return self.species.index(element)


def param_init(compounds, band_model, source=None, spinless_view=True,
               include_strain=True, deck_records=None):
    """Build parmat/strnmat for a list of compound names (e.g. ['GaAs',
    'InAs']).  Species are indexed in first-seen order."""
    if not include_strain:
        raise ValueError("only include_strain_in_Hamiltonian=true is "
                         "supported (every reference deck sets it)")
    species = []
    pairs = []
    for comp in compounds:
        c, a = split_compound(comp)
        for el in (c, a):
            if el not in species:
                species.append(el)
        pairs.append((comp, c, a))
    S = len(species)
    parmat = np.zeros((S, S, N_PARAM))
    strnmat = np.zeros((S, S, N_STRAIN))
    have_strn = np.zeros((S, S), dtype=bool)
    records = {}
    for comp, c_el, a_el in pairs:
        rec = find_material(comp, band_model, source=source,
                            spinless_view=spinless_view, records=deck_records)
        records[comp] = rec
        c, a = species.index(c_el), species.index(a_el)
        pv = np.array(rec["mb_pv"], dtype=np.float64)
        st = np.array(rec["mb_strain"], dtype=np.float64)
        if pv.size != N_PARAM or st.size != N_STRAIN:
            raise ValueError("record %s is not an sp3d5s* record" % comp)
        shift_c = st[15]
        shift_a = st[16]
        strnmat[c, a] = st
        have_strn[c, a] = True
        if not have_strn[a, c]:
            strnmat[a, c] = st
            have_strn[a, c] = True
        p = pv.copy()
        ev = rec.get("Ev_offset", 0.0)
        for k in (0, 1, 4, 6):  # anion s, p, s*, d
            p[k] += ev - shift_a
        for k in (2, 3, 5, 7):  # cation s, p, s*, d
            p[k] += ev - shift_c
        parmat[c, a] = p
    # AB <-> BA copy keyed on the lattice constant (mat3d.c:293-306)
    for i in range(S):
        for j in range(i + 1, S):
            if parmat[i, j, 31] == parmat[j, i, 31]:
                continue
            if parmat[i, j, 31] == 0.0:
                parmat[i, j] = parmat[j, i]
            elif parmat[j, i, 31] == 0.0:
                parmat[j, i] = parmat[i, j]
            else:
                raise ValueError("lattice constants of AB and BA differ")
    return MaterialTables(species, parmat, strnmat, records)
