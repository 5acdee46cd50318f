"""NEMO-3D input decks: the XML reader and the structure a deck describes.

Reader.  A deck is the ``<group>`` / ``<param>`` tree of
NEMO_3D/XMLtoC/xml/ioNemo3D.xml, which the generated ``top_struct``
classes read (XMLtoC/apps/ioNemo3D/src).  ``read`` turns it into nested
dictionaries keyed by ``cTag``: ``struct`` groups become dicts, ``array``
groups lists of their ``obj`` entries (the ``newbranch`` template is not an
entry: the generated setters, e.g. Dev_struct::set in
XMLtoC/apps/ioNemo3D/src/Dev_struct.C:28-40, take only ``obj`` groups),
params typed values.  Numbers
convert as C's atoi / atof do (a leading number, else 0: an integer param
holding ``true`` reads as 0).  A tag repeated in one group (the material
records) becomes a list.

Structure.  ``build_structure`` follows the reference's set-up of a
zincblende crystal (setup_crystal_structure, NEMO_3D/nemo3d/src/base/
ham3d_final.c:622-730):

* the first shape, whatever its kind, fixes the cell range
  [1 + int(origin / a), 1 + int(origin / a) + int(length / a)) per axis,
  cells ordered x slowest, z fastest (FillFirstShape, ham3d.c:745-857);
* every shape, the first included, marks the atoms whose ideal position
  lattice * (cell + offset) lies inside it (the isNotIn* tests of
  ham3d.c:57-281 with the shape dimensions of ham3d.c:420-488: bounds
  inclusive; Dome and Ellipsoid as ellipsoids, Pyramid as four planes of
  slope 2 z / x and 2 z / y; Cylinder along z; Cylinder_x as written in
  the reference, its centre's y and z entries crossed); with cell
  granularity a cell is marked only when all eight of its atoms are
  (FillShape, ham3d.c:499-576);
* species are drawn with glibc rand() / RAND_MAX after srand(RandomSeed)
  (set_random_seed, nemo3d/src/math/random.c:50-58): for every cell of the
  domain one draw, and with AlloyTreatment RandomInUnitCell one more draw
  per marked atom, cations (atoms 0-3) before anions; the first species of
  a role is taken when the draw is <= its weight (ham3d.c:666-718);
* a shape's cations are its ``mat`` entries with pol >= 0, its anions
  those with pol <= 0; a role with two species weighs the first by its
  ``mol``, a role with one species weighs 1 (set_sMatList,
  nemo3d/src/io/i_mat.c:248-540);
* atoms without a bond are removed (remove_few_bonded_atoms(0),
  ham3d_final.c:693).

Not supported (NotImplementedError, never approximated): VCA alloys,
AlloyFixConcentration, strain-only shapes (use_for_strain_and_electronic
false), TiltedBox, Rhombohedron (wurtzite), RemoveSingleBonds and
HomogenizeSurface, StructFile input, periodic electronic boundaries.  The
strain block (StrainModel, Keating boundary) is reported by
``strain_options``; relaxing the structure is the caller's choice
(``vff.relax`` is the harmonic Keating model with fixed surface atoms).
"""
import ctypes
import re
import xml.etree.ElementTree as ET

import numpy as np

from . import materials, model, native, structure

_INT = re.compile(r"\s*[+-]?\d+")
_FLOAT = re.compile(r"\s*[+-]?(\d+\.?\d*([eE][+-]?\d+)?|\.\d+([eE][+-]?\d+)?)")


def _c_int(s):
    m = _INT.match(s)
    return int(m.group()) if m else 0


def _c_float(s):
    m = _FLOAT.match(s)
    return float(m.group()) if m else 0.0


def _items(text):
    return text.replace("{", " ").replace("}", " ").split()


def _param(p):
    t = p.get("type", "string")
    v = (p.findtext("value") or "").strip()
    if t == "real":
        return _c_float(v)
    if t in ("integer", "int"):
        return _c_int(v)
    if t == "boolean":
        return v.lower() == "true"
    if t == "array_real":
        return [_c_float(x) for x in _items(v)]
    if t == "array_integer":
        return [_c_int(x) for x in _items(v)]
    if t == "array_string":
        return _items(v)
    return v


def _group(g):
    if g.get("type") == "array":
        # the entries are the <group type="obj"> children; the newbranch
        # param only holds the template for new entries
        return [_group(ch) for ch in g if ch.tag == "group" and ch.get("type") == "obj"]
    d = {}
    for ch in g:
        if ch.tag not in ("param", "group"):
            continue
        if ch.tag == "param" and ch.get("type") in ("deletebranch", "newbranch"):
            continue
        tag = (ch.findtext("cTag") or "").strip()
        val = _param(ch) if ch.tag == "param" else _group(ch)
        if tag in d:
            if not isinstance(d[tag], list) or ch.tag == "param":
                d[tag] = [d[tag]]
            d[tag].append(val)
        else:
            d[tag] = val
    return d


def read(path):
    """The deck's top group as nested dictionaries."""
    root = ET.parse(path).getroot()
    top = root.find("input/group")
    if top is None:
        top = root if root.tag == "group" else root.find("group")
    if top is None:
        raise ValueError("%s: no <group> tree" % path)
    return _group(top)


def find(tree, *tags):
    """First value reached through the tags, searching nested groups depth
    first in deck order."""
    if not tags:
        return tree
    if isinstance(tree, dict):
        if tags[0] in tree:
            r = find(tree[tags[0]], *tags[1:])
            if r is not None:
                return r
        for v in tree.values():
            if isinstance(v, (dict, list)):
                r = find(v, *tags)
                if r is not None:
                    return r
    elif isinstance(tree, list):
        for v in tree:
            r = find(v, *tags)
            if r is not None:
                return r
    return None


def material_records(tree):
    """The deck's material records (the material_list entries holding
    mb_pv), in deck order."""
    recs = find(tree, "material_list") or []
    return [r for r in recs if isinstance(r, dict) and "mb_pv" in r]


def lanczos_options(tree):
    """The Lanczos block (Lanczos_struct) as eigensolve_lanczos keyword
    arguments."""
    lz = find(tree, "Lanczos")
    if lz is None:
        return None
    cyc = lz.get("CyclicString")
    return dict(num_vb=lz["NumEigReq_vb"], num_cb=lz["NumEigReq_cb"],
                erange_vb=tuple(lz["Erange_vb"]), erange_cb=tuple(lz["Erange_cb"]),
                max_iter=lz["MaxIter"], tol=lz["ConvTolerance"],
                start_iter=lz["ConvCheckStartIter"], skip=lz["ConvCheckSkipRate"],
                start=lz.get("Start", "CyclicRandom"),
                cyclic=tuple(cyc) if cyc else None)


def strain_options(tree):
    """StrainModel, Keating block and strain periodicity of the deck."""
    return dict(model=find(tree, "StrainModel"), keating=find(tree, "Keating"),
                periodicity=find(tree, "Periodicity_strain"))


# ---- shapes --------------------------------------------------------------

def _dims(sh):
    """shDim of FillShape (ham3d.c:420-488) for the supported shapes."""
    d = np.zeros(15)
    d[0], d[1], d[2] = sh["xorigin"], sh["yorigin"], sh["zorigin"]
    d[3], d[4], d[5] = (sh["xorigin"] + sh["x"], sh["yorigin"] + sh["y"],
                        sh["zorigin"] + sh["z"])
    kind = sh["Shape"]
    if kind == "Box":
        pass
    elif kind == "Cylinder_x":
        d[6] = (d[1] + d[4]) * 0.5
        d[7] = (d[2] + d[5]) * 0.5
        d[8] = sh["y"] * sh["y"] * 0.25
    elif kind == "Cylinder":
        d[6] = (d[0] + d[3]) * 0.5
        d[7] = (d[1] + d[4]) * 0.5
        d[8] = sh["x"] * sh["x"] * 0.25
    elif kind == "Pyramid":
        d[6] = 2.0 * sh["z"] / sh["x"]
        d[7] = 2.0 * sh["z"] / sh["y"]
    elif kind in ("Dome", "Ellipsoid"):
        d[6] = sh["xorigin"] + sh["x"] / 2.0
        d[7] = sh["yorigin"] + sh["y"] / 2.0
        d[8] = sh["zorigin"] + (sh["z"] / 2.0 if kind == "Ellipsoid" else 0.0)
        d[9] = 0.25 * sh["x"] * sh["x"]
        d[10] = 0.25 * sh["y"] * sh["y"]
        d[11] = (0.25 * sh["z"] * sh["z"] if kind == "Ellipsoid" else
                 sh["z"] * sh["z"])
    else:
        raise NotImplementedError("shape %s" % kind)
    return d


def outside(sh, pos):
    """isNotIn<Shape> (ham3d.c:57-281) of each position (nm): True where
    the atom is outside the shape."""
    d = _dims(sh)
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    kind = sh["Shape"]

    def out_of_box(ax):
        p = pos[:, ax]
        return (p < d[ax]) | (p > d[3 + ax])

    if kind == "Box":
        return out_of_box(0) | out_of_box(1) | out_of_box(2)
    if kind == "Cylinder_x":
        s = (z - d[6]) ** 2 + (y - d[7]) ** 2
        return out_of_box(0) | ((s - d[8]) > -1e-6)
    if kind == "Cylinder":
        s = (x - d[6]) ** 2 + (y - d[7]) ** 2
        return out_of_box(2) | (s > d[8])
    if kind == "Pyramid":
        return ((z < d[2]) | (z > d[2] + d[6] * (x - d[0])) |
                (z > d[2] - d[6] * (x - d[3])) |
                (z > d[2] + d[7] * (y - d[1])) |
                (z > d[2] - d[7] * (y - d[4])))
    # Dome / Ellipsoid
    s = (x - d[6]) ** 2 / d[9] + (y - d[7]) ** 2 / d[10] + (z - d[8]) ** 2 / d[11]
    return out_of_box(0) | out_of_box(1) | out_of_box(2) | (s > 1.0)


def composition(sh):
    """(cations, cation weight, anions, anion weight) of a shape."""
    mats, mols, pols = list(sh["mat"]), list(sh["mol"]), list(sh["pol"])
    if len(mats) == 1:
        return [mats[0]], 1.0, [mats[0]], 1.0
    cat = [(m, x) for m, x, p in zip(mats, mols, pols) if p >= 0]
    an = [(m, x) for m, x, p in zip(mats, mols, pols) if p <= 0]
    if len(cat) > 2 or len(an) > 2 or not cat or not an:
        raise ValueError("a shape holds one or two cations and anions: %s" % mats)
    npolar = sum(1 for p in pols if p != 0)
    if npolar != 2 and sh.get("AlloyTreatment") == "VCA":
        raise NotImplementedError("VCA alloys")
    wc = cat[0][1] if len(cat) == 2 else 1.0
    wa = an[0][1] if len(an) == 2 else 1.0
    return [m for m, _ in cat], wc, [m for m, _ in an], wa


def compounds(tree):
    """Binary materials the deck's shapes need (cation + anion names, one
    name for a monoatomic pair), in first-seen order (set_sMatList)."""
    out = []
    for sh in find(tree, "ShapeList"):
        cats, _, ans, _ = composition(sh)
        for c in cats:
            for a in ans:
                name = c + a if c != a else c
                if name not in out:
                    out.append(name)
    return out


class _Rand:
    """glibc rand() / RAND_MAX after srand(seed), drawn in order (the
    library's generator, nh_random_start)."""

    def __init__(self, seed):
        self.seed = int(seed)
        self.vals = np.zeros(0)
        self.used = 0

    def take(self, n):
        need = self.used + n
        if need > self.vals.size:
            m = max(need, 2 * self.vals.size, 1 << 20)
            out = np.empty(m, dtype=np.complex128)
            s = np.array([self.seed], dtype=np.int32)
            native.check(native.load().nh_random_start(
                native.RANDOM_RAND, native.ptr(s), ctypes.c_int64(m),
                native.ptr(out)))
            self.vals = out.real.copy()
        v = self.vals[self.used:need]
        self.used = need
        return v


def _check_supported(tree):
    if find(tree, "StructFile") not in (None, "NULL"):
        raise NotImplementedError("StructFile input")
    if find(tree, "CrystalStruct") not in (None, "Zincblende"):
        raise NotImplementedError("crystal structure %s" % find(tree, "CrystalStruct"))
    per = find(tree, "Periodicity")
    if per not in (None, "Closed"):
        raise NotImplementedError("periodicity %s" % per)
    for flag in ("RemoveSingleBonds", "HomogenizeSurface"):
        if find(tree, flag):
            raise NotImplementedError(flag)
    for sh in find(tree, "ShapeList") or []:
        if not sh.get("use_for_strain_and_electronic", True):
            raise NotImplementedError("strain-only shapes")
        if sh.get("AlloyFixConcentration", False):
            raise NotImplementedError("AlloyFixConcentration")


def build_structure(tree):
    """Atoms and species of the deck's shapes (zincblende decks)."""
    _check_supported(tree)
    a = np.array([find(tree, "a_lattice_x"), find(tree, "a_lattice_y"),
                  find(tree, "a_lattice_z")], dtype=np.float64)
    shapes = find(tree, "ShapeList")
    if not shapes:
        raise ValueError("the deck has no shapes")
    box = shapes[0]
    org = np.array([box["xorigin"], box["yorigin"], box["zorigin"]])
    ext = np.array([box["x"], box["y"], box["z"]])
    lo = 1 + np.trunc(org / a).astype(np.int64)
    hi = lo + np.trunc(ext / a).astype(np.int64)
    if (hi <= lo).any():
        raise ValueError("domain shape has zero volume")
    elements = []
    for sh in shapes:
        cats, _, ans, _ = composition(sh)
        for e in cats + ans:
            if e not in elements:
                elements.append(e)
    ii, jj, kk = np.meshgrid(np.arange(lo[0], hi[0]), np.arange(lo[1], hi[1]),
                             np.arange(lo[2], hi[2]), indexing="ij")
    cells = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1)
    nc = cells.shape[0]
    pos = ((np.repeat(cells, 8, axis=0) + structure.CELL_POS[np.tile(np.arange(8), nc)])
           * a)
    species = np.full(nc * 8, -1, dtype=np.int64)
    seed = find(tree, "RandomSeed")
    draws = _Rand(0 if seed is None else seed)
    slot_is_cation = np.arange(8) < 4
    for sh in shapes:
        cats, wc, ans, wa = composition(sh)
        ins = ~outside(sh, pos).reshape(nc, 8)
        if sh.get("has_cell_granularity", True):
            ins &= ins.all(axis=1, keepdims=True)
        per_atom = sh.get("AlloyTreatment") == "RandomInUnitCell"
        # one draw per cell, then (per-atom alloys) one per marked atom of
        # the cell in atom order
        n_cell = 1 + (ins.sum(axis=1) if per_atom else 0)
        vals = draws.take(int(np.sum(n_cell)))
        first = np.concatenate([[0], np.cumsum(n_cell)[:-1]])
        if per_atom:
            idx = first[:, None] + np.cumsum(ins, axis=1)
        else:
            idx = np.broadcast_to(first[:, None], ins.shape)
        rnd = vals[idx]
        ci = [elements.index(e) for e in cats] * (2 // len(cats))
        ai = [elements.index(e) for e in ans] * (2 // len(ans))
        pick = np.where(slot_is_cation[None, :],
                        np.where(rnd <= wc, ci[0], ci[1]),
                        np.where(rnd <= wa, ai[0], ai[1]))
        species = np.where(ins.ravel(), pick.ravel(), species)

    def occupancy(p, m, cell):
        return species

    return structure.build(lo, hi, occupancy, elements, a)


def build_model(tree, records=None):
    """TBModel of a deck: its band model, passivation, structure and
    material tables (the deck's own material records unless ``records`` is
    given)."""
    s = build_structure(tree)
    band = find(tree, "band_model")
    recs = material_records(tree) if records is None else records
    passivate = find(tree, "SurfacePassivation") == "PassivateDanglingBond"
    shift = find(tree, "DanglingBondEnergyShift")
    return model.TBModel(s, compounds(tree), band, passivate=passivate,
                         db_shift=30.0 if shift is None else shift, deck_records=recs)
