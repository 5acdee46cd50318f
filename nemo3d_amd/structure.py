"""Zincblende structure builder: atoms, species, neighbour table, bonds.

Produces the geometry arrays the Hamiltonian consumes, in the reference's
row order: atoms are numbered cell by cell (cells ordered with x slowest and
z fastest, the order of ``Reduce``/``cell__ijk``), atom m = 0..7 inside a
cell, absent atoms skipped (``get_total_hamiltonian_dimension_and_offsets``,
nemo3d/src/base/ham3d_final.c:547-568).  Cell geometry follows
nemo3d/src/base/Cell.c:184-210 (positions) and :599-713 (the neighbour
table ``nnmap``; its slot order decides the order of ``nbr`` below and is
kept so that neighbour indices agree with the reference's ``nd_nbrIndx``
output).

Shapes (box, lens/dome, truncated pyramid, wetting layer) and seeded
random alloys are this package's own synthetic inputs.
"""
import numpy as np

# Fractional positions of the 8 atoms of the conventional cell; 0-3 are the
# cation sublattice, 4-7 the anion sublattice (Cell.c:184-210).
CELL_POS = np.array([
    [0.0, 0.0, 0.0], [0.0, -0.5, -0.5], [-0.5, -0.5, 0.0], [-0.5, 0.0, -0.5],
    [-0.25, -0.25, -0.25], [-0.25, -0.75, -0.75], [-0.75, -0.75, -0.25],
    [-0.75, -0.25, -0.75]])

# Neighbour table: for atom m, slot n -> (di, dj, dk, atom) (Cell.c:599-713).
NNMAP = np.array([
    [[0, 0, 0, 4], [1, 0, 1, 7], [1, 1, 0, 6], [0, 1, 1, 5]],
    [[0, 0, 0, 5], [0, 0, 0, 4], [1, 0, 0, 6], [1, 0, 0, 7]],
    [[0, 0, 0, 4], [0, 0, 0, 6], [0, 0, 1, 5], [0, 0, 1, 7]],
    [[0, 0, 0, 4], [0, 0, 0, 7], [0, 1, 0, 5], [0, 1, 0, 6]],
    [[0, 0, 0, 0], [0, 0, 0, 1], [0, 0, 0, 2], [0, 0, 0, 3]],
    [[0, 0, 0, 1], [0, -1, 0, 3], [0, 0, -1, 2], [0, -1, -1, 0]],
    [[0, 0, 0, 2], [0, -1, 0, 3], [-1, 0, 0, 1], [-1, -1, 0, 0]],
    [[0, 0, 0, 3], [0, 0, -1, 2], [-1, 0, 0, 1], [-1, 0, -1, 0]]],
    dtype=np.int64)

# Ideal this->neighbour vector of each slot in units of a/4.
BOND_Q = np.rint(4.0 * (NNMAP[:, :, :3] + CELL_POS[NNMAP[:, :, 3]] -
                        CELL_POS[:, None, :])).astype(np.int32)


class Structure:
    """Atoms of a zincblende crystal in Hamiltonian row order.

    Attributes (A atoms):
      species   (A,) int32   index into `elements`
      is_cation (A,) int32   1 on the cation sublattice role
      nbr       (A,4) int32  neighbour atom per slot, -1 when dangling
      bond_q    (A,4,3) int32 ideal this->neighbour vector, units a/4
      disp      (A,3) float64 displacement from the ideal site (nm)
      cell      (A,3) int64  cell index (i, j, k)
      m         (A,) int64   atom index inside the cell
      lattice   (3,) float64 cubic cell lengths (nm)
      periodic  (3,) bool
      wrapped   (A,4,3) int8 cell-boundary crossings of each bond (periodic)
    """

    def __init__(self, elements, species, is_cation, nbr, bond_q, disp, cell,
                 m, lattice, periodic, wrapped):
        self.elements = list(elements)
        self.species = species
        self.is_cation = is_cation
        self.nbr = nbr
        self.bond_q = bond_q
        self.disp = disp
        self.cell = cell
        self.m = m
        self.lattice = np.asarray(lattice, dtype=np.float64)
        self.periodic = np.asarray(periodic, dtype=bool)
        self.wrapped = wrapped
        self.dot_region = None  # per atom: inside the dot / wetting layer

    @property
    def n_atoms(self):
        return int(self.species.shape[0])

    def positions(self, strained=True):
        p = (self.cell + CELL_POS[self.m]) * self.lattice
        return p + self.disp if strained else p

    def surface_mask(self):
        return (self.nbr < 0).any(axis=1)

    def cation_sites(self):
        return np.nonzero(self.is_cation)[0]


def build(cell_lo, cell_hi, occupancy, elements, lattice,
          periodic=(False, False, False), remove_unbonded=True,
          group_iv_split=None):
    """Build a structure over the cell range [cell_lo, cell_hi).

    occupancy(pos_nm (K,3), m (K,), cell (K,3)) -> element index per atom,
    -1 for absent.  Cation role: atoms whose element is not an anion; for
    group-IV elements the first four atoms of the cell
    (geo_struct::isCation, nemo3d/src/base/geo_struct.h:598-610).
    """
    from .materials import ANIONS, GROUP_IV
    lo = np.asarray(cell_lo, dtype=np.int64)
    hi = np.asarray(cell_hi, dtype=np.int64)
    dims = hi - lo
    lattice = np.asarray(lattice, dtype=np.float64)
    ii, jj, kk = np.meshgrid(np.arange(lo[0], hi[0]), np.arange(lo[1], hi[1]),
                             np.arange(lo[2], hi[2]), indexing="ij")
    cells = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1)
    ncell = cells.shape[0]
    cell_all = np.repeat(cells, 8, axis=0)
    m_all = np.tile(np.arange(8, dtype=np.int64), ncell)
    pos = (cell_all + CELL_POS[m_all]) * lattice
    el = np.asarray(occupancy(pos, m_all, cell_all), dtype=np.int64)
    present = el >= 0

    def _index_grid(present):
        idx = np.full(ncell * 8, -1, dtype=np.int64)
        idx[present] = np.arange(int(present.sum()))
        return idx.reshape(dims[0], dims[1], dims[2], 8)

    def _neighbours(present):
        grid = _index_grid(present)
        sel = np.nonzero(present)[0]
        c = cell_all[sel] - lo
        mm = m_all[sel]
        nb = np.full((sel.size, 4), -1, dtype=np.int64)
        wrapped = np.zeros((sel.size, 4, 3), dtype=np.int8)
        for n in range(4):
            off = NNMAP[mm, n, :3]
            am = NNMAP[mm, n, 3]
            t = c + off
            ok = np.ones(sel.size, dtype=bool)
            for d in range(3):
                if periodic[d]:
                    wrapped[:, n, d] = np.where(t[:, d] < 0, -1,
                                                np.where(t[:, d] >= dims[d], 1, 0))
                    t[:, d] %= dims[d]
                else:
                    ok &= (t[:, d] >= 0) & (t[:, d] < dims[d])
            tt = np.where(ok[:, None], t, 0)
            v = grid[tt[:, 0], tt[:, 1], tt[:, 2], am]
            nb[:, n] = np.where(ok, v, -1)
        return sel, nb, wrapped

    sel, nb, wrapped = _neighbours(present)
    if remove_unbonded:
        lonely = (nb < 0).all(axis=1)
        if lonely.any():
            present[sel[lonely]] = False
            sel, nb, wrapped = _neighbours(present)

    species = el[sel].astype(np.int32)
    mm = m_all[sel]
    is_cat = np.array([e not in ANIONS for e in elements])[species]
    iv = np.array([e in GROUP_IV for e in elements])[species]
    is_cat = np.where(iv, mm < 4, is_cat).astype(np.int32)
    bond_q = BOND_Q[mm].astype(np.int32)
    return Structure(elements, species, is_cat, nb.astype(np.int32), bond_q,
                     np.zeros((sel.size, 3)), cell_all[sel], mm, lattice,
                     periodic, wrapped)


def binary_box(cells, cation, anion, lattice, periodic=(False,) * 3,
               cell_lo=(1, 1, 1)):
    """Closed (or periodic) box of `cells` conventional cells of one binary
    compound (or one element when cation == anion)."""
    elements = [cation] if cation == anion else [cation, anion]
    ci, ai = 0, (0 if cation == anion else 1)

    def occ(pos, m, cell):
        return np.where(m < 4, ci, ai)

    lo = np.asarray(cell_lo)
    return build(lo, lo + np.asarray(cells), occ, elements, lattice, periodic)


def dot_structure(box_nm, lattice, shape="dome", base_nm=20.0, height_nm=5.0,
                  top_nm=None, wl_ml=1, dot_cation="In", host_cation="Ga",
                  anion="As", alloy_x=1.0, alloy_cation=None, seed=0,
                  base_z_nm=None):
    """Quantum dot (+ wetting layer) of dot_cation_x alloy_cation_(1-x) anion
    inside a host_cation-anion box.  Ideal (unrelaxed) positions.

    shape: 'dome' (spherical cap, base diameter base_nm, height height_nm)
    or 'pyramid' (square-based truncated pyramid, base side base_nm, top side
    top_nm).  The dot sits on a wl_ml-monolayer wetting layer whose bottom is
    at base_z_nm (default: 40 % of the box height).
    """
    lattice = np.asarray(lattice, dtype=np.float64)
    ncell = np.maximum(1, np.floor(np.asarray(box_nm) / lattice + 1e-9)).astype(int)
    elements = [host_cation, anion, dot_cation]
    if alloy_cation is not None and alloy_cation not in elements:
        elements.append(alloy_cation)
    box = ncell * lattice
    cx, cy = 0.5 * box[0], 0.5 * box[1]
    z0 = 0.4 * box[2] if base_z_nm is None else base_z_nm
    wl_top = z0 + wl_ml * 0.5 * lattice[2]
    rng = np.random.default_rng(seed)

    def in_region(pos):
        """Dot or wetting layer; pos measured from the cell-range origin
        (cells start at 1)."""
        p = pos - lattice * 1.0
        x, y, z = p[:, 0] - cx, p[:, 1] - cy, p[:, 2]
        if shape == "dome":
            R = 0.5 * base_nm
            rho = (R * R + height_nm * height_nm) / (2.0 * height_nm)
            zc = wl_top + height_nm - rho
            inside = (z >= wl_top) & (x * x + y * y + (z - zc) ** 2 <= rho * rho)
        elif shape == "pyramid":
            t = base_nm if top_nm is None else top_nm
            frac = np.clip((z - wl_top) / height_nm, 0.0, 1.0)
            half = 0.5 * (base_nm + (t - base_nm) * frac)
            inside = ((z >= wl_top) & (z <= wl_top + height_nm) &
                      (np.abs(x) <= half) & (np.abs(y) <= half))
        else:
            raise ValueError(shape)
        wl = (z >= z0) & (z < wl_top)
        return inside | wl

    def occ(pos, m, cell):
        cat = m < 4
        out = np.where(cat, 0, 1)
        in_dot = in_region(pos) & cat
        if alloy_cation is not None and alloy_x < 1.0:
            draw = rng.random(pos.shape[0]) < alloy_x
            out = np.where(in_dot & draw, 2, out)
            out = np.where(in_dot & ~draw, elements.index(alloy_cation), out)
        else:
            out = np.where(in_dot, 2, out)
        return out

    st = build((1, 1, 1), 1 + ncell, occ, elements, lattice)
    st.dot_region = in_region(st.positions(strained=False))
    return st
