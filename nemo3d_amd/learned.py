"""Learned warm start: features, grid mapping and the network driver.

The network is a 3-D convolutional net on the cation sublattice of the
zincblende crystal.  Grid spacing is a/2: the cation of cell (i, j, k),
atom m in 0..3, sits at grid point 2*(i, j, k) + 2*CELL_POS[m], and the
anion m + 4 of the same cell (its neighbour along -(1,1,1)a/4) shares that
point.  Every point therefore carries up to 2 atoms x 20 orbitals.

Inputs per grid point (C_IN channels): cation species one-hot over the
cation elements of the model, anion present, the 6 components of the local
strain tensor at the cation (least-squares fit of its four bond vectors to
the ideal ones), a dot indicator, and the position relative to the dot
centroid scaled by the dot size (3 channels).
Outputs per grid point: N_STATES x 40 complex amplitudes (cation orbitals
0-19, anion orbitals 0-19), the start block of the block eigensolvers.

The start-vector hooks it replaces are the reference's random starts
(lanczos_blk.c:2214-2229, qtracemin_driver.c:112-116) and the file start
of Lanczos (Start = ReadFile, eigsys3d_par.c:154-188).
"""
import ctypes

import numpy as np

from . import native
from .structure import CELL_POS

N_STATES = 8
N_AMP = 40  # 2 atoms x 20 orbitals per grid point


class GridMap:
    """Atom <-> grid-point mapping for one structure."""

    def __init__(self, structure):
        s = structure
        m = s.m % 4
        g = 2 * s.cell + np.rint(2 * CELL_POS[m]).astype(np.int64)
        self.origin = g.min(axis=0)
        g = g - self.origin
        self.dims = tuple(int(d) for d in g.max(axis=0) + 1)
        gx, gy, gz = self.dims
        self.point = ((g[:, 0] * gy + g[:, 1]) * gz + g[:, 2]).astype(np.int32)
        self.slot = (s.m >= 4).astype(np.int32)  # 0 cation, 1 anion
        self.n_points = gx * gy * gz


def local_strain(structure):
    """Strain tensor at each cation from its bonds: R = (1 + e) R0,
    least squares over the present bonds (directions no bond constrains are
    left unstrained), symmetrised.  Returns (A, 6):
    exx, eyy, ezz, exy, exz, eyz (zero for anions)."""
    s = structure
    A = s.n_atoms
    p = s.positions()
    out = np.zeros((A, 6))
    cat = np.nonzero(s.is_cation)[0]
    nb = s.nbr[cat]
    ok = nb >= 0
    R0 = 0.25 * s.bond_q[cat] * s.lattice  # ideal host-lattice bond vectors
    R = np.where(ok[..., None], p[np.where(ok, nb, 0)] - p[cat][:, None, :], 0.0)
    R0 = np.where(ok[..., None], R0, 0.0)
    M = np.einsum("bni,bnj->bij", R, R0)
    G = np.einsum("bni,bnj->bij", R0, R0)
    # pseudo-inverse: directions without bonds (surface atoms) get no strain
    Gp = np.linalg.pinv(G, rcond=1e-9, hermitian=True)
    F = M @ Gp + (np.eye(3) - G @ Gp)
    e = 0.5 * (F + np.swapaxes(F, 1, 2)) - np.eye(3)
    out[cat] = np.stack([e[:, 0, 0], e[:, 1, 1], e[:, 2, 2], e[:, 0, 1],
                         e[:, 0, 2], e[:, 1, 2]], axis=1)
    return out


def features(tbmodel, gmap, dot_mask=None):
    """(n_points, C_IN) float32 features.  dot_mask: per-atom bool (atoms
    of the dot/wetting layer); default: the structure's dot region, or
    else every cation that is not Ga."""
    s = tbmodel.structure
    elements = s.elements
    cat_el = sorted({elements[k] for k in np.unique(s.species)
                     if s.is_cation[np.nonzero(s.species == k)[0][0]]})
    C = len(cat_el) + 1 + 6 + 1 + 3
    f = np.zeros((gmap.n_points, C), dtype=np.float32)
    cat = s.is_cation.astype(bool)
    for k, el in enumerate(cat_el):
        sel = cat & (np.array(elements)[s.species] == el)
        f[gmap.point[sel], k] = 1.0
    f[gmap.point[~cat], len(cat_el)] = 1.0
    e = local_strain(s)
    f[gmap.point[cat], len(cat_el) + 1:len(cat_el) + 7] = (e[cat] * 10.0).astype(np.float32)
    if dot_mask is None:
        if s.dot_region is not None:
            dot_mask = cat & s.dot_region
        else:
            dot_mask = cat & (np.array(elements)[s.species] != "Ga")
    f[gmap.point[dot_mask], len(cat_el) + 7] = 1.0
    pos = s.positions()
    if dot_mask.any():
        c0 = pos[dot_mask].mean(axis=0)
        ext = np.maximum(pos[dot_mask].max(axis=0) - pos[dot_mask].min(axis=0), 1.0)
    else:
        c0 = pos.mean(axis=0)
        ext = np.ones(3)
    rel = (pos - c0) / ext
    f[gmap.point, len(cat_el) + 8:len(cat_el) + 11] = rel.astype(np.float32)
    return f


def kramers_partner(psi, nb):
    """Time-reversed state (spin up u, spin down d per atom) -> (-d*, u*):
    degenerate with psi and orthogonal to it for the spin-orbit
    Hamiltonian (20-band model)."""
    assert nb == 20
    v = np.asarray(psi).reshape(-1, 2, 10)
    out = np.empty_like(v)
    out[:, 0] = -np.conj(v[:, 1])
    out[:, 1] = np.conj(v[:, 0])
    return out.reshape(np.shape(psi))


class CNN:
    """Handle of the device network (libnemohip nh_cnn_*)."""

    def __init__(self, owner, c_in, width=32, layers=4, c_out=N_STATES * N_AMP * 2,
                 seed=0):
        self.owner = owner
        self.lib = native.load()
        self.h = ctypes.c_void_p()
        native.check(self.lib.nh_cnn_create(owner.ctx, c_in, width, layers,
                                            c_out, seed, ctypes.byref(self.h)))
        self.c_in, self.width, self.layers, self.c_out = c_in, width, layers, c_out
        self.n_params = int(self.lib.nh_cnn_num_params(self.h))

    def bind(self, owner):
        """Use the network on another structure (context) of the device."""
        native.check(self.lib.nh_cnn_bind(self.h, owner.ctx))
        self.owner = owner

    def close(self):
        if self.h:
            self.lib.nh_cnn_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        w = np.zeros(self.n_params, dtype=np.float32)
        native.check(self.lib.nh_cnn_get_params(self.h, native.ptr(w)))
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.n_params
        native.check(self.lib.nh_cnn_set_params(self.h, native.ptr(w)))

    def forward(self, feats, dims):
        out = np.zeros((int(np.prod(dims)), self.c_out), dtype=np.float32)
        native.check(self.lib.nh_cnn_forward(
            self.h, native.ptr(np.ascontiguousarray(feats, dtype=np.float32)),
            dims[0], dims[1], dims[2], native.ptr(out)))
        return out

    def start_block(self, feats, gmap, n_states=N_STATES, device=False):
        """Predicted start block (n_states, N) for the owner's structure; a
        native.DeviceArray when device=True."""
        n = self.owner.n
        if device:
            blk = native.DeviceArray(self.owner, (n_states, n))
            kind = native.PTR_DEVICE
        else:
            blk = np.zeros((n_states, n), dtype=np.complex128)
            kind = native.PTR_HOST
        native.check(self.lib.nh_cnn_start_block(
            self.h, native.ptr(np.ascontiguousarray(feats, dtype=np.float32)),
            gmap.dims[0], gmap.dims[1], gmap.dims[2], native.ptr(gmap.point),
            native.ptr(gmap.slot), n_states, native.ptr(blk), kind))
        return blk

    def train_step(self, feats, gmap, targets, cluster, out_cluster, lr,
                   want_grad=False):
        """One Adam step (lr > 0) on the subspace loss; returns the loss
        (and the parameter gradient when want_grad).  targets: host array
        or native.DeviceArray, (n_targets, N) complex128."""
        if isinstance(targets, native.DeviceArray):
            assert targets.dtype == np.complex128 and len(targets.shape) == 2
            t = targets
        else:
            t = np.ascontiguousarray(targets, dtype=np.complex128)
        cl = np.ascontiguousarray(cluster, dtype=np.int32)
        oc = np.ascontiguousarray(out_cluster, dtype=np.int32)
        loss = ctypes.c_double()
        g = np.zeros(self.n_params, dtype=np.float32) if want_grad else None
        native.check(self.lib.nh_cnn_train_step(
            self.h, native.ptr(np.ascontiguousarray(feats, dtype=np.float32)),
            gmap.dims[0], gmap.dims[1], gmap.dims[2], native.ptr(gmap.point),
            native.ptr(gmap.slot), native.ptr(t), t.shape[0], native.ptr(cl),
            oc.size, native.ptr(oc), float(lr), ctypes.byref(loss),
            native.ptr(g) if g is not None else None))
        return (loss.value, g) if want_grad else loss.value

    def grad(self, feats, dims, dout):
        """Parameter gradient of sum(out * dout) (backward pass)."""
        g = np.zeros(self.n_params, dtype=np.float32)
        native.check(self.lib.nh_cnn_backward(
            self.h, native.ptr(np.ascontiguousarray(feats, dtype=np.float32)),
            dims[0], dims[1], dims[2],
            native.ptr(np.ascontiguousarray(dout, dtype=np.float32)),
            native.ptr(g)))
        return g
