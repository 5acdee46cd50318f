"""Learned warm start: training set from the plain solver, training on the
device, and the start block for a new structure.

Training structures are seeded dome dots (base 12-24 nm, height 3-7 nm, In
fraction 0.6-1.0 drawn per cation site of the dot and wetting layer) in a
GaAs box, relaxed by the device Keating VFF.  Their targets are the levels
nearest the gap from the plain FP64 Lanczos solver (the reference's
eigensolve_lanczos), each with its Kramers partner; the network learns to
span each level's two-dimensional subspace.

    python -m nemo3d_amd.warmstart train --dots 16 --out FILE.npz

A larger training set is built in rounds: --first K takes the dots K,
K + 1, ... of the seeded sequence, --resume FILE continues from a saved
network (its weights; Adam restarts) and keeps its list of dots.
"""
import argparse
import ctypes
import json
import time

import numpy as np

from . import learned, materials, model, native, structure, vff
from .hamiltonian import Hamiltonian
from .solvers import eigensolve_lanczos

A_GAAS = 0.56532  # nm, GaAs unstrnd_cubic_cell_length of the material file
COMPOUNDS = ["GaAs", "InAs"]


def dot_model(box_nm, base_nm, height_nm, x_in=1.0, seed=0, shape="dome",
              top_nm=None, db_shift=30.0):
    """Dot + 1-ML wetting layer of In_x Ga_(1-x) As in a GaAs box."""
    s = structure.dot_structure(
        box_nm, (A_GAAS,) * 3, shape=shape, base_nm=base_nm,
        height_nm=height_nm, top_nm=top_nm, wl_ml=1, alloy_x=x_in,
        alloy_cation="Ga" if x_in < 1.0 else None, seed=seed)
    return model.TBModel(s, COMPOUNDS, materials.BANDS_20_SPIN, passivate=True,
                         db_shift=db_shift)


def training_params(k, seed0=7000):
    rng = np.random.default_rng(seed0 + k)
    return dict(base_nm=float(rng.uniform(12, 24)),
                height_nm=float(rng.uniform(3, 7)),
                x_in=float(rng.uniform(0.6, 1.0)), seed=seed0 + k)


def find_gap(H, iters=2000, lo=-1.0, hi=3.0, tol=1e-6):
    """Band-gap edges from a Lanczos run: the largest gap between its
    converged Ritz values (trailing eigenvector component below tol) in
    [lo, hi] eV.  Converged Ritz values are eigenvalues of H, so the gap
    found is a gap of the spectrum."""
    from scipy.linalg import eigh_tridiagonal
    r = eigensolve_lanczos(H, num_vb=0, num_cb=1, erange_cb=(1e3, 1e3 + 1),
                           max_iter=iters, start_iter=10 * iters,
                           want_vectors=False, return_tridiag=True)
    a, b = r["alpha"], r["beta"]
    w, z = eigh_tridiagonal(a, b[:len(a) - 1], select="v",
                            select_range=(lo, hi))
    conv = w[np.abs(z[-1]) < tol]
    # interior Ritz values of a large structure converge late: with too few
    # converged ones, use all of them (Ritz values rarely sit in the gap)
    w = conv if len(conv) >= 8 else w
    if len(w) < 2:
        raise RuntimeError("find_gap: too few Ritz values in the range")
    k = int(np.argmax(np.diff(w)))
    return float(w[k]), float(w[k + 1])


def plain_levels(H, n_vb=4, n_cb=4, erange_vb=None, erange_cb=None, tol=1e-8,
                 max_iter=20000, width=0.6):
    """The n_vb highest valence and n_cb lowest conduction levels of the
    plain FP64 Lanczos solver with their eigenvectors (host).  Without
    windows, they end at the gap edges found by find_gap (the gap interior
    stays outside, so no unconverged Ritz value wandering in the gap is
    counted)."""
    if erange_vb is None or erange_cb is None:
        ev, ec = find_gap(H)
        m = 0.25 * (ec - ev)
        erange_vb = (ev - width, ev + m)
        erange_cb = (ec - m, ec + width)
    r = eigensolve_lanczos(H, num_vb=n_vb, num_cb=n_cb, erange_vb=erange_vb,
                           erange_cb=erange_cb, max_iter=max_iter, tol=tol,
                           start_iter=200, skip=50, stop_rule="nearest",
                           return_vectors=True, capacity=n_vb + n_cb)
    r["windows"] = (erange_vb, erange_cb)
    return r


def sample(M, energies, vectors, owner=None):
    """Features, grid map and targets (levels + Kramers partners, uploaded
    to the device) of one training structure."""
    g = learned.GridMap(M.structure)
    f = learned.features(M, g)
    ne = len(energies)
    T = np.empty((2 * ne, M.n), dtype=np.complex128)
    for e in range(ne):
        T[2 * e] = vectors[e]
        T[2 * e + 1] = learned.kramers_partner(vectors[e], M.nb)
    owner = owner if owner is not None else Hamiltonian(M, layout=None)
    dT = native.DeviceArray.from_host(owner, T)
    return dict(owner=owner, feats=f, gmap=g, targets=dT,
                cluster=np.repeat(np.arange(ne), 2).astype(np.int32),
                out_cluster=np.arange(ne, dtype=np.int32),
                energies=np.asarray(energies))


def new_network(owner, c_in, seed=0):
    return learned.CNN(owner, c_in, width=32, layers=4,
                       c_out=learned.N_STATES * learned.N_AMP * 2, seed=seed)


def train(net, samples, epochs=60, lr=2e-3, log=print):
    """Adam over the training structures; returns the loss history."""
    hist = []
    for ep in range(epochs):
        tot = 0.0
        for smp in samples:
            net.bind(smp["owner"])
            tot += net.train_step(smp["feats"], smp["gmap"], smp["targets"],
                                  smp["cluster"], smp["out_cluster"], lr)
        hist.append(tot / len(samples))
        if log and (ep % 10 == 0 or ep == epochs - 1):
            log("epoch %d loss %.4f" % (ep, hist[-1]))
    return hist


def save(net, path, meta):
    np.savez(path, params=net.get_params(), c_in=net.c_in, width=net.width,
             layers=net.layers, c_out=net.c_out, meta=json.dumps(meta))


def load(owner, path):
    d = np.load(path, allow_pickle=False)
    net = learned.CNN(owner, int(d["c_in"]), int(d["width"]), int(d["layers"]),
                      int(d["c_out"]))
    net.set_params(d["params"])
    return net, json.loads(str(d["meta"]))


def structure_inputs(H):
    """Grid map and per-atom features of H's structure (computed once per
    structure, like the Hamiltonian itself)."""
    g = learned.GridMap(H.model.structure)
    return g, learned.features(H.model, g)


def start_block(net, H, n_states=learned.N_STATES, inputs=None, device=False):
    """Predicted start block (n_states x N) for H's structure, orthonormal
    columns; a native.DeviceArray when device=True."""
    net.bind(H)
    g, f = inputs if inputs is not None else structure_inputs(H)
    return net.start_block(f, g, n_states, device=device)


def start_vector_device(H, block, download=True):
    """start_vector of a device block (native.DeviceArray, n_states x N): the
    sum of the normalised columns formed on the device in the block's first
    column (overwritten).  Returned on the host (N values), or with
    download=False the block itself, whose first row the Lanczos solver
    takes as a device start vector."""
    lib = native.load()
    k, n = (int(v) for v in block.shape)
    base = int(block.ptr.value)
    col = [ctypes.c_void_p(base + c * n * 16) for c in range(k)]
    one, zero = native.NHCplx(1.0, 0.0), native.NHCplx(0.0, 0.0)
    for c in range(k):
        nrm = ctypes.c_double()
        native.check(lib.nh_nrm2(H.ctx, n, col[c], ctypes.byref(nrm)))
        s = native.NHCplx(1.0 / nrm.value, 0.0)
        if c == 0:  # column 0 <- column 0 / |column 0|
            native.check(lib.nh_axpby(H.ctx, n, zero, col[0], s, col[0]))
        else:
            native.check(lib.nh_axpby(H.ctx, n, s, col[c], one, col[0]))
    if not download:
        return block
    out = np.empty(n, dtype=np.complex128)
    native.check(lib.nh_memcpy_d2h(H.ctx, native.ptr(out), col[0], n * 16))
    return out


def start_vector(block):
    """A single Lanczos start vector from a start block: the sum of its
    normalised columns."""
    b = np.asarray(block)
    return (b / np.linalg.norm(b, axis=1, keepdims=True)).sum(axis=0)


def subspace_overlap(block, vectors, nb=20):
    """Per target level: weight of the level's Kramers pair in the
    orthonormalised block (1 = contained)."""
    Q = np.linalg.qr(np.asarray(block).T)[0]
    out = []
    for v in vectors:
        P = np.stack([v, learned.kramers_partner(v, nb)])
        P = np.linalg.qr(P.T)[0]
        out.append(float(np.linalg.norm(np.conj(P.T) @ Q) ** 2 / 2))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--dots", type=int, default=16)
    ap.add_argument("--box", type=float, nargs=3, default=(24.0, 24.0, 16.0))
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--vb-window", type=float, nargs=2, default=(0.3, 0.95),
                    help="valence window (eV); its top lies in the gap")
    ap.add_argument("--cb-window", type=float, nargs=2, default=(1.0, 1.8),
                    help="conduction window (eV); its bottom lies in the gap")
    ap.add_argument("--out", required=True)
    ap.add_argument("--first", type=int, default=0,
                    help="index of the first dot of the seeded sequence")
    ap.add_argument("--resume", default=None,
                    help="continue from this saved network")
    a = ap.parse_args()
    samples, meta = [], []
    for k in range(a.first, a.first + a.dots):
        p = training_params(k)
        t0 = time.time()
        M = dot_model(tuple(a.box), p["base_nm"], p["height_nm"], p["x_in"], p["seed"])
        info = vff.relax(M)
        H = Hamiltonian(M)
        r = plain_levels(H, erange_vb=tuple(a.vb_window),
                         erange_cb=tuple(a.cb_window))
        # distinct levels only (a converged copy of a level counts once)
        ev, keep = r["eigenvalues"], []
        for i in np.argsort(ev):
            if all(abs(ev[i] - ev[k]) >= 1e-7 for k in keep):
                keep.append(i)
        if not r["converged"] or len(keep) < 6:
            print("dot %d: plain solver did not converge, skipped" % k, flush=True)
            continue
        r["eigenvalues"] = r["eigenvalues"][keep]
        r["vectors"] = r["vectors"][keep]
        samples.append(sample(M, r["eigenvalues"], r["vectors"], owner=H))
        meta.append(dict(p, atoms=M.n_atoms, vff_iter=info["iterations"],
                         levels=[float(e) for e in r["eigenvalues"]],
                         lanczos_iter=r["iterations"]))
        print("dot %d: %d atoms, levels %s, %.1fs" % (
            k, M.n_atoms, np.round(r["eigenvalues"], 4), time.time() - t0), flush=True)
    prev = {}
    if a.resume:
        net, prev = load(samples[0]["owner"], a.resume)
    else:
        net = new_network(samples[0]["owner"], samples[0]["feats"].shape[1])
    hist = train(net, samples, a.epochs, a.lr,
                 log=lambda m: print(m, flush=True))
    save(net, a.out, dict(dots=prev.get("dots", []) + meta,
                          loss=prev.get("loss", []) + hist, box_nm=list(a.box),
                          rounds=prev.get("rounds", []) + [
                              dict(first=a.first, dots=len(meta), epochs=a.epochs)]))


if __name__ == "__main__":
    main()
