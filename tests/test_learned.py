"""Learned warm start: CPU oracle of the network, features, and the device
network against the oracle."""
import numpy as np
import pytest

from nemo3d_amd import learned, materials, model, structure
from oracle import cnn_oracle

SHAPE = dict(c_in=5, width=32, layers=2, c_out=6)


def tiny():
    rng = np.random.default_rng(3)
    offs, n = cnn_oracle.layout(**SHAPE)
    params = rng.normal(0, 0.2, n)
    dims = (3, 4, 3)
    feats = rng.normal(size=(np.prod(dims), SHAPE["c_in"]))
    dout = rng.normal(size=(np.prod(dims), SHAPE["c_out"]))
    return params, feats, dims, dout


def test_oracle_backward_matches_finite_differences():
    params, feats, dims, dout = tiny()
    g = cnn_oracle.backward(params, feats, dims, dout=dout, round_bf16=False, **SHAPE)
    rng = np.random.default_rng(4)
    h = 1e-6
    for k in rng.choice(params.size, 40, replace=False):
        pp, pm = params.copy(), params.copy()
        pp[k] += h
        pm[k] -= h
        fp = (cnn_oracle.forward(pp, feats, dims, round_bf16=False, **SHAPE) * dout).sum()
        fm = (cnn_oracle.forward(pm, feats, dims, round_bf16=False, **SHAPE) * dout).sum()
        fd = (fp - fm) / (2 * h)
        assert abs(g[k] - fd) <= 1e-6 * (1 + abs(fd)), (k, g[k], fd)


def test_bf16_rounding():
    # 1 + 2^-8 and 1 + 3*2^-8 are ties: nearest even is 1.0 and 1 + 2^-6
    x = np.array([1.0, 1.00390625, 1.01171875, -3.14159265, 65504.0], dtype=np.float32)
    y = cnn_oracle.bf16(x)
    assert y[0] == 1.0 and y[1] == 1.0 and y[2] == 1.015625
    assert abs(y[3] + 3.140625) < 1e-12


def test_local_strain_hydrostatic():
    a = 0.56532
    s = structure.binary_box((3, 3, 3), "Ga", "As", (a, a, a))
    s.disp = 0.02 * s.positions(strained=False)  # uniform 2 % expansion
    e = learned.local_strain(s)
    bulk = s.is_cation.astype(bool) & ~s.surface_mask()
    assert np.allclose(e[bulk][:, :3], 0.02, atol=1e-12)
    assert np.allclose(e[bulk][:, 3:], 0.0, atol=1e-12)
    assert np.abs(e).max() <= 0.02 + 1e-12


def test_grid_map_pairs_anion_with_cation():
    a = 0.56532
    s = structure.binary_box((2, 2, 2), "Ga", "As", (a, a, a))
    g = learned.GridMap(s)
    # each point holds at most one cation and one anion; the anion of a
    # point is the cation's neighbour along -(1,1,1) a/4
    key = g.point * 2 + g.slot
    assert np.unique(key).size == s.n_atoms
    p = s.positions()
    for i in np.nonzero(s.is_cation)[0][:20]:
        j = np.nonzero((g.point == g.point[i]) & (g.slot == 1))[0]
        assert j.size == 1
        assert np.allclose(p[i] - p[j[0]], 0.25 * a, atol=1e-12)


@pytest.mark.gpu
def test_device_network_matches_oracle():
    import _models
    from nemo3d_amd.hamiltonian import Hamiltonian
    M = _models.si_box(cells=(2, 2, 2))
    H = Hamiltonian(M, layout=None)
    params, feats, dims, dout = tiny()
    net = learned.CNN(H, SHAPE["c_in"], SHAPE["width"], SHAPE["layers"],
                      SHAPE["c_out"], seed=1)
    net.set_params(params.astype(np.float32))
    p32 = params.astype(np.float32).astype(np.float64)
    out = net.forward(feats.astype(np.float32), dims)
    ref = cnn_oracle.forward(p32, feats.astype(np.float32), dims, **SHAPE)
    assert np.abs(out - ref).max() <= 1e-5 * np.abs(ref).max()
    g = net.grad(feats.astype(np.float32), dims, dout.astype(np.float32))
    gref = cnn_oracle.backward(p32, feats.astype(np.float32), dims,
                               dout=dout.astype(np.float32), **SHAPE)
    assert np.abs(g - gref).max() <= 1e-4 * np.abs(gref).max()


def _targets(n, rng, levels=4):
    """Orthonormal targets: `levels` states and their Kramers partners
    (orthonormalised together, as eigenvectors of distinct levels are)."""
    v = rng.normal(size=(levels, n)) + 1j * rng.normal(size=(levels, n))
    t = []
    for e in range(levels):
        t += [v[e], learned.kramers_partner(v[e], 20)]
    t = np.linalg.qr(np.array(t).T)[0].T
    return t, np.repeat(np.arange(levels), 2), np.arange(levels)


def test_subspace_loss_gradient_and_invariances():
    rng = np.random.default_rng(7)
    n = 60
    V, cl, oc = _targets(n, rng)
    Y = rng.normal(size=(4, n)) + 1j * rng.normal(size=(4, n))
    L, G = cnn_oracle.subspace_loss(Y, V, cl, oc)
    h = 1e-6
    for e, i in [(0, 3), (2, 17), (3, 59), (1, 0)]:
        for unit in (1.0, 1j):
            Yp, Ym = Y.copy(), Y.copy()
            Yp[e, i] += h * unit
            Ym[e, i] -= h * unit
            fd = (cnn_oracle.subspace_loss(Yp, V, cl, oc)[0] -
                  cnn_oracle.subspace_loss(Ym, V, cl, oc)[0]) / (2 * h)
            got = G[e, i].real if unit == 1.0 else G[e, i].imag
            assert abs(got - fd) < 1e-7, (e, i, unit, got, fd)
    # invariant under phases of the columns and unitary mixing of a pair
    ph = np.exp(1j * rng.uniform(0, 6.3, 4))
    assert abs(cnn_oracle.subspace_loss(Y * ph[:, None], V, cl, oc)[0] - L) < 1e-12
    U = np.linalg.qr(rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2)))[0]
    V2 = V.copy()
    V2[0:2] = U @ V[0:2]
    assert abs(cnn_oracle.subspace_loss(Y, V2, cl, oc)[0] - L) < 1e-12
    # zero exactly on the target subspaces, one on their complement
    Yin = np.array([V[2 * e] * 0.3 + V[2 * e + 1] * (0.2 - 1j) for e in range(4)])
    assert abs(cnn_oracle.subspace_loss(Yin, V, cl, oc)[0]) < 1e-12
    Yout = Y - (np.conj(V) @ Y.T).T @ V
    assert abs(cnn_oracle.subspace_loss(Yout, V, cl, oc)[0] - 1.0) < 1e-12


def test_kramers_partner_commutes_with_H():
    import sys, os
    sys.path.insert(0, os.path.dirname(__file__))
    import _models
    from oracle import oracle
    M = _models.gaas_box(cells=(2, 2, 2), strain_nm=0.02, seed=3, alloy=True)
    ref = oracle.assemble(M.arrays())
    rng = np.random.default_rng(1)
    x = rng.normal(size=M.n) + 1j * rng.normal(size=M.n)
    tx = learned.kramers_partner(x, 20)
    assert abs(np.vdot(x, tx)) < 1e-12 * np.vdot(x, x).real
    hx = ref.hv(x)
    assert np.linalg.norm(ref.hv(tx) - learned.kramers_partner(hx, 20)) <= \
        1e-13 * np.linalg.norm(hx)


@pytest.mark.gpu
def test_device_train_step_and_start_block_match_oracle():
    import _models
    from nemo3d_amd.hamiltonian import Hamiltonian
    M = _models.gaas_box(cells=(2, 2, 2), strain_nm=0.02, seed=8)
    H = Hamiltonian(M, layout=None)
    g = learned.GridMap(M.structure)
    rng = np.random.default_rng(9)
    ne, c_in = 4, 6
    shape = dict(c_in=c_in, width=32, layers=2, c_out=ne * 2 * 20 * 2)
    feats = rng.normal(size=(g.n_points, c_in)).astype(np.float32)
    net = learned.CNN(H, c_in, 32, 2, shape["c_out"], seed=2)
    params = net.get_params().astype(np.float64)
    out = cnn_oracle.forward(params, feats, g.dims, **shape)
    Yref = cnn_oracle.map_out(out, g.point, g.slot, 20, ne)
    Y = net.start_block(feats, g, ne)
    assert np.abs(Y - Yref).max() <= 1e-5 * np.abs(Yref).max()
    V, cl, oc = _targets(H.n, rng)
    L, grad = net.train_step(feats, g, V, cl, oc, lr=0.0, want_grad=True)
    Lref, G = cnn_oracle.subspace_loss(Yref, V, cl, oc)
    assert abs(L - Lref) <= 1e-6 * abs(Lref)
    dout = cnn_oracle.map_grad(G, g.point, g.slot, 20, ne, g.n_points, shape["c_out"])
    gref = cnn_oracle.backward(params, feats, g.dims, dout=dout, **shape)
    assert np.abs(grad - gref).max() <= 1e-3 * np.abs(gref).max()
    # lr = 0 leaves the network unchanged; a real step lowers the loss
    assert np.array_equal(net.get_params(), params.astype(np.float32))
    losses = [net.train_step(feats, g, V, cl, oc, lr=3e-3) for _ in range(30)]
    assert losses[-1] < losses[0], losses
