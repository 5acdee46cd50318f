"""Learned warm start end to end on small dots: plain-solver targets,
training with the device kernels, and the gain over a random start on a
dot the network has not seen."""
import numpy as np
import pytest

from nemo3d_amd import vff, warmstart
from nemo3d_amd.hamiltonian import Hamiltonian
from nemo3d_amd.solvers import eigensolve_lanczos
from scipy.linalg import eigh_tridiagonal

pytestmark = pytest.mark.gpu

BOX = (7.0, 7.0, 5.0)


def small_dot(k):
    rng = np.random.default_rng(500 + k)
    M = warmstart.dot_model(BOX, float(rng.uniform(3.5, 5.0)),
                            float(rng.uniform(1.5, 2.5)),
                            float(rng.uniform(0.6, 1.0)), 500 + k)
    vff.relax(M)
    return M


def test_learned_start_gains_over_random():
    samples = []
    for k in range(4):
        M = small_dot(k)
        H = Hamiltonian(M)
        r = warmstart.plain_levels(H, 4, 4)
        assert r["converged"], ("training dot", k, r)
        samples.append(warmstart.sample(M, r["eigenvalues"], r["vectors"], owner=H))
        print("dot", k, M.n_atoms, np.round(r["eigenvalues"], 4), r["iterations"], flush=True)
    net = warmstart.new_network(samples[0]["owner"], samples[0]["feats"].shape[1], seed=3)
    hist = warmstart.train(net, samples, epochs=120, lr=2e-3, log=None)
    print("loss", hist[0], "->", hist[-1], flush=True)
    assert hist[-1] < 0.8 * hist[0]

    # held-out dot
    M = small_dot(10)
    H = Hamiltonian(M)
    ref = warmstart.plain_levels(H, 4, 4)
    assert ref["converged"]
    print("reference levels", repr(np.sort(ref["eigenvalues"])), flush=True)
    block = warmstart.start_block(net, H)
    ov = warmstart.subspace_overlap(block, ref["vectors"])
    rng = np.random.default_rng(0)
    rand = rng.normal(size=block.shape) + 1j * rng.normal(size=block.shape)
    ov_rand = warmstart.subspace_overlap(rand, ref["vectors"])
    print("overlap learned", np.round(ov, 3), "random", np.round(ov_rand, 6), flush=True)
    assert np.mean(ov) > 100 * np.mean(ov_rand)

    # iterations until every reference level has a converged Ritz value
    # (bottom-component residual bound <= 1e-8 ||H||), from the tridiagonal
    levels = np.unique(np.round(ref["eigenvalues"], 8))
    (vlo, _), (_, chi) = ref["windows"]

    def iterations_to_converge(start_vector):
        r = eigensolve_lanczos(H, num_vb=0, num_cb=1, erange_cb=(1e3, 1e3 + 1),
                               max_iter=6000, start_iter=10 ** 6,
                               start_vector=start_vector, want_vectors=False,
                               return_tridiag=True)
        a, b = r["alpha"], r["beta"]
        for m in range(50, len(a) + 1, 10):
            w, z = eigh_tridiagonal(a[:m], b[:m - 1], select="v",
                                    select_range=(vlo, chi))
            ok = np.abs(b[m - 1] * z[-1]) <= 1e-8 * 40.0
            if all(np.any(ok & (np.abs(w - e) < 1e-7)) for e in levels):
                return m
        return None

    # the device-side start vector equals the host one
    sv_host = warmstart.start_vector(block)
    sv_dev = warmstart.start_vector_device(H, warmstart.start_block(net, H, device=True))
    assert np.linalg.norm(sv_dev - sv_host) <= 1e-12 * np.linalg.norm(sv_host)
    # a device start vector (start = 3) gives the run of the same host vector
    blk = warmstart.start_block(net, H, device=True)
    dev = warmstart.start_vector_device(H, blk, download=False)
    host = dev.download()[0]
    opts = dict(num_vb=0, num_cb=1, erange_cb=(1e3, 1e3 + 1), max_iter=60,
                start_iter=10 ** 6, want_vectors=False, return_tridiag=True)
    r_d = eigensolve_lanczos(H, start_vector=dev, **opts)
    r_h = eigensolve_lanczos(H, start_vector=host, **opts)
    assert np.array_equal(r_d["alpha"], r_h["alpha"])
    assert np.array_equal(r_d["beta"], r_h["beta"])

    m_rand = iterations_to_converge(None)
    m_learn = iterations_to_converge(warmstart.start_vector(block))
    print("iterations to converge: random start", m_rand, "learned start",
          m_learn, flush=True)
    assert m_rand is not None and m_learn is not None
    assert m_learn < m_rand
