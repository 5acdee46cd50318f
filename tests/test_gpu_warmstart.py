"""Learned warm start end to end on small dots: plain-solver targets,
training with the device kernels, and the gain over a random start on a
dot the network has not seen."""
import numpy as np
import pytest

from nemo3d_amd import vff, warmstart
from nemo3d_amd.hamiltonian import Hamiltonian
from nemo3d_amd.solvers import eigensolve_lanczos

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
    vb_w, cb_w = ref["windows"]
    block = warmstart.start_block(net, H)
    ov = warmstart.subspace_overlap(block, ref["vectors"])
    rng = np.random.default_rng(0)
    rand = rng.normal(size=block.shape) + 1j * rng.normal(size=block.shape)
    ov_rand = warmstart.subspace_overlap(rand, ref["vectors"])
    print("overlap learned", np.round(ov, 3), "random", np.round(ov_rand, 6), flush=True)
    assert np.mean(ov) > 100 * np.mean(ov_rand)

    opts = dict(num_vb=4, num_cb=4, erange_vb=vb_w, erange_cb=cb_w, tol=1e-8,
                max_iter=40000, start_iter=50, skip=25, stop_rule="nearest",
                want_vectors=False)
    r_rand = eigensolve_lanczos(H, **opts)
    r_learn = eigensolve_lanczos(H, start_vector=warmstart.start_vector(block), **opts)
    print("iterations: random start", r_rand["iterations"], "learned start",
          r_learn["iterations"], flush=True)
    assert r_rand["converged"] and r_learn["converged"]
    assert np.abs(np.sort(r_learn["eigenvalues"]) - np.sort(ref["eigenvalues"])).max() < 1e-7
    assert r_learn["iterations"] < r_rand["iterations"]
