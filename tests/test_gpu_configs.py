"""The benchmark's structures at full size, every eigenpair checked against
the FP64 oracle.

* 786,520-atom InAs dome dot (base 20 nm, height 5 nm, 1-ML wetting layer,
  30 x 30 x 20 nm GaAs box, VFF positions, sp3d5s* + SO, N = 15.7M): plain
  FP64 Lanczos; mixed precision (FP32 products and Lanczos vectors, FP64
  step arithmetic and refinement) from the learned start, in the stored
  and the matrix-free layouts; learned against CyclicRandom start.
* The same geometry with a seeded In0.8Ga0.2As alloy in the dot and
  wetting layer: plain and mixed.
* 4,336,640-atom truncated-pyramid In0.7Ga0.3As dot in a 50 x 50 x 40 nm
  GaAs box (N = 86.7M): mixed precision from the learned start.

For each eigenpair (E, psi) of the GPU the oracle (C restatement, FP64,
its own assembly of the same structure) gives ||H psi - E psi|| / ||psi||,
which must be <= 1e-8 ||H|| (||H|| = 40 eV, the spectral radius), and the
Rayleigh quotient psi^H H psi / psi^H psi, which must equal E to 1e-9 eV
(the residual bound makes the quotient's error second order); the mixed
eigenvalues must lie within 1e-6 eV of the plain FP64 ones.  This is the
reference's own residual check (compute_ritz_residual_norm_deviation,
NEMO_3D/nemo3d/src/base/h_cvectr_mult.c:78-103) done in FP64."""
import os

import numpy as np
import pytest

from nemo3d_amd import native, vff, warmstart
from nemo3d_amd.hamiltonian import Hamiltonian
from nemo3d_amd.solvers import eigensolve_lanczos, eigensolve_mixed
from oracle import oracle

pytestmark = pytest.mark.gpu

HNORM = 40.0
WEIGHTS = os.path.join(os.path.dirname(__file__), "..", "nemo3d_amd", "data",
                       "warmstart.npz")
# the bench's windows: each holds exactly the 4 wanted levels.  With a wider
# valence window the reference's rule (done once the window holds NumEigReq
# converged values) can stop on the 5th valence level, 0.7092 eV, 0.9 meV
# below the 4th, when that one converges first (seen with FP32 products).
DOT_WINDOWS = dict(erange_vb=(0.7096, 0.7300), erange_cb=(1.4900, 1.5200))
PYRAMID_WINDOWS = dict(erange_vb=(0.50, 0.56), erange_cb=(1.43, 1.48))


def oracle_check(ref, eigenvalues, vectors, label):
    """Relative residuals and Rayleigh-quotient deviations under the
    oracle's operator."""
    res, dev = [], []
    for e, v in zip(eigenvalues, vectors):
        hv = ref.hv(v)
        nv = np.vdot(v, v).real
        rq = np.vdot(v, hv).real / nv
        res.append(np.linalg.norm(hv - e * v) / np.sqrt(nv) / HNORM)
        dev.append(abs(rq - e))
    print(label, "oracle rel. residuals", np.array(res), "|RQ - E|", np.array(dev),
          flush=True)
    assert max(res) <= 1e-8, (label, res)
    assert max(dev) <= 1e-9, (label, dev)
    return res


def plain(H, windows, max_iter=30000):
    r = eigensolve_lanczos(H, num_vb=4, num_cb=4, max_iter=max_iter, tol=1e-8,
                           start_iter=200, skip=25, stop_rule="window",
                           return_vectors=True, capacity=8, **windows)
    print("plain: %d iterations, %.2f s, E = %s" % (
        r["iterations"], r["seconds"], np.round(r["eigenvalues"], 7)), flush=True)
    assert r["converged"] and len(r["eigenvalues"]) == 8
    return r


def learned_start(H):
    if not os.path.exists(WEIGHTS):
        pytest.skip("no trained network")
    net, _ = warmstart.load(H, WEIGHTS)
    blk = warmstart.start_block(net, H, device=True)
    return warmstart.start_vector_device(H, blk, download=False)


def mixed(H, windows, start=None, max_iter=30000):
    r = eigensolve_mixed(H, num_vb=4, num_cb=4, tol=1e-8, tol32=3e-8,
                         hnorm=HNORM, max_iter=max_iter, start_iter=200, skip=25,
                         start_vector=start, stop_rule="window", refine_depth=12,
                         refine_rounds=4, return_vectors=True, **windows)
    print("mixed: FP32 stage %d iterations, refinement %d rounds, %.2f s, E = %s" % (
        r["stage32"]["iterations"], r["refine"]["rounds"], r["seconds"],
        np.round(r["eigenvalues"], 7)), flush=True)
    assert r["converged"] and len(r["eigenvalues"]) == 8
    return r


class Case:
    def __init__(self, M):
        print("atoms %d, N %d" % (M.n_atoms, M.n), flush=True)
        info = vff.relax(M)
        print("vff", info, flush=True)
        self.M = M
        self.H = Hamiltonian(M, layout=native.LAYOUT_SK)
        self._ref = None
        self.plain = None

    @property
    def ref(self):
        if self._ref is None:
            self._ref = oracle.assemble(self.M.arrays())
        return self._ref


@pytest.fixture(scope="module")
def dot():
    c = Case(warmstart.dot_model((30.0, 30.0, 20.0), 20.0, 5.0, 1.0, 0))
    assert c.M.n_atoms == 786520
    yield c
    c.H.close()


@pytest.mark.timeout(300)
def test_dot_plain_fp64_against_oracle(dot):
    dot.plain = plain(dot.H, DOT_WINDOWS)
    oracle_check(dot.ref, dot.plain["eigenvalues"], dot.plain["vectors"], "plain")


@pytest.mark.timeout(300)
def test_dot_mixed_learned_against_oracle(dot):
    if dot.plain is None:
        dot.plain = plain(dot.H, DOT_WINDOWS)
    r = mixed(dot.H, DOT_WINDOWS, learned_start(dot.H))
    oracle_check(dot.ref, r["eigenvalues"], r["vectors"], "mixed+learned")
    d = np.sort(r["eigenvalues"]) - np.sort(dot.plain["eigenvalues"])
    print("mixed - plain (eV)", d, flush=True)
    assert np.abs(d).max() <= 1e-6


@pytest.mark.timeout(300)
def test_dot_learned_start_needs_fewer_iterations(dot):
    """The learned start against the reference's CyclicRandom start, both
    with FP32 products to the same tolerance on the same dot: fewer
    Lanczos iterations from the learned start."""
    rl = mixed(dot.H, DOT_WINDOWS, learned_start(dot.H))
    rr = mixed(dot.H, DOT_WINDOWS, None)
    il, ir = rl["stage32"]["iterations"], rr["stage32"]["iterations"]
    print("iterations: learned %d, CyclicRandom %d" % (il, ir), flush=True)
    assert il < ir
    assert np.abs(np.sort(rl["eigenvalues"]) - np.sort(rr["eigenvalues"])).max() <= 1e-6


@pytest.mark.timeout(300)
def test_dot_matrix_free_mixed_against_oracle(dot):
    """The same dot with the matrix-free layout (bond records recomputed
    from the positions in every product, the reference's
    HamCompute=Recompute): the mixed solve (FP32 products and Lanczos
    vectors, 8 lanes per atom) from the learned start against the FP64
    oracle, and its eigenvalues against the stored-layout plain run."""
    if dot.plain is None:
        dot.plain = plain(dot.H, DOT_WINDOWS)
    Hm = Hamiltonian(dot.M, layout=native.LAYOUT_MATRIX_FREE)
    try:
        r = mixed(Hm, DOT_WINDOWS, learned_start(Hm))
        oracle_check(dot.ref, r["eigenvalues"], r["vectors"], "matrix-free mixed")
        d = np.sort(r["eigenvalues"]) - np.sort(dot.plain["eigenvalues"])
        print("matrix-free mixed - plain (eV)", d, flush=True)
        assert np.abs(d).max() <= 1e-6
    finally:
        Hm.close()


def tight_windows(eigenvalues, split, pad=2e-5):
    """Windows holding exactly the given levels (levels of these dots are
    >= 0.2 meV apart)."""
    e = np.sort(eigenvalues)
    return dict(erange_vb=(e[0] - pad, e[split - 1] + pad),
                erange_cb=(e[split] - pad, e[-1] + pad))


@pytest.mark.timeout(420)
def test_alloy_dot_plain_and_mixed_against_oracle():
    """In0.8Ga0.2As (species drawn per cation site of the dot and wetting
    layer from the seed) on the same geometry.  The windows come from the
    gap of a Lanczos scan; the mixed run is asked for the plain run's
    levels."""
    c = Case(warmstart.dot_model((30.0, 30.0, 20.0), 20.0, 5.0, 0.8, 11))
    try:
        # 6000 steps leave converged Ritz values on both sides of the gap
        # (3000 leave none below it in this alloy)
        ev, ec = warmstart.find_gap(c.H, iters=6000, lo=0.2, hi=2.0)
        m = 0.25 * (ec - ev)
        w = dict(erange_vb=(ev - 0.03, ev + m), erange_cb=(ec - m, ec + 0.05))
        print("gap", ev, ec, "windows", w, flush=True)
        p = plain(c.H, w)
        oracle_check(c.ref, p["eigenvalues"], p["vectors"], "alloy plain")
        r = mixed(c.H, tight_windows(p["eigenvalues"], 4), learned_start(c.H))
        oracle_check(c.ref, r["eigenvalues"], r["vectors"], "alloy mixed")
        assert np.abs(np.sort(r["eigenvalues"]) - np.sort(p["eigenvalues"])).max() <= 1e-6
    finally:
        c.H.close()


@pytest.mark.timeout(600)
def test_pyramid4m_mixed_against_oracle():
    """The large single-GPU shape, mixed precision from the learned start;
    every eigenpair's residual and Rayleigh quotient under the oracle (its
    assembly of 4.3M atoms and eight FP64 products on the host)."""
    c = Case(warmstart.dot_model((50.0, 50.0, 40.0), 30.0, 8.0, 0.7, 4,
                                 shape="pyramid", top_nm=12.0))
    try:
        assert c.M.n_atoms == 4336640
        r = mixed(c.H, PYRAMID_WINDOWS, learned_start(c.H))
        vecs = r.pop("vectors")
        c.H.close()
        oracle_check(c.ref, r["eigenvalues"], vecs, "pyramid4m mixed")
    finally:
        c.H.close()
