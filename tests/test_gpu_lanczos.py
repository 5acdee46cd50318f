"""Device Lanczos against dense diagonalisation of the oracle Hamiltonian."""
import numpy as np
import pytest

import _models
from nemo3d_amd import native
from nemo3d_amd.hamiltonian import Hamiltonian
from nemo3d_amd.solvers import eigensolve_lanczos
from oracle import oracle

pytestmark = pytest.mark.gpu


def gap_windows(w, width=1.5):
    """Band-gap windows from a dense spectrum: largest gap inside [-1, 3] eV."""
    sel = w[(w > -1.0) & (w < 3.0)]
    k = np.argmax(np.diff(sel))
    ev, ec = sel[k], sel[k + 1]
    return (ev - width, ev + 1e-6), (ec - 1e-6, ec + width)


def distinct(w, tol=1e-8):
    out = []
    for x in np.sort(w):
        if not out or x - out[-1] > tol:
            out.append(x)
    return np.array(out)


@pytest.mark.parametrize("layout", ["blocks", "sk"])
@pytest.mark.parametrize("case", ["si10", "gaas20"])
def test_lanczos_matches_dense(case, layout):
    if case == "si10":
        M = _models.si_box(cells=(3, 3, 3), strain_nm=0.01, seed=21)
    else:
        M = _models.gaas_box(cells=(2, 2, 2), strain_nm=0.01, seed=22)
    print("case", case, "atoms", M.n_atoms, flush=True)
    ref = oracle.assemble(M.arrays())
    w = distinct(np.linalg.eigvalsh(ref.dense()))
    wv, wc = gap_windows(w)
    H = Hamiltonian(M, layout=native.LAYOUT_SK if layout == "sk"
                    else native.LAYOUT_BLOCKS)
    r = eigensolve_lanczos(H, num_vb=3, num_cb=3, erange_vb=wv, erange_cb=wc,
                           max_iter=6000, tol=1e-10, start_iter=100, skip=25,
                           return_vectors=True, stop_rule="nearest")
    assert r["converged"]
    got = np.sort(r["eigenvalues"])
    top_vb = w[(w > wv[0]) & (w <= wv[1])][-3:]
    bot_cb = w[(w > wc[0]) & (w <= wc[1])][:3]
    want = np.sort(np.concatenate([top_vb, bot_cb]))
    assert got.shape == want.shape
    assert np.abs(got - want).max() < 1e-9, (got, want)
    # eigenvectors: residual of the oracle H
    for e, v in zip(r["eigenvalues"], r["vectors"]):
        res = np.linalg.norm(ref.hv(v) - e * v)
        assert res < 1e-6, res
    assert (r["residuals"] < 1e-6).all()


def test_lanczos_window_rule():
    """Reference stop rule: stop once the windows hold the requested number
    of converged values; windows bracketing exactly the wanted states."""
    M = _models.gaas_box(cells=(2, 2, 2), strain_nm=0.01, seed=23)
    ref = oracle.assemble(M.arrays())
    w = distinct(np.linalg.eigvalsh(ref.dense()))
    sel = np.nonzero((w > -1.0) & (w < 3.0))[0]
    k = sel[np.argmax(np.diff(w[sel]))]  # w[k]: top of the valence band
    vb, cb = w[k - 1:k + 1], w[k + 1:k + 3]
    H = Hamiltonian(M, layout=native.LAYOUT_SK)
    r = eigensolve_lanczos(H, num_vb=2, num_cb=2,
                           erange_vb=(0.5 * (w[k - 2] + w[k - 1]), w[k] + 1e-3),
                           erange_cb=(w[k + 1] - 1e-3, 0.5 * (w[k + 2] + w[k + 3])),
                           max_iter=6000, tol=1e-10, start_iter=100, skip=25,
                           stop_rule="window")
    assert r["converged"]
    got = np.sort(r["eigenvalues"])
    assert np.abs(got - np.sort(np.concatenate([vb, cb]))).max() < 1e-9
    assert (r["residuals"] < 1e-6).all()


def tight_windows(w, nv, nc):
    """Windows holding exactly the nv highest valence and nc lowest
    conduction levels of the distinct spectrum w."""
    sel = np.nonzero((w > -1.0) & (w < 3.0))[0]
    k = sel[np.argmax(np.diff(w[sel]))]  # w[k]: top of the valence band
    vb, cb = w[k - nv + 1:k + 1], w[k + 1:k + 1 + nc]
    wv = (0.5 * (w[k - nv] + w[k - nv + 1]), w[k] + 1e-3)
    wc = (w[k + 1] - 1e-3, 0.5 * (w[k + nc] + w[k + nc + 1]))
    return vb, cb, wv, wc


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", ["si10", "gaas20"])
def test_block_lanczos_matches_dense(case, precision):
    """Block Lanczos (block 8) with the reference window rule on windows
    that hold exactly the wanted levels.  The boxes are large enough that
    the block Krylov space does not exhaust the space (N / 8 steps) before
    convergence; FP32 products converge to an FP32-level tolerance."""
    from nemo3d_amd.solvers import eigensolve_block_lanczos
    if case == "si10":
        M = _models.si_box(cells=(3, 3, 3), strain_nm=0.01, seed=31)
        nv = nc = 4
    else:
        M = _models.gaas_box(cells=(3, 2, 2), strain_nm=0.01, seed=32)
        nv = nc = 2
    print("case", case, precision, flush=True)
    ref = oracle.assemble(M.arrays())
    w = distinct(np.linalg.eigvalsh(ref.dense()))
    vb, cb, wv, wc = tight_windows(w, nv, nc)
    H = Hamiltonian(M, layout=native.LAYOUT_SK,
                    precision=native.PREC_FP64 if precision == "fp64"
                    else native.PREC_FP32)
    r = eigensolve_block_lanczos(H, block_size=8, num_vb=nv, num_cb=nc,
                                 erange_vb=wv, erange_cb=wc,
                                 max_iter=M.n // 8 - 10,
                                 tol=1e-8 if precision == "fp64" else 1e-6,
                                 start_iter=5, skip=5, return_vectors=True,
                                 stop_rule="window")
    assert r["converged"], r
    got = np.sort(r["eigenvalues"])
    want = np.sort(np.concatenate([vb, cb]))
    tol_e = 1e-9 if precision == "fp64" else 1e-5
    assert np.abs(got - want).max() < tol_e, (got, want)
    # FP32 products: the Ritz vectors are those of the perturbed operator
    # (they seed the FP64 refinement of the mixed solver, which meets 1e-6)
    for e, v in zip(r["eigenvalues"], r["vectors"]):
        assert np.linalg.norm(ref.hv(v) - e * v) < (1e-6 if precision == "fp64" else 1e-2)


def test_block_lanczos_extreme_levels():
    """The lowest levels of the spectrum (fast Krylov convergence): the
    recurrence, Ritz extraction and eigenvector pass on a 20-band box,
    nearest rule with the window's lower end acting as the gap edge."""
    from nemo3d_amd.solvers import eigensolve_block_lanczos
    M = _models.gaas_box(cells=(2, 2, 2), strain_nm=0.01, seed=34)
    ref = oracle.assemble(M.arrays())
    w = distinct(np.linalg.eigvalsh(ref.dense()))
    H = Hamiltonian(M, layout=native.LAYOUT_SK)
    r = eigensolve_block_lanczos(H, block_size=8, num_vb=0, num_cb=4,
                                 erange_vb=(0.0, 0.0),
                                 erange_cb=(w[0] - 1.0, w[6] + 1e-3),
                                 max_iter=120, tol=1e-9, start_iter=5, skip=5,
                                 return_vectors=True, stop_rule="nearest")
    assert r["converged"], r
    want = w[:4]
    assert np.abs(np.sort(r["eigenvalues"]) - want).max() < 1e-9
    for e, v in zip(r["eigenvalues"], r["vectors"]):
        assert np.linalg.norm(ref.hv(v) - e * v) < 1e-6


def test_tracemin_matches_dense():
    """PQ-TRACEMIN: the eigenvalues nearest sigma of a small strained Si box."""
    from nemo3d_amd.solvers import eigensolve_tracemin
    M = _models.si_box(cells=(2, 2, 2), strain_nm=0.01, seed=41)
    ref = oracle.assemble(M.arrays())
    w = np.linalg.eigvalsh(ref.dense())
    H = Hamiltonian(M, layout=native.LAYOUT_SK)
    sigma = 1.0
    r = eigensolve_tracemin(H, num_eig=4, block_size=8, ref_eig=sigma,
                            tol=1e-9, hnorm=40.0, cgtol=1.0, max_outer=300,
                            cg_max=60, return_vectors=True)
    assert r["converged"], r
    want = np.sort(w[np.argsort(np.abs(w - sigma))[:4]])
    assert np.abs(np.sort(r["eigenvalues"]) - want).max() < 1e-9
    for e, v in zip(r["eigenvalues"], r["vectors"]):
        assert np.linalg.norm(ref.hv(v) - e * v) < 1e-7


def test_tracemin_matches_dense_spin_orbit():
    """PQ-TRACEMIN on a 20-band + spin-orbit strained GaAs box (complex
    Hermitian H, every level a Kramers pair): the shift just above the top
    valence level, the two eigenvalues nearest it (both members of that
    pair) against dense eigh, FP64 residuals under the oracle operator.
    A shift at mid-gap of this small box (a 3 eV gap against a 40 eV
    spectrum) leaves (H - sigma)^2 too ill-conditioned for 400 outer steps."""
    from nemo3d_amd.solvers import eigensolve_tracemin
    M = _models.gaas_box(cells=(2, 2, 2), strain_nm=0.01, seed=29)
    ref = oracle.assemble(M.arrays())
    w = np.linalg.eigvalsh(ref.dense())
    d = distinct(w)
    sel = np.nonzero((d > -1.0) & (d < 4.0))[0]
    k = sel[np.argmax(np.diff(d[sel]))]  # d[k]: top of the valence band
    sigma = d[k] + 0.02
    H = Hamiltonian(M, layout=native.LAYOUT_SK)
    r = eigensolve_tracemin(H, num_eig=2, block_size=8, ref_eig=sigma,
                            tol=1e-9, hnorm=40.0, cgtol=1.0, max_outer=400,
                            cg_max=80, return_vectors=True)
    print("sigma", sigma, "outer", r["iterations"], r["eigenvalues"], flush=True)
    assert r["converged"], r
    want = np.sort(w[np.argsort(np.abs(w - sigma))[:2]])
    assert abs(want[1] - want[0]) < 1e-12  # a Kramers pair
    assert np.abs(np.sort(r["eigenvalues"]) - want).max() < 1e-9
    for e, v in zip(r["eigenvalues"], r["vectors"]):
        assert np.linalg.norm(ref.hv(v) - e * v) / np.linalg.norm(v) < 1e-7


def test_tracemin_on_the_reference_example():
    """PQ-TRACEMIN on the structure behind the reference's TRACEMIN example
    (TRACEMIN/Examples/pqexample2_full: 48-atom periodic Si, 20-band +
    spin-orbit, shift 0): the six eigenvalues nearest 0 (a four-fold and a
    two-fold level; the next lies 1.25 eV away) against the dense
    eigenvalues of the example's own matrix."""
    from nemo3d_amd.solvers import eigensolve_tracemin
    M = _models.pq_si48()
    w = np.linalg.eigvalsh(_models.pq_example_matrix())
    want = np.sort(w[np.argsort(np.abs(w))[:6]])
    H = Hamiltonian(M, layout=native.LAYOUT_SK)
    r = eigensolve_tracemin(H, num_eig=6, block_size=8, ref_eig=0.0, tol=1e-9,
                            hnorm=40.0, cgtol=1.0, max_outer=300, cg_max=60)
    print("outer", r["iterations"], np.round(r["eigenvalues"], 6), flush=True)
    assert r["converged"], r
    assert np.abs(np.sort(r["eigenvalues"]) - want).max() < 1e-9


@pytest.mark.parametrize("vectors", ["fp64", "fp32"])
def test_mixed_precision_matches_dense(vectors):
    """FP32-product Lanczos (FP64 recurrence; its vectors stored in FP64 or
    FP32) + FP64 refinement: eigenvalues to the FP64 bar, FP64 residuals
    below tol * hnorm, operator back in FP64 afterwards."""
    from nemo3d_amd.solvers import eigensolve_mixed
    M = _models.gaas_box(cells=(2, 2, 2), strain_nm=0.01, seed=23)
    ref = oracle.assemble(M.arrays())
    w = distinct(np.linalg.eigvalsh(ref.dense()))
    sel = np.nonzero((w > -1.0) & (w < 3.0))[0]
    k = sel[np.argmax(np.diff(w[sel]))]
    H = Hamiltonian(M, layout=native.LAYOUT_SK)
    H.set(native.SET_LZ_VECTORS, 1 if vectors == "fp32" else 0)
    r = eigensolve_mixed(H, num_vb=2, num_cb=2,
                         erange_vb=(0.5 * (w[k - 2] + w[k - 1]), w[k] + 1e-3),
                         erange_cb=(w[k + 1] - 1e-3, 0.5 * (w[k + 2] + w[k + 3])),
                         tol=1e-8, tol32=1e-7, max_iter=6000, start_iter=100,
                         skip=25, refine_depth=12, refine_rounds=8,
                         return_vectors=True)
    print("stage32", r["stage32"]["iterations"], r["stage32"]["eigenvalues"],
          "refine", r["refine"], flush=True)
    assert r["converged"], r
    want = np.sort(np.concatenate([w[k - 1:k + 1], w[k + 1:k + 3]]))
    assert np.abs(r["eigenvalues"] - want).max() < 1e-9
    # the FP32 stage alone is off by FP32 rounding, the refinement is not
    assert np.abs(np.sort(r["stage32"]["eigenvalues"]) - want).max() > 1e-11
    for e, v, res in zip(r["eigenvalues"], r["vectors"], r["residuals"]):
        true = np.linalg.norm(ref.hv(v) - e * v) / np.linalg.norm(v)
        assert true < 1e-8 * 40.0 * 1.01 and abs(true - res) < 1e-12
    assert H.precision == native.PREC_FP64


@pytest.mark.parametrize("model,layout,vectors,lanes", [
    ("gaas20", "sk", "fp64", -1), ("gaas20", "sk", "fp32", 4), ("gaas20", "sk", "fp32", 8),
    ("gaas20", "mf", "fp32", 4), ("gaas20", "mf", "fp32", 8), ("si10", "sk", "fp32", -1),
    ("si10", "mf", "fp32", -1)])
def test_fp32_lanczos_vectors_at_fp32_level(model, layout, vectors, lanes):
    """Lanczos with FP32 products: the eigenvector pass accumulates the FP32
    copies of r (its ring) with 1/beta folded into the coefficients; the
    vectors' FP64 residuals and eigenvalues sit at the FP32 rounding of H·v,
    far below the gap to the next level, and the FP64 run stays exact.
    With the Lanczos vectors themselves stored in FP32 (the FP64 step's
    result rounded once) the eigenvector pass must still retrace the first
    pass exactly, or its Ritz vectors would not be eigenvectors."""
    if model == "gaas20":
        M = _models.gaas_box(cells=(2, 2, 2), strain_nm=0.01, seed=24)
    else:  # 10-band spinless (one spin block per atom)
        M = _models.si_box(cells=(3, 3, 3), strain_nm=0.01, seed=25)
    ref = oracle.assemble(M.arrays())
    wd, vd = np.linalg.eigh(ref.dense())
    w = distinct(wd)
    wv, wc = gap_windows(w)
    H = Hamiltonian(M, layout=native.LAYOUT_SK if layout == "sk" else native.LAYOUT_MATRIX_FREE)
    H.set(native.SET_LZ_VECTORS, 1 if vectors == "fp32" else 0)
    H.set(native.SET_SK_LANES, lanes)
    out = {}
    for prec in (native.PREC_FP32, native.PREC_FP64):
        H.set_precision(prec)
        out[prec] = eigensolve_lanczos(H, num_vb=2, num_cb=2, erange_vb=wv,
                                       erange_cb=wc, max_iter=6000, tol=1e-7,
                                       start_iter=100, skip=25,
                                       return_vectors=True, stop_rule="nearest")
    H.set_precision(native.PREC_FP64)
    # bars (relative to ||H|| = 40 eV): the FP64 run to its Lanczos
    # tolerance; the FP32 run to the FP32 rounding of H·v
    # (measured: FP32 2.3-2.8e-8, eigenvalues within 2e-7 eV of FP64)
    for prec, bar_e, bar_r in ((native.PREC_FP32, 1e-6, 2e-7), (native.PREC_FP64, 1e-9, 1e-7)):
        r = out[prec]
        assert r["converged"]
        for e, v in zip(r["eigenvalues"], r["vectors"]):
            v = v / np.linalg.norm(v)
            assert np.abs(wd - e).min() < bar_e, (prec, e)
            res = np.linalg.norm(ref.hv(v) - e * v)
            print(model, layout, vectors, lanes, "precision", prec, "E", e, "residual / ||H||",
                  res / 40.0, "iterations", r["iterations"], flush=True)
            assert res < bar_r * 40.0, (prec, e, res)
            # the vector lies in the eigenspace of its level (Kramers pair)
            sel = np.abs(wd - e) < 1e-6
            P = vd[:, sel]
            assert np.linalg.norm(P.conj().T @ v) ** 2 > 1 - 1e-6, (prec, e)
    print("fp32 vs fp64 eigenvalues",
          np.sort(out[native.PREC_FP32]["eigenvalues"]) - np.sort(out[native.PREC_FP64]["eigenvalues"]),
          flush=True)


@pytest.mark.parametrize("layout,precision", [("sk", "fp64"), ("sk", "fp32"),
                                               ("blocks", "fp64")])
def test_eigenvector_ring_length(layout, precision):
    """The eigenvector pass adds its ring into the eigenvectors every K
    steps (K = 128, fewer when HBM is short): K = 8 gives the same
    eigenpairs to rounding (SK: rings of r, FP64 or FP32 copies; stored
    blocks: ring of q)."""
    M = _models.gaas_box(cells=(2, 2, 2), strain_nm=0.01, seed=24)
    ref = oracle.assemble(M.arrays())
    w = distinct(np.linalg.eigvalsh(ref.dense()))
    wv, wc = gap_windows(w)
    H = Hamiltonian(M, layout=native.LAYOUT_SK if layout == "sk" else native.LAYOUT_BLOCKS,
                    precision=native.PREC_FP64 if precision == "fp64" else native.PREC_FP32)
    out = {}
    for K in (0, 8):
        H.set(native.SET_RING_STEPS, K)
        out[K] = eigensolve_lanczos(H, num_vb=2, num_cb=2, erange_vb=wv,
                                    erange_cb=wc, max_iter=6000, tol=1e-7,
                                    start_iter=100, skip=25,
                                    return_vectors=True, stop_rule="nearest")
    H.set(native.SET_RING_STEPS, 0)
    a, b = out[0], out[8]
    print(layout, precision, "iterations", a["iterations"], "eigenvalues",
          a["eigenvalues"], flush=True)
    assert a["converged"] and b["converged"]
    assert a["iterations"] == b["iterations"]
    assert np.abs(a["eigenvalues"] - b["eigenvalues"]).max() < 1e-12
    for va, vb in zip(a["vectors"], b["vectors"]):
        assert np.linalg.norm(va - vb) < 1e-10 * np.linalg.norm(va)


def test_refine_eigenpairs_from_perturbed_vectors():
    """FP64 refinement of dense eigenvectors perturbed by noise: residuals
    fall below the tolerance and the values reach the dense ones."""
    from nemo3d_amd.solvers import refine_eigenpairs
    M = _models.si_box(cells=(3, 3, 3), strain_nm=0.01, seed=33)
    ref = oracle.assemble(M.arrays())
    wd, vd = np.linalg.eigh(ref.dense())
    H = Hamiltonian(M, layout=native.LAYOUT_SK)
    pick = [40, 300, 1000]
    rng = np.random.default_rng(5)
    X = vd[:, pick].T + 1e-10 * (rng.normal(size=(3, M.n)) + 1j * rng.normal(size=(3, M.n)))
    res0 = [np.linalg.norm(ref.hv(x) - e * x) / np.linalg.norm(x) for x, e in zip(X, wd[pick])]
    r = refine_eigenpairs(H, X, wd[pick] + 1e-6, tol=0.5 * min(res0) / 40.0,
                          hnorm=40.0, depth=8, max_rounds=8)
    print("refine", r["rounds"], r["matvecs"], res0, r["residuals"], flush=True)
    assert r["converged"]
    assert np.abs(r["eigenvalues"] - wd[pick]).max() < 1e-10
    for e, v, rr in zip(r["eigenvalues"], r["vectors"], r["residuals"]):
        assert np.linalg.norm(ref.hv(v) - e * v) < 0.5 * min(res0) * 1.01
        assert abs(np.linalg.norm(ref.hv(v) - e * v) - rr) < 1e-3 * rr


def test_mixed_block_matches_dense():
    """FP32 block Lanczos + FP64 refinement: eigenvalues to the FP64 bar."""
    from nemo3d_amd.solvers import eigensolve_mixed_block
    M = _models.si_box(cells=(3, 3, 3), strain_nm=0.01, seed=33)
    ref = oracle.assemble(M.arrays())
    w = distinct(np.linalg.eigvalsh(ref.dense()))
    H = Hamiltonian(M, layout=native.LAYOUT_SK)
    vb, cb, wv, wc = tight_windows(w, 2, 2)
    r = eigensolve_mixed_block(H, num_vb=2, num_cb=2, erange_vb=wv, erange_cb=wc,
                               tol=1e-8, tol32=1e-5, max_iter32=M.n // 8 - 10,
                               max_iter64=M.n // 8 - 10, stop_rule="window",
                               return_vectors=True)
    assert r["converged"], r
    want = np.sort(np.concatenate([vb, cb]))
    assert np.abs(np.sort(r["eigenvalues"]) - want).max() < 1e-9
    # Ritz vectors of the unreorthogonalised block recurrence (second
    # pass): true residuals of a few 1e-6 eV, reported by the solver
    for e, v, res in zip(r["eigenvalues"], r["vectors"], r["residuals"]):
        true = np.linalg.norm(ref.hv(v) - e * v) / np.linalg.norm(v)
        assert true < 1e-5 and abs(true - res) < 1e-9
    assert H.precision == native.PREC_FP64


def test_pctracemin_matches_dense():
    """PC-TRACEMIN: the eigenpairs in a window holding the top valence level
    (both members of its Kramers pair) match dense eigh, with FP64
    residuals under the oracle operator.  The window reaches half way into
    the 3.5 eV gap of this small box: the quadratic map separates the
    wanted levels only when the window edges sit in wide gaps of the
    spectrum."""
    from nemo3d_amd.solvers import eigensolve_pctracemin
    M = _models.gaas_box(cells=(2, 2, 2), strain_nm=0.01, seed=23)
    ref = oracle.assemble(M.arrays())
    w = np.linalg.eigvalsh(ref.dense())
    d = distinct(w)
    sel = np.nonzero((d > -2.0) & (d < 5.0))[0]
    k = sel[np.argmax(np.diff(d[sel]))]  # d[k]: top of the valence band
    a, b = 0.5 * (d[k - 1] + d[k]), 0.5 * (d[k] + d[k + 1])
    want = w[(w > a) & (w < b)]
    print("window", a, b, "wanted", want, flush=True)
    assert len(want) == 2
    H = Hamiltonian(M, layout=native.LAYOUT_SK)
    r = eigensolve_pctracemin(H, num_eig=2, want_range=(a, b),
                              all_range=(w[0] - 0.5, w[-1] + 0.5), block_size=8,
                              cheb_degree=100, tol=1e-9, max_outer=200,
                              return_vectors=True)
    print("iterations", r["iterations"], "matvecs", r["matvecs"], r["eigenvalues"], flush=True)
    assert r["converged"], r
    assert np.abs(np.sort(r["eigenvalues"]) - np.sort(want)).max() < 1e-9
    hn = max(abs(w[0] - 0.5), abs(w[-1] + 0.5))
    for e, v in zip(r["eigenvalues"], r["vectors"]):
        assert np.linalg.norm(ref.hv(v) - e * v) / np.linalg.norm(v) < 1e-9 * hn * 1.01
