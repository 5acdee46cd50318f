"""H·v / H·V on the GPU through the C ABI against the CPU oracle."""
import ctypes

import numpy as np
import pytest

import _models
from nemo3d_amd import native
from nemo3d_amd.hamiltonian import Hamiltonian
from oracle import oracle

pytestmark = pytest.mark.gpu

CASES = {
    "gaas20_strained": dict(builder="gaas", cells=(3, 2, 2), strain_nm=0.02, seed=3),
    "ingaas20_alloy_phi": dict(builder="gaas", cells=(2, 3, 2), strain_nm=0.015,
                               seed=4, alloy=True, phi=True),
    "gaas20_unpassivated": dict(builder="gaas", cells=(2, 2, 2), seed=5,
                                passivate=False),
    "si10_strained": dict(builder="si", cells=(3, 3, 3), strain_nm=0.02, seed=6),
}


def make(case):
    kw = dict(CASES[case])
    b = kw.pop("builder")
    return _models.gaas_box(**kw) if b == "gaas" else _models.si_box(**kw)


@pytest.mark.parametrize("case", sorted(CASES))
def test_assembly_matches_oracle(case):
    M = make(case)
    print("case", case, "atoms", M.n_atoms, flush=True)
    ref = oracle.assemble(M.arrays())
    H = Hamiltonian(M)
    onsite, blocks = H.export()
    scale = np.abs(ref.onsite[:, :10]).max()
    assert np.abs(onsite - ref.onsite).max() <= 1e-12 * scale
    bscale = np.abs(ref.blocks).max()
    assert np.abs(blocks - ref.blocks).max() <= 1e-12 * bscale


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("shift", [0.0, 1.3])
def test_hv_fp64_matches_oracle(case, shift):
    M = make(case)
    print("case", case, "shift", shift, flush=True)
    ref = oracle.assemble(M.arrays())
    H = Hamiltonian(M)
    X = _models.random_vectors(H.n, 3, seed=7)
    for x in X:
        y = H.hv(x, shift)
        yr = ref.hv(x, shift)
        assert np.linalg.norm(y - yr) <= 1e-12 * np.linalg.norm(yr)


@pytest.mark.parametrize("case", ["gaas20_strained", "si10_strained"])
def test_hV_block_equals_columns(case):
    M = make(case)
    H = Hamiltonian(M)
    X = _models.random_vectors(H.n, 8, seed=8)
    Y = H.hV(X, 0.4)
    for k in range(8):
        assert np.array_equal(Y[k], H.hv(X[k], 0.4))


@pytest.mark.parametrize("case", ["gaas20_strained", "si10_strained"])
def test_hv_fp32_error_bound(case):
    """FP32 stored values with FP32 products per block: error relative to
    the FP64 product stays at single-precision level."""
    M = make(case)
    ref = oracle.assemble(M.arrays())
    H = Hamiltonian(M, precision=native.PREC_FP32)
    X = _models.random_vectors(H.n, 2, seed=9)
    for x in X:
        y = H.hv(x, 0.0)
        yr = ref.hv(x, 0.0)
        err = np.linalg.norm(y - yr) / np.linalg.norm(yr)
        print(case, "FP32 rel. error %.2e" % err, flush=True)
        assert 1e-10 < err < 3e-7, err


def test_vector_kernels():
    M = make("si10_strained")
    H = Hamiltonian(M)
    lib = native.load()
    rng = np.random.default_rng(2)
    for n in (1, 17, 1000, 300001):
        x = rng.normal(size=n) + 1j * rng.normal(size=n)
        y = rng.normal(size=n) + 1j * rng.normal(size=n)
        tx = native.DeviceArray.from_host(H, x)
        ty = native.DeviceArray.from_host(H, y)
        out = native.NHCplx()
        native.check(lib.nh_dot(H.ctx, n, native.ptr(tx), native.ptr(ty),
                                ctypes.byref(out)))
        ref = np.vdot(x, y)
        assert abs(complex(out.re, out.im) - ref) <= 1e-12 * n
        nrm = ctypes.c_double()
        native.check(lib.nh_nrm2(H.ctx, n, native.ptr(tx), ctypes.byref(nrm)))
        assert abs(nrm.value - np.linalg.norm(x)) <= 1e-13 * np.linalg.norm(x)
        a = native.NHCplx(0.5, -1.25)
        b = native.NHCplx(-2.0, 0.75)
        native.check(lib.nh_axpby(H.ctx, n, a, native.ptr(tx), b, native.ptr(ty)))
        want = (0.5 - 1.25j) * x + (-2.0 + 0.75j) * y
        assert np.allclose(ty.download(), want, rtol=1e-14, atol=1e-14)


def test_device_buffers_hv():
    """H·v on device-resident buffers equals the host-pointer path."""
    M = make("gaas20_strained")
    H = Hamiltonian(M)
    X = _models.random_vectors(H.n, 4, seed=11)
    dX = native.DeviceArray.from_host(H, X)
    dY = H.hV(dX, 0.7)
    assert np.array_equal(dY.download(), H.hV(X, 0.7))


@pytest.mark.parametrize("case", sorted(CASES))
def test_hv_sk_layout_matches_oracle(case):
    """SK-compressed layout (block rebuilt from cosines + scaled integrals
    on every product) reproduces the oracle to the FP64 bar."""
    M = make(case)
    print("case", case, flush=True)
    ref = oracle.assemble(M.arrays())
    H = Hamiltonian(M, layout=native.LAYOUT_SK)
    X = _models.random_vectors(H.n, 3, seed=12)
    Y = H.hV(X, 0.9)
    for k in range(3):
        yr = ref.hv(X[k], 0.9)
        assert np.linalg.norm(Y[k] - yr) <= 1e-12 * np.linalg.norm(yr)


@pytest.mark.parametrize("case", ["gaas20_strained", "si10_strained"])
def test_hv_sk_fp32_error_bound(case):
    M = make(case)
    ref = oracle.assemble(M.arrays())
    H = Hamiltonian(M, layout=native.LAYOUT_SK, precision=native.PREC_FP32)
    x = _models.random_vectors(H.n, 1, seed=13)[0]
    err = np.linalg.norm(H.hv(x) - ref.hv(x)) / np.linalg.norm(ref.hv(x))
    print(case, "FP32 SK rel. error %.2e" % err, flush=True)
    assert 1e-10 < err < 3e-7, err


@pytest.mark.parametrize("layout", ["blocks", "sk"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_hV_mfma_fp32_matches_oracle(case, layout):
    """Block of 8 vectors on the matrix cores (FP32 builds): every column
    matches the oracle at single-precision level; device buffers with a
    padded leading dimension, padding left untouched."""
    M = make(case)
    print("case", case, layout, flush=True)
    ref = oracle.assemble(M.arrays())
    H = Hamiltonian(M, layout=native.LAYOUT_SK if layout == "sk"
                    else native.LAYOUT_BLOCKS, precision=native.PREC_FP32)
    n, ld = H.n, H.n + 40
    X = np.zeros((8, ld), dtype=np.complex128)
    X[:, :n] = _models.random_vectors(n, 8, seed=14)
    dX = native.DeviceArray.from_host(H, X)
    dY = native.DeviceArray.from_host(H, np.full((8, ld), 7.0 + 7.0j))
    native.check(native.load().nh_hV(H.ctx, 0.6, native.ptr(dX), native.ptr(dY),
                                     8, ld, native.PTR_DEVICE))
    Y = dY.download()
    assert np.all(Y[:, n:] == 7.0 + 7.0j)
    for k in range(8):
        yr = ref.hv(X[k, :n], 0.6)
        err = np.linalg.norm(Y[k, :n] - yr) / np.linalg.norm(yr)
        print(case, layout, k, "MFMA rel. error %.2e" % err, flush=True)
        assert 1e-10 < err < 3e-7, (k, err)
        # the single-vector FP32 product agrees to the same level
        y1 = H.hv(X[k, :n], 0.6)
        assert np.linalg.norm(Y[k, :n] - y1) / np.linalg.norm(yr) < 3e-7


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", ["gaas20_strained", "si10_strained", "gaas20_unpassivated"])
def test_hv_staged_and_cached_variants_agree(case, precision):
    """The LDS-staged and the cache-gather stored-block kernels give the
    same products (same arithmetic order per row) and match the oracle."""
    M = make(case)
    ref = oracle.assemble(M.arrays())
    H = Hamiltonian(M, precision=native.PREC_FP64 if precision == "fp64"
                    else native.PREC_FP32)
    X = _models.random_vectors(H.n, 3, seed=15)
    H.set(native.SET_HV_STAGED, 1)
    Ys = H.hV(X, 0.2)
    H.set(native.SET_HV_STAGED, 0)
    Yc = H.hV(X, 0.2)
    assert np.array_equal(Ys, Yc)
    for k in range(3):
        yr = ref.hv(X[k], 0.2)
        err = np.linalg.norm(Ys[k] - yr) / np.linalg.norm(yr)
        assert err <= (1e-12 if precision == "fp64" else 3e-7), err


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_hv_sk_quad_gather_matches_oracle(case, precision):
    """SK layout with the quad-cooperative LDS gather of the neighbours'
    entries: products equal to the direct gather's to FP64 / FP32 rounding
    (the compiler contracts the block's multiply-adds differently around
    the staged loads), oracle to the FP64 / FP32 bar, single and 3-column
    products (equal to rounding)."""
    M = make(case)
    print("case", case, precision, flush=True)
    ref = oracle.assemble(M.arrays())
    H = Hamiltonian(M, layout=native.LAYOUT_SK,
                    precision=native.PREC_FP64 if precision == "fp64" else native.PREC_FP32)
    X = _models.random_vectors(H.n, 3, seed=17)
    H.set(native.SET_HV_STAGED, 1)
    Ys = H.hV(X, 0.3)
    y1 = H.hv(X[1], 0.3)
    H.set(native.SET_HV_STAGED, 0)
    Yc = H.hV(X, 0.3)
    bar = 1e-14 if precision == "fp64" else 1e-6
    assert np.linalg.norm(Ys - Yc) <= bar * np.linalg.norm(Yc)
    # the 3 columns go through the block kernel (records read once): the
    # same products, contracted differently by the compiler
    assert np.linalg.norm(y1 - Ys[1]) <= bar * np.linalg.norm(y1)
    tol = 1e-12 if precision == "fp64" else 3e-7
    for k in range(3):
        yr = ref.hv(X[k], 0.3)
        assert np.linalg.norm(Ys[k] - yr) <= tol * np.linalg.norm(yr)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_lanczos_sk_quad_gather_same_run(precision):
    """The fused Lanczos kernels with the quad LDS gather give the same
    tridiagonal as the direct gather up to the rounding of the products:
    the same extreme Ritz values (to 1e-10 eV in FP64, 1e-5 eV in FP32)."""
    from nemo3d_amd.solvers import eigensolve_lanczos
    M = _models.gaas_box(cells=(2, 2, 2), strain_nm=0.01, seed=23)
    H = Hamiltonian(M, layout=native.LAYOUT_SK,
                    precision=native.PREC_FP64 if precision == "fp64" else native.PREC_FP32)
    opts = dict(num_vb=0, num_cb=1, erange_cb=(1e3, 1e3 + 1), max_iter=400,
                start_iter=10 ** 6, want_vectors=False, return_tridiag=True)
    H.set(native.SET_HV_STAGED, 0)
    r0 = eigensolve_lanczos(H, **opts)
    H.set(native.SET_HV_STAGED, 1)
    r1 = eigensolve_lanczos(H, **opts)
    H.set(native.SET_HV_STAGED, -1)
    from scipy.linalg import eigh_tridiagonal
    w0 = eigh_tridiagonal(r0["alpha"], r0["beta"][:-1], eigvals_only=True)
    w1 = eigh_tridiagonal(r1["alpha"], r1["beta"][:-1], eigvals_only=True)
    bar = 1e-10 if precision == "fp64" else 1e-5
    assert np.abs(w0[[0, 1, -2, -1]] - w1[[0, 1, -2, -1]]).max() < bar
    if precision == "fp64":  # the first steps agree to rounding
        assert np.abs(r0["alpha"][:20] - r1["alpha"][:20]).max() < 1e-12


@pytest.mark.parametrize("layout", ["sk", "mf"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", ["gaas20_strained", "ingaas20_alloy_phi", "si10_strained"])
def test_block_hv_variants_agree(case, precision, layout):
    """Block H·V in the SK and matrix-free layouts: all columns in one
    launch (bond records read once), one launch per column and, for FP32
    with 8 columns, the matrix-core kernel give the same products to
    rounding and match the oracle; device buffers with a padded leading
    dimension, padding untouched."""
    M = make(case)
    ref = oracle.assemble(M.arrays())
    H = Hamiltonian(M, layout=native.LAYOUT_SK if layout == "sk" else native.LAYOUT_MATRIX_FREE,
                    precision=native.PREC_FP64 if precision == "fp64" else native.PREC_FP32)
    n, ld, k = H.n, H.n + 24, 8
    X = np.zeros((k, ld), dtype=np.complex128)
    X[:, :n] = _models.random_vectors(n, k, seed=61)
    dX = native.DeviceArray.from_host(H, X)
    out = {}
    for mode in (0, 1, 2):
        if mode == 2 and (precision == "fp64" or layout == "mf"):
            continue
        H.set(native.SET_BLOCK_HV, mode)
        dY = native.DeviceArray.from_host(H, np.full((k, ld), 3.0 - 1.0j))
        native.check(native.load().nh_hV(H.ctx, 0.4, native.ptr(dX), native.ptr(dY),
                                         k, ld, native.PTR_DEVICE))
        Y = dY.download()
        assert np.all(Y[:, n:] == 3.0 - 1.0j)
        out[mode] = Y[:, :n]
    H.set(native.SET_BLOCK_HV, -1)
    tol = 1e-12 if precision == "fp64" else 3e-7
    for c in range(k):
        yr = ref.hv(X[c, :n], 0.4)
        for mode, Y in out.items():
            err = np.linalg.norm(Y[c] - yr) / np.linalg.norm(yr)
            assert err <= tol, (mode, c, err)
    same = 1e-15 if precision == "fp64" else 1e-7
    assert np.linalg.norm(out[1] - out[0]) <= same * np.linalg.norm(out[0])


@pytest.mark.parametrize("layout", ["sk", "mf"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sk_workgroup_roles_give_identical_products(layout, precision):
    """Both workgroup mappings of the SK / matrix-free products (one role
    per workgroup, or cations and anions of the same cells together) run
    the same per-lane arithmetic: products are bit-identical; the first 30
    Lanczos coefficients agree to rounding (the workgroup partial sums of
    alpha group the atoms differently, and without reorthogonalisation the
    difference grows once Ritz values converge, from about step 40 on this
    small box)."""
    from nemo3d_amd.solvers import eigensolve_lanczos
    M = make("ingaas20_alloy_phi")
    H = Hamiltonian(M, layout=native.LAYOUT_SK if layout == "sk" else native.LAYOUT_MATRIX_FREE,
                    precision=native.PREC_FP64 if precision == "fp64" else native.PREC_FP32)
    X = _models.random_vectors(H.n, 3, seed=71)
    out, tri = {}, {}
    for roles in (0, 1):
        H.set(native.SET_SK_ROLES, roles)
        out[roles] = H.hV(X, 0.3)
        r = eigensolve_lanczos(H, num_vb=0, num_cb=1, erange_cb=(1e3, 1e3 + 1),
                               max_iter=30, start_iter=10**6, want_vectors=False,
                               return_tridiag=True)
        tri[roles] = (r["alpha"], r["beta"])
    H.set(native.SET_SK_ROLES, -1)
    assert np.array_equal(out[0], out[1])
    for k in (0, 1):
        assert np.abs(tri[0][k] - tri[1][k]).max() <= 1e-10 * np.abs(tri[0][k]).max()


@pytest.mark.parametrize("layout", ["sk", "mf"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sk_split_lanes_lanczos(layout, precision):
    """The fused Lanczos kernels with 8 lanes per atom (a lane per bond and
    spin block) against the 4-lane ones: the tridiagonals agree to
    rounding over the first 30 steps (before Ritz values converge and
    rounding differences grow), and a solve with eigenvectors (eigenvector pass with the
    fused three-term step) gives eigenpairs whose residuals under the
    oracle's operator are at the precision's level."""
    from nemo3d_amd.solvers import eigensolve_lanczos
    M = make("ingaas20_alloy_phi")
    ref = oracle.assemble(M.arrays())
    H = Hamiltonian(M, layout=native.LAYOUT_SK if layout == "sk" else native.LAYOUT_MATRIX_FREE,
                    precision=native.PREC_FP64 if precision == "fp64" else native.PREC_FP32)
    tri = {}
    for lanes in (4, 8):
        H.set(native.SET_SK_LANES, lanes)
        r = eigensolve_lanczos(H, num_vb=0, num_cb=1, erange_cb=(1e3, 1e3 + 1),
                               max_iter=30, start_iter=10**6, want_vectors=False,
                               return_tridiag=True)
        tri[lanes] = (r["alpha"], r["beta"])
    tol = 1e-10 if precision == "fp64" else 1e-5
    for k in (0, 1):
        d = np.abs(tri[4][k] - tri[8][k]).max()
        print(layout, precision, "max tridiagonal difference", d, flush=True)
        assert d <= tol * np.abs(tri[4][k]).max()
    H.set(native.SET_SK_LANES, 8)
    r = eigensolve_lanczos(H, num_vb=0, num_cb=2, erange_cb=(-100.0, 100.0),
                           max_iter=3000, tol=1e-7, start_iter=50, skip=25,
                           return_vectors=True, capacity=4)
    H.set(native.SET_SK_LANES, -1)
    assert r["converged"]
    bound = 1e-7 * 40 if precision == "fp64" else 1e-5 * 40
    for e, v in zip(r["eigenvalues"], r["vectors"]):
        assert np.linalg.norm(ref.hv(v) - e * v) / np.linalg.norm(v) < bound


@pytest.mark.parametrize("layout", ["blocks", "sk", "mf"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_hv_reproduces_tracemin_example(layout, precision):
    """The device products of the structure behind the reference's TRACEMIN
    example (48-atom periodic Si, 20-band + spin-orbit) against the
    example's own matrix times x: FP64 to 1e-12 relative (the matrix is
    printed to 13 digits), FP32 to the FP32 bound."""
    M = _models.pq_si48()
    Hc = _models.pq_example_matrix()
    lay = {"blocks": native.LAYOUT_BLOCKS, "sk": native.LAYOUT_SK,
           "mf": native.LAYOUT_MATRIX_FREE}[layout]
    H = Hamiltonian(M, layout=lay,
                    precision=native.PREC_FP64 if precision == "fp64" else native.PREC_FP32)
    X = _models.random_vectors(H.n, 3, seed=81)
    tol = 1e-12 if precision == "fp64" else 3e-7
    for x in X:
        yr = Hc @ x
        err = np.linalg.norm(H.hv(x) - yr) / np.linalg.norm(yr)
        assert err <= tol, err


def test_context_settings_are_checked():
    """nh_context_set refuses values outside each setting's range and
    unknown settings with NH_ERR_ARG, and leaves the context usable."""
    M = make("ingaas20_alloy_phi")
    H = Hamiltonian(M, layout=native.LAYOUT_SK)
    for what, bad in ((native.SET_LZ_VECTORS, 2), (native.SET_SK_LANES, 5),
                      (native.SET_SK_ROLES, 3), (native.SET_RING_STEPS, 7),
                      (native.SET_BLOCK_HV, 3), (native.SET_HV_STAGED, 2), (99, 0)):
        with pytest.raises(native.NHError):
            H.set(what, bad)
    for what in (native.SET_LZ_VECTORS, native.SET_SK_LANES, native.SET_SK_ROLES):
        H.set(what, -1)
    X = _models.random_vectors(H.n, 1, seed=5)[0]
    ref = oracle.assemble(M.arrays())
    assert np.linalg.norm(H.hv(X, 0.0) - ref.hv(X)) <= 1e-12 * np.linalg.norm(ref.hv(X))
