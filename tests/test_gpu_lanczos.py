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
                           return_vectors=True)
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
