"""The plain method at the reference's CPU-sized shape: a passivated GaAs
box of 10x10x10 conventional cells (8,000 atoms), ideal positions,
sp3d5s* without spin-orbit (N = 80,000), 4 highest valence and 4 lowest
conduction levels with the reference Lanczos settings (window rule,
CyclicRandom start, tolerance 1e-8).  Checked at full size against the
oracle: FP64 residuals of the GPU eigenvectors under the oracle's operator,
and the oracle's own Lanczos tridiagonal (same start, same length) holding
every GPU eigenvalue among its Ritz values."""
import numpy as np
import pytest

import _models
from nemo3d_amd import materials, native, warmstart
from nemo3d_amd.hamiltonian import Hamiltonian
from nemo3d_amd.solvers import DEFAULT_CYCLIC, eigensolve_lanczos
from oracle import oracle

pytestmark = pytest.mark.gpu


def cyclic_start(n, cyc=DEFAULT_CYCLIC):
    """CyclicRandom start vector (lanczos_initialize, eigsys3d_par.c:113-215):
    entry j = (cyc[2j mod L], cyc[2j+1 mod L]), normalised."""
    c = np.asarray(cyc, dtype=np.float64)
    j = np.arange(n)
    v = c[(2 * j) % len(c)] + 1j * c[(2 * j + 1) % len(c)]
    return v / np.linalg.norm(v)


def test_config0_plain_lanczos_against_oracle():
    M = _models.gaas_box(cells=(10, 10, 10), band=materials.BANDS_10_NOSPIN)
    print("atoms", M.n_atoms, "N", M.n, flush=True)
    assert M.n_atoms == 8000 and M.n == 80000
    H = Hamiltonian(M, layout=native.LAYOUT_SK)
    ev, ec = warmstart.find_gap(H)
    m = 0.25 * (ec - ev)
    wv, wc = (ev - 0.6, ev + m), (ec - m, ec + 0.6)
    r = eigensolve_lanczos(H, num_vb=4, num_cb=4, erange_vb=wv, erange_cb=wc,
                           max_iter=8000, tol=1e-8, start_iter=200, skip=50,
                           stop_rule="window", return_vectors=True, capacity=8)
    print("gap", ev, ec, "iterations", r["iterations"], "E", r["eigenvalues"], flush=True)
    assert r["converged"] and len(r["eigenvalues"]) == 8
    ref = oracle.assemble(M.arrays())
    for e, v in zip(r["eigenvalues"], r["vectors"]):
        res = np.linalg.norm(ref.hv(v) - e * v) / np.linalg.norm(v)
        assert res <= 1e-8 * 40.0, res
    a, b = ref.lanczos_tridiag(cyclic_start(M.n), r["iterations"])
    w = oracle.tridiag_eigs(a, b[:len(a) - 1], wv[0], wc[1])
    for e in r["eigenvalues"]:
        assert np.abs(w - e).min() < 1e-9, e
