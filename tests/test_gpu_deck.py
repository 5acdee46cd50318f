"""End to end from the reference's own test deck: InGaAs_QD
(NEMO_3D/nemo3d/example/TestRun/inputs/InGaAs_QD.xml, parsed into
tests/golden/ingaas_qd_deck.json) through the deck's structure (a GaAs
dome holding an In0.8Ga0.2As box, species drawn with glibc rand() from
RandomSeed), its strain model (VFF_keating_strained, AllFree), the
20-band Hamiltonian and the deck's Lanczos settings (one level in each
window, tolerance 1e-7, CyclicRandom start), against the eigenvalues the
reference's test run printed (TestRun/InGaAs_QD.old_nd_Ek_ascii):
measured agreement 1.7e-8 and 7.8e-8 eV (the reference's Lanczos
tolerance is 1e-7), asserted at 5e-7 eV.  The relaxation follows the
reference's gradient (each atom differentiates its own copy of a bond
term); with the full derivative of the energy instead, the levels move by
11 and 10 meV."""
import json
import os
import time

import numpy as np
import pytest

from nemo3d_amd import deck, native, vff
from nemo3d_amd.hamiltonian import Hamiltonian
from nemo3d_amd.solvers import eigensolve_lanczos

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_levels():
    txt = open(os.path.join(GOLDEN, "InGaAs_QD.old_nd_Ek_ascii")).read()
    rows = txt.split("<\\HDR>")[1].split("\n")
    return np.array([float(r.split()[3]) for r in rows if r.strip()])


@pytest.mark.timeout(400)
def test_ingaas_qd_deck_against_reference_eigenvalues():
    with open(os.path.join(GOLDEN, "ingaas_qd_deck.json")) as f:
        t = json.load(f)
    M = deck.build_model(t)
    so = deck.strain_options(t)
    print("atoms", M.n_atoms, "strain", so["model"], so["keating"]["BoundCond"], flush=True)
    t0 = time.time()
    info = vff.relax(M, max_iter=20000, gtol=1e-11, model=so["model"],
                     boundary=so["keating"]["BoundCond"])
    print("relaxation %.2f s" % (time.time() - t0), info, flush=True)
    assert info["grad"] <= 1e-11 * 1.0001
    H = Hamiltonian(M, layout=native.LAYOUT_SK)
    r = eigensolve_lanczos(H, **deck.lanczos_options(t))
    got = np.sort(r["eigenvalues"])
    want = golden_levels()
    print("iterations", r["iterations"], "E", got, "reference", want,
          "difference", got - want, flush=True)
    assert r["converged"] and len(got) == 2
    assert np.abs(got - want).max() < 5e-7
    H.close()
