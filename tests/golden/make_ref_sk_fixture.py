"""Writes tests/golden/ref_sk_bonds.npz and tests/golden/ref_onsite.npz:
strained Slater-Koster blocks with their Boykin diagonal corrections, and
onsite blocks (diagonal energies + spin-orbit), computed by the reference's
own routines (h3d_offdiag_diag_corr and h3d_diag, built from the reference
sources by oracle/ref_build.sh into oracle/_ref/ref_sk).  Run in a
container that has the reference tree:
  sh oracle/ref_build.sh && python tests/golden/make_ref_sk_fixture.py
"""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from nemo3d_amd import materials  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_sk")
BOND_DIRS = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float)


def cases(seed=0, n=24):
    """Random strained bonds of every cation/anion pair of the In/Ga/As
    tables, both band models, both row roles."""
    rng = np.random.default_rng(seed)
    out = []
    for band in (materials.BANDS_20_SPIN, materials.BANDS_10_NOSPIN):
        t = materials.param_init(["GaAs", "InAs"], band)
        nb = 20 if band == materials.BANDS_20_SPIN else 10
        As = t.species.index("As")
        for k in range(n):
            c = t.species.index("Ga" if k % 2 == 0 else "In")
            par, st = t.parmat[c][As], t.strnmat[c][As]
            a0 = par[31]
            nnv0 = 0.25 * a0 * BOND_DIRS[k % 4]
            nnv = nnv0 * (1 + rng.normal(0, 0.04, 3)) + rng.normal(0, 0.003, 3)
            out.append((nb, k % 2 == 0, a0, nnv, nnv0, par.copy(), st.copy()))
    return out


def run(cs, harness=HARNESS):
    lines = ["%d %d %.17g %s %s %s %s" % (
        nb, int(anion), a0, " ".join("%.17g" % v for v in nnv),
        " ".join("%.17g" % v for v in nnv0), " ".join("%.17g" % v for v in par),
        " ".join("%.17g" % v for v in st)) for nb, anion, a0, nnv, nnv0, par, st in cs]
    res = subprocess.run([harness], input="\n".join(lines) + "\n",
                         capture_output=True, text=True, check=True)
    rows = [np.array(l.split(), float) for l in res.stdout.strip().split("\n")]
    ho = np.array([r[:100].reshape(10, 10) for r in rows])
    hd = np.array([r[100:] for r in rows])
    return ho, hd


def onsite_cases(seed=1):
    """Onsite parameter sets: the GaAs and InAs records, the mean of the
    two (an As atom with Ga and In neighbours) and random vectors (every
    entry distinct, so a swapped index shows), both band models, both
    roles."""
    rng = np.random.default_rng(seed)
    out = []
    for band in (materials.BANDS_20_SPIN, materials.BANDS_10_NOSPIN):
        t = materials.param_init(["GaAs", "InAs"], band)
        nb = 20 if band == materials.BANDS_20_SPIN else 10
        As = t.species.index("As")
        ga = t.parmat[t.species.index("Ga")][As]
        inn = t.parmat[t.species.index("In")][As]
        pars = [ga, inn, 0.5 * (ga + inn)] + [rng.normal(0, 5, 35) for _ in range(3)]
        for par in pars:
            for anion in (0, 1):
                out.append((nb, anion, np.array(par, float)))
    return out


def run_onsite(cs, harness=HARNESS):
    lines = ["%d %d %s" % (nb, anion, " ".join("%.17g" % v for v in par))
             for nb, anion, par in cs]
    res = subprocess.run([harness, "diag"], input="\n".join(lines) + "\n",
                         capture_output=True, text=True, check=True)
    rows = [np.array(l.split(), float) for l in res.stdout.strip().split("\n")]
    h = np.zeros((len(rows), 20, 20), dtype=np.complex128)
    for k, (r, (nb, _, _)) in enumerate(zip(rows, cs)):
        v = r.reshape(nb, nb, 2)
        h[k, :nb, :nb] = v[..., 0] + 1j * v[..., 1]
    return h


if __name__ == "__main__":
    oc = onsite_cases()
    h = run_onsite(oc)
    np.savez_compressed(
        os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_onsite.npz"),
        nb=np.array([c[0] for c in oc]), anion=np.array([c[1] for c in oc]),
        par=np.array([c[2] for c in oc]), h=h)
    print("wrote", len(oc), "onsite blocks")
    cs = cases()
    ho, hd = run(cs)
    np.savez_compressed(
        os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_sk_bonds.npz"),
        nb=np.array([c[0] for c in cs]), anion_row=np.array([c[1] for c in cs]),
        a0=np.array([c[2] for c in cs]), nnv=np.array([c[3] for c in cs]),
        nnv0=np.array([c[4] for c in cs]), par=np.array([c[5] for c in cs]),
        strain=np.array([c[6] for c in cs]), ho=ho, hd=hd)
    print("wrote", len(cs), "bonds")
