"""Write the small golden fixtures used by tests/ from the reference tree.

Reads data files of the reference's own tests (never its code) and stores
compact copies beside this script:

* gaas_bulk_ek.npz  - nemo3d/example/TestRun/GaAs_bulk.old_nd_Ek_ascii:
  20-band sp3d5s* + spin-orbit spectrum of periodic bulk GaAs (8-atom
  cell) at 40 k-points along [100] (k in units of pi/a).
* silicon4_meta.npz - PostNEMO3D/bin/test/silicon4.nd_Ek eigenvalues and
  checksums of the neighbour table / positions of the 64,000-atom Si box.

Usage: python make_fixtures.py [/root/reference/NEMO_3D]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def nd_payload(path):
    b = open(path, "rb").read()
    i = b.find(b"<\\HDR>")
    return b[i + 6:]


def main(root):
    g = np.loadtxt(os.path.join(root, "nemo3d/example/TestRun/"
                                "GaAs_bulk.old_nd_Ek_ascii"), skiprows=5)
    np.savez_compressed(os.path.join(HERE, "gaas_bulk_ek.npz"),
                        kx=g[:, 0], ky=g[:, 1], kz=g[:, 2], E=g[:, 3])
    t = os.path.join(root, "PostNEMO3D/bin/test/")
    ek = np.frombuffer(nd_payload(t + "silicon4.nd_Ek"), "<f8").reshape(-1, 4)
    nbr = np.frombuffer(nd_payload(t + "silicon4.nd_nbrIndx"), "<i4").reshape(-1, 4)
    r = np.frombuffer(nd_payload(t + "silicon4.nd_rAtom"), "<f8").reshape(-1, 3)
    np.savez_compressed(os.path.join(HERE, "silicon4_meta.npz"),
                        E=ek[:, 3], nbr=nbr, r_first=r[:64], r_last=r[-64:],
                        r_sum=r.sum(axis=0), n_dangling=(nbr < 0).sum())


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/NEMO_3D")
