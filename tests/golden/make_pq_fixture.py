"""Writes tests/golden/pqexample2_coo.npz: the Hamiltonian of the
reference's TRACEMIN example (NEMO_3D/TRACEMIN/Examples/pqexample2_full, a
960 x 960 complex Hermitian matrix written by NEMO-3D in coordinate form:
one line per nonzero, row and column from 1, real and imaginary part),
stored as int16 row / column indices from 0 and float64 values.  Run in a
container that has the reference tree:
  python tests/golden/make_pq_fixture.py
"""
import os

import numpy as np

SRC = "/root/reference/NEMO_3D/TRACEMIN/Examples/pqexample2_full"


def main():
    lines = open(SRC).read().split("\n")
    nnz, n = int(lines[0].split()[0]), int(lines[0].split()[1])
    rows = [l.split() for l in lines[1:] if l.strip()]
    assert len(rows) == nnz
    i = np.array([int(r[0]) - 1 for r in rows], dtype=np.int16)
    j = np.array([int(r[1]) - 1 for r in rows], dtype=np.int16)
    re = np.array([float(r[2]) for r in rows])
    im = np.array([float(r[3]) for r in rows])
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pqexample2_coo.npz")
    np.savez_compressed(dst, i=i, j=j, re=re, im=im, n=n)
    print("wrote", dst, os.path.getsize(dst), "bytes")


if __name__ == "__main__":
    main()
