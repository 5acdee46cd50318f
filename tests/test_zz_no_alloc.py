"""nh_dot_noalloc (the allocation-free dot product, scratch in a file-scope
device array) against nh_dot through the C ABI: the same kernels on the
same inputs give bit-identical results, on two calls in a row, the second
with a smaller input."""
import ctypes

import numpy as np
import pytest

import _models
from nemo3d_amd import native
from nemo3d_amd.hamiltonian import Hamiltonian

pytestmark = pytest.mark.gpu


def _dot(H, fn, n, x, y):
    out = native.NHCplx()
    native.check(getattr(native.load(), fn)(H.ctx, n, native.ptr(x), native.ptr(y),
                                            ctypes.byref(out)))
    return complex(out.re, out.im)


def test_dot_noalloc_matches_dot():
    M = _models.gaas_box(cells=(2, 2, 2))
    H = Hamiltonian(M, layout=None)
    rng = np.random.default_rng(71)
    for n in (2_500_003, 1_201):  # the second call smaller (fewer partials)
        x = native.DeviceArray.from_host(H, rng.normal(size=n) + 1j * rng.normal(size=n))
        y = native.DeviceArray.from_host(H, rng.normal(size=n) + 1j * rng.normal(size=n))
        a = _dot(H, "nh_dot_noalloc", n, x, y)
        b = _dot(H, "nh_dot", n, x, y)
        print("n", n, "noalloc", a, "dot", b, flush=True)
        assert a == b
        ref = np.vdot(x.download(), y.download())
        assert abs(a - ref) <= 1e-12 * n
    # a second context on the same device shares the scratch address
    H2 = Hamiltonian(M, layout=None)
    x = native.DeviceArray.from_host(H2, np.ones(10, dtype=np.complex128))
    assert _dot(H2, "nh_dot_noalloc", 10, x, x) == 10.0
