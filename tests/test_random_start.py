"""The reference's random start vectors (nh_random_start, host code of the
library): glibc rand() against the C library's own rand(), and LAPACK
zlarnv(idist = 1) against its published generator."""
import ctypes
import ctypes.util
import os

import numpy as np
import pytest

from nemo3d_amd import native

pytestmark = pytest.mark.skipif(not os.path.exists(native.LIB_PATH),
                                reason="library not built")


def start(kind, seed, n):
    out = np.empty(n, dtype=np.complex128)
    s = np.ascontiguousarray(seed, dtype=np.int32)
    native.check(native.load().nh_random_start(kind, native.ptr(s), n, native.ptr(out)))
    return out


@pytest.mark.parametrize("seed", [1, 0, 1001, 123456789])
def test_rand_start_is_glibc_rand(seed):
    """(rand() / RAND_MAX, 0) after srand(seed), entry for entry, equal to
    the C library's rand() of this host (glibc)."""
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.rand.restype = ctypes.c_int
    libc.srand.argtypes = [ctypes.c_uint]
    n = 5000
    libc.srand(seed)
    want = np.array([1.0 * libc.rand() / 2147483647 for _ in range(n)])
    got = start(native.RANDOM_RAND, [seed], n)
    assert np.array_equal(got.real, want)
    assert not got.imag.any()


def test_zlarnv_is_the_fishman_generator():
    """zlarnv(1, iseed, n): real and imaginary parts are consecutive values
    s_k / 2^48 of the congruential sequence s_k = a^k s_0 mod 2^48,
    a = 33952834046453 (dlaruv's multiplier; the chunking into 64 complex
    numbers per dlaruv call continues the sequence), each exactly
    representable, with the reference's rank-0 seed {103, 0, 7, 57}."""
    a, M = 33952834046453, 2 ** 48
    for seed in ([103, 0, 7, 57], [104, 101, 7, 57], [0, 0, 0, 1]):
        s = (seed[0] << 36) | (seed[1] << 24) | (seed[2] << 12) | seed[3]
        n = 300  # several dlaruv chunks
        got = start(native.RANDOM_ZLARNV, seed, n)
        u = []
        for _ in range(2 * n):
            s = s * a % M
            u.append(s / M)
        u = np.array(u)
        assert np.array_equal(got.real, u[0::2]) and np.array_equal(got.imag, u[1::2])
        assert (got.real > 0).all() and (got.real < 1).all()


def test_zlarnv_rejects_bad_seed():
    with pytest.raises(native.NHError):
        start(native.RANDOM_ZLARNV, [103, 0, 7, 56], 4)  # seed[3] must be odd
    with pytest.raises(native.NHError):
        start(native.RANDOM_ZLARNV, [4096, 0, 7, 57], 4)
