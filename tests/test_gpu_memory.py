"""Device memory of the library: the solvers' workspace pool is handed back
when the device fills up, and device arrays outlive nothing they need."""
import numpy as np
import pytest
import ctypes

import _models
from nemo3d_amd import native
from nemo3d_amd.hamiltonian import Hamiltonian
from nemo3d_amd.solvers import refine_eigenpairs

pytestmark = pytest.mark.gpu


def mem_info(H):
    fr, idle = ctypes.c_int64(), ctypes.c_int64()
    native.check(native.load().nh_context_mem_info(H.ctx, ctypes.byref(fr),
                                                   ctypes.byref(idle)))
    return fr.value, idle.value


def test_allocation_takes_back_idle_workspace():
    """After a refinement (a 104-vector Krylov basis, ~0.8 GB, kept idle in
    the context's pool), an allocation the device cannot serve first
    returns the idle blocks of every context on the device, then fails
    cleanly (NH_ERR_NOMEM); an allocation of the device's free memory plus
    the idle pool minus 128 MB succeeds; the solver works afterwards."""
    M = _models.gaas_box(cells=(15, 15, 14), seed=3)
    H = Hamiltonian(M, layout=native.LAYOUT_SK)
    print("N =", H.n, flush=True)
    X = _models.random_vectors(H.n, 8, seed=4)
    refine_eigenpairs(H, X, np.linspace(0.5, 1.5, 8), depth=12, max_rounds=1)
    fr, idle = mem_info(H)
    print("free %.1f GB, idle pool %.2f GB" % (fr / 1e9, idle / 1e9), flush=True)
    assert idle > 512 << 20
    # a second context on the device sees the first one's idle blocks
    H2 = Hamiltonian(_models.gaas_box(cells=(2, 2, 2)), layout=None)
    assert mem_info(H2)[1] == idle
    with pytest.raises(native.NHError):
        native.DeviceArray(H2, 4 * (fr + idle), dtype=np.uint8)
    assert mem_info(H)[1] == 0  # the failed request took the idle blocks back
    H2.close()
    # refill the pool, then allocate across it
    refine_eigenpairs(H, X, np.linspace(0.5, 1.5, 8), depth=12, max_rounds=1)
    fr, idle = mem_info(H)
    big = native.DeviceArray(H, fr + idle - (128 << 20), dtype=np.uint8)
    big.free()
    r = refine_eigenpairs(H, X[:2], np.linspace(0.5, 1.5, 2), depth=4, max_rounds=1)
    assert np.isfinite(r["residuals"]).all()
    H.close()


def test_device_array_after_context_close():
    """Arrays made from a Hamiltonian are freed when it closes; freeing them
    again, or dropping them, touches no destroyed context, and using one
    raises."""
    M = _models.gaas_box(cells=(2, 2, 2))
    H = Hamiltonian(M, layout=native.LAYOUT_SK)
    a = native.DeviceArray.from_host(H, np.arange(H.n, dtype=np.complex128))
    assert np.array_equal(a.download(), np.arange(H.n))
    H.close()
    assert not a.ptr
    a.free()
    with pytest.raises(RuntimeError):
        a.download()
    del a
