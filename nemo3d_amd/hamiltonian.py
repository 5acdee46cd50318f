"""Device Hamiltonian: the drop-in for the reference's operator boundary.

`Hamiltonian.hv` is the counterpart of ``Hmult_spds_complete(y, d, shift,
x)`` (nemo3d/src/base/h_cvectr_mult.c:262): y = (H - shift I) x with y fully
overwritten; `hV` is ``Hmatmult_spds_complete`` (h_cmat_mult.c:85).  Vectors
may be numpy arrays (host, copied) or native.DeviceArray buffers (device,
used in place).
"""
import ctypes
import weakref

import numpy as np

from . import native


class Hamiltonian:
    def __init__(self, tbmodel, device=0, layout=native.LAYOUT_BLOCKS,
                 precision=native.PREC_FP64):
        self.model = tbmodel
        self._device_arrays = weakref.WeakSet()
        lib = native.load()
        arr = tbmodel.arrays()
        self._keep = {}
        m = native.NHModel()
        m.nb = arr["nb"]
        m.n_atoms = arr["n_atoms"]
        m.n_species = arr["n_species"]
        for k, dt in [("parmat", np.float64), ("strnmat", np.float64),
                      ("species", np.int32), ("is_cation", np.int32),
                      ("nbr", np.int32), ("bond_q", np.int32),
                      ("disp", np.float64)]:
            a = np.ascontiguousarray(arr[k], dtype=dt)
            self._keep[k] = a
            setattr(m, k, native.ptr(a))
        for d in range(3):
            m.lattice[d] = float(arr["lattice"][d])
        m.passivate = arr["passivate"]
        m.db_shift = arr["db_shift"]
        if arr.get("phi") is not None:
            a = np.ascontiguousarray(arr["phi"], dtype=np.float64)
            self._keep["phi"] = a
            m.phi = native.ptr(a)
        self._ctx = ctypes.c_void_p()
        native.check(lib.nh_context_create(ctypes.byref(m), device,
                                           ctypes.byref(self._ctx)))
        self.nb = arr["nb"]
        self.n_atoms = arr["n_atoms"]
        self.n = int(lib.nh_context_dim(self._ctx))
        self.layout = None
        self.precision = None
        if layout is not None:
            self.build(layout, precision)

    def close(self):
        if self._ctx:
            # device arrays made from this context are freed first
            for a in list(self._device_arrays):
                a.free()
            native.load().nh_context_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ctx(self):
        return self._ctx

    def build(self, layout=native.LAYOUT_BLOCKS, precision=native.PREC_FP64):
        native.check(native.load().nh_ham_build(self._ctx, layout, precision))
        self.layout, self.precision = layout, precision

    def set_precision(self, precision):
        """Arithmetic of later products (FP64 values are always kept)."""
        native.check(native.load().nh_ham_set_precision(self._ctx, int(precision)))
        self.precision = precision

    def set(self, what, value):
        """Kernel variant switch (nh_context_set)."""
        native.check(native.load().nh_context_set(self._ctx, int(what), int(value)))

    def hv(self, x, shift=0.0, out=None):
        """y = (H - shift I) x for one vector (numpy host or DeviceArray)."""
        return self.hV(x, shift, out)

    def hV(self, X, shift=0.0, out=None):
        """Y = (H - shift I) X for a (ncol, N) block or a single (N,) vector."""
        lib = native.load()
        if isinstance(X, np.ndarray):
            Xc = np.ascontiguousarray(X, dtype=np.complex128)
            ncol = 1 if Xc.ndim == 1 else Xc.shape[0]
            assert Xc.shape[-1] == self.n
            Y = np.empty_like(Xc) if out is None else out
            native.check(lib.nh_hV(self._ctx, float(shift), native.ptr(Xc),
                                   native.ptr(Y), ncol, self.n,
                                   native.PTR_HOST))
            return Y
        assert isinstance(X, native.DeviceArray)
        assert X.dtype == np.complex128 and X.shape[-1] == self.n
        ncol = 1 if len(X.shape) == 1 else X.shape[0]
        Y = native.DeviceArray(self, X.shape) if out is None else out
        native.check(lib.nh_hV(self._ctx, float(shift), native.ptr(X),
                               native.ptr(Y), ncol, self.n,
                               native.PTR_DEVICE))
        return Y

    def export(self):
        """Assembled FP64 Hamiltonian: onsite (A,17), blocks (A,4,10,10)."""
        onsite = np.zeros((self.n_atoms, 17))
        blocks = np.zeros((self.n_atoms, 4, 10, 10))
        native.check(native.load().nh_ham_export(
            self._ctx, native.ptr(onsite), native.ptr(blocks)))
        return onsite, blocks
