"""ctypes binding of libnemohip.so (the HIP implementation).

There is no CPU fallback: if the library is missing or the device cannot
be used, every entry point raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnemohip.so")

NH_OK = 0
LAYOUT_BLOCKS, LAYOUT_SK, LAYOUT_MATRIX_FREE = 0, 1, 2
RANDOM_RAND, RANDOM_ZLARNV = 0, 1
PREC_FP64, PREC_FP32 = 0, 1
PTR_HOST, PTR_DEVICE = 0, 1
SET_HV_STAGED = 1
SET_RING_STEPS = 2
SET_BLOCK_HV = 3
SET_SK_ROLES = 4
SET_SK_LANES = 5
SET_LZ_VECTORS = 6

_STATUS = {0: "NH_OK", 1: "NH_ERR_ARG", 2: "NH_ERR_HIP", 3: "NH_ERR_STATE",
           4: "NH_ERR_NOMEM", 5: "NH_ERR_NOCONV"}


class NHError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (_STATUS.get(status, status), msg))
        self.status = status


class NHModel(ctypes.Structure):
    _fields_ = [
        ("nb", ctypes.c_int32),
        ("n_atoms", ctypes.c_int32),
        ("n_species", ctypes.c_int32),
        ("parmat", ctypes.c_void_p),
        ("strnmat", ctypes.c_void_p),
        ("species", ctypes.c_void_p),
        ("is_cation", ctypes.c_void_p),
        ("nbr", ctypes.c_void_p),
        ("bond_q", ctypes.c_void_p),
        ("disp", ctypes.c_void_p),
        ("lattice", ctypes.c_double * 3),
        ("passivate", ctypes.c_int32),
        ("db_shift", ctypes.c_double),
        ("phi", ctypes.c_void_p),
    ]


class NHCplx(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


class NHLanczosOpts(ctypes.Structure):
    _fields_ = [
        ("max_iter", ctypes.c_int32),
        ("conv_tolerance", ctypes.c_double),
        ("num_eig_vb", ctypes.c_int32),
        ("num_eig_cb", ctypes.c_int32),
        ("erange_vb", ctypes.c_double * 2),
        ("erange_cb", ctypes.c_double * 2),
        ("conv_check_start", ctypes.c_int32),
        ("conv_check_skip", ctypes.c_int32),
        ("start", ctypes.c_int32),
        ("cyclic_len", ctypes.c_int32),
        ("cyclic_string", ctypes.c_void_p),
        ("start_vector", ctypes.c_void_p),
        ("stop_rule", ctypes.c_int32),
        ("want_vectors", ctypes.c_int32),
        ("random_seed", ctypes.c_uint32),
        ("restore_alpha", ctypes.c_void_p),
        ("restore_beta", ctypes.c_void_p),
        ("restore_len", ctypes.c_int32),
    ]


class NHBLanczosOpts(ctypes.Structure):
    _fields_ = [
        ("block_size", ctypes.c_int32),
        ("max_iter", ctypes.c_int32),
        ("conv_tolerance", ctypes.c_double),
        ("num_eig_vb", ctypes.c_int32),
        ("num_eig_cb", ctypes.c_int32),
        ("erange_vb", ctypes.c_double * 2),
        ("erange_cb", ctypes.c_double * 2),
        ("conv_check_start", ctypes.c_int32),
        ("conv_check_skip", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("start_block", ctypes.c_void_p),
        ("stop_rule", ctypes.c_int32),
        ("want_vectors", ctypes.c_int32),
        ("deflation_on", ctypes.c_int32),
        ("deflation_tol", ctypes.c_double),
    ]


class NHPCTraceminOpts(ctypes.Structure):
    _fields_ = [
        ("num_eig", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("want_range", ctypes.c_double * 2),
        ("all_range", ctypes.c_double * 2),
        ("cheb_degree", ctypes.c_int32),
        ("tolerance", ctypes.c_double),
        ("hnorm", ctypes.c_double),
        ("max_outer", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("start_block", ctypes.c_void_p),
    ]


class NHRefineOpts(ctypes.Structure):
    _fields_ = [
        ("depth", ctypes.c_int32),
        ("max_rounds", ctypes.c_int32),
        ("tolerance", ctypes.c_double),
        ("hnorm", ctypes.c_double),
    ]


class NHTraceminOpts(ctypes.Structure):
    _fields_ = [
        ("num_eig", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("ref_eig", ctypes.c_double),
        ("tolerance", ctypes.c_double),
        ("hnorm", ctypes.c_double),
        ("cgtol", ctypes.c_double),
        ("max_outer", ctypes.c_int32),
        ("cg_max", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("start_block", ctypes.c_void_p),
    ]


class NHVffParams(ctypes.Structure):
    _fields_ = [("alpha", ctypes.c_void_p), ("beta", ctypes.c_void_p),
                ("a0", ctypes.c_void_p), ("aa", ctypes.c_void_p),
                ("ba", ctypes.c_void_p), ("ca", ctypes.c_void_p),
                ("model", ctypes.c_int32), ("fix_surface", ctypes.c_int32)]


VFF_KEATING = 0
VFF_KEATING_STRAINED = 1


class NHEigResult(ctypes.Structure):
    _fields_ = [
        ("n_found", ctypes.c_int32),
        ("iterations", ctypes.c_int32),
        ("matvecs", ctypes.c_int32),
        ("eigenvalues", ctypes.c_void_p),
        ("residuals", ctypes.c_void_p),
        ("eigenvectors", ctypes.c_void_p),
        ("capacity", ctypes.c_int32),
        ("seconds", ctypes.c_double),
        ("hv_ms", ctypes.c_double),
        ("hv_timed", ctypes.c_int32),
        ("tridiag_alpha", ctypes.c_void_p),
        ("tridiag_beta", ctypes.c_void_p),
        ("tridiag_cap", ctypes.c_int32),
        ("hv_ms_vec", ctypes.c_double),
        ("hv_timed_vec", ctypes.c_int32),
    ]


# Every symbol declared in include/nemohip.h: name -> (restype, argtypes)
P = ctypes.c_void_p
SIGNATURES = {
    "nh_context_create": (ctypes.c_int, [P, ctypes.c_int, P]),
    "nh_context_destroy": (None, [P]),
    "nh_context_set": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int]),
    "nh_ham_set_precision": (ctypes.c_int, [P, ctypes.c_int]),
    "nh_context_dim": (ctypes.c_int64, [P]),
    "nh_ham_build": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int]),
    "nh_hv": (ctypes.c_int, [P, ctypes.c_double, P, P, ctypes.c_int]),
    "nh_hV": (ctypes.c_int, [P, ctypes.c_double, P, P, ctypes.c_int,
                             ctypes.c_int64, ctypes.c_int]),
    "nh_ham_export": (ctypes.c_int, [P, P, P]),
    "nh_context_mem_info": (ctypes.c_int, [P, P, P]),
    "nh_random_start": (ctypes.c_int, [ctypes.c_int, P, ctypes.c_int64, P]),
    "nh_device_alloc": (ctypes.c_int, [P, ctypes.c_int64, P]),
    "nh_device_free": (ctypes.c_int, [P, P]),
    "nh_memcpy_h2d": (ctypes.c_int, [P, P, P, ctypes.c_int64]),
    "nh_memcpy_d2h": (ctypes.c_int, [P, P, P, ctypes.c_int64]),
    "nh_dot": (ctypes.c_int, [P, ctypes.c_int64, P, P, P]),
    "nh_dot_noalloc": (ctypes.c_int, [P, ctypes.c_int64, P, P, P]),
    "nh_nrm2": (ctypes.c_int, [P, ctypes.c_int64, P, P]),
    "nh_axpby": (ctypes.c_int, [P, ctypes.c_int64, NHCplx, P, NHCplx, P]),
    "nh_eigensolve_lanczos": (ctypes.c_int, [P, P, P]),
    "nh_eigensolve_block_lanczos": (ctypes.c_int, [P, P, P]),
    "nh_eigensolve_tracemin": (ctypes.c_int, [P, P, P]),
    "nh_eigensolve_pctracemin": (ctypes.c_int, [P, P, P]),
    "nh_refine_eigenpairs": (ctypes.c_int, [P, P, ctypes.c_int32, P, ctypes.c_int,
                                            P, P, P, P]),
    "nh_vff_energy_grad": (ctypes.c_int, [P, P, P, P, P, P, P]),
    "nh_vff_energy_grad_model": (ctypes.c_int, [P, P, P, P, P]),
    "nh_vff_relax_model": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.c_double, P, P]),
    "nh_vff_relax": (ctypes.c_int, [P, P, P, P, ctypes.c_int, ctypes.c_double,
                                    P, P]),
    "nh_cnn_create": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_uint64, P]),
    "nh_cnn_destroy": (None, [P]),
    "nh_cnn_num_params": (ctypes.c_int64, [P]),
    "nh_cnn_get_params": (ctypes.c_int, [P, P]),
    "nh_cnn_set_params": (ctypes.c_int, [P, P]),
    "nh_cnn_forward": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, P]),
    "nh_cnn_backward": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, P, P]),
    "nh_cnn_train_step": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, P, P, P, ctypes.c_int,
                                         P, ctypes.c_int, P, ctypes.c_double,
                                         P, P]),
    "nh_cnn_bind": (ctypes.c_int, [P, P]),
    "nh_cnn_start_block": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, P, P, ctypes.c_int, P,
                                          ctypes.c_int]),
    "nh_shim_bind": (ctypes.c_int, [P]),
    "nh_shim_set_wavefunctions": (ctypes.c_int, [ctypes.c_int, P, P]),
    "nh_matvect_mult": (None, [P, P, P]),
    "nh_matmat_mult": (None, [P, P, P]),
    "nh_Hmult_spds_complete": (None, [P, P, ctypes.c_double, P]),
    "nh_Hmatmult_spds_complete": (None, [P, P, ctypes.c_double, P,
                                         ctypes.c_int]),
    "nh_last_error": (ctypes.c_char_p, []),
}

_LIB = None


def load():
    """Load libnemohip.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libnemohip.so not built: run "
                               "`python -m nemo3d_amd.build` (HIP path only, "
                               "no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _LIB = lib
    return _LIB


def check(status):
    if status != NH_OK:
        raise NHError(status, load().nh_last_error().decode())


def ptr(a):
    """Address of a numpy array (host) or a DeviceArray (device)."""
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(ctypes.c_void_p)
    if isinstance(a, DeviceArray):
        return a.ptr
    raise TypeError(type(a))


class DeviceArray:
    """Device allocation made through a libnemohip context.

    `owner` is the object holding the context (it must expose `.ctx`); the
    array keeps a reference to it so the context outlives the buffer.  An
    owner with a `_device_arrays` set (Hamiltonian) frees its live arrays
    before it destroys the context; an array whose owner's context is gone
    is not freed again."""

    def __init__(self, owner, shape, dtype=np.complex128):
        self.owner = owner
        ctx = owner.ctx
        self.shape = tuple(np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = ctypes.c_void_p()
        check(load().nh_device_alloc(ctx, self.nbytes, ctypes.byref(self.ptr)))
        reg = getattr(owner, "_device_arrays", None)
        if reg is not None:
            reg.add(self)

    @property
    def ctx(self):
        ctx = self.owner.ctx
        if not ctx:
            raise RuntimeError("DeviceArray: its context has been destroyed")
        return ctx

    @classmethod
    def from_host(cls, owner, a):
        a = np.ascontiguousarray(a)
        d = cls(owner, a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(load().nh_memcpy_h2d(self.ctx, self.ptr, ptr(a), self.nbytes))

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(load().nh_memcpy_d2h(self.ctx, ptr(out), self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr:
            ctx = self.owner.ctx
            if ctx:
                check(load().nh_device_free(ctx, self.ptr))
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
