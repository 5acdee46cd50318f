"""Calibration run for the PMC traffic counters.

nh_axpby (y = a x + b y) is a streaming kernel that reads x and y once and
writes y once.  On two device vectors far larger than the 256 MB Infinity
Cache every byte crosses HBM, so the FETCH_SIZE / WRITE_SIZE counters of
its dispatches, against 2 n 16 bytes read and n 16 bytes written, give the
counters' scale for the kernels profiled in the same way:

    rocprofv3 --kernel-include-regex k_axpby --pmc FETCH_SIZE -- \\
        python3 -m nemo3d_amd.calibrate
    python3 -m nemo3d_amd.calibrate [--gib 2] [--repeat 5]
"""
import argparse
import ctypes

from . import materials, model, native, structure
from .hamiltonian import Hamiltonian


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gib", type=float, default=2.0, help="size of each vector")
    ap.add_argument("--repeat", type=int, default=5)
    a = ap.parse_args()
    s = structure.binary_box((1, 1, 1), "Ga", "As", (0.56532,) * 3)
    H = Hamiltonian(model.TBModel(s, ["GaAs"], materials.BANDS_20_SPIN), layout=None)
    n = int(a.gib * (1 << 30)) // 16
    x = native.DeviceArray(H, n)
    y = native.DeviceArray(H, n)
    lib = native.load()
    zero, one = native.NHCplx(0.0, 0.0), native.NHCplx(1.0, 0.0)
    for _ in range(a.repeat):
        native.check(lib.nh_axpby(H.ctx, ctypes.c_int64(n), zero, x.ptr, zero, y.ptr))
        native.check(lib.nh_axpby(H.ctx, ctypes.c_int64(n), one, x.ptr, one, y.ptr))
    out = native.NHCplx()  # a one-entry dot product: waits for the stream
    native.check(lib.nh_dot(H.ctx, ctypes.c_int64(1), x.ptr, x.ptr, ctypes.byref(out)))
    print("axpby over %d complex entries (%.2f GB per vector), %d launches: "
          "%.3f GB read and %.3f GB written per launch" % (
              n, n * 16 / 1e9, 2 * a.repeat, 2 * n * 16 / 1e9, n * 16 / 1e9))
    x.free()
    y.free()
    H.close()


if __name__ == "__main__":
    main()
