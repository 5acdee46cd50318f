"""Solver drivers on the device Hamiltonian.

`eigensolve_lanczos` keeps the knobs of the reference's Lanczos deck block
(``Lanczos_struct``: MaxIter, ConvTolerance, NumEigReq_vb/cb, Erange_vb/cb,
ConvCheckStartIter, ConvCheckSkipRate, Start, CyclicString) and the
semantics of ``eigensolve_lanczos`` (NEMO_3D/eigen/src/lanczos.c:1823-2089);
the iteration runs in libnemohip on the GPU.
"""
import ctypes
import time

import numpy as np

from . import native

DEFAULT_CYCLIC = (1.2, 0.8, -1.4, 0.1, -0.9, -0.2, 0.6)
START = {"CyclicRandom": 0, "Const1": 1, "Vector": 2, "DeviceVector": 3,
         "Random": 4}


def eigensolve_lanczos(H, num_vb=4, num_cb=4, erange_vb=(-0.6, 0.2),
                       erange_cb=(1.4, 2.0), max_iter=8000, tol=1e-7,
                       start_iter=200, skip=50, start="CyclicRandom",
                       cyclic=DEFAULT_CYCLIC, start_vector=None,
                       want_vectors=True, capacity=64, return_vectors=False,
                       return_tridiag=False, stop_rule="window",
                       vectors_out=None, random_seed=1, restore=None):
    """stop_rule "window": stop once num_vb / num_cb converged values lie in
    the windows (the reference's rule, lanczos.c:1231-1232); "nearest":
    stop once the num_vb / num_cb Ritz values nearest the gap are all
    converged.  start "Random": (rand() / RAND_MAX, 0) per entry after
    srand(random_seed), glibc's rand() (eigsys3d_par.c:125-131).
    restore=(alpha, beta): restart from the tridiagonal of an earlier run
    with the same start (the result's "alpha"/"beta" with
    return_tridiag=True, or ndio.read_tridiagonal of a saved one): no
    eigenvalue pass, only the analysis and the eigenvector pass
    (restoreState, lanczos.c:1905-1960)."""
    lib = native.load()
    o = native.NHLanczosOpts()
    o.max_iter = int(max_iter)
    o.conv_tolerance = float(tol)
    o.num_eig_vb = int(num_vb)
    o.num_eig_cb = int(num_cb)
    o.erange_vb[0], o.erange_vb[1] = map(float, erange_vb)
    o.erange_cb[0], o.erange_cb[1] = map(float, erange_cb)
    o.conv_check_start = int(start_iter)
    o.conv_check_skip = int(skip)
    cyc = np.ascontiguousarray(cyclic, dtype=np.float64)
    o.cyclic_len = cyc.size
    o.cyclic_string = native.ptr(cyc)
    if start_vector is None:
        o.start = START[start]
    elif isinstance(start_vector, native.DeviceArray):  # stays on the device
        assert start_vector.shape[-1] == H.n and start_vector.dtype == np.complex128
        o.start = START["DeviceVector"]
        o.start_vector = start_vector.ptr
    else:
        o.start = START["Vector"]
        sv = np.ascontiguousarray(start_vector, dtype=np.complex128)
        o.start_vector = native.ptr(sv)
    o.stop_rule = {"window": 0, "nearest": 1}[stop_rule]
    o.want_vectors = int(bool(want_vectors))
    o.random_seed = int(random_seed)
    if restore is not None:
        ra = np.ascontiguousarray(restore[0], dtype=np.float64)
        rb = np.ascontiguousarray(restore[1], dtype=np.float64)
        assert ra.shape == rb.shape and ra.size <= max_iter
        o.restore_alpha, o.restore_beta = native.ptr(ra), native.ptr(rb)
        o.restore_len = ra.size
    ev = np.zeros(capacity)
    rs = np.full(capacity, -1.0)
    vec = np.zeros((capacity, H.n), dtype=np.complex128) if return_vectors else None
    if vectors_out is not None:  # device array (capacity, N): vectors stay on the GPU
        assert vectors_out.shape == (capacity, H.n) and vec is None
        vec = vectors_out
    r = native.NHEigResult()
    r.eigenvalues = native.ptr(ev)
    r.residuals = native.ptr(rs)
    r.eigenvectors = native.ptr(vec) if vec is not None else None
    r.capacity = capacity
    if return_tridiag:
        ta = np.zeros(max_iter + 1)
        tb = np.zeros(max_iter + 1)
        r.tridiag_alpha = native.ptr(ta)
        r.tridiag_beta = native.ptr(tb)
        r.tridiag_cap = max_iter + 1
    st = lib.nh_eigensolve_lanczos(H.ctx, ctypes.byref(o), ctypes.byref(r))
    if st not in (native.NH_OK, 5):
        native.check(st)
    n = r.n_found
    out = {"eigenvalues": ev[:n].copy(), "residuals": rs[:n].copy(),
           "iterations": r.iterations, "matvecs": r.matvecs,
           "seconds": r.seconds, "converged": st == native.NH_OK,
           "hv_ms": r.hv_ms, "hv_timed": r.hv_timed,
           "hv_ms_vec": r.hv_ms_vec, "hv_timed_vec": r.hv_timed_vec}
    if return_tridiag:
        out["alpha"] = ta[:r.iterations].copy()
        out["beta"] = tb[:r.iterations].copy()
    if vec is not None:
        out["vectors"] = vec if vectors_out is not None else vec[:n]
    return out


def eigensolve_block_lanczos(H, block_size=8, num_vb=4, num_cb=4,
                             erange_vb=(-0.6, 0.2), erange_cb=(1.4, 2.0),
                             max_iter=2000, tol=1e-8, start_iter=20, skip=10,
                             seed=1, start_block=None, want_vectors=True,
                             capacity=64, return_vectors=False,
                             stop_rule="window", deflation=False,
                             deflation_tol=1e-7, vectors_out=None):
    """Block Lanczos (eigensolve_blk_lanczos, NEMO_3D/eigen/src/
    lanczos_blk.c:2145-...) with BL_size = block_size.  num_vb / num_cb
    count distinct levels; stop_rule as in eigensolve_lanczos.  Without a
    start block the reference's random block (rand() after srand(seed)).
    deflation / deflation_tol: DeflationOn / DefTolerance (a block vector
    whose norm after orthogonalisation is below the tolerance is dropped
    and the block shrinks, deflation_test, lanczos_blk.c:225-268)."""
    lib = native.load()
    o = native.NHBLanczosOpts()
    o.block_size = int(block_size)
    o.max_iter = int(max_iter)
    o.conv_tolerance = float(tol)
    o.num_eig_vb = int(num_vb)
    o.num_eig_cb = int(num_cb)
    o.erange_vb[0], o.erange_vb[1] = map(float, erange_vb)
    o.erange_cb[0], o.erange_cb[1] = map(float, erange_cb)
    o.conv_check_start = int(start_iter)
    o.conv_check_skip = int(skip)
    o.seed = int(seed)
    if start_block is not None:
        sb = np.ascontiguousarray(start_block, dtype=np.complex128)
        assert sb.shape == (block_size, H.n)
        o.start_block = native.ptr(sb)
    o.stop_rule = {"window": 0, "nearest": 1}[stop_rule]
    o.want_vectors = int(bool(want_vectors))
    o.deflation_on = int(bool(deflation))
    o.deflation_tol = float(deflation_tol)
    ev = np.zeros(capacity)
    rs = np.full(capacity, -1.0)
    vec = np.zeros((capacity, H.n), dtype=np.complex128) if return_vectors else None
    if vectors_out is not None:  # device array (capacity, N): vectors stay on the GPU
        assert vectors_out.shape == (capacity, H.n) and vec is None
        vec = vectors_out
        o.want_vectors = 1
    r = native.NHEigResult()
    r.eigenvalues = native.ptr(ev)
    r.residuals = native.ptr(rs)
    r.eigenvectors = native.ptr(vec) if vec is not None else None
    r.capacity = capacity
    st = lib.nh_eigensolve_block_lanczos(H.ctx, ctypes.byref(o), ctypes.byref(r))
    if st not in (native.NH_OK, 5):
        native.check(st)
    n = r.n_found
    out = {"eigenvalues": ev[:n].copy(), "residuals": rs[:n].copy(),
           "iterations": r.iterations, "matvecs": r.matvecs,
           "seconds": r.seconds, "converged": st == native.NH_OK,
           "hv_ms": r.hv_ms, "hv_timed": r.hv_timed}
    if vec is not None:
        out["vectors"] = vec if vectors_out is not None else vec[:n]
    return out


def eigensolve_tracemin(H, num_eig=4, ref_eig=0.7, block_size=0, tol=1e-8,
                        hnorm=40.0, cgtol=1.0, max_outer=40, cg_max=100,
                        seed=0, start_block=None, capacity=64,
                        return_vectors=False):
    """PQ-TRACEMIN on (H - ref_eig)^2 (qtracemin_driver,
    NEMO_3D/nemo3d/src/base/qtracemin_driver.c:9-947).  Without a start
    block the reference's zlarnv block of rank `seed` (iseed = {103 +
    seed, 101 seed, 7, 57})."""
    lib = native.load()
    o = native.NHTraceminOpts()
    o.num_eig = int(num_eig)
    o.block_size = int(block_size)
    o.ref_eig = float(ref_eig)
    o.tolerance = float(tol)
    o.hnorm = float(hnorm)
    o.cgtol = float(cgtol)
    o.max_outer = int(max_outer)
    o.cg_max = int(cg_max)
    o.seed = int(seed)
    if start_block is not None:
        sb = np.ascontiguousarray(start_block, dtype=np.complex128)
        o.start_block = native.ptr(sb)
    ev = np.zeros(capacity)
    rs = np.full(capacity, -1.0)
    vec = np.zeros((capacity, H.n), dtype=np.complex128) if return_vectors else None
    r = native.NHEigResult()
    r.eigenvalues = native.ptr(ev)
    r.residuals = native.ptr(rs)
    r.eigenvectors = native.ptr(vec) if vec is not None else None
    r.capacity = capacity
    st = lib.nh_eigensolve_tracemin(H.ctx, ctypes.byref(o), ctypes.byref(r))
    if st not in (native.NH_OK, 5):
        native.check(st)
    n = r.n_found
    out = {"eigenvalues": ev[:n].copy(), "residuals": rs[:n].copy(),
           "iterations": r.iterations, "matvecs": r.matvecs,
           "seconds": r.seconds, "converged": st == native.NH_OK}
    if vec is not None:
        out["vectors"] = vec[:n]
    return out


def eigensolve_pctracemin(H, num_eig=4, want_range=(0.6, 0.7),
                          all_range=(-13.0, 40.0), block_size=0, cheb_degree=20,
                          tol=1e-8, hnorm=40.0, max_outer=200, seed=0,
                          start_block=None, capacity=16, return_vectors=False):
    """PC-TRACEMIN (ctracemin_driver, NEMO_3D/nemo3d/src/base/
    ctracemin_driver.c): subspace iteration on T_k(1 - gamma (H-a)(H-b)) + 2I
    with Rayleigh-Ritz of H; the num_eig Ritz pairs in want_range."""
    lib = native.load()
    o = native.NHPCTraceminOpts()
    o.num_eig, o.block_size = int(num_eig), int(block_size)
    o.want_range[0], o.want_range[1] = map(float, want_range)
    o.all_range[0], o.all_range[1] = map(float, all_range)
    o.cheb_degree, o.tolerance, o.hnorm = int(cheb_degree), float(tol), float(hnorm)
    o.max_outer, o.seed = int(max_outer), int(seed)
    p = block_size if block_size > 0 else min(16, num_eig + 4)
    if start_block is not None:
        sb = np.ascontiguousarray(start_block, dtype=np.complex128)
        assert sb.shape == (p, H.n)
        o.start_block = native.ptr(sb)
    ev = np.zeros(capacity)
    rs = np.zeros(capacity)
    vec = np.zeros((capacity, H.n), dtype=np.complex128) if return_vectors else None
    r = native.NHEigResult()
    r.eigenvalues = native.ptr(ev)
    r.residuals = native.ptr(rs)
    r.eigenvectors = native.ptr(vec) if vec is not None else None
    r.capacity = capacity
    st = lib.nh_eigensolve_pctracemin(H.ctx, ctypes.byref(o), ctypes.byref(r))
    if st not in (native.NH_OK, 5):
        native.check(st)
    n = r.n_found
    out = {"eigenvalues": ev[:n].copy(), "residuals": rs[:n].copy(),
           "iterations": r.iterations, "matvecs": r.matvecs,
           "seconds": r.seconds, "converged": st == native.NH_OK}
    if vec is not None:
        out["vectors"] = vec[:n]
    return out


def refine_eigenpairs(H, X, eigenvalues, tol=1e-8, hnorm=40.0, depth=6,
                      max_rounds=4, row0=0):
    """FP64 refinement of approximate eigenpairs (nh_refine_eigenpairs):
    block Krylov [X, H X, ..., H^depth X] and refined Ritz vectors per round
    until every ||H x - e x|| <= tol * hnorm.  X: (k, N) host array (a
    refined copy is returned) or native.DeviceArray of shape (>= row0 + k,
    N) whose rows row0 .. row0 + k - 1 are refined in place (k =
    len(eigenvalues))."""
    lib = native.load()
    ev = np.array(eigenvalues, dtype=np.float64, copy=True)
    k = ev.shape[0]
    if isinstance(X, native.DeviceArray):
        assert X.dtype == np.complex128 and X.shape[0] >= row0 + k and X.shape[1] == H.n
        Xc, kind = X, native.PTR_DEVICE
        xp = ctypes.c_void_p(X.ptr.value + row0 * H.n * 16)
    else:
        Xc = np.array(X, dtype=np.complex128, order="C", copy=True)
        assert Xc.shape == (k, H.n) and row0 == 0
        kind = native.PTR_HOST
        xp = native.ptr(Xc)
    assert 1 <= k <= 16
    rs = np.zeros(k)
    o = native.NHRefineOpts()
    o.depth, o.max_rounds = int(depth), int(max_rounds)
    o.tolerance, o.hnorm = float(tol), float(hnorm)
    rounds, mv = ctypes.c_int32(), ctypes.c_int32()
    t0 = time.perf_counter()
    st = lib.nh_refine_eigenpairs(H.ctx, ctypes.byref(o), k, xp, kind, native.ptr(ev), native.ptr(rs),
                                  ctypes.byref(rounds), ctypes.byref(mv))
    if st not in (native.NH_OK, 5):
        native.check(st)
    return {"eigenvalues": ev, "vectors": Xc, "residuals": rs,
            "rounds": rounds.value, "matvecs": mv.value,
            "seconds": time.perf_counter() - t0,
            "converged": st == native.NH_OK}


def eigensolve_mixed(H, num_vb=4, num_cb=4, erange_vb=(-0.6, 0.2),
                     erange_cb=(1.4, 2.0), tol=1e-8, tol32=None, hnorm=40.0,
                     max_iter=30000, start_iter=200, skip=50, start_vector=None,
                     stop_rule="window", refine_depth=12, refine_rounds=4,
                     return_vectors=False, refine_split=True):
    """Mixed-precision Lanczos: the Lanczos run and its eigenvector pass
    with FP32 H·v (FP32 values and products inside each block, FP64 sums
    over the bonds) and, with the SK and matrix-free layouts, FP32 Lanczos
    vectors (nh_context_set NH_SET_LZ_VECTORS, on by default), while the
    step's arithmetic, the dots and the tridiagonal stay FP64; then FP64
    residual refinement (refine_eigenpairs) of its Ritz
    pairs to ||H x - e x|| <= tol * hnorm.  return_vectors: True gives the
    eigenvectors on the host, "device" the native.DeviceArray that holds
    them (rows in the order of `order` in the result).  tol32: convergence tolerance of
    the FP32 stage (default tol).  refine_split: refine the valence and the
    conduction pairs as two blocks (default) instead of one.  start_vector:
    optional start, e.g. from
    the learned start block."""
    from . import native as nt
    prec0 = H.precision
    cap = num_vb + num_cb
    # the eigenvectors stay in device memory from the FP32 run through the
    # refinement; only the requested output goes to the host
    dv = nt.DeviceArray(H, (cap, H.n))
    try:
        H.set_precision(nt.PREC_FP32)
        r1 = eigensolve_lanczos(
            H, num_vb=num_vb, num_cb=num_cb, erange_vb=erange_vb,
            erange_cb=erange_cb, max_iter=max_iter,
            tol=tol if tol32 is None else tol32, start_iter=start_iter,
            skip=skip, start_vector=start_vector, want_vectors=True,
            capacity=cap, stop_rule=stop_rule, vectors_out=dv)
        H.set_precision(nt.PREC_FP64)
        n = len(r1["eigenvalues"])
        # refined per side of the gap (valence levels, conduction levels):
        # the two groups are far apart in energy, so one joint block helps
        # neither and its Gram-Schmidt costs twice the two blocks'
        ev32 = np.asarray(r1["eigenvalues"])
        split = int(np.searchsorted(ev32, 0.5 * (erange_vb[1] + erange_cb[0])))
        groups = [(a, b) for a, b in ((0, split), (split, n)) if b > a]
        if not refine_split:
            groups = [(0, n)] if n else []
        rf = {"eigenvalues": np.zeros(n), "residuals": np.zeros(n), "rounds": 0,
              "matvecs": 0, "seconds": 0.0, "converged": n > 0}
        for a, b in groups:
            g = refine_eigenpairs(H, dv, ev32[a:b], tol=tol, hnorm=hnorm,
                                  depth=refine_depth, max_rounds=refine_rounds,
                                  row0=a)
            rf["eigenvalues"][a:b] = g["eigenvalues"]
            rf["residuals"][a:b] = g["residuals"]
            rf["rounds"] = max(rf["rounds"], g["rounds"])
            rf["matvecs"] += g["matvecs"]
            rf["seconds"] += g["seconds"]
            rf["converged"] = rf["converged"] and g["converged"]
    finally:
        if prec0 is not None:
            H.set_precision(prec0)
    order = np.argsort(rf["eigenvalues"])
    out = {"eigenvalues": rf["eigenvalues"][order],
           "residuals": rf["residuals"][order],
           "iterations": r1["iterations"],
           "matvecs": r1["matvecs"] + rf["matvecs"],
           "seconds": r1["seconds"] + rf["seconds"],
           "converged": bool(r1["converged"] and rf["converged"]),
           "hv_ms": r1["hv_ms"], "hv_timed": r1["hv_timed"],
           "stage32": {k: r1[k] for k in ("iterations", "matvecs", "seconds",
                                          "converged", "hv_ms", "hv_timed",
                                          "hv_ms_vec", "hv_timed_vec",
                                          "eigenvalues", "residuals")},
           "refine": {k: rf[k] for k in ("rounds", "matvecs", "seconds",
                                         "converged")}}
    if return_vectors == "device":
        out["device_vectors"] = dv
        out["order"] = order
    elif return_vectors:
        out["vectors"] = dv.download()[:len(order)][order] if n else np.zeros((0, H.n))
    return out


def eigensolve_mixed_block(H, block_size=8, num_vb=4, num_cb=4, erange_vb=(-0.6, 0.2),
                           erange_cb=(1.4, 2.0), tol=1e-8, tol32=1e-6,
                           max_iter32=2000, max_iter64=400, seed=1, start_block=None,
                           stop_rule="window", return_vectors=False, capacity=64,
                           stage2="block", hnorm=40.0, refine_depth=12,
                           refine_rounds=4, deflation=False):
    """Mixed-precision block Lanczos: the bulk iterations with FP32 H·V
    (FP32 values, FP32 products inside each block; on the matrix cores for
    a block of 8) and FP64 recurrences, dots and Ritz solves, to an
    FP32-level tolerance `tol32`; then, stage2 = "block", FP64 block
    Lanczos started from those Ritz vectors to `tol`, with FP64 Rayleigh
    quotients and residuals, or, stage2 = "refine", the FP64 residual
    refinement of eigensolve_mixed (refine_eigenpairs per side of the gap,
    ||H x - e x|| <= tol * hnorm) on the device-resident Ritz vectors.
    start_block: optional (block_size, N) start, e.g. the learned one."""
    from . import native as nt
    prec0 = H.precision
    t_all = 0.0
    if stage2 == "refine":
        cap = num_vb + num_cb
        dv = nt.DeviceArray(H, (cap, H.n))
        try:
            H.set_precision(nt.PREC_FP32)
            r1 = eigensolve_block_lanczos(
                H, block_size=block_size, num_vb=num_vb, num_cb=num_cb,
                erange_vb=erange_vb, erange_cb=erange_cb, max_iter=max_iter32,
                tol=tol32, start_iter=4, skip=4, seed=seed, start_block=start_block,
                capacity=cap, stop_rule=stop_rule, deflation=deflation,
                vectors_out=dv)
            H.set_precision(nt.PREC_FP64)
            n = len(r1["eigenvalues"])
            ev32 = np.asarray(r1["eigenvalues"])
            split = int(np.searchsorted(ev32, 0.5 * (erange_vb[1] + erange_cb[0])))
            rf = {"eigenvalues": np.zeros(n), "residuals": np.zeros(n), "rounds": 0,
                  "matvecs": 0, "seconds": 0.0, "converged": n > 0}
            for a, b in ((0, split), (split, n)):
                if b <= a:
                    continue
                g = refine_eigenpairs(H, dv, ev32[a:b], tol=tol, hnorm=hnorm,
                                      depth=refine_depth, max_rounds=refine_rounds,
                                      row0=a)
                rf["eigenvalues"][a:b] = g["eigenvalues"]
                rf["residuals"][a:b] = g["residuals"]
                rf["rounds"] = max(rf["rounds"], g["rounds"])
                rf["matvecs"] += g["matvecs"]
                rf["seconds"] += g["seconds"]
                rf["converged"] = rf["converged"] and g["converged"]
        finally:
            if prec0 is not None:
                H.set_precision(prec0)
        order = np.argsort(rf["eigenvalues"])
        out = {"eigenvalues": rf["eigenvalues"][order],
               "residuals": rf["residuals"][order],
               "iterations": r1["iterations"],
               "matvecs": r1["matvecs"] + rf["matvecs"],
               "seconds": r1["seconds"] + rf["seconds"],
               "converged": bool(r1["converged"] and rf["converged"]),
               "hv_ms": r1["hv_ms"], "hv_timed": r1["hv_timed"],
               "stage32": {k: r1[k] for k in ("iterations", "matvecs", "seconds",
                                              "converged", "hv_ms", "hv_timed",
                                              "eigenvalues")},
               "refine": {k: rf[k] for k in ("rounds", "matvecs", "seconds",
                                             "converged")}}
        if return_vectors:
            out["vectors"] = dv.download()[:len(order)][order] if n else np.zeros((0, H.n))
        return out
    try:
        H.set_precision(nt.PREC_FP32)
        r1 = eigensolve_block_lanczos(
            H, block_size=block_size, num_vb=num_vb, num_cb=num_cb,
            erange_vb=erange_vb, erange_cb=erange_cb, max_iter=max_iter32,
            tol=tol32, start_iter=4, skip=4, seed=seed, start_block=start_block,
            return_vectors=True, capacity=capacity, stop_rule=stop_rule,
            deflation=deflation)
        t_all += r1["seconds"]
        V = r1["vectors"]
        if len(V) < block_size:
            rng = np.random.default_rng(seed + 1)
            extra = rng.random((block_size - len(V), H.n)) + 1j * rng.random(
                (block_size - len(V), H.n))
            V = np.concatenate([V, extra]) if len(V) else extra
        V = V[:block_size]
        H.set_precision(nt.PREC_FP64)
        r2 = eigensolve_block_lanczos(
            H, block_size=block_size, num_vb=num_vb, num_cb=num_cb,
            erange_vb=erange_vb, erange_cb=erange_cb, max_iter=max_iter64,
            tol=tol, start_iter=1, skip=1, start_block=V,
            return_vectors=return_vectors, capacity=capacity,
            stop_rule=stop_rule, deflation=deflation)
        t_all += r2["seconds"]
    finally:
        if prec0 is not None:
            H.set_precision(prec0)
    out = dict(r2)
    out["seconds"] = t_all
    out["stage32"] = {k: r1[k] for k in ("iterations", "matvecs", "seconds",
                                          "converged", "hv_ms", "hv_timed")}
    out["stage32"]["eigenvalues"] = r1["eigenvalues"]
    out["stage64"] = {k: r2[k] for k in ("iterations", "matvecs", "seconds",
                                          "converged", "hv_ms", "hv_timed")}
    out["matvecs"] = r1["matvecs"] + r2["matvecs"]
    return out
