"""ctypes driver of the host kernel simulator (see build_sim.py)."""
import ctypes
import os

import numpy as np

from . import build_sim

_lib = None
# Plain build for in-process use (AddressSanitizer needs its runtime preloaded:
# run build_sim.py directly and LD_PRELOAD libasan for that).
NOSAN = os.path.join(build_sim.OUT, "libhpdex_cpusim_nosan.so")
PAD = 16   # kMatrixPad: entries a resident matrix keeps past its end


def lib(path=None):
    global _lib
    if _lib is None:
        path = path or NOSAN
        srcs = [os.path.join(build_sim.CSRC, f) for f in ("hpdex_kernels.hip", "hpdex_device.h",
                                                          "hpdex_internal.h")]
        srcs += [os.path.join(build_sim.HERE, f) for f in ("sim_main.cpp", "sim_runtime.h")]
        def stale():
            return not os.path.exists(path) or any(os.path.getmtime(x) > os.path.getmtime(path)
                                                   for x in srcs)
        if stale():
            import fcntl
            os.makedirs(build_sim.OUT, exist_ok=True)
            with open(os.path.join(build_sim.OUT, ".lib.lock"), "w") as lk:
                fcntl.flock(lk, fcntl.LOCK_EX)
                if stale():   # another worker may have built it meanwhile
                    build_sim.build(sanitize=False, opt="-O2", lib=path)
        L = ctypes.CDLL(path)
        vp, ci = ctypes.c_void_p, ctypes.c_int
        L.sim_mwu.argtypes = [ci, vp, vp, vp, ci, ci, vp, ci, ci, ci, ci, ci, ci, ci, ci,
                              vp, vp, vp, vp, vp, vp, vp, ctypes.c_double]
        L.sim_mwu.restype = ci
        L.sim_count_blocks.restype = ci
        L.sim_set_shrink_range.argtypes = [ci]
        L.sim_set_shrink_range.restype = None
        L.sim_set_pack.argtypes = [ci]
        L.sim_set_pack.restype = None
        L.sim_set_split_arrival.argtypes = [ci]
        L.sim_set_split_arrival.restype = None
        L.sim_set_ref_cap.argtypes = [ci]
        L.sim_set_ref_cap.restype = None
        L.sim_packed_blocks.restype = ci
        L.sim_set_refcache.argtypes = [ci]
        L.sim_set_refcache.restype = None
        L.sim_cached_columns.restype = ci
        L.sim_set_global_split.argtypes = [ci]
        L.sim_set_global_split.restype = None
        L.sim_set_members.argtypes = [ci]
        L.sim_set_members.restype = None
        _lib = L
    return _lib


def mwu(data, indices, indptr, group_id, n_targets, n_rows, mode=1, alt=2, tie=True, cc=True,
        count_lds=40448, block=256, hist_max=8192, f64=False, timeout_s=30.0,
        shrink_range=False, pack=True, split_arrival=True, ref_cap=8192, refcache=False,
        global_split=False, members=True):
    """Run the launch sequence on the host.  count_lds: dynamic LDS of the
    count kernel (0 = every column takes the split/reference/target kernels).
    refcache: run the job, keep its reference tables with the fill kernel,
    then run it again with the job's own tables wiped, ranking every kept
    column from the cache (mwu.last_cached: the kept columns).  global_split:
    the split with its group counters in memory (the form taken when the
    groups' counters do not fit LDS), whatever the group count.  members: the
    job's member list (columns in row order gather their groups' rows through
    it; without it they are read entry by entry)."""
    # laid out like a resident matrix: PAD padding entries (value 0, row -1)
    vals = np.concatenate([np.asarray(data, dtype=np.float64 if f64 else np.float32),
                           np.zeros(PAD, dtype=np.float64 if f64 else np.float32)])
    rows = np.concatenate([np.where((indices >= 0) & (indices < n_rows), indices, -1),
                           np.full(PAD, -1)]).astype(np.int32)
    indptr = np.ascontiguousarray(indptr, dtype=np.int64)
    gid = np.ascontiguousarray(group_id, dtype=np.int32)
    C = len(indptr) - 1
    n = C * n_targets
    outs = [np.zeros(max(n, 1)) for _ in range(5)]
    err = np.zeros(1, dtype=np.uint64)
    status = np.zeros(1, dtype=np.uint32)
    lib().sim_set_shrink_range(int(shrink_range))   # test hook: metadata that does not match
    lib().sim_set_pack(int(pack))   # packed entries for the count kernel where they fit
    lib().sim_set_split_arrival(int(split_arrival))
    lib().sim_set_ref_cap(int(ref_cap))   # the LDS/global sort cutover of the reference
    lib().sim_set_refcache(int(refcache))
    lib().sim_set_global_split(int(global_split))
    lib().sim_set_members(int(members))
    counted = lib().sim_mwu(int(f64), vals.ctypes.data, rows.ctypes.data, indptr.ctypes.data, n_rows, C,
                  gid.ctypes.data, n_targets, mode, alt, int(tie), int(cc), count_lds, block, hist_max,
                  *[o.ctypes.data for o in outs], err.ctypes.data, status.ctypes.data,
                  float(timeout_s))
    mwu.last_counted = counted   # columns routed to the count kernel
    mwu.last_blocks = lib().sim_count_blocks()   # its workgroups (column x target slice)
    mwu.last_packed = lib().sim_packed_blocks()   # of them, those that read packed entries
    mwu.last_cached = lib().sim_cached_columns()
    return tuple(o[:n] for o in outs) + (int(err[0]), int(status[0]))
