"""ctypes loader of the CPU harness of the pair kernel's run walk (tests/native).

The harness runs the kernel's lane program (picca_amd/csrc/xi_walk.h) on the host under
bounds checks; it is test infrastructure, built here with hipcc when missing.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCE = os.path.join(HERE, "native", "xi_walk_host.hip")
CSRC = os.path.join(os.path.dirname(HERE), "picca_amd", "csrc")
DEPENDS = [SOURCE] + [os.path.join(CSRC, h) for h in ("xi_walk.h", "xi_const.h", "engine.h")]
LIBRARY = os.path.join(HERE, "native", "libxi_walk_host.so")
STATS = ("edge", "overflow", "walk_tasks", "pair_tasks", "runs", "probes", "oob", "cover")

SCHEDULE_AUTO = 0
SCHEDULE_PAIRS = 1
SCHEDULE_RUNS = 2

_lib = None


def build():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    hipcc = os.path.join(rocm, "bin", "hipcc")
    # a per-process temporary name: parallel test workers may build at the same time
    cmd = [hipcc if os.path.exists(hipcc) else "hipcc", "--offload-arch=gfx950", "-O2",
           "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Wall",
           "-o", LIBRARY + ".tmp%d" % os.getpid(), SOURCE]
    subprocess.run(cmd, check=True)
    os.replace(LIBRARY + ".tmp%d" % os.getpid(), LIBRARY)
    return LIBRARY


def lib():
    global _lib
    if _lib is None:
        if (not os.path.exists(LIBRARY)
                or os.path.getmtime(LIBRARY) < max(os.path.getmtime(f) for f in DEPENDS)):
            build()
        _lib = ctypes.CDLL(LIBRARY)
        _lib.xi_walk_host.restype = ctypes.c_int
        _lib.xi_walk_rcp_guess.restype = ctypes.c_double
        _lib.xi_walk_rcp_guess.argtypes = [ctypes.c_double]
    return _lib


def rcp_guess(x):
    """The walk's reciprocal guess of x > 0 (the same code the kernel runs)."""
    return lib().xi_walk_rcp_guess(float(x))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def xi_walk(params, store, pair_a, pair_b, pair_ang, pair_row, num_rows,
            schedule=SCHEDULE_RUNS):
    """Un-normalised (we, xi, rp, rt, z, nb) rows [num_rows, nb] and a stats dict."""
    nb = params.num_bins_r_par * params.num_bins_r_trans
    arr = lambda x, t: np.ascontiguousarray(x, dtype=t)
    off = arr(store.offsets, np.int64)
    cols = [arr(getattr(store, k), np.float64)
            for k in ("z", "r_comov", "dist_m", "weights", "delta")]
    a = arr(pair_a, np.int32)
    b = arr(pair_b, np.int32)
    ang = arr(pair_ang, np.float64)
    row = arr(pair_row, np.int32)
    out = [np.zeros((num_rows, nb)) for _ in range(5)]
    nbs = np.zeros((num_rows, nb), dtype=np.int64)
    stats = np.zeros(len(STATS), dtype=np.int64)
    status = lib().xi_walk_host(ctypes.byref(params), ctypes.c_int64(len(off) - 1), _p(off),
                                *[_p(c) for c in cols], ctypes.c_int64(len(a)), _p(a), _p(b),
                                _p(ang), _p(row), ctypes.c_int32(num_rows),
                                ctypes.c_int32(schedule), *[_p(o) for o in out], _p(nbs),
                                _p(stats))
    if status != 0:
        raise ValueError("xi_walk_host: parameters outside the specialised case")
    return out + [nbs], dict(zip(STATS, (int(s) for s in stats)))
