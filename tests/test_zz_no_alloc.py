"""picca_xi_execute_noalloc against picca_xi_execute, through the C-ABI.

The variant writes the sums into a device buffer the caller owns (allocated here with the
HIP runtime, as such a caller would) and keeps its queue and counters in the library's
per-GPU scratch array.  Two calls in a row, the second over fewer pairs and rows, must
each match the allocating entry point: num_pairs bit-exact, the FP64 sums within 1e-12
relative (the order of the atomic adds differs between launches), the counters equal.
"""
import ctypes

import numpy as np
import pytest

import walk_cases as W
from picca_amd import _native, batch
from picca_amd.engine import get_engine, params_from_globals

pytestmark = pytest.mark.gpu

RECORD_BYTES = 48   # we, xi, rp, rt, z (FP64), nb (uint64)


class DeviceBuffer:
    """Device memory of the caller, from the HIP runtime."""

    def __init__(self, nbytes):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(nbytes, 1))) == 0
        self.nbytes = nbytes

    def to_host(self, nbytes):
        out = np.zeros(nbytes, dtype=np.uint8)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr,
                                  ctypes.c_size_t(nbytes), 2) == 0   # device to host
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def records(raw, rows, nb):
    rec = raw.view(np.float64).reshape(rows, nb, 6)
    counts = raw.view(np.uint64).reshape(rows, nb, 6)[:, :, 5].astype(np.int64)
    return [rec[:, :, k] for k in range(5)] + [counts]


def test_noalloc_matches_execute_twice_second_smaller():
    case = W.survey_case("survey_small", 0.05)
    params = params_from_globals(case.g)
    nb = case.g.num_bins_r_par * case.g.num_bins_r_trans
    eng = get_engine(0)
    lib = _native.load()
    full = case.pairs
    cut = full.pair_row == 0           # the second call: the pairs of row 0 only
    small = batch.XiPairs(full.pair_a[cut], full.pair_b[cut], full.pair_ang[cut],
                          full.pair_flags[cut], full.pair_row[cut], full.healpixs[:1])
    assert 0 < small.num_pairs < full.num_pairs
    buf = DeviceBuffer(full.num_rows * nb * RECORD_BYTES)
    try:
        for pairs in (full, small):
            ref = batch.run_xi(case.store, pairs, params)
            report = _native.XiReport()
            status = lib.picca_xi_execute_noalloc(eng._ctx, ctypes.byref(params), buf.ptr,
                                                  pairs.num_rows * nb, ctypes.byref(report))
            _native.check(status, "picca_xi_execute_noalloc")
            mine = records(buf.to_host(pairs.num_rows * nb * RECORD_BYTES), pairs.num_rows, nb)
            assert np.array_equal(mine[5], ref.num_pairs)
            for k, arr in enumerate((ref.weights, ref.xi, ref.r_par, ref.r_trans, ref.z)):
                scale = float(np.abs(arr).max())
                np.testing.assert_allclose(mine[k], arr, rtol=1e-12, atol=1e-13 * scale)
            rep = report.as_dict()
            for key in ("accepted_pairs", "edge_pairs", "overflow_pairs", "tasks", "runs",
                        "probes", "walk_tasks"):
                assert rep[key] == ref.report[key], key
            assert rep["accepted_pairs"] > 0
        # a buffer shorter than num_rows * np * nt records is rejected before any launch
        report = _native.XiReport()
        status = lib.picca_xi_execute_noalloc(eng._ctx, ctypes.byref(params), buf.ptr,
                                              small.num_rows * nb - 1, ctypes.byref(report))
        assert status == _native.PICCA_E_VALUE
    finally:
        buf.free()
