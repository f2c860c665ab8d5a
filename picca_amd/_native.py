"""ctypes binding of ``libpicca_hip.so`` (C-ABI declared in ``include/picca_hip.h``).

The library is the only compute path of this package: when it is missing or does not
load, every entry point raises ``RuntimeError`` (there is no CPU fallback).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "libpicca_hip.so")

PICCA_OK = 0
PICCA_E_VALUE = -1
PICCA_E_RUNTIME = -2
PICCA_E_INDEX = -3
PICCA_E_HIP = -4
PICCA_E_NOMEM = -5

PICCA_PAIR_SAME_HALF_PLATE = 1

# schedules of the pair kernel's specialised path (picca_xi_set_schedule)
XI_SCHEDULE_AUTO = 0
XI_SCHEDULE_PAIRS = 1
XI_SCHEDULE_RUNS = 2
PICCA_PAIR_XCF_COUNT = 2
PICCA_METAL_AUTO = 0
PICCA_METAL_CROSS = 1
PICCA_METAL_XCF = 2

PAIR_SAME_HALF_PLATE = 1


class CfParams(ctypes.Structure):
    """``picca_cf_params``: the cf module globals, None -> has_* = 0."""

    _fields_ = [
        ("num_bins_r_par", ctypes.c_int32),
        ("num_bins_r_trans", ctypes.c_int32),
        ("num_model_bins_r_par", ctypes.c_int32),
        ("num_model_bins_r_trans", ctypes.c_int32),
        ("r_par_min", ctypes.c_double),
        ("r_par_max", ctypes.c_double),
        ("r_trans_max", ctypes.c_double),
        ("z_min_pairs", ctypes.c_double),
        ("z_max_pairs", ctypes.c_double),
        ("zerr_cut_deg", ctypes.c_double),
        ("zerr_cut_kms", ctypes.c_double),
        ("z_ref", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("alpha2", ctypes.c_double),
        ("edge_tol", ctypes.c_double),
        ("has_z_min_pairs", ctypes.c_int32),
        ("has_z_max_pairs", ctypes.c_int32),
        ("has_zerr_cut", ctypes.c_int32),
        ("x_correlation", ctypes.c_int32),
        ("rmu_binning", ctypes.c_int32),
        ("ang_correlation", ctypes.c_int32),
        ("remove_same_half_plate_close_pairs", ctypes.c_int32),
        ("redshift_evolution", ctypes.c_int32),
    ]


class XiReport(ctypes.Structure):
    _fields_ = [
        ("accepted_pairs", ctypes.c_int64),
        ("edge_pairs", ctypes.c_int64),
        ("overflow_pairs", ctypes.c_int64),
        ("visited_pairs", ctypes.c_int64),
        ("wave_steps", ctypes.c_int64),
        ("tasks", ctypes.c_int64),
        ("empty_tasks", ctypes.c_int64),
        ("kernel_ms", ctypes.c_double),
        ("launches", ctypes.c_int64),
        ("runs", ctypes.c_int64),
        ("probes", ctypes.c_int64),
        ("walk_tasks", ctypes.c_int64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class DmatReport(ctypes.Structure):
    """``picca_dmat_report``."""
    _fields_ = [
        ("pass_ms", ctypes.c_double),
        ("fill_ms", ctypes.c_double),
        ("batches", ctypes.c_int64),
        ("mfma_flops", ctypes.c_int64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class WickReport(ctypes.Structure):
    """``picca_wick_report``."""
    _fields_ = [
        ("in_range_pairs", ctypes.c_int64),
        ("products", ctypes.c_int64),
        ("overflow_pairs", ctypes.c_int64),
        ("batches", ctypes.c_int64),
        ("kernel_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class MetalReport(ctypes.Structure):
    """``picca_metal_report``."""
    _fields_ = [
        ("data_pairs", ctypes.c_int64),
        ("model_pairs", ctypes.c_int64),
        ("overflow_pairs", ctypes.c_int64),
        ("visited_pairs", ctypes.c_int64),
        ("kernel_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_F64 = ctypes.c_double

# name -> argument types (every function returns int)
class ContParams(ctypes.Structure):
    """``picca_cont_params``."""
    _fields_ = [("max_iter", ctypes.c_int32), ("use_constant_weight", ctypes.c_int32),
                ("tol", ctypes.c_double), ("tol_abs", ctypes.c_double)]


_PROTOTYPES = {
    "picca_abi_version": [],
    "picca_device_count": [_P],
    "picca_device_info": [_I32, ctypes.c_char_p, _I32, _P, _P, _P],
    "picca_create": [_I32, _P],
    "picca_destroy": [_P],
    "picca_set_forests": [_P, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "picca_xi_set_pairs": [_P, _I64, _P, _P, _P, _P, _P, _I32],
    "picca_xi_execute": [_P, _P, _P],
    "picca_xi_download": [_P, _P, _P, _P, _P, _P, _P],
    "picca_xi_set_schedule": [_P, _I32],
    "picca_xi_use_graph": [_P, _I32],
    "picca_xi_execute_noalloc": [_P, _P, _P, _I64, _P],
    "picca_dmat_run": [_P, _P, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "picca_dmat_last_report": [_P, _P],
    "picca_diag_copy": [_P, _I64, _P],
    "picca_cov_subsample": [_P, _I64, _I64, _P, _P, _P, _P],
    "picca_wick_t123": [_P, _P, _I64, _P, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _P],
    "picca_metal_set_absorber": [_P, _I32, _P, _P, _P, _P, _P],
    "picca_metal_pass": [_P, _P, _I32, _I32, _I32, _F64, _I64, _P, _P, _P, _P, _P, _P, _P, _P,
                         _P, _P, _P],
    "picca_xcf_set_pairs": [_P, _I64, _P, _P, _P, _P, _P, _P, _P, _I32, _I32],
    "picca_xcf_execute": [_P, _P, _P],
    "picca_cont_set_forests": [_P, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "picca_cont_fit": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "picca_cont_set_rest_frame": [_P, _P, _P, _P, _I32],
    "picca_cont_stack": [_P, _P, _P, _P],
    "picca_cont_set_mean_cont": [_P, _I32, _P, _P, _P],
    "picca_mlp_init": [_P, _I32, _I32, _P, _P, _P, _P],
    "picca_mlp_get": [_P, _P, _P, _P, _P],
    "picca_mlp_forward": [_P, _I64, _P, _P, _P],
    "picca_mlp_gradients": [_P, _I64, _P, _P, _P, _P, _P, _P, _P],
    "picca_mlp_set_data": [_P, _I64, _P, _P],
    "picca_mlp_train": [_P, _I64, _I32, ctypes.c_float, ctypes.c_uint64, _P],
    "picca_neighbour_candidates": [_I64, _P, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _P,
                                   _P, _F64, _F64, _F64, _I32, _P, _P, _P, _P],
}

_lib = None
_load_error = None


# PICCA_HIP_ABI_VERSION of include/picca_hip.h that these structures mirror
ABI_VERSION = 4


def exported_symbols():
    """Names of the C-ABI functions this binding expects the library to export."""
    return sorted(_PROTOTYPES) + ["picca_last_error"]


def load():
    """Load the library once; raise RuntimeError when it is not built or not loadable."""
    global _lib, _load_error
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBRARY_PATH):
        raise RuntimeError(
            "libpicca_hip.so is not built (expected at %s); run `python -m picca_amd.build`"
            % LIBRARY_PATH)
    try:
        lib = ctypes.CDLL(LIBRARY_PATH)
    except OSError as err:
        _load_error = err
        raise RuntimeError("cannot load %s: %s" % (LIBRARY_PATH, err)) from err
    for name, argtypes in _PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int
    lib.picca_last_error.argtypes = []
    lib.picca_last_error.restype = ctypes.c_char_p
    if lib.picca_abi_version() != ABI_VERSION:
        raise RuntimeError("%s has ABI version %d, these bindings expect %d: rebuild it"
                           % (LIBRARY_PATH, lib.picca_abi_version(), ABI_VERSION))
    _lib = lib
    return lib


def check(status, what=""):
    """Map a C-ABI status code to the exception type picca raises for that condition."""
    if status == PICCA_OK:
        return
    lib = load()
    message = lib.picca_last_error().decode("utf-8", "replace")
    text = "%s: %s" % (what, message) if what else message
    if status == PICCA_E_VALUE:
        raise ValueError(text)
    if status == PICCA_E_INDEX:
        raise IndexError(text)
    if status == PICCA_E_NOMEM:
        raise MemoryError(text)
    raise RuntimeError(text)


def ptr(array):
    """Address of a C-contiguous numpy array (None -> NULL)."""
    if array is None:
        return None
    if not array.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return array.ctypes.data


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def as_i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def as_u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)
