"""ctypes binding of libpfc_hip.so (declared in include/pfc_hip.h).

The library is built in-tree (`pfc_hip/_build.py`) and loaded from this
package directory only.  There is no fallback: if the library is missing or
fails to load, `load()` raises and nothing on the device path runs.
"""
from __future__ import annotations

import ctypes
import os

LIB_NAME = "libpfc_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

PFC_MODEL_STD = 0
PFC_MODEL_LOG = 1
PFC_SCHEME_ETD1 = 1
PFC_SCHEME_ETDRK2 = 2
PFC_STATUS_NONFINITE = 1

ERRORS = {-1: "invalid argument", -2: "unsupported size", -3: "HIP error", -4: "out of device memory"}


class PfcParams(ctypes.Structure):
    """Mirror of `pfc_params` (include/pfc_hip.h)."""
    _fields_ = [
        ("model", ctypes.c_int),
        ("Nx", ctypes.c_int),
        ("Ny", ctypes.c_int),
        ("Lx", ctypes.c_double),
        ("Ly", ctypes.c_double),
        ("Gamma", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("q0", ctypes.c_double),
        ("epsilon", ctypes.c_double),
        ("g", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("minPhi", ctypes.c_double),
        ("smallZTol", ctypes.c_double),
        ("M", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("scheme", ctypes.c_int),
        ("clamp_log", ctypes.c_int),
        ("pin_mean", ctypes.c_int),
    ]


# (name, restype, argtypes) of every exported entry point
_P = ctypes.c_void_p
_SIGNATURES = [
    ("pfc_last_error", ctypes.c_char_p, []),
    ("pfc_version", ctypes.c_int, []),
    ("pfc_create", ctypes.c_int, [ctypes.POINTER(PfcParams), ctypes.POINTER(_P)]),
    ("pfc_destroy", ctypes.c_int, [_P]),
    ("pfc_set_field", ctypes.c_int, [_P, _P, ctypes.c_int, _P]),
    ("pfc_get_field", ctypes.c_int, [_P, _P, ctypes.c_int, _P]),
    ("pfc_step", ctypes.c_int, [_P, ctypes.c_longlong, _P]),
    ("pfc_reduce", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double), _P]),
    ("pfc_set_shift", ctypes.c_int, [_P, ctypes.c_double]),
    ("pfc_get_status", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int), _P]),
    ("pfc_chemical_potential", ctypes.c_int, [_P, _P, _P, ctypes.c_int, _P]),
    ("pfc_local_field", ctypes.c_int, [_P, _P, ctypes.c_int, _P]),
    ("pfc_step_profile", ctypes.c_int, [_P, ctypes.c_longlong, _P, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_longlong)]),
    ("pfc_plan_schedule", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int)]),
    ("pfc_plan_stores", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int)]),
    ("pfc_plan_graph_stats", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_longlong)]),
    ("pfc_set_auto_shift", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double, ctypes.c_double, _P]),
    ("pfc_get_shift", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double), _P]),
    ("pfc_batch_set_auto_shift", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double, ctypes.c_double, _P]),
    ("pfc_batch_get_shift", ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_double), _P]),
    ("pfc_run_tracked", ctypes.c_int, [_P, ctypes.c_longlong, ctypes.c_longlong, ctypes.POINTER(ctypes.c_double), _P]),
    ("pfc_batch_run_tracked", ctypes.c_int, [_P, ctypes.c_longlong, ctypes.c_longlong, ctypes.POINTER(ctypes.c_double),
                                             _P]),
    ("pfc_group_create", ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int, ctypes.POINTER(_P)]),
    ("pfc_group_destroy", ctypes.c_int, [_P]),
    ("pfc_group_step", ctypes.c_int, [_P, ctypes.c_longlong, _P]),
    ("pfc_group_graph_stats", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_longlong)]),
    ("pfc_group_chunks", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int)]),
    ("pfc_group_set_flow", ctypes.c_int, [_P, ctypes.c_int]),
    ("pfc_group_flow_stats", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_longlong)]),
    ("pfc_batch_graph_stats", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_longlong)]),
    ("pfc_plan_bytes", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_longlong)]),
    ("pfc_batch_create", ctypes.c_int, [ctypes.POINTER(PfcParams), ctypes.c_int, ctypes.POINTER(_P)]),
    ("pfc_batch_destroy", ctypes.c_int, [_P]),
    ("pfc_batch_set_slot", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double, ctypes.c_double, _P]),
    ("pfc_batch_set_slot_box", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                              _P]),
    ("pfc_batch_get_fields", ctypes.c_int, [_P, _P, ctypes.c_int, _P]),
    ("pfc_batch_step_counts", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_longlong), _P]),
    ("pfc_batch_set_external_potential", ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_int, _P]),
    ("pfc_set_external_potential", ctypes.c_int, [_P, _P, ctypes.c_int, _P]),
    ("pfc_batch_diagnose", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                          ctypes.c_int, ctypes.POINTER(ctypes.c_double), _P]),
    ("pfc_diagnose", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_double), _P]),
    ("pfc_batch_set_field", ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_int, _P]),
    ("pfc_batch_get_field", ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_int, _P]),
    ("pfc_batch_set_active", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int), ctypes.c_int, _P]),
    ("pfc_batch_step", ctypes.c_int, [_P, ctypes.c_longlong, _P]),
    ("pfc_batch_reduce", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double), _P]),
    ("pfc_batch_get_status", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int), _P]),
    ("pfc_batch_bytes", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_longlong)]),
    ("pfc_slot_bytes", ctypes.c_int, [ctypes.POINTER(PfcParams), ctypes.POINTER(ctypes.c_longlong)]),
    ("pfc_rfft2", ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P]),
    ("pfc_irfft2", ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P]),
    ("pfc_irfft2_noalloc", ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P]),
]

EXPORTED = [s[0] for s in _SIGNATURES]

_lib = None


def load(path: str | None = None):
    """Load and bind the library (once).  Raises OSError/RuntimeError if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{LIB_NAME} not built at {p}; run __graft_entry__.build() (no CPU fallback exists)")
    lib = ctypes.CDLL(p)
    for name, res, args in _SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class PfcError(RuntimeError):
    """A non-zero status of a library call; `code` is the PFC_E* value."""
    INVALID, UNSUPPORTED, HIP, NOMEM = -1, -2, -3, -4

    def __init__(self, code, message):
        super().__init__(f"libpfc_hip: {ERRORS.get(code, code)}: {message}")
        self.code = code


def check(rc: int):
    if rc != 0:
        raise PfcError(rc, load().pfc_last_error().decode(errors="replace"))
