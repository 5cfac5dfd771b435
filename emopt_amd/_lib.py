"""ctypes binding of ``libemfdtd.so`` (the C ABI declared in include/emfdtd.h).

Replaces ``emopt/fdtd_ctypes.py``: the same loader pattern (a shared library
next to the package, ``ctypes`` argtypes per symbol), but every entry point
returns an int status that is turned into :class:`EmfError` here.

The library is the only compute path.  There is no CPU fallback: if the
library is missing or fails to load this module raises, and the GPU entry
points raise when no HIP device is visible.
"""
from __future__ import annotations

import ctypes
import os
import threading
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_void_p

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libemfdtd.so")


class EmfError(RuntimeError):
    """A libemfdtd call returned a non-zero status."""


class SolveInfo(ctypes.Structure):
    _fields_ = [
        ("n0", c_int64),
        ("Tn", c_int64),
        ("steps", c_int64),
        ("status", c_int32),
        ("nchecks", c_int32),
        ("a_change", c_double),
        ("phi_change", c_double),
        ("seconds", c_double),
    ]


# symbol -> (argtypes, restype); status-returning calls are checked
_P = c_void_p
_SIGS = {
    "emf_last_error": ([], c_char_p),
    "emf_version": ([], c_char_p),
    "emf_device_count": ([], c_int),
    "emf_new": ([c_int, POINTER(c_void_p)], c_int),
    "emf_delete": ([_P], c_int),
    "emf_set_wavelength": ([_P, c_double], c_int),
    "emf_set_physical_dims": ([_P] + [c_double] * 6, c_int),
    "emf_set_grid_dims": ([_P, c_int, c_int, c_int], c_int),
    "emf_set_local_grid": ([_P] + [c_int] * 6, c_int),
    "emf_set_dt": ([_P, c_double], c_int),
    "emf_set_complex_eps": ([_P, c_int], c_int),
    "emf_set_mat_arrays": ([_P] + [_P] * 6, c_int),
    "emf_update_eps_box": ([_P, c_int, _P] + [c_int] * 6, c_int),
    "emf_update_mu_box": ([_P, c_int, _P] + [c_int] * 6, c_int),
    "emf_get_mat_arrays": ([_P] + [_P] * 6, c_int),
    "emf_set_compression": ([_P, c_int], c_int),
    "emf_set_bc": ([_P, c_char_p], c_int),
    "emf_set_pml_widths": ([_P] + [c_int] * 6, c_int),
    "emf_set_pml_properties": ([_P] + [c_double] * 4, c_int),
    "emf_build_pml": ([_P], c_int),
    "emf_reset_pml": ([_P], c_int),
    "emf_add_source": ([_P] + [_P] * 6 + [c_int] * 7, c_int),
    "emf_clear_sources": ([_P], c_int),
    "emf_set_source_properties": ([_P, c_double, c_double], c_int),
    "emf_src_func_t": ([_P, c_int, c_double, c_double], c_double),
    "emf_update_H": ([_P, c_int, c_double], c_int),
    "emf_update_E": ([_P, c_int, c_double], c_int),
    "emf_zero_fields": ([_P], c_int),
    "emf_get_time_fields": ([_P] + [_P] * 6, c_int),
    "emf_set_time_fields": ([_P] + [_P] * 6, c_int),
    "emf_capture_t0_fields": ([_P, c_int], c_int),
    "emf_capture_t1_fields": ([_P, c_int], c_int),
    "emf_calc_complex_fields_2T": ([_P, c_int, c_double, c_double], c_int),
    "emf_calc_complex_fields_3T": ([_P, c_int, c_double, c_double, c_double], c_int),
    "emf_calc_phase_2T": ([c_double] * 4, c_double),
    "emf_calc_amplitude_2T": ([c_double] * 5, c_double),
    "emf_calc_phase_3T": ([c_double] * 6, c_double),
    "emf_calc_amplitude_3T": ([c_double] * 7, c_double),
    "emf_set_nconv": ([_P, c_int], c_int),
    "emf_solve": ([_P, c_int, c_double, c_double, c_int64, c_int64, POINTER(SolveInfo)], c_int),
    "emf_get_history": ([_P, _P, c_int], c_int),
    "emf_get_field": ([_P, c_int, c_int] + [c_int] * 6 + [_P], c_int),
    "emf_get_field_interp": ([_P, c_int, c_int] + [c_int] * 12 + [_P], c_int),
    "emf_interp_flux": ([_P, c_int, c_int] + [c_int] * 12 + [_P], c_int),
    "emf_set_plane_weights": ([_P, c_int, c_int64] + [_P] * 6, c_int),
    "emf_interp_dot": ([_P, c_int, c_int] + [c_int] * 12 + [_P], c_int),
    "emf_time_steps": ([_P, c_int64, c_int64, c_int64, c_int, _P], c_int),
    "emf_set_tuning": ([_P, c_int, c_int, c_int, c_int], c_int),
    "emf_set_streaming": ([_P, c_int], c_int),
    "emf_set_h_shape": ([_P, c_int, c_int], c_int),
    "emf_sample_structured": ([c_int, _P, c_int, _P, _P, c_int, c_double, c_double, c_double,
                               c_int, c_double, c_double, c_int, c_int, c_int, c_int, c_int, c_int,
                               c_double, c_double, c_double, _P], c_int),
    "emf_copy_bandwidth": ([c_int, c_int64, c_int, _P], c_int),
    "emf_mem_info": ([c_int, POINTER(c_int64), POINTER(c_int64)], c_int),
    "emf_calc_ydAx": ([_P] + [_P] * 6 + [c_int] * 6 + [_P], c_int),
    "emf_grad_assemble": ([_P, c_int64, _P, _P, _P, _P, c_int, c_double, _P], c_int),
    "emf_grad_assemble_noalloc": ([_P, c_int64, _P, _P, _P, _P, c_int, c_double, _P], c_int),
    "emf_pml_stretch": ([_P, c_int64, _P, _P, _P], c_int),
    "emf_batch_new": ([c_int, POINTER(_P)], c_int),
    "emf_batch_delete": ([_P], c_int),
    "emf_batch_set_tuning": ([_P, c_int, c_int, c_int], c_int),
    "emf_batch_set_grouping": ([_P, c_int64, c_int], c_int),
    "emf_batch_set_graph": ([_P, c_int], c_int),
    "emf_set_wide_offsets": ([_P, c_int], c_int),
    "emf_field_products": ([_P, c_int] + [c_int] * 6 + [_P], c_int),
    "emf_stage_clear": ([_P], c_int),
    "emf_stage_count": ([_P, _P], c_int),
    "emf_stage_interp_T": ([_P, c_int, _P, c_int, c_double] + [c_int] * 6 + [_P], c_int),
    "emf_stage_power_face": ([_P, c_int, _P, c_double, c_double, c_double] + [c_int] * 12 + [_P],
                             c_int),
    "emf_stage_register": ([_P], c_int),
    "emf_stage_get": ([_P, c_int, _P, _P], c_int),
    "emf_batch_set_lanes": ([_P, c_int], c_int),
    "emf_batch_add": ([_P, _P, c_int, c_double, c_double, c_int64, c_int64, POINTER(c_int)], c_int),
    "emf_batch_run": ([_P, c_int, c_int, POINTER(c_int)], c_int),
    "emf_batch_poll": ([_P, c_int, POINTER(SolveInfo)], c_int),
    "emf_batch_remove": ([_P, c_int], c_int),
    "emf_batch_count": ([_P, POINTER(c_int), POINTER(c_int)], c_int),
}

# calls whose int result is a count, not a status
_COUNT_RESULT = {"emf_device_count", "emf_get_history", "emf_batch_poll"}

_lock = threading.Lock()
_lib = None


class _Checked:
    """Wraps the raw CDLL so that a negative status raises EmfError."""

    def __init__(self, raw):
        self._raw = raw

    def __getattr__(self, name):
        fn = getattr(self._raw, name)
        argt, rest = _SIGS[name]
        if rest is not c_int:
            return fn
        raw = self._raw

        def call(*args):
            rc = fn(*args)
            if rc < 0 or (rc != 0 and name not in _COUNT_RESULT):
                msg = raw.emf_last_error()
                raise EmfError("%s failed (%d): %s" % (name, rc, msg.decode() if msg else ""))
            return rc

        call.__name__ = name
        setattr(self, name, call)
        return call


def load():
    """Load libemfdtd.so (once) and return the checked wrapper."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libemfdtd.so not found at %s; build it with `python -m emopt_amd.build` "
                "(there is no CPU fallback)" % LIB_PATH)
        raw = ctypes.CDLL(LIB_PATH)
        for name, (argt, rest) in _SIGS.items():
            fn = getattr(raw, name)
            fn.argtypes = argt
            fn.restype = rest
        _lib = _Checked(raw)
        return _lib


def exported_symbols():
    return list(_SIGS)


def ptr(a, dtype=None):
    """Raw data pointer of a C-contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    if dtype is not None and a.dtype != dtype:
        raise TypeError("expected %s array, got %s" % (np.dtype(dtype), a.dtype))
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(c_void_p)


def c128(a):
    """Contiguous complex128 copy/view suitable for the ABI."""
    return np.ascontiguousarray(a, dtype=np.complex128)
