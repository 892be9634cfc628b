"""ctypes binding of the C ABI declared in include/nyxpm.h.

The library is the product path: if ``libnyxpm.so`` is missing, loading raises
-- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnyxpm.so")

NYXPM_OK = 0
NYXPM_ERR_ARG = 1
NYXPM_ERR_HIP = 2
NYXPM_ERR_NOMEM = 3
NYXPM_ERR_NOT_CONVERGED = 4
NYXPM_ERR_UNSUPPORTED = 5
NYXPM_ERR_STATE = 6
NYXPM_FP64 = 0
NYXPM_MIXED = 1

_ERRNAMES = {1: "ERR_ARG", 2: "ERR_HIP", 3: "ERR_NOMEM", 4: "ERR_NOT_CONVERGED",
             5: "ERR_UNSUPPORTED", 6: "ERR_STATE"}


class NyxPMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nyxpm {_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class Geom(ctypes.Structure):
    _fields_ = [("n_cell", ctypes.c_int * 3), ("prob_lo", ctypes.c_double * 3),
                ("prob_hi", ctypes.c_double * 3)]


class Opts(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("nu1", ctypes.c_int), ("nu2", ctypes.c_int),
                ("max_iter", ctypes.c_int), ("tile", ctypes.c_int * 3), ("verbose", ctypes.c_int),
                ("smoother", ctypes.c_int), ("kick", ctypes.c_int), ("sweep_tx", ctypes.c_int),
                ("sweep_ty", ctypes.c_int), ("sweep_zchunk", ctypes.c_int), ("kick_chunk", ctypes.c_int),
                ("separate_kicks", ctypes.c_int), ("coarse", ctypes.c_int), ("particle_cfl", ctypes.c_double),
                ("resort_slack", ctypes.c_double), ("phi_guess", ctypes.c_int), ("residual", ctypes.c_int),
                ("graph_cycles", ctypes.c_int), ("ir_first", ctypes.c_int),
                ("sweep2", ctypes.c_int), ("sweep2_tx", ctypes.c_int), ("sweep2_ty", ctypes.c_int)]


# (name, restype, argtypes) for every function in include/nyxpm.h
_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_i64 = ctypes.c_int64
_s = ctypes.c_char_p
SIGNATURES = [
    ("nyxpm_create", _i, [_vp, _vp, _vp]),
    ("nyxpm_create_batch", _i, [_vp, _i, _vp, _vp]),
    ("nyxpm_destroy", _i, [_vp]),
    ("nyxpm_set_stream", _i, [_vp, _vp]),
    ("nyxpm_sync", _i, [_vp]),
    ("nyxpm_last_error", _s, [_vp]),
    ("nyxpm_num_boxes", _i, [_vp]),
    ("nyxpm_set_precision", _i, [_vp, _i]),
    ("nyxpm_particles_upload", _i, [_vp, _i, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d]),
    ("nyxpm_particles_upload_device", _i, [_vp, _i, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d]),
    ("nyxpm_particles_load_binary", _i, [_vp, _i, _s, _vp]),
    ("nyxpm_particles_load_ascii", _i, [_vp, _i, _s, _vp]),
    ("nyxpm_particles_write_amrex", _i, [_vp, _i, _s, _s, _i]),
    ("nyxpm_particles_read_amrex", _i, [_vp, _i, _s, _s, _vp]),
    ("nyxpm_particles_download", _i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("nyxpm_num_particles", _i64, [_vp, _i]),
    ("nyxpm_assign_density", _i, [_vp, _i, _vp]),
    ("nyxpm_sum_particle_mass", _i, [_vp, _i, _vp]),
    ("nyxpm_assign_density_and_vels", _i, [_vp, _i, _vp]),
    ("nyxpm_solve_for_phi", _i, [_vp, _d, _d, _d, _d, _i, _vp, _vp]),
    ("nyxpm_solve_poisson", _i, [_vp, _i, _vp, _vp, _d, _d, _i, _vp, _vp]),
    ("nyxpm_get_grav_vector", _i, [_vp, _i, _vp]),
    ("nyxpm_gather_accel", _i, [_vp, _i, _vp]),
    ("nyxpm_gather_accel_noalloc", _i, [_vp, _i, _vp]),
    ("nyxpm_est_time_step", _i, [_vp, _d, _d, _vp]),
    ("nyxpm_get_phi", _i, [_vp, _i, _vp]),
    ("nyxpm_set_phi", _i, [_vp, _i, _vp]),
    ("nyxpm_set_density", _i, [_vp, _i, _vp]),
    ("nyxpm_get_rhs", _i, [_vp, _i, _vp]),
    ("nyxpm_move_kick_drift", _i, [_vp, _d, _d, _d]),
    ("nyxpm_move_kick", _i, [_vp, _d, _d, _d]),
    ("nyxpm_redistribute", _i, [_vp]),
    ("nyxpm_pm_init", _i, [_vp, _d, _d, _d, _d]),
    ("nyxpm_pm_steps", _i, [_vp, _i, _d, _d, _d, _d, _d, _d, _i, _i, _vp]),
    ("nyxpm_integrate_comoving_a", _i, [_d, _d, _d, _d, _d, _i, _vp]),
    ("nyxpm_pm_steps_comoving", _i, [_vp, _i, _d, _vp, _d, _d, _d, _i, _d, _d, _d, _i, _vp]),
    ("nyxpm_autotune", _i, [_vp, _i, _vp, _i]),
    ("nyxpm_timing_enable", _i, [_vp, _i]),
    ("nyxpm_timing_select", _i, [_vp, _s]),
    ("nyxpm_timing_get", _i, [_vp, _s, _vp, _vp, _vp]),
    ("nyxpm_timing_reset", _i, [_vp]),
    ("nyxpm_device_bytes", _i64, [_vp]),
    ("nyxpm_sort_stats", _i, [_vp, _vp]),
]

_LIB = None


def load():
    """Load libnyxpm.so (built by nyx_amd.build / __graft_entry__.build)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m nyx_amd.build` "
                          "(the HIP library is required; there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def ptr(a):
    """Data pointer of a C-contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.flags.c_contiguous
    return a.ctypes.data


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
