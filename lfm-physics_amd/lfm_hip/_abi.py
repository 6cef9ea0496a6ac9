"""ctypes binding of liblfm_hip.so (the C ABI declared in include/lfm_hip.h).

Loading never falls back to anything: if the library is missing the import
error says how to build it.  torch is imported first so that the library
binds to the same HIP runtime (libamdhip64.so.7) torch already loaded.
"""
from __future__ import annotations

import ctypes
import os

from . import _build

c_float_p = ctypes.c_void_p  # device pointers travel as integers
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f32 = ctypes.c_float
_f64 = ctypes.c_double
_vp = ctypes.c_void_p

LFM_F_MASK = 0x01
LFM_F_DAMP_PREV = 0x02
LFM_F_FREEZE_CHI = 0x04
LFM_F_DIRECT = 0x08
LFM_F_CHI_DELTA = 0x10
LFM_F_SA_SCV = 0x20
LFM_F_KEEP_PSI = 0x40
LFM_F_SOURCE = 0x80
LFM_F_DETECTOR = 0x100
LFM_ABI_VERSION = 6
LFM_TILE_VARIANTS = 6
LFM_RELAX_SLOT_WORDS = 2048

# (name, restype, argtypes) for every entry point of include/lfm_hip.h
_STEP_REAL = [_vp] * 9 + [_vp, _vp]
_STEP_COMPLEX = [_vp] * 13 + [_vp, _vp]
_STEP_COLOR = [_vp] * 16 + [_vp, _vp]
SIGNATURES = [
    ("lfm_abi_version", ctypes.c_int, []),
    ("lfm_last_error", ctypes.c_char_p, []),
    ("lfm_step_real_f32", ctypes.c_int, _STEP_REAL),
    ("lfm_step_complex_f32", ctypes.c_int, _STEP_COMPLEX),
    ("lfm_step_slab_f32", ctypes.c_int, [_i32, _i32] + [_vp] * 12 + [_vp, _vp]),
    ("lfm_step2_real_f32", ctypes.c_int, [_vp] * 9 + [_vp, _vp]),
    ("lfm_step2_complex_f32", ctypes.c_int, [_vp] * 13 + [_vp, _vp]),
    ("lfm_step_color_f32", ctypes.c_int, _STEP_COLOR),
    ("lfm_step_color_fused_f32", ctypes.c_int, [_vp] * 11 + [_vp, _vp]),
    ("lfm_scv_prepare_f32", ctypes.c_int, [_vp, _vp, _vp, _f32, _f32, _vp]),
    ("lfm_scv_columns_f32", ctypes.c_int, [_vp, _f32, _f32, _vp]),
    ("lfm_scv_plan_create", ctypes.c_int, [_i32, _i32, ctypes.POINTER(_vp)]),
    ("lfm_scv_plan_create_ex", ctypes.c_int, [_i32, _i32, _i32, ctypes.POINTER(_vp)]),
    ("lfm_scv_plan_method", _i32, [_vp]),
    ("lfm_scv_plan_destroy", ctypes.c_int, [_vp]),
    ("lfm_scv_apply_psisq", ctypes.c_int, [_vp, _vp, _f32, _f32, _vp, _vp]),
    ("lfm_scv_apply", ctypes.c_int, [_vp, _vp, _vp, _f32, _f32, _vp, _vp]),
    ("lfm_scv_apply_scv_f32", ctypes.c_int, [_vp, _vp, _vp, _vp, _f32, _f32, _vp, _vp]),
    ("lfm_poisson_chi_f32", ctypes.c_int, [_vp, _vp, _vp, _i32, _f32, _f32, _f32, _f64, _vp, _vp]),
    ("lfm_clip_nonneg_f32", ctypes.c_int, [_vp, _vp, _i64, _vp]),
    ("lfm_batch_plan_create", ctypes.c_int, [_i32, _i32, _vp, _i32, _i32, ctypes.POINTER(_vp)]),
    ("lfm_batch_plan_destroy", ctypes.c_int, [_vp]),
    ("lfm_batch_total_sites", _i64, [_vp]),
    ("lfm_batch_workgroups", _i32, [_vp]),
    ("lfm_batch_step_f32", ctypes.c_int, [_vp] * 13 + [_i32, _vp]),
    ("lfm_batch_step2_f32", ctypes.c_int, [_vp] * 14 + [_i32, _vp]),
    ("lfm_adjoint_partials_len", _i64, [_i32]),
    ("lfm_calib_stream_f32", ctypes.c_int, [_vp, _i32, _vp, _i32, ctypes.c_int64, _i32, _i32, _vp]),
    ("lfm_relax_step_f32", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_float, _i32,
                                          ctypes.c_float, _vp]),
    ("lfm_relax_scale_f32", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_float, ctypes.c_int64,
                                           _vp]),
    ("lfm_adjoint_complex_seed_f32", ctypes.c_int, [_vp] * 8 + [_vp, _vp]),
    ("lfm_adjoint_complex_step_f32", ctypes.c_int, [_vp] * 13 + [_i32] + [_vp] * 4),
    ("lfm_reduce_partials_len", _i64, [_i32]),
    ("lfm_energy_f64", ctypes.c_int,
     [_vp, _vp, _vp, _vp, _vp, _f64, _vp, _i32, _i32, _i32, _f64, _f64, _vp, _vp, _vp]),
    ("lfm_energy_f64_noalloc", ctypes.c_int,
     [_vp, _vp, _vp, _vp, _vp, _f64, _vp, _i32, _i32, _i32, _f64, _f64, _vp, _vp]),
    ("lfm_chi_stats_f64", ctypes.c_int, [_vp, _f64, _i32, _f64, _f64, _f64, _vp, _vp, _vp]),
]


class LfmParams(ctypes.Structure):
    """Mirror of lfm_params_t."""

    _fields_ = [("N", _i32), ("flags", _i32),
                ("dt2", _f32), ("kappa", _f32), ("lambda_self", _f32), ("chi0", _f32),
                ("e0_sq", _f32), ("epsilon_w", _f32), ("kappa_c", _f32),
                ("epsilon_cc", _f32), ("kappa_string", _f32), ("kappa_tube", _f32),
                ("tile", _i32), ("chunk", _i32), ("keep", _vp),
                ("step_counter", _vp), ("step_off", _i32),
                ("src_axis", _i32), ("src_plane", _i32),
                ("det_axis", _i32), ("det_plane", _i32), ("det_every", _i32), ("det_field", _i32),
                ("reserved", _i32),
                ("src_env", _vp), ("src_amp", _vp), ("src_len", _i64),
                ("det_frames", _vp), ("det_first", _i64), ("det_len", _i64)]


class LfmError(RuntimeError):
    """A C-ABI call returned a negative status."""


_lib = None


def load(path: str | None = None):
    """Load (once) and return the ctypes handle of liblfm_hip.so."""
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  (binds libamdhip64 before our library)

    path = path or _build.library_path()
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: the HIP library has not been built. "
            "Run `python -c 'import __graft_entry__ as g; g.build()'` from the repository root.")
    lib = ctypes.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.lfm_abi_version() != LFM_ABI_VERSION:
        raise ImportError(f"liblfm_hip ABI version {lib.lfm_abi_version()} != {LFM_ABI_VERSION}: "
                          "rebuild the library")
    _lib = lib
    return lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = _lib.lfm_last_error().decode(errors="replace") if _lib else ""
        raise LfmError(f"{what} failed with status {rc}: {msg}")
