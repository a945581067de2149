"""ctypes binding of ``libisce2_amd.so`` (the C ABI in include/isce2_amd.h).

The library is loaded from the package directory only; there is no fallback:
if it is missing the import of the compute entry points raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libisce2_amd.so"

AMPCOR_OK = 0
ERRORS = {
    -1: ValueError,      # AMPCOR_E_ARG
    -2: IndexError,      # AMPCOR_E_RANGE
    -3: RuntimeError,    # AMPCOR_E_DEVICE
    -4: MemoryError,     # AMPCOR_E_NOMEM
    -5: NotImplementedError,  # AMPCOR_E_UNSUPPORTED
}
MODE_AUTO, MODE_STAGED, MODE_FUSED = 0, 1, 2
FLAG_NOSYNC = 0x100

EXPORTS = ("ampcor_create", "ampcor_destroy", "ampcor_setup", "ampcor_run_windows",
           "ampcor_run_windows_debug", "ampcor_set_precision", "ampcor_precision_stats",
           "ampcor_profile", "ampcor_profile_report", "ampcor_last_error", "ampcor_version",
           "ampcor_set_fft_radices", "ampcor_set_option", "ampcor_profile_only",
           "ampcor_start_pixels", "ampcor_check_range", "ampcor_run_windows_noalloc",
           "ampcor_select_near_ties")


class Params(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "window_h", "window_w", "hs_down", "hs_across", "raw_ovs", "zoom", "surf_ovs",
        "surf_method", "corr_stat", "deramp", "algorithm", "image_complex")]


class Sizes(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "zoom", "half_zoom", "search_h", "search_w", "zoom_in_h", "zoom_in_w", "os_win_h",
        "os_win_w", "os_search_h", "os_search_w", "raw_corr_h", "raw_corr_w", "os_corr_h",
        "os_corr_w", "surf_out", "corr_raw_zoom_h", "corr_raw_zoom_w", "fused",
        "fft_corr_raw_h", "fft_corr_raw_w", "fft_corr_os_h", "fft_corr_os_w")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


_lib = None


def load() -> C.CDLL:
    """Load the HIP library (raises ``OSError`` with a build hint if it is absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise OSError(f"{LIB_PATH} is missing: build it with `python -m isce2_amd.build` "
                      "(the engine has no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | getattr(os, "RTLD_GLOBAL", 0))
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    lib.ampcor_create.argtypes = [i32, C.POINTER(vp)]
    lib.ampcor_destroy.argtypes = [vp]
    lib.ampcor_setup.argtypes = [vp, C.POINTER(Params), C.POINTER(Sizes)]
    run_args = [vp, vp, i64, i64, i64, vp, i64, i64, i64, vp, vp, vp, vp, i32, vp, vp, vp]
    lib.ampcor_run_windows.argtypes = run_args + [i32, i32, vp]
    lib.ampcor_run_windows_debug.argtypes = run_args + [vp, vp, vp, vp, vp, vp, vp, vp, i32, i32, vp]
    lib.ampcor_run_windows_noalloc.argtypes = run_args + [vp]
    lib.ampcor_set_precision.argtypes = [vp, i32, C.c_float, C.c_float]
    lib.ampcor_precision_stats.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
    lib.ampcor_set_fft_radices.argtypes = [vp, i32, vp, i32]
    lib.ampcor_set_option.argtypes = [vp, C.c_char_p, i64]
    lib.ampcor_profile_only.argtypes = [vp, C.c_char_p]
    lib.ampcor_profile.argtypes = [vp, i32]
    lib.ampcor_profile_report.argtypes = [vp, C.c_char_p, i64]
    lib.ampcor_start_pixels.argtypes = [C.POINTER(Params), i32, i32, i32, i32, i32, i32, i32, i32,
                                        vp, vp, vp, vp, vp, vp]
    lib.ampcor_check_range.argtypes = [C.POINTER(Params), vp, vp, vp, vp, i32, i64, i64, i64, i64,
                                       C.POINTER(i32), C.POINTER(i32)]
    lib.ampcor_select_near_ties.argtypes = [vp, vp, vp, i32, C.c_float, C.c_float, i32, vp, vp]
    lib.ampcor_last_error.argtypes = [vp]
    lib.ampcor_last_error.restype = C.c_char_p
    lib.ampcor_version.restype = C.c_char_p
    for name in ("ampcor_create", "ampcor_destroy", "ampcor_setup", "ampcor_run_windows",
                 "ampcor_run_windows_debug", "ampcor_profile", "ampcor_profile_report",
                 "ampcor_set_precision", "ampcor_precision_stats", "ampcor_set_fft_radices",
                 "ampcor_set_option", "ampcor_profile_only", "ampcor_start_pixels", "ampcor_check_range",
                 "ampcor_run_windows_noalloc", "ampcor_select_near_ties"):
        getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def check(code: int, ctx=None):
    if code == AMPCOR_OK:
        return
    msg = load().ampcor_last_error(ctx).decode(errors="replace")
    raise ERRORS.get(code, RuntimeError)(f"ampcor error {code}: {msg}")
