"""Device-resident offset-tracking engine over the C ABI.

:class:`Engine` owns one ``ampcor_ctx`` on one HIP device. Images and window
start pixels are torch tensors in device memory (torch is the allocator and
stream provider here; all arithmetic runs in the HIP kernels of
``libisce2_amd.so``).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._lib import FLAG_NOSYNC, MODE_AUTO, MODE_FUSED, MODE_STAGED, Params, Sizes, check

__all__ = ["Engine", "WindowParams", "MODE_AUTO", "MODE_STAGED", "MODE_FUSED"]


@dataclass
class WindowParams:
    """Per-window knobs (names of cuAmpcorParameter, contrib/PyCuAmpcor/src/cuAmpcorParameter.h:32-158)."""
    window_h: int = 64
    window_w: int = 64
    hs_down: int = 20
    hs_across: int = 20
    raw_ovs: int = 2
    zoom: int = 16
    surf_ovs: int = 16
    surf_method: int = 0
    corr_stat: int = 21
    deramp: int = 1
    algorithm: int = 0
    image_complex: int = 1

    def to_c(self) -> Params:
        return Params(**{k: int(getattr(self, k)) for k, _ in Params._fields_})


def _ptr(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


class Engine:
    """One engine context on ``device`` (an int or ``torch.device``)."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        if isinstance(device, torch.device):
            device = 0 if device.index is None else device.index
        self.device = int(device)
        h = C.c_void_p()
        check(self.lib.ampcor_create(self.device, C.byref(h)), None)
        self._h = h
        self.params: WindowParams | None = None
        self.precision = 1
        self.sizes: dict | None = None

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.ampcor_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setup(self, params: WindowParams) -> dict:
        s = Sizes()
        cp = params.to_c()
        check(self.lib.ampcor_setup(self._h, C.byref(cp), C.byref(s)), self._h)
        self.params = params
        self.sizes = s.as_dict()
        return self.sizes

    def run_windows(self, ref, sec, ref_d, ref_a, sec_d, sec_a, *, mode=MODE_AUTO, chunk=0,
                    debug=False, stream=None, out=None, nosync=False):
        """Run the per-window pipeline; every tensor argument lives on this engine's device.

        ``ref``/``sec`` are complex64 (or float32) 2-D tensors; start arrays are
        int32 1-D tensors of length n. Returns a dict of device tensors
        ``offset [n,2]``, ``snr [n]``, ``cov [n,3]`` (+ debug stage outputs).
        ``out`` may supply any of these tensors. ``nosync`` makes the call free
        of host synchronization (device-side bounds check with NaN outputs for
        bad windows, no FP64 re-run), as needed under graph capture.
        """
        if self.params is None:
            raise RuntimeError("setup() must be called first")
        dev = torch.device("cuda", self.device)
        n = int(ref_d.numel())
        for t in (ref_d, ref_a, sec_d, sec_a):
            if t.dtype != torch.int32 or t.device != dev or not t.is_contiguous() or t.numel() != n:
                raise ValueError("start arrays must be contiguous int32 device tensors of equal length")
        for img in (ref, sec):
            if img.device != dev or img.dim() != 2 or img.stride(1) != 1:
                raise ValueError("images must be 2-D row-contiguous device tensors")
            want = torch.complex64 if self.params.image_complex else torch.float32
            if img.dtype != want:
                raise ValueError(f"image dtype {img.dtype} does not match image_complex={self.params.image_complex}")
        o = out or {}
        offset = o["offset"] if "offset" in o else torch.empty((n, 2), dtype=torch.float32, device=dev)
        snr = o["snr"] if "snr" in o else torch.empty((n,), dtype=torch.float32, device=dev)
        cov = o["cov"] if "cov" in o else torch.empty((n, 3), dtype=torch.float32, device=dev)
        res = dict(offset=offset, snr=snr, cov=cov)
        st = C.c_void_p(stream.cuda_stream if stream is not None else torch.cuda.current_stream(dev).cuda_stream)
        common = [self._h, _ptr(ref), ref.shape[0], ref.shape[1], ref.stride(0),
                  _ptr(sec), sec.shape[0], sec.shape[1], sec.stride(0),
                  _ptr(ref_d), _ptr(ref_a), _ptr(sec_d), _ptr(sec_a), n,
                  _ptr(offset), _ptr(snr), _ptr(cov)]
        if nosync:
            mode = int(mode) | FLAG_NOSYNC
        if debug:
            for k in ("maxloc", "start", "shift", "zoom_idx"):
                res[k] = o[k] if k in o else torch.empty((n, 2), dtype=torch.int32, device=dev)
            for k in ("maxval", "margin_coarse", "margin_fine"):
                res[k] = o[k] if k in o else torch.full((n,), float("nan"), dtype=torch.float32, device=dev)
            code = self.lib.ampcor_run_windows_debug(
                *common, _ptr(res["maxloc"]), _ptr(res["start"]), _ptr(res["shift"]),
                _ptr(res["zoom_idx"]), _ptr(res["maxval"]), _ptr(res["margin_coarse"]),
                _ptr(res["margin_fine"]), int(chunk), int(mode), st)
        else:
            code = self.lib.ampcor_run_windows(*common, int(chunk), int(mode), st)
        check(code, self._h)
        return res

    def set_precision(self, mode: int, tau_coarse: float = -1.0, tau_fine: float = -1.0):
        """0: FP32 only; 1: FP32 + FP64 re-run of near-tie windows (default); 2: FP64 for all."""
        check(self.lib.ampcor_set_precision(self._h, int(mode), float(tau_coarse), float(tau_fine)), self._h)
        self.precision = int(mode)

    def set_option(self, name: str, value: int):
        """Launch-shape option (``fine_threads``: 512 or 1024)."""
        check(self.lib.ampcor_set_option(self._h, name.encode(), int(value)), self._h)
        self.options = dict(getattr(self, "options", {}), **{name: int(value)})

    def set_fft_radices(self, n: int, radices):
        """Radix schedule of the staged pipeline's length-n FFTs (autotuning)."""
        arr = (C.c_int32 * len(radices))(*[int(r) for r in radices])
        check(self.lib.ampcor_set_fft_radices(self._h, int(n), C.cast(arr, C.c_void_p), len(radices)), self._h)

    def precision_stats(self) -> tuple:
        """(windows checked, windows re-run in FP64) since the previous call."""
        a, b = C.c_int64(), C.c_int64()
        check(self.lib.ampcor_precision_stats(self._h, C.byref(a), C.byref(b)), self._h)
        return a.value, b.value

    def profile(self, enable: bool):
        """Reset and enable (or disable) per-kernel HIP event timing."""
        check(self.lib.ampcor_profile(self._h, int(bool(enable))), self._h)

    def profile_only(self, name: str | None):
        """Time only kernel ``name`` (None: every kernel)."""
        check(self.lib.ampcor_profile_only(self._h, (name or "").encode()), self._h)

    def profile_report(self) -> dict:
        """{kernel: {"launches", "ms", "windows"}} accumulated since profile(True)."""
        buf = C.create_string_buffer(1 << 16)
        check(self.lib.ampcor_profile_report(self._h, buf, len(buf)), self._h)
        out = {}
        for line in buf.value.decode().splitlines():
            name, launches, ms, windows = line.split()
            out[name] = dict(launches=int(launches), ms=float(ms), windows=int(windows))
        return out

    @staticmethod
    def version() -> str:
        return _lib.load().ampcor_version().decode()


def to_device_starts(arrays, device):
    """int32 contiguous device tensors from host integer arrays."""
    return [torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32)).to(device) for a in arrays]
