"""Device-resident offset-tracking engine over the C ABI.

:class:`Engine` owns one ``ampcor_ctx`` on one HIP device. Images and window
start pixels are torch tensors in device memory (torch is the allocator and
stream provider here; all arithmetic runs in the HIP kernels of
``libisce2_amd.so``).
"""
from __future__ import annotations

import ctypes as C
import os
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._lib import FLAG_NOSYNC, MODE_AUTO, MODE_FUSED, MODE_STAGED, Params, Sizes, check

TAU_DEFAULT = 2e-6   # the engine's default precision-1 margin thresholds

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
        self.tau = [TAU_DEFAULT, TAU_DEFAULT]   # precision-1 margin thresholds (coarse, fine)
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
                    debug=False, refine=False, stream=None, out=None, nosync=False):
        """Run the per-window pipeline; every tensor argument lives on this engine's device.

        ``ref``/``sec`` are complex64 (or float32) 2-D tensors; start arrays are
        int32 1-D tensors of length n. Returns a dict of device tensors
        ``offset [n,2]``, ``snr [n]``, ``cov [n,3]`` (+ debug stage outputs).
        ``refine`` adds ``offset_fp64 [n,2]`` (float64): the offset refined by
        an FP64 parabola through the oversampled surface's peak (an extension;
        ``offset`` stays the reference's grid value).
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
        if refine:
            res["offset_fp64"] = o["offset_fp64"] if "offset_fp64" in o else \
                torch.empty((n, 2), dtype=torch.float64, device=dev)
        if debug or refine:
            if debug:
                for k in ("maxloc", "start", "shift", "zoom_idx"):
                    res[k] = o[k] if k in o else torch.empty((n, 2), dtype=torch.int32, device=dev)
                for k in ("maxval", "margin_coarse", "margin_fine"):
                    res[k] = o[k] if k in o else torch.full((n,), float("nan"), dtype=torch.float32, device=dev)
            g = res.get
            code = self.lib.ampcor_run_windows_debug(
                *common, _ptr(g("maxloc")), _ptr(g("start")), _ptr(g("shift")),
                _ptr(g("zoom_idx")), _ptr(g("maxval")), _ptr(g("margin_coarse")),
                _ptr(g("margin_fine")), _ptr(g("offset_fp64")), int(chunk), int(mode), st)
        else:
            code = self.lib.ampcor_run_windows(*common, int(chunk), int(mode), st)
        check(code, self._h)
        return res

    def run_windows_from_host(self, ref, sec, ref_d, ref_a, sec_d, sec_a, *, mode=MODE_AUTO, chunk=0,
                              bands=8, threads=None, refine=False):
        """Run on host-resident images, overlapping their upload with the computation.

        ``ref``/``sec`` are 2-D NumPy arrays (memory maps included) or CPU
        tensors; pinned contiguous tensors of the image dtype are used in
        place, anything else is staged into pinned buffers by row band on
        ``threads`` host threads. The windows (host integer start arrays) are
        split into ``bands`` consecutive groups; the rows each group reads are
        copied to the device on a copy stream, and each group runs as soon as
        its rows have arrived, so later uploads overlap earlier groups'
        kernels. Windows in row order give the most overlap. The engine's
        bounds rule is checked on the host first (IndexError), so the groups
        run without host synchronization; under precision 1 the near-tie
        windows of all groups are then re-run once through the FP64 pipeline,
        with the rule ``run_windows`` applies: every re-run window takes the
        FP64 offset, and SNR and covariance come from the FP64 run only where
        the raw-surface margin was small (so both methods return the same
        outputs). Returns host arrays ``offset [n,2]``,
        ``snr [n]``, ``cov [n,3]`` (and ``offset_fp64 [n,2]`` with ``refine``).
        """
        if self.params is None:
            raise RuntimeError("setup() must be called first")
        dev = torch.device("cuda", self.device)
        ref_d, ref_a, sec_d, sec_a = (np.ascontiguousarray(a, dtype=np.int64).ravel()
                                      for a in (ref_d, ref_a, sec_d, sec_a))
        n = len(ref_d)
        if not (len(ref_a) == len(sec_d) == len(sec_a) == n):
            raise ValueError("start arrays must have equal length")
        dt = torch.complex64 if self.params.image_complex else torch.float32

        def pinned(img):
            if isinstance(img, torch.Tensor):
                if img.device.type != "cpu":
                    raise ValueError("images must be host-resident")
                if img.is_pinned() and img.dtype == dt and img.dim() == 2 and img.is_contiguous():
                    return img, None
                img = img.numpy()
            if img.ndim != 2:
                raise ValueError("images must be 2-D")
            return torch.empty(img.shape, dtype=dt, pin_memory=True), img

        p_ref, s_ref = pinned(ref)
        p_sec, s_sec = pinned(sec)
        comp = torch.cuda.current_stream(dev)
        d_ref = torch.empty(p_ref.shape, dtype=dt, device=dev)
        d_sec = torch.empty(p_sec.shape, dtype=dt, device=dev)
        out = dict(offset=torch.empty((n, 2), dtype=torch.float32, device=dev),
                   snr=torch.empty(n, dtype=torch.float32, device=dev),
                   cov=torch.empty((n, 3), dtype=torch.float32, device=dev))
        if refine:
            out["offset_fp64"] = torch.empty((n, 2), dtype=torch.float64, device=dev)
        if n == 0:
            return {k: v.cpu().numpy() for k, v in out.items()}
        # the engine's bounds rule, checked here so that the bands run without host syncs
        wh, ww, sh, sw = self.params.window_h, self.params.window_w, self.sizes["search_h"], self.sizes["search_w"]
        outside = ((ref_d < 0) | (ref_a < 0) | (ref_d + wh > p_ref.shape[0]) | (ref_a + ww > p_ref.shape[1])
                   | (sec_d < 0) | (sec_a < 0) | (sec_d + sh > p_sec.shape[0]) | (sec_a + sw > p_sec.shape[1]))
        if outside.any():
            raise IndexError(f"window {int(np.argmax(outside))} reads outside its image")
        st = to_device_starts((ref_d, ref_a, sec_d, sec_a), dev)
        nb = max(1, min(int(bands), n))
        edges = np.linspace(0, n, nb + 1).astype(np.int64)
        h_ref, h_sec = p_ref.shape[0], p_sec.shape[0]
        need_r = np.maximum.accumulate(ref_d + wh)
        need_s = np.maximum.accumulate(sec_d + sh)
        copy = torch.cuda.Stream(device=dev)
        copy.wait_stream(comp)              # the device images' memory may be in use on comp until now
        pool, nthr = None, threads or min(16, os.cpu_count() or 1)
        if (s_ref is not None or s_sec is not None) and nthr > 1:
            pool = ThreadPoolExecutor(nthr)
        events, up_r, up_s = [], 0, 0
        for b in range(nb):
            last = edges[b + 1] - 1
            tr = h_ref if b == nb - 1 else max(up_r, int(need_r[last]))
            ts = h_sec if b == nb - 1 else max(up_s, int(need_s[last]))
            for src, pin, dst, r0, r1 in ((s_ref, p_ref, d_ref, up_r, tr), (s_sec, p_sec, d_sec, up_s, ts)):
                if r1 <= r0:
                    continue
                if src is not None:
                    fill_rows(pin.numpy(), src, r0, r1, pool, nthr)
                with torch.cuda.stream(copy):
                    dst[r0:r1].copy_(pin[r0:r1], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(copy)
            events.append(ev)
            up_r, up_s = tr, ts
        if pool is not None:
            pool.shutdown()
        d_ref.record_stream(copy)
        d_sec.record_stream(copy)
        # the bands run without host synchronization (the starts are checked above);
        # under precision 1 the near-tie windows of all bands are re-run in FP64 once
        rerun = self.precision == 1
        if rerun:
            for k in ("margin_coarse", "margin_fine"):
                out[k] = torch.empty(n, dtype=torch.float32, device=dev)
        for b in range(nb):
            i0, i1 = int(edges[b]), int(edges[b + 1])
            comp.wait_event(events[b])
            if i1 > i0:
                self.run_windows(d_ref, d_sec, *(s[i0:i1] for s in st), mode=mode, chunk=chunk, debug=rerun,
                                 refine=refine, nosync=True, out={k: v[i0:i1] for k, v in out.items()})
        if rerun:
            mc, mf = out.pop("margin_coarse").cpu().numpy(), out.pop("margin_fine").cpu().numpy()
            coarse = ~(mc >= self.tau[0])
            sel = np.nonzero(coarse | ~(mf >= self.tau[1]))[0]
            if len(sel):
                t = torch.as_tensor(sel, dtype=torch.int64).to(dev)
                self.set_precision(2)
                try:
                    o = self.run_windows(d_ref, d_sec, *(s.index_select(0, t).contiguous() for s in st),
                                         mode=MODE_STAGED, refine=refine)
                finally:
                    self.set_precision(1)
                # the same rule as the engine's own re-run: a window whose raw-surface
                # argmax is safe keeps its FP32 SNR and covariance (its raw peak cannot
                # change) and takes only the FP64 offset; the others take every output
                out["offset"].index_copy_(0, t, o["offset"])
                if refine:
                    out["offset_fp64"].index_copy_(0, t, o["offset_fp64"])
                keep = torch.as_tensor(coarse[sel], device=dev)
                for k in ("snr", "cov"):
                    cur = out[k].index_select(0, t)
                    mask = keep.view(-1, *([1] * (cur.dim() - 1)))
                    out[k].index_copy_(0, t, torch.where(mask, o[k], cur))
        return {k: v.cpu().numpy() for k, v in out.items()}

    def set_precision(self, mode: int, tau_coarse: float = -1.0, tau_fine: float = -1.0):
        """0: FP32 only; 1: FP32 + FP64 re-run of near-tie windows (default); 2: FP64 for all."""
        check(self.lib.ampcor_set_precision(self._h, int(mode), float(tau_coarse), float(tau_fine)), self._h)
        self.precision = int(mode)
        for i, t in enumerate((tau_coarse, tau_fine)):
            if t >= 0:
                self.tau[i] = float(t)

    def select_near_ties(self, margin_coarse, margin_fine, cap: int, tau_coarse: float, tau_fine: float,
                         out=None, stream=None):
        """int32 [cap] device tensor: the first ``cap`` windows whose margins fall below
        the thresholds (NaN counts as below), in order, then the first that do not
        (``ampcor_select_near_ties``; capturable, no host synchronization)."""
        n = int(margin_coarse.numel())
        dev = torch.device("cuda", self.device)
        for t in (margin_coarse, margin_fine):
            if t.dtype != torch.float32 or t.device != dev or not t.is_contiguous() or t.numel() != n:
                raise ValueError("margins must be contiguous float32 device tensors of equal length")
        idx = out if out is not None else torch.empty(cap, dtype=torch.int32, device=dev)
        st = C.c_void_p(stream.cuda_stream if stream is not None else torch.cuda.current_stream(dev).cuda_stream)
        check(self.lib.ampcor_select_near_ties(self._h, _ptr(margin_coarse), _ptr(margin_fine), n, float(tau_coarse),
                                               float(tau_fine), int(cap), _ptr(idx), st), self._h)
        return idx

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


def fill_rows(dst, src, r0, r1, pool=None, nthr=1):
    """dst[r0:r1] = src[r0:r1] for 2-D arrays; with a thread pool, large copies
    run as row bands on its threads (NumPy copies release the GIL, and reading
    a memory-mapped file is bound by page faults and memcpy rather than by one
    core). ``nthr`` is the pool's thread count."""
    rows = r1 - r0
    if pool is None or nthr <= 1 or rows < 2 * nthr or rows * dst[0].nbytes < (16 << 20):
        dst[r0:r1] = src[r0:r1]
        return
    step = -(-rows // nthr)

    def band(a):
        b = min(a + step, r1)
        dst[a:b] = src[a:b]

    list(pool.map(band, range(r0, r1, step)))


def to_device_starts(arrays, device):
    """int32 contiguous device tensors from host integer arrays."""
    return [torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32)).to(device) for a in arrays]
