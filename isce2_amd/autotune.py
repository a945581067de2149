"""Per-size-class launch autotuner for the engine.

For one window configuration (window, half search) the tuner times candidate
schedules on a sample of real windows and keeps the fastest:

* execution path: the fused per-window kernels (where they cover the size) or
  the staged pipeline;
* chunk: windows per launch;
* FFT radix order of every staged transform length (coordinate descent over
  the orderings of each length's factorization into supported radices).

Timing uses events on the launch stream, after one untimed call per candidate
(which also builds plans and buffers). Results are cached per process and can
be saved to / loaded from a JSON file.
"""
from __future__ import annotations

import json
import os
from dataclasses import asdict, dataclass, field

import torch

from .engine import MODE_FUSED, MODE_STAGED, Engine, WindowParams

@dataclass
class Schedule:
    mode: int = MODE_STAGED
    chunk: int = 0
    radices: dict = field(default_factory=dict)     # FFT length -> list of radices (staged only)
    options: dict = field(default_factory=dict)     # launch-shape options (fused only)
    ms_per_window: float = float("nan")

    def apply(self, eng: Engine):
        for n, r in self.radices.items():
            eng.set_fft_radices(int(n), r)
        for k, v in self.options.items():
            eng.set_option(k, int(v))


def factorizations(n: int, limit: int = 4):
    """Up to ``limit`` radix orders for an n-point transform (supported radices only)."""
    rest, pow2 = n, 0
    while rest % 2 == 0:
        rest //= 2
        pow2 += 1
    odd = []
    for p in (3, 5, 7, 11, 13, 17):
        while rest % p == 0:
            odd.append(p)
            rest //= p
    if rest != 1:
        return []
    pow2_splits = []
    for first in (16, 8, 4):
        parts, e = [], pow2
        k = {16: 4, 8: 3, 4: 2}[first]
        while e >= k:
            parts.append(first)
            e -= k
        parts += {3: [8], 2: [4], 1: [2], 0: []}[e]
        pow2_splits.append(parts)
    orders = []
    for parts in pow2_splits:
        for seq in (parts + odd, odd + parts, sorted(parts + odd, reverse=True)):
            if seq and seq not in orders and len(seq) <= 14:
                orders.append(seq)
    return orders[:limit]


def staged_lengths(sizes: dict, params: WindowParams):
    """Transform lengths the staged pipeline uses for a configuration."""
    lens = {sizes["fft_corr_raw_h"], sizes["fft_corr_raw_w"], params.window_h, params.window_w,
            sizes["os_win_h"], sizes["os_win_w"], sizes["zoom_in_h"], sizes["zoom_in_w"],
            sizes["os_search_h"], sizes["os_search_w"], sizes["fft_corr_os_h"], sizes["fft_corr_os_w"]}
    if params.surf_method or sizes["zoom"] != 16 or sizes["surf_out"] != 256:
        lens |= {sizes["zoom"], sizes["surf_out"]}
    return sorted(lens)


class Autotuner:
    def __init__(self, device=0, reps=3, sample=2048, radix_sample=256, cache_file=None):
        self.device = device
        self.reps = reps
        self.sample = sample
        self.radix_sample = radix_sample
        self.cache = {}
        self.cache_file = cache_file
        if cache_file and os.path.exists(cache_file):
            with open(cache_file) as fh:
                for k, v in json.load(fh).items():
                    v["radices"] = {int(n): r for n, r in v["radices"].items()}
                    self.cache[k] = Schedule(**v)

    def save(self):
        if self.cache_file:
            with open(self.cache_file, "w") as fh:
                json.dump({k: asdict(v) for k, v in self.cache.items()}, fh, indent=1)

    @staticmethod
    def key(params: WindowParams) -> str:
        """Every field of the configuration: a schedule tuned for one surface
        method, zoom or oversampling factor is not reused for another."""
        p = params
        return (f"W{p.window_h}x{p.window_w}_hs{p.hs_down}x{p.hs_across}_d{p.deramp}_ro{p.raw_ovs}_z{p.zoom}"
                f"_so{p.surf_ovs}_sm{p.surf_method}_cs{p.corr_stat}_a{p.algorithm}_c{p.image_complex}")

    def _time(self, eng, args, mode, chunk, n):
        dev = torch.device("cuda", self.device)
        eng.run_windows(*args, mode=mode, chunk=chunk)        # untimed: plans, buffers
        s = torch.cuda.current_stream(dev)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        for _ in range(self.reps):
            eng.run_windows(*args, mode=mode, chunk=chunk)
        b.record(s)
        b.synchronize()
        return a.elapsed_time(b) / (self.reps * n)

    def tune(self, eng: Engine, params: WindowParams, args, n: int) -> Schedule:
        """Choose a schedule for ``params`` from ``args`` = (ref, sec, ref_d, ref_a, sec_d, sec_a)
        (device tensors holding at least a sample of real windows of this class)."""
        k = self.key(params)
        sizes = eng.setup(params)
        if k in self.cache:
            if self.cache[k].mode != MODE_FUSED or sizes["fused"]:
                return self.cache[k]
            del self.cache[k]           # a fused schedule for a configuration the fused kernels do not cover
        m = min(n, self.sample)
        sub = list(args[:2]) + [t[:m].contiguous() for t in args[2:]]
        prev = eng.precision
        eng.set_precision(0)
        best = None
        try:
            cands = []
            if sizes["fused"]:
                cands += [(MODE_FUSED, c) for c in (512, 1024, 2048, 4096)]
            cands += [(MODE_STAGED, c) for c in (64, 256, 1024)]
            for mode, chunk in cands:
                t = self._time(eng, sub, mode, min(chunk, max(m, 1)), m)
                if best is None or t < best.ms_per_window:
                    best = Schedule(mode=mode, chunk=chunk, ms_per_window=t)
            if best.mode == MODE_FUSED and params.window_h == 64:
                times = {}
                for fk, ft in ((2, 1024), (1, 512), (1, 1024)):
                    eng.set_option("fine_kernel", fk)
                    eng.set_option("fine_threads", ft)
                    times[(fk, ft)] = self._time(eng, sub, MODE_FUSED, min(best.chunk, m), m)
                fk, ft = min(times, key=times.get)
                eng.set_option("fine_kernel", fk)
                eng.set_option("fine_threads", ft)
                best.options.update(fine_kernel=fk, fine_threads=ft)
                best.ms_per_window = min(best.ms_per_window, times[(fk, ft)])
            if best.mode == MODE_STAGED:
                m = min(m, self.radix_sample)
                sub = list(args[:2]) + [t[:m].contiguous() for t in args[2:]]
                times = {}
                for fe in (1024, 2048, 4096, 8192):
                    eng.set_option("fft_block_elems", fe)
                    times[fe] = self._time(eng, sub, MODE_STAGED, min(best.chunk, m), m)
                fe = min(times, key=times.get)
                eng.set_option("fft_block_elems", fe)
                best.options["fft_block_elems"] = fe
                times = {}
                for cw in (0, 8, 16):
                    eng.set_option("fft_cols_width", cw)
                    times[cw] = self._time(eng, sub, MODE_STAGED, min(best.chunk, m), m)
                cw = min(times, key=times.get)
                eng.set_option("fft_cols_width", cw)
                best.options["fft_cols_width"] = cw
                for n_fft in staged_lengths(sizes, params):
                    orders = factorizations(n_fft)
                    if len(orders) < 2:
                        continue
                    times = []
                    for r in orders:
                        eng.set_fft_radices(n_fft, r)
                        times.append(self._time(eng, sub, MODE_STAGED, min(best.chunk, m), m))
                    i = min(range(len(orders)), key=times.__getitem__)
                    eng.set_fft_radices(n_fft, orders[i])
                    best.radices[n_fft] = orders[i]
                    best.ms_per_window = min(best.ms_per_window, times[i])
        finally:
            eng.set_precision(prev)
        self.cache[k] = best
        self.save()
        return best
