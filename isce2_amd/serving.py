"""Offset-tracking request server with continuous batching on one GPU.

Requests (offset-field jobs: a reference/secondary image pair plus a window
grid) arrive on a queue. One scheduler thread owns the device:

* admission: the oldest waiting job is admitted when its image rows fit in the
  device image arena and its footprint fits the HBM budget; its images go to
  the arena on a copy stream (H2D overlaps the compute of admitted jobs);
* batching: every iteration takes the oldest loaded job's window class
  ``(window, half search, deramp)`` and fills one launch with pending windows
  of every loaded job of that class, oldest first — windows of different
  requests share a launch because all images live in one arena with a common
  row pitch (a window's start pixel is its job's arena row plus its row);
* completion: a job whose last window has run copies its outputs to the host,
  frees its arena rows (so waiting jobs can be admitted) and resolves its
  future with the outputs and its latency (arrival to host-visible result).

Each window class keeps a resident engine context (FFT plans, buffers). The
per-window results do not depend on which other windows share a launch, so a
job's outputs are the same alone and in a mix.
"""
from __future__ import annotations

import queue
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np
import torch

from .engine import MODE_AUTO, MODE_FUSED, Engine, WindowParams


@dataclass
class OffsetJob:
    """One offset-field request (PyCuAmpcor's window-grid attributes)."""
    ref: object                     # complex64 [H, Wd] host array / CPU tensor (rows may be wider: ``width``)
    sec: object
    window: int = 64
    half_search: int = 20
    skip_down: int = 64
    skip_across: int = 64
    start_down: int | None = None   # reference start pixel of window (0, 0); default half_search
    start_across: int | None = None
    n_down: int | None = None       # window grid; default: as many as fit
    n_across: int | None = None
    gross_down: int = 0
    gross_across: int = 0
    deramp: int = 1
    height: int | None = None       # valid image size when the arrays carry padding columns
    width: int | None = None
    job_id: int = -1
    arrival: float = field(default_factory=time.perf_counter)


def job_windows(job: OffsetJob):
    """Start pixels of a job's windows (reference and secondary search chip)."""
    H = job.height or int(job.ref.shape[0])
    Wd = job.width or int(job.ref.shape[1])
    W, hs = job.window, job.half_search
    s0d = hs if job.start_down is None else job.start_down
    s0a = hs if job.start_across is None else job.start_across
    nd = job.n_down if job.n_down is not None else (H - s0d - W - hs - max(job.gross_down, 0)) // job.skip_down + 1
    na = job.n_across if job.n_across is not None else (Wd - s0a - W - hs - max(job.gross_across, 0)) // job.skip_across + 1
    if nd <= 0 or na <= 0:
        raise ValueError(f"job {job.job_id}: no window fits a {H}x{Wd} image")
    r, c = np.meshgrid(np.arange(nd), np.arange(na), indexing="ij")
    rd = (s0d + r * job.skip_down).ravel()
    ra = (s0a + c * job.skip_across).ravel()
    sd = rd + job.gross_down - hs
    sa = ra + job.gross_across - hs
    S = W + 2 * hs
    if rd.min() < 0 or ra.min() < 0 or sd.min() < 0 or sa.min() < 0 or rd.max() + W > H or ra.max() + W > Wd \
            or sd.max() + S > H or sa.max() + S > Wd:
        raise IndexError(f"job {job.job_id}: a window reads outside its {H}x{Wd} image")
    return nd, na, rd.astype(np.int64), ra.astype(np.int64), sd.astype(np.int64), sa.astype(np.int64)


class RowAllocator:
    """First-fit allocator of row ranges ``[a, a + n)`` in ``[0, rows)`` (coalescing frees)."""

    def __init__(self, rows):
        self.rows = rows
        self.free = [(0, rows)]

    def alloc(self, n):
        for i, (a, b) in enumerate(self.free):
            if b - a >= n:
                if a + n == b:
                    self.free.pop(i)
                else:
                    self.free[i] = (a + n, b)
                return a
        return None

    def release(self, a, n):
        self.free.append((a, a + n))
        self.free.sort()
        merged = []
        for x, y in self.free:
            if merged and merged[-1][1] > x:
                raise ValueError("double free of arena rows")
            if merged and merged[-1][1] == x:
                merged[-1] = (merged[-1][0], y)
            else:
                merged.append((x, y))
        self.free = merged

    def free_rows(self):
        return sum(b - a for a, b in self.free)


class _RowArena(RowAllocator):
    """A [rows, pitch] pair of device images whose rows are handed out to jobs."""

    def __init__(self, rows, pitch, device):
        super().__init__(rows)
        self.pitch = pitch
        self.ref = torch.empty((rows, pitch), dtype=torch.complex64, device=device)
        self.sec = torch.empty((rows, pitch), dtype=torch.complex64, device=device)


class _Active:
    def __init__(self, job, fut, row0, starts, nd, na, device):
        self.job, self.fut, self.row0 = job, fut, row0
        self.nd, self.na = nd, na
        self.n = nd * na
        rd, ra, sd, sa = starts
        dev = torch.device("cuda", device)
        self.starts = torch.as_tensor(np.stack([rd + row0, ra, sd + row0, sa]).astype(np.int32), device=dev)
        self.out = dict(offset=torch.empty((self.n, 2), dtype=torch.float32, device=dev),
                        snr=torch.empty(self.n, dtype=torch.float32, device=dev),
                        cov=torch.empty((self.n, 3), dtype=torch.float32, device=dev))
        self.cursor = 0
        self.loaded = None          # CUDA event: images are in the arena
        self.key = (job.window, job.half_search, job.deramp)
        self.admitted_at = time.perf_counter()


class OffsetServer:
    """Continuous-batching offset-tracking server (see module docstring)."""

    def __init__(self, device=0, hbm_budget=200e9, arena_rows=None, pitch=4096, max_batch=16384, precision=1,
                 tuner=None):
        self.device = device
        self.dev = torch.device("cuda", device)
        self.pitch = pitch
        self.max_batch = max_batch
        self.precision = precision
        self.budget = float(hbm_budget)
        if arena_rows is None:
            arena_rows = int(self.budget // (2 * pitch * 8))
        self.arena = _RowArena(arena_rows, pitch, self.dev)
        self.used = 0.0
        self.copy_stream = torch.cuda.Stream(device=self.dev)
        self.compute_stream = torch.cuda.Stream(device=self.dev)
        self.engines = {}
        self.inbox = queue.Queue()
        self.waiting = []
        self.active = []
        self.stats = dict(batches=0, windows=0, jobs=0)
        self._stop = threading.Event()
        self._dead = None           # the exception that ended the scheduler thread
        self._lock = threading.Lock()
        self._thread = None
        self._next_id = 0
        self.tuner = tuner
        self.schedules = {}

    # -- client side ------------------------------------------------------
    def submit(self, job: OffsetJob) -> Future:
        fut = Future()
        with self._lock:
            if self._dead is not None:
                raise RuntimeError(f"offset server stopped after an error: {self._dead!r}")
            job.arrival = time.perf_counter()
            if job.job_id < 0:
                job.job_id = self._next_id
                self._next_id += 1
            self.inbox.put((job, fut))
        return fut

    def start(self):
        self._thread = threading.Thread(target=self._loop, name="offset-server", daemon=True)
        self._thread.start()
        return self

    def stop(self):
        self._stop.set()
        if self._thread is not None:
            self._thread.join()
        for e in self.engines.values():
            e.close()
        self.engines.clear()

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        self.stop()

    # -- scheduler --------------------------------------------------------
    def _footprint(self, job, n):
        rows = int(job.height or job.ref.shape[0])
        return 2.0 * rows * self.pitch * 8 + n * (24 + 16) + 64e6

    def _engine(self, key):
        if key not in self.engines:
            W, hs, deramp = key
            e = Engine(self.device)
            try:
                e.setup(WindowParams(window_h=W, window_w=W, hs_down=hs, hs_across=hs, deramp=deramp))
                e.set_precision(self.precision)
            except Exception:
                e.close()
                raise
            self.engines[key] = e
        return self.engines[key]

    def _presize(self, eng, mode, chunk):
        """Size a new class engine's device buffers for its largest launch now
        (one throw-away pass over windows at the arena origin), so that a later
        large job does not pay for growing them (a free synchronizes the device)."""
        if self.arena.ref.shape[0] < eng.sizes["search_h"] or self.arena.pitch < eng.sizes["search_w"]:
            return
        fused = mode == MODE_FUSED or (mode == MODE_AUTO and eng.sizes["fused"])
        m = min(self.max_batch, chunk if chunk > 0 else (self.max_batch if fused else 1024))
        z = torch.zeros(m, dtype=torch.int32, device=self.dev)
        prec = eng.precision
        eng.set_precision(0)
        try:
            eng.run_windows(self.arena.ref, self.arena.sec, z, z, z, z, mode=mode, chunk=chunk)
        finally:
            eng.set_precision(prec)

    def _admit(self):
        while True:
            try:
                self.waiting.append(self.inbox.get_nowait())
            except queue.Empty:
                break
        while self.waiting:
            job, fut = self.waiting[0]
            try:
                nd, na, *starts = self._validate(job)
            except Exception as exc:   # bad request: fail it, keep serving
                self.waiting.pop(0)
                fut.set_exception(exc)
                continue
            rows = int(job.height or job.ref.shape[0])
            need = self._footprint(job, nd * na)
            if self.used + need > self.budget or need > self.budget:
                if need > self.budget:      # can never fit: fail it rather than block the queue
                    self.waiting.pop(0)
                    fut.set_exception(MemoryError(f"job {job.job_id}: footprint {need / 1e9:.2f} GB exceeds "
                                                  f"the HBM budget {self.budget / 1e9:.2f} GB"))
                    continue
                return
            if rows > self.arena.rows:
                self.waiting.pop(0)
                fut.set_exception(MemoryError(f"job {job.job_id}: {rows} rows exceed the image arena's "
                                              f"{self.arena.rows}"))
                continue
            row0 = self.arena.alloc(rows)
            if row0 is None:
                return
            self.waiting.pop(0)
            self.used += need
            try:
                a = _Active(job, fut, row0, starts, nd, na, self.device)
                a.footprint = need
                with torch.cuda.stream(self.copy_stream):
                    for src, dst in ((job.ref, self.arena.ref), (job.sec, self.arena.sec)):
                        t = self._as_tensor(src)
                        if t.shape[1] == self.pitch:
                            dst[row0:row0 + rows].copy_(t[:rows], non_blocking=True)
                        else:
                            dst[row0:row0 + rows, :t.shape[1]].copy_(t[:rows], non_blocking=True)
                    ev = torch.cuda.Event()
                    ev.record(self.copy_stream)
            except Exception as exc:   # the upload failed: give its rows and budget back
                self.copy_stream.synchronize()
                self.arena.release(row0, rows)
                self.used -= need
                fut.set_exception(exc)
                continue
            a.loaded = ev
            self.active.append(a)

    @staticmethod
    def _as_tensor(src):
        return src if isinstance(src, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(src))

    def _validate(self, job):
        """Everything that can be wrong with a request, checked before it takes
        arena rows or budget: window placement, array shapes and dtypes, and the
        window class (its engine is built here, so a class the engine cannot run
        fails this request only)."""
        if int(job.window) <= 0 or int(job.half_search) <= 0:
            raise ValueError(f"job {job.job_id}: window {job.window} and half search {job.half_search} "
                             "must be positive")
        if job.skip_down <= 0 or job.skip_across <= 0:
            raise ValueError(f"job {job.job_id}: skips must be positive")
        nd, na, *starts = job_windows(job)
        rows = int(job.height or job.ref.shape[0])
        width = int(job.width or job.ref.shape[1])
        for name, arr in (("ref", job.ref), ("sec", job.sec)):
            shape = tuple(arr.shape)
            dtype = arr.dtype
            ok_dtype = dtype == torch.complex64 if isinstance(arr, torch.Tensor) else dtype == np.complex64
            if len(shape) != 2 or not ok_dtype:
                raise ValueError(f"job {job.job_id}: {name} must be a 2-D complex64 array, got {dtype} {shape}")
            if shape[0] < rows or shape[1] < width:
                raise ValueError(f"job {job.job_id}: {name} {shape} is smaller than the declared {rows}x{width}")
            if shape[1] > self.pitch:
                raise ValueError(f"job {job.job_id}: {name} rows of {shape[1]} exceed the arena pitch {self.pitch}")
        if job.window + 2 * job.half_search > self.pitch:
            raise ValueError(f"job {job.job_id}: search chip wider than the arena pitch {self.pitch}")
        self._engine((job.window, job.half_search, job.deramp))
        return (nd, na, *starts)

    def _pick_batch(self):
        ready = [a for a in self.active if a.cursor < a.n and a.loaded.query()]
        if not ready:
            return None, []
        key = ready[0].key
        parts, total = [], 0
        for a in ready:
            if a.key != key:
                continue
            take = min(a.n - a.cursor, self.max_batch - total)
            if take <= 0:
                break
            parts.append((a, a.cursor, take))
            total += take
        return key, parts

    def _run_batch(self, key, parts):
        eng = self._engine(key)
        for a, _, _ in parts:
            self.compute_stream.wait_event(a.loaded)
        with torch.cuda.stream(self.compute_stream):
            st = torch.cat([a.starts[:, c0:c0 + k] for a, c0, k in parts], dim=1)
            rd, ra, sd, sa = (st[i].contiguous() for i in range(4))
            if key not in self.schedules:
                if self.tuner is not None:    # first batch of a class: tune on its windows
                    sch = self.tuner.tune(eng, eng.params, (self.arena.ref, self.arena.sec, rd, ra, sd, sa), len(rd))
                    sch.apply(eng)
                    self.schedules[key] = (sch.mode, sch.chunk)
                else:
                    self.schedules[key] = (MODE_AUTO, 0)
                self._presize(eng, *self.schedules[key])
            mode, chunk = self.schedules[key]
            out = eng.run_windows(self.arena.ref, self.arena.sec, rd, ra, sd, sa, mode=mode, chunk=chunk)
            pos = 0
            for a, c0, k in parts:
                for name in ("offset", "snr", "cov"):
                    a.out[name][c0:c0 + k].copy_(out[name][pos:pos + k])
                pos += k
                a.cursor = c0 + k
        self.stats["batches"] += 1
        self.stats["windows"] += pos

    def _finish(self):
        done = [a for a in self.active if a.cursor >= a.n]
        if not done:
            return
        self.compute_stream.synchronize()
        for a in done:
            res = {k: v.cpu().numpy() for k, v in a.out.items()}
            t = time.perf_counter()
            res["offset"] = res["offset"].reshape(a.nd, a.na, 2)
            res["snr"] = res["snr"].reshape(a.nd, a.na)
            res["cov"] = res["cov"].reshape(a.nd, a.na, 3)
            res.update(job_id=a.job.job_id, latency_s=t - a.job.arrival, queue_s=a.admitted_at - a.job.arrival,
                       windows=a.n)
            self.active.remove(a)
            self.arena.release(a.row0, int(a.job.height or a.job.ref.shape[0]))
            self.used -= a.footprint
            self.stats["jobs"] += 1
            a.fut.set_result(res)

    def _drop(self, a, exc):
        """Fail one admitted job and give back its arena rows and budget."""
        if a in self.active:
            self.active.remove(a)
            self.arena.release(a.row0, int(a.job.height or a.job.ref.shape[0]))
            self.used -= a.footprint
        if not a.fut.done():
            a.fut.set_exception(exc)

    def _fail_all(self, exc):
        """The scheduler cannot go on: resolve every future it holds or will
        receive, so that no client waits forever, and refuse new requests."""
        with self._lock:
            self._dead = exc
        for a in list(self.active):
            if not a.fut.done():
                a.fut.set_exception(exc)
        self.active.clear()
        pending = list(self.waiting)
        self.waiting.clear()
        while True:
            try:
                pending.append(self.inbox.get_nowait())
            except queue.Empty:
                break
        for _, fut in pending:
            if not fut.done():
                fut.set_exception(exc)

    def _loop(self):
        try:
            torch.cuda.set_device(self.dev)
            while not self._stop.is_set() or self.active or self.waiting or not self.inbox.empty():
                self._admit()
                key, parts = self._pick_batch()
                if parts:
                    try:
                        self._run_batch(key, parts)
                    except Exception as exc:    # this launch failed: fail its jobs, keep serving the rest
                        self.compute_stream.synchronize()
                        for a, _, _ in parts:
                            self._drop(a, exc)
                        continue
                    self._finish()
                elif self.active or self.waiting:
                    time.sleep(0.0002)
                else:
                    try:
                        item = self.inbox.get(timeout=0.01)
                        self.waiting.append(item)
                    except queue.Empty:
                        pass
        except BaseException as exc:    # fail every job rather than hang its client
            self._fail_all(exc)
            raise
