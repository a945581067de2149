"""Batched chip pairs of uneven window size and search range.

Each pair is a reference chip ``[W, W]`` and a search chip ``[W + 2hs]^2``
(the search chip's origin is ``hs`` pixels up-left of the reference chip, as
for one window of an offset field). Pairs are grouped into size classes
``(W, hs)``; a class's chips are stacked into two device arenas (reference
``[m*W, W]``, search ``[m*S, S]``) so the engine runs them as ordinary windows
with start pixels ``(i*W, 0)`` and ``(i*S, 0)``. Each class has its own engine
context (resident plans and buffers) and a schedule chosen by the autotuner
(fused kernels or staged pipeline, chunk, FFT radix orders).

One step = every class's launches in no-sync mode, each class on a side
stream of its own (``concurrent=True``; forked from and joined to the step's
stream) so that small classes run beside large ones; with ``graph=True`` the
step is captured once into a HIP graph (``torch.cuda.CUDAGraph``) and
replayed. Under precision 1 the windows whose argmax margins fall below the
thresholds are re-run through the FP64 pipeline inside the same step: right
after a class's launches, its near ties are picked on the device (a stable
sort puts them first, in window order) and up to ``cap`` of them (ceil(m/64),
at least 8) run through the class's FP64 engine context on its own stream,
whose results are scattered into the class's outputs. The FP64 chains thus
overlap the other classes' FP32 launches, and nothing waits on the host. The
capacity's spare slots re-run the first windows that are not near ties, whose
integer results FP64 leaves as they are. After the step one small copy
brings the near-tie counts back; a class with more than ``cap`` near ties has
the rest re-run from the host.
"""
from __future__ import annotations

import numpy as np
import torch

from .autotune import Autotuner, Schedule
from .engine import MODE_FUSED, MODE_STAGED, Engine, WindowParams

TAU_DEFAULT = 2e-6


class _Class:
    def __init__(self, key, ref, sec, idx, device, deramp):
        self.W, self.hs = key
        self.S = self.W + 2 * self.hs
        self.idx = np.asarray(idx)
        m = ref.shape[0]
        self.m = m
        dev = torch.device("cuda", device)
        self.ref = ref.reshape(m * self.W, self.W).contiguous()
        self.sec = sec.reshape(m * self.S, self.S).contiguous()
        i = torch.arange(m, dtype=torch.int32, device=dev)
        z = torch.zeros(m, dtype=torch.int32, device=dev)
        self.starts = [(i * self.W).contiguous(), z, (i * self.S).contiguous(), z.clone()]
        self.params = WindowParams(window_h=self.W, window_w=self.W, hs_down=self.hs, hs_across=self.hs,
                                   deramp=deramp)
        self.eng = Engine(device)
        self.sizes = self.eng.setup(self.params)
        self.out = dict(offset=torch.empty((m, 2), dtype=torch.float32, device=dev),
                        snr=torch.empty(m, dtype=torch.float32, device=dev),
                        cov=torch.empty((m, 3), dtype=torch.float32, device=dev))
        for k in ("maxloc", "start", "shift", "zoom_idx"):
            self.out[k] = torch.empty((m, 2), dtype=torch.int32, device=dev)
        for k in ("maxval", "margin_coarse", "margin_fine"):
            self.out[k] = torch.full((m,), float("nan"), dtype=torch.float32, device=dev)
        self.sched = Schedule(mode=MODE_FUSED if self.sizes["fused"] else MODE_STAGED,
                              chunk=2048 if self.sizes["fused"] else 256)
        self.rerun_eng = None      # FP64 re-run context and stream (precision 1)
        self.rerun_stream = None
        self.overflow_eng = None   # FP64 context of the host-driven re-runs (never captured)
        self.cap = min(m, max(8, -(-m // 64)))            # FP64 re-runs inside the step

    def args(self):
        return (self.ref, self.sec, *self.starts)

    def launch(self, nosync):
        self.eng.run_windows(*self.args(), mode=self.sched.mode, chunk=self.sched.chunk, debug=True,
                             out=self.out, nosync=nosync)


class ChipPairBatch:
    """Uneven chip pairs run as one graph-captured step (see module docstring).

    ``groups`` maps ``(W, hs)`` to ``{"ref": complex64[m, W, W], "sec":
    complex64[m, S, S], "idx": int[m]}`` device tensors (``idx`` = position of
    each pair in the caller's order); :func:`ChipPairBatch.from_pairs` builds it
    from a list of ``{"W", "hs", "ref", "sec"}`` host chips.
    """

    def __init__(self, groups, device=0, deramp=1, precision=1, tau=TAU_DEFAULT, graph=True,
                 tuner: Autotuner | None = None, concurrent: bool = True):
        self.device = device
        self.dev = torch.device("cuda", device)
        self.precision = precision
        self.tau = tau
        self.use_graph = graph
        self.classes = [_Class(k, g["ref"], g["sec"], g["idx"], device, deramp) for k, g in sorted(groups.items())]
        self.n = sum(c.m for c in self.classes)
        for c in self.classes:
            c.eng.set_precision(2 if precision == 2 else 0)
            if tuner is not None:
                c.sched = tuner.tune(c.eng, c.params, c.args(), c.m)
                c.sched.apply(c.eng)
                c.eng.setup(c.params)
            c.eng.set_precision(2 if precision == 2 else 0)
        self.concurrent = concurrent
        for c in self.classes:
            c.main_stream = torch.cuda.Stream(device=self.dev) if concurrent else None
        if precision == 1:
            for c in self.classes:
                c.rerun_eng = Engine(device)
                c.rerun_eng.setup(c.params)
                c.rerun_eng.set_precision(2)
                c.rerun_stream = torch.cuda.Stream(device=self.dev)
        self.graph = None
        self.stream = torch.cuda.Stream(device=self.dev)
        self.rerun_count = 0

    @classmethod
    def from_pairs(cls, pairs, device=0, **kw):
        dev = torch.device("cuda", device)
        groups = {}
        for i, p in enumerate(pairs):
            groups.setdefault((int(p["W"]), int(p["hs"])), []).append(i)
        g2 = {}
        for k, idx in groups.items():
            ref = torch.from_numpy(np.stack([np.asarray(pairs[i]["ref"], np.complex64) for i in idx]))
            sec = torch.from_numpy(np.stack([np.asarray(pairs[i]["sec"], np.complex64) for i in idx]))
            g2[k] = dict(ref=ref.to(dev), sec=sec.to(dev), idx=np.asarray(idx))
        return cls(g2, device=device, **kw)

    _RERUN_KEYS = ("offset", "snr", "cov", "maxloc", "start", "shift", "zoom_idx", "maxval")

    def _rerun_in_step(self, c):
        """FP64 re-run of the class's first ``cap`` near ties (then non-ties), on the device."""
        idx = c.rerun_eng.select_near_ties(c.out["margin_coarse"], c.out["margin_fine"], c.cap, self.tau,
                                           self.tau).long()
        st = [x.index_select(0, idx).contiguous() for x in c.starts]
        o = c.rerun_eng.run_windows(c.ref, c.sec, *st, mode=MODE_STAGED, debug=True, nosync=True)
        for k in self._RERUN_KEYS:
            c.out[k].index_copy_(0, idx, o[k])

    def _step(self, nosync):
        main = torch.cuda.current_stream(self.dev)
        # largest windows first: their FP64 re-run chains, the longest, then overlap the
        # smaller classes' launches instead of trailing the step
        order = sorted(self.classes, key=lambda k: (-k.W, -k.hs))
        if self.concurrent:
            # every class on a stream of its own (its re-run after it on another), so
            # the small classes' launches fill the GPU beside the large ones
            for c in order:
                c.main_stream.wait_stream(main)
                with torch.cuda.stream(c.main_stream):
                    c.launch(nosync)
                if self.precision == 1:
                    c.rerun_stream.wait_stream(c.main_stream)
                    with torch.cuda.stream(c.rerun_stream):
                        self._rerun_in_step(c)
            for c in order:
                main.wait_stream(c.rerun_stream if self.precision == 1 else c.main_stream)
            return
        for c in order:
            c.launch(nosync)
            if self.precision == 1:
                s = c.rerun_stream
                s.wait_stream(main)
                with torch.cuda.stream(s):
                    self._rerun_in_step(c)
        if self.precision == 1:
            for c in self.classes:
                main.wait_stream(c.rerun_stream)

    def run(self):
        """One pass over every pair; results stay on the device until :meth:`results`."""
        torch.cuda.current_stream(self.dev).synchronize()
        with torch.cuda.stream(self.stream):
            if self.use_graph:
                if self.graph is None:
                    self._step(True)                     # allocations, plans, tables
                    self.stream.synchronize()
                    g = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(g, stream=self.stream):
                        self._step(True)
                    self.graph = g
                self.graph.replay()
            else:
                self._step(True)
            if self.precision == 1:
                self._rerun_overflow()
        self.stream.synchronize()

    def _rerun_overflow(self):
        """Host-driven FP64 re-run of the near ties beyond each class's in-step capacity."""
        both = torch.cat([c.out["margin_coarse"] for c in self.classes] +
                         [c.out["margin_fine"] for c in self.classes]).cpu().numpy()   # one copy back
        mc, mf = both[:self.n], both[self.n:]
        near = ~(mc >= self.tau) | ~(mf >= self.tau)
        self.rerun_count = 0
        main = torch.cuda.current_stream(self.dev)
        used = []
        o0 = 0
        for c in self.classes:
            pos = np.nonzero(near[o0:o0 + c.m])[0]
            o0 += c.m
            self.rerun_count += len(pos)
            if len(pos) <= c.cap:
                continue
            sel = torch.as_tensor(pos[c.cap:], dtype=torch.int64).to(self.dev)
            # a context of its own: the in-step context's buffers are held by the graph
            # at the capacity's size and must not grow
            if c.overflow_eng is None:
                c.overflow_eng = Engine(self.device)
                c.overflow_eng.setup(c.params)
                c.overflow_eng.set_precision(2)
            s = c.rerun_stream
            s.wait_stream(main)
            with torch.cuda.stream(s):
                st = [x.index_select(0, sel).contiguous() for x in c.starts]
                o = c.overflow_eng.run_windows(c.ref, c.sec, *st, mode=MODE_STAGED, debug=True)
                for k in self._RERUN_KEYS:
                    c.out[k].index_copy_(0, sel, o[k])
            used.append(s)
        for s in used:
            main.wait_stream(s)

    def results(self):
        """Host arrays in the caller's pair order: offset [n,2], snr [n], cov [n,3] (+ zoom_idx, maxloc)."""
        res = dict(offset=np.empty((self.n, 2), np.float32), snr=np.empty(self.n, np.float32),
                   cov=np.empty((self.n, 3), np.float32), maxloc=np.empty((self.n, 2), np.int32),
                   zoom_idx=np.empty((self.n, 2), np.int32))
        for c in self.classes:
            for k in res:
                res[k][c.idx] = c.out[k].cpu().numpy()
        return res

    def close(self):
        self.graph = None
        for c in self.classes:
            c.eng.close()
            for e in (c.rerun_eng, c.overflow_eng):
                if e is not None:
                    e.close()
