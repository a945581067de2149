"""ampcor_run_windows_noalloc against ampcor_run_windows, through the C ABI.

Two calls in a row on one context, the second with fewer windows than the
first (the first spans two chunks of the static scratch array), each compared
with ampcor_run_windows at precision 0 (the same FP32 kernels): offsets, SNR and
covariance identical. A window whose search chip leaves the image gets NaN
outputs instead of an error.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from isce2_amd import _lib
from isce2_amd.synth import speckle_pair_shift
from oracle import ampcor_oracle as O

pytestmark = pytest.mark.gpu


def _p(t):
    return C.c_void_p(t.data_ptr())


def _run(lib, ctx, fn, ref, sec, st, n):
    dev = ref.device
    off = torch.full((n, 2), -7.0, dtype=torch.float32, device=dev)
    snr = torch.full((n,), -7.0, dtype=torch.float32, device=dev)
    cov = torch.full((n, 3), -7.0, dtype=torch.float32, device=dev)
    stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    args = [ctx, _p(ref), ref.shape[0], ref.shape[1], ref.stride(0), _p(sec), sec.shape[0], sec.shape[1],
            sec.stride(0), *map(_p, st), n, _p(off), _p(snr), _p(cov)]
    if fn == "noalloc":
        code = lib.ampcor_run_windows_noalloc(*args, stream)
    else:
        code = lib.ampcor_run_windows(*args, 0, _lib.MODE_FUSED, stream)
    _lib.check(code, ctx)
    torch.cuda.synchronize(dev)
    return off.cpu().numpy(), snr.cpu().numpy(), cov.cpu().numpy()


def test_noalloc_equals_run_windows_twice():
    lib = _lib.load()
    dev = torch.device("cuda", 0)
    ref_np, sec_np = speckle_pair_shift(1024, 1024, shift=(2.3, -4.6), snr_db=10.0, seed=4)
    ref = torch.from_numpy(ref_np).to(dev)
    sec = torch.from_numpy(sec_np).to(dev)
    ctx = C.c_void_p()
    _lib.check(lib.ampcor_create(0, C.byref(ctx)), None)
    try:
        p = _lib.Params(window_h=64, window_w=64, hs_down=20, hs_across=20, raw_ovs=2, zoom=16, surf_ovs=16,
                        surf_method=0, corr_stat=21, deramp=1, algorithm=0, image_complex=1)
        s = _lib.Sizes()
        _lib.check(lib.ampcor_setup(ctx, C.byref(p), C.byref(s)), ctx)
        assert s.fused == 1
        _lib.check(lib.ampcor_set_precision(ctx, 0, -1.0, -1.0), ctx)
        big = O.start_pixels(57, 57, 20, 20, 16, 16, 0, 0, 20, 20)[:4]        # 3249 windows: two scratch chunks
        rng = np.random.default_rng(9)
        k = np.sort(rng.choice(3249, size=400, replace=False))
        small = [np.asarray(a)[k] for a in big]
        small[2] = small[2].copy()
        small[2][-1] = 1024 - 50                                           # search chip past the last row
        for name, starts in (("first", big), ("second", small)):
            st = [torch.as_tensor(np.asarray(a, np.int32)).to(dev) for a in starts]
            n = len(starts[0])
            a = _run(lib, ctx, "noalloc", ref, sec, st, n)
            good = np.ones(n, bool)
            if name == "second":
                good[-1] = False
                assert np.isnan(a[0][-1]).all() and np.isnan(a[1][-1]) and np.isnan(a[2][-1]).all()
                st = [t[:-1].contiguous() for t in st]
            b = _run(lib, ctx, "run", ref, sec, st, int(good.sum()))
            print(f"\n[noalloc {name}] n={n} max|doff|={np.abs(a[0][good] - b[0]).max()}", flush=True)
            assert np.array_equal(a[0][good], b[0]), name
            assert np.array_equal(a[1][good], b[1]), name
            assert np.array_equal(a[2][good], b[2]), name
        # configurations outside the fused kernels are refused
        p.surf_method = 1
        _lib.check(lib.ampcor_setup(ctx, C.byref(p), C.byref(s)), ctx)
        st = [torch.as_tensor(np.asarray(a, np.int32)).to(dev) for a in small]
        with pytest.raises(NotImplementedError):
            _run(lib, ctx, "noalloc", ref, sec, st, len(small[0]))
    finally:
        lib.ampcor_destroy(ctx)
