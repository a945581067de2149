"""GPU precision layer on a warped-speckle field (window 64, half search 20).

* precision 2 (FP64 pipeline) equals the FP64 oracle: integer results
  bit-exact, offsets identical, SNR / covariance / peak to FP64 rounding;
* precision 0 (FP32 only) differs from FP64 only where the FP32 argmax margin
  (top-1 minus top-2 of the normalized surface) is at rounding level, for
  both the fused and the staged pipelines;
* precision 1 (default: FP32 plus an FP64 re-run of windows whose margins fall
  below the thresholds) equals the FP64 pipeline on every integer result and
  offset.
"""
import numpy as np
import pytest
import torch

from isce2_amd.engine import MODE_FUSED, MODE_STAGED, Engine, WindowParams
from isce2_amd.synth import speckle_pair_warp
from oracle import ampcor_oracle as O

from .gpu_helpers import compare, run_both

pytestmark = pytest.mark.gpu

# FP32 surfaces carry errors of a few 1e-7 on values of order 1; a flip with a
# margin above this bound would mean an accuracy defect, not a rounding tie
FP32_TIE_BOUND = 2e-6
INT_KEYS = ("maxloc", "start", "shift", "zoom_idx")


@pytest.fixture(scope="module")
def field():
    dev = torch.device("cuda", 0)
    ref, sec, _, _ = speckle_pair_warp(1024, 1024, max_amp=6.0, jitter=0.05, tile=256, seed=11, device=dev)
    W, hs, skip = 64, 20, 16
    n1 = (1024 - 2 * hs - W) // skip
    rows, cols = np.meshgrid(np.arange(n1), np.arange(n1), indexing="ij")
    rd = (hs + rows * skip).ravel().astype(np.int64)
    ra = (hs + cols * skip).ravel().astype(np.int64)
    return ref, sec, (rd, ra, rd - hs, ra - hs)


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


def _run(eng, field, precision, mode, deramp=1):
    ref, sec, starts = field
    dev = torch.device("cuda", 0)
    eng.setup(WindowParams(window_h=64, window_w=64, hs_down=20, hs_across=20, deramp=deramp))
    eng.set_precision(precision)
    st = [torch.as_tensor(s.astype(np.int32)).to(dev) for s in starts]
    out = eng.run_windows(torch.from_numpy(ref).to(dev), torch.from_numpy(sec).to(dev), *st, mode=mode, debug=True)
    torch.cuda.synchronize(dev)
    stats = eng.precision_stats()
    eng.set_precision(1)
    return {k: v.cpu().numpy() for k, v in out.items()}, stats


@pytest.mark.parametrize("deramp", [0, 1])
def test_fp64_pipeline_equals_oracle(eng, field, deramp):
    ref, sec, starts = field
    sel = np.random.default_rng(deramp).choice(len(starts[0]), 192, replace=False)
    cfg = O.AmpcorConfig(hs_down=20, hs_across=20, deramp=deramp)
    eng.set_precision(2)
    try:
        g, o = run_both(eng, ref, sec, cfg, [s[sel] for s in starts], MODE_STAGED)
    finally:
        eng.set_precision(1)
    c = compare(g, o)
    print(f"\n[fp64 deramp={deramp}] {c}", flush=True)
    assert c["maxloc_mismatch"] == 0 and c["zoom_mismatch"] == 0, c
    assert c["start_mismatch"] == 0 and c["shift_mismatch"] == 0, c
    assert c["offset_max_err"] == 0.0, c
    assert c["cov_sentinel_mismatch"] == 0, c
    # outputs are stored as float32: equal up to one float32 rounding
    assert c["snr_max_rel"] < 1e-6 and c["cov_max_rel"] < 1e-6, c
    assert c["maxval_max_abs"] < 1e-6, c


@pytest.mark.parametrize("mode", [MODE_FUSED, MODE_STAGED])
def test_fp32_flips_only_at_rounding_ties(eng, field, mode):
    a, _ = _run(eng, field, 0, mode)
    b, _ = _run(eng, field, 2, MODE_STAGED)
    flip_c = np.any(a["maxloc"] != b["maxloc"], axis=1)
    flip_f = np.any(a["zoom_idx"] != b["zoom_idx"], axis=1) & ~flip_c
    n = len(flip_c)
    print(f"\n[fp32 mode={mode}] n={n} coarse_flips={int(flip_c.sum())} fine_flips={int(flip_f.sum())} "
          f"max_flip_margin_coarse={a['margin_coarse'][flip_c].max() if flip_c.any() else None} "
          f"max_flip_margin_fine={a['margin_fine'][flip_f].max() if flip_f.any() else None}", flush=True)
    assert np.all(a["margin_coarse"][flip_c] < FP32_TIE_BOUND)
    assert np.all(a["margin_fine"][flip_f] < FP32_TIE_BOUND)
    assert flip_c.sum() + flip_f.sum() <= max(2, n // 200)
    same = ~flip_c
    rel = np.abs(a["snr"][same] - b["snr"][same]) / np.abs(b["snr"][same])
    assert rel.max() < 1e-4


@pytest.mark.parametrize("mode", [MODE_FUSED, MODE_STAGED])
def test_default_precision_equals_fp64(eng, field, mode):
    a, (checked, rerun) = _run(eng, field, 1, mode)
    b, _ = _run(eng, field, 2, MODE_STAGED)
    n = len(a["offset"])
    print(f"\n[precision 1 mode={mode}] n={n} checked={checked} rerun={rerun}", flush=True)
    assert checked == n and 0 <= rerun < n // 4
    for k in INT_KEYS:
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["offset"], b["offset"])


def test_mixed_fp64_rerun_kinds(eng, field):
    """Thresholds that send about half of the windows through the whole FP64
    pipeline and all the others through its fine stage only, in one staged pass
    whose chunks hold either kind or both: integer results and offsets equal the
    FP64 pipeline's, and the fine-only windows keep their FP32 SNR."""
    a0, _ = _run(eng, field, 0, MODE_FUSED)
    b, _ = _run(eng, field, 2, MODE_STAGED)
    tc = float(np.median(a0["margin_coarse"]))
    ref, sec, starts = field
    dev = torch.device("cuda", 0)
    eng.setup(WindowParams(window_h=64, window_w=64, hs_down=20, hs_across=20, deramp=1))
    eng.set_precision(1, tc, float("inf"))
    try:
        st = [torch.as_tensor(s.astype(np.int32)).to(dev) for s in starts]
        out = eng.run_windows(torch.from_numpy(ref).to(dev), torch.from_numpy(sec).to(dev), *st, mode=MODE_FUSED,
                              debug=True)
        torch.cuda.synchronize(dev)
        checked, rerun = eng.precision_stats()
    finally:
        eng.set_precision(1, 2e-6, 2e-6)
    a = {k: v.cpu().numpy() for k, v in out.items()}
    n = len(a["offset"])
    fine_only = a0["margin_coarse"] >= tc
    print(f"\n[mixed re-run] n={n} rerun={rerun} fine_only={int(fine_only.sum())}", flush=True)
    assert checked == n and rerun == n and 0 < fine_only.sum() < n
    for k in INT_KEYS:
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["offset"], b["offset"])
    assert np.array_equal(a["snr"][fine_only], a0["snr"][fine_only])
    full = ~fine_only
    np.testing.assert_allclose(a["snr"][full], b["snr"][full], rtol=1e-6)
