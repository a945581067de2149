"""Staged x2 chip oversampling: separable half-sample form against the FP64 oracle.

The staged pipeline oversamples even-sided chips by one half-sample
interpolation per axis (option ``oversampler`` 2, the default: in-place
transforms for the chip sides of windows 32/64/128/256 and their zoom chips;
option 3: the general out-of-place transforms) instead of the
quadrant-padded inverse 2-D FFT the oracle restates (option 1). They are
equal in exact arithmetic, so:

* the FP64 staged pipeline with the separable form equals the oracle on every
  integer result and offset (square and non-square even windows, deramp 0
  and 1, windows 32 to 128);
* in FP32 the separable forms give the 2-D form's coarse peaks (the coarse
  stage does not oversample) and its fine peaks except where the fine argmax
  margin is below the near-tie threshold.
"""
import numpy as np
import pytest
import torch

from isce2_amd.engine import MODE_STAGED, Engine
from isce2_amd.synth import speckle_pair_shift
from oracle import ampcor_oracle as O

from .gpu_helpers import compare, run_both

pytestmark = pytest.mark.gpu

TAU = 2e-6
# (window_h, window_w, hs_down, hs_across, deramp)
CASES = [(32, 32, 8, 8, 1), (48, 64, 6, 10, 1), (64, 40, 12, 5, 0), (128, 128, 16, 16, 1)]


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def scene():
    return speckle_pair_shift(768, 768, shift=(3.4, -5.7), snr_db=12.0, seed=21)


def _starts(cfg, shape, n=6):
    H, W = shape
    wh, ww, hd, ha = cfg.window_h, cfg.window_w, cfg.hs_down, cfg.hs_across
    nd = min(n, (H - wh - 2 * hd - 16) // (wh + 4))
    na = min(n, (W - ww - 2 * ha - 16) // (ww + 4))
    rd, ra, sd, sa, _, _ = O.start_pixels(nd, na, hd + 8, ha + 8, wh + 4, ww + 4, 0, 0, hd, ha)
    return rd, ra, sd, sa


@pytest.mark.parametrize("form", [2, 3])
@pytest.mark.parametrize("case", CASES)
def test_fp64_separable_equals_oracle(eng, scene, case, form):
    wh, ww, hd, ha, deramp = case
    ref, sec = scene
    cfg = O.AmpcorConfig(window_h=wh, window_w=ww, hs_down=hd, hs_across=ha, deramp=deramp)
    eng.set_option("oversampler", form)
    eng.set_precision(2)
    try:
        g, o = run_both(eng, ref, sec, cfg, _starts(cfg, ref.shape), MODE_STAGED)
    finally:
        eng.set_precision(1)
        eng.set_option("oversampler", 2)
    c = compare(g, o)
    print(f"\n[separable fp64 form={form} {case}] {c}", flush=True)
    assert c["n"] >= 16
    assert c["maxloc_mismatch"] == 0 and c["zoom_mismatch"] == 0, c
    assert c["offset_max_err"] == 0.0, c
    assert c["snr_max_rel"] < 1e-6, c


@pytest.mark.parametrize("case", CASES)
def test_fp32_forms_agree_up_to_ties(eng, scene, case):
    wh, ww, hd, ha, deramp = case
    ref, sec = scene
    cfg = O.AmpcorConfig(window_h=wh, window_w=ww, hs_down=hd, hs_across=ha, deramp=deramp)
    dev = torch.device("cuda", 0)
    st = [torch.as_tensor(np.asarray(s).astype(np.int32)).to(dev) for s in _starts(cfg, ref.shape, n=12)]
    from .gpu_helpers import params_from_cfg
    eng.setup(params_from_cfg(cfg))
    ref_t, sec_t = torch.from_numpy(ref).to(dev), torch.from_numpy(sec).to(dev)
    out = {}
    eng.set_precision(0)
    try:
        for form in (1, 2, 3):
            eng.set_option("oversampler", form)
            r = eng.run_windows(ref_t, sec_t, *st, mode=MODE_STAGED, debug=True)
            torch.cuda.synchronize()
            out[form] = {k: v.cpu().numpy() for k, v in r.items()}
    finally:
        eng.set_option("oversampler", 2)
        eng.set_precision(1)
    a = out[1]
    for form in (2, 3):
        b = out[form]
        assert np.array_equal(a["maxloc"], b["maxloc"])
        flip = np.any(a["zoom_idx"] != b["zoom_idx"], axis=1)
        print(f"\n[fp32 forms 1 vs {form} {case}] n={len(flip)} fine_flips={int(flip.sum())} max_flip_margin="
              f"{float(np.maximum(a['margin_fine'], b['margin_fine'])[flip].max()) if flip.any() else None}", flush=True)
        assert np.all(np.minimum(a["margin_fine"], b["margin_fine"])[flip] < TAU)
        same = ~flip
        assert np.array_equal(a["offset"][same], b["offset"][same])
