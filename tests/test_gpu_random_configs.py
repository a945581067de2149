"""Seeded random window configurations through the HIP engine against the FP64 oracle.

Each case draws window sides (equal or not, odd or even), half search ranges
down and across, the deramp method (0 amplitude chips, 1 deramp, 2 complex
without deramp), complex or real images and a window grid inside a 384x384
cut of the shifted-speckle scene. The staged pipeline runs every case; the
cases the fused kernels cover (square windows of 32 or 64, zoom 16) run them too. Bar as in test_gpu_parity:
integer peaks bit-exact, offsets within 0.01 px, SNR and covariance within
1e-3 relative.
"""
import numpy as np
import pytest
import torch

from isce2_amd.engine import MODE_FUSED, MODE_STAGED, Engine
from isce2_amd.synth import speckle_pair_shift
from oracle import ampcor_oracle as O

from .gpu_helpers import compare, params_from_cfg, run_both

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def scene():
    ref, sec = speckle_pair_shift(384, 384, shift=(-2.6, 1.3), snr_db=15.0, seed=4)
    return ref, sec


def _case(seed):
    rng = np.random.default_rng(seed)
    if rng.uniform() < 0.4:
        wh = ww = int(rng.choice([32, 64]))
    else:
        wh, ww = (int(v) for v in rng.integers(12, 72, size=2))
    hd, ha = (int(v) for v in rng.integers(3, 25, size=2))
    deramp = int(rng.choice([0, 1, 2]))
    real = bool(rng.uniform() < 0.25)
    return wh, ww, hd, ha, deramp, real


@pytest.mark.parametrize("seed", range(40))
def test_random_config_matches_oracle(eng, scene, seed):
    wh, ww, hd, ha, deramp, real = _case(seed)
    ref, sec = scene
    if real:
        ref, sec = np.abs(ref).astype(np.float32), (np.abs(sec) + 0.05).astype(np.float32)
    cfg = O.AmpcorConfig(window_h=wh, window_w=ww, hs_down=hd, hs_across=ha, deramp=deramp)
    H, W = ref.shape
    nd = min(4, (H - wh - 2 * hd) // (wh + 8))
    na = min(4, (W - ww - 2 * ha) // (ww + 8))
    rd, ra, sd, sa, _, _ = O.start_pixels(nd, na, hd + 3, ha + 5, wh + 8, ww + 8, 0, 0, hd, ha)
    modes = [MODE_STAGED]
    if eng.setup(params_from_cfg(cfg, not real))["fused"]:
        modes.append(MODE_FUSED)
    for mode in modes:
        g, o = run_both(eng, ref, sec, cfg, (rd, ra, sd, sa), mode, chunk=5)
        c = compare(g, o)
        print(f"\n[random seed={seed} {wh}x{ww} hs={hd}x{ha} deramp={deramp} real={real} mode={mode}] "
              + " ".join(f"{k}={v}" for k, v in c.items()), flush=True)
        assert c["maxloc_mismatch"] == 0, c
        assert c["start_mismatch"] == 0 and c["shift_mismatch"] == 0, c
        assert c["zoom_mismatch"] == 0 and c["offset_over_tol"] == 0, c
        assert c["snr_max_rel"] < 1e-3, c
        assert c["cov_sentinel_mismatch"] == 0 and c["cov_max_rel"] < 1e-3, c


EXOTIC = [dict(window_h=48, window_w=40, hs_down=10, hs_across=12, zoom=8),
          dict(window_h=48, window_w=48, hs_down=12, hs_across=10, zoom=32),
          dict(window_h=40, window_w=56, hs_down=9, hs_across=7, corr_stat=9),
          dict(window_h=64, window_w=64, hs_down=16, hs_across=16, surf_ovs=8),
          dict(window_h=32, window_w=32, hs_down=8, hs_across=8, raw_ovs=4),
          dict(window_h=36, window_w=44, hs_down=11, hs_across=9, surf_method=1),
          dict(window_h=128, window_w=128, hs_down=8, hs_across=10)]


@pytest.mark.parametrize("kw", EXOTIC, ids=[",".join(f"{k}={v}" for k, v in d.items()) for d in EXOTIC])
def test_non_default_settings_match_oracle(eng, scene, kw):
    """Zoom windows of 8 and 32 (a non-16 zoom takes the general FFT surface
    path), a smaller SNR patch, surface oversampling 8, raw oversampling 4, the
    sinc surface on unequal sides and a 128 window, through the staged pipeline."""
    ref, sec = scene
    cfg = O.AmpcorConfig(**kw)
    wh, ww, hd, ha = cfg.window_h, cfg.window_w, cfg.hs_down, cfg.hs_across
    H, W = ref.shape
    kd, ka = wh // 2 + 8, ww // 2 + 8
    nd = min(4, (H - wh - 2 * hd - 8) // kd + 1)
    na = min(4, (W - ww - 2 * ha - 8) // ka + 1)
    rd, ra, sd, sa, _, _ = O.start_pixels(nd, na, hd + 3, ha + 5, kd, ka, 0, 0, hd, ha)
    g, o = run_both(eng, ref, sec, cfg, (rd, ra, sd, sa), MODE_STAGED, chunk=7)
    c = compare(g, o)
    print(f"\n[settings {kw}] " + " ".join(f"{k}={v}" for k, v in c.items()), flush=True)
    assert c["maxloc_mismatch"] == 0, c
    assert c["start_mismatch"] == 0 and c["shift_mismatch"] == 0, c
    assert c["zoom_mismatch"] == 0 and c["offset_over_tol"] == 0, c
    assert c["snr_max_rel"] < 1e-3, c
    assert c["cov_sentinel_mismatch"] == 0 and c["cov_max_rel"] < 1e-3, c
