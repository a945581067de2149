"""FP64 sub-pixel peak refinement on the GPU against the oracle's.

The engine evaluates the FFT-interpolated surface in FP64 at the peak and its
axis neighbours directly (a Dirichlet-kernel sum over the trimmed surface); the
oracle takes the same three values per axis from its FP64 oversampled surface.
Bar: where the 1/32-pixel peaks agree (every window here), refined offsets
within 1e-4 px of the oracle's (the surfaces differ by FP32 rounding, about
1e-7 relative; the vertex moves by ~1e-6 px per 1e-7), and never more than half
a grid cell from the grid offset.
"""
import numpy as np
import pytest
import torch

from isce2_amd.engine import MODE_FUSED, MODE_STAGED, Engine
from isce2_amd.synth import speckle_pair_shift
from oracle import ampcor_oracle as O

from .gpu_helpers import params_from_cfg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def scene():
    return speckle_pair_shift(512, 512, shift=(2.61, -1.37), snr_db=10.0, seed=5)


def _check(eng, ref, sec, cfg, starts, mode):
    dev = torch.device("cuda", 0)
    eng.setup(params_from_cfg(cfg))
    st = [torch.as_tensor(np.asarray(s, np.int32)).to(dev) for s in starts]
    g = eng.run_windows(torch.from_numpy(ref).to(dev), torch.from_numpy(sec).to(dev), *st, mode=mode, debug=True,
                        refine=True)
    torch.cuda.synchronize()
    g = {k: v.cpu().numpy() for k, v in g.items()}
    o = O.run_windows(ref, sec, *starts, cfg)
    same = np.all(g["zoom_idx"] == o["zoom_idx"], axis=1) & np.all(g["start"] == o["start"], axis=1)
    err = np.abs(g["offset_fp64"] - o["offset_fp64"]).max(axis=1)
    cell = 1.0 / (cfg.surf_ovs * cfg.raw_ovs)
    print(f"\n[refine mode={mode} W={cfg.window_h} zoom={cfg.zoom}] windows={len(err)} same_peaks={int(same.sum())} "
          f"max|d fp64|={err[same].max():.3g} px", flush=True)
    assert same.all()
    assert err.max() < 1e-4
    assert (np.abs(g["offset_fp64"] - g["offset"]) <= 0.5 * cell + 1e-6).all()
    return g


@pytest.mark.parametrize("mode", [MODE_FUSED, MODE_STAGED])
def test_refine_default_config(eng, scene, mode):
    ref, sec = scene
    cfg = O.AmpcorConfig(hs_down=20, hs_across=20)
    starts = O.start_pixels(5, 5, 30, 30, 80, 80, 0, 0, 20, 20)[:4]
    _check(eng, ref, sec, cfg, starts, mode)


def test_refine_window32_and_zoom8(eng, scene):
    ref, sec = scene
    cfg = O.AmpcorConfig(window_h=32, window_w=32, hs_down=12, hs_across=12)
    _check(eng, ref, sec, cfg, O.start_pixels(6, 6, 20, 20, 70, 70, 0, 0, 12, 12)[:4], MODE_FUSED)
    cfg = O.AmpcorConfig(hs_down=16, hs_across=16, zoom=8, surf_ovs=8)
    _check(eng, ref, sec, cfg, O.start_pixels(4, 4, 30, 30, 100, 100, 0, 0, 16, 16)[:4], MODE_STAGED)


def test_refine_through_fp64_rerun_and_host_runner(eng, scene):
    """Windows re-run in FP64 (forced for all) carry the re-run's refinement;
    the host-input runner returns the same refined offsets."""
    ref, sec = scene
    cfg = O.AmpcorConfig(hs_down=20, hs_across=20)
    starts = O.start_pixels(5, 5, 30, 30, 80, 80, 0, 0, 20, 20)[:4]
    eng.set_precision(1, float("inf"), float("inf"))
    try:
        g = _check(eng, ref, sec, cfg, starts, MODE_FUSED)
        h = eng.run_windows_from_host(ref, sec, *starts, refine=True)
    finally:
        eng.set_precision(1, 2e-6, 2e-6)
    assert np.array_equal(h["offset_fp64"], g["offset_fp64"])


def test_refine_needs_fft_surface(eng, scene):
    ref, sec = scene
    cfg = O.AmpcorConfig(hs_down=20, hs_across=20, surf_method=1)
    eng.setup(params_from_cfg(cfg))
    dev = torch.device("cuda", 0)
    st = [torch.as_tensor(np.asarray(s, np.int32)).to(dev) for s in O.start_pixels(2, 2, 30, 30, 80, 80, 0, 0, 20, 20)[:4]]
    with pytest.raises(NotImplementedError):
        eng.run_windows(torch.from_numpy(ref).to(dev), torch.from_numpy(sec).to(dev), *st, refine=True)
