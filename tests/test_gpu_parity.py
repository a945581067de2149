"""GPU parity: the HIP engine against the FP64 oracle on the same seeded windows.

Bar: integer stage results (raw-surface argmax, zoom start/shift, oversampled
argmax) bit-exact; sub-pixel offsets within 0.01 px; SNR and covariance
within 1e-3 relative (FP32 arithmetic vs FP64) where the integer peaks agree.
"""
import numpy as np
import pytest
import torch

from isce2_amd.engine import MODE_STAGED, MODE_FUSED, MODE_AUTO, Engine
from isce2_amd.synth import speckle_pair_shift
from oracle import ampcor_oracle as O

from .gpu_helpers import compare, run_both

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def scene1():
    # 512x512 cut of the CPU-comparison scene: speckle, shift (+3.25, -1.75), 10 dB noise
    return speckle_pair_shift(512, 512, shift=(3.25, -1.75), snr_db=10.0, seed=0)


def _grid(nd, na, s0, skip, hs):
    rd, ra, sd, sa, _, _ = O.start_pixels(nd, na, s0, s0, skip, skip, 0, 0, hs, hs)
    return rd, ra, sd, sa


def _report(name, c):
    print(f"\n[{name}] " + " ".join(f"{k}={v}" for k, v in c.items()), flush=True)


def _assert_parity(c, zoom_allow=0):
    assert c["maxloc_mismatch"] == 0, c
    assert c["start_mismatch"] == 0 and c["shift_mismatch"] == 0, c
    assert c["zoom_mismatch"] <= zoom_allow, c
    assert c["offset_over_tol"] <= zoom_allow, c
    assert c["snr_max_rel"] < 1e-3, c
    assert c["cov_sentinel_mismatch"] == 0 and c["cov_max_rel"] < 1e-3, c


@pytest.mark.parametrize("deramp", [0, 1])
def test_staged_default_config(eng, scene1, deramp):
    ref, sec = scene1
    cfg = O.AmpcorConfig(hs_down=16, hs_across=16, deramp=deramp)
    starts = _grid(6, 6, 16, 64, 16)
    g, o = run_both(eng, ref, sec, cfg, starts, MODE_STAGED, chunk=7)
    c = compare(g, o)
    _report(f"staged deramp={deramp}", c)
    _assert_parity(c)


def test_staged_hs20_sinc_and_time_domain(eng, scene1):
    ref, sec = scene1
    cfg = O.AmpcorConfig(hs_down=20, hs_across=20, surf_method=1, algorithm=1)
    starts = _grid(4, 4, 20, 100, 20)
    g, o = run_both(eng, ref, sec, cfg, starts, MODE_STAGED)
    c = compare(g, o)
    _report("staged sinc", c)
    _assert_parity(c)


def test_staged_real_image_and_small_window(eng, scene1):
    ref, sec = scene1
    ref_r = np.abs(ref).astype(np.float32)
    sec_r = np.abs(sec).astype(np.float32)
    cfg = O.AmpcorConfig(window_h=32, window_w=32, hs_down=8, hs_across=8)
    starts = _grid(5, 5, 8, 90, 8)
    g, o = run_both(eng, ref_r, sec_r, cfg, starts, MODE_STAGED)
    c = compare(g, o)
    _report("staged real", c)
    _assert_parity(c)


@pytest.mark.parametrize("deramp", [0, 1])
def test_staged_odd_rectangular_windows(eng, scene1, deramp):
    """Odd and unequal chip sides: the quadrant padding drops the middle row/column
    of an odd spectrum, the row-pair spectra end on an unpaired row."""
    ref, sec = scene1
    cfg = O.AmpcorConfig(window_h=33, window_w=40, hs_down=7, hs_across=9, deramp=deramp)
    rd, ra, sd, sa, _, _ = O.start_pixels(5, 4, 7, 9, 80, 96, 0, 0, 7, 9)
    g, o = run_both(eng, ref, sec, cfg, (rd, ra, sd, sa), MODE_STAGED, chunk=6)
    c = compare(g, o)
    _report(f"staged odd 33x40 deramp={deramp}", c)
    _assert_parity(c)


def test_staged_edges_and_errors(eng, scene1):
    ref, sec = scene1
    cfg = O.AmpcorConfig(window_h=64, window_w=64, hs_down=16, hs_across=16)
    H, W = ref.shape
    S = 64 + 32
    # windows touching the first and last rows/columns of the images
    rd = np.array([16, H - 64, 16, H - 64 - 16])
    ra = np.array([16, W - 64, W - 64, 40])
    sd = np.array([0, H - S, 0, H - S])
    sa = np.array([0, W - S, W - S, 10])
    g, o = run_both(eng, ref, sec, cfg, (rd, ra, sd, sa), MODE_STAGED)
    c = compare(g, o)
    _report("staged edges", c)
    _assert_parity(c)
    # a chip one pixel past the edge is rejected before any kernel runs
    dev = torch.device("cuda", 0)
    t = torch.from_numpy(ref).to(dev)
    bad = [torch.tensor(v, dtype=torch.int32, device=dev) for v in ([H - 63], [0], [0], [0])]
    with pytest.raises(IndexError):
        eng.run_windows(t, t, *bad, mode=MODE_STAGED)
    # zero windows is a no-op
    empty = [torch.zeros(0, dtype=torch.int32, device=dev) for _ in range(4)]
    r = eng.run_windows(t, t, *empty, mode=MODE_STAGED)
    assert r["offset"].shape == (0, 2)


@pytest.mark.parametrize("W,hs,deramp", [(64, 16, 1), (64, 20, 0), (64, 32, 1), (32, 8, 1), (32, 16, 0),
                                         (32, 32, 1)])
def test_fused_matches_oracle(eng, scene1, W, hs, deramp):
    ref, sec = scene1
    cfg = O.AmpcorConfig(window_h=W, window_w=W, hs_down=hs, hs_across=hs, deramp=deramp)
    S = W + 2 * hs
    nd = (512 - S) // 72
    starts = _grid(nd, nd, hs, 72, hs)
    g, o = run_both(eng, ref, sec, cfg, starts, MODE_FUSED, chunk=5)
    c = compare(g, o)
    _report(f"fused W={W} hs={hs} deramp={deramp}", c)
    _assert_parity(c)


@pytest.mark.parametrize("W,hd,ha,deramp", [(64, 12, 24, 1), (64, 24, 12, 0), (32, 20, 8, 1)])
def test_fused_unequal_half_search(eng, scene1, W, hd, ha, deramp):
    """Different half search ranges down and across (search chip and raw
    surface not square) through the fused kernels."""
    ref, sec = scene1
    cfg = O.AmpcorConfig(window_h=W, window_w=W, hs_down=hd, hs_across=ha, deramp=deramp)
    rd, ra, sd, sa, _, _ = O.start_pixels(4, 4, hd, ha, 100, 100, 0, 0, hd, ha)
    g, o = run_both(eng, ref, sec, cfg, (rd, ra, sd, sa), MODE_FUSED, chunk=6)
    c = compare(g, o)
    _report(f"fused W={W} hs={hd}x{ha} deramp={deramp}", c)
    _assert_parity(c)


def test_fused_real_image_and_edges(eng, scene1):
    ref, sec = scene1
    ref_r = np.abs(ref).astype(np.float32)
    sec_r = (np.abs(sec) + 0.1).astype(np.float32)
    cfg = O.AmpcorConfig(hs_down=20, hs_across=20)
    H, W = ref.shape
    S = 104
    rd = np.array([20, H - 64, 20, H - 84, 200])
    ra = np.array([20, W - 64, W - 64, 60, 200])
    sd = np.array([0, H - S, 0, H - S, 170])
    sa = np.array([0, W - S, W - S, 20, 190])
    g, o = run_both(eng, ref_r, sec_r, cfg, (rd, ra, sd, sa), MODE_FUSED)
    c = compare(g, o)
    _report("fused real/edges", c)
    _assert_parity(c)


def test_fused_equals_staged_integers(eng, scene1):
    ref, sec = scene1
    cfg = O.AmpcorConfig(hs_down=20, hs_across=20)
    starts = _grid(5, 5, 20, 80, 20)
    from .gpu_helpers import params_from_cfg
    eng.setup(params_from_cfg(cfg))
    dev = torch.device("cuda", 0)
    rt, stt = torch.from_numpy(ref).to(dev), torch.from_numpy(sec).to(dev)
    st = [torch.as_tensor(np.asarray(s, np.int32)).to(dev) for s in starts]
    a = eng.run_windows(rt, stt, *st, mode=MODE_FUSED, debug=True)
    b = eng.run_windows(rt, stt, *st, mode=MODE_STAGED, debug=True)
    torch.cuda.synchronize()
    for k in ("maxloc", "start", "shift", "zoom_idx"):
        assert torch.equal(a[k], b[k]), k
    torch.testing.assert_close(a["snr"], b["snr"], rtol=1e-4, atol=0)


def test_staged_chunk_invariance(eng, scene1):
    ref, sec = scene1
    cfg = O.AmpcorConfig(hs_down=16, hs_across=16)
    starts = _grid(6, 6, 16, 64, 16)
    from .gpu_helpers import params_from_cfg
    eng.setup(params_from_cfg(cfg))
    dev = torch.device("cuda", 0)
    rt, stt = torch.from_numpy(ref).to(dev), torch.from_numpy(sec).to(dev)
    st = [torch.as_tensor(np.asarray(s, np.int32)).to(dev) for s in starts]
    a = eng.run_windows(rt, stt, *st, mode=MODE_STAGED, chunk=1)
    b = eng.run_windows(rt, stt, *st, mode=MODE_STAGED, chunk=36)
    torch.cuda.synchronize()
    for k in ("offset", "snr", "cov"):
        assert torch.equal(a[k], b[k]), k


def test_launch_options_do_not_change_results(eng, scene1):
    """fine_threads (fused), fft_block_elems and fft_cols_width (staged) change only launch shapes."""
    ref, sec = scene1
    cfg = O.AmpcorConfig(hs_down=20, hs_across=20)
    starts = _grid(5, 5, 20, 80, 20)
    from .gpu_helpers import params_from_cfg
    eng.setup(params_from_cfg(cfg))
    dev = torch.device("cuda", 0)
    rt, stt = torch.from_numpy(ref).to(dev), torch.from_numpy(sec).to(dev)
    st = [torch.as_tensor(np.asarray(s, np.int32)).to(dev) for s in starts]
    res = {}
    eng.set_option("fine_kernel", 1)          # fine_threads shapes the whole-template fine kernel
    for mode, name, vals in ((MODE_FUSED, "fine_threads", (512, 1024)),
                             (MODE_FUSED, "fine_kernel", (1, 2)),
                             (MODE_STAGED, "fft_block_elems", (1024, 2048, 4096, 8192)),
                             (MODE_STAGED, "fft_cols_width", (0, 16, 8))):
        for v in vals:
            eng.set_option(name, v)
            r = eng.run_windows(rt, stt, *st, mode=mode, debug=True)
            torch.cuda.synchronize()
            res[(name, v)] = {k: t.cpu().numpy() for k, t in r.items()}
        eng.set_option(name, {"fft_block_elems": 2048, "fine_threads": 1024, "fft_cols_width": 8,
                              "fine_kernel": 2}[name])
    base_f, base_s = res[("fine_threads", 1024)], res[("fft_block_elems", 2048)]
    for (name, v), r in res.items():
        base = base_f if name in ("fine_threads", "fine_kernel") else base_s
        for k in ("maxloc", "zoom_idx", "offset"):
            assert np.array_equal(r[k], base[k]), (name, v, k)
        np.testing.assert_allclose(r["snr"], base["snr"], rtol=1e-5)
    with pytest.raises(ValueError):
        eng.set_option("fine_threads", 256)
    with pytest.raises(ValueError):
        eng.set_option("no_such_option", 1)


@pytest.mark.parametrize("pin", [True, False])
def test_host_inputs_banded_equals_device_resident(eng, scene1, pin):
    """run_windows_from_host (row bands uploaded on a copy stream while earlier
    bands run) gives the device-resident call's offsets bit for bit (SNR and
    covariance of re-run near ties come from a full FP64 run there)."""
    ref, sec = scene1
    cfg = O.AmpcorConfig(hs_down=20, hs_across=20)
    rd, ra, sd, sa = _grid(5, 5, 20, 80, 20)
    from .gpu_helpers import params_from_cfg
    eng.setup(params_from_cfg(cfg))
    dev = torch.device("cuda", 0)
    st = [torch.as_tensor(np.asarray(s, np.int32)).to(dev) for s in (rd, ra, sd, sa)]
    a = eng.run_windows(torch.from_numpy(ref).to(dev), torch.from_numpy(sec).to(dev), *st)
    torch.cuda.synchronize()
    h_ref, h_sec = (torch.from_numpy(ref).pin_memory(), torch.from_numpy(sec).pin_memory()) if pin else (ref, sec)
    b = eng.run_windows_from_host(h_ref, h_sec, rd, ra, sd, sa, bands=3)
    assert np.array_equal(a["offset"].cpu().numpy(), b["offset"])
    for k in ("snr", "cov"):
        np.testing.assert_allclose(b[k], a[k].cpu().numpy(), rtol=1e-5, atol=0)
    bad = np.array(sd)
    bad[-1] = ref.shape[0]             # a search chip past the last row
    with pytest.raises(IndexError):
        eng.run_windows_from_host(h_ref, h_sec, rd, ra, bad, sa, bands=2)


@pytest.mark.parametrize("taus", [(2e-6, float("inf")), (float("inf"), float("inf"))])
def test_host_inputs_fp64_rerun_of_every_window(eng, scene1, taus):
    """The host-input runner's FP64 re-run, forced for every window: from the
    zoom stage (every fine margin below tau, raw-surface margins at the default
    threshold) and whole-pipeline. Its outputs equal
    the device-resident call's (which re-runs inside the engine) and the FP64
    oracle's integer peaks and offsets."""
    from .gpu_helpers import params_from_cfg
    ref, sec = scene1
    cfg = O.AmpcorConfig(hs_down=20, hs_across=20)
    rd, ra, sd, sa = _grid(5, 5, 20, 80, 20)
    eng.setup(params_from_cfg(cfg))
    dev = torch.device("cuda", 0)
    st = [torch.as_tensor(np.asarray(s, np.int32)).to(dev) for s in (rd, ra, sd, sa)]
    eng.set_precision(1, *taus)
    try:
        eng.precision_stats()
        a = eng.run_windows(torch.from_numpy(ref).to(dev), torch.from_numpy(sec).to(dev), *st)
        torch.cuda.synchronize()
        checked, rerun = eng.precision_stats()
        b = eng.run_windows_from_host(ref, sec, rd, ra, sd, sa, bands=2)
    finally:
        eng.set_precision(1, 2e-6, 2e-6)
    print(f"\n[host rerun taus={taus}] checked={checked} rerun={rerun}", flush=True)
    assert rerun == len(rd)
    assert np.array_equal(a["offset"].cpu().numpy(), b["offset"])
    for k in ("snr", "cov"):
        np.testing.assert_allclose(b[k], a[k].cpu().numpy(), rtol=1e-6, atol=0)
    o = O.run_windows(ref, sec, rd, ra, sd, sa, cfg)
    assert np.abs(b["offset"] - o["offset"]).max() < 0.01
    rel = np.abs(b["snr"] - o["snr"]) / np.abs(o["snr"])
    assert rel.max() < 1e-3


@pytest.mark.parametrize("mode", [MODE_FUSED, MODE_STAGED])
def test_pitched_images_of_different_sizes(eng, scene1, mode):
    """Reference and secondary of different shapes, the secondary a column view
    of a wider image (row pitch 512 > width 470): chips are read through the pitch."""
    ref, sec = scene1
    sec_small = np.ascontiguousarray(sec[10:490, 20:490])          # 480 x 470
    cfg = O.AmpcorConfig(hs_down=16, hs_across=16)
    rd, ra, _, _, _, _ = O.start_pixels(4, 4, 40, 40, 96, 96, 0, 0, 16, 16)
    sd, sa = rd - 10 - 16, ra - 20 - 16                              # same scene points in the crop
    from .gpu_helpers import params_from_cfg
    eng.setup(params_from_cfg(cfg))
    dev = torch.device("cuda", 0)
    wide = torch.zeros((480, 512), dtype=torch.complex64, device=dev)
    wide[:, :470] = torch.from_numpy(sec_small).to(dev)
    sec_view = wide[:, :470]
    assert sec_view.stride(0) == 512
    st = [torch.as_tensor(np.asarray(s, np.int32)).to(dev) for s in (rd, ra, sd, sa)]
    g = eng.run_windows(torch.from_numpy(ref).to(dev), sec_view, *st, mode=mode, debug=True)
    torch.cuda.synchronize()
    g = {k: v.cpu().numpy() for k, v in g.items()}
    o = O.run_windows(ref, sec_small, rd, ra, sd, sa, cfg)
    c = compare(g, o)
    _report(f"pitched mode={mode}", c)
    _assert_parity(c)
