"""Shared helpers of the GPU parity tests: run engine and oracle on the same windows."""
import numpy as np
import torch

from isce2_amd.engine import Engine, WindowParams
from oracle import ampcor_oracle as O


def params_from_cfg(cfg: O.AmpcorConfig, complex_image=True) -> WindowParams:
    return WindowParams(window_h=cfg.window_h, window_w=cfg.window_w, hs_down=cfg.hs_down,
                        hs_across=cfg.hs_across, raw_ovs=cfg.raw_ovs, zoom=cfg.zoom, surf_ovs=cfg.surf_ovs,
                        surf_method=cfg.surf_method, corr_stat=cfg.corr_stat, deramp=cfg.deramp,
                        algorithm=cfg.algorithm, image_complex=int(complex_image))


def run_both(eng: Engine, ref, sec, cfg: O.AmpcorConfig, starts, mode, chunk=0):
    """Run the engine (device) and the oracle (host) on the same windows."""
    ref_d, ref_a, sec_d, sec_a = [np.asarray(s, np.int64) for s in starts]
    dev = torch.device("cuda", eng.device)
    eng.setup(params_from_cfg(cfg, np.iscomplexobj(ref)))
    ref_t = torch.from_numpy(np.ascontiguousarray(ref)).to(dev)
    sec_t = torch.from_numpy(np.ascontiguousarray(sec)).to(dev)
    st = [torch.as_tensor(s.astype(np.int32)).to(dev) for s in (ref_d, ref_a, sec_d, sec_a)]
    g = eng.run_windows(ref_t, sec_t, *st, mode=mode, chunk=chunk, debug=True)
    torch.cuda.synchronize(dev)
    g = {k: v.cpu().numpy() for k, v in g.items()}
    o = O.run_windows(ref, sec, ref_d, ref_a, sec_d, sec_a, cfg)
    return g, o


def compare(g, o, off_tol=0.01, rel=1e-3):
    """Return a dict of mismatch counts / error maxima between engine and oracle."""
    n = len(o["offset"])
    same_loc = np.all(g["maxloc"] == o["maxloc"], axis=1)
    same_zoom = np.all(g["zoom_idx"] == o["zoom_idx"], axis=1)
    off_err = np.abs(g["offset"] - o["offset"]).max(axis=1)
    snr_rel = np.abs(g["snr"] - o["snr"]) / np.maximum(np.abs(o["snr"]), 1e-30)
    sentinel_o = np.all(o["cov"] == [99, 99, 0], axis=1)
    sentinel_g = np.all(g["cov"] == [99, 99, 0], axis=1)
    cov_rel = np.abs(g["cov"] - o["cov"]).max(axis=1) / np.maximum(np.abs(o["cov"]).max(axis=1), 1e-30)
    return dict(
        n=n,
        maxloc_mismatch=int((~same_loc).sum()),
        start_mismatch=int((~np.all(g["start"] == o["start"], axis=1)).sum()),
        shift_mismatch=int((~np.all(g["shift"] == o["shift"], axis=1)).sum()),
        zoom_mismatch=int((~same_zoom).sum()),
        offset_max_err=float(off_err.max()),
        offset_over_tol=int((off_err > off_tol).sum()),
        snr_max_rel=float(snr_rel[same_loc].max()) if same_loc.any() else 0.0,
        cov_sentinel_mismatch=int((sentinel_o != sentinel_g).sum()),
        cov_max_rel=float(cov_rel[same_loc & ~sentinel_o & ~sentinel_g].max())
        if (same_loc & ~sentinel_o & ~sentinel_g).any() else 0.0,
        maxval_max_abs=float(np.abs(g["maxval"] - o["maxval"]).max()),
    )
