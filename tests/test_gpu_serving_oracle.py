"""Serving at the shape of the Poisson-arrival workload, checked against the FP64 oracle.

Jobs arrive as a seeded Poisson stream (20 jobs/s) of crops of an 8192x4096
warped speckle pair: one full 8192x4096 job on a sparse window grid, the rest
log-uniform in size, with windows of all four sizes 32/64/128/256 and half
searches 8/16/32 (skip = window/2). A fixed sample of windows of EVERY job is
recomputed by the oracle: offsets within 0.01 px (so the integer and 1/32-pixel
peaks are the oracle's), SNR and covariance within 1e-3 relative. Two jobs are
also re-run alone on a fresh server: bit-identical to their results in the mix.
"""
import time

import numpy as np
import pytest
import torch

from isce2_amd.serving import OffsetJob, OffsetServer, job_windows
from isce2_amd.synth import speckle_pair_warp
from oracle import ampcor_oracle as O

pytestmark = pytest.mark.gpu

SAMPLE = 5          # oracle windows per job


@pytest.fixture(scope="module")
def pair():
    ref, sec, _, _ = speckle_pair_warp(8192, 4096, max_amp=8.0, jitter=0.05, tile=256, seed=0,
                                       device=torch.device("cuda", 0))
    return ref, sec


def _jobs(ref, sec, n=14, seed=11):
    rng = np.random.default_rng(seed)
    H, Wd = ref.shape
    classes = [(W, hs) for W in (32, 64, 128, 256) for hs in (8, 16, 32)]
    order = rng.permutation(len(classes))
    jobs = [OffsetJob(ref=ref, sec=sec, window=64, half_search=20, skip_down=1024, skip_across=1024)]
    for i in range(1, n):
        s = rng.uniform(0, 0.25)                        # 512..1024 rows, 512..861 columns
        h, w = int(round(512 * 16 ** s)), int(round(512 * 8 ** s))
        W, hs = classes[order[(i - 1) % len(classes)]]
        if W + 4 * hs > min(h, w):
            h = w = max(h, w, W + 4 * hs + 64)
        r0 = int(rng.integers(0, H - h + 1))
        c0 = int(rng.integers(0, Wd - w + 1))
        jobs.append(OffsetJob(ref=np.ascontiguousarray(ref[r0:r0 + h, c0:c0 + w]),
                              sec=np.ascontiguousarray(sec[r0:r0 + h, c0:c0 + w]),
                              window=W, half_search=hs, skip_down=W // 2, skip_across=W // 2))
    return jobs, rng.exponential(1.0 / 20.0, size=n)


def _serve(jobs, gaps=None):
    with OffsetServer(device=0, hbm_budget=8e9, pitch=4096, max_batch=8192) as srv:
        futs, t_next = [], time.perf_counter()
        for k, j in enumerate(jobs):
            if gaps is not None:
                t_next += gaps[k]
                dt = t_next - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
            futs.append(srv.submit(j))
        return [f.result(timeout=600) for f in futs]


def test_serving_mix_against_oracle(pair):
    ref, sec = pair
    jobs, gaps = _jobs(ref, sec)
    res = _serve(jobs, gaps)
    lat = np.array([r["latency_s"] for r in res]) * 1e3
    print(f"\n[serving oracle] jobs={len(jobs)} windows={sum(r['windows'] for r in res)} "
          f"p50={np.percentile(lat, 50):.1f} ms p99={np.percentile(lat, 99):.1f} ms", flush=True)
    assert {(j.window, j.half_search) for j in jobs[1:]} >= {(W, hs) for W in (32, 64, 128, 256) for hs in (8, 16, 32)}
    rng = np.random.default_rng(3)
    worst = 0.0
    for j, r in zip(jobs, res):
        nd, na, rd, ra, sd, sa = job_windows(j)
        k = np.sort(rng.choice(nd * na, size=min(SAMPLE, nd * na), replace=False))
        cfg = O.AmpcorConfig(window_h=j.window, window_w=j.window, hs_down=j.half_search,
                             hs_across=j.half_search, deramp=j.deramp)
        o = O.run_windows(j.ref, j.sec, rd[k], ra[k], sd[k], sa[k], cfg)
        off = r["offset"].reshape(-1, 2)[k]
        snr = r["snr"].reshape(-1)[k]
        cov = r["cov"].reshape(-1, 3)[k]
        err = np.abs(off - o["offset"]).max()
        worst = max(worst, float(err))
        assert err < 0.01, (j.window, j.half_search, err)
        np.testing.assert_allclose(snr, o["snr"], rtol=1e-3)
        sent = np.all(o["cov"] == [99, 99, 0], axis=1)
        assert np.array_equal(sent, np.all(cov == [99, 99, 0], axis=1))
        scale = np.abs(o["cov"]).max(axis=1, keepdims=True)
        assert (np.abs(cov - o["cov"]) <= 1e-3 * scale).all(), (j.window, j.half_search)
    print(f"[serving oracle] max offset error vs FP64 oracle {worst:.3g} px", flush=True)
    # alone == mixed for the full-size job and one W=256 job
    pick = [0] + [i for i, j in enumerate(jobs) if j.window == 256][:1]
    alone = _serve([jobs[i] for i in pick])
    for i, a in zip(pick, alone):
        for key in ("offset", "snr", "cov"):
            assert np.array_equal(a[key], res[i][key], equal_nan=True), (i, key)
