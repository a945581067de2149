"""Serving: continuous batching across requests on one GPU.

Each job's result is the same alone and in a mix (bit-identical offsets, SNR
and covariance), equals a direct engine call on its own images, and jobs that
must wait for HBM budget are admitted as earlier ones finish.
"""
import numpy as np
import pytest
import torch

from isce2_amd.engine import Engine, WindowParams
from isce2_amd.serving import OffsetJob, OffsetServer, job_windows
from isce2_amd.synth import speckle_pair_warp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene():
    ref, sec, _, _ = speckle_pair_warp(1024, 1024, max_amp=4.0, jitter=0.05, tile=256, seed=21,
                                       device=torch.device("cuda", 0))
    return ref, sec


def _job(scene, r0, c0, h, w, W, hs, skip):
    ref, sec = scene
    return OffsetJob(ref=np.ascontiguousarray(ref[r0:r0 + h, c0:c0 + w]),
                     sec=np.ascontiguousarray(sec[r0:r0 + h, c0:c0 + w]),
                     window=W, half_search=hs, skip_down=skip, skip_across=skip)


def _direct(job):
    e = Engine(0)
    e.setup(WindowParams(window_h=job.window, window_w=job.window, hs_down=job.half_search,
                         hs_across=job.half_search, deramp=job.deramp))
    nd, na, *st = job_windows(job)
    dev = torch.device("cuda", 0)
    st = [torch.as_tensor(s.astype(np.int32), device=dev) for s in st]
    o = e.run_windows(torch.from_numpy(job.ref).to(dev), torch.from_numpy(job.sec).to(dev), *st)
    r = {k: v.cpu().numpy() for k, v in o.items()}
    e.close()
    return r["offset"].reshape(nd, na, 2), r["snr"].reshape(nd, na), r["cov"].reshape(nd, na, 3)


def test_alone_equals_mixed_equals_direct(scene):
    a = dict(r0=100, c0=60, h=600, w=700, W=64, hs=16, skip=32)
    with OffsetServer(device=0, hbm_budget=2e9, pitch=1024) as srv:
        alone = srv.submit(_job(scene, **a)).result(timeout=120)
    mix = [dict(r0=0, c0=0, h=512, w=512, W=32, hs=8, skip=16), dict(r0=300, c0=200, h=700, w=800, W=64, hs=16, skip=32),
           dict(r0=50, c0=10, h=900, w=900, W=128, hs=16, skip=64), a, dict(r0=10, c0=300, h=520, w=700, W=64, hs=32, skip=32),
           dict(r0=200, c0=100, h=800, w=600, W=256, hs=8, skip=128)]
    with OffsetServer(device=0, hbm_budget=2e9, pitch=1024, max_batch=700) as srv:
        futs = [srv.submit(_job(scene, **m)) for m in mix]
        res = [f.result(timeout=300) for f in futs]
        stats = dict(srv.stats)
    mixed = res[3]
    print(f"\n[serving mix] jobs={len(res)} batches={stats['batches']} windows={stats['windows']} "
          f"latencies={[round(r['latency_s'], 3) for r in res]}", flush=True)
    for k in ("offset", "snr", "cov"):
        assert np.array_equal(alone[k], mixed[k], equal_nan=True), k
    d_off, d_snr, d_cov = _direct(_job(scene, **a))
    assert np.array_equal(alone["offset"], d_off) and np.array_equal(alone["snr"], d_snr)
    assert np.array_equal(alone["cov"], d_cov, equal_nan=True)
    for m, r in zip(mix, res):
        d = _direct(_job(scene, **m))
        assert np.array_equal(r["offset"], d[0]), m


def test_bad_requests_fail_alone(scene):
    """Malformed requests fail their own future at admission; good requests
    submitted around them complete and match a direct engine call, every
    arena row and budget byte comes back, and the server keeps accepting work."""
    good = dict(r0=100, c0=60, h=600, w=700, W=64, hs=16, skip=64)
    bad = []
    j = _job(scene, **good)
    j.window = 0                                   # no window
    bad.append(j)
    j = _job(scene, **good)
    j.half_search = 0                              # no search range
    bad.append(j)
    j = _job(scene, **good)
    j.window, j.half_search = 600, 40              # the window grid does not fit the image
    bad.append(j)
    j = _job(scene, **good)
    j.ref = np.zeros((600, 1100), np.complex64)    # rows wider than the arena pitch
    j.sec = np.zeros((600, 1100), np.complex64)
    j.width = 700
    bad.append(j)
    j = _job(scene, **good)
    j.ref = j.ref.real.copy()                      # not complex64
    bad.append(j)
    j = _job(scene, **good)
    j.skip_down = 0                                # no window grid
    bad.append(j)
    j = _job(scene, **good)
    j.n_down = 50                                  # explicit grid running off the image
    bad.append(j)
    with OffsetServer(device=0, hbm_budget=2e9, pitch=1024) as srv:
        futs = []
        for b in bad:
            futs.append(srv.submit(_job(scene, **good)))
            futs.append(srv.submit(b))
        res = []
        for i, f in enumerate(futs):
            if i % 2 == 0:
                res.append(f.result(timeout=120))
            else:
                with pytest.raises(Exception):
                    f.result(timeout=120)
        later = srv.submit(_job(scene, **good)).result(timeout=120)
        assert srv.used < 1.0 and srv.arena.free_rows() == srv.arena.rows
    d_off, d_snr, _ = _direct(_job(scene, **good))
    for r in res + [later]:
        assert np.array_equal(r["offset"], d_off) and np.array_equal(r["snr"], d_snr)


def test_admission_waits_for_budget(scene):
    jobs = [dict(r0=0, c0=0, h=600, w=600, W=64, hs=16, skip=64) for _ in range(5)]
    # budget for about two jobs at a time: 2 images x 600 rows x 1024 pitch x 8 B + margin
    budget = 2 * (2 * 600 * 1024 * 8 + 64e6 + 1e6)
    with OffsetServer(device=0, hbm_budget=budget, pitch=1024) as srv:
        futs = [srv.submit(_job(scene, **m)) for m in jobs]
        res = [f.result(timeout=120) for f in futs]
        assert srv.used < 1.0 and srv.arena.free_rows() == srv.arena.rows
    for r in res[1:]:
        assert np.array_equal(r["offset"], res[0]["offset"])
    assert max(r["queue_s"] for r in res) > 0.0
