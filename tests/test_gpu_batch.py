"""Uneven chip-pair batches: graph-captured multi-class step vs the FP64 oracle.

Pairs of window 32/64/128/256 and half search 8/16/32 (sub-pixel shifts up
to 0.8 of the search range) run as one ChipPairBatch step. Bar: integer peaks
bit-exact against the oracle and offsets identical (precision 1: FP32 kernels
plus FP64 re-run of near ties); graph replays bit-identical to eager steps;
autotuned schedules give the same results.
"""
import numpy as np
import pytest
import torch

from isce2_amd.autotune import Autotuner
from isce2_amd.batch import ChipPairBatch
from isce2_amd.engine import MODE_STAGED
from isce2_amd.synth import chip_pairs_uneven_grouped
from oracle import ampcor_oracle as O

pytestmark = pytest.mark.gpu


def _oracle(groups):
    """The oracle over each class at once: the class's chips stacked into one tall
    reference and one tall secondary image, windows at (i*W, 0) and (i*S, 0)."""
    out = {}
    for (W, hs), g in groups.items():
        cfg = O.AmpcorConfig(window_h=W, window_w=W, hs_down=hs, hs_across=hs, deramp=1)
        S = W + 2 * hs
        m = len(g["idx"])
        ref = g["ref"].cpu().numpy().reshape(m * W, W)
        sec = g["sec"].cpu().numpy().reshape(m * S, S)
        i = np.arange(m, dtype=np.int64)
        z = np.zeros(m, np.int64)
        o = O.run_windows(ref, sec, i * W, z, i * S, z, cfg)
        for j, p in enumerate(g["idx"]):
            out[int(p)] = {k: v[j:j + 1] for k, v in o.items()}
    return out


@pytest.fixture(scope="module")
def groups():
    # 300 pairs: at least 21 in each of the 12 (window, half search) classes
    g = chip_pairs_uneven_grouped(n=300, seed=5, device=torch.device("cuda", 0))
    assert len(g) == 12 and min(len(v["idx"]) for v in g.values()) >= 16
    return g


@pytest.fixture(scope="module")
def oracle_out(groups):
    return _oracle(groups)


def test_uneven_batch_equals_oracle(groups, oracle_out):
    b = ChipPairBatch(groups, device=0, graph=True)
    b.run()
    r = b.results()
    o = oracle_out
    bad = []
    for i, oi in o.items():
        same = (np.array_equal(r["maxloc"][i], oi["maxloc"][0]) and np.array_equal(r["zoom_idx"][i], oi["zoom_idx"][0])
                and np.array_equal(r["offset"][i], oi["offset"][0].astype(np.float32)))
        if not same:
            bad.append((i, r["maxloc"][i], oi["maxloc"][0], r["zoom_idx"][i], oi["zoom_idx"][0]))
        assert abs(r["snr"][i] - oi["snr"][0]) <= 1e-3 * abs(oi["snr"][0])
    print(f"\n[uneven batch] pairs={len(o)} classes={len(groups)} rerun={b.rerun_count} mismatches={len(bad)}",
          flush=True)
    b.close()
    assert not bad, bad


def test_forced_fp64_rerun_equals_oracle(groups, oracle_out):
    """tau = inf sends every pair through the per-class FP64 re-run (own stream and
    context per class, overlapping): results equal the oracle's, twice in a row."""
    b = ChipPairBatch(groups, device=0, graph=True, tau=float("inf"))
    o = oracle_out
    for rep in range(2):
        b.run()
        r = b.results()
        assert b.rerun_count == len(o)
        for i, oi in o.items():
            assert np.array_equal(r["maxloc"][i], oi["maxloc"][0]), (rep, i)
            assert np.array_equal(r["zoom_idx"][i], oi["zoom_idx"][0]), (rep, i)
            assert np.array_equal(r["offset"][i], oi["offset"][0].astype(np.float32)), (rep, i)
            assert abs(r["snr"][i] - oi["snr"][0]) <= 1e-5 * abs(oi["snr"][0]), (rep, i)
    b.close()


def test_graph_replays_equal_eager(groups):
    eager = ChipPairBatch(groups, device=0, graph=False)
    eager.run()
    ref = eager.results()
    g = ChipPairBatch(groups, device=0, graph=True)
    x = torch.randn(512, 512, device="cuda")
    for rep in range(3):
        g.run()
        r = g.results()
        for k in ("offset", "snr", "cov", "maxloc", "zoom_idx"):
            assert np.array_equal(r[k], ref[k], equal_nan=True), (rep, k)
        eager.run()                      # eager engine work and torch work between replays
        (x @ x).sum().item()
    eager.close()
    g.close()


def test_shift_recovery_and_autotuned_schedule():
    grp = chip_pairs_uneven_grouped(n=96, seed=9, windows=(64, 128), ranges=(8, 16),
                                    device=torch.device("cuda", 0))
    base = ChipPairBatch(grp, device=0, graph=False)
    base.run()
    r0 = base.results()
    tuner = Autotuner(device=0, reps=1, sample=64, radix_sample=16)
    tuned = ChipPairBatch(grp, device=0, graph=True, tuner=tuner)
    tuned.run()
    r1 = tuned.results()
    for k in tuner.cache:
        s = tuner.cache[k]
        assert s.chunk > 0 and np.isfinite(s.ms_per_window)
        for n, rad in s.radices.items():
            assert int(np.prod(rad)) == int(n)
    assert np.array_equal(r0["maxloc"], r1["maxloc"]) and np.array_equal(r0["zoom_idx"], r1["zoom_idx"])
    assert np.array_equal(r0["offset"], r1["offset"])
    truth = np.concatenate([g["shift"] for _, g in sorted(grp.items())])
    idx = np.concatenate([g["idx"] for _, g in sorted(grp.items())])
    err = np.abs(r1["offset"][idx] - truth)
    print(f"\n[shift recovery] n={len(idx)} median_err={np.median(err):.4f} p95={np.percentile(err, 95):.4f} "
          f"schedules={ {k: (v.mode, v.chunk) for k, v in tuner.cache.items()} }", flush=True)
    assert np.median(err) < 0.05
    base.close()
    tuned.close()


@pytest.mark.parametrize("n,cap", [(1, 1), (37, 8), (700, 11), (2500, 40), (3000, 3000)])
def test_select_near_ties_matches_numpy(n, cap):
    """The device selection of the in-step FP64 re-runs: near ties first (NaN margins
    count as near), in window order, then the first windows that are not."""
    from isce2_amd.engine import Engine
    rng = np.random.default_rng(n)
    mc = rng.uniform(0, 1e-5, n).astype(np.float32)
    mf = rng.uniform(0, 1e-5, n).astype(np.float32)
    mc[rng.uniform(size=n) < 0.02] = np.nan
    e = Engine(0)
    try:
        for tau in (2e-6, 0.0, float("inf")):
            got = e.select_near_ties(torch.from_numpy(mc).cuda(), torch.from_numpy(mf).cuda(), cap, tau, tau)
            near = ~(mc >= tau) | ~(mf >= tau)
            want = np.concatenate([np.nonzero(near)[0], np.nonzero(~near)[0]])[:cap]
            assert np.array_equal(got.cpu().numpy(), want), tau
    finally:
        e.close()


def test_captured_context_refuses_to_grow(groups):
    """After a step is captured, a larger call on a captured context fails on the host
    (its buffers are held by the graph) instead of moving them."""
    b = ChipPairBatch(groups, device=0, graph=True)
    b.run()
    c = max(b.classes, key=lambda k: k.W)
    st = [x.contiguous() for x in c.starts]
    with pytest.raises(Exception, match="captured"):
        c.rerun_eng.run_windows(c.ref, c.sec, *st, mode=MODE_STAGED, debug=True, nosync=True)
    b.close()


def test_concurrent_classes_equal_sequential(groups):
    """Classes launched on side streams of their own give the sequential step's results."""
    a = ChipPairBatch(groups, device=0, graph=True, concurrent=True)
    b = ChipPairBatch(groups, device=0, graph=True, concurrent=False)
    for _ in range(2):
        a.run()
        b.run()
        ra, rb = a.results(), b.results()
        for k in ra:
            assert np.array_equal(ra[k], rb[k], equal_nan=True), k
    a.close()
    b.close()
