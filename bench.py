#!/usr/bin/env python3
"""Throughput benchmark: dense offset tracking on one MI355X.

Workload (one "step" = one pass over the whole offset field): a complex64
8192x4096 speckle pair whose secondary is warped by a smooth offset field
(2nd-order polynomial, max 8 px, plus N(0, 0.05) px per 256x256 tile), window
64x64, half search 20, skip 32 (252 x 124 = 31,248 windows), raw oversampling
2, 16x16 zoom surface oversampled 16x, derampMethod 1. Images are resident in
HBM before the timed region. Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# model FLOPs per window at the reference's transform sizes (5 n log2 n per
# complex FFT, 2.5 n log2 n per real FFT, plus pointwise terms), split over the
# engine's kernels; they sum to the 12.364 MFLOP/window of the whole pipeline
MFLOP_PER_WINDOW = 12.364
KERNEL_MFLOP = {"k_coarse": 1.235, "k_oversample_ref": 1.434, "k_oversample_search": 1.817, "k_fine": 7.878}
FP32_PEAK_TFLOPS = 157.3
# HBM bytes per window of each fused kernel, from rocprofv3 FETCH_SIZE + WRITE_SIZE (KiB per dispatch of a
# 31248-window launch) scaled by a 1 GiB device copy's counts (FETCH x2, WRITE x1); re-measure after kernel changes
HBM_BYTES_PER_WINDOW = {k: v * 1024 / 31248 for k, v in {
    "k_fine": 2266579.0 * 2 + 698.5, "k_coarse": 652045.4 * 2 + 1466.8,
    "k_oversample_ref": 453804.7 * 2 + 2500350.4, "k_oversample_search": 560834.7 * 2 + 2531088.0}.items()}
HBM_PEAK_GBS = 8000.0


def bench_offset_qa(gpu_out, ref_d, ref_a, W, skip):
    """Host quality control of the field's offsets: SNR cull + offoutliers, and the
    iterative affine fit (L2 on the whole field; L1 on a 4096-offset subsample)."""
    from isce2_amd.offset_qa import OffsetField, fitoff, offoutliers
    off = gpu_out["offset"]
    f = OffsetField(ref_a + W // 2, off[:, 1], ref_d + W // 2, off[:, 0], gpu_out["snr"])
    t0 = time.perf_counter()
    oo = offoutliers(f, distance=2.0, snr_threshold=2.0)
    t1 = time.perf_counter()
    l2 = fitoff(f, nsig=1.5, maxrms=0.08, minpoint=50, l1norm=False)
    t2 = time.perf_counter()
    sub = f.take(np.random.default_rng(3).choice(len(f), size=min(4096, len(f)), replace=False))
    l1 = fitoff(sub, l1norm=True)
    t3 = time.perf_counter()
    return {"offsets": len(f), "offoutliers_ms": round((t1 - t0) * 1e3, 2), "offoutliers_kept": int(len(oo["index"])),
            "fitoff_l2_ms": round((t2 - t1) * 1e3, 2), "fitoff_l2_status": l2["status"],
            "fitoff_l2_kept": int(len(l2["index"])), "fitoff_l1_4096_ms": round((t3 - t2) * 1e3, 2),
            "fitoff_l1_status": l1["status"], "note": "host numpy/scipy; the field is warped, so the affine fit "
            "culls toward the best affine subset"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", default="auto", choices=["auto", "fused", "staged"])
    ap.add_argument("--deramp", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the staged-pipeline comparison")
    ap.add_argument("--cpu-windows", type=int, default=12288)
    ap.add_argument("--no-deramp0", action="store_true", help="skip the derampMethod 0 leg")
    ap.add_argument("--oracle-windows", type=int, default=384)
    ap.add_argument("--no-host", action="store_true", help="skip the host-resident-input (PCIe-inclusive) leg")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven chip-pair batch leg")
    ap.add_argument("--uneven-pairs", type=int, default=8192)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-jobs", type=int, default=60)
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate, jobs/s")
    ap.add_argument("--no-stress", action="store_true", help="skip the precision-stress leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (scale reference for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this engine runs on one GPU per process; bench.py measures N=1 only "
                                   "(windows are independent: each GPU would take a band of window rows plus its image halo)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from isce2_amd.engine import MODE_AUTO, MODE_FUSED, MODE_STAGED, Engine, WindowParams
    from isce2_amd.synth import speckle_pair_warp

    dev = torch.device("cuda", 0)
    if args.copy_calibration:
        n = 1 << 28
        x = torch.ones(n, dtype=torch.float32, device=dev)
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize(dev)
        print(json.dumps({"copy_bytes_read_per_dispatch": 4 * n, "copy_bytes_written_per_dispatch": 4 * n}))
        return
    H, Wd, W, HS, SKIP = 8192, 4096, 64, 20, 32
    t0 = time.time()
    ref_np, sec_np, dd, da = speckle_pair_warp(H, Wd, max_amp=8.0, jitter=0.05, tile=256, seed=0, device=dev)
    gen_s = time.time() - t0
    nd = (H - 2 * HS - W) // SKIP
    na = (Wd - 2 * HS - W) // SKIP
    rows, cols = np.meshgrid(np.arange(nd), np.arange(na), indexing="ij")
    ref_d = (HS + rows * SKIP).ravel()
    ref_a = (HS + cols * SKIP).ravel()
    sec_d, sec_a = ref_d - HS, ref_a - HS
    n = nd * na
    ref = torch.from_numpy(ref_np).to(dev)
    sec = torch.from_numpy(sec_np).to(dev)
    starts = [torch.as_tensor(a.astype(np.int32)).to(dev) for a in (ref_d, ref_a, sec_d, sec_a)]

    eng = Engine(0)
    params = WindowParams(window_h=W, window_w=W, hs_down=HS, hs_across=HS, deramp=args.deramp)
    sizes = eng.setup(params)
    mode = {"auto": MODE_AUTO, "fused": MODE_FUSED, "staged": MODE_STAGED}[args.mode]
    out = None
    for _ in range(args.warmup):
        out = eng.run_windows(ref, sec, *starts, mode=mode, out=out)
    torch.cuda.synchronize(dev)
    # one untimed pass with every kernel timed: the per-kernel table and the dominant kernel
    eng.profile(True)
    out = eng.run_windows(ref, sec, *starts, mode=mode, out=out)
    all_prof = eng.profile_report()
    dom_name = max(all_prof, key=lambda k: all_prof[k]["ms"]) if all_prof else None
    # timed region: HIP events around the dominant kernel only (fewest records in the region)
    eng.profile(True)
    eng.profile_only(dom_name)
    eng.precision_stats()
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = eng.run_windows(ref, sec, *starts, mode=mode, out=out)
    torch.cuda.synchronize(dev)
    elapsed = time.perf_counter() - t0
    prof = eng.profile_report()
    eng.profile(False)
    eng.profile_only(None)
    checked, rerun = eng.precision_stats()
    wps = n * args.steps / elapsed
    ms_step = 1000 * elapsed / args.steps
    gpu_out = {k: v.cpu().numpy() for k, v in out.items()}

    # host-resident variant: pinned host images -> HBM -> field -> host outputs, per pass;
    # row bands upload on a copy stream while earlier bands of windows run
    pcie = None
    if not args.no_host:
        pcie = bench_host_inputs(eng, ref_np, sec_np, (ref_d, ref_a, sec_d, sec_a), mode, gpu_out, n)

    # roofline of the dominant kernel, from the events recorded inside the timed region
    roof = None
    kernels = {}
    for k, v in all_prof.items():
        avg_ms = v["ms"] / max(1, v["launches"])
        per_launch = v["windows"] / max(1, v["launches"])
        kernels[k] = dict(avg_ms=round(avg_ms, 4), launches=v["launches"], windows_per_launch=round(per_launch, 1),
                          model_tflops=round(KERNEL_MFLOP.get(k, 0) * 1e6 * per_launch / (avg_ms * 1e-3) / 1e12, 2))
    if prof and dom_name in prof:
        v = prof[dom_name]
        avg_s = v["ms"] / v["launches"] * 1e-3
        wpl = v["windows"] / v["launches"]
        achieved = KERNEL_MFLOP[dom_name] * 1e6 * wpl / avg_s / 1e12
        traffic = HBM_BYTES_PER_WINDOW.get(dom_name)
        roof = {"bound": "valu", "kernel": dom_name, "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                "traffic": round(traffic * wpl) if traffic else None,
                "traffic_note": "HBM bytes per average launch: rocprofv3 FETCH_SIZE/WRITE_SIZE per window "
                                "(copy-scaled) x windows per launch",
                "per_unit": f"{KERNEL_MFLOP[dom_name]} MFLOP/window (reference-size FFT model)",
                "windows_per_launch": round(wpl, 1), "avg_launch_ms": round(v["ms"] / v["launches"], 4),
                "timed_launches": v["launches"]}

    # plain (staged, one batched kernel per stage) for comparison
    plain = None
    if not args.no_plain and args.mode != "staged":
        torch.cuda.synchronize(dev)
        o2 = eng.run_windows(ref, sec, *starts, mode=MODE_STAGED, chunk=1024)
        torch.cuda.synchronize(dev)
        t1 = time.perf_counter()
        o2 = eng.run_windows(ref, sec, *starts, mode=MODE_STAGED, chunk=1024)
        torch.cuda.synchronize(dev)
        t_plain = time.perf_counter() - t1
        o2n = {k: v.cpu().numpy() for k, v in o2.items()}
        plain = {"windows_per_s": round(n / t_plain, 1), "ms_per_step": round(1000 * t_plain, 2),
                 "speedup_fused_over_staged": round(wps / (n / t_plain), 2),
                 "max_abs_offset_diff_vs_fused": float(np.abs(o2n["offset"] - gpu_out["offset"]).max())}
        # the reference controller's schedule: staged stages over chunks of 64 windows
        # (1 x 64, TopsProc/runDenseOffsets.py:234-235) with a stream sync per chunk
        torch.cuda.synchronize(dev)
        t1 = time.perf_counter()
        for c0 in range(0, n, 64):
            sl = [t[c0:c0 + 64] for t in starts]
            eng.run_windows(ref, sec, *sl, mode=MODE_STAGED, chunk=64)
            torch.cuda.current_stream(dev).synchronize()
        t_ref = time.perf_counter() - t1
        plain["reference_chunking_windows_per_s"] = round(n / t_ref, 1)
        plain["speedup_fused_over_reference_chunking"] = round(wps / (n / t_ref), 2)

    # CPU baseline: the C++/OpenMP restatement of mroipac/ampcor (ampcor.F) on a bounded
    # random sample of the same windows; the FP64 oracle checks the GPU offsets on a subset
    cpu = None
    accuracy = None
    config1 = None
    if not args.no_cpu:
        from oracle import ampcor_oracle as O
        rng = np.random.default_rng(7)
        cpu = bench_cpu_baseline(ref_np, sec_np, (ref_d, ref_a, sec_d, sec_a), W, HS, args.cpu_windows, wps, rng)
        config1 = bench_config1(dev, cpu["cores"])
        sel = np.sort(rng.choice(n, size=min(args.oracle_windows, n), replace=False))
        cfg = O.AmpcorConfig(window_h=W, window_w=W, hs_down=HS, hs_across=HS, deramp=args.deramp)
        o = O.run_windows(ref_np, sec_np, ref_d[sel], ref_a[sel], sec_d[sel], sec_a[sel], cfg)
        g_off = gpu_out["offset"][sel]
        err = np.abs(g_off - o["offset"]).max(axis=1)
        accuracy = {"windows_checked_vs_fp64_oracle": int(len(sel)), "max_offset_err_px": float(err.max()),
                    "p99_offset_err_px": float(np.percentile(err, 99)),
                    "windows_over_0.01px": int((err > 0.01).sum()),
                    "snr_max_rel_err": float(np.max(np.abs(gpu_out["snr"][sel] - o["snr"]) / np.abs(o["snr"])))}
        cr = ref_d[sel] + W // 2
        cc = ref_a[sel] + W // 2
        truth = np.stack([dd[cr, cc], da[cr, cc]], axis=1)
        accuracy["median_abs_err_vs_true_field_px"] = float(np.median(np.abs(g_off - truth)))

    deramp0 = None
    if not args.no_deramp0 and args.deramp != 0 and args.mode == "auto":
        deramp0 = bench_deramp0(dev, ref, sec, starts, args.steps)
    qa = bench_offset_qa(gpu_out, ref_d, ref_a, W, SKIP)
    uneven = None if args.no_uneven else bench_uneven(dev, args.uneven_pairs, max(1, args.steps // 2))
    serving = None if args.no_serve else bench_serving(dev, ref_np, sec_np, args.serve_jobs, args.serve_rate)
    stress = None if args.no_stress else bench_precision_stress(dev)

    line = {
        "metric": "windows/sec @64x64 ±20px, offset err<0.01px vs FP64; serve p50/p99 latency",
        "value": round(wps, 1), "unit": "windows/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (FP64 window sums)", "data": "synthetic",
        "config": {"workload": "plain single-GPU offset field: complex64 8192x4096 speckle pair, polynomial warp "
                               "+-8 px, window 64x64, half search 20, skip 32, 31248 windows, raw ovs 2, "
                               "surface ovs 16, derampMethod %d" % args.deramp,
                   "windows": n, "mode": args.mode, "engine_path": "fused" if sizes["fused"] and args.mode != "staged" else "staged"},
        "precision": {"mode": "FP32 kernels, FP64 window sums; FP64 re-run of windows whose argmax margin < 2e-6",
                      "windows_rerun_fp64_per_step": rerun // max(1, args.steps),
                      "rerun_fraction": round(rerun / max(1, checked), 5)},
        "roofline": roof, "cpu_baseline": cpu, "kernels_untimed_pass": kernels, "plain_staged": plain, "accuracy": accuracy,
        "plain_gpu_host_to_host": pcie, "deramp0": deramp0, "config1_cpu_case": config1,
        "uneven_batch": uneven, "serving": serving, "precision_stress": stress,
        "offset_qa": qa,
        "scene_gen_s": round(gen_s, 1),
    }
    print(json.dumps(line), flush=True)


def bench_host_inputs(eng, ref_np, sec_np, host_starts, mode, gpu_out, n):
    """The plain-GPU figure of the measurement protocol: pinned host images -> HBM -> field -> host
    outputs (PCIe-inclusive); one warm-up, then the median of 5 passes."""
    import torch
    dev = torch.device("cuda", eng.device)
    h_ref = torch.from_numpy(ref_np).pin_memory()
    h_sec = torch.from_numpy(sec_np).pin_memory()
    host_out = eng.run_windows_from_host(h_ref, h_sec, *host_starts, mode=mode)
    torch.cuda.synchronize(dev)
    times = []
    for _ in range(5):
        t1 = time.perf_counter()
        host_out = eng.run_windows_from_host(h_ref, h_sec, *host_starts, mode=mode)
        times.append(time.perf_counter() - t1)
    t_host = float(np.median(times))
    return {"windows_per_s": round(n / t_host, 1), "ms_per_field": round(1e3 * t_host, 2),
            "timing": "median of 5 after one warm-up",
            "max_abs_offset_diff_vs_device_resident": float(np.abs(host_out["offset"] - gpu_out["offset"]).max()),
            "note": "pinned host images uploaded to HBM in 8 row bands on a copy stream overlapping the "
                    "earlier bands' windows, outputs copied back to host memory"}


def host_cpu_info():
    """What the CPU baseline runs on: logical CPUs, this process's affinity, the cgroup CPU
    quota (the box's share for a one-GPU job) and the CPU model."""
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)), "cgroup_quota_cpus": None,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "model": None}
    try:
        q, p = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            info["cgroup_quota_cpus"] = round(int(q) / int(p), 2)
    except (OSError, ValueError):
        pass
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    if info["model"] is None:
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    info["model"] = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
    return info


def bench_cpu_baseline(ref_np, sec_np, starts, W, HS, n_windows, wps_gpu, rng):
    """mroipac/ampcor restated in C++/OpenMP (oracle/ampcor_cpu.cpp) on a random sample of the
    field's windows, at every CPU this process may use and at the cgroup's CPU share when that is
    smaller; the faster of the two is the reported value."""
    from oracle import build_cpu
    ref_d, ref_a, sec_d, sec_a = starts
    info = host_cpu_info()
    counts = {info["affinity_cpus"]}
    if info["cgroup_quota_cpus"]:
        counts.add(max(1, min(info["affinity_cpus"], int(info["cgroup_quota_cpus"]))))
    sel = np.sort(rng.choice(len(ref_d), size=min(n_windows, len(ref_d)), replace=False))
    runs = {}
    for threads in sorted(counts):
        build_cpu.run(ref_np, sec_np, ref_d[:64], ref_a[:64], sec_d[:64], sec_a[:64], W, HS, threads)   # warm
        tc = time.perf_counter()
        _, _, _, acc = build_cpu.run(ref_np, sec_np, ref_d[sel], ref_a[sel], sec_d[sel], sec_a[sel], W, HS, threads)
        runs[threads] = (len(sel) / (time.perf_counter() - tc), acc)
    best = max(runs, key=lambda t: runs[t][0])
    wps, acc = runs[best]
    return {"value": round(wps, 2), "unit": "windows/s", "cores": best, "kind": "port",
            "sample": f"{len(sel)} random windows of the 8192x4096 field (window 64, half search 20) through "
                      f"a C++/OpenMP restatement of mroipac/ampcor (ampcor.F), {best} threads, "
                      f"{len(sel) / wps:.1f} s; {acc} windows passed its SNR/edge gates",
            "rates_by_threads": {str(t): round(v[0], 1) for t, v in runs.items()},
            "host": info, "gpu_speedup": round(wps_gpu / wps, 1)}


def bench_config1(dev, threads):
    """The CPU ampcor case: complex64 1024x1024 speckle pair shifted by (+3.25, -1.75) px with 10 dB
    noise, window 64, half search 16, skip 64, a 14x14 grid from (16, 16). CPU restatement: median
    of 3; GPU: device-resident and pinned-host-to-host-outputs, each the median of 5 after a warm-up."""
    import torch
    from isce2_amd.engine import Engine, WindowParams
    from isce2_amd.synth import speckle_pair_shift
    from oracle import ampcor_oracle as O
    from oracle import build_cpu
    ref, sec = speckle_pair_shift(1024, 1024, shift=(3.25, -1.75), snr_db=10.0, seed=0)
    rd, ra, sd, sa, _, _ = O.start_pixels(14, 14, 16, 16, 64, 64, 0, 0, 16, 16)
    n = len(rd)
    t = []
    for _ in range(3):
        tc = time.perf_counter()
        c_off, _, _, acc = build_cpu.run(ref, sec, rd, ra, sd, sa, 64, 16, threads)
        t.append(time.perf_counter() - tc)
    cpu_wps = n / float(np.median(t))
    e = Engine(dev.index or 0)
    e.setup(WindowParams(hs_down=16, hs_across=16))
    st = [torch.as_tensor(np.asarray(s, np.int32), device=dev) for s in (rd, ra, sd, sa)]
    rt, stt = torch.from_numpy(ref).to(dev), torch.from_numpy(sec).to(dev)
    g = e.run_windows(rt, stt, *st)
    torch.cuda.synchronize(dev)
    dt = []
    for _ in range(5):
        t1 = time.perf_counter()
        g = e.run_windows(rt, stt, *st)
        torch.cuda.synchronize(dev)
        dt.append(time.perf_counter() - t1)
    h_ref, h_sec = torch.from_numpy(ref).pin_memory(), torch.from_numpy(sec).pin_memory()
    e.run_windows_from_host(h_ref, h_sec, rd, ra, sd, sa)
    ht = []
    for _ in range(5):
        t1 = time.perf_counter()
        h = e.run_windows_from_host(h_ref, h_sec, rd, ra, sd, sa)
        ht.append(time.perf_counter() - t1)
    g_off = g["offset"].cpu().numpy()
    e.close()
    truth = np.array([3.25, -1.75])
    return {"windows": n, "cpu_windows_per_s": round(cpu_wps, 1), "cpu_threads": threads,
            "cpu_median_abs_err_vs_true_shift_px": round(float(np.median(np.abs(c_off - truth))), 4),
            "gpu_device_resident_windows_per_s": round(n / float(np.median(dt)), 1),
            "gpu_host_to_host_windows_per_s": round(n / float(np.median(ht)), 1),
            "gpu_median_abs_err_vs_true_shift_px": round(float(np.median(np.abs(g_off - truth))), 4),
            "host_equals_device_offsets": bool(np.array_equal(h["offset"], g_off)),
            "note": "196 windows are far too few to fill the GPU: the GPU figures here are launch/latency bound"}


def bench_deramp0(dev, ref, sec, starts, steps):
    """The same field with derampMethod 0 (amplitude chips, the dense-offsets command line's default)."""
    import torch
    from isce2_amd.engine import Engine, WindowParams
    e = Engine(dev.index or 0)
    e.setup(WindowParams(deramp=0))
    out = e.run_windows(ref, sec, *starts)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(steps):
        out = e.run_windows(ref, sec, *starts, out=out)
    torch.cuda.synchronize(dev)
    el = (time.perf_counter() - t0) / steps
    e.close()
    return {"windows_per_s": round(starts[0].numel() / el, 1), "ms_per_step": round(el * 1e3, 3), "steps": steps}


# model MFLOP per chip pair by (window, half search) at the reference's transform sizes
PAIR_MFLOP = {(32, 8): 7.05, (32, 16): 7.25, (32, 32): 7.88, (64, 8): 11.82, (64, 16): 12.16, (64, 32): 13.08,
              (128, 8): 31.90, (128, 16): 32.54, (128, 32): 34.08, (256, 8): 118.23, (256, 16): 119.52,
              (256, 32): 122.38}


def bench_uneven(dev, n_pairs, steps):
    """Uneven chip pairs (window 32/64/128/256, half search 8/16/32, sub-pixel shifts up to 0.8 of
    the search): one autotuned, graph-captured multi-class step per pass."""
    import torch
    from isce2_amd.autotune import Autotuner
    from isce2_amd.batch import ChipPairBatch
    from isce2_amd.synth import chip_pairs_uneven_grouped
    t0 = time.time()
    groups = chip_pairs_uneven_grouped(n=n_pairs, seed=0, device=dev)
    gen_s = time.time() - t0
    t0 = time.time()
    tuner = Autotuner(device=dev.index or 0, reps=2, sample=1024, radix_sample=64)
    b = ChipPairBatch(groups, device=dev.index or 0, graph=True, tuner=tuner)
    b.run()                                   # capture + first replay
    tune_s = time.time() - t0
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run()
    el = (time.perf_counter() - t0) / steps
    r = b.results()
    truth = np.empty((b.n, 2))
    mflop = 0.0
    for k, g in groups.items():
        truth[g["idx"]] = g["shift"]
        mflop += PAIR_MFLOP.get(k, 0.0) * len(g["idx"])
    err = np.abs(r["offset"] - truth).max(axis=1)
    res = {"pairs": b.n, "classes": len(groups), "pairs_per_s": round(b.n / el, 1), "ms_per_step": round(el * 1e3, 2),
           "model_tflops": round(mflop * 1e6 / el / 1e12, 2), "fp64_rerun_pairs": b.rerun_count,
           "schedules": {k: [v.mode, v.chunk, {int(n): r for n, r in v.radices.items()}] for k, v in tuner.cache.items()},
           "median_err_vs_true_shift_px": round(float(np.median(err)), 4), "setup_tune_capture_s": round(tune_s, 1),
           "generate_s": round(gen_s, 1)}
    b.close()
    return res


def bench_serving(dev, ref_np, sec_np, n_jobs, rate, seed=5, oracle_per_job=3):
    """Poisson arrivals of offset-field jobs over crops of the 8192x4096 warped pair: log-uniform
    sizes 512x512..8192x4096, window/half search from the uneven grid, skip = window/2."""
    import torch
    from isce2_amd.autotune import Autotuner
    from isce2_amd.serving import OffsetJob, OffsetServer
    H, Wd = ref_np.shape
    # the server's 200 GB arena plus its class engines take most of the 288 GB:
    # hand the blocks earlier legs left in torch's cache back to the device first
    torch.cuda.synchronize(dev)
    torch.cuda.empty_cache()
    pool_ref = torch.from_numpy(ref_np).pin_memory()
    pool_sec = torch.from_numpy(sec_np).pin_memory()
    rng = np.random.default_rng(seed)

    def make(i):
        s = rng.uniform()
        h, w = int(round(512 * 16 ** s)), int(round(512 * 8 ** s))
        W = int(rng.choice([32, 64, 128, 256]))
        hs = int(rng.choice([8, 16, 32]))
        if W + 2 * hs + 2 * hs > min(h, w):
            W = 64
        r0 = int(rng.integers(0, H - h + 1))
        return OffsetJob(ref=pool_ref[r0:r0 + h], sec=pool_sec[r0:r0 + h], height=h, width=w, window=W,
                         half_search=hs, skip_down=W // 2, skip_across=W // 2, job_id=i)

    srv = OffsetServer(device=dev.index or 0, hbm_budget=200e9, pitch=Wd,
                       tuner=Autotuner(device=dev.index or 0, reps=2, sample=1024, radix_sample=64))
    srv.start()
    # warm-up: one small job per window class (engine contexts, plans, tuning)
    warm = [srv.submit(OffsetJob(ref=pool_ref[:600], sec=pool_sec[:600], height=600, width=600, window=W,
                                 half_search=hs, skip_down=W // 2, skip_across=W // 2))
            for W in (32, 64, 128, 256) for hs in (8, 16, 32) if W + 4 * hs <= 600]
    for f in warm:
        f.result()
    gaps = rng.exponential(1.0 / rate, size=n_jobs)
    jobs = [make(i) for i in range(n_jobs)]
    t_start = time.perf_counter()
    futs, t_next = [], t_start
    for j, gap in zip(jobs, gaps):
        t_next += gap
        dt = t_next - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        futs.append(srv.submit(j))
    res = [f.result() for f in futs]
    t_end = time.perf_counter()
    srv.stop()
    lat = np.array([r["latency_s"] for r in res])
    wins = sum(r["windows"] for r in res)
    # a fixed sample of every job's windows against the FP64 oracle
    from isce2_amd.serving import job_windows
    from oracle import ampcor_oracle as O
    errs, snr_rel, checked = [], [], 0
    orng = np.random.default_rng(seed + 1)
    for j, r in zip(jobs, res):
        nd, na, rd, ra, sd, sa = job_windows(j)
        k = np.sort(orng.choice(nd * na, size=min(oracle_per_job, nd * na), replace=False))
        img_r = j.ref.numpy()[:j.height, :j.width]
        img_s = j.sec.numpy()[:j.height, :j.width]
        cfg = O.AmpcorConfig(window_h=j.window, window_w=j.window, hs_down=j.half_search,
                             hs_across=j.half_search, deramp=j.deramp)
        o = O.run_windows(img_r, img_s, rd[k], ra[k], sd[k], sa[k], cfg)
        errs.append(np.abs(r["offset"].reshape(-1, 2)[k] - o["offset"]).max(axis=1))
        snr_rel.append(np.abs(r["snr"].reshape(-1)[k] - o["snr"]) / np.abs(o["snr"]))
        checked += len(k)
    errs = np.concatenate(errs)
    snr_rel = np.concatenate(snr_rel)
    return {"jobs": n_jobs, "arrival_rate_per_s": rate, "windows": int(wins),
            "oracle_check": {"jobs": len(res), "windows_per_job": oracle_per_job, "windows": checked,
                             "max_offset_err_px": float(errs.max()), "windows_over_0.01px": int((errs > 0.01).sum()),
                             "snr_max_rel_err": float(snr_rel.max())},
            "windows_per_s": round(wins / (t_end - t_start), 1),
            "p50_latency_ms": round(float(np.percentile(lat, 50)) * 1e3, 1),
            "p99_latency_ms": round(float(np.percentile(lat, 99)) * 1e3, 1),
            "max_latency_ms": round(float(lat.max()) * 1e3, 1),
            "mean_queue_ms": round(float(np.mean([r["queue_s"] for r in res])) * 1e3, 2),
            "makespan_s": round(t_end - t_start, 2), "batches": srv.stats["batches"],
            "hbm_budget_gb": 200}


def bench_precision_stress(dev):
    """4096x4096 precision-stress pair (log-normal sigma 2 amplitude, 200 +60 dB points, 20% of
    windows at amplitude 1e-6, per-128x128-tile shifts up to +-4 px), a 32x32 grid of 64x64 windows
    (half search 20) centred in the tiles: error of every output against the FP64 oracle, FP32-only
    and default precision."""
    import torch
    from isce2_amd.engine import Engine, WindowParams
    from isce2_amd.synth import precision_stress_scene
    from oracle import ampcor_oracle as O
    ref, sec, shifts, tiny = precision_stress_scene(4096, seed=0, device=dev)
    ty, tx = np.meshgrid(np.arange(32), np.arange(32), indexing="ij")
    rd = (ty * 128 + 32).ravel()
    ra = (tx * 128 + 32).ravel()
    starts = (rd, ra, rd - 20, ra - 20)
    t0 = time.time()
    o = O.run_windows(ref, sec, *starts, O.AmpcorConfig())
    oracle_s = time.time() - t0
    e = Engine(dev.index or 0)
    e.setup(WindowParams())
    st = [torch.as_tensor(s_.astype(np.int32), device=dev) for s_ in starts]
    rt, stt = torch.from_numpy(ref).to(dev), torch.from_numpy(sec).to(dev)
    out = {"windows": len(rd), "tiny_windows": int(tiny.sum()), "oracle_s": round(oracle_s, 1)}
    for prec, name in ((0, "fp32_only"), (1, "default")):
        e.set_precision(prec)
        g = e.run_windows(rt, stt, *st, debug=True)
        torch.cuda.synchronize(dev)
        _, rerun = e.precision_stats()
        g = {k: v.cpu().numpy() for k, v in g.items()}
        err = np.abs(g["offset"] - o["offset"]).max(axis=1)
        snr_rel = np.abs(g["snr"] - o["snr"]) / np.abs(o["snr"])
        sent_o = np.all(o["cov"] == [99, 99, 0], axis=1)
        sent_g = np.all(g["cov"] == [99, 99, 0], axis=1)
        ok = ~sent_o & ~sent_g
        cov_rel = np.abs(g["cov"] - o["cov"]).max(axis=1) / np.abs(o["cov"]).max(axis=1)
        out[name] = {"offset_max_err_px": float(err.max()), "windows_over_0.01px": int((err > 0.01).sum()),
                     "peak_mismatches": int(np.any(g["maxloc"] != o["maxloc"], axis=1).sum()
                                            + np.any(g["zoom_idx"] != o["zoom_idx"], axis=1).sum()),
                     "snr_max_rel_err": float(snr_rel.max()),
                     "cov_max_rel_err": float(cov_rel[ok].max()) if ok.any() else 0.0,
                     "cov_sentinel_mismatches": int((sent_o != sent_g).sum()), "fp64_rerun_windows": int(rerun)}
    e.close()
    return out


if __name__ == "__main__":
    main()
