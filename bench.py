#!/usr/bin/env python3
"""Benchmark of the calibrate-and-geocode path on one MI355X.

One step = one pass of the path over one full synthetic IW scene whose inputs
are already resident in HBM:
  fused deburst + LUT calibration (+NESZ) + IW merge  →  backward RD terrain
  correction (FP64 Newton, bilinear, 10 m UTM grid, shadow/layover mask +
  cos_lia sidecars)  →  Refined Lee 7x7.
Workload: 3 subswaths x 9 bursts x 1500 lines x 21000 samples CInt16 (burst
overlap 50 lines), fBm DEM (Hurst 0.8, 0-3000 m) at 1 arcsec, output a 10 m
UTM grid of about 25000 (east) x 18000 (north) cells around the scene centre.
value = output grid pixels per second (Mpx/s).

Printed JSON line (rank 0): metric/value/unit, timing, the dominant kernel's
roofline (its duration measured with HIP events on its launch stream), and a
CPU baseline: the FP64 C++ oracle restatement of the reference timed on a
bounded sample of the same workload on the host cores.
"""
from __future__ import annotations

import argparse
import ctypes as C
import gc
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "geocoded output Mpx/s (cal+RD-TC+resample) & %HBM roofline; p50/p99 latency"

# Algorithmic FP64 operation counts of the reference algorithm per unit of
# work, counted from the oracle restatement (+,-,*,/,sqrt and each sin/cos as
# one operation):
F_ECEF = 22        # geodetic_to_ecef (geodesy.rs:38-55)
F_DEM = 20         # bilinear_on_grid + tile addressing (dem.rs:385-409)
F_ITER = 439       # one Newton iteration: Lagrange-8 pos/vel/accel (417) + f, f', δ (22)
F_TAIL = 40        # slant range, radar coordinates, bilinear gather
F_GEOM = 4 * F_ECEF + 4 * F_DEM + 40  # cardinal DEM/ECEF + facet normal, look, w, cos_lia

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["iw_full", "small_plain", "batch64", "serve", "fp32", "cog", "multilook",
                                                "grd", "slc", "dem"],
                    default="iw_full",
                    help="iw_full: the headline step; batch64: 64 AOI tiles batched vs one call per tile; "
                         "serve: Poisson AOI requests on 4 resident scenes, served vs one call per request; "
                         "fp32: the FP32 scene-centred variant's errors and speed against FP64; "
                         "cog: the COG export (pyramid + tile pack, and the whole file write) of the "
                         "full-size dB product; multilook: the pre-TC multilook of the full merged image; "
                         "grd: the ground-range projection of the full merged image; "
                         "slc: a full-size swath measurement TIFF read into HBM; "
                         "dem: SRTM tile mosaics and a GLO-30 GeoTIFF tile loaded into HBM")
    ap.add_argument("--requests", type=int, default=600, help="serve: requests per run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="target CPU-baseline work (all runs)")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--tc-segment", type=int, default=0, help="columns per lane segment (0 = automatic)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 1 GiB device-to-device copy first (reference kernel for HBM byte counters)")
    return ap.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def settle_gc() -> None:
    """Collect once and move every object created so far out of the cyclic
    garbage collector's view (gc.freeze): a full pass over the set-up's
    objects takes tens of ms and would otherwise land inside a timed step
    (timeit turns the collector off for the same reason)."""
    gc.collect()
    gc.freeze()


def recent_kernel_ms(L, family: int, n: int) -> list:
    """Durations of the last n launches of a kernel family (HIP events)."""
    buf = (C.c_double * max(1, min(n, 64)))()
    k = L.sardine_recent_kernel_ms(family, C.cast(buf, C.c_void_p), len(buf))
    return list(buf[:k]) or [float("nan")]


def tc_flops(img, need_geom: bool) -> float:
    hist = img.newton_iter_histogram
    n_conv = sum(hist)
    iters = sum(i * h for i, h in enumerate(hist))
    dem_ok = img.total_pixel_count - img.dem_missing_count
    f = dem_ok * (F_ECEF + F_DEM) + iters * F_ITER + n_conv * F_TAIL
    f += img.non_converged_count * 10 * F_ITER
    if need_geom:
        f += (img.valid_pixel_count + img.flat_masked_count + img.noise_masked_count) * F_GEOM
    return float(f)


def measured_valu(kernel: str, workload: str, launch_ms: float):
    """Hardware FP64 VALU use of `kernel` from the committed rocprofv3 --pmc
    summary of this bench command (SQ_INSTS_VALU_*_F64, SQ_INSTS_VALU_FLOPS_FP64,
    SQ_ACTIVE_INST_VALU, ...): the FP64 lane flops one launch issues, divided
    by this run's launch time, against the FP64 peak; plus the counter run's
    VALU-busy share and FP64 share of VALU instructions.  None without a
    summary for this workload."""
    f = ROOT / "profiles" / f"valu_counters_{workload}.json"
    if not f.exists():
        return None
    k = json.loads(f.read_text())["kernels"].get(kernel)
    if k is None:
        return None
    flops = k["fp64_lane_flops"]
    return {"fp64_lane_flops_per_launch": flops,
            "tflops": round(flops / (launch_ms / 1e3) / 1e12, 4),
            "frac": round(flops / (launch_ms / 1e3) / 1e12 / FP64_PEAK_TFLOPS, 5),
            "valu_busy": round(k["valu_busy"], 4),
            "f64_share_of_valu_instructions": round(k["f64_share_of_valu_instructions"], 4),
            "vgprs": k["vgprs"], "waves_per_simd": round(k["waves_resident_per_simd"], 2),
            "source": str(f.relative_to(ROOT))}


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE passes of this bench command, each
    scaled by what the counter reads for a copy kernel of known size), or
    None when no summary for this workload exists."""
    f = ROOT / "profiles" / f"{kernel}_{workload}_traffic.json"
    if not f.exists():
        return None, None
    d = json.loads(f.read_text())
    return int(d["traffic_bytes_per_launch"]), str(f.relative_to(ROOT))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path shards by output row "
              f"bands with no exchange, but only the single-GPU figure is measured here.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from sardine_amd import _native, abi, api, synthetic
    from sardine_amd.workload import DeviceScene

    dev = torch.device("cuda:0")
    L = _native.lib()
    if args.workload == "batch64":
        return bench_batch64(args, dev)
    if args.workload == "serve":
        return bench_serve(args, dev)
    if args.workload == "fp32":
        return bench_fp32(args, dev)
    if args.workload == "cog":
        return bench_cog(args, dev)
    if args.workload == "multilook":
        return bench_multilook(args, dev)
    if args.workload == "grd":
        return bench_grd(args, dev)
    if args.workload == "slc":
        return bench_slc(args, dev)
    if args.workload == "dem":
        return bench_dem(args, dev)
    L.sardine_set_tc_segment(args.tc_segment)
    if args.copy_calibration:
        copy_calibration(dev)
    stream = torch.cuda.current_stream(dev)

    t0 = time.time()
    if args.workload == "iw_full":
        sc = synthetic.scene_iw_full(1, out_rows=18000, out_cols=25000)
        workload = "iw_full_3x9x1500x21000_utm10m_25000x18000_bilinear_refinedlee7"
    else:
        sc = synthetic.scene_small_plain(0)
        workload = "small_plain_2048x4096_wgs84_2048x2048_bilinear"
    ds = DeviceScene.build(sc, dev)
    info = ds.merged_layout()
    cfg = ds.tc_config(compute_lia=args.workload == "iw_full")
    rows, cols, gt = api.output_grid(ds.scene, cfg)
    log(f"scene built in {time.time() - t0:.1f}s: merged {info.lines}x{info.samples}, grid {rows}x{cols}, "
        f"dem {sc.dem_spec.height}x{sc.dem_spec.width}")
    s0 = torch.empty((info.lines, info.samples), dtype=torch.float32, device=dev)
    nz = torch.empty_like(s0)
    out = torch.empty((rows, cols), dtype=torch.float32, device=dev)
    cl = torch.empty_like(out) if cfg.compute_lia else None
    mk = torch.empty((rows, cols), dtype=torch.uint8, device=dev) if cfg.compute_lia else None
    sp = torch.empty(rows * cols, dtype=torch.float32, device=dev)
    filt = api.SpeckleFilter.RefinedLee(1.0) if args.workload == "iw_full" else None

    def step():
        m = ds.prepare_merged(stream=stream, out=(s0, nz))
        img = ds.terrain_correction(m, cfg, stream=stream, out=(out, cl, mk))
        if filt is not None:
            filt.apply(img.data, img.cols, img.rows, stream=stream, out=sp)
        return img

    # the kernel timers (created here, before the warm-up, so that the GPU is
    # busy right up to the timed region) record HIP events on each kernel's
    # launch stream without waiting; they are read after the timed region
    L.sardine_set_profiling(1)
    for i in range(args.warmup):
        img = step()
        log(f"warmup {i}: valid {img.valid_pixel_count}/{img.total_pixel_count}, redo {img.chain_redo_rounds}, "
            f"newton iterations {img.newton_iter_histogram[:8]}")

    # timed steps: nothing waits on the host between steps
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    settle_gc()
    torch.cuda.synchronize()
    ev[0].record(stream)
    t_wall = time.perf_counter()
    for i in range(args.steps):
        img = step()
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    lat = [ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)]
    k_tc, k_seg, k_cal, k_sp = (recent_kernel_ms(L, f, args.steps) for f in (
        abi.KERNEL_TC, abi.KERNEL_TC_SEGMENT, abi.KERNEL_CALIBRATE, abi.KERNEL_SPECKLE))
    L.sardine_set_profiling(0)
    total_ms = ev[0].elapsed_time(ev[-1])
    ms_per_step = total_ms / args.steps
    px = rows * cols
    value = px / (ms_per_step / 1e3) / 1e6
    log(f"steps: {['%.1f' % x for x in lat]} ms (wall {wall:.2f}s); kernels tc {np.mean(k_tc):.2f} "
        f"(segment {np.mean(k_seg):.2f}) "
        f"cal {np.mean(k_cal):.2f} speckle {np.mean(k_sp) if filt else 0:.2f} ms")

    # dominant kernel roofline: the per-pixel chain kernel (every output
    # pixel's full computation runs in it once; the fix-up kernel only
    # replays pixels whose chain seed changed)
    tc_ms = float(np.mean(k_tc))
    seg_ms = float(np.mean(k_seg))
    flops = tc_flops(img, cfg.compute_lia)
    tflops = flops / (seg_ms / 1e3) / 1e12
    merged_bytes = info.lines * info.samples * 4
    dem_bytes = sc.dem_spec.height * sc.dem_spec.width * 4
    tc_bytes = px * (4 + (5 if cfg.compute_lia else 0)) + merged_bytes + dem_bytes
    cal_ms = float(np.mean(k_cal))
    cal_bytes = info.lines * info.samples * 12
    sp_ms = float(np.mean(k_sp)) if filt else None
    traffic, traffic_src = measured_traffic("tc_segment_kernel", workload)
    roofline = {
        "kernel": "tc_segment_kernel",
        "bound": "fp64_valu",
        "achieved": round(tflops, 4),
        "peak": FP64_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(tflops / FP64_PEAK_TFLOPS, 5),
        "traffic": traffic,
        "traffic_source": traffic_src,
        "flops_per_launch": flops,
        "launch_ms": round(seg_ms, 4),
        "tc_total_ms": round(tc_ms, 4),
        "note": "FP64 VALU-bound (no dense contraction on the path); peak is MI355X FP64 78.6 TFLOP/s; "
                "achieved counts the reference algorithm's flops (product-form Lagrange), hw_fp64 the FP64 "
                "lane flops the kernel issues (hardware counters)",
        "hw_fp64": measured_valu("tc_segment_kernel", workload, seg_ms),
    }
    other = {
        "calibrate_merge": {"ms": round(cal_ms, 4), "algorithmic_bytes": cal_bytes,
                            "GBps": round(cal_bytes / (cal_ms / 1e3) / 1e9, 1),
                            "hbm_frac": round(cal_bytes / (cal_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
        "tc_hbm": {"ms": round(tc_ms, 4), "algorithmic_bytes": tc_bytes,
                   "GBps": round(tc_bytes / (tc_ms / 1e3) / 1e9, 1)},
    }
    if filt:
        sb = px * 8
        other["refined_lee7"] = {"ms": round(sp_ms, 4), "algorithmic_bytes": sb,
                                 "GBps": round(sb / (sp_ms / 1e3) / 1e9, 1),
                                 "hbm_frac": round(sb / (sp_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                                 "hw_fp64": measured_valu("refined_lee_walk_kernel", workload, sp_ms)}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sc, ds, info, cfg, rows, cols, s0, nz, filt, args.cpu_seconds)

    srt = sorted(lat)
    p50 = srt[len(srt) // 2]
    p99 = srt[min(len(srt) - 1, int(round(0.99 * (len(srt) - 1))))]
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpx/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload, "output_grid": [rows, cols], "merged": [info.lines, info.samples],
                   "parallelism": "single_gpu"},
        "latency_ms": {"p50": round(p50, 3), "p99": round(p99, 3)},
        "valid_pixels": img.valid_pixel_count,
        "roofline": roofline,
        "stages": other,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def host_cpu_share():
    """Threads the CPU baseline runs on, and a description of the host.

    All CPUs this process may run on (its affinity set), lowered by the job's
    thread budget when one is set (OMP_NUM_THREADS: a one-GPU job's share of
    a shared host) and by SARDINE_CPU_THREADS; either variable can only lower
    the count."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    cores = affinity
    for var in ("OMP_NUM_THREADS", "SARDINE_CPU_THREADS"):
        v = os.environ.get(var, "")
        if v.strip().isdigit() and int(v) > 0:
            cores = min(cores, int(v))
    model = "unknown"
    try:
        for ln in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return cores, {"cpu_model": model, "host_logical_cpus": os.cpu_count(), "affinity_cpus": affinity}


def cpu_baseline(sc, ds, info, cfg, rows, cols, s0, nz, filt, target_s, runs=3):
    """Oracle (FP64 C++ restatement of the reference, 64-row chunks on every
    host thread of the job) on a bounded sample of the same workload, the
    median of `runs` timings of each stage."""
    import numpy as np
    import torch

    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    from sardine_amd import abi

    cores, host = host_cpu_share()
    t = time.time()
    s0h = s0.cpu().numpy()
    log(f"cpu baseline: merged copied to host in {time.time() - t:.1f}s, {cores} threads on {host}")
    keep = []
    scene_c = sc.scene_c(keep)
    meta = dict(lines=info.lines, samples=info.samples, azimuth_start_time_us=info.azimuth_start_time_us,
                near_slant_range_time_s=info.near_slant_range_time_s, range_pixel_spacing_m=info.range_pixel_spacing_m)
    dem = sc.dem()

    def tc_sample(r0, nr):
        c = sc.tc_config(compute_lia=int(cfg.compute_lia), aoi_row0=r0, aoi_col0=0, aoi_rows=nr, aoi_cols=cols)
        t1 = time.perf_counter()
        res = O.terrain_correction(s0h, None, meta, scene_c, dem, sc.dem_spec, c, threads=cores)
        return time.perf_counter() - t1, res

    r_mid = rows // 2
    probe_rows = 64 * cores  # one 64-row chunk per thread, as the full run uses them
    dt, _ = tc_sample(r_mid - probe_rows // 2, probe_rows)
    per_row = dt / probe_rows
    n_tc = int(max(64 * cores, min(rows // 2, 0.6 * target_s / runs / max(per_row, 1e-6))))
    tc_runs = []
    for _ in range(runs):
        dt_tc, res = tc_sample(max(0, r_mid - n_tc // 2), n_tc)
        tc_runs.append(dt_tc)
    tc_per_px = float(np.median(tc_runs)) / (n_tc * cols)
    # calibration on the first merged lines of swath 0 (raw DN copied back)
    sw = sc.swaths[0]
    n_cal = 4096
    raw = ds.raws[0][:n_cal].cpu().numpy()
    fv = int(max(sw.first_valid))
    lv = int(min(sw.last_valid))
    deb = np.ascontiguousarray(raw[:, fv:lv])
    cal_runs = []
    for _ in range(runs):
        t1 = time.perf_counter()
        st, *_ = O.apply_calibration(deb, fv, sw.cal.as_c(), sw.noise.as_c(), sw.noise_az.as_c(), 0, threads=cores)
        cal_runs.append(time.perf_counter() - t1)
    cal_per_px = float(np.median(cal_runs)) / deb.shape[0] / deb.shape[1]
    sp_per_px = 0.0
    sp_runs = []
    if filt is not None:
        for _ in range(runs):
            t1 = time.perf_counter()
            O.speckle(res.data, abi.SPECKLE_REFINED_LEE, 7, filt.enl, 1.0, threads=cores)
            sp_runs.append(time.perf_counter() - t1)
        sp_per_px = float(np.median(sp_runs)) / res.data.size
    total_s = cal_per_px * info.lines * info.samples + (tc_per_px + sp_per_px) * rows * cols
    value = rows * cols / total_s / 1e6
    sample = (f"median of {runs} runs each: TC on {n_tc} full-width grid rows ({n_tc * cols} px), calibration on "
              f"{deb.shape[0]}x{deb.shape[1]} DN, Refined Lee on the TC sample; whole job extrapolated per pixel: "
              f"cal {cal_per_px * 1e9:.1f} ns/merged px, TC {tc_per_px * 1e9:.1f} ns/px, speckle "
              f"{sp_per_px * 1e9:.1f} ns/px -> {total_s:.1f} s per scene; {cores} threads on {host['cpu_model']}")
    log("cpu baseline:", sample)
    return {"value": round(value, 4), "unit": "Mpx/s", "cores": cores, "kind": "port", "sample": sample,
            "runs": runs, **host,
            "stage_seconds_runs": {"tc": [round(x, 3) for x in tc_runs], "cal": [round(x, 3) for x in cal_runs],
                                   "speckle": [round(x, 3) for x in sp_runs]},
            "seconds_measured": round(dt + sum(tc_runs) + sum(cal_runs) + sum(sp_runs), 2)}


def copy_calibration(dev) -> None:
    """One plain copy kernel of known size (1 GiB read, 1 GiB written), whose
    FETCH_SIZE / WRITE_SIZE counter readings scale the counters of the
    measured kernels to bytes."""
    import torch

    a = torch.empty(1 << 28, dtype=torch.float32, device=dev)
    a.fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    log("copy calibration: 1 GiB read + 1 GiB written by one copy kernel")
    del a, b


def bench_batch64(args, dev):
    """64 square AOI tiles of the full IW scene's UTM grid (log-uniform sides
    256..4096, random placement, random resampler and speckle filter): one
    batched run (persistent launches replayed from a HIP graph, launch shape
    from the autotuner) against the same 64 tiles as one plain
    terrain_correction + speckle call each."""
    import math

    import numpy as np
    import torch

    from sardine_amd import api, synthetic
    from sardine_amd.workload import DeviceScene

    t0 = time.time()
    sc = synthetic.scene_iw_full(1, out_rows=18000, out_cols=25000)
    ds = DeviceScene.build(sc, dev)
    m = ds.prepare_merged()
    torch.cuda.synchronize()
    rows, cols, _ = api.output_grid(ds.scene, ds.tc_config())
    rs = api.ResidentScene(m, ds.scene, ds.dem)
    rng = np.random.default_rng(64)
    F = api.SpeckleFilter
    filters = [None, F.Lee(5, 4.0), F.Frost(5, 1.0), F.GammaMap(7, 4.0), F.RefinedLee(4.0)]
    jobs = []
    for _ in range(64):
        side = int(round(math.exp(rng.uniform(math.log(256), math.log(4096)))))
        r0, c0 = int(rng.integers(0, rows - side + 1)), int(rng.integers(0, cols - side + 1))
        cfg = ds.tc_config(aoi=(r0, c0, side, side), resampling=int(rng.integers(0, 3)))
        jobs.append(api.BatchJob(0, cfg, filters[int(rng.integers(0, len(filters)))]))
    px = sum(j.config.aoi[2] * j.config.aoi[3] for j in jobs)
    log(f"scene built in {time.time() - t0:.1f}s; 64 tiles, {px / 1e6:.1f} Mpx")
    stream = torch.cuda.current_stream(dev)

    # plain: one terrain_correction (+ speckle) call per tile, outputs preallocated
    outs = [(torch.empty(j.config.aoi[2:], dtype=torch.float32, device=dev),
             torch.empty(j.config.aoi[2] * j.config.aoi[3], dtype=torch.float32, device=dev)) for j in jobs]

    def plain():
        for j, (o, so) in zip(jobs, outs):
            img = api.terrain_correction(m, ds.scene, ds.dem, j.config, stream=stream, out=(o, None, None))
            if j.speckle is not None:
                j.speckle.apply(img.data, img.cols, img.rows, stream=stream, out=so)

    plain()
    torch.cuda.synchronize()
    settle_gc()
    reps = max(1, args.steps)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    ev[0].record(stream)
    for i in range(reps):
        plain()
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    plain_lat = [ev[i].elapsed_time(ev[i + 1]) for i in range(reps)]
    plain_ms = float(np.median(plain_lat))
    log(f"plain: {['%.1f' % x for x in plain_lat]} ms")

    b = api.Batch(max_jobs=64, use_graph=True)
    best, tune_ms = b.autotune([rs], jobs, stream=stream)
    log(f"autotuner: segment {best.segment_columns}, tc groups/CU {best.tc_groups_per_cu}, speckle groups/CU "
        f"{best.speckle_groups_per_cu} -> {tune_ms:.2f} ms device")
    prep = b.prepare([rs], jobs)
    b.execute(prep, stream=stream)
    torch.cuda.synchronize()
    settle_gc()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    ev[0].record(stream)
    dev_ms = []
    for i in range(reps):
        b.execute(prep, stream=stream)
        ev[i + 1].record(stream)
        dev_ms.append(b.last_device_ms)
    torch.cuda.synchronize()
    batch_lat = [ev[i].elapsed_time(ev[i + 1]) for i in range(reps)]
    batch_ms = float(np.median(batch_lat))
    log(f"batched: {['%.1f' % x for x in batch_lat]} ms (device {['%.1f' % x for x in dev_ms]})")
    srt = sorted(batch_lat)
    line = {
        "metric": "batched AOI geocode Mpx/s (64 tiles, one persistent graph launch) vs one call per tile",
        "value": round(px / (batch_ms / 1e3) / 1e6, 3),
        "unit": "Mpx/s",
        "n_gpus": 1,
        "steps": reps,
        "warmup": 1,
        "ms_per_step": round(batch_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batch64_iw_full_utm10m_tiles256-4096_mixed_resample_speckle", "tiles": 64,
                   "tile_px": px, "grid": [rows, cols]},
        "latency_ms": {"p50": round(srt[len(srt) // 2], 3), "p99": round(srt[-1], 3)},
        "plain_one_call_per_tile": {"ms_per_step": round(plain_ms, 3),
                                    "value": round(px / (plain_ms / 1e3) / 1e6, 3)},
        "speedup_vs_plain": round(plain_ms / batch_ms, 3),
        "tuning": {"segment_columns": best.segment_columns, "tc_groups_per_cu": best.tc_groups_per_cu,
                   "speckle_groups_per_cu": best.speckle_groups_per_cu},
    }
    print(json.dumps(line), flush=True)
    b.close()
    rs.close()


def bench_serve(args, dev):
    """Four full IW scenes resident in HBM (distinct seeds and orbits); AOI
    requests with log-uniform square sides 256..8192, random scene, CRS
    (UTM or EPSG:4326), resampler and speckle filter.  The native server
    (HBM-budget admission, batches formed on arrival, two batches in
    flight) against one plain terrain_correction (+ speckle) call per
    request, FIFO: saturated throughput of both, latency at the same Poisson
    arrival rate (80 % of the plain loop's capacity), and the highest rate
    the server sustains within the plain loop's p99 there."""
    import math
    import threading

    import numpy as np
    import torch

    from sardine_amd import abi, api, synthetic
    from sardine_amd.workload import DeviceScene

    t0 = time.time()
    scenes = []
    for seed in (1, 2, 3, 4):
        sc = synthetic.scene_iw_full(seed, out_rows=18000, out_cols=25000)
        ds = DeviceScene.build(sc, dev)
        m = ds.prepare_merged()
        torch.cuda.synchronize()
        ds.raws, ds._swc = [], None  # the CInt16 rasters are not needed once calibrated
        grids = {}
        for crs in ("utm", "wgs84"):
            cfg = ds.tc_config() if crs == "utm" else ds.tc_config(crs=abi.CRS_WGS84, utm_zone=0,
                                                                     pixel_spacing_deg=10.0 / 111320.0)
            grids[crs] = (cfg, api.output_grid(ds.scene, cfg)[:2])
        scenes.append((ds, m, api.ResidentScene(m, ds.scene, ds.dem), grids))
    torch.cuda.empty_cache()
    log(f"4 scenes resident in {time.time() - t0:.1f}s")

    rng = np.random.default_rng(2024)
    F = api.SpeckleFilter
    filters = [None, F.Lee(5, 4.0), F.Frost(5, 1.0), F.GammaMap(7, 4.0), F.RefinedLee(4.0)]

    def make_requests(n):
        out = []
        for _ in range(n):
            si = int(rng.integers(0, 4))
            ds, _, _, grids = scenes[si]
            base, (rows, cols) = grids["utm" if rng.random() < 0.5 else "wgs84"]
            side = int(round(math.exp(rng.uniform(math.log(256), math.log(8192)))))
            h, w = min(side, rows), min(side, cols)
            r0, c0 = int(rng.integers(0, rows - h + 1)), int(rng.integers(0, cols - w + 1))
            cfg = api.TerrainCorrectionConfig(pixel_spacing_deg=base.pixel_spacing_deg, crs=base.crs,
                                              utm_zone=base.utm_zone, aoi=(r0, c0, h, w),
                                              resampling=int(rng.integers(0, 3)))
            out.append((si, cfg, filters[int(rng.integers(0, len(filters)))], h * w))
        return out

    n = max(8, args.requests)
    reqs = make_requests(n)
    px = sum(r[3] for r in reqs)
    log(f"{n} requests, {px / 1e9:.2f} Gpx")
    stream = torch.cuda.current_stream(dev)

    def plain_one(r):
        si, cfg, filt, _ = r
        ds, m, _, _ = scenes[si]
        img = api.terrain_correction(m, ds.scene, ds.dem, cfg, stream=stream)
        if filt is not None:
            filt.apply(img.data, img.cols, img.rows, stream=stream)
        torch.cuda.synchronize()

    def plain_run(arrivals):
        start = time.perf_counter()
        lat = []
        for r, a in zip(reqs, arrivals):
            now = time.perf_counter() - start
            if now < a:
                time.sleep(a - now)
            plain_one(r)
            lat.append((time.perf_counter() - start - a) * 1e3)
        return time.perf_counter() - start, lat

    srv = api.Server([s[2] for s in scenes], hbm_budget_bytes=64 << 30, max_batch=64, slots=2)
    settle_gc()

    def served_run(arrivals):
        tickets = [None] * n
        start = time.perf_counter()

        def submitter():
            for i, (r, a) in enumerate(zip(reqs, arrivals)):
                now = time.perf_counter() - start
                if now < a:
                    time.sleep(a - now)
                tickets[i] = srv.submit(r[0], r[1], r[2])

        th = threading.Thread(target=submitter)
        th.start()
        lat, sizes = [], []
        for i in range(n):
            while tickets[i] is None:
                time.sleep(20e-6)
            res = srv.wait(tickets[i])
            assert res.status == 0, res.status
            lat.append(res.latency_ms)
            sizes.append(res.batch_size)
            srv.release(tickets[i])
        th.join()
        return time.perf_counter() - start, lat, sizes

    for r in reqs[:8]:  # warm-up of both paths
        plain_one(r)
    zero = [0.0] * n
    t_plain_sat, _ = plain_run(zero)
    t_srv_sat, _, sizes_sat = served_run(zero)
    cap_plain = n / t_plain_sat
    rate = 0.8 * cap_plain
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n)).tolist()
    t_plain, lat_plain = plain_run(arrivals)
    t_srv, lat_srv, sizes = served_run(arrivals)

    def pct(v, q):
        v = sorted(v)
        return v[min(len(v) - 1, int(round(q * (len(v) - 1))))]

    # throughput at equal p99: the plain loop's p99 at 80 % of its capacity
    # is the latency bound; the same request sequence is offered to the
    # server at rising rates (arrival times compressed by f) up to its
    # saturated capacity, and the highest rate whose p99 stays within the
    # bound is the served throughput at equal p99
    slo = pct(lat_plain, .99)
    sweep, best_f = [], None
    f = 1.0
    while f * rate <= 1.05 * n / t_srv_sat:
        t_f, lat_f, sz_f = served_run([a / f for a in arrivals])
        p99_f = pct(lat_f, .99)
        sweep.append({"rate_req_s": round(f * rate, 2), "mpx_s": round(px / t_f / 1e6, 3),
                      "p50_ms": round(pct(lat_f, .5), 3), "p99_ms": round(p99_f, 3),
                      "mean_batch": float(np.mean(sz_f))})
        log(f"served at {f * rate:.0f} req/s ({f:.2f}x): p50 {pct(lat_f, .5):.1f} p99 {p99_f:.1f} ms, "
            f"{px / t_f / 1e6:.0f} Mpx/s, mean batch {np.mean(sz_f):.1f}")
        if p99_f > slo:
            break
        best_f = f
        f += 0.25
    srv.close()

    log(f"saturated: plain {px / t_plain_sat / 1e6:.0f} Mpx/s, served {px / t_srv_sat / 1e6:.0f} Mpx/s "
        f"(mean batch {np.mean(sizes_sat):.1f})")
    log(f"Poisson {rate:.0f} req/s: plain p50 {pct(lat_plain, .5):.1f} p99 {pct(lat_plain, .99):.1f} ms; served "
        f"p50 {pct(lat_srv, .5):.1f} p99 {pct(lat_srv, .99):.1f} ms (mean batch {np.mean(sizes):.1f})")
    line = {
        "metric": "served AOI geocode Mpx/s (4 resident IW scenes, Poisson requests) & p50/p99 latency",
        "value": round(px / t_srv_sat / 1e6, 3),
        "unit": "Mpx/s",
        "n_gpus": 1,
        "steps": n,
        "warmup": 8,
        "ms_per_step": round(t_srv_sat * 1e3 / n, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "serve_4x_iw_full_aoi256-8192_utm_or_wgs84_mixed", "requests": n,
                   "request_px": px, "hbm_budget_gb": 64, "max_batch": 64, "slots": 2},
        "saturated": {"served_mpx_s": round(px / t_srv_sat / 1e6, 3),
                      "plain_mpx_s": round(px / t_plain_sat / 1e6, 3),
                      "speedup": round(t_plain_sat / t_srv_sat, 3), "served_mean_batch": float(np.mean(sizes_sat))},
        "poisson": {"rate_req_s": round(rate, 2),
                    "served": {"mpx_s": round(px / t_srv / 1e6, 3), "p50_ms": round(pct(lat_srv, .5), 3),
                               "p99_ms": round(pct(lat_srv, .99), 3), "mean_batch": float(np.mean(sizes))},
                    "plain": {"mpx_s": round(px / t_plain / 1e6, 3), "p50_ms": round(pct(lat_plain, .5), 3),
                              "p99_ms": round(pct(lat_plain, .99), 3)}},
        "equal_p99": {"p99_bound_ms": round(slo, 3), "plain_rate_req_s": round(rate, 2),
                      "served_max_rate_req_s": round(best_f * rate, 2) if best_f else None,
                      "throughput_ratio": best_f, "served_sweep": sweep},
    }
    print(json.dumps(line), flush=True)
    for s in scenes:
        s[2].close()


def bench_fp32(args, dev):
    """The full IW scene geocoded by the FP32 scene-centred variant, the mixed
    form (FP64 Newton, FP32 resampling and terrain geometry) and the FP64
    path (which the parity tests hold to the oracle), per resampler and with
    flattening: max/RMS azimuth-time error (s), range-pixel error and σ0 dB
    error of each variant over the pixels valid in both, mask agreement, and
    the chain kernel time of each."""
    import numpy as np
    import torch

    from sardine_amd import _native, abi, api, synthetic
    from sardine_amd.workload import DeviceScene

    L = _native.lib()
    sc = synthetic.scene_iw_full(1, out_rows=18000, out_cols=25000)
    ds = DeviceScene.build(sc, dev)
    m = ds.prepare_merged()
    torch.cuda.synchronize()
    ds.raws, ds._swc = [], None
    rows, cols, _ = api.output_grid(ds.scene, ds.tc_config())
    ati = sc.ati
    variants = [("bilinear", dict(resampling=abi.RESAMPLE_BILINEAR, compute_lia=True)),
                ("bicubic", dict(resampling=abi.RESAMPLE_BICUBIC, compute_lia=True)),
                ("lanczos3", dict(resampling=abi.RESAMPLE_LANCZOS3, compute_lia=True)),
                ("bilinear_flatten", dict(resampling=abi.RESAMPLE_BILINEAR, compute_lia=True, flatten=True))]
    out = {}

    def run(prec, kw, coords):
        cfg = ds.tc_config(precision=prec, **kw)
        L.sardine_set_profiling(1)
        img = api.terrain_correction(m, ds.scene, ds.dem, cfg, radar_coords=coords)
        torch.cuda.synchronize()
        ms = L.sardine_last_kernel_ms(abi.KERNEL_TC_SEGMENT)
        L.sardine_set_profiling(0)
        return img, ms

    def errors(ref, v):
        a, b = ref.radar_coords, v.radar_coords
        both = torch.isfinite(a[..., 0]) & torch.isfinite(b[..., 0])
        dl = (b[..., 0] - a[..., 0])[both].double()
        ds_ = (b[..., 1] - a[..., 1])[both].double()
        s64, s32 = ref.data, v.data
        ok = torch.isfinite(s64) & torch.isfinite(s32) & (s64 > 0) & (s32 > 0)
        ddb = (10.0 * torch.log10(s32[ok].double() / s64[ok].double())).abs()
        return {
            "azimuth_time_err_s": {"max": float(dl.abs().max() * ati), "rms": float(dl.pow(2).mean().sqrt() * ati)},
            "range_px_err": {"max": float(ds_.abs().max()), "rms": float(ds_.pow(2).mean().sqrt())},
            "sigma0_db_err": {"max": float(ddb.max()), "rms": float(ddb.pow(2).mean().sqrt()),
                              "p99": float(torch.quantile(ddb[:: max(1, ddb.numel() // 2000000)].float(), 0.99)),
                              "above_1e-3_db": int((ddb > 1e-3).sum())},
            "mask_agreement": float((ref.mask == v.mask).double().mean()),
            "valid": v.valid_pixel_count,
        }

    for name, kw in variants:
        ref, _ = run(abi.PRECISION_FP64, kw, True)
        v32, _ = run(abi.PRECISION_FP32_CENTRED, kw, True)
        e32 = errors(ref, v32)
        del v32
        vmx, _ = run(abi.PRECISION_MIXED, kw, True)
        emx = errors(ref, vmx)
        valid64 = ref.valid_pixel_count
        del ref, vmx
        torch.cuda.empty_cache()
        _, ms64 = run(abi.PRECISION_FP64, kw, False)
        _, ms32 = run(abi.PRECISION_FP32_CENTRED, kw, False)
        _, msmx = run(abi.PRECISION_MIXED, kw, False)
        out[name] = {
            "fp32_centred": e32, "mixed": emx, "valid_fp64": valid64,
            "tc_kernel_ms": {"fp64": round(ms64, 3), "fp32": round(ms32, 3), "mixed": round(msmx, 3)},
            "mpx_s": {"fp64": round(rows * cols / ms64 / 1e3, 1), "fp32": round(rows * cols / ms32 / 1e3, 1),
                      "mixed": round(rows * cols / msmx / 1e3, 1)},
        }
        log(f"{name}: {json.dumps(out[name])}")
        torch.cuda.empty_cache()
    line = {
        "metric": "FP32 scene-centred and mixed-precision variants vs FP64: errors and TC Mpx/s",
        "value": out["bilinear"]["mpx_s"]["fp32"],
        "unit": "Mpx/s",
        "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": out["bilinear"]["tc_kernel_ms"]["fp32"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "fp32_iw_full_utm10m_25000x18000", "grid": [rows, cols]},
        "variants": out,
    }
    print(json.dumps(line), flush=True)



def bench_cog(args, dev):
    """COG export of the full-size geocoded dB product (tile 512, as the
    reference pipeline writes it, run.rs:433-441): the device pyramid + pack
    (value, HBM roofline), then the whole file write (device pack, PCIe,
    file system) against the oracle's writer on the same raster, whose file
    must be identical byte for byte."""
    import hashlib

    import numpy as np
    import torch

    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    from sardine_amd import _native, abi, api, synthetic
    from sardine_amd.workload import DeviceScene

    L = _native.lib()
    stream = torch.cuda.current_stream(dev)
    if args.copy_calibration:
        copy_calibration(dev)
    t0 = time.time()
    sc = synthetic.scene_iw_full(1, out_rows=18000, out_cols=25000)
    ds = DeviceScene.build(sc, dev)
    cfg = ds.tc_config()
    img = ds.terrain_correction(ds.prepare_merged(stream=stream), cfg, stream=stream)
    api.to_db_inplace(img.data, 0.0, stream=stream)
    rows, cols, gt = img.rows, img.cols, list(img.geotransform)
    data = img.data.reshape(rows, cols).contiguous()
    del img
    torch.cuda.synchronize()
    log(f"dB product {rows}x{cols} ready in {time.time() - t0:.1f}s")
    ts = 512
    lay = api.cog_layout(cols, rows, ts, cfg.crs)
    tiles = torch.empty(lay.tile_bytes // 4, dtype=torch.float32, device=dev)
    # algorithmic bytes of one pack: the full raster read once, every tile
    # written once, and every overview but the coarsest read once (by the
    # launch that makes the next level)
    lvl_bytes = [lay.level[k].tiles_x * lay.level[k].tiles_y * ts * ts * 4 for k in range(lay.n_levels)]
    alg_bytes = rows * cols * 4 + lay.tile_bytes + sum(lvl_bytes[1:lay.n_levels - 1])
    L.sardine_set_profiling(1)
    for _ in range(max(1, args.warmup)):
        api.cog_pack(data, lay, tiles=tiles, stream=stream)
    steps = max(args.steps, 1)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    settle_gc()
    torch.cuda.synchronize()
    ev[0].record(stream)
    for _ in range(steps):
        api.cog_pack(data, lay, tiles=tiles, stream=stream)
    ev[1].record(stream)
    torch.cuda.synchronize()
    k_ms = recent_kernel_ms(L, abi.KERNEL_COG_PACK, steps)
    L.sardine_set_profiling(0)
    ms = ev[0].elapsed_time(ev[1]) / steps
    pack_ms = float(np.mean(k_ms))
    gbs = alg_bytes / (pack_ms / 1e3) / 1e9
    log(f"pack: {['%.3f' % x for x in k_ms]} ms per pack ({lay.n_levels} levels, {lay.tile_bytes / 1e9:.2f} GB "
        f"of tiles), {gbs:.0f} GB/s algorithmic")
    del tiles
    traffic, traffic_src = measured_traffic("cog_level_kernel", "cog_iw_full_25000x18000_ts512")

    # the whole file, device path vs the oracle writer (one thread)
    tmp = pathlib.Path("/dev/shm") if pathlib.Path("/dev/shm").is_dir() else pathlib.Path("/tmp")
    gpu_path, ref_path = tmp / f"sardine_bench_gpu_{os.getpid()}.tif", tmp / f"sardine_bench_ref_{os.getpid()}.tif"

    def digest(p):
        h = hashlib.sha256()
        with open(p, "rb") as f:
            while True:
                b = f.read(64 << 20)
                if not b:
                    break
                h.update(b)
        return h.hexdigest()

    try:
        api.write_cog_with_crs(gpu_path, data, cols, rows, gt, crs=cfg.crs, utm_zone=cfg.utm_zone, tile_size=ts)
        t = time.perf_counter()
        api.write_cog_with_crs(gpu_path, data, cols, rows, gt, crs=cfg.crs, utm_zone=cfg.utm_zone, tile_size=ts)
        gpu_write_s = time.perf_counter() - t
        cpu_write_s, same = None, None
        if not args.no_cpu_baseline:
            host = data.cpu().numpy()
            t = time.perf_counter()
            st = O.write_cog(ref_path, host, gt, crs=cfg.crs, utm_zone=cfg.utm_zone, tile_size=ts)
            cpu_write_s = time.perf_counter() - t
            assert st == 0, st

        if cpu_write_s is not None:
            same = gpu_path.stat().st_size == ref_path.stat().st_size and digest(gpu_path) == digest(ref_path)
        # the striped GeoTIFF of the same raster (the pipeline's default output)
        api.write_geotiff_with_crs(gpu_path, data, cols, rows, gt, crs=cfg.crs, utm_zone=cfg.utm_zone)
        t = time.perf_counter()
        api.write_geotiff_with_crs(gpu_path, data, cols, rows, gt, crs=cfg.crs, utm_zone=cfg.utm_zone)
        tif_gpu_s = time.perf_counter() - t
        tif_bytes = gpu_path.stat().st_size
        tif_cpu_s, tif_same = None, None
        if cpu_write_s is not None:
            t = time.perf_counter()
            assert O.write_geotiff(ref_path, host, gt, crs=cfg.crs, utm_zone=cfg.utm_zone) == 0
            tif_cpu_s = time.perf_counter() - t
            tif_same = tif_bytes == ref_path.stat().st_size and digest(gpu_path) == digest(ref_path)
    finally:
        for p in (gpu_path, ref_path):
            if p.exists():
                p.unlink()
    log(f"striped GeoTIFF {tif_bytes / 1e9:.2f} GB: device path {tif_gpu_s:.2f}s, oracle writer {tif_cpu_s}s, "
        f"identical={tif_same}")
    fb = lay.file_bytes
    log(f"file {fb / 1e9:.2f} GB: device path {gpu_write_s:.2f}s, oracle writer {cpu_write_s}s, identical={same}")
    px = rows * cols
    line = {
        "metric": "COG export Mpx/s (overview pyramid + tile pack, inputs in HBM)",
        "value": round(px / (pack_ms / 1e3) / 1e6, 1),
        "unit": "Mpx/s",
        "n_gpus": 1, "steps": steps, "warmup": max(1, args.warmup), "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "cog_iw_full_25000x18000_ts512", "raster": [rows, cols], "tile_size": ts,
                   "levels": lay.n_levels, "bigtiff": bool(lay.bigtiff), "file_bytes": fb},
        "roofline": {"kernel": "cog_level_kernel (all launches of one pack)", "bound": "hbm",
                     "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes": alg_bytes, "launch_ms": round(pack_ms, 4)},
        "write_file": {"device_path_s": round(gpu_write_s, 3), "GBps": round(fb / gpu_write_s / 1e9, 2),
                       "identical_to_oracle": same, "fs": str(tmp)},
        "striped_geotiff": {"file_bytes": tif_bytes, "device_path_s": round(tif_gpu_s, 3),
                            "GBps": round(tif_bytes / tif_gpu_s / 1e9, 2),
                            "oracle_writer_s": None if tif_cpu_s is None else round(tif_cpu_s, 3),
                            "identical_to_oracle": tif_same},
        "cpu_baseline": None if cpu_write_s is None else {
            "value": round(px / cpu_write_s / 1e6, 1), "unit": "Mpx/s (whole file write)", "cores": 1,
            "kind": "port",
            "sample": "the oracle's COG writer (restatement of export.rs:329-760) on the whole raster, file "
                      "included; compare with write_file",
            "seconds": round(cpu_write_s, 3), "device_path_speedup": round(cpu_write_s / gpu_write_s, 2)},
    }
    print(json.dumps(line), flush=True)



def bench_multilook(args, dev):
    """Pre-TC multilook (apply_multilook, scene_prep.rs:1060-1161) of the
    benchmark scene's merged σ⁰ + NESZ (13100 x 60814), inputs in HBM.  value:
    input Mpx/s at 4 x 1 (range x azimuth) looks; other shapes alongside."""
    import numpy as np
    import torch

    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    from sardine_amd import _native, abi, api, synthetic
    from sardine_amd.workload import DeviceScene

    L = _native.lib()
    if args.copy_calibration:
        copy_calibration(dev)
    stream = torch.cuda.current_stream(dev)
    sc = synthetic.scene_iw_full(1, out_rows=18000, out_cols=25000)
    ds = DeviceScene.build(sc, dev)
    m = ds.prepare_merged(stream=stream)
    ati = sc.ati
    torch.cuda.synchronize()
    lines, samples = m.lines, m.samples
    log(f"merged {lines}x{samples}")
    shapes = [(4, 1), (2, 2), (5, 1), (10, 2), (3, 3)]
    res = {}
    L.sardine_set_profiling(1)
    for r, a in shapes:
        ol, os_ = lines // a, samples // r
        od = torch.empty((ol, os_), dtype=torch.float32, device=dev)
        on = torch.empty_like(od)
        for _ in range(max(1, args.warmup)):
            api.apply_multilook(m, r, a, ati, stream=stream, out=(od, on))
        settle_gc()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            api.apply_multilook(m, r, a, ati, stream=stream, out=(od, on))
        e1.record(stream)
        torch.cuda.synchronize()
        k = recent_kernel_ms(L, abi.KERNEL_MULTILOOK, args.steps)
        kms = float(np.mean(k))
        alg = ol * a * os_ * r * 8 + ol * os_ * 8
        res[f"{r}x{a}"] = {"kernel_ms": round(kms, 4), "step_ms": round(e0.elapsed_time(e1) / args.steps, 4),
                           "algorithmic_bytes": alg, "GBps": round(alg / (kms / 1e3) / 1e9, 1),
                           "hbm_frac": round(alg / (kms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"multilook {r}x{a}: {json.dumps(res[f'{r}x{a}'])}")
        if (r, a) == shapes[0]:
            keep = (od, on)
        else:
            del od, on
    L.sardine_set_profiling(0)
    od, on = keep
    cpu = None
    same = None
    if not args.no_cpu_baseline:
        # the oracle (one thread) on the whole image (about a second of CPU
        # work), which also checks the device result bit for bit at full size
        band = lines
        hd = m.data[:band].cpu().numpy()
        hn = m.nesz[:band].cpu().numpy()
        t = time.perf_counter()
        st, rd, rn, _ = O.multilook(hd, hn, m.info, 4, 1, ati)
        cs = time.perf_counter() - t
        assert st == 0
        gd = od[:band].cpu().numpy()
        gn = on[:band].cpu().numpy()
        same = bool(np.array_equal(gd.view(np.uint32), rd[:band].view(np.uint32)) and
                    np.array_equal(gn.view(np.uint32), rn[:band].view(np.uint32)))
        cpu = {"value": round(band * samples / cs / 1e6, 1), "unit": "input Mpx/s", "cores": 1, "kind": "port",
               "sample": f"the oracle's apply_multilook (4 x 1) on the whole merged image ({band} x {samples})",
               "seconds": round(cs, 3), "device_result_bit_exact": same}
        log(f"cpu: {json.dumps(cpu)}")
    main = res["4x1"]
    px = lines * samples
    traffic, traffic_src = measured_traffic("multilook_kernel", "multilook_iw_full_13100x60814")
    line = {
        "metric": "pre-TC multilook input Mpx/s (sigma0 + NESZ, 4x1 looks, inputs in HBM)",
        "value": round(px / (main["kernel_ms"] / 1e3) / 1e6, 1), "unit": "Mpx/s",
        "n_gpus": 1, "steps": args.steps, "warmup": max(1, args.warmup), "ms_per_step": main["step_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "multilook_iw_full_13100x60814", "merged": [lines, samples], "looks": "4x1"},
        "roofline": {"kernel": "multilook_kernel<4>", "bound": "hbm", "achieved": main["GBps"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": main["hbm_frac"], "traffic": traffic,
                     "traffic_source": traffic_src,
                     "algorithmic_bytes": main["algorithmic_bytes"], "launch_ms": main["kernel_ms"]},
        "shapes": res,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def bench_grd(args, dev):
    """Ground-range projection (to_ground_range, ground_range.rs:303-442) of
    the benchmark scene's merged σ⁰ (13100 x 60814) to 10 m, with a synthetic
    3-swath geolocation grid over its range (incidence 30-46°), inputs in
    HBM.  Azimuth spacing 14.08 m (1 look, the IW case) and 2.0 m (5 looks)."""
    import numpy as np
    import torch

    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    from sardine_amd import _native, abi, api, synthetic
    from sardine_amd.workload import DeviceScene

    L = _native.lib()
    if args.copy_calibration:
        copy_calibration(dev)
    stream = torch.cuda.current_stream(dev)
    sc = synthetic.scene_iw_full(1, out_rows=18000, out_cols=25000)
    ds = DeviceScene.build(sc, dev)
    m = ds.prepare_merged(stream=stream)
    torch.cuda.synchronize()
    lines, samples = m.lines, m.samples
    per = samples // 3
    grids = []
    for k in range(3):
        c0, c1 = max(k * per - 40, 0), ((k + 1) * per - 1 if k < 2 else samples - 1)
        grids.append([(ln, m.near_slant_range_time_s + c * m.range_pixel_spacing_s, 30.0 + 16.0 * c / samples)
                      for ln in (0, lines // 2, lines - 1) for c in np.linspace(c0, c1, 21)])
    res = {}
    L.sardine_set_profiling(1)
    keep = None
    for az in (14.08, 2.0):
        g = api.to_ground_range(m, grids, 10.0, az, stream=stream)
        for _ in range(max(1, args.warmup)):
            api.to_ground_range(m, grids, 10.0, az, stream=stream, out=g.data)
        settle_gc()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            api.to_ground_range(m, grids, 10.0, az, stream=stream, out=g.data)
        e1.record(stream)
        torch.cuda.synchronize()
        k = float(np.mean(recent_kernel_ms(L, abi.KERNEL_GRD, args.steps)))
        # two taps per output column and input line are read (the map makes
        # them adjacent columns), one value per output pixel written
        taps = g.lines * g.azimuth_looks * g.samples * 8 + g.lines * g.samples * 4
        # output columns sit ~2.6 input columns (~10 B) apart, so every 64 B
        # DRAM segment of every input line read holds taps: the least HBM
        # traffic is those lines whole plus the output
        alg = g.lines * g.azimuth_looks * samples * 4 + g.lines * g.samples * 4
        res[f"az{az}"] = {"out": [g.lines, g.samples], "azimuth_looks": g.azimuth_looks, "kernel_ms": round(k, 4),
                          "call_ms": round(e0.elapsed_time(e1) / args.steps, 4), "algorithmic_bytes": alg,
                          "tap_bytes": taps, "GBps": round(alg / (k / 1e3) / 1e9, 1),
                          "hbm_frac": round(alg / (k / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"grd az {az}: {json.dumps(res[f'az{az}'])}")
        if keep is None:
            keep = g
    L.sardine_set_profiling(0)
    cpu = None
    if not args.no_cpu_baseline:
        hd = m.data.cpu().numpy()
        t = time.perf_counter()
        st, ref, img = O.to_ground_range(hd, m.info, grids, 10.0, 14.08)
        cs = time.perf_counter() - t
        assert st == 0
        same = bool(np.array_equal(keep.data.cpu().numpy().view(np.uint32), ref.view(np.uint32)))
        cpu = {"value": round(lines * samples / cs / 1e6, 1), "unit": "input Mpx/s", "cores": 1, "kind": "port",
               "sample": f"the oracle's to_ground_range on the whole merged image ({lines} x {samples}, 14.08 m)",
               "seconds": round(cs, 3), "device_result_bit_exact": same}
        log(f"cpu: {json.dumps(cpu)}")
    main = res["az14.08"]
    traffic, traffic_src = measured_traffic("grd_kernel", "grd_iw_full_13100x60814_10m")
    line = {
        "metric": "ground-range projection input Mpx/s (10 m, 1 azimuth look, inputs in HBM)",
        "value": round(lines * samples / (main["kernel_ms"] / 1e3) / 1e6, 1), "unit": "Mpx/s",
        "n_gpus": 1, "steps": args.steps, "warmup": max(1, args.warmup), "ms_per_step": main["call_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "grd_iw_full_13100x60814_10m", "merged": [lines, samples], "out": main["out"]},
        "roofline": {"kernel": "grd_kernel", "bound": "hbm", "achieved": main["GBps"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": main["hbm_frac"], "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes": main["algorithmic_bytes"], "launch_ms": main["kernel_ms"]},
        "cases": res,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def _write_slc_tiff(path, raw) -> None:
    """A Sentinel-1-style measurement TIFF (little-endian classic TIFF,
    uncompressed CInt16, one row per strip, contiguous rows) of raw
    (height, width, 2) int16."""
    import struct

    h, w = raw.shape[:2]
    tags = [(256, 4, 1, w), (257, 4, 1, h), (258, 3, 1, 32), (259, 3, 1, 1), (273, 4, h, None), (277, 3, 1, 1),
            (278, 4, 1, 1), (279, 4, h, None), (339, 3, 1, 5)]
    table = 8 + 2 + 12 * len(tags) + 4
    counts = table + 4 * h
    data = counts + 4 * h
    with open(path, "wb") as f:
        f.write(b"II" + struct.pack("<HI", 42, 8) + struct.pack("<H", len(tags)))
        for tag, typ, cnt, val in tags:
            v = table if tag == 273 else counts if tag == 279 else val
            f.write(struct.pack("<HHI", tag, typ, cnt) + (struct.pack("<HH", v, 0) if typ == 3 else struct.pack("<I", v)))
        f.write(struct.pack("<I", 0))
        f.write(struct.pack(f"<{h}I", *[data + k * w * 4 for k in range(h)]))
        f.write(struct.pack(f"<{h}I", *([w * 4] * h)))
        f.write(raw.tobytes())


def bench_slc(args, dev):
    """A full-size IW swath measurement TIFF (9 bursts x 1500 lines x 21000
    samples CInt16, 1.13 GB, in the page cache) read into HBM: the whole
    swath in one call and burst by burst as deburst reads it
    (deburst.rs:315-334), against the oracle's read_burst_raw into host
    memory (one thread, as the reference reads)."""
    import numpy as np
    import torch

    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    from sardine_amd import api

    h, w, lpb = 13500, 21000, 1500
    t = torch.empty((h, w, 2), dtype=torch.int16, device=dev)
    api.synth_cint16(t, 7)
    host = t.cpu().numpy()
    tmp = pathlib.Path("/dev/shm") if pathlib.Path("/dev/shm").is_dir() else pathlib.Path("/tmp")
    path = tmp / f"sardine_bench_slc_{os.getpid()}.tiff"
    try:
        _write_slc_tiff(path, host)
        r = api.SlcReader.open(path)
        out = torch.empty_like(t)
        r.read_burst_raw(0, h, out=out)
        torch.cuda.synchronize()
        steps = max(args.steps, 1)
        settle_gc()
        t0 = time.perf_counter()
        for _ in range(steps):
            r.read_burst_raw(0, h, out=out)
        torch.cuda.synchronize()
        whole_s = (time.perf_counter() - t0) / steps
        same = bool(torch.equal(out, t))
        out.zero_()
        t0 = time.perf_counter()
        for _ in range(steps):
            for b in range(h // lpb):
                r.read_burst_raw(b * lpb, lpb, out=out[b * lpb:(b + 1) * lpb])
        torch.cuda.synchronize()
        burst_s = (time.perf_counter() - t0) / steps
        same = same and bool(torch.equal(out, t))
        cpu = None
        if not args.no_cpu_baseline:
            st, tt, _ = O.slc_open(path)
            t0 = time.perf_counter()
            st2, ref = O.slc_read(path, tt, 0, h)
            cs = time.perf_counter() - t0
            assert st == 0 and st2 == 0 and np.array_equal(ref, host)
            cpu = {"value": round(host.nbytes / cs / 1e9, 2), "unit": "GB/s (file → host memory)", "cores": 1,
                   "kind": "port", "sample": "the oracle's read_burst_raw of the whole swath (1.13 GB, page cache)",
                   "seconds": round(cs, 3)}
    finally:
        if path.exists():
            path.unlink()
    gb = host.nbytes / 1e9
    log(f"slc: whole {whole_s * 1e3:.1f} ms ({gb / whole_s:.1f} GB/s), per burst {burst_s * 1e3:.1f} ms, "
        f"identical={same}, cpu {cpu}")
    line = {
        "metric": "SLC measurement TIFF → HBM GB/s (whole swath, page cache)",
        "value": round(gb / whole_s, 2), "unit": "GB/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": round(whole_s * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "i16", "data": "synthetic",
        "config": {"workload": "slc_iw_swath_13500x21000", "bytes": host.nbytes},
        "roofline": None,
        "note": "I/O path with no kernel: parallel preads into pinned chunks, host→device copies overlapped; "
                "bounded by host memory copy and PCIe (Gen5 x16 ≈ 64 GB/s per direction)",
        "per_burst": {"ms": round(burst_s * 1e3, 3), "GBps": round(gb / burst_s, 2)},
        "device_result_bit_exact": same,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)

def _srtm_heights(la: int, lo: int):
    """3601 x 3601 int16 heights of a synthetic SRTM tile (rolling terrain,
    0.1 % voids), rows from the north edge."""
    import numpy as np

    y, x = np.mgrid[0:3601, 0:3601].astype(np.float32) / 3600.0
    z = 800.0 + 600.0 * np.sin(2.0 * np.pi * (x + lo) * 1.3) * np.cos(2.0 * np.pi * (y + la) * 0.9)
    z += np.random.default_rng(la * 1000 + lo).normal(0.0, 4.0, z.shape)
    h = np.clip(np.rint(z), -500, 8000).astype(np.int16)
    h[np.random.default_rng(lo * 1000 + la).random(h.shape) < 1e-3] = -32768
    return h


def _write_glo30_tiff(path, a, la: int, lo: int, tile: int = 1024, nodata: str = "-9999") -> None:
    """A Copernicus-style GLO-30 tile: little-endian classic TIFF, one f32
    band in tile x tile tiles, Deflate with the floating-point predictor
    (bytes of each tile row split into planes, most significant first, then
    differenced), ModelPixelScale + ModelTiepoint (corner of the north-west
    cell at (lo, la + 1)) and GDAL_NODATA."""
    import struct
    import zlib

    import numpy as np

    h, w = a.shape
    nx, ny = -(-w // tile), -(-h // tile)
    blocks = []
    for by in range(ny):
        for bx in range(nx):
            t = np.zeros((tile, tile), np.float32)
            part = a[by * tile:(by + 1) * tile, bx * tile:(bx + 1) * tile]
            t[:part.shape[0], :part.shape[1]] = part
            planes = t.view(np.uint8).reshape(tile, tile, 4)[:, :, ::-1].transpose(0, 2, 1).reshape(tile, 4 * tile)
            enc = np.diff(planes, axis=1, prepend=np.zeros((tile, 1), np.uint8))
            blocks.append(zlib.compress(enc.tobytes(), 6))
    n = len(blocks)
    scale = struct.pack("<3d", 1.0 / w, 1.0 / h, 0.0)
    tie = struct.pack("<6d", 0.0, 0.0, 0.0, float(lo), float(la + 1), 0.0)
    nd = nodata.encode() + b"\0"
    # tag, type, count, inline value or bytes placed after the IFD
    tags = [(256, 4, 1, w), (257, 4, 1, h), (258, 3, 1, 32), (259, 3, 1, 8), (262, 3, 1, 1), (277, 3, 1, 1),
            (284, 3, 1, 1), (317, 3, 1, 3), (322, 3, 1, tile), (323, 3, 1, tile), (324, 4, n, None),
            (325, 4, n, None), (339, 3, 1, 3), (33550, 12, 3, scale), (33922, 12, 6, tie), (42113, 2, len(nd), nd)]
    ifd = 8
    extra = ifd + 2 + 12 * len(tags) + 4
    blobs = {324: None, 325: struct.pack(f"<{n}I", *[len(b) for b in blocks]), 33550: scale, 33922: tie, 42113: nd}
    data0 = extra + 4 * n + len(blobs[325]) + len(scale) + len(tie) + len(nd)
    offs, o = [], data0
    for b in blocks:
        offs.append(o)
        o += len(b)
    blobs[324] = struct.pack(f"<{n}I", *offs)
    out = [b"II" + struct.pack("<HI", 42, ifd) + struct.pack("<H", len(tags))]
    tail, at = [], extra
    for tag, typ, cnt, val in tags:
        if tag in blobs:
            out.append(struct.pack("<HHII", tag, typ, cnt, at))
            tail.append(blobs[tag])
            at += len(blobs[tag])
        elif typ == 3:
            out.append(struct.pack("<HHIHH", tag, typ, cnt, val, 0))
        else:
            out.append(struct.pack("<HHII", tag, typ, cnt, val))
    out.append(struct.pack("<I", 0))
    with open(path, "wb") as f:
        f.write(b"".join(out) + b"".join(tail) + b"".join(blocks))


def bench_dem(args, dev):
    """DEM tiles loaded from files (page cache) into one device raster:
    DemMosaic::load_srtm_files of a 2 x 2 SRTM 1" mosaic (dem.rs:628-640; 4 x
    25.9 MB big-endian tiles → 7201 x 7201 f32) and Glo30Tile::load of a
    3600 x 3600 Copernicus-style tile (dem.rs:248-352; Deflate, floating-point
    predictor, 1024 tiles).  value: SRTM input Mpx/s, file → HBM; the
    oracle's SrtmTile::load of the same files (one thread) beside it."""
    import numpy as np
    import torch

    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    from sardine_amd import _native, abi, api

    L = _native.lib()
    stream = torch.cuda.current_stream(dev)
    tmp = pathlib.Path("/dev/shm") if pathlib.Path("/dev/shm").is_dir() else pathlib.Path("/tmp")
    work = tmp / f"sardine_bench_dem_{os.getpid()}"
    work.mkdir()
    steps = max(args.steps, 1)
    try:
        tiles, paths = {}, []
        for la, lo in ((50, 10), (50, 11), (51, 10), (51, 11)):
            tiles[(la, lo)] = _srtm_heights(la, lo)
            paths.append(work / f"N{la:02d}E{lo:03d}.hgt")
            paths[-1].write_bytes(tiles[(la, lo)].astype(">i2").tobytes())
        glo = work / "Copernicus_DSM_COG_10_N45_00_E006_00_DEM.tif"
        gz = (_srtm_heights(45, 6)[:3600, :3600].astype(np.float32) + np.float32(0.25)).astype(np.float32)
        gz[gz < -400] = np.float32(-9999.0)
        _write_glo30_tiff(glo, gz, 45, 6)
        log(f"files: 4 SRTM tiles ({4 * 3601 * 3601 * 2 / 1e6:.1f} MB), GLO-30 tile {glo.stat().st_size / 1e6:.1f} MB")
        res = {}
        L.sardine_set_profiling(1)
        for name, fn in (("srtm_2x2", lambda: api.load_srtm_mosaic(paths, device=dev, stream=stream)),
                         ("glo30_3600", lambda: api.load_glo30_tile(glo, device=dev, stream=stream))):
            for _ in range(max(1, args.warmup)):
                r = fn()
            torch.cuda.synchronize()
            settle_gc()
            t0 = time.perf_counter()
            for _ in range(steps):
                r = fn()
            torch.cuda.synchronize()
            call_s = (time.perf_counter() - t0) / steps
            per = 4 if name == "srtm_2x2" else 1
            k = recent_kernel_ms(L, abi.KERNEL_DEM, min(steps * per, 64))
            kms = float(np.mean(k))
            if name == "srtm_2x2":
                px_in = 4 * 3601 * 3601
                alg = 3601 * 3601 * 6  # one tile's decode: 2 B read, 4 B written per sample
                got = r.data.cpu().numpy()
                want = np.full((7201, 7201), np.nan, np.float32)
                for (la, lo), t in tiles.items():
                    v = t.astype(np.float32)
                    v[t == -32768] = np.nan
                    want[(51 - la) * 3600:(51 - la) * 3600 + 3601, (lo - 10) * 3600:(lo - 10) * 3600 + 3601] = v
                same = bool(np.array_equal(got, want, equal_nan=True))
            else:
                px_in = 3600 * 3600
                alg = 3600 * 3600 * 8  # 4 B read, 4 B written per sample
                want = gz.copy()
                want[want == np.float32(-9999.0)] = np.nan
                same = bool(np.array_equal(r.data.cpu().numpy(), want, equal_nan=True))
            res[name] = {"call_ms": round(call_s * 1e3, 3), "input_Mpx_per_s": round(px_in / call_s / 1e6, 1),
                         "kernel_ms": round(kms, 4), "kernel_algorithmic_bytes": alg,
                         "kernel_GBps": round(alg / (kms / 1e3) / 1e9, 1),
                         "kernel_hbm_frac": round(alg / (kms / 1e3) / 1e9 / HBM_PEAK_GBS, 4), "identical": same}
            log(f"{name}: {json.dumps(res[name])}")
        L.sardine_set_profiling(0)
        cpu = None
        if not args.no_cpu_baseline:
            t0 = time.perf_counter()
            st, _, _ = O.srtm_sample([str(p) for p in paths], np.zeros(0), np.zeros(0))
            cs = time.perf_counter() - t0
            assert st == 0
            cpu = {"value": round(4 * 3601 * 3601 / cs / 1e6, 1), "unit": "input Mpx/s", "cores": 1, "kind": "port",
                   "sample": "the oracle's SrtmTile::load of the same 4 tiles (one thread)", "seconds": round(cs, 3)}
            log(f"cpu: {json.dumps(cpu)}")
    finally:
        for p in work.iterdir():
            p.unlink()
        work.rmdir()
    main = res["srtm_2x2"]
    line = {
        "metric": "SRTM 1\" mosaic load, input Mpx/s (2 x 2 tiles, file in page cache → HBM)",
        "value": main["input_Mpx_per_s"], "unit": "Mpx/s", "n_gpus": 1, "steps": steps,
        "warmup": max(1, args.warmup), "ms_per_step": main["call_ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "i16", "data": "synthetic",
        "config": {"workload": "dem_srtm_2x2_glo30_3600", "srtm_tiles": 4, "mosaic": [7201, 7201],
                   "glo30": [3600, 3600, "deflate, floating-point predictor, 1024 tiles"]},
        "roofline": {"kernel": "srtm_tile_kernel", "bound": "hbm", "achieved": main["kernel_GBps"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": main["kernel_hbm_frac"], "traffic": None,
                     "algorithmic_bytes": main["kernel_algorithmic_bytes"], "launch_ms": main["kernel_ms"]},
        "note": "I/O path: the files' bytes go through pinned chunks to HBM (parallel preads overlapped with the "
                "copies); the device decode kernel is a small part of a call",
        "cases": res,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()



