"""Benchmark of the path-tracer hot path on one MI355X.

Workload (one step = one drawOnce frame): the seeded heightfield scene -
a 1024 x 1024 vertex grid of fBm noise (2,093,058 triangles) plus 64
spheres with GGX / Beckmann / SGGX-surface materials, constant environment
light, 1920 x 1080 pixels, 64 camera samples per pixel, max depth 8, FP32
paths, FP64 accumulation.  Inputs (scene, BVH, path pool) are resident in
HBM before the timed region.

Prints one JSON line: Msamples/s over the timed steps, the roofline of the
dominant kernel (closest-hit BVH traversal, HBM-bound; algorithmic bytes per
ray segment = one node record (128 B) per node fetched + 48 B per primitive
tested + 64 B of ray state and hit, counted by the kernels in a separate
counting frame),
and the CPU oracle timed on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--small]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E datasheet


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads():
    """Host cores of this job for the CPU baselines: the cores the process
    may run on (sched_getaffinity), limited to the job's CPU share where the
    environment states one (OMP_NUM_THREADS; 16 per GPU on the MI355X
    boxes), one oracle thread per core."""
    n = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def gpu_material(model, fields):
    from tempestrenderer_amd.scene import Material
    return Material(model, **fields).setup()


def load_traffic(path):
    """Measured HBM bytes per trace launch from a committed rocprofv3 PMC
    summary (profiles/), or None."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(desc, rows, threads):
    """The CPU oracle (FP32, reference estimator) on `rows` rows centred in
    the same frame, with `threads` host threads."""
    import oracle
    o = oracle.OracleScene(desc.width, desc.height, desc.vpi)
    desc.populate(o, oracle.material)
    o.commitScene()
    r0 = (desc.height - rows) // 2
    _, _, secs = o.renderRows(r0, r0 + rows, frame=0, fp64=False, threads=threads)
    samples = rows * desc.width * desc.settings["samples_camera"]
    return samples / secs / 1e6, secs, samples, r0


def gonio_bench(cpu_paths=2.0e8, threads=None, cpu=True):
    """The batched virtual gonioreflectometer: all 6561 requests in one
    tr_scene_render_requests call (graph replay, autotuned), timed after a
    small warm-up call; beside it the CPU oracle rendering the first
    requests one after another, as the reference's request loop does, until
    about `cpu_paths` paths."""
    threads = threads or host_threads()
    import copy
    from tempestrenderer_amd import workloads
    from tempestrenderer_amd.scene import RayTracerScene
    d, reqs = workloads.gonio()
    g = RayTracerScene(32, 32, d.vpi)
    d.populate(g, gpu_material)
    g.commitScene()
    g.renderRequests(reqs[:64], use_graph=True)
    t0 = time.perf_counter()
    _, _, st = g.renderRequests(reqs, use_graph=True, autotune=True)
    secs = time.perf_counter() - t0
    g.close()
    out = {"workload": "gonioreflectometer: 64x64 fBm patch (7938 tris), GGX alpha 0.3, 81 x 81 light/view "
                       "requests, 32^2-256^2 px, 4-64 spp (Philox seed 1), depth 1, RR 0.9",
           "requests": len(reqs), "paths": st["paths"], "seconds": round(secs, 4),
           "requests_per_s": round(len(reqs) / secs, 1), "msamples_per_s": round(st["paths"] / secs / 1e6, 1),
           "batches": st["batches"], "tuned": {"pool_slots": st["pool_slots"],
                                               "trace_waves_per_cu": st["trace_waves_per_cu"],
                                               "candidates": st["candidates"], "trial_ms": round(st["tune_ms"], 2)},
           "graph": True}
    if cpu:
        import oracle
        done, paths, t_cpu, base = 0, 0, 0.0, 0
        while done < len(reqs) and paths < cpu_paths:
            q = reqs[done]
            d2 = copy.copy(d)
            d2.lights = [dict(d.lights[0], direction=tuple(q["light_direction"]))]
            o = oracle.OracleScene(q["width"], q["height"], q["view_proj_inv"])
            d2.populate(o, oracle.material)
            o.setSamplesCamera(q["samples_camera"])
            o.setPixelKeyBase(base)
            o.commitScene()
            _, _, secs_r = o.renderRows(0, q["height"], frame=0, fp64=False, threads=threads)
            t_cpu += secs_r
            paths += q["width"] * q["height"] * q["samples_camera"]
            base += q["width"] * q["height"]
            done += 1
        out["cpu_baseline"] = {"value": round(paths / t_cpu / 1e6, 3), "unit": "Msamples/s", "cores": threads,
                               "kind": "port", "requests_per_s": round(done / t_cpu, 2),
                               "sample": "the first %d requests one after another (%d paths), %.1f s" % (
                                   done, paths, t_cpu)}
    return out


def pseudo_volume_bench(width=1920, height=1080, repeats=5, cpu=True):
    """The reference's pseudo-volume microbenchmark (ComputeImage,
    tests/benchmark/pseudo-volume-performance-test.cu): one SGGX pseudo-volume
    BRDF evaluation per pixel of a 1080p image, SampleCount 256, Depth 1,
    Specular 0.75, StandardDeviation (0.5, 0.5), identity basis, fixed 45
    degree directions."""
    import ctypes
    from tempestrenderer_amd import _abi
    from tempestrenderer_amd.scene import Material
    m = Material(25, diffuse=(0.0, 0.0, 0.0), specular=(0.75, 0.75, 0.75), specular_power=(0.5, 0.5),
                 sggx_basis=(0.0, 0.0, 0.0, 1.0), sggx_depth=1, sggx_sample_count=256).setup()
    img = np.zeros(width * height, dtype=np.uint32)
    probe = np.zeros(4, dtype=np.float32)
    ms = ctypes.c_float()
    _abi.check(_abi.load().tr_bench_pseudo_volume_image(
        ctypes.byref(m.rec), width, height, repeats, img.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
        probe.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(ms)))
    evals = width * height
    out = {"workload": "SGGX pseudo-volume BRDF, %dx%d evaluations, SampleCount 256, Depth 1" % (width, height),
           "ms_per_image": round(ms.value, 3), "mevals_per_s": round(evals / ms.value / 1e3, 2),
           "walks_per_s": round(evals * 256 / ms.value * 1e3, 0), "brdf": [float(v) for v in probe[:3]]}
    if cpu:
        # the oracle's FP32 evaluation of the same BRDF, one thread
        import time
        import oracle
        lib = oracle.load()
        pf = ctypes.POINTER(ctypes.c_float)
        lib.orc_bsdf_eval.argtypes = [ctypes.POINTER(_abi.tr_material), pf, pf, pf, pf, pf, ctypes.c_float, pf]
        om = oracle.material(25, dict(diffuse=(0.0, 0.0, 0.0), specular=(0.75, 0.75, 0.75), specular_power=(0.5, 0.5),
                                      sggx_basis=(0.0, 0.0, 0.0, 1.0), sggx_depth=1, sggx_sample_count=256))
        c = float(np.sqrt(0.5))
        vec = [np.array(v, dtype=np.float32) for v in ((0, 0, 1), (1, 0, 0), (0, 1, 0), (0, -c, c), (0, c, c))]
        res = np.zeros(4, dtype=np.float32)
        ptrs = [v.ctypes.data_as(pf) for v in vec]
        n = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < 2.0:
            for _ in range(200):
                lib.orc_bsdf_eval(ctypes.byref(om.rec), *ptrs, ctypes.c_float(float(probe[3])), res.ctypes.data_as(pf))
            n += 200
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(n / dt / 1e6, 5), "unit": "Mevals/s", "cores": 1, "kind": "port",
                               "sample": "%d evaluations of the same BRDF, one after another, %.1f s" % (n, dt)}
    return out


def volume_bench(cpu_rows=128, threads=None, cpu=True):
    """The heterogeneous SGGX volume: 256^3 binary16 grid (mean extinction
    4), 1024x1024, 128 camera samples, depth 32, delta tracking.  One timed
    frame; a counting frame (8 camera samples) gives tentative collisions
    per path for the gather-byte numerator (8 corners x 8 B each)."""
    threads = threads or host_threads()
    from tempestrenderer_amd import workloads
    from tempestrenderer_amd.scene import RayTracerScene
    d = workloads.volume()
    st = d.settings
    g = RayTracerScene(d.width, d.height, d.vpi)
    d.populate(g, gpu_material)
    g.commitScene()
    g_info = g.info()
    g.setSamplesCamera(8)
    g.setCounting(True)
    g.drawOnce()
    cnt = g.stats()
    g.setCounting(False)
    g.setSamplesCamera(st["samples_camera"])
    g.repaint()
    g.drawOnce()                                 # warm-up
    g.repaint()
    g.setTiming(True)
    g.drawOnce()
    tim = g.stats()
    g.setTiming(False)
    g.repaint()
    t0 = time.perf_counter()
    g.drawOnce()
    secs = time.perf_counter() - t0
    g.close()
    paths = d.width * d.height * st["samples_camera"]
    steps_per_path = cnt["volume_steps"] / max(1, cnt["paths"])
    gather = steps_per_path * paths * 64.0
    # roofline of the closest-hit + delta-tracking kernel: algorithmic bytes
    # of the counted frame (node records, primitives, 64 B of ray state and
    # hit per segment, 8 corners x 8 B per tentative collision), scaled from
    # its 8 camera samples to the timed frame's, over the timing frame's
    # summed launch time of that kernel
    scale = st["samples_camera"] / 8.0
    info = g_info
    node_bytes = info["bvh_bytes"] / max(1, info["bvh_nodes"])
    alg = scale * (node_bytes * cnt["node_visits"] + 48.0 * cnt["prim_tests"] + 64.0 * cnt["segments"] +
                   64.0 * cnt["volume_steps"])
    achieved = alg / (tim["trace_ms"] / 1e3) / 1e9
    out = {"workload": "SGGX fibre volume 256^3 binary16 (mean extinction 4), %dx%d, %d spp, depth %d, albedo 0.9, "
                       "delta tracking, env light" % (d.width, d.height, st["samples_camera"], st["max_depth"]),
           "msamples_per_s": round(paths / secs / 1e6, 1), "ms_per_frame": round(secs * 1e3, 2),
           "tentative_collisions_per_path": round(steps_per_path, 2),
           "trace_ms": round(tim["trace_ms"], 2), "shade_ms": round(tim["shade_ms"], 2),
           "gather_gbs": round(gather / (tim["trace_ms"] / 1e3) / 1e9, 1),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 5),
                        "traffic": load_traffic(os.path.join(ROOT, "profiles", "volume_kernels_pmc.txt")),
                        "kernel": "k_trace_vol<false> (closest hit + delta tracking, ray refill)",
                        "alg_bytes_per_frame": alg, "launches_per_frame": tim["trace_launches"],
                        "avg_launch_us": round(tim["trace_ms"] * 1e3 / max(1, tim["trace_launches"]), 2)}}
    if cpu:
        import oracle
        o = oracle.OracleScene(d.width, d.height, d.vpi)
        d.populate(o, oracle.material)
        o.commitScene()
        r0 = (d.height - cpu_rows) // 2
        _, _, cs = o.renderRows(r0, r0 + cpu_rows, frame=0, fp64=False, threads=threads)
        cp = cpu_rows * d.width * st["samples_camera"]
        out["cpu_baseline"] = {"value": round(cp / cs / 1e6, 3), "unit": "Msamples/s", "cores": threads, "kind": "port",
                               "sample": "rows %d-%d of the frame (%d paths), %.1f s" % (r0, r0 + cpu_rows - 1, cp, cs)}
    return out


REF_IMAGE = os.path.join(ROOT, "tests", "golden", "box_unbiased_fp64.npz")


OFF_EMITTER_TARGET = 2.5e-3


def warm_up_rmse_kernels():
    """One untimed plain frame and one untimed cached frame of the box
    workload on a throwaway scene, so that neither timed run pays the
    process's first launches (code-object loading) of the kernels it
    shares with the other."""
    from tempestrenderer_amd import workloads
    from tempestrenderer_amd.scene import RayTracerScene
    desc = workloads.box(256, 256, samples_camera=16, max_depth=4)
    desc.settings.update(estimator=1)
    sc = RayTracerScene(desc.width, desc.height, desc.vpi)
    desc.populate(sc, gpu_material)
    sc.commitScene()
    sc.drawOnce()
    sc.setRadianceCache(True, train=True, train_paths=65536, batch=16384, rough_alpha=0.2, seed=1)
    sc.drawOnce()
    sc.close()


def time_to_rmse(cached, target=1e-2, off_target=OFF_EMITTER_TARGET, max_seconds=30.0, max_frames=400,
                 warmup_frames=8, frame_spp=16, query_nee=False):
    """Wall time of drawOnce calls until the de-biased RMSE of the running
    image against the converged FP64 CPU render of the box workload
    (256 x 256, depth 4, unbiased estimator; tests/golden/make_converged.py)
    reaches `target`: RMSE^2 = mean over pixels and RGB of (I - I_ref)^2,
    de-biased by the reference's own squared standard error.  The same is
    tracked over the off-emitter pixels (off_emitter_mask) against
    `off_target`.  The cached run first trains for `warmup_frames` frames
    whose image is discarded; that time counts."""
    from tempestrenderer_amd import workloads
    from tempestrenderer_amd.scene import RayTracerScene
    ref = np.load(REF_IMAGE)
    ref_mean, ref_se = ref["mean"].astype(np.float64), ref["stderr"].astype(np.float64)
    mask = workloads.off_emitter_mask(ref_mean)
    ref_var = float(np.mean(ref_se * ref_se))
    ref_var_off = float(np.mean((ref_se * ref_se)[mask]))
    desc = workloads.box(256, 256, samples_camera=frame_spp, max_depth=4)
    desc.settings.update(estimator=1)
    sc = RayTracerScene(desc.width, desc.height, desc.vpi)
    desc.populate(sc, gpu_material)
    sc.commitScene()
    elapsed = 0.0
    if cached:
        sc.setRadianceCache(True, train=True, train_paths=65536, batch=16384, rough_alpha=0.2, seed=1,
                            query_nee=query_nee)
        for _ in range(warmup_frames):
            t0 = time.perf_counter()
            sc.drawOnce()
            elapsed += time.perf_counter() - t0
        sc.repaint()
    frames = 0
    rmse = rmse_off = float("inf")
    hit = {"full": None, "off": None}
    trace = []
    while elapsed < max_seconds and frames < max_frames:
        t0 = time.perf_counter()
        sc.drawOnce()
        elapsed += time.perf_counter() - t0
        frames += 1
        acc, n = sc.accumulation()
        err2 = (acc / n - ref_mean) ** 2
        rmse = math.sqrt(max(float(np.mean(err2)) - ref_var, 0.0))
        rmse_off = math.sqrt(max(float(np.mean(err2[mask])) - ref_var_off, 0.0))
        trace.append((round(elapsed, 4), n, round(rmse, 5), round(rmse_off, 6)))
        if hit["full"] is None and rmse <= target:
            hit["full"] = (round(elapsed, 4), frames)
        if hit["off"] is None and rmse_off <= off_target:
            hit["off"] = (round(elapsed, 4), frames)
        if hit["full"] and hit["off"]:
            break
    out = {"seconds": hit["full"][0] if hit["full"] else None,
           "spp": hit["full"][1] * frame_spp if hit["full"] else None,
           "off_emitter": {"target_rmse": off_target, "pixels": int(mask.sum()),
                           "seconds": hit["off"][0] if hit["off"] else None,
                           "spp": hit["off"][1] * frame_spp if hit["off"] else None,
                           "final_rmse": round(rmse_off, 6)},
           "frames": frames, "final_rmse": round(rmse, 5),
           "ms_per_frame": round(elapsed * 1e3 / max(1, frames + (warmup_frames if cached else 0)), 3)}
    if cached:
        out["warmup_frames"] = warmup_frames
        out["query_nee"] = bool(query_nee)
        out["last_loss"] = round(sc.stats()["cache_loss"], 5)
    sc.close()
    return out, trace


def calibrate_copy(mib=512, reps=4):
    """A plain device-to-device copy of known size, run under
    `rocprofv3 --pmc FETCH_SIZE` / `WRITE_SIZE` to scale those counters into
    bytes: each copy reads and writes exactly `mib` MiB."""
    import torch
    n = mib << 20
    src = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--small", action="store_true", help="reduced scene for a quick check (not the metric)")
    ap.add_argument("--cpu-rows", type=int, default=1080, help="rows of the frame the CPU oracle renders (capped at the height)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-rmse", action="store_true", help="skip the time-to-RMSE runs (plain vs radiance cache)")
    ap.add_argument("--rmse-only", action="store_true", help="only the time-to-RMSE runs, printed with their traces")
    ap.add_argument("--no-gonio", action="store_true", help="skip the batched gonioreflectometer run")
    ap.add_argument("--gonio-only", action="store_true", help="only the batched gonioreflectometer run")
    ap.add_argument("--no-volume", action="store_true", help="skip the volume run")
    ap.add_argument("--no-pseudo-volume", action="store_true", help="skip the pseudo-volume BRDF microbenchmark")
    ap.add_argument("--volume-only", action="store_true", help="only the volume run")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 512 MiB device copy (HBM counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the tracer runs on a single MI355X in this build "
              "(multi-GPU sharding by sample index is described, not built).", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    args.steps = max(1, args.steps)
    if args.gonio_only:
        log(json.dumps(gonio_bench(cpu=not args.no_cpu)))
        return
    if args.volume_only:
        log(json.dumps(volume_bench(cpu=not args.no_cpu)))
        return
    if args.rmse_only:
        warm_up_rmse_kernels()
        for cached, qn in ((False, False), (True, False), (True, True)):
            res, trace = time_to_rmse(cached, query_nee=qn)
            log(("cached query_nee=%d" % qn) if cached else "plain", json.dumps(res))
            log("  trace (seconds, spp, rmse, off-emitter rmse):", trace)
        return

    from tempestrenderer_amd import workloads
    from tempestrenderer_amd.scene import RayTracerScene

    if args.small:
        desc = workloads.heightfield(480, 270, samples_camera=16, grid=256)
    else:
        desc = workloads.heightfield()
    st = desc.settings
    t0 = time.time()
    gpu = RayTracerScene(desc.width, desc.height, desc.vpi)
    desc.populate(gpu, gpu_material)
    gpu.commitScene()
    info = gpu.info()
    log("scene: %d prims, %d BVH8 nodes (%.1f MB), depth %d, build %.0f ms, setup %.1f s" % (
        info["primitives"], info["bvh_nodes"], info["bvh_bytes"] / 1e6, info["bvh_depth"], info["build_ms"],
        time.time() - t0))

    # counting frame (not timed): node / primitive visits for the roofline numerator
    gpu.setCounting(True)
    gpu.drawOnce()
    cnt = gpu.stats()
    gpu.setCounting(False)
    gpu.repaint()
    log("counting frame: %d segments, %.2f nodes/seg, %.2f prims/seg" % (
        cnt["segments"], cnt["node_visits"] / max(1, cnt["segments"]), cnt["prim_tests"] / max(1, cnt["segments"])))

    # timing frame (not part of the throughput steps): every kernel bracketed
    # by HIP events on the tracer's stream and waited for
    gpu.setTiming(True)
    gpu.drawOnce()
    tim = gpu.stats()
    gpu.setTiming(False)
    gpu.repaint()
    log("timing frame: trace %.2f ms over %d launches, shade %.2f ms, shadow %.2f ms, other %.2f ms" % (
        tim["trace_ms"], tim["trace_launches"], tim["shade_ms"], tim["shadow_ms"], tim["other_ms"]))

    for _ in range(args.warmup):
        gpu.drawOnce()
    t_start = time.perf_counter()
    for i in range(args.steps):
        gpu.drawOnce()
        log("step %d: %.1f ms" % (i, gpu.stats()["frame_ms"]))
    elapsed = time.perf_counter() - t_start
    samples = desc.pixels * st["samples_camera"]
    value = samples * args.steps / elapsed / 1e6
    ms_per_step = elapsed * 1e3 / args.steps

    segs = cnt["segments"]
    node_bytes = info["bvh_bytes"] / max(1, info["bvh_nodes"])
    alg_bytes = node_bytes * cnt["node_visits"] + 48.0 * cnt["prim_tests"] + 64.0 * segs
    trace_s = tim["trace_ms"] / 1e3
    achieved = alg_bytes / trace_s / 1e9 if trace_s > 0 else 0.0
    traffic = load_traffic(os.path.join(ROOT, "profiles", "trace_kernel_heightfield_pmc.txt"))
    acc, spp = gpu.accumulation()
    img = acc / spp
    log("image mean %s, finite %s" % (img.mean(axis=(0, 1)), bool(np.isfinite(img).all())))

    ttr = None
    if not args.no_rmse and os.path.exists(REF_IMAGE):
        warm_up_rmse_kernels()
        plain, _ = time_to_rmse(False)
        log("time-to-RMSE plain: %s" % json.dumps(plain))
        cached, _ = time_to_rmse(True, query_nee=True)
        log("time-to-RMSE cached: %s" % json.dumps(cached))
        ref = np.load(REF_IMAGE)
        ttr = {"workload": "box 256x256, depth 4, unbiased estimator, 16 camera samples per frame",
               "target_rmse": 0.01, "reference": "FP64 CPU oracle render, %d spp, seed %d" % (int(ref["spp"]),
                                                                                           int(ref["seed"])),
               "plain": plain, "cached": cached,
               "speedup": (round(plain["seconds"] / cached["seconds"], 3)
                           if plain["seconds"] and cached["seconds"] else None),
               "speedup_off_emitter": (round(plain["off_emitter"]["seconds"] / cached["off_emitter"]["seconds"], 3)
                                       if plain["off_emitter"]["seconds"] and cached["off_emitter"]["seconds"]
                                       else None)}

    pv = None
    if not args.no_pseudo_volume:
        pv = pseudo_volume_bench(cpu=not args.no_cpu)
        log("pseudo_volume_brdf: %s" % json.dumps(pv))

    vol = None
    if not args.no_volume:
        vol = volume_bench(cpu=not args.no_cpu, threads=host_threads())
        log("volume: %s" % json.dumps(vol))

    gonio = None
    if not args.no_gonio:
        gonio = gonio_bench(cpu=not args.no_cpu, threads=host_threads())
        log("gonioreflectometer: %s" % json.dumps(gonio))

    cpu = None
    if not args.no_cpu:
        threads = host_threads()
        args.cpu_rows = min(args.cpu_rows, desc.height)
        cpu_value, cpu_secs, cpu_samples, r0 = cpu_baseline(desc, args.cpu_rows, threads)
        cpu = {"value": round(cpu_value, 4), "unit": "Msamples/s", "cores": threads, "kind": "port",
               "sample": "rows %d-%d of the same frame (%d px x %d camera samples = %d paths), %.1f s wall" % (
                   r0, r0 + args.cpu_rows - 1, args.cpu_rows * desc.width, st["samples_camera"], cpu_samples,
                   cpu_secs)}

    out = {
        "metric": "Msamples/s (plain path tracer, heightfield scene, 1 GPU)",
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded fBm heightfield + 64 spheres, Philox seed 1)",
        "config": {
            "workload": "heightfield %dx%d grid (%d tris) + 64 spheres, %dx%d, %d spp, depth %d, env light" % (
                1024 if not args.small else 256, 1024 if not args.small else 256, info["primitives"] - 64,
                desc.width, desc.height, st["samples_camera"], st["max_depth"]),
            "samples_per_step": samples,
            "estimator": "reference",
        },
        "mrays_per_s": round(segs / (ms_per_step / 1e3) / 1e6, 2),
        "kernel_ms_per_frame": {"trace": round(tim["trace_ms"], 3), "shade": round(tim["shade_ms"], 3),
                                "shadow": round(tim["shadow_ms"], 3), "raygen_accumulate": round(tim["other_ms"], 3)},
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "kernel": "closest-hit BVH8 traversal: k_trace_packet (camera rays) + k_trace_refill (later vertices)",
            "alg_bytes_per_step": alg_bytes,
            "segments_per_step": segs,
            "launches_per_frame": tim["trace_launches"],
            "avg_launch_us": round(tim["trace_ms"] * 1e3 / max(1, tim["trace_launches"]), 2),
            # HBM bytes the counters measured per launch over the launch time: the
            # algorithmic bytes count every node fetch per ray, while the camera
            # rays of one wave share theirs through the caches
            "traffic_frac": (round(traffic / (tim["trace_ms"] * 1e-3 / max(1, tim["trace_launches"])) / 1e9
                                   / HBM_PEAK_GBS, 5) if traffic and tim["trace_ms"] > 0 else None),
        },
        "cpu_baseline": cpu,
        "time_to_rmse": ttr,
        "gonioreflectometer": gonio,
        "volume": vol,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
