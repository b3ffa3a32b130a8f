"""Benchmark: micropolygon-sample tests/s and frames/s of the HIP hider.

Workload (default "hd"): the plain hider at HD -- 1920x1080, 16x16 buckets,
PixelSamples 4x4, catmull-rom 3x3, ShadingRate 1, 10,000 bicubic Bezier
patches (seed 1, control-point jitter N(0, 0.1) of the patch size), opaque,
constant Cs per patch.  A "step" is one frame on the device with the scene's
grids and the frame's sample plan resident in HBM: dice + bust + bin +
sample + combine + filter + expose + quantise.  The host split/dice schedule
and upload run once before the timed region and are reported separately
(`host_schedule_ms`, `e2e_ms_per_frame`).

The CPU baseline is the oracle (a C++ restatement of the reference hider,
occlusion culling on as in the reference) on the host cores: the whole frame
in one crop band per thread, and one thread on a quarter of the frame.

Other workloads: "clustered" (2048x2048, micropolygons per bucket log-normal),
"motion" (256x256, PixelSamples 8x8, transparent deforming patches, shutter
0-1), "spheres" (256x256, 64 spheres) and "serving" (a Poisson request
stream).

The default run (the HD workload) adds short runs of the motion, clustered
and serving configurations in "other_workloads" (--no-workloads or
--no-targets skips them).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload hd|clustered|motion|spheres|serving]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published peaks
PEAK_FP32_TFLOPS = 157.3        # vector FP32 (packed FMA counted as 4 flops/lane-cycle)
PEAK_VALU_TLANEOPS = 39.3       # 256 CU x 64 lanes x 2.4 GHz, one plain FP32 op per lane-cycle
PEAK_HBM_GBS = 8000.0


def workload(name):
    from aqsis_amd import scenes
    if name == "hd":
        opts, prims = scenes.patches_scene(seed=1, n=10000, res=(1920, 1080))
        desc = {"workload": "plain hider HD: 1920x1080, 16x16 buckets, PixelSamples 4x4, catmull-rom 3x3, "
                            "ShadingRate 1, 10000 bicubic Bezier patches (seed 1), opaque",
                "width": 1920, "height": 1080, "pixel_samples": [4, 4], "filter": ["catmull-rom", 3, 3],
                "patches": 10000, "bucket": [16, 16]}
    elif name == "clustered":
        opts, prims = scenes.clustered_scene(seed=1)
        desc = {"workload": "batched hider: 2048x2048, 16x16 buckets, PixelSamples 4x4, gaussian 2x2, ShadingRate 1, "
                            "bicubic patches clustered so micropolygons per bucket ~ LogNormal(ln 2000, 1.5) "
                            "clipped at 200k (seed 1), opaque",
                "width": 2048, "height": 2048, "pixel_samples": [4, 4], "filter": ["gaussian", 2, 2],
                "patches": len(prims), "bucket": [16, 16]}
    elif name == "motion":
        opts, prims = scenes.motion_scene(seed=1, n=600, res=(256, 256))
        desc = {"workload": "plain hider with transparency and motion blur: 256x256, 16x16 buckets, PixelSamples 8x8, "
                            "gaussian 2x2, ShadingRate 1, 600 bicubic Bezier patches (seed 1), Os uniform in "
                            "[0.2,0.8] per patch, two motion keys at times 0 and 1 (deforming surfaces moved by "
                            "N(0, 0.05 x width) pixels per axis), shutter 0-1",
                "width": 256, "height": 256, "pixel_samples": [8, 8], "filter": ["gaussian", 2, 2],
                "patches": 600, "bucket": [16, 16], "motion_keys": 2, "shutter": [0, 1]}
    elif name == "spheres":
        opts, prims = scenes.spheres_scene(seed=1)
        desc = {"workload": "plain hider CPU-runnable case: 256x256, 16x16 buckets, PixelSamples 4x4, "
                            "gaussian 2x2, 64 spheres (seed 1)",
                "width": 256, "height": 256, "pixel_samples": [4, 4], "filter": ["gaussian", 2, 2],
                "spheres": 64, "bucket": [16, 16]}
    else:
        raise SystemExit(f"unknown workload {name}")
    return opts, prims, desc


def algorithmic_ops(st, motion=False):
    """FP32 lane-operations of the sampling kernel, counted from the
    reference's arithmetic.  Static micropolygons: bound, occlusion and LOD
    tests, 7 per candidate; four edge functions, 24 per bound hit; one Newton
    step + bilinear depth + store test, 53 per point-in-quad hit; a second
    Newton step, 39.  Moving micropolygons (RenderMPG_MBOrDof and the motion
    Sample): the time window adds 2 per visit; a bound hit adds the key span
    (5), the tight bound at the sample's time (16), the vertices at that time
    (36) and the edge and bilinear set-up (22) to the 24 of the edges."""
    if motion:
        return (9 * st["spl_count"] + 103 * st["spl_bound_hits"] + 53 * st["inside"]
                + 39 * st["inside_nonlinear"])
    return (7 * st["spl_count"] + 24 * st["spl_bound_hits"] + 53 * st["inside"]
            + 39 * st["inside_nonlinear"])


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` at the HD workload, as measured by
    rocprofv3 FETCH_SIZE/WRITE_SIZE (scaled by the copy-kernel calibration)
    and recorded in profiles/hd_patches_frame_kernels.md; None if absent."""
    path = os.path.join(ROOT, "profiles", "hd_patches_frame_kernels.md")
    try:
        for line in open(path):
            if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def host_cpus():
    """The host cores this process may use: the affinity set, narrowed by a
    cgroup CPU quota when one is set; with the count, nproc and the CPU
    model name (/proc/cpuinfo, the lscpu "Model name")."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    threads = aff if quota is None else max(1, min(aff, int(quota)))
    return {"threads": threads, "nproc": os.cpu_count(), "affinity": aff, "cgroup_cpu_quota": quota,
            "model": model}


def cpu_baseline(opts, prims, quarter_single=True):
    """The oracle (a C++ restatement of the reference hider, occlusion
    culling on as in the reference) on the host cores: (i) all the cores
    this process may use, the whole frame cut into one horizontal CropWindow
    band per thread (ctypes releases the GIL, the oracle has no shared
    state), timed wall to wall -- the reported value; (ii) one thread, as the
    reference's single-threaded bucket loop runs, on the CropWindow
    [0,0.5]x[0,0.5] quarter of the frame (or the whole frame when
    quarter_single is False)."""
    from oracle import loader
    import copy
    import threading
    cpus = host_cpus()
    o = copy.copy(opts)
    if quarter_single:
        o.crop[:] = [0.0, 0.5, 0.0, 0.5]
    loader.render(o, prims, occlusion_cull=True, samples=False)   # warm-up / page-in
    t0 = time.perf_counter()
    r = loader.render(o, prims, occlusion_cull=True, samples=False)
    wall1 = time.perf_counter() - t0
    st = r["stats"]
    one_frame_s = st["ms_device"] / 1e3
    one_rate = st["spl_count"] / one_frame_s
    threads = cpus["threads"]
    bands = []
    for i in range(threads):
        ob = copy.copy(opts)
        ob.crop[:] = [0.0, 1.0, i / threads, (i + 1) / threads]
        bands.append(ob)
    results = [None] * threads

    def work(i):
        results[i] = loader.render(bands[i], prims, occlusion_cull=True, samples=False)["stats"]

    t0 = time.perf_counter()
    pool = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
    for t in pool:
        t.start()
    for t in pool:
        t.join()
    wall_all = time.perf_counter() - t0
    tests_all = sum(x["spl_count"] for x in results)
    part = "the CropWindow [0,0.5]x[0,0.5] quarter" if quarter_single else "the whole frame"
    return {
        "value": tests_all / wall_all,
        "unit": "sample-tests/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle (C++ restatement of the reference hider, occlusion culling on): the whole frame as "
                  f"{threads} horizontal CropWindow bands, one thread each, wall time; one thread on {part} beside it",
        "host": cpus,
        "frames_per_s": 1.0 / wall_all,
        "wall_s": wall_all,
        "single_thread": {"value": one_rate, "frame_ms": st["ms_device"],
                          "frames_per_s_extrapolated": (0.25 if quarter_single else 1.0) / one_frame_s,
                          "spl_count": st["spl_count"], "wall_s": wall1, "part": part},
    }


def e2e_frames(h, prims, steps):
    """A new scene per frame on a resident context: surfaces posted, the
    split/dice schedule (occlusion culling included), upload and the device
    frame, timed wall to wall.  Returns the median and the parts of the last."""
    times, parts = [], None
    for _ in range(steps):
        h.clear_surfaces()
        t0 = time.perf_counter()
        h.post_surfaces(prims)
        h.render()
        times.append((time.perf_counter() - t0) * 1e3)
        st = h.stats()
        parts = {"schedule_ms": st["ms_schedule"], "cull_ms": st["ms_cull"], "upload_ms": st["ms_upload"],
                 "device_ms": st["ms_device"], "cull_passes": st["cull_passes"], "cull_probes": st["cull_probes"],
                 "forest_ms": st["ms_forest"], "simulate_ms": st["ms_simulate"],
                 "cull_depth_ms": st["ms_cull_depth"], "cull_probes_ms": st["ms_cull_probes"]}
    times.sort()
    med = times[len(times) // 2]
    return {"ms_per_frame": med, "frames_per_s": 1e3 / med, "steps": steps, "last": parts}


def config1_target(args):
    """The 20x target on the CPU-runnable configuration (256x256, 64
    spheres, 4x4, gaussian 2x2): the device hider against the oracle on all
    host cores, end to end (new scene per frame) and on resident inputs, with
    the frame's parity checked in the same run."""
    import numpy as np
    from aqsis_amd import scenes
    from aqsis_amd.hider import Hider
    from oracle import loader
    opts, prims = scenes.spheres_scene(seed=1)
    with Hider(opts) as h:
        h.post_surfaces(prims)
        h.render()
        ids, depth, _ = h.samples()
        ch, rgba = h.channels(), h.rgba8()
        e2e = e2e_frames(h, prims, max(3, args.e2e_steps))
        for _ in range(3):
            h.render_device()
        t0 = time.perf_counter()
        n = 20
        for _ in range(n):
            h.render_device()
        dev_fps = n / (time.perf_counter() - t0)
    ref = loader.render(opts, prims, occlusion_cull=True)
    exact = bool(np.array_equal(ids, ref["mp_id"]) and np.array_equal(depth.view(np.uint32), ref["depth"].view(np.uint32)))
    err = float(np.max(np.abs(ch.astype(np.float64) - ref["channels"])))
    cb = cpu_baseline(opts, prims, quarter_single=False)
    return {"config": "256x256, 16x16 buckets, PixelSamples 4x4, gaussian 2x2, 64 spheres (seed 1)",
            "gpu_frames_per_s_e2e": e2e["frames_per_s"], "gpu_frames_per_s_device": dev_fps,
            "cpu_frames_per_s_all_cores": cb["frames_per_s"], "cpu_cores": cb["cores"],
            "cpu_frames_per_s_1thread": 1.0 / cb["single_thread"]["wall_s"],
            "ratio_e2e_vs_cpu_all_cores": e2e["frames_per_s"] / cb["frames_per_s"],
            "ratio_device_vs_cpu_all_cores": dev_fps / cb["frames_per_s"],
            "ratio_e2e_vs_cpu_1thread": e2e["frames_per_s"] * cb["single_thread"]["wall_s"],
            "visibility_bit_exact": exact, "max_pixel_error": err,
            "rgba8_identical": bool(np.array_equal(rgba, ref["rgba8"])), "e2e_parts": e2e["last"]}


def serving(args):
    """The serving form: a Poisson stream of render requests of mixed side
    (256..2048), PixelSamples (2x2, 4x4, 8x8), patch count (100..20,000) and
    opacity (0.2..1) submitted to one rh_server as they arrive; the server's
    scheduler threads compute each job's split/dice schedule, and jobs run
    concurrently on the GPU, admitted by device-memory footprint.  Latency is
    client-side, arrival to result."""
    import threading
    from aqsis_amd import scenes
    from aqsis_amd.hider import Server
    reqs = list(scenes.serving_requests(seed=1, count=args.requests, rate_hz=args.rate))
    lat, infos, errs = [None] * len(reqs), [None] * len(reqs), []
    with Server(slots=args.slots, memory_budget=args.budget_gb << 30) as srv:
        # One thread submits each request at its arrival time (submission
        # returns at once); this thread collects finished jobs by polling.
        ids = {}
        lock = threading.Lock()
        t_start = time.perf_counter()

        def arrivals():
            for k, (opts, prims, arrival) in enumerate(reqs):
                delay = t_start + arrival - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                jid = srv.submit(opts, prims)
                with lock:
                    ids[jid] = k

        sub = threading.Thread(target=arrivals)
        sub.start()
        finished = 0
        while finished < len(reqs):
            got = srv.poll()
            now = time.perf_counter()
            for info in got:
                with lock:
                    k = ids[info["id"]]
                finished += 1
                if finished % 100 == 0:
                    print(f"serving: {finished} of {len(reqs)} jobs done at {now - t_start:.1f} s", file=sys.stderr,
                          flush=True)
                if info["status"] != 0:
                    try:
                        srv.result(info["id"], info["width"], info["height"])
                        msg = ""
                    except Exception as e:          # the job's own error message
                        msg = str(e)
                    errs.append(f"job {k}: status {info['status']} {msg}")
                else:
                    lat[k] = (now - (t_start + reqs[k][2])) * 1e3   # from the scheduled arrival
                    infos[k] = info
                srv.release(info["id"])
            if not got:
                time.sleep(0.001)
        sub.join()
        wall = time.perf_counter() - t_start
        resident, high = srv.memory()
    done = [i for i in range(len(reqs)) if infos[i] is not None]
    ls = sorted(lat[i] for i in done)
    tests = sum(infos[i]["spl_count"] for i in done)
    dev = sum(infos[i]["ms_device"] for i in done)
    pct = lambda q: ls[min(len(ls) - 1, int(q * len(ls)))] if ls else None
    mean = lambda v: sum(v) / len(v) if v else None
    return {
        "metric": "micropolygon-sample tests/s",
        "value": tests / wall,
        "unit": "sample-tests/s",
        "n_gpus": 1,
        "steps": len(done),
        "warmup": 0,
        "ms_per_step": wall * 1e3 / max(1, len(done)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "serving: Poisson stream of render requests, side in {256,512,1024,2048}, "
                               "PixelSamples in {2x2,4x4,8x8}, 100-20,000 bicubic patches, Os in [0.2,1], "
                               "gaussian 2x2", "requests": len(reqs), "rate_hz": args.rate, "slots": args.slots,
                   "memory_budget_gb": args.budget_gb},
        "frames_per_s": len(done) / wall,
        "latency_ms_p50": pct(0.5),
        "latency_ms_p99": pct(0.99),
        "device_busy_fraction": dev / (wall * 1e3),
        # where a request's time goes (means over the finished ones): the
        # split/dice schedule and the wait for admission, the run (context
        # set-up, upload, frame, read-back) and the device frame
        "mean_ms": {"queue_incl_schedule": mean([infos[i]["ms_queue"] for i in done]),
                    "run": mean([infos[i]["ms_run"] for i in done]),
                    "run_setup": mean([infos[i]["ms_setup"] for i in done]),
                    "run_upload": mean([infos[i]["ms_upload"] for i in done]),
                    "run_render": mean([infos[i]["ms_render"] for i in done]),
                    "run_readback": mean([infos[i]["ms_readback"] for i in done]),
                    "device": mean([infos[i]["ms_device"] for i in done])},
        "contexts_created": sum(1 for i in done if infos[i]["ms_setup"] > 1.0),
        "device_memory_high_water_gb": high / 2 ** 30,
        "failed": len(errs),
        "errors": errs[:5],
    }


def other_workloads(args):
    """The other configurations, short, beside the default line: the motion
    and clustered workloads (first frame with the host schedule, then device
    frames on resident inputs in the batched form) and a 200-request serving
    stream at the default rate."""
    from aqsis_amd.hider import Hider
    res = {}
    for name in ("motion", "clustered"):
        opts, prims, desc = workload(name)
        with Hider(opts) as h:
            h.post_surfaces(prims)
            t0 = time.perf_counter()
            h.render()
            first = (time.perf_counter() - t0) * 1e3
            h.set_tuning(mode="batched")
            for _ in range(2):
                h.render_device()
            n = 5
            t0 = time.perf_counter()
            for _ in range(n):
                h.render_device()
            wall = time.perf_counter() - t0
            st = h.stats()
        res[name] = {"config": desc["workload"], "ms_per_frame": wall * 1e3 / n, "frames_per_s": n / wall,
                     "sample_tests_per_s": st["spl_count"] * n / wall, "spl_count": st["spl_count"],
                     "first_frame_ms_with_schedule": first, "mode": "batched (216-record chunks, 4 blocks a wave)"}
    sargs = argparse.Namespace(**vars(args))
    sargs.requests = 200
    sv = serving(sargs)
    res["serving"] = {"config": sv["config"]["workload"], "requests": 200, "rate_hz": sargs.rate,
                      "frames_per_s": sv["frames_per_s"], "latency_ms_p50": sv["latency_ms_p50"],
                      "latency_ms_p99": sv["latency_ms_p99"], "sample_tests_per_s": sv["value"],
                      "failed": sv["failed"]}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="hd")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--mode", default="auto", choices=["auto", "grid", "plain", "batched"],
                    help="device layout: auto = autotuned batched form")
    ap.add_argument("--plain-steps", type=int, default=3,
                    help="frames of the one-launch-per-bucket form timed beside (0: skip)")
    ap.add_argument("--requests", type=int, default=2000, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=20.0,
                    help="serving: Poisson arrival rate (1/s); 20/s keeps the GPU about 90%% busy")
    ap.add_argument("--slots", type=int, default=4, help="serving: jobs in flight")
    ap.add_argument("--budget-gb", type=int, default=160, help="serving: device memory budget")
    ap.add_argument("--e2e-steps", type=int, default=5, help="new-scene frames timed end to end")
    ap.add_argument("--no-targets", action="store_true",
                    help="skip the CPU-runnable configuration's target line (and the other configurations' runs)")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the other configurations' short runs beside the default workload")
    ap.add_argument("--copy-kernel", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_kernel:
        import torch
        n = args.copy_kernel // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4}))
        return
    if args.gpus != 1:
        print("bench.py: measures one GPU only; the hider's multi-GPU sharding "
              "(screen-space bucket rows per GPU) is a design note, not code.", file=sys.stderr)
        raise SystemExit(2)

    if args.workload == "serving":
        print(json.dumps(serving(args)))
        return
    from aqsis_amd.hider import Hider
    opts, prims, desc = workload(args.workload)
    h = Hider(opts)
    h.post_surfaces(prims)
    t0 = time.perf_counter()
    h.render()                     # host schedule + upload + first frame
    first_ms = (time.perf_counter() - t0) * 1e3
    st0 = h.stats()
    # per-stage times of the static-grid form, one frame
    grid_stages = {"dice_bust": st0["ms_dice"], "bin": st0["ms_bin"], "sample": st0["ms_sample"],
                   "filter_quantise": st0["ms_filter"]}
    if args.mode == "auto":
        tuning = h.autotune()
    else:
        h.set_tuning(mode=args.mode)
        tuning = h.tuning()
    for _ in range(args.warmup):
        h.render_device()
    dev_ms, sample_ms, dice_ms, bin_ms, filt_ms, comp_ms = [], [], [], [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        h.render_device()          # returns after the frame's last kernel completed
        st = h.stats()
        dev_ms.append(st["ms_device"]); sample_ms.append(st["ms_sample"])
        dice_ms.append(st["ms_dice"]); bin_ms.append(st["ms_bin"]); filt_ms.append(st["ms_filter"])
        comp_ms.append(st["ms_composite"])
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    fps = args.steps / wall
    tests_per_s = st["spl_count"] * args.steps / wall
    dev_sorted = sorted(dev_ms)
    motion = args.workload == "motion"
    ops = algorithmic_ops(st, motion)
    if tuning["mode"] == 2:
        # The batched form replays a frame graph, which carries no stage
        # timestamps: the same launches issued one by one give the stage times.
        dice_ms, bin_ms, sample_ms, filt_ms, comp_ms = [], [], [], [], []
        for _ in range(3):
            h.render_stages()
            sst = h.stats()
            dice_ms.append(sst["ms_dice"]); bin_ms.append(sst["ms_bin"])
            sample_ms.append(sst["ms_sample"]); filt_ms.append(sst["ms_filter"])
            comp_ms.append(sst["ms_composite"])
    sample_avg_ms = sum(sample_ms) / len(sample_ms)
    # the sampler's launch time: HIP events on its stream around it (the
    # sampling stage less the transparent-hit combine that follows it)
    comp_avg_ms = sum(comp_ms) / len(comp_ms)
    roof_ms = sample_avg_ms - comp_avg_ms
    achieved_tops = ops / (roof_ms * 1e-3) / 1e12
    out = {
        "metric": "micropolygon-sample tests/s",
        "value": tests_per_s,
        "unit": "sample-tests/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": desc,
        "frames_per_s": fps,
        "frames_per_s_note": "device frames on resident inputs (schedule and upload outside); "
                             "frames_per_s_e2e has them in",
        "frame_ms_p50": dev_sorted[len(dev_sorted) // 2],
        "frame_ms_p99": dev_sorted[min(len(dev_sorted) - 1, int(0.99 * len(dev_sorted)))],
        "stage_ms": {"dice_bust": sum(dice_ms) / len(dice_ms), "bin": sum(bin_ms) / len(bin_ms),
                     "sample": sample_avg_ms, "filter_quantise": sum(filt_ms) / len(filt_ms),
                     "composite_in_sample": comp_avg_ms},
        "stage_ms_grid_first_frame": grid_stages,
        "tuning": {"mode": ["grid", "plain", "batched"][tuning["mode"]], "lds_records": tuning["lds_records"],
                   "blocks_per_wave": tuning["blocks_per_wave"], "split_cost": tuning["split_cost"]},
        "host_schedule_ms": st0["ms_schedule"],
        "host_upload_ms": st0["ms_upload"],
        "e2e_ms_per_frame": st0["ms_schedule"] + st0["ms_upload"] + ms_per_step,
        "first_frame_ms": first_ms,
        "counts": {k: st[k] for k in ("spl_count", "spl_bound_hits", "spl_hits", "inside", "inside_nonlinear",
                                      "grids", "micropolygons", "mp_bucket_refs", "occlusion_culled")},
        "counts_note": "occlusion culling on, as the reference's default: SPL_count in the reference's units",
        "roofline": {
            "bound": "valu",
            "kernel": "rh_sample_mb" if motion else "rh_sample",
            "achieved": achieved_tops,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tops / PEAK_FP32_TFLOPS,
            "peak_plain_valu": PEAK_VALU_TLANEOPS,
            "frac_plain_valu": achieved_tops / PEAK_VALU_TLANEOPS,
            "algorithmic_ops_per_launch": ops,
            "launch_ms": roof_ms,
            "traffic": profiled_traffic("rh_sample") if args.workload == "hd" else None,
            "traffic_source": "rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0 (copy-kernel calibrated), "
                              "profiles/hd_patches_frame_kernels.md",
        },
    }
    if args.plain_steps > 0:
        # The one-bucket-per-launch form on the same frame, for comparison.
        h.set_tuning(mode="plain")
        h.render_device()
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            h.render_device()
        plain_wall = time.perf_counter() - t0
        out["plain_form"] = {"frames_per_s": args.plain_steps / plain_wall,
                             "ms_per_frame": plain_wall * 1e3 / args.plain_steps}
        out["speedup_vs_plain_form_frames"] = fps / out["plain_form"]["frames_per_s"]
    if args.e2e_steps > 0:
        h.set_tuning(mode=["grid", "plain", "batched"][tuning["mode"]], lds_records=tuning["lds_records"],
                     blocks_per_wave=tuning["blocks_per_wave"], split_cost=tuning["split_cost"])
        e2e = e2e_frames(h, prims, args.e2e_steps)
        out["frames_per_s_e2e"] = e2e["frames_per_s"]
        out["e2e"] = e2e
    if not args.no_cpu:
        cb = cpu_baseline(opts, prims)
        out["cpu_baseline"] = cb
        sp = {"sample_tests": tests_per_s / cb["value"],
              "frames_device": fps / cb["frames_per_s"],
              "sample_tests_vs_1thread": tests_per_s / cb["single_thread"]["value"]}
        if "frames_per_s_e2e" in out:
            sp["frames_e2e"] = out["frames_per_s_e2e"] / cb["frames_per_s"]
        out["speedup_vs_cpu"] = sp
        out["speedup_note"] = ("both sides count sample tests with occlusion culling on; frames_e2e puts the "
                               "host schedule and upload into the device side")
    h.close()
    if not args.no_targets and not args.no_cpu:
        t = {"config1": config1_target(args)}
        if "speedup_vs_plain_form_frames" in out:
            t["batched_vs_plain_frames"] = out["speedup_vs_plain_form_frames"]
        out["targets"] = t
    if args.workload == "hd" and not args.no_workloads and not args.no_targets:
        out["other_workloads"] = other_workloads(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
