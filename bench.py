"""Benchmark of the MI355X LiDAR hot path (one JSON line on stdout).

Workload: synthetic 64x2048 HDL-64E-like scans of the full single-frame
pipeline configuration (40 cars, 60 poles, 2000 rain points, ~128k points),
each frame through the DROR noise filter, segmentation (projection, RECM,
RANSAC, JCP) of the kept points, the obstacle split, curved-voxel
clustering, per-cluster convex hulls and a rotating-calipers box per hull.
A step is one batch of B frames already resident in HBM, run as one
captured hipGraph.  `value` is frames/s of the batched form; the plain
per-frame path (B = 1, one graph per frame) and the host-CPU oracle are
reported beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "lidar-processing-v2_amd")]

import numpy as np  # noqa: E402

METRIC = "frames/s + points/s at 64x2048 (batched vs plain GPU vs CPU); p50/p99 ms"
# Profiled stages that consist of exactly one kernel launch.
SINGLE_KERNEL_STAGES = {"noise_count": "k_noise_count", "noise_search": "k_noise_search", "project": "k_project",
                        "recm": "k_recm",
                        "pixels": "k_pixel_label", "jcp_slots": "k_jcp_slots", "jcp_solve": "k_jcp_solve",
                        "labels": "k_seg_labels"}
HBM_PEAK_GBS = 8000.0


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def frames_c2(count, offset=0):
    from lidar_amd import synth
    return [synth.workload_c2(offset + i) for i in range(count)]


def algorithmic_bytes(n_raw, n, p, n_obs, voxels, hull_vertices, clusters):
    """Per-frame algorithmic HBM bytes of the whole path: noise filter (read
    xyz 12 B + write a label 1 B per raw point, write the kept xyz+ring 14 B),
    read xyz+ring (14 B per point), write labels (4), range image write+read
    (10 B per pixel), read obstacle xyz (12) + write cluster labels (4), voxel
    key+value (8 per voxel), read hull input x,y (8 per obstacle point), write
    hull vertices (16 per vertex), read them for the box (16) and write a box
    (80 per cluster)."""
    return (13 * n_raw + 14 * n + 14 * n + 4 * n + 10 * p + 12 * n_obs + 4 * n_obs + 8 * voxels + 8 * n_obs
            + 16 * hull_vertices + 16 * hull_vertices + 80 * clusters)


def stage_bytes(stage, n, p, q, n_obs, voxels, hull_vertices, n_cand, n_raw=0, clusters=0, searched=0):
    """Algorithmic bytes one launch of a stage moves (per frame)."""
    return {
        "noise": 12 * n_raw + 4 * n_raw + 16 * n_raw,  # xyz in, cell slot out, cell-sorted points out
        "noise_count": 25 * n_raw,                   # cell-sorted point + index in, label + keep flag out
        "noise_search": 29 * searched,               # list entry + point in, label + keep flag out
        "noise_compact": 4 * n_raw + 14 * n,         # keep scan in, kept xyz+ring out
        "project": 14 * n + 8 * n,                  # xyz+ring in, cell+pixel out
        "cells": 4 * n * 3 + 4 * 18000 * 2,          # scatter + sort of point indices, zmin out
        "recm": 4 * 18000 * 2,
        "ransac": 12 * n_cand * 2,                   # candidate compaction + one scoring pass
        "pixels": 8 * p + 21 * p,                    # winner keys in, state/mapping/xyz out
        "jcp_queue": p + 4 * p,
        "jcp_slots": q * (24 * 21 + 192),            # 24 neighbours (index, state, xyz) in, slots out
        "jcp_solve": q * 193 + q,                    # slot weights/sources in, states out
        "labels": 4 * n + 4 * n,
        "split": 4 * n + 16 * n_obs,
        "voxels": 12 * n_obs + 12 * n_obs + 8 * voxels,
        "components": 8 * voxels * 14 + 12 * n_obs,
        "hull": 8 * n_obs + 16 * hull_vertices,
        "boxes": 16 * hull_vertices + 80 * clusters,
    }.get(stage, 0)


def timed_batch(N, args, frames, streams):
    """Graph replays of one staged batch of args.batch frames; returns
    (frames/s, ms per step, points/s, outputs of an eager run, context)."""
    B = args.batch
    maxn = max(f.n for f in frames)
    scfg = N.segmenter_config(image_height=64, image_width=2048)
    ctx = N.Context(max_frames=B, max_points=maxn, max_height=64, max_width=2048, streams=streams)
    ctx.set_noise_filter(N.noise_config())
    ctx.set_oriented_boxes(True)
    ctx.stage([(f.x, f.y, f.z, f.ring) for f in frames[:B]], scfg)
    ctx.run(use_graph=False)
    outs = ctx.fetch()
    bad = [i for i, o in enumerate(outs) if o["status"] != 0]
    if bad:
        raise SystemExit(f"frames {bad} failed: {[outs[i]['status'] for i in bad]}")
    for _ in range(args.warmup):
        ctx.run(use_graph=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.run(use_graph=True)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    return (B * args.steps / dt, dt / args.steps * 1e3, sum(f.n for f in frames[:B]) * args.steps / dt, outs,
            ctx)


def run_gpu(args, frames):
    from lidar_amd import native as N
    maxn = max(f.n for f in frames)
    scfg = N.segmenter_config(image_height=64, image_width=2048)
    ncfg = N.noise_config()
    # Headline: the batch as args.streams contiguous shares on as many
    # streams (lp_capacity.streams), so one share's latency-bound kernels
    # overlap the others' work.
    fps, ms_per_step, pts, outs, ctx = timed_batch(N, args, frames, args.streams)
    ctx.close()
    # One stream, one graph: the same batch, and the per-stage HIP event
    # timings of its kernels (profiled launches).
    fps1, ms1, _, _, ctx = timed_batch(N, args, frames, 1)
    stages = ctx.profile(runs=5)
    ctx.close()

    # Plain per-frame path: one frame, one graph launch, synchronous.
    ctx1 = N.Context(max_frames=1, max_points=maxn, max_height=64, max_width=2048)
    ctx1.set_noise_filter(ncfg)
    ctx1.set_oriented_boxes(True)
    lat = []
    lat_e2e = []
    for i in range(args.plain_frames):
        f = frames[i % len(frames)]
        ctx1.stage([(f.x, f.y, f.z, f.ring)], scfg)
        ctx1.synchronize()
        if i < 3:
            ctx1.run(use_graph=True)
            ctx1.synchronize()
        t = time.perf_counter()
        ctx1.run(use_graph=True)
        ctx1.synchronize()
        lat.append(time.perf_counter() - t)
    for i in range(min(args.plain_frames, 60)):
        f = frames[i % len(frames)]
        t = time.perf_counter()
        ctx1.process([(f.x, f.y, f.z, f.ring)], scfg, use_graph=True)
        lat_e2e.append(time.perf_counter() - t)
    ctx1.close()
    lat = np.array(lat) * 1e3
    lat_e2e = np.array(lat_e2e[3:]) * 1e3
    return dict(fps=fps, pts=pts, ms_per_step=ms_per_step, stages=stages, outs=outs,
                one_stream=dict(frames_per_s=fps1, ms_per_batch=ms1),
                plain=dict(frames_per_s=1e3 / lat.mean(), p50_ms=float(np.percentile(lat, 50)),
                           p99_ms=float(np.percentile(lat, 99)), frames=len(lat)),
                plain_e2e=dict(frames_per_s=1e3 / lat_e2e.mean(), p50_ms=float(np.percentile(lat_e2e, 50)),
                               p99_ms=float(np.percentile(lat_e2e, 99)),
                               note="host arrays in, labels/clusters/hulls out (PCIe included)"))


def run_batched(frames, steps, noise=False, boxes=False, streams=2, load=0.5):
    """Frames/s of one batch of `frames` (any sizes) replayed as graphs on
    `streams` streams."""
    from lidar_amd import native as N
    maxn = max(f.n for f in frames)
    H = max(f.height for f in frames)
    ctx = N.Context(max_frames=len(frames), max_points=maxn, max_height=H, max_width=2048, streams=streams,
                    hash_load_factor=load)
    if noise:
        ctx.set_noise_filter(N.noise_config())
    ctx.set_oriented_boxes(boxes)
    cfgs = [N.segmenter_config(image_height=f.height, image_width=f.width, elevation_up_deg=f.fov_up_deg,
                               elevation_down_deg=f.fov_down_deg) for f in frames]
    ctx.stage([(f.x, f.y, f.z, f.ring) for f in frames], cfgs)
    ctx.run(use_graph=True)
    outs = ctx.fetch()
    bad = [o["status"] for o in outs if o["status"] != 0]
    if bad:
        raise SystemExit(f"batched frames failed: {bad}")
    for _ in range(2):
        ctx.run(use_graph=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.run(use_graph=True)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    ctx.close()
    return dict(frames=len(frames), streams=streams, hash_load_factor=load,
                frames_per_s=round(len(frames) * steps / dt, 2),
                points_per_s=round(sum(f.n for f in frames) * steps / dt, 1),
                ms_per_batch=round(dt / steps * 1e3, 4))


def run_serving(seconds):
    """Eight simulated sensors (4x 32x1024, 3x 64x2048, 1x 128x2048) at 10 Hz
    with +-5 ms jitter, served with continuous batching."""
    from lidar_amd import native as N
    from lidar_amd import serving
    pools = serving.sensor_pools(pool_frames=3)
    schedule = serving.arrival_schedule(seconds)
    srv = N.Server(slots=3, max_batch_frames=8)
    # Warm the per-batch-size graphs outside the measured schedule.
    serving.serve(srv, pools, serving.arrival_schedule(0.3, rate_hz=200.0, seed=11))
    before = srv.stats()
    run = serving.serve(srv, pools, schedule)
    st = srv.stats()
    srv.close()
    out = serving.summarize(schedule, run["latency_s"])
    batches = st["batches"] - before["batches"]
    out.update(wall_s=round(run["wall_s"], 3), frames_per_s=round(len(schedule) / run["wall_s"], 2),
               slots=st["slots"], batches=int(batches), mean_batch=round(len(schedule) / max(batches, 1), 3),
               slot_bytes=int(st["slot_bytes"]))
    return out


def run_cpu(frames, budget_s):
    """The oracle's full frame with the reference NoiseRemover's KD-tree
    search (its data structures and its stack behaviour) in front."""
    from oracle import lp_oracle as O
    proc = O.Processor(O.seg_config(), O.cluster_config())
    nr = O.NoiseRemover()
    bufs = None
    times = []
    t_end = time.perf_counter() + budget_s
    k = 0
    while time.perf_counter() < t_end or k < 3:
        f = frames[k % len(frames)]
        if bufs is None or len(bufs["seg"]) < f.n:
            n = max(ff.n for ff in frames)
            bufs = dict(noise=np.zeros(n, np.uint8), seg=np.zeros(n, np.uint32), clab=np.zeros(n, np.int32),
                        hoff=np.zeros(n + 2, np.int32), hidx=np.zeros(2 * n + 16, np.int32),
                        boxes=(O.BoundingBox * (n + 2))())
        t = time.perf_counter()
        r = O.full_frame(proc, f.x, f.y, f.z, f.ring, noise_remover=nr, bufs=bufs)
        times.append(time.perf_counter() - t)
        assert r["status"] == 0
        k += 1
    t = np.array(times)
    return dict(fps=1.0 / t.mean(), pts=np.mean([f.n for f in frames]) / t.mean(), frames=len(t),
                p50_ms=float(np.percentile(t * 1e3, 50)))


def run_cpu_parallel(frames, budget_s, threads):
    """The same oracle frames on `threads` host threads, one oracle processor
    and noise remover per thread (the reference objects are per instance);
    the oracle's C calls release the interpreter lock."""
    import threading
    from oracle import lp_oracle as O
    n = max(ff.n for ff in frames)
    done = [0] * threads
    t_end = [0.0]

    def worker(w):
        proc = O.Processor(O.seg_config(), O.cluster_config())
        nr = O.NoiseRemover()
        bufs = dict(noise=np.zeros(n, np.uint8), seg=np.zeros(n, np.uint32), clab=np.zeros(n, np.int32),
                    hoff=np.zeros(n + 2, np.int32), hidx=np.zeros(2 * n + 16, np.int32),
                    boxes=(O.BoundingBox * (n + 2))())
        k = w
        while time.perf_counter() < t_end[0]:
            f = frames[k % len(frames)]
            r = O.full_frame(proc, f.x, f.y, f.z, f.ring, noise_remover=nr, bufs=bufs)
            assert r["status"] == 0
            done[w] += 1
            k += threads

    t0 = time.perf_counter()
    t_end[0] = t0 + budget_s
    ts = [threading.Thread(target=worker, args=(w,)) for w in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    dt = time.perf_counter() - t0
    return dict(fps=sum(done) / dt, frames=sum(done), threads=threads, seconds=dt)


def run_probe():
    """Counter-collection workload: 8 full-pipeline frames run eagerly three
    times in one batch, then a 256 MiB device-to-device copy whose known size
    calibrates FETCH_SIZE / WRITE_SIZE."""
    import torch
    from lidar_amd import native as N
    frames = frames_c2(8)
    ctx = N.Context(max_frames=8, max_points=max(f.n for f in frames), max_height=64, max_width=2048)
    ctx.set_noise_filter(N.noise_config())
    ctx.set_oriented_boxes(True)
    ctx.stage([(f.x, f.y, f.z, f.ring) for f in frames], N.segmenter_config())
    for _ in range(3):
        ctx.run(use_graph=False)
    ctx.synchronize()
    log("probe statuses", [o["status"] for o in ctx.fetch()])
    ctx.close()
    a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log("copy bytes", a.numel() * 4)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--streams", type=int, default=2, help="streams the batch is split over")
    ap.add_argument("--plain-frames", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--no-extra", action="store_true", help="skip the c3/c4 batches and the serving run")
    ap.add_argument("--probe", action="store_true", help="only the counter-collection workload (no JSON line)")
    args = ap.parse_args()
    if args.probe:
        run_probe()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (frames shard across GPUs "
              "with no exchange, so multi-GPU throughput is N independent copies of this run)", file=sys.stderr)
        raise SystemExit(2)
    log("generating frames")
    frames = frames_c2(args.batch)
    log("gpu")
    g = run_gpu(args, frames)
    extra = {}
    if not args.no_extra:
        from lidar_amd import synth
        from lidar_amd import autotune
        from lidar_amd import native as N
        c3 = [synth.workload_c3(i) for i in range(32)]
        log("autotune c3 (streams x hash load factor)")
        c3_cfgs = [N.segmenter_config(image_height=f.height, image_width=f.width, elevation_up_deg=f.fov_up_deg,
                                      elevation_down_deg=f.fov_down_deg) for f in c3]
        tuned = autotune.tune(c3, c3_cfgs, streams=(1, 2, 3), loads=(0.5, 0.7), steps=5, noise=False, boxes=False)
        extra["c3_autotune"] = tuned
        log("batched c3 (mixed widths, tuned)")
        extra["batched_c3"] = run_batched(c3, 5, streams=tuned["best"]["streams"],
                                          load=tuned["best"]["hash_load_factor"])
        log("batched c4 (128x2048)")
        extra["batched_c4"] = run_batched([synth.workload_c4(i) for i in range(16)], 5)
        if args.serve_seconds > 0:
            log("serving")
            extra["serving"] = run_serving(args.serve_seconds)
    cpu = None
    cpu_par = None
    if not args.no_cpu:
        log("cpu oracle")
        cpu = run_cpu(frames[:8], args.cpu_seconds)
        threads = min(len(os.sched_getaffinity(0)), 16)
        if threads > 1:
            log(f"cpu oracle, {threads} threads")
            cpu_par = run_cpu_parallel(frames[:8], args.cpu_seconds / 2, threads)

    # Dominant kernel stage and its roofline.
    outs = g["outs"]
    B = args.batch
    mean = lambda key: float(np.mean([o[key] if not isinstance(o[key], np.ndarray) else len(o[key]) for o in outs]))
    n_raw = float(np.mean([f.n for f in frames[:B]]))
    n = float(np.mean([o["num_points"] for o in outs]))
    n_obs = float(np.mean([len(o["obstacle_index"]) for o in outs]))
    ncl = mean("num_clusters")
    vox = mean("num_voxels")
    hv = float(np.mean([len(o["hull_indices"]) for o in outs]))
    q = mean("jcp_queue")
    ncand = float(np.mean([o["ransac"]["candidates"] for o in outs]))
    searched = mean("num_noise_searched")
    P = 64 * 2048
    # Dominant kernel: the slowest stage that is a single kernel launch (its
    # HIP-event time is then that kernel's duration, comparable to rocprof).
    single = {s: k for s, k in SINGLE_KERNEL_STAGES.items()}
    stages = [(s, ms) for s, ms in g["stages"] if s in single]
    dom, dom_ms = max(stages, key=lambda t: t[1])
    dom_bytes = stage_bytes(dom, n, P, q, n_obs, vox, hv, ncand, n_raw, ncl, searched) * B
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        try:
            t = json.load(open(tpath)).get(single[dom])
            if t:
                traffic = round(t["hbm_bytes_per_frame"] * B)
        except (OSError, ValueError, KeyError):
            traffic = None
    alg_frame = algorithmic_bytes(n_raw, n, P, n_obs, vox, hv, ncl)
    line = {
        "metric": METRIC,
        "value": round(g["fps"], 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(g["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 64x2048 scans generated on the host, resident in HBM)",
        "config": {
            "workload": "64x2048 full single-frame pipeline: 40 cars, 60 poles, 2000 rain points "
                        "(~128k points), DROR filter + segment + split + cluster + hull + rotating-calipers box, "
                        "batch of %d frames per step" % B,
            "batch_frames": B,
            "points_per_frame": round(n_raw),
            "form": "batched (CSR-packed frames; %d contiguous shares on %d streams, one hipGraph each, "
                    "per step)" % (args.streams, args.streams),
        },
        "points_per_s": round(g["pts"], 1),
        "plain": {k: round(v, 4) if isinstance(v, float) else v for k, v in g["plain"].items()},
        "plain_end_to_end": {k: round(v, 4) if isinstance(v, float) else v for k, v in g["plain_e2e"].items()},
        "batched_one_stream": {k: round(v, 4) for k, v in g["one_stream"].items()},
        "batched_over_plain": round(g["fps"] / g["plain"]["frames_per_s"], 3),
        "stage_ms_per_batch": {s: round(ms, 4) for s, ms in g["stages"]},
        "roofline": {
            "bound": "hbm",
            "kernel": single[dom],
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": round(dom_bytes),
            "launch_ms": round(dom_ms, 4),
        },
        "pipeline_roofline": {
            "algorithmic_bytes_per_frame": round(alg_frame),
            "achieved_GBps": round(alg_frame * g["fps"] / 1e9, 3),
            "frac": round(alg_frame * g["fps"] / 1e9 / HBM_PEAK_GBS, 6),
        },
    }
    line.update(extra)
    if cpu:
        line["cpu_baseline"] = {"value": round(cpu["fps"], 3), "unit": "frames/s", "cores": 1, "kind": "port",
                                "sample": f"{cpu['frames']} frames of the same workload (8 distinct seeds) through "
                                          "the oracle's KD-tree DROR + segment+split+cluster+gather+hull+box, "
                                          "single thread",
                                "points_per_s": round(cpu["pts"], 1), "p50_ms": round(cpu["p50_ms"], 3)}
        line["speedup_vs_cpu"] = {"batched": round(g["fps"] / cpu["fps"], 2),
                                  "plain": round(g["plain"]["frames_per_s"] / cpu["fps"], 2)}
    if cpu_par:
        line["cpu_all_cores"] = {"value": round(cpu_par["fps"], 3), "unit": "frames/s", "cores": cpu_par["threads"],
                                 "kind": "port", "sample": f"{cpu_par['frames']} frames in {cpu_par['seconds']:.1f} s, "
                                 "frame-parallel, one oracle processor and noise remover per thread",
                                 "cpu_model": cpu_model()}
        line["speedup_vs_cpu"]["batched_vs_all_cores"] = round(g["fps"] / cpu_par["fps"], 2)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
