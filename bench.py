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
# Profiled stages that consist of exactly one kernel launch (a name that
# recurs, like the three hull rounds, is summed over its launches).
SINGLE_KERNEL_STAGES = {"project": "k_project", "recm": "k_recm", "pixels": "k_pixel_label",
                        "jcp_slots": "k_jcp_slots", "jcp_solve": "k_jcp_solve", "labels": "k_seg_labels",
                        "hull_round": "k_hull_round", "hull_big": "k_hull_big", "hull_ref": "k_hull_ref",
                        "hull_local": "k_hull_local"}
# The noise filter's stages that are single kernels (radius rule search).
NOISE_STAGES = {"noise_count": "k_noise_count", "noise_search": "k_noise_search"}
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


def bytes_8d(n, p, n_obs, voxels, hull_vertices):
    """Per-frame algorithmic HBM bytes of the hot path as the scope table
    defines them: read x, y, z f32 + ring u16 (14 B per point), write a
    Label u32 (4), range image write + read of winner i32 + state u8 (10 B
    per pixel), read obstacle xyz (12) and write a ClusterLabel i32 (4) per
    obstacle point, hash key + value (8 per voxel), read hull input x, y f32
    (8 per obstacle point), write hull vertices (16 per vertex, two f64)."""
    return 14 * n + 4 * n + 10 * p + 12 * n_obs + 4 * n_obs + 8 * voxels + 8 * n_obs + 16 * hull_vertices


def bytes_extended(n_raw, n, p, n_obs, voxels, hull_vertices, clusters):
    """bytes_8d plus the stages around it in the headline pipeline: the noise
    filter (read xyz 12 B + write a label 1 B per raw point, write the kept
    xyz+ring 14 B per kept point) and the boxes (read the hull vertices 16 B,
    write an 80 B box per cluster)."""
    return (bytes_8d(n, p, n_obs, voxels, hull_vertices) + 13 * n_raw + 14 * n + 16 * hull_vertices
            + 80 * clusters)


def kernel_bytes_8d(kernel, n, p, n_obs, voxels, hull_vertices):
    """The share of bytes_8d each kernel moves (per frame): the projection
    reads the points, the pixel labels write the range image and the row
    solve reads it, the labels kernel writes the labels, the spherical
    conversion reads the obstacle points and the component labels write
    their cluster labels, the voxel insert writes the hash, and the hull
    rounds read the hull inputs and produce the vertices."""
    return {"k_project": 14 * n, "k_pixel_label": 5 * p, "k_jcp_solve": 5 * p, "k_seg_labels": 4 * n,
            "k_spherical": 12 * n_obs, "k_component_label": 4 * n_obs, "k_voxel_insert": 8 * voxels,
            "k_hull_round": 8 * n_obs + 16 * hull_vertices}.get(kernel, 0)


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


def plain_latencies(N, ctx, frames, scfgs, count):
    """One frame per graph launch, synchronous: device-resident latency (s)."""
    lat = []
    for i in range(count):
        k = i % len(frames)
        f = frames[k]
        ctx.stage([(f.x, f.y, f.z, f.ring)], scfgs[k])
        ctx.synchronize()
        if i < 3:
            ctx.run(use_graph=True)
            ctx.synchronize()
        t = time.perf_counter()
        ctx.run(use_graph=True)
        ctx.synchronize()
        lat.append(time.perf_counter() - t)
    return np.array(lat)


def lat_summary(lat_s):
    ms = np.asarray(lat_s) * 1e3
    return dict(frames_per_s=1e3 / ms.mean(), p50_ms=float(np.percentile(ms, 50)),
                p99_ms=float(np.percentile(ms, 99)), frames=len(ms))


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
    # timings of its kernels (profiled launches on the context stream).
    fps1, ms1, _, _, ctx = timed_batch(N, args, frames, 1)
    stages = ctx.profile(runs=5)
    ctx.close()

    # Plain per-frame path: one frame, one graph launch, synchronous.
    ctx1 = N.Context(max_frames=1, max_points=maxn, max_height=64, max_width=2048)
    ctx1.set_noise_filter(ncfg)
    ctx1.set_oriented_boxes(True)
    lat = plain_latencies(N, ctx1, frames, [scfg] * len(frames), args.plain_frames)
    lat_e2e = []
    for i in range(min(args.plain_frames, 60)):
        f = frames[i % len(frames)]
        t = time.perf_counter()
        ctx1.process([(f.x, f.y, f.z, f.ring)], scfg, use_graph=True)
        lat_e2e.append(time.perf_counter() - t)
    ctx1.close()
    e2e = lat_summary(lat_e2e[3:])
    e2e["note"] = "host arrays in, labels/clusters/hulls/boxes out (PCIe included)"
    return dict(fps=fps, pts=pts, ms_per_step=ms_per_step, stages=stages, outs=outs,
                one_stream=dict(frames_per_s=fps1, ms_per_batch=ms1), plain=lat_summary(lat), plain_e2e=e2e)


def run_c1_plain(args):
    """The plain method on its own (segment + split + cluster + gather +
    hull, no noise filter, no boxes) on 64x2048 frames with 15 cars, 20
    poles and 2 walls: GPU one frame per launch, and the oracle on one host
    thread."""
    from lidar_amd import native as N
    from lidar_amd import synth
    from oracle import lp_oracle as O
    frames = [synth.workload_c1(i) for i in range(8)]
    scfg = N.segmenter_config(image_height=64, image_width=2048)
    ctx = N.Context(max_frames=1, max_points=max(f.n for f in frames), max_height=64, max_width=2048)
    gpu = lat_summary(plain_latencies(N, ctx, frames, [scfg] * len(frames), args.plain_frames))
    ctx.close()
    gpu["points_per_s"] = gpu["frames_per_s"] * float(np.mean([f.n for f in frames]))
    out = dict(workload="64x2048 plain method: 15 cars, 20 poles, 2 walls, 10% dropout (~118k points), "
                        "segment + split + cluster + gather + hull", gpu_plain=gpu)
    if not args.no_cpu:
        proc = O.Processor(O.seg_config(), O.cluster_config())
        times = []
        t_end = time.perf_counter() + args.cpu_seconds / 2
        k = 0
        while time.perf_counter() < t_end or k < 3:
            f = frames[k % len(frames)]
            t = time.perf_counter()
            r = proc.frame(f.x, f.y, f.z, f.ring)
            times.append(time.perf_counter() - t)
            assert r["status"] == 0
            k += 1
        t = np.array(times)
        out["cpu_1_thread"] = dict(frames_per_s=1.0 / t.mean(), p50_ms=float(np.percentile(t * 1e3, 50)),
                                   frames=len(t), kind="port")
        out["gpu_over_cpu"] = gpu["frames_per_s"] / out["cpu_1_thread"]["frames_per_s"]
    return out


def run_batched(frames, steps, noise=False, boxes=False, streams=2, load=0.5, point_workgroup=256, mark_tile=256,
                persistent=False):
    """Frames/s of one batch of `frames` (any sizes) replayed as graphs on
    `streams` streams with the given launch knobs."""
    from lidar_amd import native as N
    maxn = max(f.n for f in frames)
    H = max(f.height for f in frames)
    ctx = N.Context(max_frames=len(frames), max_points=maxn, max_height=H, max_width=2048, streams=streams,
                    hash_load_factor=load)
    if noise:
        ctx.set_noise_filter(N.noise_config())
    ctx.set_oriented_boxes(boxes)
    ctx.set_launch_shape(point_workgroup, mark_tile)
    ctx.set_persistent(persistent)
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
    return dict(frames=len(frames), streams=streams, hash_load_factor=load, point_workgroup=point_workgroup,
                mark_tile=mark_tile, persistent=persistent,
                frames_per_s=round(len(frames) * steps / dt, 2),
                points_per_s=round(sum(f.n for f in frames) * steps / dt, 1),
                ms_per_batch=round(dt / steps * 1e3, 4))


def run_serving(seconds, rate_hz=10.0, jitter_ms=5.0):
    """Eight simulated sensors (4x 32x1024, 3x 64x2048, 1x 128x2048) at
    `rate_hz` with +-jitter_ms, served with continuous batching."""
    from lidar_amd import native as N
    from lidar_amd import serving
    pools = serving.sensor_pools(pool_frames=3)
    schedule = serving.arrival_schedule(seconds, rate_hz=rate_hz, jitter_ms=jitter_ms)
    srv = N.Server(slots=3, max_batch_frames=8)
    # Warm the per-batch-size graphs outside the measured schedule.
    serving.serve(srv, pools, serving.arrival_schedule(0.3, rate_hz=200.0, seed=11))
    before = srv.stats()
    run = serving.serve(srv, pools, schedule)
    st = srv.stats()
    srv.close()
    out = serving.summarize(schedule, run["latency_s"])
    batches = st["batches"] - before["batches"]
    out.update(rate_hz_per_sensor=rate_hz, wall_s=round(run["wall_s"], 3),
               frames_per_s=round(len(schedule) / run["wall_s"], 2), slots=st["slots"], batches=int(batches),
               mean_batch=round(len(schedule) / max(batches, 1), 3), slot_bytes=int(st["slot_bytes"]))
    return out


def run_cpu(frames, budget_s):
    """The oracle's full frame with the reference NoiseRemover's KD-tree
    search (its data structures and its stack behaviour; a fresh object per
    frame) in front."""
    from oracle import lp_oracle as O
    proc = O.Processor(O.seg_config(), O.cluster_config())
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
        nr = O.NoiseRemover()
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
    per thread and a fresh noise remover per frame (the reference objects are
    per instance); the oracle's C calls release the interpreter lock."""
    import threading
    from oracle import lp_oracle as O
    n = max(ff.n for ff in frames)
    done = [0] * threads
    t_end = [0.0]

    def worker(w):
        proc = O.Processor(O.seg_config(), O.cluster_config())
        bufs = dict(noise=np.zeros(n, np.uint8), seg=np.zeros(n, np.uint32), clab=np.zeros(n, np.int32),
                    hoff=np.zeros(n + 2, np.int32), hidx=np.zeros(2 * n + 16, np.int32),
                    boxes=(O.BoundingBox * (n + 2))())
        k = w
        while time.perf_counter() < t_end[0]:
            f = frames[k % len(frames)]
            r = O.full_frame(proc, f.x, f.y, f.z, f.ring, noise_remover=O.NoiseRemover(), bufs=bufs)
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


CALIBRATION_BYTES = 256 * 1024 * 1024


def calibration_copy():
    """A device-to-device copy of CALIBRATION_BYTES through the HIP runtime
    (three times), whose known size calibrates FETCH_SIZE / WRITE_SIZE."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(CALIBRATION_BYTES)) != 0:
            raise SystemExit("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(CALIBRATION_BYTES))
    for _ in range(3):
        hip.hipMemcpy(b, a, ctypes.c_size_t(CALIBRATION_BYTES), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def run_probe(args):
    """Counter-collection workload: the headline batch (args.batch frames as
    args.streams shares, one graph each) replayed three times, then the
    calibration copy."""
    from lidar_amd import native as N
    frames = frames_c2(args.batch)
    fps, _, _, outs, ctx = timed_batch(N, argparse.Namespace(**{**vars(args), "steps": 3, "warmup": 0}), frames,
                                       args.streams)
    ctx.close()
    log("probe statuses", [o["status"] for o in outs])
    calibration_copy()
    log("copy bytes", CALIBRATION_BYTES)


# ---------------------------------------------------------------- profiles
def _short(name):
    import re
    m = re.search(r"(k_\w+)", name)
    if m:
        return m.group(1)
    return name.replace("void ", "")[:40]


def summarize_trace(db, frames=None, per_frame_grid=False):
    """Kernel durations of a rocprofv3 kernel trace (rocpd SQLite): calls,
    total, average, min, max; with `frames`, only launches covering that many
    frames (grid y / block y), or with per_frame_grid the launches of one
    workgroup per frame (grid y = 1, `frames` workgroups in x)."""
    import sqlite3
    con = sqlite3.connect(db)
    agg = {}
    rows = con.execute("select name, duration, grid_x, workgroup_x, grid_y, workgroup_y from kernels")
    for name, dur, gx, wx, gy, wy in rows:
        nb = gy // max(wy, 1) if gy else 1
        if per_frame_grid:
            if nb != 1 or gx // max(wx, 1) != frames:
                continue
        elif frames is not None and nb != frames:
            continue
        a = agg.setdefault(_short(name), [0, 0.0, float("inf"), 0.0])
        a[0] += 1
        a[1] += dur
        a[2] = min(a[2], dur)
        a[3] = max(a[3], dur)
    tot = sum(v[1] for v in agg.values()) or 1.0
    lines = [f"{'kernel':32s} {'calls':>6s} {'total_ms':>9s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s} {'pct':>6s}"]
    for k, (c, t, mn, mx) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        lines.append(f"{k[:32]:32s} {c:6d} {t/1e6:9.3f} {t/c/1e3:9.2f} {mn/1e3:9.2f} {mx/1e3:9.2f} {100*t/tot:6.2f}")
    return "\n".join(lines)


def _per_dispatch(db, counter):
    import sqlite3
    from collections import defaultdict
    con = sqlite3.connect(db)
    per = defaultdict(float)
    disp = defaultdict(set)
    rows = con.execute("select dispatch_id, kernel_name, counter_name, value, grid_size from counters_collection")
    for did, name, cn, v, gs in rows:
        if cn != counter:
            continue
        k = _short(name) if "k_" in name else f"copy {gs}" if "copyBuffer" in name else None
        if k is None:
            continue
        per[k] += v
        disp[k].add(did)
    return {k: (per[k] / len(disp[k]), len(disp[k])) for k in per}


def summarize_traffic(fetch_db, write_db, frames, copy_bytes):
    """HBM bytes per frame of every kernel: FETCH_SIZE and WRITE_SIZE (KiB
    per dispatch, one run each) scaled so that the device-to-device copy of
    copy_bytes in the same runs reads and writes exactly its size."""
    fe = _per_dispatch(fetch_db, "FETCH_SIZE")
    wr = _per_dispatch(write_db, "WRITE_SIZE")
    big = max((k for k in fe if k.startswith("copy ")), key=lambda k: int(k.split()[1]))
    fe["copy"], wr["copy"] = fe[big], wr[big]
    fs = copy_bytes / 1024.0 / fe["copy"][0]
    ws = copy_bytes / 1024.0 / wr["copy"][0]
    out = {"calibration": {"copy_bytes": copy_bytes, "fetch_kib": round(fe["copy"][0], 3),
                           "write_kib": round(wr["copy"][0], 3), "fetch_scale": round(fs, 4),
                           "write_scale": round(ws, 4)}}
    for k in sorted(set(fe) & set(wr)):
        if k.startswith("copy"):
            continue
        f_kib, nd = fe[k]
        w_kib = wr[k][0]
        out[k] = {"fetch_kib_per_dispatch": round(f_kib, 1), "write_kib_per_dispatch": round(w_kib, 1),
                  "dispatches": nd, "frames_per_dispatch": frames,
                  "hbm_bytes_per_frame": round((f_kib * 1024 * fs + w_kib * 1024 * ws) / frames)}
    return out


def summarize_pmc(db):
    """Per-kernel counter values averaged per dispatch."""
    import sqlite3
    from collections import defaultdict
    con = sqlite3.connect(db)
    per = defaultdict(lambda: defaultdict(float))
    disp = defaultdict(set)
    dur = defaultdict(dict)
    rows = con.execute("select dispatch_id, kernel_name, counter_name, value, duration, grid_size "
                       "from counters_collection")
    for did, name, cn, v, du, gs in rows:
        k = _short(name) + ("" if "k_" in name else f"[grid {gs}]")
        per[k][cn] += v
        disp[k].add(did)
        dur[k][did] = du
    names = sorted({cn for k in per for cn in per[k]})
    lines = [f"{'kernel':40s} {'disp':>5s} {'avg_us':>9s} " + " ".join(f"{n:>14s}" for n in names)]
    for k in sorted(per, key=lambda k: -sum(dur[k].values())):
        nd = len(disp[k])
        lines.append(f"{k[:40]:40s} {nd:5d} {sum(dur[k].values())/nd/1e3:9.1f} " +
                     " ".join(f"{per[k][n]/nd:14.1f}" for n in names))
    return "\n".join(lines)


def summarize(argv):
    mode, db = argv[0], argv[1]
    if mode == "trace":
        print(summarize_trace(db, int(argv[2]) if len(argv) > 2 else None))
    elif mode == "trace_wg":
        print(summarize_trace(db, int(argv[2]), per_frame_grid=True))
    elif mode == "traffic":
        print(json.dumps(summarize_traffic(db, argv[2], int(argv[3]), int(argv[4])), indent=1))
    elif mode == "pmc":
        print(summarize_pmc(db))
    else:
        raise SystemExit(f"unknown summary {mode}")


# ---------------------------------------------------------------- main
def stage_totals(stages):
    tot = {}
    for name, ms in stages:
        tot[name] = tot.get(name, 0.0) + ms
    return tot


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize":
        summarize(sys.argv[2:])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--streams", type=int, default=4, help="streams the batch is split over")
    ap.add_argument("--plain-frames", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--no-extra", action="store_true", help="skip the c1/c3/c4 runs and the serving runs")
    ap.add_argument("--probe", action="store_true", help="only the counter-collection workload (no JSON line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (frames shard across GPUs "
              "with no exchange, so multi-GPU throughput is N independent copies of this run)", file=sys.stderr)
        raise SystemExit(2)
    if args.probe:
        run_probe(args)
        return
    log("generating frames")
    frames = frames_c2(args.batch)
    log("gpu")
    g = run_gpu(args, frames)
    extra = {}
    if not args.no_extra:
        from lidar_amd import synth
        from lidar_amd import autotune
        from lidar_amd import native as N
        log("c1 plain method")
        extra["c1_plain"] = run_c1_plain(args)
        # Per shape bucket: c3 (32 frames of mixed widths) and c4 (16 frames
        # of 128x2048), tuned over streams x load factor, then the point
        # workgroup size, the marking tile and the launch form (the tuner's
        # result list holds the persistent / grid A/B at the best setting).
        def cfgs_of(frames):
            return [N.segmenter_config(image_height=f.height, image_width=f.width, elevation_up_deg=f.fov_up_deg,
                                       elevation_down_deg=f.fov_down_deg) for f in frames]

        knobs = ("streams", "hash_load_factor", "point_workgroup", "mark_tile", "persistent")
        for name, bucket in (("c3", [synth.workload_c3(i) for i in range(32)]),
                             ("c4", [synth.workload_c4(i) for i in range(16)])):
            log(f"autotune {name}")
            tuned = autotune.tune(bucket, cfgs_of(bucket), streams=(1, 2, 4), loads=(0.5, 0.7),
                                  workgroups=(64, 128, 256), tiles=(64, 128, 256), persistent=(False, True), steps=5,
                                  noise=False, boxes=False)
            extra[f"{name}_autotune"] = tuned
            log(f"batched {name} (tuned)")
            best = {k: tuned["best"][k] for k in knobs}
            extra[f"batched_{name}"] = run_batched(bucket, 5, streams=best["streams"], load=best["hash_load_factor"],
                                                   point_workgroup=best["point_workgroup"],
                                                   mark_tile=best["mark_tile"], persistent=best["persistent"])
        if args.serve_seconds > 0:
            log("serving (10 Hz per sensor)")
            extra["serving"] = run_serving(args.serve_seconds)
            log("serving, saturated (1 kHz per sensor, above the service rate)")
            extra["serving_saturated"] = run_serving(min(args.serve_seconds, 2.0), rate_hz=1000.0, jitter_ms=0.5)
    cpu = None
    cpu_par = None
    if not args.no_cpu:
        log("cpu oracle")
        cpu = run_cpu(frames[:8], args.cpu_seconds)
        threads = min(len(os.sched_getaffinity(0)), 16)
        if threads > 1:
            log(f"cpu oracle, {threads} threads")
            cpu_par = run_cpu_parallel(frames[:8], args.cpu_seconds / 2, threads)

    # Per-frame counts of the headline batch.
    outs = g["outs"]
    B = args.batch
    mean = lambda key: float(np.mean([o[key] if not isinstance(o[key], np.ndarray) else len(o[key]) for o in outs]))
    n_raw = float(np.mean([f.n for f in frames[:B]]))
    n = float(np.mean([o["num_points"] for o in outs]))
    n_obs = float(np.mean([len(o["obstacle_index"]) for o in outs]))
    ncl = mean("num_clusters")
    vox = mean("num_voxels")
    hv = float(np.mean([len(o["hull_indices"]) for o in outs]))
    P = 64 * 2048
    # Dominant kernel: the single-kernel stage with the most time per step
    # (summed over its launches), from HIP events on the context stream.
    single = dict(SINGLE_KERNEL_STAGES)
    single.update(NOISE_STAGES)
    totals = stage_totals(g["stages"])
    launches = {}
    for name, _ in g["stages"]:
        launches[name] = launches.get(name, 0) + 1
    # Among the stages on the scope's path (kernels with a share of its
    # per-frame bytes); the DROR and box stages beside it have none.
    cand = [(s, ms) for s, ms in totals.items()
            if s in single and kernel_bytes_8d(single[s], 1, 1, 1, 1, 1) > 0]
    dom, dom_ms = max(cand, key=lambda t: t[1])
    kern = single[dom]
    dom_bytes = kernel_bytes_8d(kern, n, P, n_obs, vox, hv) * B
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        try:
            t = json.load(open(tpath)).get(kern)
            if t:
                # hbm_bytes_per_frame is per dispatch (the average over the
                # kernel's launch sites), so one launch over B frames moves
                # B times it.
                traffic = round(t["hbm_bytes_per_frame"] * B)
        except (OSError, ValueError, KeyError):
            traffic = None
    b8 = bytes_8d(n, P, n_obs, vox, hv)
    bext = bytes_extended(n_raw, n, P, n_obs, vox, hv, ncl)
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
                        "(~128k points), noise filter + segment + split + cluster + hull + rotating-calipers "
                        "box, batch of %d frames per step" % B,
            "batch_frames": B,
            "points_per_frame": round(n_raw),
            "noise_search": "radius rule",
            "form": "batched (CSR-packed frames; %d contiguous shares on %d streams, one hipGraph each, "
                    "per step)" % (args.streams, args.streams),
        },
        "points_per_s": round(g["pts"], 1),
        "plain": {k: round(v, 4) if isinstance(v, float) else v for k, v in g["plain"].items()},
        "plain_end_to_end": {k: round(v, 4) if isinstance(v, float) else v for k, v in g["plain_e2e"].items()},
        "batched_one_stream": {k: round(v, 4) for k, v in g["one_stream"].items()},
        "batched_over_plain": round(g["fps"] / g["plain"]["frames_per_s"], 3),
        "stage_ms_per_batch": {s: round(ms, 4) for s, ms in totals.items()},
        "roofline": {
            "bound": "hbm",
            "kernel": kern,
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6),
            "traffic": traffic,
            "launches_per_step": launches[dom],
            "algorithmic_bytes_per_step": round(dom_bytes),
            "algorithmic_bytes_per_launch": round(dom_bytes / launches[dom]),
            "ms_per_step": round(dom_ms, 4),
            "launch_ms": round(dom_ms / launches[dom], 4),
        },
        "pipeline_roofline": {
            "bytes_per_frame": round(b8),
            "achieved_GBps": round(b8 * g["fps"] / 1e9, 3),
            "frac": round(b8 * g["fps"] / 1e9 / HBM_PEAK_GBS, 6),
            "extended_bytes_per_frame": round(bext),
            "extended_frac": round(bext * g["fps"] / 1e9 / HBM_PEAK_GBS, 6),
        },
    }
    line.update(extra)
    if cpu:
        line["cpu_baseline"] = {"value": round(cpu["fps"], 3), "unit": "frames/s", "cores": 1, "kind": "port",
                                "sample": f"{cpu['frames']} frames of the same workload (8 distinct seeds) through "
                                          "the oracle: a fresh reference NoiseRemover per frame (KD-tree search) + "
                                          "segment + split + cluster + gather + hull + box, single thread",
                                "points_per_s": round(cpu["pts"], 1), "p50_ms": round(cpu["p50_ms"], 3)}
        line["speedup_vs_cpu"] = {"batched": round(g["fps"] / cpu["fps"], 2),
                                  "plain": round(g["plain"]["frames_per_s"] / cpu["fps"], 2)}
    if cpu_par:
        line["cpu_all_cores"] = {"value": round(cpu_par["fps"], 3), "unit": "frames/s", "cores": cpu_par["threads"],
                                 "kind": "port", "sample": f"{cpu_par['frames']} frames in {cpu_par['seconds']:.1f} s, "
                                 "frame-parallel, one oracle processor per thread",
                                 "cpu_model": cpu_model()}
        line["speedup_vs_cpu"]["batched_vs_all_cores"] = round(g["fps"] / cpu_par["fps"], 2)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
