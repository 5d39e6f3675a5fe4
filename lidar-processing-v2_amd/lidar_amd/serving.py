"""Multi-sensor serving on the GPU: simulated sensors feed lp_server.

Eight sensors (four 32x1024, three 64x2048, one 128x2048) each deliver a
frame every 100 ms (10 Hz) with a seeded arrival jitter uniform in
[-5, +5] ms.  Frames are submitted at their arrival time (wall clock) and the
server is polled continuously; the latency of a frame is the time from its
scheduled arrival to the poll that returns it with its outputs written.
Frame contents come from a small seeded pool per sensor, cycled, so the
host does not have to synthesise thousands of scans.
"""
import time

import numpy as np

from . import native as N
from . import synth

SENSORS = [32, 32, 32, 32, 64, 64, 64, 128]


def seg_config_for(f):
    return N.segmenter_config(image_height=f.height, image_width=f.width, elevation_up_deg=f.fov_up_deg,
                              elevation_down_deg=f.fov_down_deg)


def sensor_pools(pool_frames=3, sensors=SENSORS):
    """Per sensor: a list of frames and the segmenter configuration of its geometry."""
    pools = []
    for s, kind in enumerate(sensors):
        frames = [synth.workload_sensor(kind, k, s) for k in range(pool_frames)]
        pools.append(dict(kind=kind, frames=frames, cfg=seg_config_for(frames[0])))
    return pools


def arrival_schedule(seconds, sensors=SENSORS, rate_hz=10.0, jitter_ms=5.0, seed=7):
    """Sorted (time_s, sensor, frame_number) arrivals."""
    rng = np.random.default_rng(seed)
    period = 1.0 / rate_hz
    out = []
    for s in range(len(sensors)):
        phase = rng.uniform(0, period)
        n = int(seconds * rate_hz)
        t = phase + np.arange(n) * period + rng.uniform(-jitter_ms, jitter_ms, n) * 1e-3
        out.extend((float(max(ti, 0.0)), s, k) for k, ti in enumerate(t))
    out.sort()
    return out


def serve(server, pools, schedule, keep_results=False, on_done=None):
    """Runs the schedule in real time.  Returns latencies (s) per arrival and,
    with keep_results, each frame's outputs."""
    max_points = max(f.n for p in pools for f in p["frames"])
    free = []
    busy = {}
    lat = np.zeros(len(schedule))
    results = {} if keep_results else None
    t0 = time.perf_counter()
    nxt = 0
    completed = 0
    while completed < len(schedule):
        now = time.perf_counter() - t0
        while nxt < len(schedule) and schedule[nxt][0] <= now:
            _, s, k = schedule[nxt]
            f = pools[s]["frames"][k % len(pools[s]["frames"])]
            slot = free.pop() if free else N.FrameSlot(max_points, server.boxes)
            server.submit(slot, f.x, f.y, f.z, f.ring, pools[s]["cfg"], nxt)
            busy[nxt] = slot
            nxt += 1
        for tag in server.poll():
            t_done = time.perf_counter() - t0
            lat[tag] = t_done - schedule[tag][0]
            slot = busy.pop(tag)
            if keep_results:
                results[tag] = slot.result(server.boxes, server.vehicles)
            if on_done is not None:
                on_done(tag, slot)
            free.append(slot)
            completed += 1
        if nxt < len(schedule) and not busy:
            wait = schedule[nxt][0] - (time.perf_counter() - t0)
            if wait > 0.0005:
                time.sleep(wait - 0.0003)
    wall = time.perf_counter() - t0
    return dict(latency_s=lat, wall_s=wall, results=results)


def summarize(schedule, lat, sensors=SENSORS):
    ms = lat * 1e3
    by_kind = {}
    for kind in sorted(set(sensors)):
        sel = np.array([sensors[s] == kind for _, s, _ in schedule])
        by_kind[f"{kind}-beam"] = dict(frames=int(sel.sum()), p50_ms=round(float(np.percentile(ms[sel], 50)), 3),
                                       p99_ms=round(float(np.percentile(ms[sel], 99)), 3))
    return dict(frames=len(schedule), p50_ms=round(float(np.percentile(ms, 50)), 3),
                p99_ms=round(float(np.percentile(ms, 99)), 3), max_ms=round(float(ms.max()), 3), by_sensor=by_kind)
