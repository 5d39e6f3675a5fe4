"""Launch-shape tuning of a batch: which context layout runs a given batch
fastest on this device.

The knobs are the ones the library exposes at run time:

* ``streams``: the batch runs as that many contiguous shares, each captured as
  its own graph on its own stream, so one share's latency-bound kernels
  (one workgroup per frame or per cluster) overlap the other shares' work;
* ``hash_load_factor``: worst-case load of the voxel and noise-cell hash
  tables; a smaller table is cleared and probed with fewer cache misses.

Workgroup sizes and tiles are compile-time properties of the kernels (256
threads for the per-point kernels, one workgroup per frame for the row
solve, one wavefront per chunk, cluster or listed point elsewhere), chosen
for 64-wide wavefronts and LDS sizes; they are not run-time knobs.

Every candidate is timed the same way: the batch is staged once, its graph
captured by a first run, then ``steps`` replays are timed between two
synchronisations.  Outputs do not depend on the knobs.
"""
import itertools
import time

from lidar_amd import native as N

DEFAULT_STREAMS = (1, 2, 4)
DEFAULT_LOADS = (0.5, 0.7)


def time_batch(frames, cfgs, streams, load, steps=8, noise=True, boxes=True, max_width=2048):
    """ms per batch of `frames` replayed as graphs with the given knobs."""
    maxn = max(f.n for f in frames)
    H = max(f.height for f in frames)
    with N.Context(max_frames=len(frames), max_points=maxn, max_height=H, max_width=max_width, streams=streams,
                   hash_load_factor=load) as ctx:
        if noise:
            ctx.set_noise_filter(N.noise_config())
        ctx.set_oriented_boxes(boxes)
        ctx.stage([(f.x, f.y, f.z, f.ring) for f in frames], cfgs)
        ctx.run(use_graph=True)
        bad = [o["status"] for o in ctx.fetch() if o["status"] != 0]
        if bad:
            raise RuntimeError(f"batch failed under streams={streams} load={load}: {bad}")
        ctx.run(use_graph=True)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.run(use_graph=True)
        ctx.synchronize()
        return (time.perf_counter() - t0) / steps * 1e3


def tune(frames, cfgs, streams=DEFAULT_STREAMS, loads=DEFAULT_LOADS, steps=8, noise=True, boxes=True):
    """Times every (streams, hash_load_factor) pair on the batch; returns
    dict(results=[...], best=dict(streams, hash_load_factor, ms_per_batch))."""
    results = []
    for s, lf in itertools.product(streams, loads):
        if s > len(frames):
            continue
        ms = time_batch(frames, cfgs, s, lf, steps, noise, boxes)
        results.append(dict(streams=s, hash_load_factor=lf, ms_per_batch=round(ms, 4),
                            frames_per_s=round(len(frames) / ms * 1e3, 1)))
    best = min(results, key=lambda r: r["ms_per_batch"])
    return dict(results=results, best=best)
