"""Launch-shape tuning of a batch: which context layout runs a given batch
fastest on this device.

The knobs are the ones the library exposes at run time:

* ``streams``: the batch runs as that many contiguous shares, each captured as
  its own graph on its own stream, so one share's latency-bound kernels
  (one workgroup per frame or per cluster) overlap the other shares' work;
* ``hash_load_factor``: worst-case load of the voxel and noise-cell hash
  tables; a smaller table is cleared and probed with fewer cache misses;
* ``point_workgroup``: workgroup size (64, 128 or 256) of the per-point
  kernels (projection, labels, spherical coordinates, voxel insert);
* ``mark_tile``: columns per workgroup (64, 128 or 256) of the JCP queue
  marking, whose workgroup stages a 5-row LDS tile of that width plus a
  two-pixel halo on each side;
* ``persistent``: the per-point kernels as persistent launches over per-XCD
  (frame, tile) work queues instead of capacity-sized grids.

Every candidate is timed the same way: the batch is staged once, its graph
captured by a first run, then ``steps`` replays are timed between two
synchronisations.  Outputs do not depend on the knobs.  ``tune`` searches
coordinate by coordinate (streams x load factor first, then the workgroup
size, the tile width and the launch form, each at the best of the earlier
ones), so a shape bucket costs about a dozen timings instead of the full
product.
"""
import itertools
import time

from lidar_amd import native as N

DEFAULT_STREAMS = (1, 2, 4)
DEFAULT_LOADS = (0.5, 0.7)
DEFAULT_WORKGROUPS = (64, 128, 256)
DEFAULT_TILES = (64, 128, 256)


def time_batch(frames, cfgs, streams, load, steps=8, noise=True, boxes=True, max_width=2048, point_workgroup=256,
               mark_tile=256, persistent=False):
    """ms per batch of `frames` replayed as graphs with the given knobs."""
    maxn = max(f.n for f in frames)
    H = max(f.height for f in frames)
    with N.Context(max_frames=len(frames), max_points=maxn, max_height=H, max_width=max_width, streams=streams,
                   hash_load_factor=load) as ctx:
        if noise:
            ctx.set_noise_filter(N.noise_config())
        ctx.set_oriented_boxes(boxes)
        ctx.set_launch_shape(point_workgroup, mark_tile)
        ctx.set_persistent(persistent)
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


def tune(frames, cfgs, streams=DEFAULT_STREAMS, loads=DEFAULT_LOADS, workgroups=(256,), tiles=(256,),
         persistent=(False,), steps=8, noise=True, boxes=True):
    """Times (streams, hash_load_factor) pairs on the batch, then each
    workgroup size, tile width and launch form at the best setting so far;
    returns dict(results=[...], best=dict(...knobs..., ms_per_batch))."""
    results = []

    def measure(**knobs):
        ms = time_batch(frames, cfgs, knobs["streams"], knobs["hash_load_factor"], steps, noise, boxes,
                        point_workgroup=knobs["point_workgroup"], mark_tile=knobs["mark_tile"],
                        persistent=knobs["persistent"])
        r = dict(knobs, ms_per_batch=round(ms, 4), frames_per_s=round(len(frames) / ms * 1e3, 1))
        results.append(r)
        return r

    base = dict(point_workgroup=256, mark_tile=256, persistent=False)
    cands = [measure(streams=s, hash_load_factor=lf, **base)
             for s, lf in itertools.product(streams, loads) if s <= len(frames)]
    best = min(cands, key=lambda r: r["ms_per_batch"])
    for key, values in (("point_workgroup", workgroups), ("mark_tile", tiles), ("persistent", persistent)):
        for v in values:
            if v == best[key]:
                continue
            knobs = {k: best[k] for k in ("streams", "hash_load_factor", "point_workgroup", "mark_tile",
                                          "persistent")}
            knobs[key] = v
            r = measure(**knobs)
            if r["ms_per_batch"] < best["ms_per_batch"]:
                best = r
    return dict(results=results, best=best)
