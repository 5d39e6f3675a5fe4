"""Continuous-batching server (GPU): every frame's outputs are the same alone
and in the mix, HBM-budget admission, capacity checks."""
import numpy as np
import pytest

from lidar_amd import native as N
from lidar_amd import serving

from tests.test_gpu_noise_boxes import box_equal

pytestmark = pytest.mark.gpu


def alone(pools, noise, boxes, extras=False):
    """Each pool frame processed on its own context, one frame per batch."""
    ref = {}
    with N.Context(max_frames=1, max_points=262144, max_height=128, max_width=2048) as ctx:
        if noise:
            ctx.set_noise_filter(N.noise_config())
        ctx.set_oriented_boxes(boxes)
        if extras:
            ctx.set_label_clouds(True)
            ctx.set_vehicle_boxes(True)
        for s, p in enumerate(pools):
            for k, f in enumerate(p["frames"]):
                ref[s, k] = ctx.process([(f.x, f.y, f.z, f.ring)], p["cfg"])[0]
    return ref


@pytest.mark.parametrize("noise,boxes,extras", [(False, False, False), (True, True, False), (True, True, True)])
def test_alone_equals_in_mix(noise, boxes, extras):
    pools = serving.sensor_pools(pool_frames=2)
    ref = alone(pools, noise, boxes, extras)
    # 200 Hz per sensor for 0.4 s: frames pile up while batches run, so the
    # slots take mixed batches of all three sensor geometries.
    schedule = serving.arrival_schedule(0.4, rate_hz=200.0, jitter_ms=2.0, seed=3)
    srv = N.Server(slots=3, max_batch_frames=8, noise_cfg=N.noise_config() if noise else None, boxes=boxes,
                   label_clouds=extras, vehicles=extras)
    print(f"serving {len(schedule)} frames noise={noise} boxes={boxes} extras={extras}", flush=True)
    out = serving.serve(srv, pools, schedule, keep_results=True)
    st = srv.stats()
    srv.close()
    assert st["frames"] == len(schedule) and st["batches"] < len(schedule)  # some batches held several frames
    for tag, (_, s, k) in enumerate(schedule):
        got = out["results"][tag]
        want = ref[s, k % len(pools[s]["frames"])]
        assert got["status"] == 0
        for key in ("seg_labels", "obstacle_index", "cluster_labels", "hull_offsets", "hull_indices"):
            assert np.array_equal(got[key], want[key]), (tag, key)
        assert got["num_points"] == want["num_points"]
        if boxes:
            assert len(got["boxes"]) == len(want["boxes"])
            assert all(box_equal(a, b) for a, b in zip(got["boxes"], want["boxes"]))
        if extras:
            for key in ("ground_index", "unknown_index", "vehicle_class"):
                assert np.array_equal(got[key], want[key]), (tag, key)
            assert got["num_vehicles"] == want["num_vehicles"]


def test_hbm_budget_admission():
    srv = N.Server(slots=4, max_batch_frames=2)
    per_slot = srv.stats()["slot_bytes"]
    assert srv.stats()["slots"] == 4
    srv.close()
    srv = N.Server(slots=4, max_batch_frames=2, hbm_budget_bytes=int(per_slot * 2.5))
    assert srv.stats()["slots"] == 2
    srv.close()
    with pytest.raises(N.LidarError):
        N.Server(slots=4, max_batch_frames=2, hbm_budget_bytes=per_slot // 2)


def test_oversized_frame_rejected_at_submit():
    srv = N.Server(slots=1, max_batch_frames=1, max_points=1000, max_height=64, max_width=2048)
    f = serving.sensor_pools(pool_frames=1, sensors=[64])[0]
    x = f["frames"][0]
    slot = N.FrameSlot(x.n)
    with pytest.raises(N.LidarError):
        srv.submit(slot, x.x, x.y, x.z, x.ring, f["cfg"], 0)
    assert srv.poll() == []
    srv.close()


def test_invalid_frame_rejected_at_submit_keeps_the_batch():
    """Frames that staging would refuse (empty image, fewer than four radial
    rings) are refused by submit, so the valid frames around them still run
    in one batch and give their single-frame results."""
    pools = serving.sensor_pools(pool_frames=2, sensors=[32, 64])
    ref = alone(pools, False, False)
    srv = N.Server(slots=1, max_batch_frames=8)
    slots, tags = [], []
    bad_cfgs = [N.segmenter_config(image_height=0), N.segmenter_config(max_distance_m=6.0)]
    tag = 0
    for s, p in enumerate(pools):
        for k, f in enumerate(p["frames"]):
            slot = N.FrameSlot(f.n)
            srv.submit(slot, f.x, f.y, f.z, f.ring, p["cfg"], tag)
            slots.append((slot, s, k))
            tags.append(tag)
            tag += 1
            bad = N.FrameSlot(f.n)
            with pytest.raises(N.LidarError):
                srv.submit(bad, f.x, f.y, f.z, f.ring, bad_cfgs[k % 2], 1000 + tag)
    import time
    done = []
    t_end = time.monotonic() + 60.0
    while len(done) < len(tags) and time.monotonic() < t_end:
        done += srv.poll()
        time.sleep(0.0005)
    st = srv.stats()
    srv.close()
    assert sorted(done) == tags
    assert st["batches"] == 1 and st["frames"] == len(tags)
    for (slot, s, k) in slots:
        got = slot.result()
        want = ref[s, k]
        assert got["status"] == 0
        for key in ("seg_labels", "obstacle_index", "cluster_labels", "hull_offsets", "hull_indices"):
            assert np.array_equal(got[key], want[key]), key
