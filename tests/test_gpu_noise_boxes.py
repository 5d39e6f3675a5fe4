"""Noise filter (DROR) and oriented boxes on the GPU vs the oracle (GPU).

Noise labels are bit-exact against the radius rule of NoiseRemover::filter
(noise_remover.cpp:38-68); boxes are bit-exact (corners, float area, float
angle, validity) against the restatement of boundingBoxRotatingCalipers
(polygonizer.cpp:163-278).  The fused path (filter -> keep VALID in input
order -> segment -> cluster -> hull -> box) is compared stage by stage.
"""
import math

import numpy as np
import pytest

from lidar_amd import native as N
from lidar_amd import synth

from tests.test_gpu_parity import cfgs, hull_coords
from tests.test_oracle_dror_obb import cloud, convex_polygon, strictly_convex_ccw

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(max_frames=4, max_points=262144, max_height=128, max_width=2048)
    yield c
    c.close()


def box_equal(a, b):
    return (a["is_valid"] == b["is_valid"] and np.array_equal(a["corners"].view(np.uint64), b["corners"].view(np.uint64))
            and a["area"].view(np.uint32) == b["area"].view(np.uint32)
            and a["angle_rad"].view(np.uint32) == b["angle_rad"].view(np.uint32))


@pytest.mark.parametrize("fi", [0, 1])
def test_noise_labels_bit_exact_on_full_frames(ctx, oracle, fi):
    f = synth.workload_c2(fi)
    print(f"noise c2 frame {fi} n={f.n}", flush=True)
    st, lab = ctx.remove_noise(f.x, f.y, f.z)
    assert st == 0
    ref, _ = oracle.noise_definition(f.x, f.y, f.z)
    assert np.array_equal(lab, ref)
    assert 1500 < lab.sum() < 4000


@pytest.mark.parametrize("seed", [1, 2])
def test_noise_labels_on_sparse_clouds_and_configs(ctx, oracle, seed):
    x, y, z = cloud(seed)
    for cfgkw in ({}, dict(min_neighbours=2), dict(min_neighbours=0), dict(min_search_radius_m=0.5),
                  dict(radius_multiplier_m_per_m=0.05)):
        st, lab = ctx.remove_noise(x, y, z, N.noise_config(**cfgkw))
        assert st == 0
        ref, _ = oracle.noise_definition(x, y, z, oracle.noise_config(**cfgkw))
        assert np.array_equal(lab, ref), cfgkw


def test_noise_edge_cases(ctx, oracle):
    st, lab = ctx.remove_noise(np.zeros(0, np.float32), np.zeros(0, np.float32), np.zeros(0, np.float32))
    assert st == 0 and len(lab) == 0
    one = np.array([3.0], np.float32)
    st, lab = ctx.remove_noise(one, one, one)
    assert st == 0 and lab.tolist() == [1]
    # Far points (radius beyond the cell cube), duplicates and a NaN.
    x = np.array([400.0, 400.0, 400.5, 401.0, 1e6, 0.0, 0.0, 0.0, 0.0, np.nan], np.float32)
    y = np.array([0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0], np.float32)
    z = np.zeros(len(x), np.float32)
    st, lab = ctx.remove_noise(x, y, z)
    assert st == 0
    ref, _ = oracle.noise_definition(x, y, z)
    assert np.array_equal(lab, ref)


def test_boxes_bit_exact_on_random_polygons(ctx, oracle):
    rng = np.random.default_rng(5)
    polys = [convex_polygon(rng, int(rng.integers(3, 80))) for _ in range(300)]
    w, h, t = 4.5, 1.8, 0.7
    loc = np.array([[-w / 2, -h / 2], [w / 2, -h / 2], [w / 2, h / 2], [-w / 2, h / 2]])
    R = np.array([[math.cos(t), -math.sin(t)], [math.sin(t), math.cos(t)]])
    P = loc @ R.T + [3.0, 4.0]
    polys.append(strictly_convex_ccw(P[:, 0], P[:, 1]))
    polys.append((np.array([0.0, 1.0]), np.array([0.0, 0.0])))  # fewer than 3 vertices: invalid
    st, boxes = ctx.oriented_boxes(polys)
    assert st == 0
    for k, (px, py) in enumerate(polys):
        ref = oracle.obb_rotating_calipers(px, py)
        assert box_equal(boxes[k], ref), k


def test_boxes_bit_exact_on_large_and_symmetric_polygons(ctx, oracle):
    """Hundreds of vertices (long antipodal runs), regular polygons with exact
    symmetric coordinates (equal triangle areas take the walk's equality
    branch, equal edge areas test the running-minimum rule) and lattice
    polygons."""
    rng = np.random.default_rng(9)
    polys = [convex_polygon(rng, int(rng.integers(200, 1000))) for _ in range(12)]
    for m in (4, 6, 8, 12, 16, 64, 360):
        a = np.arange(m) * (2 * math.pi / m)
        polys.append(strictly_convex_ccw(np.round(np.cos(a) * 5.0, 6) + 2.0, np.round(np.sin(a) * 5.0, 6) - 1.0))
    polys.append(strictly_convex_ccw(np.array([0.0, 4.0, 4.0, 0.0]), np.array([0.0, 0.0, 2.0, 2.0])))
    polys.append(strictly_convex_ccw(np.array([0.0, 2.0, 3.0, 2.0, 0.0, -1.0]), np.array([0.0, 0.0, 1.0, 2.0, 2.0, 1.0])))
    polys.append(strictly_convex_ccw(np.array([0.0, 1.0, 1.0, 0.0]), np.array([0.0, 0.0, 1.0, 1.0])))
    st, boxes = ctx.oriented_boxes(polys)
    assert st == 0
    for k, (px, py) in enumerate(polys):
        ref = oracle.obb_rotating_calipers(px, py)
        assert box_equal(boxes[k], ref), (k, len(px))


def full_frame_reference(oracle, f, sco):
    noise, _ = oracle.noise_definition(f.x, f.y, f.z)
    keep = noise == 0
    x, y, z, ring = f.x[keep], f.y[keep], f.z[keep], f.ring[keep]
    ref = oracle.process_frame(x, y, z, ring, scfg=sco)
    obs = np.nonzero(ref["seg_labels"] == 2)[0]
    ox, oy = x[obs].astype(np.float64), y[obs].astype(np.float64)
    hulls = hull_coords(ox, oy, ref["cluster_labels"], ref["num_clusters"], ref["hull_offsets"],
                        ref["hull_indices"])
    boxes = [oracle.obb_rotating_calipers(h[:, 0].copy(), h[:, 1].copy()) for h in hulls]
    return noise, (x, y), ref, hulls, boxes


def test_fused_filter_segment_cluster_hull_boxes(ctx, oracle):
    frames = [synth.workload_c2(0), synth.workload_c2(3)]
    ctx.set_noise_filter(N.noise_config())
    ctx.set_oriented_boxes(True)
    try:
        for use_graph in (False, True, True):
            outs = ctx.process([(f.x, f.y, f.z, f.ring) for f in frames], [cfgs(oracle, f)[1] for f in frames],
                               use_graph=use_graph)
            for f, out in zip(frames, outs):
                sco, _ = cfgs(oracle, f)
                noise, (x, y), ref, hulls, boxes = full_frame_reference(oracle, f, sco)
                assert out["status"] == 0
                assert np.array_equal(out["noise_labels"], noise)
                assert out["num_points"] == len(x) and out["num_noise"] == int(noise.sum())
                assert np.array_equal(out["seg_labels"], ref["seg_labels"])
                assert np.array_equal(out["cluster_labels"], ref["cluster_labels"])
                obs = out["obstacle_index"]
                gh = hull_coords(x[obs].astype(np.float64), y[obs].astype(np.float64), out["cluster_labels"],
                                 out["num_clusters"], out["hull_offsets"], out["hull_indices"])
                assert len(gh) == len(hulls)
                assert all(np.array_equal(a, b) for a, b in zip(gh, hulls))
                assert len(out["boxes"]) == len(boxes)
                assert all(box_equal(a, b) for a, b in zip(out["boxes"], boxes))
    finally:
        ctx.set_noise_filter(None)
        ctx.set_oriented_boxes(False)


def _all_noise_frame(like, n=60):
    """Points on a widening spiral, each metres from the next: all DROR noise."""
    ang = np.arange(n) * (2 * np.pi / n)
    rad = 10.0 + 1.5 * np.arange(n)
    x = (rad * np.cos(ang)).astype(np.float32)
    y = (rad * np.sin(ang)).astype(np.float32)
    z = np.full(n, -1.0, np.float32)
    ring = (np.arange(n) % like.height).astype(np.uint16)
    return synth.Frame(x, y, z, z.copy(), ring, like.height, like.width, like.fov_up_deg, like.fov_down_deg)


def test_fused_batch_with_empty_and_all_noise_frames(ctx, oracle):
    """Frames with no input and with no kept point share a batch with a full
    frame: the keep, obstacle, queue and cluster compactions see zero items
    for them, and the full frame's results do not change."""
    full = synth.workload_c2(1)
    e = np.zeros(0, np.float32)
    empty = synth.Frame(e, e, e, e, np.zeros(0, np.uint16), full.height, full.width, full.fov_up_deg,
                        full.fov_down_deg)
    frames = [empty, full, _all_noise_frame(full)]
    ctx.set_noise_filter(N.noise_config())
    ctx.set_oriented_boxes(True)
    try:
        for use_graph in (False, True, True):
            outs = ctx.process([(f.x, f.y, f.z, f.ring) for f in frames], [cfgs(oracle, f)[1] for f in frames],
                               use_graph=use_graph)
            for f, out in zip(frames, outs):
                sco, _ = cfgs(oracle, f)
                noise, (x, y), ref, hulls, boxes = full_frame_reference(oracle, f, sco)
                assert out["status"] == 0
                assert np.array_equal(out["noise_labels"], noise)
                assert out["num_points"] == len(x) and out["num_noise"] == int(noise.sum())
                assert np.array_equal(out["seg_labels"], ref["seg_labels"])
                assert np.array_equal(out["cluster_labels"], ref["cluster_labels"])
                assert out["num_clusters"] == ref["num_clusters"]
                obs = out["obstacle_index"]
                gh = hull_coords(x[obs].astype(np.float64), y[obs].astype(np.float64), out["cluster_labels"],
                                 out["num_clusters"], out["hull_offsets"], out["hull_indices"])
                assert len(gh) == len(hulls)
                assert all(np.array_equal(a, b) for a, b in zip(gh, hulls))
                assert all(box_equal(a, b) for a, b in zip(out["boxes"], boxes))
            assert outs[0]["num_points"] == 0 and outs[2]["num_points"] == 0
            assert outs[2]["num_noise"] == frames[2].n
    finally:
        ctx.set_noise_filter(None)
        ctx.set_oriented_boxes(False)


def test_filter_off_again_restores_plain_path(ctx, oracle):
    f = synth.workload_c1(0)
    sco, scg = cfgs(oracle, f)
    out = ctx.process([(f.x, f.y, f.z, f.ring)], [scg])[0]
    ref = oracle.process_frame(f.x, f.y, f.z, f.ring, scfg=sco)
    assert out["num_points"] == f.n and out["num_noise"] == 0 and out["boxes"] == []
    assert np.array_equal(out["seg_labels"], ref["seg_labels"])


def test_shared_hash_stays_clean_across_stages(ctx, oracle):
    """The noise cells and the voxels share one hash table per frame, which
    each stage leaves clean for the next (only its own slots are reset).
    Alternating the stages on one context must give what each gives alone."""
    f2, f1 = synth.workload_c2(1), synth.workload_c1(2)
    ref_noise, _ = oracle.noise_definition(f2.x, f2.y, f2.z)
    sco, scg = cfgs(oracle, f1)
    ref_plain = oracle.process_frame(f1.x, f1.y, f1.z, f1.ring, scfg=sco)
    for round_ in range(2):
        print(f"shared hash round {round_}", flush=True)
        st, lab = ctx.remove_noise(f2.x, f2.y, f2.z)
        assert st == 0 and np.array_equal(lab, ref_noise)
        out = ctx.process([(f1.x, f1.y, f1.z, f1.ring)], [scg], use_graph=bool(round_))[0]
        assert out["status"] == 0
        assert np.array_equal(out["seg_labels"], ref_plain["seg_labels"])
        assert np.array_equal(out["cluster_labels"], ref_plain["cluster_labels"])
        ctx.set_noise_filter(N.noise_config())
        try:
            fused = ctx.process([(f2.x, f2.y, f2.z, f2.ring)], [cfgs(oracle, f2)[1]], use_graph=bool(round_))[0]
        finally:
            ctx.set_noise_filter(None)
        assert fused["status"] == 0 and np.array_equal(fused["noise_labels"], ref_noise)


def test_cpp_dropin_full_pipeline(tmp_path, oracle):
    """NoiseRemover -> Segmenter -> Clusterer -> Polygonizer (hulls and boxes)
    through the drop-in C++ classes (cpp/processor_example.cpp, mode full)
    against the oracle's full frame with the radius rule."""
    import os
    import subprocess
    f = synth.workload_c2(5)
    rec = np.zeros(f.n, dtype=np.dtype({"names": ["x", "y", "z", "intensity", "ring"],
                                        "formats": ["<f4", "<f4", "<f4", "<f4", "<u2"],
                                        "offsets": [0, 4, 8, 16, 20], "itemsize": 32}))
    rec["x"], rec["y"], rec["z"], rec["intensity"], rec["ring"] = f.x, f.y, f.z, f.intensity, f.ring
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(src, "wb") as fh:
        fh.write(np.int32(f.n).tobytes())
        fh.write(rec.tobytes())
    exe = os.path.join(os.path.dirname(N.library_path()), "processor_example")
    subprocess.run([exe, str(src), str(dst), "full"], check=True, timeout=120)
    raw = open(dst, "rb").read()
    pos = 0

    def take(dtype, count):
        nonlocal pos
        a = np.frombuffer(raw, dtype, count, pos)
        pos += a.nbytes
        return a

    n = int(take(np.int32, 1)[0])
    noise = take(np.uint8, f.n)
    seg = take(np.uint32, n)
    no = int(take(np.int32, 1)[0])
    clab = take(np.int32, no)
    k = int(take(np.int32, 1)[0])
    hoff = take(np.int32, k + 1)
    hidx = take(np.int32, int(hoff[-1]))
    boxes_raw = take(np.dtype([("c", "<f8", 8), ("a", "<f4"), ("g", "<f4"), ("v", "<i4")]), k)
    assert pos == len(raw)
    proc = oracle.Processor()
    ref = oracle.full_frame(proc, f.x, f.y, f.z, f.ring)
    assert np.array_equal(noise, ref["noise_labels"])
    assert n == ref["num_points"]
    assert np.array_equal(seg, ref["seg_labels"])
    assert np.array_equal(clab, ref["cluster_labels"])
    keep = noise == 0
    x, y = f.x[keep], f.y[keep]
    obs = np.flatnonzero(seg == 2)
    ox, oy = x[obs].astype(np.float64), y[obs].astype(np.float64)
    hg = hull_coords(ox, oy, clab, k, hoff, hidx)
    hr = hull_coords(ox, oy, ref["cluster_labels"], k, ref["hull_offsets"], ref["hull_indices"])
    assert all(np.array_equal(a, b) for a, b in zip(hg, hr))
    for c in range(k):
        rb = ref["boxes"][c]
        assert boxes_raw[c]["v"] == rb.is_valid
        assert np.array_equal(boxes_raw[c]["c"].view(np.uint64), np.array(rb.corners[:], np.float64).view(np.uint64))
        assert boxes_raw[c]["a"] == np.float32(rb.area) and boxes_raw[c]["g"] == np.float32(rb.angle_rad)
