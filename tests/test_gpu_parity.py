"""GPU path vs the oracle, through the C ABI (GPU).

Integer and label outputs must be bit-exact; the RANSAC plane (float) is
required to be bit-exact too, which is stricter than the 1e-5 relative bound
the plane is specified with.  Hull vertex lists are compared as coordinate
sequences: among points with identical (x, y) the reference keeps whichever
index its unstable std::sort happens to place (polygonizer.cpp:52-59).
"""
import numpy as np
import pytest

from lidar_amd import native as N
from lidar_amd import synth

from tests.test_oracle_segmenter import build

pytestmark = pytest.mark.gpu

UNKNOWN, GROUND, OBSTACLE = 0, 1, 2


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(max_frames=6, max_points=262144, max_height=128, max_width=2048)
    yield c
    c.close()


def cfgs(oracle, f):
    kw = dict(image_height=f.height, image_width=f.width, elevation_up_deg=f.fov_up_deg,
              elevation_down_deg=f.fov_down_deg)
    return oracle.seg_config(**kw), N.segmenter_config(**kw)


_FRAMES = {}


def frame(name):
    if name not in _FRAMES:
        _FRAMES[name] = {"c1": lambda: synth.workload_c1(0), "c2": lambda: synth.workload_c2(0),
                         "c4": lambda: synth.workload_c4(0), "c1b": lambda: synth.workload_c1(7)}[name]()
    return _FRAMES[name]


def hull_coords(x, y, labels, K, hoff, hidx):
    out = []
    order = np.argsort(labels, kind="stable")
    starts = np.searchsorted(labels[order], np.arange(K + 1))
    for c in range(K):
        m = order[starts[c]:starts[c + 1]]
        idx = hidx[hoff[c]:hoff[c + 1]]
        out.append(np.stack([x[m][idx], y[m][idx]], 1))
    return out


@pytest.mark.parametrize("name", ["c1", "c2", "c4"])
def test_segment_stages_bit_exact(ctx, oracle, name):
    f = frame(name)
    sco, scg = cfgs(oracle, f)
    st, lab_o, dbg = oracle.segment(f.x, f.y, f.z, f.ring, cfg=sco, debug=True)
    stg, lab_g, rr = ctx.segment(f.x, f.y, f.z, f.ring, cfg=scg)
    assert st == 0 and stg == 0
    assert np.array_equal(ctx.debug(0, N.DEBUG_ELEVATION_MAP).view(np.uint32), dbg["elevation_map"].view(np.uint32))
    assert np.array_equal(ctx.debug(0, N.DEBUG_MAPPING), dbg["mapping"])
    assert np.array_equal(ctx.debug(0, N.DEBUG_STATE_RECM), dbg["state_recm"])
    assert np.array_equal(ctx.debug(0, N.DEBUG_JCP_QUEUE), dbg["queue"])
    assert np.array_equal(ctx.debug(0, N.DEBUG_STATE_FINAL), dbg["state_final"])
    assert np.array_equal(lab_g, lab_o)
    assert rr.candidates == dbg["ransac_candidates"] and rr.best_inliers == dbg["ransac_best_inliers"]
    assert rr.applied == dbg["ransac_applied"]
    assert np.array_equal(np.array(list(rr.plane), np.float32).view(np.uint32), dbg["ransac_plane"].view(np.uint32))


@pytest.mark.parametrize("name", ["c1", "c2", "c4", "c1b"])
def test_full_frame_bit_exact(ctx, oracle, name):
    f = frame(name)
    sco, scg = cfgs(oracle, f)
    ref = oracle.process_frame(f.x, f.y, f.z, f.ring, scfg=sco)
    out = ctx.process([(f.x, f.y, f.z, f.ring)], scg)[0]
    assert ref["status"] == 0 and out["status"] == 0
    assert np.array_equal(out["seg_labels"], ref["seg_labels"])
    assert np.array_equal(out["obstacle_index"], np.flatnonzero(ref["seg_labels"] == OBSTACLE))
    assert out["num_clusters"] == ref["num_clusters"]
    assert np.array_equal(out["cluster_labels"], ref["cluster_labels"])
    obs = out["obstacle_index"]
    ox, oy = f.x[obs].astype(np.float64), f.y[obs].astype(np.float64)
    K = ref["num_clusters"]
    hg = hull_coords(ox, oy, out["cluster_labels"], K, out["hull_offsets"], out["hull_indices"])
    hr = hull_coords(ox, oy, ref["cluster_labels"], K, ref["hull_offsets"], ref["hull_indices"])
    bad = [c for c in range(K) if not np.array_equal(hg[c], hr[c])]
    assert not bad, f"{len(bad)} hulls differ, first {bad[:5]}"


def _posts():
    """Vertical posts: every beam of a column returns the same float (x, y).
    (azimuth deg, distance m, columns, first ring, last ring)."""
    xs, ys, zs, rs = [], [], [], []
    for az, d, ncol, r0, r1 in [(30.0, 12.0, 1, 20, 50), (75.0, 9.0, 2, 25, 45), (130.0, 15.0, 4, 22, 40),
                                (200.0, 7.0, 3, 30, 31), (250.0, 11.0, 1, 24, 44), (300.0, 6.0, 1, 33, 34)]:
        for c in range(ncol):
            a = np.deg2rad(az + 0.4 * c)
            x, y = np.float32(d * np.cos(a)), np.float32(d * np.sin(a))
            for r in range(r0, r1 + 1):
                xs.append(x)
                ys.append(y)
                zs.append(np.float32(-1.2 + 0.08 * (r - r0)))
                rs.append(r)
    return np.array(xs, np.float32), np.array(ys, np.float32), np.array(zs, np.float32), np.array(rs, np.uint16)


def test_posts_of_repeated_points_bit_exact(ctx, oracle):
    """Clusters whose points repeat one (x, y) down a range-image column: a
    one-column post's hull is its point listed twice, as in the reference's
    chain; posts of 2-4 columns give 2-4 vertices."""
    f = frame("c1")
    px, py, pz, pr = _posts()
    x, y, z, ring = np.r_[f.x, px], np.r_[f.y, py], np.r_[f.z, pz], np.r_[f.ring, pr]
    sco, scg = cfgs(oracle, f)
    ref = oracle.process_frame(x, y, z, ring, scfg=sco)
    out = ctx.process([(x, y, z, ring)], scg)[0]
    assert ref["status"] == 0 and out["status"] == 0
    assert np.array_equal(out["seg_labels"], ref["seg_labels"])
    assert np.array_equal(out["cluster_labels"], ref["cluster_labels"])
    obs = out["obstacle_index"]
    ox, oy = x[obs].astype(np.float64), y[obs].astype(np.float64)
    K = ref["num_clusters"]
    hg = hull_coords(ox, oy, out["cluster_labels"], K, out["hull_offsets"], out["hull_indices"])
    hr = hull_coords(ox, oy, ref["cluster_labels"], K, ref["hull_offsets"], ref["hull_indices"])
    bad = [c for c in range(K) if not np.array_equal(hg[c], hr[c])]
    assert not bad, f"{len(bad)} hulls differ, first {bad[:5]}"
    # The posts reach the hull stage: one hull is a single point twice.
    assert any(len(h) == 2 and np.array_equal(h[0], h[1]) for h in hr)


def _known_answer_cases():
    return [
        # in-order updates (Gauss-Seidel), expected labels hand-derived
        ([(10, 600, 20.0, -1.73), (9, 602, 20.0, -1.73), (11, 602, 20.0, -1.33), (10, 598, 20.0, -1.73)],
         [OBSTACLE, OBSTACLE, OBSTACLE, GROUND]),
        # stale border slots
        ([(20, 5, 8.2, -1.73), (20, 3, 8.2, -1.52), (30, 0, 50.0, -1.73), (28, 2, 50.0, -1.58), (32, 2, 50.0, 1.5)],
         [OBSTACLE, OBSTACLE, OBSTACLE, GROUND, OBSTACLE]),
        # unresolved queue pixel stays UNKNOWN
        ([(40, 1000, 20.0, -1.73), (41, 1000, 20.0, 1.27), (42, 1000, 20.0, 1.30)], [UNKNOWN, OBSTACLE, OBSTACLE]),
    ]


@pytest.mark.parametrize("case", range(3))
def test_jcp_known_answers(ctx, case):
    pts, expect = _known_answer_cases()[case]
    x, y, z, r = build(pts)
    st, lab, _ = ctx.segment(x, y, z, r)
    assert st == 0
    assert list(lab) == expect


def test_ransac_single_candidate_is_an_error(ctx):
    x, y, z, r = build([(2, 300, 5.0, -1.73), (30, 900, 20.0, -1.73), (31, 900, 20.0, -1.70)])
    st, lab, _ = ctx.segment(x, y, z, r)
    assert st == N.ERR_RANSAC_SINGLE_CANDIDATE


def test_empty_and_tiny_clouds(ctx, oracle):
    e = np.zeros(0, np.float32)
    st, lab, rr = ctx.segment(e, e, e, np.zeros(0, np.uint16))
    assert st == 0 and len(lab) == 0 and rr.applied == 0
    x, y, z, r = build([(10, 700, 20.0, -1.73), (10, 700, 20.5, -1.73)])
    st, lab, _ = ctx.segment(x, y, z, r)
    assert list(lab) == [GROUND, UNKNOWN]


def test_batch_with_uneven_frames(ctx, oracle):
    frames = [synth.workload_c3(i) for i in range(5)]
    assert len({f.width for f in frames}) >= 2
    pairs = [cfgs(oracle, f) for f in frames]
    outs = ctx.process([(f.x, f.y, f.z, f.ring) for f in frames], [p[1] for p in pairs])
    for f, (sco, _), out in zip(frames, pairs, outs):
        ref = oracle.process_frame(f.x, f.y, f.z, f.ring, scfg=sco)
        assert out["status"] == 0
        assert np.array_equal(out["seg_labels"], ref["seg_labels"])
        assert np.array_equal(out["cluster_labels"], ref["cluster_labels"])
        assert out["num_clusters"] == ref["num_clusters"]


@pytest.mark.parametrize("streams,nframes", [(2, 5), (3, 2), (4, 7)])
def test_multi_stream_batch_matches_single(ctx, oracle, streams, nframes):
    """A batch split over several streams (contiguous shares, uneven and
    with idle lanes) gives the single-stream results for every frame, with
    the noise filter and the boxes on, eager and replayed."""
    frames = [synth.workload_c3(i) for i in range(nframes)]
    scgs = [cfgs(oracle, f)[1] for f in frames]
    ins = [(f.x, f.y, f.z, f.ring) for f in frames]
    keys = ("seg_labels", "cluster_labels", "hull_offsets", "hull_indices", "noise_labels")
    ref = N.Context(max_frames=8, max_points=max(f.n for f in frames), max_height=64, max_width=2048)
    ms = N.Context(max_frames=8, max_points=max(f.n for f in frames), max_height=64, max_width=2048,
                   streams=streams)
    try:
        for c in (ref, ms):
            c.set_noise_filter(N.noise_config())
            c.set_oriented_boxes(True)
            c.stage(ins, scgs)
        ref.run(use_graph=False)
        want = ref.fetch()
        for use_graph in (False, True, True):
            ms.run(use_graph=use_graph)
            got = ms.fetch()
            assert len(got) == nframes
            for a, b in zip(want, got):
                assert a["status"] == 0 and b["status"] == 0
                for k in keys:
                    assert np.array_equal(a[k], b[k]), k
                assert all(box_a == box_b for box_a, box_b in zip(
                    [tuple(np.asarray(x["corners"]).tolist()) for x in a["boxes"]],
                    [tuple(np.asarray(x["corners"]).tolist()) for x in b["boxes"]]))
    finally:
        ref.close()
        ms.close()


def test_graph_replay_matches_eager(ctx, oracle):
    frames = [frame("c1"), frame("c1b")]
    scg = cfgs(oracle, frames[0])[1]
    ctx.stage([(f.x, f.y, f.z, f.ring) for f in frames], scg)
    ctx.run(use_graph=False)
    eager = ctx.fetch()
    for _ in range(3):
        ctx.run(use_graph=True)
        got = ctx.fetch()
        for a, b in zip(eager, got):
            for k in ("seg_labels", "cluster_labels", "hull_offsets", "hull_indices"):
                assert np.array_equal(a[k], b[k]), k
        # eager work between replays
        ctx.run(use_graph=False)
        ctx.synchronize()


def test_repeatable(ctx, oracle):
    f = frame("c2")
    scg = cfgs(oracle, f)[1]
    a = ctx.process([(f.x, f.y, f.z, f.ring)], scg)[0]
    b = ctx.process([(f.x, f.y, f.z, f.ring)], scg)[0]
    for k in ("seg_labels", "cluster_labels", "hull_offsets", "hull_indices"):
        assert np.array_equal(a[k], b[k])


def test_aos_point_records(ctx, oracle):
    """pcl::PointXYZIR layout: x, y, z, pad, intensity, ring (u16), 32-byte records."""
    f = frame("c1")
    rec = np.zeros(f.n, dtype=np.dtype({"names": ["x", "y", "z", "intensity", "ring"],
                                        "formats": ["<f4", "<f4", "<f4", "<f4", "<u2"],
                                        "offsets": [0, 4, 8, 16, 20], "itemsize": 32}))
    rec["x"], rec["y"], rec["z"], rec["intensity"], rec["ring"] = f.x, f.y, f.z, f.intensity, f.ring
    cloud = N.aos_cloud(rec)
    st, lab, _ = ctx.segment(None, None, None, cfg=cfgs(oracle, f)[1], cloud=cloud)
    st2, lab2, _ = ctx.segment(f.x, f.y, f.z, f.ring, cfg=cfgs(oracle, f)[1])
    assert st == 0 and np.array_equal(lab, lab2)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_standalone_cluster(ctx, oracle, seed):
    rng = np.random.default_rng(seed)
    n = 5000
    blobs = rng.uniform(-40, 40, (30, 2))
    idx = rng.integers(0, 30, n)
    x = (blobs[idx, 0] + rng.normal(0, 1.0, n)).astype(np.float32)
    y = (blobs[idx, 1] + rng.normal(0, 1.0, n)).astype(np.float32)
    z = rng.uniform(-1.5, 2.0, n).astype(np.float32)
    st, lab = ctx.cluster(x, y, z)
    sto, labo = oracle.cluster(x, y, z)
    assert st == 0 and sto == 0
    assert np.array_equal(lab, labo)


def test_standalone_cluster_many_voxels(oracle):
    """More occupied voxels than the LDS union-find holds (32768): the
    global-memory union-find and its in-place compression run."""
    rng = np.random.default_rng(7)
    blobs = rng.uniform(-50, 50, (40, 2))
    idx = rng.integers(0, 40, 30000)
    bx = blobs[idx, 0] + rng.normal(0, 1.5, 30000)
    by = blobs[idx, 1] + rng.normal(0, 1.5, 30000)
    sx, sy = rng.uniform(-70, 70, 60000), rng.uniform(-70, 70, 60000)
    x = np.concatenate([bx, sx]).astype(np.float32)
    y = np.concatenate([by, sy]).astype(np.float32)
    z = rng.uniform(-1.5, 3.0, len(x)).astype(np.float32)
    perm = rng.permutation(len(x))
    x, y, z = x[perm], y[perm], z[perm]
    print(f"many voxels: {len(x)} points", flush=True)
    with N.Context(max_points=131072) as c:
        st, lab = c.cluster(x, y, z)
    sto, labo = oracle.cluster(x, y, z)
    assert st == 0 and sto == 0
    assert np.array_equal(lab, labo)
    assert lab.max() > 10


def test_standalone_cluster_wrap_and_overflow(oracle):
    x = np.array([-10.0, -10.0, -10.0, 10.0, 10.0, 10.0], np.float32)
    y = np.array([0.0, 0.0, 0.0, 0.0873, 0.0873, 0.0873], np.float32)
    z = np.array([0.0, 0.05, 0.1, 0.0, 0.05, 0.1], np.float32)
    with N.Context(max_points=4096, max_voxels=10) as c:
        st, lab = c.cluster(x, y, z)
        assert st == 0 and list(lab) == [0] * 6
        xs = np.linspace(5, 60, 40).astype(np.float32)
        st, _ = c.cluster(xs, np.ones(40, np.float32), np.zeros(40, np.float32))
        assert st == N.ERR_VOXEL_OVERFLOW


def test_standalone_hulls(ctx, oracle):
    rng = np.random.default_rng(11)
    lists = []
    for k in range(40):
        n = int(rng.integers(1, 3000))
        x = rng.normal(0, 3, n).astype(np.float32).astype(np.float64)
        y = rng.normal(0, 1, n).astype(np.float32).astype(np.float64)
        if k % 5 == 0:
            y = x * 0.5  # collinear
        if k % 7 == 0:
            x[:] = x[0]
            y[:] = y[0]  # identical
        if k % 4 == 1:
            x = x + rng.normal(0, 1e-3, n)  # full double precision (not floats)
        lists.append((x, y))
    st, hulls = ctx.convex_hulls(lists)
    assert st == 0
    for (x, y), h in zip(lists, hulls):
        r = oracle.convex_hull(x, y)
        assert np.array_equal(np.stack([x[h], y[h]], 1), np.stack([x[r], y[r]], 1))


def test_standalone_hulls_duplicates(ctx, oracle):
    """Points repeated many times over (a vertical surface gives one (x, y)
    per beam): same vertex coordinates as the reference chain, and each index
    is the smallest with its coordinates."""
    rng = np.random.default_rng(12)
    lists = []
    for k in range(30):
        m = int(rng.integers(1, 200))
        rep = int(rng.integers(2, 33))
        bx = rng.normal(0, 3, m).astype(np.float32).astype(np.float64)
        by = rng.normal(0, 1, m).astype(np.float32).astype(np.float64)
        if k % 3 == 0:
            by = bx * 0.25  # collinear columns
        perm = rng.permutation(m * rep)
        lists.append((np.repeat(bx, rep)[perm], np.repeat(by, rep)[perm]))
    # Two distinct points, repeated, the larger one listed first.
    lists.append((np.array([2.0, 1.0, 2.0, 1.0]), np.zeros(4)))
    lists.append((np.array([5.0, 5.0, 5.0, 1.0]), np.ones(4)))
    # Chunks that each hold one repeated point.
    lists.append((np.r_[np.full(300, 3.0), np.full(100, -1.0)], np.r_[np.full(300, 2.0), np.full(100, 4.0)]))
    lists.append((np.r_[np.full(100, -1.0), np.full(300, 3.0)], np.r_[np.full(100, 4.0), np.full(300, 2.0)]))
    lists.append((np.full(700, 1.5), np.full(700, -2.5)))
    st, hulls = ctx.convex_hulls(lists)
    assert st == 0
    for (x, y), h in zip(lists, hulls):
        r = oracle.convex_hull(x, y)
        assert np.array_equal(np.stack([x[h], y[h]], 1), np.stack([x[r], y[r]], 1))
        for i in h:
            assert i == np.flatnonzero((x == x[i]) & (y == y[i]))[0]


def test_cpp_dropin_processor(tmp_path, oracle):
    """The reference processor's calls compiled against the drop-in C++
    classes (cpp/processor_example.cpp) give the oracle's results."""
    import os
    import subprocess
    f = frame("c1b")
    rec = np.zeros(f.n, dtype=np.dtype({"names": ["x", "y", "z", "intensity", "ring"],
                                        "formats": ["<f4", "<f4", "<f4", "<f4", "<u2"],
                                        "offsets": [0, 4, 8, 16, 20], "itemsize": 32}))
    rec["x"], rec["y"], rec["z"], rec["intensity"], rec["ring"] = f.x, f.y, f.z, f.intensity, f.ring
    src = tmp_path / "in.bin"
    dst = tmp_path / "out.bin"
    with open(src, "wb") as fh:
        fh.write(np.int32(f.n).tobytes())
        fh.write(rec.tobytes())
    exe = os.path.join(os.path.dirname(N.library_path()), "processor_example")
    subprocess.run([exe, str(src), str(dst)], check=True, timeout=120)
    buf = np.fromfile(dst, np.int32)
    seg = buf[:f.n].view(np.uint32)
    no = buf[f.n]
    clab = buf[f.n + 1:f.n + 1 + no]
    k = buf[f.n + 1 + no]
    hoff = buf[f.n + 2 + no:f.n + 3 + no + k]
    hidx = buf[f.n + 3 + no + k:]
    ref = oracle.process_frame(f.x, f.y, f.z, f.ring)
    assert np.array_equal(seg, ref["seg_labels"])
    assert np.array_equal(clab, ref["cluster_labels"])
    obs = np.flatnonzero(ref["seg_labels"] == OBSTACLE)
    ox, oy = f.x[obs].astype(np.float64), f.y[obs].astype(np.float64)
    hg = hull_coords(ox, oy, clab, k, hoff, hidx)
    hr = hull_coords(ox, oy, ref["cluster_labels"], k, ref["hull_offsets"], ref["hull_indices"])
    assert all(np.array_equal(a, b) for a, b in zip(hg, hr))
    # The same records handed over as a PointCloud2 message (CloudView, no
    # conversion to a point cloud): identical output file.
    dst2 = tmp_path / "out_msg.bin"
    subprocess.run([exe, str(src), str(dst2), "msg"], check=True, timeout=120)
    assert open(dst, "rb").read() == open(dst2, "rb").read()
    # Segmenter::image() after the same segment call: the oracle's image_.
    dst3 = tmp_path / "out_image.bin"
    subprocess.run([exe, str(src), str(dst3), "image"], check=True, timeout=120)
    body = open(dst, "rb").read()
    tail = open(dst3, "rb").read()
    assert tail[:len(body)] == body
    rows, cols = np.frombuffer(tail[len(body):len(body) + 8], np.int32)
    assert (rows, cols) == (f.height, f.width)
    _, _, dbg = oracle.segment(f.x, f.y, f.z, f.ring, debug=True)
    assert np.array_equal(np.frombuffer(tail[len(body) + 8:], np.uint8), dbg["image"])


def test_cpp_dropin_grows_past_the_reserve(tmp_path, oracle):
    """A 128x2048 scan of more than 200 000 points (the reference's reserve
    hint, segmenter.cpp:62) through the drop-in classes: the context grows
    for the call instead of failing, and the results equal the oracle's."""
    import os
    import subprocess
    f = frame("c4")
    assert f.n > 200000
    rec = np.zeros(f.n, dtype=np.dtype({"names": ["x", "y", "z", "intensity", "ring"],
                                        "formats": ["<f4", "<f4", "<f4", "<f4", "<u2"],
                                        "offsets": [0, 4, 8, 16, 20], "itemsize": 32}))
    rec["x"], rec["y"], rec["z"], rec["intensity"], rec["ring"] = f.x, f.y, f.z, f.intensity, f.ring
    src = tmp_path / "in.bin"
    dst = tmp_path / "out.bin"
    with open(src, "wb") as fh:
        fh.write(np.int32(f.n).tobytes())
        fh.write(rec.tobytes())
    exe = os.path.join(os.path.dirname(N.library_path()), "processor_example")
    subprocess.run([exe, str(src), str(dst)], check=True, timeout=120)
    buf = np.fromfile(dst, np.int32)
    seg = buf[:f.n].view(np.uint32)
    no = buf[f.n]
    clab = buf[f.n + 1:f.n + 1 + no]
    ref = oracle.process_frame(f.x, f.y, f.z, f.ring)
    assert ref["status"] == 0
    assert np.array_equal(seg, ref["seg_labels"])
    assert np.array_equal(clab, ref["cluster_labels"])
