"""Ring-less and unorganized clouds: the HIP path vs the oracle (GPU).

`Segmenter::segment` is instantiated for point types without a ring field
(`PointXYZ`, `PointXYZI`: lidar_processing_lib/src/segmenter.cpp:672-679) and,
for organized types, honours `assume_unorganized_cloud`; both take the
image row from the elevation angle (segmenter.cpp:155-193).  The C ABI
selects the first with `ring == NULL` and the second with the flag.  Every
intermediate state, the labels, the clusters and the hull coordinate
sequences must equal the oracle's.
"""
import numpy as np
import pytest

from lidar_amd import native as N
from lidar_amd import synth

from tests.test_gpu_parity import hull_coords

pytestmark = pytest.mark.gpu

OBSTACLE = 2


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(max_frames=2, max_points=262144, max_height=128, max_width=2048)
    yield c
    c.close()


_FRAMES = {}


def frame(name):
    if name not in _FRAMES:
        _FRAMES[name] = {"c1": lambda: synth.workload_c1(2), "c2": lambda: synth.workload_c2(1),
                         "c4": lambda: synth.workload_c4(1)}[name]()
    return _FRAMES[name]


def configs(oracle, f, unorganized):
    kw = dict(image_height=f.height, image_width=f.width, elevation_up_deg=f.fov_up_deg,
              elevation_down_deg=f.fov_down_deg, assume_unorganized_cloud=1 if unorganized else 0)
    return oracle.seg_config(**kw), N.segmenter_config(**kw)


MODES = [("c1", "ringless"), ("c2", "ringless"), ("c4", "ringless"), ("c1", "flag"), ("c4", "flag")]


@pytest.mark.parametrize("name,mode", MODES)
def test_unorganized_segment_stages(ctx, oracle, name, mode):
    f = frame(name)
    print(f"unorganized segment {name} {mode}: {f.n} points", flush=True)
    flag = mode == "flag"
    ring = f.ring if flag else None
    sco, scg = configs(oracle, f, flag)
    st, lab_o, dbg = oracle.segment(f.x, f.y, f.z, ring, cfg=sco, debug=True)
    stg, lab_g, rr = ctx.segment(f.x, f.y, f.z, ring, cfg=scg)
    assert st == 0 and stg == 0
    assert np.array_equal(ctx.debug(0, N.DEBUG_MAPPING), dbg["mapping"])
    assert np.array_equal(ctx.debug(0, N.DEBUG_ELEVATION_MAP).view(np.uint32), dbg["elevation_map"].view(np.uint32))
    assert np.array_equal(ctx.debug(0, N.DEBUG_STATE_RECM), dbg["state_recm"])
    assert np.array_equal(ctx.debug(0, N.DEBUG_JCP_QUEUE), dbg["queue"])
    assert np.array_equal(ctx.debug(0, N.DEBUG_STATE_FINAL), dbg["state_final"])
    assert np.array_equal(lab_g, lab_o)
    assert rr.candidates == dbg["ransac_candidates"] and rr.best_inliers == dbg["ransac_best_inliers"]
    assert np.array_equal(np.array(list(rr.plane), np.float32).view(np.uint32), dbg["ransac_plane"].view(np.uint32))


@pytest.mark.parametrize("name,mode", MODES)
def test_unorganized_full_frame(ctx, oracle, name, mode):
    f = frame(name)
    print(f"unorganized frame {name} {mode}: {f.n} points", flush=True)
    flag = mode == "flag"
    ring = f.ring if flag else None
    sco, scg = configs(oracle, f, flag)
    ref = oracle.process_frame(f.x, f.y, f.z, ring, scfg=sco)
    out = ctx.process([(f.x, f.y, f.z, ring)], scg)[0]
    assert ref["status"] == 0 and out["status"] == 0
    assert np.array_equal(out["seg_labels"], ref["seg_labels"])
    assert np.array_equal(out["obstacle_index"], np.flatnonzero(ref["seg_labels"] == OBSTACLE))
    assert out["num_clusters"] == ref["num_clusters"] and ref["num_clusters"] > 10
    assert np.array_equal(out["cluster_labels"], ref["cluster_labels"])
    obs = out["obstacle_index"]
    ox, oy = f.x[obs].astype(np.float64), f.y[obs].astype(np.float64)
    K = ref["num_clusters"]
    hg = hull_coords(ox, oy, out["cluster_labels"], K, out["hull_offsets"], out["hull_indices"])
    hr = hull_coords(ox, oy, ref["cluster_labels"], K, ref["hull_offsets"], ref["hull_indices"])
    bad = [c for c in range(K) if not np.array_equal(hg[c], hr[c])]
    assert not bad, f"{len(bad)} hulls differ, first {bad[:5]}"


def test_ringless_aos_records(ctx, oracle):
    """A PointXYZI-like AoS record (x, y, z, intensity; no ring) through the
    strided cloud view."""
    f = frame("c1")
    rec = np.zeros(f.n, dtype=[("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("intensity", "<f4")])
    rec["x"], rec["y"], rec["z"] = f.x, f.y, f.z
    sco, scg = configs(oracle, f, False)
    st, lab_o, _ = oracle.segment(f.x, f.y, f.z, None, cfg=sco)
    stg, lab_g, _ = ctx.segment(None, None, None, cfg=scg, cloud=N.aos_cloud(rec, has_ring=False))
    assert st == 0 and stg == 0
    assert np.array_equal(lab_g, lab_o)
