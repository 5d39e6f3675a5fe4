"""The reference NoiseRemover's KD-tree search on the GPU (LP_NOISE_KDTREE).

The oracle's NoiseRemover restates the reference object
(lidar_processing_lib/src/noise_remover.cpp:38-68 with
include/lidar_processing_lib/kdtree.hpp:170-213, 339-397): a KD-tree built
with libstdc++'s std::nth_element and a radius search whose stack is not
cleared when it stops at the k-th neighbour.  A fresh oracle object per
frame is the defined result; its labels differ from the radius rule on a
few hundred points per frame, which the device must reproduce bit for bit.
"""
import numpy as np
import pytest

from lidar_amd import native as N
from lidar_amd import synth

from tests.test_gpu_noise_boxes import box_equal
from tests.test_gpu_parity import hull_coords

pytestmark = pytest.mark.gpu

# The search's first device run left the GPU faulted and its cause has not
# been shown, so lp_set_noise_search rejects LP_NOISE_KDTREE and these cases
# stay off the device.
FAULTED = pytest.mark.skip(reason="KD-tree noise search switched off: its first device run faulted the GPU, "
                                  "cause not yet shown")


def test_kdtree_search_is_rejected():
    with N.Context(max_frames=1, max_points=1024, max_height=16, max_width=64) as c:
        with pytest.raises(N.LidarError):
            c.set_noise_search(N.NOISE_KDTREE)
        c.set_noise_search(N.NOISE_RULE)


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(max_frames=1, max_points=200000, max_height=64, max_width=2048)
    c.set_noise_search(N.NOISE_KDTREE)
    yield c
    c.close()


def oracle_labels(oracle, x, y, z, cfg=None):
    return oracle.NoiseRemover().filter(x, y, z, cfg or oracle.noise_config())


@FAULTED
@pytest.mark.parametrize("k", range(4))
def test_c2_frames_match_the_reference_object(ctx, oracle, k):
    f = synth.workload_c2(k)
    print(f"kdtree noise c2 frame {k}: {f.n} points", flush=True)
    want = oracle_labels(oracle, f.x, f.y, f.z)
    st, got = ctx.remove_noise(f.x, f.y, f.z)
    assert st == N.OK
    rule, _ = oracle.noise_definition(f.x, f.y, f.z)
    assert np.count_nonzero(rule != want) > 100  # the frame exercises the stale stack
    assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} labels differ"


@FAULTED
def test_kitti_scan_matches_the_reference_object(ctx, oracle):
    from tests.test_kitti_fixture import kitti
    x, y, z = (np.ascontiguousarray(a, np.float32) for a in kitti())
    print(f"kdtree noise kitti: {len(x)} points", flush=True)
    want = oracle_labels(oracle, x, y, z)
    st, got = ctx.remove_noise(x, y, z)
    assert st == N.OK
    assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} labels differ"


@FAULTED
@pytest.mark.parametrize("n,q,seed", [(1, 5, 0), (2, 5, 1), (3, 5, 2), (4, 1, 3), (17, 3, 4), (100, 4, 5),
                                      (3000, 20, 6), (3000, 1000000, 7), (50000, 200, 8), (150000, 2000, 9)])
def test_tie_heavy_random_clouds(ctx, oracle, n, q, seed):
    """Coordinates on a coarse lattice, so split keys tie constantly and the
    tree depends on introselect's exact element moves."""
    rng = np.random.default_rng(seed)
    pts = (rng.integers(0, q, (n, 3)).astype(np.float32) * np.float32(0.05)) - np.float32(0.025 * q)
    x, y, z = (np.ascontiguousarray(pts[:, i]) for i in range(3))
    print(f"kdtree noise lattice n={n} q={q}", flush=True)
    for kmin in (4, 1, 7):
        cfg = N.noise_config(min_neighbours=kmin)
        want = oracle_labels(oracle, x, y, z, oracle.noise_config(min_neighbours=kmin))
        st, got = ctx.remove_noise(x, y, z, cfg=cfg)
        assert st == N.OK
        assert np.array_equal(got, want), (kmin, np.count_nonzero(got != want))


@FAULTED
def test_errors(ctx):
    x = np.array([1.0, 2.0, np.nan], np.float32)
    st, _ = ctx.remove_noise(x, x, x)
    assert st == N.ERR_INVALID_ARGUMENT
    st, _ = ctx.remove_noise(np.ones(5, np.float32), np.ones(5, np.float32), np.ones(5, np.float32),
                             cfg=N.noise_config(min_neighbours=8))
    assert st == N.ERR_INVALID_ARGUMENT
    e = np.zeros(0, np.float32)
    st, lab = ctx.remove_noise(e, e, e)
    assert st == N.OK and len(lab) == 0
    # A good frame after the failures.
    f = synth.workload_c2(5)
    st, _ = ctx.remove_noise(f.x, f.y, f.z)
    assert st == N.OK


@FAULTED
def test_fused_c2_with_the_reference_search(oracle):
    """Noise filter (reference object) -> segment -> cluster -> hull -> box,
    on two frames in one batch, against the oracle's full frame with a fresh
    NoiseRemover each."""
    frames = [synth.workload_c2(0), synth.workload_c2(3)]
    with N.Context(max_frames=2, max_points=max(f.n for f in frames)) as ctx:
        ctx.set_noise_search(N.NOISE_KDTREE)
        ctx.set_noise_filter(N.noise_config())
        ctx.set_oriented_boxes(True)
        outs = ctx.process([(f.x, f.y, f.z, f.ring) for f in frames], N.segmenter_config(), use_graph=True)
    proc = oracle.Processor()
    for f, out in zip(frames, outs):
        ref = oracle.full_frame(proc, f.x, f.y, f.z, f.ring, noise_remover=oracle.NoiseRemover())
        assert out["status"] == 0 and ref["status"] == 0
        assert np.array_equal(out["noise_labels"], ref["noise_labels"])
        assert out["num_points"] == ref["num_points"]
        assert np.array_equal(out["seg_labels"], ref["seg_labels"])
        assert np.array_equal(out["cluster_labels"], ref["cluster_labels"])
        keep = np.flatnonzero(ref["noise_labels"] == 0)
        kx, ky = f.x[keep], f.y[keep]
        obs = out["obstacle_index"]
        ox, oy = kx[obs].astype(np.float64), ky[obs].astype(np.float64)
        K = ref["num_clusters"]
        hg = hull_coords(ox, oy, out["cluster_labels"], K, out["hull_offsets"], out["hull_indices"])
        hr = hull_coords(ox, oy, ref["cluster_labels"], K, ref["hull_offsets"], ref["hull_indices"])
        assert all(np.array_equal(a, b) for a, b in zip(hg, hr))
        assert all(box_equal(out["boxes"][c], oracle_box(ref["boxes"][c])) for c in range(K))


def oracle_box(b):
    return dict(corners=np.array(b.corners[:], np.float64).reshape(4, 2), area=np.float32(b.area),
                angle_rad=np.float32(b.angle_rad), is_valid=bool(b.is_valid))
