"""ctypes binding of liblp_amd.so, the C ABI declared in include/lp_amd.h.

The library is the product: every call below runs HIP kernels on the GPU.
Loading fails loudly if the library is missing; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

OK = 0
ERR_INVALID_ARGUMENT = -1
ERR_HIP = -2
ERR_CAPACITY = -3
ERR_VOXEL_OVERFLOW = -4
ERR_RANSAC_SINGLE_CANDIDATE = -5

LABEL_UNKNOWN, LABEL_GROUND, LABEL_OBSTACLE = 0, 1, 2

# Every symbol include/lp_amd.h declares.
EXPORTED = [
    "lp_segmenter_config_default", "lp_clusterer_config_default", "lp_clusterer_config_shipped",
    "lp_capacity_default", "lp_context_create", "lp_context_destroy", "lp_last_error", "lp_segment",
    "lp_cluster", "lp_convex_hulls", "lp_process_frames", "lp_batch_stage", "lp_batch_run", "lp_batch_fetch",
    "lp_synchronize", "lp_stage_names", "lp_profile_run", "lp_context_stream", "lp_debug_fetch",
    "lp_math_eval", "lp_math_check_range", "lp_add_ring_info", "lp_query", "lp_noise_config_default",
    "lp_remove_noise", "lp_oriented_boxes", "lp_set_noise_filter", "lp_set_oriented_boxes",
    "lp_server_config_default", "lp_server_create", "lp_server_destroy", "lp_server_submit", "lp_server_poll",
    "lp_server_stats_get", "lp_cloud_from_pointcloud2", "lp_pca_boxes", "lp_set_label_clouds",
    "lp_set_vehicle_boxes", "lp_frame_image", "lp_add_ring_info_noalloc", "lp_set_noise_search", "lp_set_persistent", "lp_set_launch_shape",
]

NOISE_RULE, NOISE_KDTREE = 0, 1

MATH_EXPF, MATH_ATANF, MATH_ATAN2F, MATH_ATAN2_APPROX = range(4)


def math_eval(which, a, b=None, on_device=True):
    """Evaluate a libm restatement on the GPU (or the host copy of its source)."""
    a = np.ascontiguousarray(a, np.float32)
    b = None if b is None else np.ascontiguousarray(b, np.float32)
    out = np.zeros(len(a), np.float32)
    check(lib().lp_math_eval(which, _p(a), _p(b), _p(out), len(a), 1 if on_device else 0))
    return out


def math_check_range(which, lo, ref_bits):
    """Count device results differing from host results over float bit
    patterns [lo, lo + len(ref_bits))."""
    ref_bits = np.ascontiguousarray(ref_bits, np.uint32)
    bad = ctypes.c_uint64(0)
    check(lib().lp_math_check_range(which, lo, len(ref_bits), _p(ref_bits), ctypes.byref(bad)))
    return bad.value

DEBUG_ELEVATION_MAP, DEBUG_MAPPING, DEBUG_STATE_RECM, DEBUG_JCP_QUEUE, DEBUG_STATE_FINAL = range(5)
_DEBUG_DTYPES = {0: np.float32, 1: np.int32, 2: np.uint8, 3: np.int32, 4: np.uint8}


class SegmenterConfig(ctypes.Structure):
    """lp_segmenter_config = SegmenterConfiguration (segmenter.hpp:87-112)."""
    _fields_ = [
        ("elevation_up_deg", ctypes.c_float),
        ("elevation_down_deg", ctypes.c_float),
        ("image_width", ctypes.c_int32),
        ("image_height", ctypes.c_int32),
        ("assume_unorganized_cloud", ctypes.c_int32),
        ("grid_radial_spacing_m", ctypes.c_float),
        ("grid_slice_resolution_deg", ctypes.c_float),
        ("ground_height_threshold_m", ctypes.c_float),
        ("road_maximum_slope_m_per_m", ctypes.c_float),
        ("min_distance_m", ctypes.c_float),
        ("max_distance_m", ctypes.c_float),
        ("sensor_height_m", ctypes.c_float),
        ("kernel_threshold_distance_m", ctypes.c_float),
        ("amplification_factor", ctypes.c_float),
        ("z_min_m", ctypes.c_float),
        ("z_max_m", ctypes.c_float),
    ]


class ClustererConfig(ctypes.Structure):
    """lp_clusterer_config = ClustererConfiguration (clusterer.hpp:61-68)."""
    _fields_ = [
        ("voxel_grid_range_resolution_m", ctypes.c_float),
        ("voxel_grid_azimuth_resolution_deg", ctypes.c_float),
        ("voxel_grid_elevation_resolution_deg", ctypes.c_float),
        ("min_cluster_size", ctypes.c_uint32),
    ]


class Capacity(ctypes.Structure):
    _fields_ = [
        ("max_frames", ctypes.c_int32),
        ("max_points_per_frame", ctypes.c_int32),
        ("max_image_height", ctypes.c_int32),
        ("max_image_width", ctypes.c_int32),
        ("max_voxels", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("streams", ctypes.c_int32),
        ("hash_load_factor", ctypes.c_float),
    ]


class Cloud(ctypes.Structure):
    _fields_ = [
        ("x", ctypes.c_void_p),
        ("y", ctypes.c_void_p),
        ("z", ctypes.c_void_p),
        ("ring", ctypes.c_void_p),
        ("xyz_stride", ctypes.c_int32),
        ("ring_stride", ctypes.c_int32),
        ("num_points", ctypes.c_int32),
        ("on_device", ctypes.c_int32),
    ]


class RansacResult(ctypes.Structure):
    _fields_ = [
        ("candidates", ctypes.c_uint32),
        ("best_inliers", ctypes.c_uint32),
        ("applied", ctypes.c_int32),
        ("plane", ctypes.c_float * 4),
    ]


class NoiseConfig(ctypes.Structure):
    """lp_noise_config = NoiseRemoverConfiguration (noise_remover.hpp:41-54)."""
    _fields_ = [
        ("radius_multiplier_m_per_m", ctypes.c_float),
        ("min_search_radius_m", ctypes.c_float),
        ("min_neighbours", ctypes.c_uint32),
    ]


class BoundingBox(ctypes.Structure):
    """lp_bounding_box = BoundingBox (polygonizer.hpp:75-81)."""
    _fields_ = [
        ("corners", ctypes.c_double * 8),
        ("area", ctypes.c_float),
        ("angle_rad", ctypes.c_float),
        ("is_valid", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class FrameOutputs(ctypes.Structure):
    _fields_ = [
        ("seg_labels", ctypes.c_void_p),
        ("obstacle_index", ctypes.c_void_p),
        ("cluster_labels", ctypes.c_void_p),
        ("hull_offsets", ctypes.c_void_p),
        ("hull_indices", ctypes.c_void_p),
        ("hull_capacity", ctypes.c_int32),
        ("noise_labels", ctypes.c_void_p),
        ("boxes", ctypes.c_void_p),
        ("ground_index", ctypes.c_void_p),
        ("unknown_index", ctypes.c_void_p),
        ("vehicle_class", ctypes.c_void_p),
    ]


class FrameResult(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("num_points", ctypes.c_int32),
        ("num_obstacles", ctypes.c_int32),
        ("num_voxels", ctypes.c_int32),
        ("num_clusters", ctypes.c_int32),
        ("num_hull_vertices", ctypes.c_int32),
        ("jcp_queue", ctypes.c_int32),
        ("ransac", RansacResult),
        ("num_noise", ctypes.c_int32),
        ("num_boxes", ctypes.c_int32),
        ("num_noise_searched", ctypes.c_int32),
        ("num_ground", ctypes.c_int32),
        ("num_unknown", ctypes.c_int32),
        ("num_vehicles", ctypes.c_int32),
        ("num_hull_inexact", ctypes.c_int32),
        ("num_hull_rounded", ctypes.c_int32),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("slots", ctypes.c_int32),
        ("max_batch_frames", ctypes.c_int32),
        ("capacity", Capacity),
        ("hbm_budget_bytes", ctypes.c_uint64),
        ("cluster", ClustererConfig),
        ("noise_filter", ctypes.c_int32),
        ("noise", NoiseConfig),
        ("oriented_boxes", ctypes.c_int32),
        ("label_clouds", ctypes.c_int32),
        ("vehicle_boxes", ctypes.c_int32),
        ("noise_search", ctypes.c_int32),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("slots", ctypes.c_int32),
        ("busy_slots", ctypes.c_int32),
        ("queued", ctypes.c_int32),
        ("in_flight", ctypes.c_int32),
        ("slot_bytes", ctypes.c_uint64),
        ("batches", ctypes.c_uint64),
        ("frames", ctypes.c_uint64),
    ]


def library_path():
    return os.path.join(_HERE, "liblp_amd.so")


def lib():
    """Load liblp_amd.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError("liblp_amd.so is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        I = ctypes.c_int32
        L.lp_segmenter_config_default.argtypes = [ctypes.POINTER(SegmenterConfig)]
        L.lp_segmenter_config_default.restype = None
        for fn in ("lp_clusterer_config_default", "lp_clusterer_config_shipped"):
            getattr(L, fn).argtypes = [ctypes.POINTER(ClustererConfig)]
            getattr(L, fn).restype = None
        L.lp_capacity_default.argtypes = [ctypes.POINTER(Capacity)]
        L.lp_capacity_default.restype = None
        L.lp_context_create.argtypes = [ctypes.POINTER(Capacity), ctypes.POINTER(P)]
        L.lp_context_destroy.argtypes = [P]
        L.lp_last_error.restype = ctypes.c_char_p
        L.lp_segment.argtypes = [P, ctypes.POINTER(SegmenterConfig), ctypes.POINTER(Cloud), P,
                                 ctypes.POINTER(RansacResult)]
        L.lp_cluster.argtypes = [P, ctypes.POINTER(ClustererConfig), ctypes.POINTER(Cloud), P]
        L.lp_convex_hulls.argtypes = [P, P, P, P, I, P, P, I]
        L.lp_process_frames.argtypes = [P, ctypes.POINTER(SegmenterConfig), I, ctypes.POINTER(ClustererConfig), I,
                                        ctypes.POINTER(Cloud), ctypes.POINTER(FrameOutputs),
                                        ctypes.POINTER(FrameResult)]
        L.lp_batch_stage.argtypes = [P, ctypes.POINTER(SegmenterConfig), I, ctypes.POINTER(ClustererConfig), I,
                                     ctypes.POINTER(Cloud)]
        L.lp_batch_run.argtypes = [P, I]
        L.lp_batch_fetch.argtypes = [P, ctypes.POINTER(FrameOutputs), ctypes.POINTER(FrameResult)]
        L.lp_synchronize.argtypes = [P]
        L.lp_query.argtypes = [P]
        L.lp_noise_config_default.argtypes = [ctypes.POINTER(NoiseConfig)]
        L.lp_noise_config_default.restype = None
        L.lp_remove_noise.argtypes = [P, ctypes.POINTER(NoiseConfig), ctypes.POINTER(Cloud), P]
        L.lp_oriented_boxes.argtypes = [P, P, P, P, I, ctypes.POINTER(BoundingBox)]
        L.lp_set_noise_filter.argtypes = [P, ctypes.POINTER(NoiseConfig)]
        L.lp_set_oriented_boxes.argtypes = [P, I]
        L.lp_pca_boxes.argtypes = [P, P, P, P, I, ctypes.POINTER(BoundingBox)]
        L.lp_set_label_clouds.argtypes = [P, I]
        L.lp_set_vehicle_boxes.argtypes = [P, I]
        L.lp_set_noise_search.argtypes = [P, I]
        L.lp_set_persistent.argtypes = [P, I]
        L.lp_set_launch_shape.argtypes = [P, I, I]
        L.lp_frame_image.argtypes = [P, I, P, I]
        L.lp_server_config_default.argtypes = [ctypes.POINTER(ServerConfig)]
        L.lp_server_config_default.restype = None
        L.lp_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(P)]
        L.lp_server_destroy.argtypes = [P]
        L.lp_server_submit.argtypes = [P, ctypes.POINTER(SegmenterConfig), ctypes.POINTER(Cloud),
                                       ctypes.POINTER(FrameOutputs), ctypes.POINTER(FrameResult), ctypes.c_uint64]
        L.lp_server_poll.argtypes = [P, ctypes.POINTER(ctypes.c_uint64), I]
        L.lp_server_stats_get.argtypes = [P, ctypes.POINTER(ServerStats)]
        L.lp_stage_names.argtypes = [P, ctypes.c_char_p, I]
        L.lp_profile_run.argtypes = [P, I, P, I, ctypes.POINTER(I)]
        L.lp_context_stream.argtypes = [P]
        L.lp_context_stream.restype = P
        L.lp_debug_fetch.argtypes = [P, I, I, P, I]
        L.lp_add_ring_info.argtypes = [P, P, P, I, P]
        L.lp_math_eval.argtypes = [I, P, P, P, I, I]
        L.lp_math_check_range.argtypes = [I, ctypes.c_uint32, ctypes.c_uint32, P, ctypes.POINTER(ctypes.c_uint64)]
        _LIB = L
    return _LIB


class LidarError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lp error {code}: {msg}")
        self.code = code


def check(code):
    if code != OK:
        raise LidarError(code, lib().lp_last_error().decode())
    return code


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def segmenter_config(**kw):
    c = SegmenterConfig()
    lib().lp_segmenter_config_default(ctypes.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def clusterer_config(shipped=True, **kw):
    """shipped=True: processor.param.yaml values (3.0 deg elevation voxels)."""
    c = ClustererConfig()
    (lib().lp_clusterer_config_shipped if shipped else lib().lp_clusterer_config_default)(ctypes.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def noise_config(**kw):
    c = NoiseConfig()
    lib().lp_noise_config_default(ctypes.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def box_dict(b):
    return dict(corners=np.array(b.corners[:], np.float64).reshape(4, 2), area=np.float32(b.area),
                angle_rad=np.float32(b.angle_rad), is_valid=bool(b.is_valid))


class PointField(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("offset", ctypes.c_uint32)]


def pointcloud2_cloud(data, fields, width, height, point_step, row_step):
    """Processor::run's PointCloud2 field dispatch as a zero-copy view:
    data is the message's byte buffer (numpy uint8), fields [(name, offset)].
    Returns (status, Cloud); the Cloud points into `data`."""
    arr = (PointField * max(len(fields), 1))()
    names = [n.encode() for n, _ in fields]
    for k, (n, off) in enumerate(fields):
        arr[k].name = names[k]
        arr[k].offset = off
    c = Cloud()
    L = lib()
    L.lp_cloud_from_pointcloud2.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                            ctypes.c_uint32, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]
    L.lp_cloud_from_pointcloud2.restype = ctypes.c_int32
    st = L.lp_cloud_from_pointcloud2(data.ctypes.data if data is not None and len(data) else None, width, height,
                                     point_step, row_step, ctypes.addressof(arr), len(fields), ctypes.byref(c))
    return st, c


def soa_cloud(x, y, z, ring=None):
    c = Cloud()
    c.x, c.y, c.z = (a.ctypes.data for a in (x, y, z))
    c.ring = ring.ctypes.data if ring is not None else None
    c.xyz_stride = 4
    c.ring_stride = 2
    c.num_points = len(x)
    c.on_device = 0
    return c


def aos_cloud(records, has_ring=True):
    """records: structured numpy array with fields x, y, z (f4) and ring (u2)."""
    base = records.ctypes.data
    c = Cloud()
    c.x = base + records.dtype.fields["x"][1]
    c.y = base + records.dtype.fields["y"][1]
    c.z = base + records.dtype.fields["z"][1]
    c.ring = base + records.dtype.fields["ring"][1] if has_ring else None
    c.xyz_stride = records.dtype.itemsize
    c.ring_stride = records.dtype.itemsize
    c.num_points = len(records)
    c.on_device = 0
    return c


class Context:
    """One device context: a HIP stream plus preallocated buffers."""

    def __init__(self, max_frames=1, max_points=200000, max_height=64, max_width=2048, max_voxels=200000, device=0,
                 streams=1, hash_load_factor=0.5):
        cap = Capacity()
        lib().lp_capacity_default(ctypes.byref(cap))
        cap.streams = streams
        cap.hash_load_factor = hash_load_factor
        cap.max_frames = max_frames
        cap.max_points_per_frame = max_points
        cap.max_image_height = max_height
        cap.max_image_width = max_width
        cap.max_voxels = max_voxels
        cap.device = device
        h = ctypes.c_void_p()
        check(lib().lp_context_create(ctypes.byref(cap), ctypes.byref(h)))
        self.h = h
        self.cap = cap
        self._boxes_on = False
        self._clouds_on = False
        self._vehicles_on = False

    def close(self):
        if getattr(self, "h", None):
            lib().lp_context_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- Segmenter / Clusterer / Polygonizer -------------------------------
    def segment(self, x, y, z, ring=None, cfg=None, cloud=None):
        """Segmenter::segment -> (status, labels u32, RansacResult)."""
        cfg = cfg or segmenter_config()
        n = len(x) if cloud is None else cloud.num_points
        labels = np.zeros(n, np.uint32)
        rr = RansacResult()
        if cloud is None:
            x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
            ring = None if ring is None else np.ascontiguousarray(ring, np.uint16)
            cloud = soa_cloud(x, y, z, ring)
        st = lib().lp_segment(self.h, ctypes.byref(cfg), ctypes.byref(cloud), _p(labels), ctypes.byref(rr))
        return st, labels, rr

    def cluster(self, x, y, z, cfg=None):
        """Clusterer::cluster -> (status, labels i32)."""
        cfg = cfg or clusterer_config()
        x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
        labels = np.zeros(len(x), np.int32)
        cloud = soa_cloud(x, y, z)
        st = lib().lp_cluster(self.h, ctypes.byref(cfg), ctypes.byref(cloud), _p(labels))
        return st, labels

    def add_ring_info(self, x, y):
        """Dataloader::addRingInfo -> (status, ring u16)."""
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        ring = np.zeros(len(x), np.uint16)
        st = lib().lp_add_ring_info(self.h, _p(x), _p(y), len(x), _p(ring))
        return st, ring

    def convex_hulls(self, lists):
        """Polygonizer::convexHull on each (x, y) double list -> (status, [indices])."""
        offs = np.zeros(len(lists) + 1, np.int32)
        for k, (lx, _) in enumerate(lists):
            offs[k + 1] = offs[k] + len(lx)
        xs = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float64) for a, _ in lists]) if lists else
                                  np.zeros(0), np.float64)
        ys = np.ascontiguousarray(np.concatenate([np.asarray(b, np.float64) for _, b in lists]) if lists else
                                  np.zeros(0), np.float64)
        hoff = np.zeros(len(lists) + 1, np.int32)
        cap = max(int(offs[-1]) + 4, 4)
        hidx = np.zeros(cap, np.int32)
        st = lib().lp_convex_hulls(self.h, _p(xs), _p(ys), _p(offs), len(lists), _p(hoff), _p(hidx), cap)
        if st != OK:
            return st, None
        return st, [hidx[hoff[k]:hoff[k + 1]].copy() for k in range(len(lists))]

    def remove_noise(self, x, y, z, cfg=None):
        """NoiseRemover::filter -> (status, labels u8: 0 VALID, 1 NOISE)."""
        cfg = cfg or noise_config()
        x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
        labels = np.zeros(len(x), np.uint8)
        cloud = soa_cloud(x, y, z)
        st = lib().lp_remove_noise(self.h, ctypes.byref(cfg), ctypes.byref(cloud), _p(labels))
        return st, labels

    def oriented_boxes(self, polygons):
        """boundingBoxRotatingCalipers on each (x, y) CCW polygon -> (status, [box dict])."""
        return self._polygon_boxes(lib().lp_oriented_boxes, polygons)

    def pca_boxes(self, polygons):
        """boundingBoxPrincipalComponentAnalysis on each (x, y) list -> (status, [box dict])."""
        return self._polygon_boxes(lib().lp_pca_boxes, polygons)

    def _polygon_boxes(self, fn, polygons):
        offs = np.zeros(len(polygons) + 1, np.int32)
        for k, (px, _) in enumerate(polygons):
            offs[k + 1] = offs[k] + len(px)
        xs = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float64) for a, _ in polygons]) if polygons
                                  else np.zeros(0), np.float64)
        ys = np.ascontiguousarray(np.concatenate([np.asarray(b, np.float64) for _, b in polygons]) if polygons
                                  else np.zeros(0), np.float64)
        boxes = (BoundingBox * max(len(polygons), 1))()
        st = fn(self.h, _p(xs), _p(ys), _p(offs), len(polygons), boxes)
        if st != OK:
            return st, None
        return st, [box_dict(boxes[k]) for k in range(len(polygons))]

    def set_noise_filter(self, cfg):
        """Noise filter ahead of the segmenter in the fused path (None: off)."""
        check(lib().lp_set_noise_filter(self.h, None if cfg is None else ctypes.byref(cfg)))
        self._noise_cfg = cfg

    def set_noise_search(self, mode):
        """NOISE_RULE (the radius rule) or NOISE_KDTREE (the reference object's search)."""
        check(lib().lp_set_noise_search(self.h, int(mode)))

    def set_persistent(self, enable):
        """Per-point kernels as persistent (frame, tile) work-queue launches."""
        check(lib().lp_set_persistent(self.h, 1 if enable else 0))

    def set_launch_shape(self, point_workgroup=256, mark_tile=256):
        """Per-point workgroup size and JCP marking tile width (64, 128 or 256)."""
        check(lib().lp_set_launch_shape(self.h, int(point_workgroup), int(mark_tile)))

    def set_oriented_boxes(self, enable):
        check(lib().lp_set_oriented_boxes(self.h, 1 if enable else 0))
        self._boxes_on = bool(enable)

    def set_label_clouds(self, enable):
        """Ground and unknown index lists beside the obstacle split."""
        check(lib().lp_set_label_clouds(self.h, 1 if enable else 0))
        self._clouds_on = bool(enable)

    def set_vehicle_boxes(self, enable):
        """Vehicle class per cluster (boxes are computed for it)."""
        check(lib().lp_set_vehicle_boxes(self.h, 1 if enable else 0))
        self._vehicles_on = bool(enable)

    def image(self, frame=0):
        """Segmenter::image of a frame of the last run: uint8 [H, W, 3] BGR."""
        need = lib().lp_frame_image(self.h, frame, ctypes.c_void_p(1), 0)
        check(min(need, 0))
        out = np.zeros(need, np.uint8)
        got = lib().lp_frame_image(self.h, frame, _p(out), need)
        check(min(got, 0))
        return out

    # -- fused frames --------------------------------------------------------
    def _clouds(self, frames):
        arr = (Cloud * len(frames))()
        keep = []
        for i, f in enumerate(frames):
            x, y, z = (np.ascontiguousarray(a, np.float32) for a in (f[0], f[1], f[2]))
            ring = None if f[3] is None else np.ascontiguousarray(f[3], np.uint16)
            keep.append((x, y, z, ring))
            arr[i] = soa_cloud(x, y, z, ring)
        return arr, keep

    def stage(self, frames, seg_cfgs, ccfg=None):
        """frames: list of (x, y, z, ring); seg_cfgs: one config or one per frame."""
        ccfg = ccfg or clusterer_config()
        seg_cfgs = seg_cfgs if isinstance(seg_cfgs, (list, tuple)) else [seg_cfgs]
        carr = (SegmenterConfig * len(seg_cfgs))(*seg_cfgs)
        clouds, keep = self._clouds(frames)
        self._staged = (frames, keep, carr, ccfg, clouds)
        check(lib().lp_batch_stage(self.h, carr, len(seg_cfgs), ctypes.byref(ccfg), len(frames), clouds))

    def run(self, use_graph=True):
        check(lib().lp_batch_run(self.h, 1 if use_graph else 0))

    def synchronize(self):
        check(lib().lp_synchronize(self.h))

    def query(self):
        """True once all queued work of this context has finished."""
        r = lib().lp_query(self.h)
        check(min(r, 0))
        return r == 1

    def fetch(self):
        frames = self._staged[0]
        B = len(frames)
        outs = (FrameOutputs * B)()
        res = (FrameResult * B)()
        bufs = []
        for i, f in enumerate(frames):
            n = len(f[0])
            b = dict(seg=np.zeros(n, np.uint32), obs=np.zeros(max(n, 1), np.int32),
                     clab=np.zeros(max(n, 1), np.int32), hoff=np.zeros(n + 2, np.int32),
                     hidx=np.zeros(2 * n + 16, np.int32), noise=np.zeros(max(n, 1), np.uint8),
                     boxes=(BoundingBox * (max(n, 1) if self._boxes_on or self._vehicles_on else 1))(),
                     gidx=np.zeros(max(n, 1), np.int32), uidx=np.zeros(max(n, 1), np.int32),
                     vcls=np.zeros(max(n, 1), np.int32))
            bufs.append(b)
            outs[i].ground_index = b["gidx"].ctypes.data
            outs[i].unknown_index = b["uidx"].ctypes.data
            outs[i].vehicle_class = b["vcls"].ctypes.data
            outs[i].noise_labels = b["noise"].ctypes.data
            outs[i].boxes = ctypes.addressof(b["boxes"])
            outs[i].seg_labels = b["seg"].ctypes.data
            outs[i].obstacle_index = b["obs"].ctypes.data
            outs[i].cluster_labels = b["clab"].ctypes.data
            outs[i].hull_offsets = b["hoff"].ctypes.data
            outs[i].hull_indices = b["hidx"].ctypes.data
            outs[i].hull_capacity = len(b["hidx"])
        lib().lp_batch_fetch(self.h, outs, res)
        out = []
        for i in range(B):
            r = res[i]
            b = bufs[i]
            no, K, V = r.num_obstacles, r.num_clusters, r.num_hull_vertices
            out.append(dict(status=r.status, seg_labels=b["seg"][:r.num_points], obstacle_index=b["obs"][:no],
                            num_points=r.num_points, num_noise=r.num_noise, noise_labels=b["noise"][:len(frames[i][0])],
                            num_noise_searched=r.num_noise_searched,
                            boxes=[box_dict(b["boxes"][k]) for k in range(K)]
                            if self._boxes_on or self._vehicles_on else [],
                            ground_index=b["gidx"][:r.num_ground], unknown_index=b["uidx"][:r.num_unknown],
                            num_ground=r.num_ground, num_unknown=r.num_unknown, num_vehicles=r.num_vehicles,
                            num_hull_inexact=r.num_hull_inexact, num_hull_rounded=r.num_hull_rounded,
                            vehicle_class=b["vcls"][:K] if self._vehicles_on else None,
                            cluster_labels=b["clab"][:no], num_clusters=K, hull_offsets=b["hoff"][:K + 1],
                            hull_indices=b["hidx"][:V], num_voxels=r.num_voxels, jcp_queue=r.jcp_queue,
                            ransac=dict(candidates=r.ransac.candidates, best_inliers=r.ransac.best_inliers,
                                        applied=r.ransac.applied, plane=np.array(list(r.ransac.plane),
                                                                                 np.float32))))
        return out

    def process(self, frames, seg_cfgs, ccfg=None, use_graph=False):
        self.stage(frames, seg_cfgs, ccfg)
        self.run(use_graph)
        return self.fetch()

    def debug(self, frame, which):
        """Intermediate device state of the last run (see lp_debug_fetch)."""
        need = lib().lp_debug_fetch(self.h, frame, which, ctypes.c_void_p(1), 0)
        if need < 0:
            check(need)
        out = np.zeros(max(need, 1), _DEBUG_DTYPES[which])
        got = lib().lp_debug_fetch(self.h, frame, which, _p(out), len(out))
        if got < 0:
            check(got)
        return out[:got]

    def profile(self, runs=5):
        """Average per-stage milliseconds of the staged batch (HIP events)."""
        ms = np.zeros(64, np.float32)
        ns = ctypes.c_int32(0)
        check(lib().lp_profile_run(self.h, runs, _p(ms), 64, ctypes.byref(ns)))
        buf = ctypes.create_string_buffer(4096)
        lib().lp_stage_names(self.h, buf, 4096)
        names = buf.value.decode().split(",")
        return list(zip(names, ms[: ns.value].tolist()))


class FrameSlot:
    """Host buffers of one in-flight frame (inputs kept alive, outputs written
    by the server)."""

    def __init__(self, max_points, boxes=False):
        n = max(max_points, 1)
        self.seg = np.zeros(n, np.uint32)
        self.obs = np.zeros(n, np.int32)
        self.clab = np.zeros(n, np.int32)
        self.hoff = np.zeros(n + 2, np.int32)
        self.hidx = np.zeros(2 * n + 16, np.int32)
        self.noise = np.zeros(n, np.uint8)
        self.boxes = (BoundingBox * (n if boxes else 1))()
        self.gidx = np.zeros(n, np.int32)
        self.uidx = np.zeros(n, np.int32)
        self.vcls = np.zeros(n, np.int32)
        self.out = FrameOutputs()
        self.out.seg_labels = self.seg.ctypes.data
        self.out.obstacle_index = self.obs.ctypes.data
        self.out.cluster_labels = self.clab.ctypes.data
        self.out.hull_offsets = self.hoff.ctypes.data
        self.out.hull_indices = self.hidx.ctypes.data
        self.out.hull_capacity = len(self.hidx)
        self.out.noise_labels = self.noise.ctypes.data
        self.out.boxes = ctypes.addressof(self.boxes)
        self.out.ground_index = self.gidx.ctypes.data
        self.out.unknown_index = self.uidx.ctypes.data
        self.out.vehicle_class = self.vcls.ctypes.data
        self.res = FrameResult()
        self.cloud = None
        self.cfg = None
        self.inputs = None

    def result(self, boxes=False, vehicles=False):
        r = self.res
        no, K, V = r.num_obstacles, r.num_clusters, r.num_hull_vertices
        return dict(status=r.status, num_points=r.num_points, num_noise=r.num_noise,
                    seg_labels=self.seg[:r.num_points].copy(), obstacle_index=self.obs[:no].copy(),
                    cluster_labels=self.clab[:no].copy(), num_clusters=K, hull_offsets=self.hoff[:K + 1].copy(),
                    hull_indices=self.hidx[:V].copy(), boxes=[box_dict(self.boxes[k]) for k in range(K)] if boxes
                    else [], num_voxels=r.num_voxels, ground_index=self.gidx[:r.num_ground].copy(),
                    unknown_index=self.uidx[:r.num_unknown].copy(), num_vehicles=r.num_vehicles,
                    vehicle_class=self.vcls[:K].copy() if vehicles else None)


class Server:
    """lp_server: continuous batching over several contexts."""

    def __init__(self, slots=3, max_batch_frames=8, max_points=262144, max_height=128, max_width=2048,
                 hbm_budget_bytes=0, noise_cfg=None, boxes=False, ccfg=None, label_clouds=False, vehicles=False,
                 noise_search=0):
        cfg = ServerConfig()
        lib().lp_server_config_default(ctypes.byref(cfg))
        cfg.slots = slots
        cfg.max_batch_frames = max_batch_frames
        cfg.capacity.max_points_per_frame = max_points
        cfg.capacity.max_image_height = max_height
        cfg.capacity.max_image_width = max_width
        cfg.hbm_budget_bytes = hbm_budget_bytes
        if ccfg is not None:
            cfg.cluster = ccfg
        if noise_cfg is not None:
            cfg.noise_filter = 1
            cfg.noise = noise_cfg
        cfg.oriented_boxes = 1 if boxes else 0
        cfg.label_clouds = 1 if label_clouds else 0
        cfg.vehicle_boxes = 1 if vehicles else 0
        cfg.noise_search = int(noise_search)
        h = ctypes.c_void_p()
        check(lib().lp_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.boxes = boxes or vehicles
        self.vehicles = vehicles
        self._done = (ctypes.c_uint64 * 1024)()

    def close(self):
        if getattr(self, "h", None):
            lib().lp_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def submit(self, slot, x, y, z, ring, seg_cfg, tag):
        """Queue one frame; `slot` (a FrameSlot) keeps the inputs alive and
        receives the outputs."""
        slot.inputs = (x, y, z, ring)
        slot.cloud = soa_cloud(x, y, z, ring)
        slot.cfg = seg_cfg
        check(lib().lp_server_submit(self.h, ctypes.byref(seg_cfg), ctypes.byref(slot.cloud),
                                     ctypes.byref(slot.out), ctypes.byref(slot.res), tag))

    def poll(self):
        n = lib().lp_server_poll(self.h, self._done, len(self._done))
        if n < 0:
            check(n)
        return [self._done[k] for k in range(n)]

    def stats(self):
        st = ServerStats()
        check(lib().lp_server_stats_get(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}
