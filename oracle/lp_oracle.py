"""ctypes binding of the CPU oracle (oracle/liblp_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker.  The product path
(lidar_amd / liblp_amd.so) never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

OK = 0
ERR_INVALID = -1
ERR_VOXEL_OVERFLOW = -4
ERR_RANSAC_SINGLE = -5


class SegConfig(ctypes.Structure):
    """Mirror of SegmenterConfiguration (segmenter.hpp:87-112)."""
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


class ClusterConfig(ctypes.Structure):
    """Mirror of ClustererConfiguration (clusterer.hpp:61-68)."""
    _fields_ = [
        ("voxel_grid_range_resolution_m", ctypes.c_float),
        ("voxel_grid_azimuth_resolution_deg", ctypes.c_float),
        ("voxel_grid_elevation_resolution_deg", ctypes.c_float),
        ("min_cluster_size", ctypes.c_uint32),
    ]


class SegDebug(ctypes.Structure):
    _fields_ = [
        ("elevation_map", ctypes.c_void_p),
        ("mapping", ctypes.c_void_p),
        ("state_recm", ctypes.c_void_p),
        ("state_final", ctypes.c_void_p),
        ("queue", ctypes.c_void_p),
        ("queue_len", ctypes.c_int32),
        ("ransac_candidates", ctypes.c_uint32),
        ("ransac_best_inliers", ctypes.c_uint32),
        ("ransac_applied", ctypes.c_int32),
        ("ransac_plane", ctypes.c_float * 4),
        ("ransac_draws", ctypes.c_uint32 * 180),
        ("image", ctypes.c_void_p),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liblp_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        L.lpo_seg_config_default.argtypes = [ctypes.POINTER(SegConfig)]
        L.lpo_cluster_config_default.argtypes = [ctypes.POINTER(ClusterConfig)]
        L.lpo_segment.argtypes = [ctypes.POINTER(SegConfig), P, P, P, P, ctypes.c_int32, P, ctypes.POINTER(SegDebug)]
        L.lpo_cluster.argtypes = [ctypes.POINTER(ClusterConfig), P, P, P, ctypes.c_int32, ctypes.c_int32, P]
        L.lpo_convex_hull.argtypes = [P, P, ctypes.c_int32, P]
        L.lpo_convex_hull.restype = ctypes.c_int32
        L.lpo_process_frame.argtypes = [ctypes.POINTER(SegConfig), ctypes.POINTER(ClusterConfig), P, P, P, P,
                                        ctypes.c_int32, ctypes.c_int32, P, P, P, P, P, ctypes.c_int32, P,
                                        ctypes.c_int32]
        L.lpo_add_ring_info.argtypes = [P, P, ctypes.c_int32, P]
        L.lpo_atan2_approx.argtypes = [ctypes.c_float, ctypes.c_float]
        L.lpo_atan2_approx.restype = ctypes.c_float
        L.lpo_std_uniform_draws.argtypes = [ctypes.c_uint32, ctypes.c_int32, P]
        L.lpo_mt19937_raw.argtypes = [ctypes.c_int32, P]
        for fn in ("lpo_host_expf", "lpo_host_atanf"):
            getattr(L, fn).argtypes = [ctypes.c_float]
            getattr(L, fn).restype = ctypes.c_float
        L.lpo_host_atan2f.argtypes = [ctypes.c_float, ctypes.c_float]
        L.lpo_host_atan2f.restype = ctypes.c_float
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def seg_config(**kw):
    c = SegConfig()
    lib().lpo_seg_config_default(ctypes.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def cluster_config(**kw):
    c = ClusterConfig()
    lib().lpo_cluster_config_default(ctypes.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def segment(x, y, z, ring=None, cfg=None, debug=False):
    """Segmenter::segment.  Returns (status, labels[u32], debug dict or None)."""
    cfg = cfg or seg_config()
    n = len(x)
    x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
    ring = None if ring is None else np.ascontiguousarray(ring, np.uint16)
    labels = np.zeros(n, np.uint32)
    dbg = None
    out = None
    if debug:
        H, W = cfg.image_height, cfg.image_width
        rings = int(np.float32(cfg.max_distance_m) / np.float32(cfg.grid_radial_spacing_m))
        slices = 360 if cfg.grid_slice_resolution_deg == 1.0 else int(
            np.float32(2 * np.pi) / (np.float32(cfg.grid_slice_resolution_deg) * np.float32(np.pi / 180)))
        out = dict(elevation_map=np.zeros(rings * slices, np.float32), mapping=np.zeros(H * W, np.int32),
                   state_recm=np.zeros(H * W, np.uint8), state_final=np.zeros(H * W, np.uint8),
                   queue=np.zeros(H * W, np.int32), image=np.zeros(H * W * 3, np.uint8))
        dbg = SegDebug()
        for k in ("elevation_map", "mapping", "state_recm", "state_final", "queue", "image"):
            setattr(dbg, k, out[k].ctypes.data)
    st = lib().lpo_segment(ctypes.byref(cfg), _p(x), _p(y), _p(z), _p(ring), n, _p(labels),
                           ctypes.byref(dbg) if dbg is not None else None)
    if debug:
        out["queue"] = out["queue"][: dbg.queue_len].copy()
        out["ransac_candidates"] = dbg.ransac_candidates
        out["ransac_best_inliers"] = dbg.ransac_best_inliers
        out["ransac_applied"] = dbg.ransac_applied
        out["ransac_plane"] = np.array(list(dbg.ransac_plane), np.float32)
        out["ransac_draws"] = np.array(list(dbg.ransac_draws), np.uint32)[:120]
    return st, labels, out


def cluster(x, y, z, cfg=None, max_voxels=200000):
    cfg = cfg or cluster_config()
    n = len(x)
    x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
    labels = np.zeros(n, np.int32)
    st = lib().lpo_cluster(ctypes.byref(cfg), _p(x), _p(y), _p(z), n, max_voxels, _p(labels))
    return st, labels


def convex_hull(x, y):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    out = np.zeros(2 * len(x) + 1, np.int32)
    k = lib().lpo_convex_hull(_p(x), _p(y), len(x), _p(out))
    return out[:k].copy()


def process_frame(x, y, z, ring=None, scfg=None, ccfg=None, max_voxels=200000):
    """Processor glue: returns dict(status, seg_labels, obstacle_count,
    cluster_labels, num_clusters, hull_offsets, hull_indices)."""
    scfg = scfg or seg_config()
    ccfg = ccfg or cluster_config()
    n = len(x)
    x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
    ring = None if ring is None else np.ascontiguousarray(ring, np.uint16)
    seg = np.zeros(n, np.uint32)
    clab = np.zeros(max(n, 1), np.int32)
    nobs = ctypes.c_int32(0)
    ncl = ctypes.c_int32(0)
    hoff = np.zeros(n + 2, np.int32)
    hidx = np.zeros(2 * n + 16, np.int32)
    st = lib().lpo_process_frame(ctypes.byref(scfg), ctypes.byref(ccfg), _p(x), _p(y), _p(z), _p(ring), n,
                                 max_voxels, _p(seg), ctypes.byref(nobs), _p(clab), ctypes.byref(ncl), _p(hoff),
                                 len(hoff), _p(hidx), len(hidx))
    k = ncl.value
    return dict(status=st, seg_labels=seg, obstacle_count=nobs.value, cluster_labels=clab[: nobs.value].copy(),
                num_clusters=k, hull_offsets=hoff[: k + 1].copy(), hull_indices=hidx[: hoff[k]].copy())


def add_ring_info(x, y):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    ring = np.zeros(len(x), np.uint16)
    lib().lpo_add_ring_info(_p(x), _p(y), len(x), _p(ring))
    return ring


def std_uniform_draws(n, count):
    out = np.zeros(count, np.uint32)
    lib().lpo_std_uniform_draws(n, count, _p(out))
    return out


def host_math_range(which, lo, count):
    """Host C library results (bit patterns) over float bits [lo, lo+count):
    which 0 = expf, 1 = atanf."""
    L = lib()
    if not getattr(L, "_range_bound", False):
        L.lpo_host_math_range.argtypes = [ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
        L._range_bound = True
    out = np.zeros(count, np.uint32)
    L.lpo_host_math_range(which, lo, count, _p(out))
    return out


def mt19937_raw(count):
    out = np.zeros(count, np.uint32)
    lib().lpo_mt19937_raw(count, _p(out))
    return out


class Processor:
    """Persistent oracle processor: one Segmenter, Clusterer and Polygonizer
    constructed once and reused per frame, as the reference's processor node
    holds them (processor.hpp:248-258)."""

    def __init__(self, scfg=None, ccfg=None, max_voxels=200000):
        L = lib()
        if not getattr(L, "_proc_bound", False):
            P = ctypes.c_void_p
            L.lpo_processor_create.argtypes = [ctypes.POINTER(SegConfig), ctypes.POINTER(ClusterConfig),
                                               ctypes.c_int32]
            L.lpo_processor_create.restype = ctypes.c_void_p
            L.lpo_processor_destroy.argtypes = [P]
            L.lpo_processor_segment.argtypes = [P, P, P, P, P, ctypes.c_int32, P, P]
            L.lpo_processor_cluster.argtypes = [P, P, P, P, ctypes.c_int32, P]
            L.lpo_processor_frame.argtypes = [P, P, P, P, P, ctypes.c_int32, P, P, P, P, P, ctypes.c_int32, P,
                                              ctypes.c_int32]
            L._proc_bound = True
        self.scfg = scfg or seg_config()
        self.ccfg = ccfg or cluster_config()
        self.h = L.lpo_processor_create(ctypes.byref(self.scfg), ctypes.byref(self.ccfg), max_voxels)
        if not self.h:
            raise ValueError("invalid oracle configuration")

    def __del__(self):
        if getattr(self, "h", None):
            lib().lpo_processor_destroy(self.h)
            self.h = None

    def segment(self, x, y, z, ring=None, labels=None):
        n = len(x)
        labels = np.zeros(n, np.uint32) if labels is None else labels
        st = lib().lpo_processor_segment(self.h, _p(x), _p(y), _p(z), _p(ring), n, _p(labels), None)
        return st, labels

    def frame(self, x, y, z, ring=None, bufs=None):
        """Full frame (segment -> split -> cluster -> gather -> hull).  `bufs`
        lets a timing loop reuse output arrays."""
        n = len(x)
        if bufs is None:
            bufs = dict(seg=np.zeros(n, np.uint32), clab=np.zeros(max(n, 1), np.int32),
                        hoff=np.zeros(n + 2, np.int32), hidx=np.zeros(2 * n + 16, np.int32))
        nobs = ctypes.c_int32(0)
        ncl = ctypes.c_int32(0)
        st = lib().lpo_processor_frame(self.h, _p(x), _p(y), _p(z), _p(ring), n, _p(bufs["seg"]),
                                       ctypes.byref(nobs), _p(bufs["clab"]), ctypes.byref(ncl), _p(bufs["hoff"]),
                                       len(bufs["hoff"]), _p(bufs["hidx"]), len(bufs["hidx"]))
        k = ncl.value
        return dict(status=st, seg_labels=bufs["seg"], obstacle_count=nobs.value,
                    cluster_labels=bufs["clab"][: nobs.value], num_clusters=k,
                    hull_offsets=bufs["hoff"][: k + 1], hull_indices=bufs["hidx"][: bufs["hoff"][k]])


# ----------------------------------------------------------- DROR and OBB
class NoiseConfig(ctypes.Structure):
    """NoiseRemoverConfiguration (noise_remover.hpp:41-54)."""
    _fields_ = [("radius_multiplier_m_per_m", ctypes.c_float), ("min_search_radius_m", ctypes.c_float),
                ("min_neighbours", ctypes.c_uint32)]


class BoundingBox(ctypes.Structure):
    """BoundingBox (polygonizer.hpp:75-81), corners x0,y0,..,x3,y3."""
    _fields_ = [("corners", ctypes.c_double * 8), ("area", ctypes.c_float), ("angle_rad", ctypes.c_float),
                ("is_valid", ctypes.c_int32)]


def _bind_ext():
    L = lib()
    if getattr(L, "_ext_bound", False):
        return L
    P = ctypes.c_void_p
    L.lpo_noise_config_default.argtypes = [ctypes.POINTER(NoiseConfig)]
    L.lpo_noise_radius_sqr.argtypes = [ctypes.POINTER(NoiseConfig), ctypes.c_float, ctypes.c_float]
    L.lpo_noise_radius_sqr.restype = ctypes.c_float
    L.lpo_noise_definition.argtypes = [ctypes.POINTER(NoiseConfig), P, P, P, ctypes.c_int32, P, P]
    L.lpo_noise_remover_create.restype = P
    L.lpo_noise_remover_destroy.argtypes = [P]
    L.lpo_noise_remover_filter.argtypes = [P, ctypes.POINTER(NoiseConfig), P, P, P, ctypes.c_int32, P]
    L.lpo_noise_remover_stack_depth.argtypes = [P]
    L.lpo_antipodal_pairs.argtypes = [P, P, ctypes.c_int32, P, ctypes.c_int32]
    L.lpo_antipodal_pairs.restype = ctypes.c_int32
    L.lpo_obb_rotating_calipers.argtypes = [P, P, ctypes.c_int32, ctypes.POINTER(BoundingBox)]
    L._ext_bound = True
    return L


def noise_config(**kw):
    c = NoiseConfig()
    _bind_ext().lpo_noise_config_default(ctypes.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def noise_radius_sqr(x, y, cfg=None):
    return _bind_ext().lpo_noise_radius_sqr(ctypes.byref(cfg or noise_config()), float(x), float(y))


def noise_definition(x, y, z, cfg=None):
    """The DROR rule: (labels u8 1 = NOISE, in-radius counts u32)."""
    L = _bind_ext()
    x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
    n = len(x)
    labels = np.zeros(n, np.uint8)
    counts = np.zeros(n, np.uint32)
    st = L.lpo_noise_definition(ctypes.byref(cfg or noise_config()), _p(x), _p(y), _p(z), n, _p(labels),
                                _p(counts))
    if st != OK:
        raise ValueError("noise_definition failed")
    return labels, counts


class NoiseRemover:
    """NoiseRemover::filter as the reference object behaves, search stack
    kept between queries and frames."""

    def __init__(self):
        self.L = _bind_ext()
        self.h = self.L.lpo_noise_remover_create()

    def __del__(self):
        if getattr(self, "h", None):
            self.L.lpo_noise_remover_destroy(self.h)
            self.h = None

    def filter(self, x, y, z, cfg=None):
        x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
        labels = np.zeros(len(x), np.uint8)
        st = self.L.lpo_noise_remover_filter(self.h, ctypes.byref(cfg or noise_config()), _p(x), _p(y), _p(z),
                                             len(x), _p(labels))
        if st != OK:
            raise ValueError("noise remover rejected the frame")
        return labels

    def stack_depth(self):
        return self.L.lpo_noise_remover_stack_depth(self.h)


def antipodal_pairs(x, y):
    L = _bind_ext()
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    cap = 4 * len(x) + 16
    out = np.zeros(2 * cap, np.int32)
    k = L.lpo_antipodal_pairs(_p(x), _p(y), len(x), _p(out), cap)
    if k < 0:
        raise ValueError("antipodal walk did not terminate")
    return out[: 2 * k].reshape(-1, 2)


def obb_rotating_calipers(x, y):
    """boundingBoxRotatingCalipers over hull vertices (CCW).  Returns dict."""
    L = _bind_ext()
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    b = BoundingBox()
    st = L.lpo_obb_rotating_calipers(_p(x), _p(y), len(x), ctypes.byref(b))
    if st != OK:
        raise ValueError("rotating calipers failed")
    return dict(corners=np.array(b.corners[:], np.float64).reshape(4, 2), area=np.float32(b.area),
                angle_rad=np.float32(b.angle_rad), is_valid=bool(b.is_valid))


def _bind_boxes():
    L = _bind_ext()
    if not getattr(L, "_boxes_bound", False):
        P = ctypes.c_void_p
        L.lpo_pca_box.argtypes = [P, P, ctypes.c_int32, ctypes.POINTER(BoundingBox)]
        L.lpo_svd2_v.argtypes = [P, P]
        L.lpo_vehicle_tables.argtypes = [P, P, P]
        L.lpo_vehicle_tables.restype = None
        L.lpo_vehicle_class.argtypes = [P, P, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_float,
                                        ctypes.POINTER(BoundingBox)]
        L._boxes_bound = True
    return L


def _box_dict(b):
    return dict(corners=np.array(b.corners[:], np.float64).reshape(4, 2), area=np.float32(b.area),
                angle_rad=np.float32(b.angle_rad), is_valid=bool(b.is_valid))


def pca_box(x, y):
    """boundingBoxPrincipalComponentAnalysis over a point list.  Returns dict."""
    L = _bind_boxes()
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    b = BoundingBox()
    L.lpo_pca_box(_p(x), _p(y), len(x), ctypes.byref(b))
    return _box_dict(b)


def svd2_v(a):
    """Right singular vectors (columns, largest first) of a 2x2 matrix."""
    L = _bind_boxes()
    a = np.ascontiguousarray(a, np.float64).reshape(4)
    v = np.zeros(4, np.float64)
    if L.lpo_svd2_v(_p(a), _p(v)) != OK:
        raise ValueError("non-finite matrix")
    return v.reshape(2, 2)


def vehicle_tables():
    """(dims[5,6], bounds[5,4], loose[4]) of processor.hpp:78-192."""
    L = _bind_boxes()
    dims, bounds, loose = np.zeros(30), np.zeros(20), np.zeros(4)
    L.lpo_vehicle_tables(_p(dims), _p(bounds), _p(loose))
    return dims.reshape(5, 6), bounds.reshape(5, 4), loose


def vehicle_class(hx, hy, cluster_size, z_min, z_max, box):
    """The processor's vehicle test for one cluster (box: a dict like pca_box's)."""
    L = _bind_boxes()
    hx = np.ascontiguousarray(hx, np.float64)
    hy = np.ascontiguousarray(hy, np.float64)
    b = BoundingBox()
    b.corners[:] = [float(v) for v in np.asarray(box["corners"], np.float64).reshape(8)]
    b.area = float(box["area"])
    b.angle_rad = float(box["angle_rad"])
    b.is_valid = 1 if box["is_valid"] else 0
    return int(L.lpo_vehicle_class(_p(hx), _p(hy), len(hx), int(cluster_size), float(z_min), float(z_max),
                                   ctypes.byref(b)))


def _bind_full():
    L = _bind_ext()
    if not getattr(L, "_full_bound", False):
        P = ctypes.c_void_p
        L.lpo_full_frame.argtypes = [P, P, ctypes.POINTER(NoiseConfig), P, P, P, P, ctypes.c_int32, P,
                                     ctypes.POINTER(ctypes.c_int32), P, ctypes.POINTER(ctypes.c_int32), P,
                                     ctypes.POINTER(ctypes.c_int32), P, ctypes.c_int32, P, ctypes.c_int32, P]
        L._full_bound = True
    return L


def full_frame(proc, x, y, z, ring=None, noise_remover=None, ncfg=None, bufs=None):
    """Noise filter -> keep VALID -> segment -> split -> cluster -> gather ->
    hull -> box per hull, on a Processor.  noise_remover: a NoiseRemover (the
    reference object's behaviour) or None (the radius rule)."""
    L = _bind_full()
    n = len(x)
    if bufs is None:
        bufs = dict(noise=np.zeros(max(n, 1), np.uint8), seg=np.zeros(max(n, 1), np.uint32),
                    clab=np.zeros(max(n, 1), np.int32), hoff=np.zeros(n + 2, np.int32),
                    hidx=np.zeros(2 * n + 16, np.int32), boxes=(BoundingBox * (n + 2))())
    kept = ctypes.c_int32(0)
    nobs = ctypes.c_int32(0)
    ncl = ctypes.c_int32(0)
    st = L.lpo_full_frame(proc.h, noise_remover.h if noise_remover is not None else None,
                          ctypes.byref(ncfg or noise_config()), _p(x), _p(y), _p(z), _p(ring), n, _p(bufs["noise"]),
                          ctypes.byref(kept), _p(bufs["seg"]), ctypes.byref(nobs), _p(bufs["clab"]),
                          ctypes.byref(ncl), _p(bufs["hoff"]), len(bufs["hoff"]), _p(bufs["hidx"]),
                          len(bufs["hidx"]), bufs["boxes"])
    k = ncl.value
    return dict(status=st, noise_labels=bufs["noise"][:n], num_points=kept.value,
                seg_labels=bufs["seg"][:kept.value], obstacle_count=nobs.value,
                cluster_labels=bufs["clab"][: nobs.value], num_clusters=k, hull_offsets=bufs["hoff"][: k + 1],
                hull_indices=bufs["hidx"][: bufs["hoff"][k]], boxes=bufs["boxes"])
