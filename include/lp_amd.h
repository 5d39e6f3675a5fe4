/*
 * lp_amd.h -- C ABI of the MI355X (gfx950) implementation of the
 * lidar_processing_lib per-frame hot path.
 *
 * The ABI is what the reference's C++ API (namespace lidar_processing_lib)
 * binds to once its bodies are replaced by HIP kernels.  Plain pointers and
 * sizes only; no exceptions cross it.  Paths are relative to the reference
 * repository root.
 *
 *   lp_segment        replaces Segmenter::segment
 *                     (lidar_processing_lib/include/lidar_processing_lib/segmenter.hpp:168-169,
 *                      lidar_processing_lib/src/segmenter.cpp:87-101)
 *   lp_cluster        replaces Clusterer::cluster
 *                     (lidar_processing_lib/include/lidar_processing_lib/clusterer.hpp:92-93,
 *                      lidar_processing_lib/src/clusterer.cpp:43-53)
 *   lp_convex_hulls   replaces a loop of Polygonizer::convexHull calls
 *                     (lidar_processing_lib/include/lidar_processing_lib/polygonizer.hpp:104-106,
 *                      lidar_processing_lib/src/polygonizer.cpp:33-91)
 *   lp_process_frames replaces the per-frame body of Processor::run from the
 *                     segment call to the last hull (src/processor/src/processor.cpp:552-670)
 *   lp_remove_noise   replaces NoiseRemover::filter
 *                     (lidar_processing_lib/include/lidar_processing_lib/noise_remover.hpp:64,
 *                      lidar_processing_lib/src/noise_remover.cpp:38-68)
 *   lp_oriented_boxes replaces a loop of Polygonizer::boundingBoxRotatingCalipers calls
 *                     (lidar_processing_lib/include/lidar_processing_lib/polygonizer.hpp:115,
 *                      lidar_processing_lib/src/polygonizer.cpp:93-278)
 *   lp_pca_boxes      replaces a loop of Polygonizer::boundingBoxPrincipalComponentAnalysis calls
 *                     (lidar_processing_lib/include/lidar_processing_lib/polygonizer.hpp:118-119,
 *                      lidar_processing_lib/src/polygonizer.cpp:280-362)
 *   lp_frame_image    replaces Segmenter::image
 *                     (lidar_processing_lib/include/lidar_processing_lib/segmenter.hpp:163-166)
 *
 * Conventions (mirroring the reference):
 *   - the caller owns inputs and outputs; every output entry is written
 *     (like labels.assign, segmenter.cpp:90 / clusterer.cpp:124);
 *   - one context = one HIP stream and its preallocated device buffers, used
 *     from one thread at a time (the reference objects are not thread safe);
 *     separate contexts may run concurrently;
 *   - errors are negative status codes; lp_last_error() gives the message of
 *     the calling thread's last failure.  The reference's exceptions map to:
 *     std::overflow_error (static_unordered_map.hpp:79-81) -> LP_ERR_VOXEL_OVERFLOW.
 */
#ifndef LP_AMD_H
#define LP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    LP_OK = 0,
    LP_ERR_INVALID_ARGUMENT = -1,
    LP_ERR_HIP = -2,
    LP_ERR_CAPACITY = -3,             /* a context capacity (points, image, cell, hull) was exceeded */
    LP_ERR_VOXEL_OVERFLOW = -4,       /* more than max_voxels occupied voxels (reference: std::overflow_error) */
    LP_ERR_RANSAC_SINGLE_CANDIDATE = -5 /* exactly one RANSAC candidate: the reference's redraw loop
                                           (segmenter.cpp:383-386) never terminates */
};

/* Label values (segmenter.hpp:69-74). */
enum { LP_LABEL_UNKNOWN = 0, LP_LABEL_GROUND = 1, LP_LABEL_OBSTACLE = 2 };

/* SegmenterConfiguration (segmenter.hpp:87-112): same fields, same defaults.
 * display_recm_with_low_confidence_points (an OpenCV window) is not carried. */
typedef struct lp_segmenter_config {
    float elevation_up_deg;
    float elevation_down_deg;
    int32_t image_width;
    int32_t image_height;
    int32_t assume_unorganized_cloud;
    float grid_radial_spacing_m;
    float grid_slice_resolution_deg;
    float ground_height_threshold_m;
    float road_maximum_slope_m_per_m;
    float min_distance_m;
    float max_distance_m;
    float sensor_height_m;
    float kernel_threshold_distance_m;
    float amplification_factor;
    float z_min_m;
    float z_max_m;
} lp_segmenter_config;

/* ClustererConfiguration (clusterer.hpp:61-68). */
typedef struct lp_clusterer_config {
    float voxel_grid_range_resolution_m;
    float voxel_grid_azimuth_resolution_deg;
    float voxel_grid_elevation_resolution_deg;
    uint32_t min_cluster_size;
} lp_clusterer_config;

/* Defaults of segmenter.hpp:91-108. */
void lp_segmenter_config_default(lp_segmenter_config* cfg);
/* Header defaults of clusterer.hpp:61-68 (elevation resolution 1.5 deg). */
void lp_clusterer_config_default(lp_clusterer_config* cfg);
/* Shipped processor.param.yaml:31-35 values (elevation resolution 3.0 deg). */
void lp_clusterer_config_shipped(lp_clusterer_config* cfg);

/* Device buffers are preallocated for these limits at context creation
 * (the reference reserves fixed capacities in its constructors:
 * segmenter.cpp:38-71, clusterer.cpp:27-41). */
typedef struct lp_capacity {
    int32_t max_frames;            /* frames per batch */
    int32_t max_points_per_frame;
    int32_t max_image_height;
    int32_t max_image_width;
    int32_t max_voxels;            /* 200000 in the reference (clusterer.cpp:36) */
    int32_t device;                /* HIP device ordinal */
    int32_t streams;               /* 1..8: a staged batch runs as this many contiguous
                                      sub-batches, each on its own stream and graph, so the
                                      latency-bound kernels of one overlap the others'
                                      work; max_frames is the total over all of them */
    float hash_load_factor;        /* worst-case load of the voxel and noise-cell hash tables
                                      (capacity = next power of two of points / factor):
                                      0.5 (default) .. 0.9; a smaller table is cleared and
                                      probed faster, a larger one has shorter probe runs */
} lp_capacity;
void lp_capacity_default(lp_capacity* cap);

typedef struct lp_context lp_context;
int lp_context_create(const lp_capacity* cap, lp_context** out);
int lp_context_destroy(lp_context* ctx);
const char* lp_last_error(void);

/* A point cloud: x, y, z (float) and optional ring (uint16) addressed with
 * byte strides, so both AoS pcl::PointXYZIR (stride sizeof(point)) and SoA
 * arrays (stride 4 / 2) are accepted without a host-side copy.  ring == NULL
 * selects the ring-less instantiations (PointXYZ / PointXYZI), whose image
 * row comes from the elevation angle (segmenter.cpp:180-193). */
typedef struct lp_cloud {
    const void* x;
    const void* y;
    const void* z;
    const void* ring;
    int32_t xyz_stride;
    int32_t ring_stride;
    int32_t num_points;
    int32_t on_device;      /* 1: the pointers are device pointers */
} lp_cloud;

/* A sensor_msgs::msg::PointCloud2 field as Processor::run reads it: its name
 * and byte offset inside a point. */
typedef struct lp_point_field {
    const char* name;
    uint32_t offset;
} lp_point_field;

/* Processor::run's field dispatch (processor.cpp:494-549) and convert<>
 * (processor.cpp:42-179) without the AoS copy.  The field names must be
 * exactly {x,y,z}, {x,y,z,intensity}, {x,y,z,ring} or {x,y,z,intensity,ring}
 * (else LP_ERR_INVALID_ARGUMENT, the reference's "Unsupported point format"
 * error); x, y, z are the floats at fields[0..2].offset and ring the uint16 at
 * the ring field's offset of every point.  Rows hold `width` points of
 * `point_step` bytes, `row_step` bytes apart.  Fills *cloud with views into
 * `data` (no copy) when the rows are packed (height <= 1 or row_step ==
 * width * point_step); other layouts return LP_ERR_INVALID_ARGUMENT (pack
 * the rows first).  No GPU work; intensity is not read by the path. */
int lp_cloud_from_pointcloud2(const uint8_t* data, uint32_t width, uint32_t height, uint32_t point_step,
                              uint32_t row_step, const lp_point_field* fields, int32_t num_fields,
                              lp_cloud* cloud);

/* The RANSAC plane the reference keeps in locals (segmenter.cpp:363-374). */
typedef struct lp_ransac_result {
    uint32_t candidates;
    uint32_t best_inliers;
    int32_t applied;        /* best_inliers > 0 */
    float plane[4];         /* a, b, c, d after the sign normalisation (c >= 0) */
} lp_ransac_result;

/* Segmenter::segment: labels[num_points] receives LP_LABEL_*; ransac may be NULL. */
int lp_segment(lp_context* ctx, const lp_segmenter_config* cfg, const lp_cloud* cloud, uint32_t* labels,
               lp_ransac_result* ransac);

/* Clusterer::cluster: labels[num_points] receives the cluster id or -1. */
int lp_cluster(lp_context* ctx, const lp_clusterer_config* cfg, const lp_cloud* cloud, int32_t* labels);

/* Polygonizer::convexHull on num_lists point lists packed by offsets
 * (list k is points offsets[k] .. offsets[k+1]-1).  Output: hull_offsets
 * [num_lists+1] and hull_indices (list-local indices, CCW from the
 * lexicographically smallest point), capacity hull_capacity entries. */
int lp_convex_hulls(lp_context* ctx, const double* x, const double* y, const int32_t* offsets, int32_t num_lists,
                    int32_t* hull_offsets, int32_t* hull_indices, int32_t hull_capacity);

/* NoiseRemoverConfiguration (noise_remover.hpp:41-54): same fields, same defaults. */
typedef struct lp_noise_config {
    float radius_multiplier_m_per_m;
    float min_search_radius_m;
    uint32_t min_neighbours;
} lp_noise_config;
void lp_noise_config_default(lp_noise_config* cfg);

/* How the noise filter searches (lp_remove_noise and the fused path's filter):
 *   LP_NOISE_RULE    the filter's stated rule: NOISE when fewer than
 *                    min_neighbours points, itself included, lie within the
 *                    point's search radius (the default);
 *   LP_NOISE_KDTREE  what a NoiseRemover object computes for the frame
 *                    (a fresh object per frame): its KD-tree
 *                    (kdtree.hpp:170-213, std::nth_element of libstdc++)
 *                    and radius_search_k_nearest (kdtree.hpp:339-397), which
 *                    stops at the k-th neighbour without clearing its stack,
 *                    so a later query short of k neighbours continues through
 *                    the earlier queries' unvisited subtrees and can count a
 *                    point again.  min_neighbours up to 7; non-finite
 *                    coordinates give LP_ERR_INVALID_ARGUMENT.  Its device
 *                    buffers (about 208 bytes per point and frame) are
 *                    allocated when it is first selected.
 * In this build LP_NOISE_KDTREE is rejected with LP_ERR_INVALID_ARGUMENT
 * ("switched off"): its kernels are compiled but not reachable. */
enum { LP_NOISE_RULE = 0, LP_NOISE_KDTREE = 1 };
int lp_set_noise_search(lp_context* ctx, int32_t mode);

/* NoiseRemover::filter: labels[num_points] receives 0 (VALID) or 1 (NOISE),
 * by the context's search (lp_set_noise_search). */
int lp_remove_noise(lp_context* ctx, const lp_noise_config* cfg, const lp_cloud* cloud, uint8_t* labels);

/* BoundingBox (polygonizer.hpp:75-81): corners as x0,y0,x1,y1,x2,y2,x3,y3. */
typedef struct lp_bounding_box {
    double corners[8];
    float area;
    float angle_rad;
    int32_t is_valid;
    int32_t reserved;
} lp_bounding_box;

/* Polygonizer::boundingBoxRotatingCalipers on num_polygons convex polygons
 * (vertices counter-clockwise, as convexHull returns them) packed by offsets. */
int lp_oriented_boxes(lp_context* ctx, const double* x, const double* y, const int32_t* offsets,
                      int32_t num_polygons, lp_bounding_box* boxes);

/* Polygonizer::boundingBoxPrincipalComponentAnalysis on num_polygons point
 * lists packed by offsets: the box of the points in the frame of the right
 * singular vectors of their covariance (largest first), corners in the order
 * (min,min), (max,min), (max,max), (min,max) of that frame.  Lists of fewer
 * than 3 points give is_valid = 0. */
int lp_pca_boxes(lp_context* ctx, const double* x, const double* y, const int32_t* offsets,
                 int32_t num_polygons, lp_bounding_box* boxes);

/* Dataloader::addRingInfo (src/dataloader/src/dataloader.cpp:68-137): ring
 * index per point of a scan-ordered cloud (63 down to 0, one step at each
 * fourth-to-first azimuth quadrant transition). */
int lp_add_ring_info(lp_context* ctx, const float* x, const float* y, int32_t n, uint16_t* ring);

/* Dataloader::addRingInfo without allocating anything: x, y (n floats) and
 * ring (n uint16) are the caller's device arrays; the work is queued on
 * `stream` (a hipStream_t, NULL = the default stream) and not waited for.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, serving loops that forbid hipMalloc).  The
 * per-tile transition counts live in a fixed device array of 16384 entries
 * (64 KiB on each device, held until the process ends), so n is limited to
 * 16777216 points (else LP_ERR_CAPACITY); calls that overlap on one device
 * race on that array.  Host pointers give LP_ERR_INVALID_ARGUMENT. */
int lp_add_ring_info_noalloc(const float* x, const float* y, int32_t n, uint16_t* ring, void* stream);

/* Fused per-frame path of the processor: segment -> obstacle split (input
 * order, processor.cpp:562-579) -> cluster -> per-cluster gather -> hull. */
typedef struct lp_frame_outputs {
    uint32_t* seg_labels;       /* [num_points] */
    int32_t* obstacle_index;    /* [num_points]: input index of each obstacle point (first num_obstacles used) */
    int32_t* cluster_labels;    /* [num_points]: cluster id per obstacle point (first num_obstacles used) */
    int32_t* hull_offsets;      /* [num_points + 1] (first num_clusters+1 used) */
    int32_t* hull_indices;      /* cluster-local hull vertex indices */
    int32_t hull_capacity;
    uint8_t* noise_labels;      /* [num_points] with the noise filter on (else unused, may be NULL) */
    lp_bounding_box* boxes;     /* [num_points] (first num_clusters used) with boxes on, else may be NULL */
    int32_t* ground_index;      /* [num_points] with label clouds on: input index of each GROUND point */
    int32_t* unknown_index;     /* [num_points] with label clouds on: input index of each UNKNOWN point */
    int32_t* vehicle_class;     /* [num_points] (first num_clusters used) with vehicle boxes on: -1 or 0..4 */
} lp_frame_outputs;

typedef struct lp_frame_result {
    int32_t status;             /* LP_OK or an error for this frame only */
    int32_t num_points;
    int32_t num_obstacles;
    int32_t num_voxels;
    int32_t num_clusters;
    int32_t num_hull_vertices;
    int32_t jcp_queue;          /* pixels resolved by JCP */
    lp_ransac_result ransac;
    int32_t num_noise;          /* points removed by the noise filter */
    int32_t num_boxes;          /* clusters with an oriented box computed */
    int32_t num_noise_searched; /* points the noise filter took past their own grid cell */
    int32_t num_ground;         /* GROUND points (label clouds on, else 0) */
    int32_t num_unknown;        /* UNKNOWN points (label clouds on, else 0) */
    int32_t num_vehicles;       /* clusters classified as a vehicle (vehicle boxes on, else 0) */
    int32_t num_hull_inexact;   /* hull orientation tests whose double value (the reference's predicate,
                                   polygonizer.cpp:45-48) had another sign than the exact one */
    int32_t num_hull_rounded;   /* clusters whose hull was recomputed as the reference computes it
                                   (full sort, rounded double predicate) because of such a test */
} lp_frame_result;

/* Optional stages of the fused path, kept by the context until changed:
 * the noise filter ahead of the segmenter (NULL switches it off; the kept
 * points, in input order, are the cloud every later output indexes, like the
 * filter -> keep VALID -> segment order of the full pipeline), and one oriented
 * box per cluster hull (boundingBoxRotatingCalipers over the hull vertices). */
int lp_set_noise_filter(lp_context* ctx, const lp_noise_config* cfg);
int lp_set_oriented_boxes(lp_context* ctx, int32_t enable);
/* The ground / obstacle / unsegmented split of processor.cpp:557-578 in full:
 * besides obstacle_index, the input-order indices of the GROUND and of the
 * UNKNOWN points (ground_cloud_ and unsegmented_cloud_). */
int lp_set_label_clouds(lp_context* ctx, int32_t enable);
/* The vehicle test of the processor's polygon simplification
 * (processor.cpp:680-760 with the tables of processor.hpp:78-192), per
 * cluster: a hull of >= 3 vertices, more than 150 points, height z_max - z_min
 * and hull volume inside the loosest vehicle bounds, a valid rotating-calipers
 * box covering the hull with area ratio > 0.4, then the first vehicle type
 * (0 compact, 1 sedan, 2 SUV, 3 truck, 4 minivan) whose length, width,
 * height, volume and area bounds all hold; -1 when none does.  A cluster with
 * class >= 0 is drawn as its box (boxes[k].corners).  Switches the boxes on
 * for the clusters it tests.  The reference's branch (disabled there by a
 * constant, processor.cpp:676) fits the box to the raw cluster points; this
 * fits it to the cluster's hull, the polygon the calipers are defined on. */
int lp_set_vehicle_boxes(lp_context* ctx, int32_t enable);

/* Launch form of the widest per-point kernels (projection, labels,
 * spherical coordinates, voxel insert): 0 = one workgroup per (tile, frame)
 * of the capacity grid (the default); 1 = a persistent launch of a fixed
 * number of workgroups taking (frame, tile) items from per-XCD work queues
 * (a frame's items stay on one XCD), so frames of uneven sizes leave no idle
 * workgroups.  Outputs are identical in both forms.  No reference
 * counterpart (the reference processes one frame on the CPU). */
int lp_set_persistent(lp_context* ctx, int32_t enable);
/* Launch shape, tuned per device and workload (lidar_amd.autotune): the
 * workgroup size of the per-point kernels (projection, labels, spherical
 * coordinates, voxel insert) and the tile width in columns of the JCP queue
 * marking (a 5-row LDS tile per workgroup); each 64, 128 or 256 (default
 * 256 and 256).  Outputs are identical under every shape. */
int lp_set_launch_shape(lp_context* ctx, int32_t point_workgroup, int32_t mark_tile_width);

/* Processes num_frames frames as one batch (CSR-packed on the device, one
 * launch sequence).  cfgs may hold one config (shared) or num_frames. */
int lp_process_frames(lp_context* ctx, const lp_segmenter_config* seg_cfgs, int32_t num_seg_cfgs,
                      const lp_clusterer_config* cluster_cfg, int32_t num_frames, const lp_cloud* clouds,
                      const lp_frame_outputs* outputs, lp_frame_result* results);

/* Split form of lp_process_frames for device-resident pipelines: stage
 * inputs, run (asynchronously on the context stream, captured once per
 * batch shape as a hipGraph and replayed), then fetch results.  A context
 * created with streams = S > 1 stages the batch as S contiguous shares of
 * ceil(num_frames / S) frames and runs them on S streams; the context
 * stream waits for all of them, so lp_synchronize / lp_query and work the
 * caller queues behind lp_context_stream see the whole batch. */
int lp_batch_stage(lp_context* ctx, const lp_segmenter_config* seg_cfgs, int32_t num_seg_cfgs,
                   const lp_clusterer_config* cluster_cfg, int32_t num_frames, const lp_cloud* clouds);
int lp_batch_run(lp_context* ctx, int32_t use_graph);
int lp_batch_fetch(lp_context* ctx, const lp_frame_outputs* outputs, lp_frame_result* results);
int lp_synchronize(lp_context* ctx);
/* Non-blocking: 1 when all work queued on the context has finished, 0 while
 * it is still running, negative on error. */
int lp_query(lp_context* ctx);

/* Serving front end for several sensors.  No reference counterpart: the
 * reference processor handles one frame per point-cloud callback on a
 * single-threaded executor (src/processor/src/processor.cpp:878-894); the
 * server feeds that per-frame path from many sensors with continuous
 * batching: frames queue as they arrive and, whenever one of `slots`
 * contexts (one HIP stream each) is free, up to max_batch_frames queued
 * frames of any sizes run on it as one batch, overlapping the batches of the
 * other slots.  The slots created are those whose preallocated buffers fit
 * hbm_budget_bytes.  A frame's outputs are the same alone and in any mix. */
typedef struct lp_server_config {
    int32_t slots;                 /* contexts (batches in flight at once) */
    int32_t max_batch_frames;      /* frames one slot runs as one batch */
    lp_capacity capacity;          /* per-frame limits of every slot (max_frames is ignored) */
    uint64_t hbm_budget_bytes;     /* device memory the slots may hold; 0 = no limit */
    lp_clusterer_config cluster;
    int32_t noise_filter;          /* 1: noise filter (below) ahead of the segmenter */
    lp_noise_config noise;
    int32_t oriented_boxes;        /* 1: a rotating-calipers box per cluster hull */
    int32_t label_clouds;          /* 1: ground and unknown index lists (lp_set_label_clouds) */
    int32_t vehicle_boxes;         /* 1: vehicle classes per cluster (lp_set_vehicle_boxes) */
    int32_t noise_search;          /* LP_NOISE_RULE or LP_NOISE_KDTREE (lp_set_noise_search) */
} lp_server_config;
void lp_server_config_default(lp_server_config* cfg);

typedef struct lp_server lp_server;
int lp_server_create(const lp_server_config* cfg, lp_server** out);
int lp_server_destroy(lp_server* s);
/* Queues one frame.  cfg is copied; the cloud's memory, the outputs and the
 * result must stay valid until lp_server_poll returns the frame's tag. */
int lp_server_submit(lp_server* s, const lp_segmenter_config* cfg, const lp_cloud* cloud,
                     const lp_frame_outputs* outputs, lp_frame_result* result, uint64_t tag);
/* Non-blocking: collects finished batches (writing outputs and results),
 * starts queued frames on free slots, and writes up to max_done tags of
 * completed frames to done.  Returns the number written, or an error. */
int lp_server_poll(lp_server* s, uint64_t* done, int32_t max_done);

typedef struct lp_server_stats {
    int32_t slots;
    int32_t busy_slots;
    int32_t queued;
    int32_t in_flight;
    uint64_t slot_bytes;           /* device bytes held by each slot */
    uint64_t batches;
    uint64_t frames;
} lp_server_stats;
int lp_server_stats_get(lp_server* s, lp_server_stats* st);

/* Timing of the dominant kernel for roofline reporting: average duration (ms)
 * of the last `runs` launches of the named stage, from HIP events recorded on
 * the context stream.  With streams > 1 it runs the first share of the
 * staged batch alone. */
int lp_stage_names(lp_context* ctx, char* buf, int32_t buf_len);
int lp_profile_run(lp_context* ctx, int32_t runs, float* stage_ms, int32_t max_stages, int32_t* num_stages);

/* Per-context device stream (hipStream_t) for callers that interleave work. */
void* lp_context_stream(lp_context* ctx);

/* Intermediate state of frame `frame` of the last run (of the first share
 * when streams > 1), for stage-by-stage parity checks.  No reference counterpart: the reference keeps these in
 * private members (segmenter.hpp:190-215).  Returns the element count
 * written (or needed, when larger than `capacity`). */
enum {
    LP_DEBUG_ELEVATION_MAP = 0, /* float[slices*rings], RECM elevation_map_ */
    LP_DEBUG_MAPPING = 1,       /* int32[H*W], cloud_mapping_indices_ (-1 = empty) */
    LP_DEBUG_STATE_RECM = 2,    /* uint8[H*W]: 0 empty, 1 ground, 2 obstacle after RECM + RANSAC */
    LP_DEBUG_JCP_QUEUE = 3,     /* int32[queue]: JCP queue pixels in processing order */
    LP_DEBUG_STATE_FINAL = 4    /* uint8[H*W]: 0 empty, 1 ground, 2 obstacle, 3 unresolved */
};
int lp_debug_fetch(lp_context* ctx, int32_t frame, int32_t which, void* dst, int32_t capacity);

/* Segmenter::image after segment() for frame `frame` of the last run: H*W
 * BGR pixels (row-major, row 0 = image row 0), black where no point landed,
 * red (0,0,255) for obstacle pixels and for empty pixels the 5x5 dilation of
 * the obstacle pixels reaches, green (0,255,0) for ground, blue (255,0,0) for
 * queue pixels JCP left unresolved (segmenter.cpp:288-318, 484-513, 626-635).
 * Built on the device from the run's pixel states.  Returns the byte count
 * written (or needed, when larger than `capacity`). */
int lp_frame_image(lp_context* ctx, int32_t frame, uint8_t* bgr, int32_t capacity);

/* The device restatements of the C library functions the reference calls
 * (expf: segmenter.cpp:588; atanf / atan2f: clusterer.cpp:74,86,
 * segmenter.cpp:157,182) and of atan2Approx (common.hpp:33-62), evaluated on
 * the GPU (on_device = 1) or from the same source on the host (0).  No
 * reference counterpart: it exists so tests can compare them with the host
 * libm bit for bit.  b is the second argument of the two-argument forms. */
enum { LP_MATH_EXPF = 0, LP_MATH_ATANF = 1, LP_MATH_ATAN2F = 2, LP_MATH_ATAN2_APPROX = 3 };
int lp_math_eval(int32_t which, const float* a, const float* b, float* out, int32_t n, int32_t on_device);
/* Device check of a contiguous range of float bit patterns [lo, lo+count)
 * against host results `ref_bits` (one-argument functions only). */
int lp_math_check_range(int32_t which, uint32_t lo, uint32_t count, const uint32_t* ref_bits, uint64_t* mismatches);

#ifdef __cplusplus
}
#endif

#endif
