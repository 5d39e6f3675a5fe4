/*
 * lp_oracle.h -- CPU restatement of the lidar_processing_lib per-frame hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the MI355X
 * implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never links or calls it.
 *
 * It restates, function by function, the semantics of (paths relative to the
 * reference repository root):
 *   lidar_processing_lib/src/segmenter.cpp:31-669      (Segmenter)
 *   lidar_processing_lib/include/lidar_processing_lib/common.hpp:33-62 (atan2Approx)
 *   lidar_processing_lib/src/clusterer.cpp:27-239      (Clusterer)
 *   lidar_processing_lib/include/lidar_processing_lib/static_unordered_map.hpp:67-97 (capacity)
 *   lidar_processing_lib/src/polygonizer.cpp:33-91     (convexHull)
 *   src/processor/src/processor.cpp:552-670            (split / gather / hull glue)
 *   src/dataloader/src/dataloader.cpp:68-137           (addRingInfo)
 *   lidar_processing_lib/src/noise_remover.cpp:38-68 + kdtree.hpp:131-397 (NoiseRemover, DROR)
 *   lidar_processing_lib/src/polygonizer.cpp:93-278    (antipodal pairs, rotating-calipers box)
 *   lidar_processing_lib/src/polygonizer.cpp:280-362   (PCA box)
 *   src/processor/src/processor.cpp:672-760 + processor.hpp:78-192 (vehicle test)
 * and mirrors the reference's data-structure choices (vector-of-vector polar
 * grid, per-cell std::sort, FIFO JCP queue, linear-probing voxel map, BFS with
 * a visited set, std::sort hulls) so that its speed is representative.
 *
 * Parity pins (what checks each function against something other than its
 * own formula) are listed in tests/test_oracle_*.py.
 */
#ifndef LP_ORACLE_H
#define LP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same fields and defaults as SegmenterConfiguration (segmenter.hpp:87-112). */
typedef struct lpo_seg_config {
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
} lpo_seg_config;

/* Same fields as ClustererConfiguration (clusterer.hpp:61-68). */
typedef struct lpo_cluster_config {
    float voxel_grid_range_resolution_m;
    float voxel_grid_azimuth_resolution_deg;
    float voxel_grid_elevation_resolution_deg;
    uint32_t min_cluster_size;
} lpo_cluster_config;

/* Intermediate results of one segment() call, for stage-by-stage parity. */
typedef struct lpo_seg_debug {
    float* elevation_map;      /* [slices*rings] or NULL */
    int32_t* mapping;          /* [H*W] winner point index or -1, or NULL */
    uint8_t* state_recm;       /* [H*W] 0 empty 1 ground 2 obstacle after RECM+RANSAC, or NULL */
    uint8_t* state_final;      /* [H*W] 0 empty 1 ground 2 obstacle 3 pending(unresolved), or NULL */
    int32_t* queue;            /* [H*W] JCP queue pixels in processing order, or NULL */
    int32_t queue_len;
    uint32_t ransac_candidates;
    uint32_t ransac_best_inliers;
    int32_t ransac_applied;     /* best_inliers > 0 */
    float ransac_plane[4];      /* a,b,c,d after the sign flip */
    uint32_t ransac_draws[180]; /* p2,p3 indices of the 60 hypotheses (+ unused) */
    uint8_t* image;             /* [H*W*3] BGR image_ after segment() (Segmenter::image), or NULL */
} lpo_seg_debug;

enum {
    LPO_OK = 0,
    LPO_ERR_INVALID = -1,
    LPO_ERR_VOXEL_OVERFLOW = -4,   /* StaticUnorderedMap throws std::overflow_error */
    LPO_ERR_RANSAC_SINGLE = -5     /* one RANSAC candidate: the reference loops forever */
};

void lpo_seg_config_default(lpo_seg_config* c);
void lpo_cluster_config_default(lpo_cluster_config* c);

/* Segmenter::segment.  ring may be NULL (PointXYZ/PointXYZI instantiation). */
int lpo_segment(const lpo_seg_config* cfg, const float* x, const float* y, const float* z,
                const uint16_t* ring, int32_t n, uint32_t* labels, lpo_seg_debug* dbg);

/* Clusterer::cluster; max_voxels is the map capacity (200000 in the reference). */
int lpo_cluster(const lpo_cluster_config* cfg, const float* x, const float* y, const float* z,
                int32_t n, int32_t max_voxels, int32_t* labels);

/* Polygonizer::convexHull on one point list (doubles).  out must hold 2n. */
int32_t lpo_convex_hull(const double* x, const double* y, int32_t n, int32_t* out);

/* Processor glue: segment -> split -> cluster -> gather -> hull per cluster.
 * Outputs: seg_labels[n]; obstacle_count; cluster_labels[n] (first
 * obstacle_count valid); num_clusters; hull_offsets[num_clusters+1];
 * hull_indices (cluster-local indices).  Capacities: hull_cap entries. */
int lpo_process_frame(const lpo_seg_config* scfg, const lpo_cluster_config* ccfg,
                      const float* x, const float* y, const float* z, const uint16_t* ring,
                      int32_t n, int32_t max_voxels, uint32_t* seg_labels,
                      int32_t* obstacle_count, int32_t* cluster_labels, int32_t* num_clusters,
                      int32_t* hull_offsets, int32_t hull_offsets_cap, int32_t* hull_indices,
                      int32_t hull_cap);

/* Persistent processor (one Segmenter + Clusterer + Polygonizer, constructed
 * once like the reference's member objects, processor.hpp:248-258). */
typedef struct lpo_processor lpo_processor;
lpo_processor* lpo_processor_create(const lpo_seg_config* scfg, const lpo_cluster_config* ccfg,
                                    int32_t max_voxels);
void lpo_processor_destroy(lpo_processor* p);
int lpo_processor_segment(lpo_processor* p, const float* x, const float* y, const float* z,
                          const uint16_t* ring, int32_t n, uint32_t* labels, lpo_seg_debug* dbg);
int lpo_processor_cluster(lpo_processor* p, const float* x, const float* y, const float* z, int32_t n,
                          int32_t* labels);
int lpo_processor_frame(lpo_processor* p, const float* x, const float* y, const float* z,
                        const uint16_t* ring, int32_t n, uint32_t* seg_labels, int32_t* obstacle_count,
                        int32_t* cluster_labels, int32_t* num_clusters, int32_t* hull_offsets,
                        int32_t hull_offsets_cap, int32_t* hull_indices, int32_t hull_cap);

/* Dataloader::addRingInfo for a scan-ordered cloud. */
void lpo_add_ring_info(const float* x, const float* y, int32_t n, uint16_t* ring);

/* NoiseRemoverConfiguration (noise_remover.hpp:41-54). */
typedef struct lpo_noise_config {
    float radius_multiplier_m_per_m;
    float min_search_radius_m;
    uint32_t min_neighbours;
} lpo_noise_config;
void lpo_noise_config_default(lpo_noise_config* c);
/* Search radius squared of a point (noise_remover.cpp:46-59). */
float lpo_noise_radius_sqr(const lpo_noise_config* cfg, float x, float y);
/* The filter's rule: labels[i] = 1 (NOISE) when fewer than min_neighbours
 * points (i included) have dist_sqr <= r_i^2, else 0 (VALID).  counts
 * (optional) receives the full in-radius count. */
int lpo_noise_definition(const lpo_noise_config* cfg, const float* x, const float* y, const float* z, int32_t n,
                         uint8_t* labels, uint32_t* counts);
/* A persistent NoiseRemover with the KD-tree's search stack carried between
 * queries and frames (kdtree.hpp:339-397). */
typedef struct lpo_noise_remover lpo_noise_remover;
lpo_noise_remover* lpo_noise_remover_create(void);
void lpo_noise_remover_destroy(lpo_noise_remover* p);
int lpo_noise_remover_filter(lpo_noise_remover* p, const lpo_noise_config* cfg, const float* x, const float* y,
                             const float* z, int32_t n, uint8_t* labels);
int32_t lpo_noise_remover_stack_depth(const lpo_noise_remover* p);

/* BoundingBox (polygonizer.hpp:75-81): corners as x0,y0,..,x3,y3. */
typedef struct lpo_bounding_box {
    double corners[8];
    float area;
    float angle_rad;
    int32_t is_valid;
} lpo_bounding_box;
/* findAntipodalPairsOfConvexHull: pair count (pairs[2k], pairs[2k+1]), -1 if the walk does not end. */
int32_t lpo_antipodal_pairs(const double* x, const double* y, int32_t n, int32_t* pairs, int32_t cap);
/* boundingBoxRotatingCalipers over the hull vertices in CCW order. */
int lpo_obb_rotating_calipers(const double* x, const double* y, int32_t n, lpo_bounding_box* out);

/* The full single-frame pipeline: noise filter (the reference object when nr
 * is given, else the radius rule) -> keep VALID in input order -> segment ->
 * split -> cluster -> gather -> hull -> boundingBoxRotatingCalipers on each
 * hull polygon (boxes[num_clusters], may be NULL).  *kept = points kept. */
int lpo_full_frame(lpo_processor* p, lpo_noise_remover* nr, const lpo_noise_config* ncfg, const float* x,
                   const float* y, const float* z, const uint16_t* ring, int32_t n, uint8_t* noise_labels,
                   int32_t* kept, uint32_t* seg_labels, int32_t* obstacle_count, int32_t* cluster_labels,
                   int32_t* num_clusters, int32_t* hull_offsets, int32_t hull_offsets_cap, int32_t* hull_indices,
                   int32_t hull_cap, lpo_bounding_box* boxes);

/* boundingBoxPrincipalComponentAnalysis over a point list. */
int lpo_pca_box(const double* x, const double* y, int32_t n, lpo_bounding_box* out);
/* Right singular vectors of a 2x2 matrix as the PCA box computes them,
 * v = {v00, v01, v10, v11} (columns are the vectors, largest value first).
 * Returns 0, or -1 for a non-finite matrix. */
int lpo_svd2_v(const double* a, double* v);
/* Vehicle size tables of processor.hpp:78-192: dims[5*6] (min/max length,
 * width, height after the tolerances), bounds[5*4] (min/max volume, min/max
 * area), loose[4] (min/max height, min/max volume over all types). */
void lpo_vehicle_tables(double* dims, double* bounds, double* loose);
/* The polygon simplification test of processor.cpp:680-757 for one cluster:
 * hull vertices (hull order), the cluster's point count and z range, and the
 * cluster's box.  Returns the vehicle type 0..4 or -1. */
int32_t lpo_vehicle_class(const double* hx, const double* hy, int32_t nv, int32_t cluster_size, float z_min,
                          float z_max, const lpo_bounding_box* box);

/* Library math helpers exposed for pinning tests. */
float lpo_atan2_approx(float y, float x);
/* First `count` draws of mt19937{42} + uniform_int_distribution<uint32_t>{0,n-1}. */
void lpo_std_uniform_draws(uint32_t n, int32_t count, uint32_t* out);
/* Raw mt19937{42} outputs. */
void lpo_mt19937_raw(int32_t count, uint32_t* out);
float lpo_host_expf(float x);
/* Host libm over float bit patterns [lo, lo+count): which 0 expf, 1 atanf. */
void lpo_host_math_range(int32_t which, uint32_t lo, uint32_t count, uint32_t* out_bits);
float lpo_host_atanf(float x);
float lpo_host_atan2f(float y, float x);

#ifdef __cplusplus
}
#endif

#endif
