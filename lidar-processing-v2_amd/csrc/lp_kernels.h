// lp_kernels.h -- launchers of the HIP stages (host side).
#pragma once

#include "lp_device.h"
#include "lp_internal.h"

namespace lp {

// Stores msg as the calling thread's lp_last_error() text; returns code.
int set_last_error(int code, const char* msg);

// The per-frame checks of lp_batch_stage (sizes, pointers, configuration,
// record layout) without staging: LP_OK or the error lp_batch_stage would give.
int validate_frame(const lp_context* c, const lp_segmenter_config& cfg, const lp_cloud& cl);

// Records a named stage boundary on the context stream when profiling.
void stage_mark(lp_context* c, const char* name);

// Which per-frame count bounds a scan.
enum ScanCount : int { kCountPoints = 0, kCountPixels = 1, kCountSlot = 2, kCountHash = 3, kCountRawPoints = 4 };

// Exclusive scan of `in` into `out` for every frame of the batch; frame f
// covers in[f*stride .. f*stride + count_f).  The total goes to
// fs[f][total_slot] (if >= 0).
void launch_scan(lp_context* c, int B, const uint32_t* in, uint32_t* out, size_t stride, ScanCount kind,
                 int count_slot, int total_slot);
// The same in one launch, one workgroup per frame: for cluster-sized counts.
void launch_scan_small(lp_context* c, int B, const uint32_t* in, uint32_t* out, size_t stride, ScanCount kind,
                       int count_slot, int total_slot);
// The same with element i of frame f read as
// in[(f * in_stride + gather[f * stride + i]) * elem] (a field of a record array).
void launch_scan_gather(lp_context* c, int B, const uint32_t* in, size_t in_stride, int elem, const int32_t* gather,
                        uint32_t* out, size_t stride, ScanCount kind, int count_slot, int total_slot);
// The same, each result written back through the gather index into another
// field of the record array (`field` is that field of record 0).
void launch_scan_gather_into(lp_context* c, int B, const uint32_t* in, size_t in_stride, int elem,
                             const int32_t* gather, uint32_t* field, size_t stride, ScanCount kind, int count_slot);

void launch_mt19937(uint32_t* out, int count, hipStream_t st);
// Resets the hash slots slots[f * N .. + fs[count_slot]) of every frame to
// their clean state (all slots when the count exceeds N).
void launch_hash_reset(lp_context* c, int B, const int32_t* slots, int count_slot);
// Clears every hash slot of the context (at creation).
void launch_hash_init(lp_context* c);
void launch_segment_front(lp_context* c, int B);
void launch_jcp(lp_context* c, int B);
void launch_cluster(lp_context* c, int B, bool from_split);
// Ring indices of the scan-ordered cloud staged as frame 0 (x, y).
void launch_ring_info(lp_context* c, int n);
// Fill kernel of the per-batch device state (first stage of every run).
void launch_clear(lp_context* c, int B);
// Per-cluster hulls; `list_xy` (standalone lists only) holds the lists'
// coordinates as (x, y) pairs for the canonical choice among identical points.
void launch_hull(lp_context* c, int B, const double2* list_xy = nullptr);
// Noise filter over the staged raw inputs; with `compact`, the kept points
// become the segmenter's inputs.  The first applies the radius rule, the
// second the reference object's KD-tree search.
void launch_noise(lp_context* c, int B, bool compact);
void launch_noise_kdtree(lp_context* c, int B, bool compact);
// Kept points (keep flags of the noise stage) into the segmenter's inputs.
void launch_noise_compact(lp_context* c, int B);
// Tree levels built range by range before the per-subtree LDS kernel.
int kd_small_level_host(int n);
// Oriented boxes of the polygons in the second hull candidate buffer: the
// hulls of the last hull stage (written there by the local-rank kernel) or,
// for lists, the polygons staged by the caller.
void launch_boxes(lp_context* c, int B);

// Segmenter::image of frame f of the last run into img (H*W*3 bytes).
void launch_label_image(lp_context* c, int f, int pixels, uint8_t* img);
// Vehicle classes of the clusters of the last box stage.
void launch_vehicles(lp_context* c, int B);
// PCA boxes of polygons staged like lp_oriented_boxes's.
void launch_pca_boxes(lp_context* c, int num_polygons);

// Standalone hull entry (lp_convex_hulls): `num_lists` lists staged in the
// round-0 hull buffers of frame 0 with offsets in cl_off.
void launch_hull_lists(lp_context* c, int num_lists, int total_points);

}  // namespace lp
