// lp_internal.h -- shared definitions of the HIP implementation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "lp_amd.h"

namespace lp {

// Pixel / point label codes (segmenter.hpp:69-74); PENDING is the JCP
// "intersection" colour of a queue pixel whose weights never resolved.
enum : uint8_t { kEmpty = 0, kGround = 1, kObstacle = 2, kPending = 3 };

constexpr int kRansacIters = 60;      // segmenter.cpp:324
constexpr int kRansacRadialBins = 4;  // segmenter.cpp:323
constexpr int kMtRawCount = 4096;     // precomputed mt19937{42} outputs
constexpr int kSmallCell = 512;       // cells up to this size are sorted by one wave
constexpr int kBigCell = 8192;        // larger cells: one workgroup; beyond: capacity error
constexpr int kHullChunk = 256;       // points per chunk-hull wave
constexpr int kHullFinal = 8192;      // candidates a final hull workgroup can hold
constexpr int kNoiseOccWords = 1024 * 1024 / 32;  // noise filter: 1024 x 1024 column bits per frame
constexpr int32_t kHashMinClean = 0x7fffffff;     // aux of a free hash slot (the voxel minimum's identity)

// Per-frame error bits (device), mapped to status codes on fetch.
enum : uint32_t {
    kErrCellCapacity = 1u << 0,
    kErrRansacSingle = 1u << 1,
    kErrRngExhausted = 1u << 2,
    kErrQueueCapacity = 1u << 3,
    kErrVoxelOverflow = 1u << 4,
    kErrHullCapacity = 1u << 5,
    kErrBoxCapacity = 1u << 6,
    kErrHullFuzzyOrder = 1u << 7,  // distinct x closer than DBL_EPSILON in a redone hull
    kErrNoiseInput = 1u << 8,      // non-finite coordinates for the KD-tree noise search
    kErrNoiseCapacity = 1u << 9,   // pending search entries beyond the KD-tree search's capacity
};

// Segmenter constants derived on the host exactly as the reference derives
// them (segmenter.cpp:42-46, 106-109, 121-122, 209-211, 359-360, 532-533).
struct SegConst {
    int32_t H, W;
    int32_t rings, slices, cells;
    int32_t organized;        // use the ring field as the image row
    float slice_res;          // rad
    float zlo, zhi;           // -h + z_min, -h + z_max
    float spacing, min_d, max_d;
    float down_rad, per_pixel;
    float w_minus_1;          // (float)(W - 1)
    float thr;                // ground_height_threshold_m
    float two_thr;            // 2.0f * thr
    float dz;                 // min(spacing * tanf(slope), thr - FLT_EPSILON)
    float e0;                 // -h + thr
    float neg_h;              // -sensor_height_m
    float cos_max;            // cosf(tanf(slope))
    float kthr2;              // kernel_threshold^2
    float neg_amp;            // -amplification_factor
};

// Noise filter constants (noise_remover.cpp:44-48): the squared radius
// multiplier (computed in double by std::pow on the host), the squared
// minimum radius (float) and the neighbour threshold.
struct NoiseConst {
    double scaling;
    float min_sqr;
    uint32_t k;
};

// Clusterer constants (clusterer.hpp:178-185).
struct CluConst {
    float res_r, res_az, res_el;
    uint32_t min_size;
    int32_t max_voxels;
};

// Per-frame descriptor of a batch (device copy).
struct FrameDesc {
    int32_t pt_off;   // CSR offset of the frame's points in the packed inputs
    int32_t n;        // points seen by the segmenter (after the noise filter, set on the device)
    int32_t n_raw;    // points staged
    int32_t pix_off;  // CSR offset of the frame's pixels
    int32_t has_ring;
    SegConst sc;
};

// One slot of the shared open-addressing hash (frame stride hash_cap), all
// fields in one 16-byte record so a touched slot costs one cache line:
// key (-1 free), count (points in the cell / voxel), dense (the occupied
// slot's dense id) and aux (voxels: minimum obstacle index, kHashMinClean
// when free; noise cells: first cell-sorted position, reset to
// kHashMinClean afterwards).
struct alignas(16) HashSlot {
    int32_t key;
    uint32_t count;
    int32_t dense;
    int32_t aux;
};

// Device pointers of one context (all sized for the capacity).
struct DevBuf {
    // packed inputs (CSR by FrameDesc::pt_off)
    float *x, *y, *z;
    uint16_t* ring;
    // staged inputs ahead of the noise filter, its labels and the cell-sorted
    // points of its grid (frame stride N)
    float *rx, *ry, *rz;
    uint16_t* rring;
    uint8_t* noise;
    float4* dsort;
    int32_t* spos;           // per input point (frame stride N): its position in dsort
    int32_t* nopen;          // = cl_members (free until clustering): points the input-order pass left open
    int32_t* nloc;           // = clabel (free until clustering): each point's place in its cell, then the listed points' input indices
    uint32_t* nocc;          // per frame: kNoiseOccWords words, one bit per occupied 1 m (x, y) column
    int32_t* ncell_slot;     // per frame (stride N): hash slot of each occupied noise cell, by dense id
    // per point
    int32_t* pt_cell;
    float4* pt_xyzc;         // per point (CSR): x, y, z and the polar cell (int bits), for gathers
    int32_t* pt_pix;
    uint32_t* labels;
    uint32_t* obs_flag;      // scan scratch over points
    int32_t* obs_src;        // [frame stride N] input index of obstacle k
    float *ox, *oy, *oz;     // obstacle xyz (frame stride N)
    float *sr, *sa, *se;     // spherical coords of obstacle points
    int32_t* vslot;          // hash slot of each obstacle point
    uint32_t* cl_flag;       // first-point flags / scan
    int32_t* clabel;         // cluster label per obstacle point
    // per cell (frame stride cell_stride)
    uint32_t* cell_count;
    uint32_t* cell_start;    // stride cell_stride + 1 used
    int32_t* pt_loc;         // per point (CSR): its place inside its polar cell
    int32_t* cell_items;     // frame stride N
    uint32_t* cell_zkey;     // frame stride N: ord_f32(z) of the cell_items entries, in fill order
    float* zmin;
    float* elev;
    uint32_t* big_cells;     // per frame list of cells needing a workgroup sort
    uint4* cell_list;        // per frame: the non-empty cells in cell order, {cell, start, count, 0}
    // RANSAC (per frame)
    uint32_t* cand_count;    // frame stride cand_stride
    uint32_t* cand_start;
    float *cx, *cy, *cz;     // candidates, frame stride N
    float* hyp;              // [B][60][4]
    uint32_t* hyp_valid;     // [B][60]
    uint32_t* hyp_count;     // [B][60]
    uint32_t* mt_raw;        // [kMtRawCount]
    // per pixel (CSR by pix_off)
    unsigned long long* pix_key;
    uint8_t* state;
    int32_t* mapping;
    float4* pxyz;            // per pixel: the winner's x, y, z and its index (int bits, -1 when empty)
    uint8_t* qflag;          // 1: the pixel joins the JCP queue
    int32_t* qmap;
    // JCP queue (frame stride pix_stride)
    int32_t* qlist;
    int32_t* qrow;           // [frame][H + 1] first queue index of each row
    float* slot_w;           // [q][24]
    int32_t* slot_src;       // [q][24]
    uint8_t* qsum_ok;
    uint8_t* fin;
    // voxel hash (frame stride hash_cap)
    HashSlot* hslot;
    // dense voxels (frame stride N)
    int32_t* vlist;          // dense id -> hash slot (insertion order)
    int32_t* vpar;           // union-find parent / component root (dense ids)
    int32_t* vnb;            // [13][N] per frame: dense id of each forward neighbour voxel, -1 none
    int32_t* vmin;           // per root: minimum obstacle index of the component
    uint32_t* vcnt;          // per root: points in the component
    int32_t* vlab;           // per root: cluster label (components of at least min_size points)
    // clusters (frame stride N)
    int32_t* cl_root;
    uint32_t* cl_size;
    uint32_t* cl_off;
    uint32_t* cl_cursor;
    uint32_t* hull_flag;     // = cl_cursor (free after the scatter): per cluster, 1 when a hull test's double sign was uncertain
    int32_t* cl_members;
    // hull
    double* hcx;             // candidate coordinates, frame stride 2N (two buffers)
    double* hcy;
    int32_t* hci;            // candidate obstacle index
    uint32_t* hc_count;      // per cluster candidates (two buffers)
    uint32_t* hc_choff;      // their exclusive scan
    // Per chunk of a hull round (written by the round's plan): its cluster,
    // with bit 31 set when the chunk is the cluster's only one (= cell_items,
    // free in the hull stage); its index in the cluster and, from bit 16, its
    // point count (= cell_zkey); the cluster's first candidate (= hull_large,
    // free until the local ranks).
    int32_t* hc_chunk_c;
    uint32_t* hc_chunk_i;
    uint32_t* hc_chunk_b;
    int32_t* hull_large;     // [B][N] clusters left to the workgroup local-rank launch
    int32_t* hull_idx;       // final hull (obstacle indices), frame stride N
    uint32_t* hull_count;    // per cluster
    uint32_t* hull_off;      // per cluster (+1)
    int32_t* hull_local;     // cluster-local indices, frame stride N
    // generic scan scratch: block sums
    uint32_t* blk;
    // per-frame scalars
    uint32_t* fs;            // [B][kFsCount]
    FrameDesc* frames;       // [B]
    CluConst* cc;            // [1]
    NoiseConst* nc;          // [1]
    lp_bounding_box* boxes;  // [B][box_stride]
    int32_t* box_large;      // [B][box_stride] polygons of more than 64 vertices (box stage)
    int32_t* vclass;         // [B][box_stride] vehicle class per cluster
    // label clouds (allocated when switched on): ground flags and the
    // ground-then-unknown index lists (frame stride N)
    uint32_t* gnd_flag;
    int32_t* split_idx;
    // KD-tree noise search (allocated when switched on, frame stride N; the
    // entry arrays have kKdEntriesPerPoint entries per point, the short-query
    // records kKdRecInts ints)
    uint32_t* wq;            // persistent-launch queue counters (kWqCounters)
    float4* kd;              // tree slots: x, y, z, input index bits
    int32_t* kd_tl;          // partition stop positions
    int32_t* kd_tr;
    uint2* kd_q;             // per query: slot of the k-th hit, pending depths
    uint32_t* kd_lcnt;       // leftover entries per query
    uint32_t* kd_loff;
    uint32_t* kd_dflag;      // 1: fewer than k hits in the own traversal
    uint32_t* kd_dpos;
    int32_t* kd_drec;        // short queries in order: index, hit count, hit slots
    uint64_t* kd_csr;        // leftover entries in query order
    uint64_t* kd_live;       // the stale stack while short queries are resolved
};

// Work-queue counters of the persistent launches (lp_queue.h): 8 per site.
constexpr int kWqCounters = 4 * 8;

constexpr bool kKdSearchDisabled = true;  // LP_NOISE_KDTREE rejected by lp_set_noise_search
constexpr int kKdEntriesPerPoint = 8;
constexpr uint32_t kKdMaxNeighbours = 7;
constexpr int kKdRecInts = 8;

// Per-frame scalar slots in DevBuf::fs.
enum FsSlot : int {
    kFsCands = 0,      // RANSAC candidates
    kFsApplied,        // RANSAC plane applied
    kFsBest,           // best inlier count
    kFsPlaneA, kFsPlaneB, kFsPlaneC, kFsPlaneD,
    kFsNumBig,         // big cells
    kFsQueue,          // JCP queue length
    kFsObstacles,
    kFsMaxR, kFsMaxA, kFsMaxE,
    kFsVoxels,
    kFsClusters,
    kFsHullTotal,
    kFsHullChunks,
    kFsHullInexact,    // predicates whose double evaluation had the wrong sign
    kFsHullRounded,    // clusters whose hull was redone with the reference's rounded predicate
    kFsCcGlobal,       // 1: too many voxels for the LDS union-find
    kFsErr,
    kFsBoxes,          // clusters with a box
    kFsBoxLarge,       // polygons listed for the workgroup box kernel
    kFsHullLarge,      // clusters listed for the workgroup local-rank launch
    kFsCellsUsed,      // non-empty polar cells
    kFsGround,         // GROUND points (label clouds on)
    kFsVehicles,       // clusters classified as vehicles
    // kept across the per-run clear: written by the noise filter before it
    kFsKept,           // points kept by the noise filter
    kFsNoise,          // points labelled NOISE
    kFsNoiseOpen,      // points the input-order pass left undecided
    kFsNoiseHard,      // points left to the neighbour-cell search
    kFsNoiseCells,     // occupied noise-filter cells
    kFsNoiseErr,       // error bits of the noise filter (kErrNoise*)
    kFsCount
};

struct Layout {
    int32_t B, N, P, H, W;  // capacity: frames, points/frame, pixels/frame
    int32_t B_active;       // frames in the current batch
    int32_t idx_bits;       // bits of the point index in a pixel winner key
    int32_t cell_stride;    // cells per frame (capacity)
    int32_t cand_stride;    // slices*4 capacity
    int32_t hash_cap;       // power of two
    int32_t blk_stride;     // scan block sums per frame
    int32_t box_stride;     // oriented boxes per frame (capacity)
};

}  // namespace lp

struct lp_context {
    int device = 0;
    hipStream_t stream = nullptr;
    lp_capacity cap{};
    lp::Layout L{};
    lp::DevBuf d{};
    void* arena = nullptr;
    size_t arena_bytes = 0;
    // host staging
    std::vector<lp::FrameDesc> frames;
    int32_t num_frames = 0;
    lp::CluConst cc{};
    int32_t staged_points = 0;
    // optional stages of the fused path
    bool noise_on = false;
    bool noise_kdtree = false;      // the reference object's KD-tree search instead of the radius rule
    bool persistent = false;        // per-point kernels as persistent (frame, tile) work-queue launches
    int32_t point_wg = 256;         // workgroup size of the per-point kernels (64, 128 or 256)
    int32_t mark_tile = 256;        // columns per workgroup of the JCP queue marking (64, 128 or 256)
    void* kd_mem = nullptr;         // its device buffers
    lp::NoiseConst nc{};
    bool boxes_on = false;
    bool clouds_on = false;
    bool vehicles_on = false;
    void* clouds_mem = nullptr;     // gnd_flag + split_idx
    uint8_t* image_dev = nullptr;   // one frame's label image (allocated on first use)
    bool staged_filtered = false;   // the staged batch went through the noise filter
    // pinned host buffers for inputs / outputs
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    // One captured graph per (batch size, optional-stage switches): every
    // kernel reads its frame descriptors from device memory and is launched
    // with capacity-sized grids, so a graph replays for any frames of that
    // batch size.
    struct GraphEntry {
        int32_t frames;
        int32_t options;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    // stage timing
    std::vector<hipEvent_t> events;
    std::vector<std::string> stage_names;
    bool profiling = false;
    std::vector<float> stage_ms_acc;
    int32_t profiled_runs = 0;
    // Further lanes of a multi-stream context (lp_capacity.streams > 1): each
    // is a context of its own (buffers, stream, graphs) that takes a
    // contiguous share of a staged batch; this context is lane 0 and its
    // stream waits for the others after every run.
    std::vector<lp_context*> lanes;
    std::vector<int32_t> lane_frames;   // frames staged per lane, lane 0 first
    std::vector<hipEvent_t> lane_done;  // one per further lane
};
