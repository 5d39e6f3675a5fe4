// lp_api.hip -- context, staging, launch sequence and the C ABI.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lp_kernels.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define LP_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) return fail(LP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr float kDegToRad = static_cast<float>(M_PI / 180.0);  // segmenter.hpp:135
constexpr float kTwoPi = static_cast<float>(2.0 * M_PI);        // segmenter.hpp:136

int ceil_log2(int64_t v) {
    int b = 0;
    while ((int64_t{1} << b) < v) ++b;
    return b;
}

}  // namespace

namespace lp {

int set_last_error(int code, const char* msg) { return fail(code, msg); }

void stage_mark(lp_context* c, const char* name) {
    if (!c->profiling) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, c->stream);
    c->events.push_back(e);
    c->stage_names.push_back(name);
}

// Derived segmenter constants, computed with host libm in float exactly as
// Segmenter::config / constructPolarGrid / RECM / RANSAC / JCP do.
static bool derive_seg(const lp_segmenter_config& g, const Layout& L, SegConst& s, std::string& why) {
    if (g.image_width <= 0 || g.image_height <= 0 || g.image_width > L.W || g.image_height > L.H) {
        why = "image size outside the context capacity";
        return false;
    }
    s.H = g.image_height;
    s.W = g.image_width;
    s.slice_res = g.grid_slice_resolution_deg * kDegToRad;
    s.rings = static_cast<int32_t>(g.max_distance_m / g.grid_radial_spacing_m);
    s.slices = static_cast<int32_t>(kTwoPi / s.slice_res);
    s.cells = s.rings * s.slices;
    if (s.rings < kRansacRadialBins || s.slices < 1) {
        why = "polar grid needs at least 4 radial rings and one slice";
        return false;
    }
    if (s.cells > L.cell_stride || s.slices * kRansacRadialBins > L.cand_stride ||
        s.slices >= (1 << (32 - L.idx_bits))) {
        why = "polar grid larger than the context capacity";
        return false;
    }
    s.organized = g.assume_unorganized_cloud ? 0 : 1;
    s.zlo = -g.sensor_height_m + g.z_min_m;
    s.zhi = -g.sensor_height_m + g.z_max_m;
    s.spacing = g.grid_radial_spacing_m;
    s.min_d = g.min_distance_m;
    s.max_d = g.max_distance_m;
    const float up = g.elevation_up_deg * kDegToRad;
    s.down_rad = g.elevation_down_deg * kDegToRad;
    s.per_pixel = (up - s.down_rad) / g.image_height;
    s.w_minus_1 = static_cast<float>(g.image_width - 1);
    s.thr = g.ground_height_threshold_m;
    s.two_thr = 2.0F * g.ground_height_threshold_m;
    s.dz = std::min(g.grid_radial_spacing_m * std::tan(g.road_maximum_slope_m_per_m),
                    g.ground_height_threshold_m - FLT_EPSILON);
    s.e0 = -g.sensor_height_m + g.ground_height_threshold_m;
    s.neg_h = -g.sensor_height_m;
    s.cos_max = std::cos(std::tan(g.road_maximum_slope_m_per_m));
    s.kthr2 = g.kernel_threshold_distance_m * g.kernel_threshold_distance_m;
    s.neg_amp = -g.amplification_factor;
    return true;
}

static void derive_clu(const lp_clusterer_config& g, int32_t max_voxels, CluConst& c) {
    c.res_r = g.voxel_grid_range_resolution_m;
    c.res_az = g.voxel_grid_azimuth_resolution_deg * kDegToRad;
    c.res_el = g.voxel_grid_elevation_resolution_deg * kDegToRad;
    c.min_size = g.min_cluster_size;
    c.max_voxels = max_voxels;
}

// De-interleave a raw AoS record buffer into the SoA inputs.
__global__ void k_deinterleave(const uint8_t* raw, int n, int stride, int ox, int oy, int oz, int oring,
                               float* x, float* y, float* z, uint16_t* ring) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint8_t* r = raw + static_cast<size_t>(i) * stride;
    float v;
    memcpy(&v, r + ox, 4);
    x[i] = v;
    memcpy(&v, r + oy, 4);
    y[i] = v;
    memcpy(&v, r + oz, 4);
    z[i] = v;
    if (oring >= 0) {
        uint16_t q;
        memcpy(&q, r + oring, 2);
        ring[i] = q;
    }
}

}  // namespace lp

using namespace lp;

namespace {

template <typename T>
T* carve(char*& p, size_t count) {
    T* r = reinterpret_cast<T*>(p);
    size_t bytes = (count * sizeof(T) + 255) & ~size_t{255};
    p += bytes;
    return r;
}

size_t layout_bytes(const Layout& L, DevBuf* d, char* base) {
    char* p = base;
    const size_t B = L.B, N = L.N, P = L.P, C = L.cell_stride, HC = L.hash_cap;
    auto go = [&](auto*& dst, size_t count) {
        using T = std::remove_reference_t<decltype(*dst)>;
        T* r = carve<T>(p, count);
        if (d) dst = r;
    };
    DevBuf tmp{};
    DevBuf& x = d ? *d : tmp;
    go(x.x, B * N);
    go(x.y, B * N);
    go(x.z, B * N);
    go(x.ring, B * N);
    go(x.rx, B * N);
    go(x.ry, B * N);
    go(x.rz, B * N);
    go(x.rring, B * N);
    go(x.noise, B * N);
    go(x.dsort, B * N);
    go(x.spos, B * N);
    go(x.nocc, B * kNoiseOccWords);
    go(x.ncell_slot, B * N);
    go(x.pt_cell, B * N);
    go(x.pt_xyzc, B * N);
    go(x.pt_pix, B * N);
    go(x.labels, B * N);
    go(x.obs_flag, B * N);
    go(x.obs_src, B * N);
    go(x.ox, B * N);
    go(x.oy, B * N);
    go(x.oz, B * N);
    go(x.sr, B * N);
    go(x.sa, B * N);
    go(x.se, B * N);
    go(x.vslot, B * N);
    go(x.cl_flag, B * N);
    go(x.clabel, B * N);
    x.nloc = x.clabel;
    go(x.cell_count, B * C);
    go(x.cell_start, B * (C + 1));
    go(x.pt_loc, B * N);
    go(x.cell_items, B * N);
    go(x.cell_zkey, B * N);
    x.hc_chunk_c = x.cell_items;
    go(x.hull_large, B * N);
    x.hc_chunk_b = reinterpret_cast<uint32_t*>(x.hull_large);
    x.hc_chunk_i = x.cell_zkey;
    go(x.zmin, B * C);
    go(x.elev, B * C);
    go(x.big_cells, B * C);
    go(x.cell_list, B * C);
    go(x.cand_count, B * L.cand_stride);
    go(x.cand_start, B * L.cand_stride);
    go(x.cx, B * N);
    go(x.cy, B * N);
    go(x.cz, B * N);
    go(x.hyp, B * kRansacIters * 4);
    go(x.hyp_valid, B * kRansacIters);
    go(x.hyp_count, B * kRansacIters);
    go(x.mt_raw, kMtRawCount);
    go(x.pix_key, B * P);
    go(x.state, B * P);
    go(x.mapping, B * P);
    go(x.pxyz, B * P);
    go(x.qflag, B * P);
    go(x.qmap, B * P);
    go(x.qlist, B * P);
    go(x.qrow, B * (L.H + 1));
    go(x.slot_w, B * P * 24);
    go(x.slot_src, B * P * 24);
    go(x.qsum_ok, B * P);
    go(x.fin, B * P);
    go(x.hslot, B * HC);
    go(x.vlist, B * N);
    go(x.vpar, B * N);
    go(x.vnb, B * N * 13);
    go(x.vmin, B * N);
    go(x.vcnt, B * N);
    go(x.vlab, B * N);
    go(x.cl_root, B * N);
    go(x.cl_size, B * N);
    go(x.cl_off, B * N);
    go(x.cl_cursor, B * N);
    x.hull_flag = x.cl_cursor;
    go(x.cl_members, B * N);
    x.nopen = x.cl_members;
    go(x.hcx, 2 * B * N);
    go(x.hcy, 2 * B * N);
    go(x.hci, 2 * B * N);
    go(x.hc_count, 2 * B * N);
    go(x.hc_choff, B * N);
    go(x.hull_idx, B * N);
    go(x.hull_count, B * N);
    go(x.hull_off, B * N);
    go(x.hull_local, B * N);
    go(x.blk, B * L.blk_stride);
    go(x.fs, B * kFsCount);
    go(x.wq, kWqCounters);
    go(x.frames, B);
    go(x.cc, 1);
    go(x.nc, 1);
    go(x.boxes, B * static_cast<size_t>(L.box_stride));
    go(x.box_large, B * static_cast<size_t>(L.box_stride));
    go(x.vclass, B * static_cast<size_t>(L.box_stride));
    return static_cast<size_t>(p - base);
}

// Field offsets of an AoS cloud relative to its lowest field.
struct AosLayout {
    const uint8_t* lo = nullptr;
    int fx = 0, fy = 0, fz = 0, fr = -1;
    int stride = 0;
    int span = 0;   // bytes from `lo` to the end of the last field of a record
};

bool aos_layout(const lp_cloud& cl, AosLayout& a, std::string& why) {
    const uint8_t* lo = static_cast<const uint8_t*>(cl.x);
    for (const void* q : {cl.y, cl.z, cl.ring})
        if (q && static_cast<const uint8_t*>(q) < lo) lo = static_cast<const uint8_t*>(q);
    auto field = [&](const void* q) { return static_cast<int64_t>(static_cast<const uint8_t*>(q) - lo); };
    const int64_t fx = field(cl.x), fy = field(cl.y), fz = field(cl.z), fr = cl.ring ? field(cl.ring) : -1;
    const int64_t stride = cl.xyz_stride;
    if (stride < 12 || fx + 4 > stride || fy + 4 > stride || fz + 4 > stride || (fr >= 0 && fr + 2 > stride) ||
        (cl.ring && cl.ring_stride != stride)) {
        why = "unsupported point layout: fields must share one record stride";
        return false;
    }
    a.lo = lo;
    a.fx = static_cast<int>(fx);
    a.fy = static_cast<int>(fy);
    a.fz = static_cast<int>(fz);
    a.fr = static_cast<int>(fr);
    a.stride = static_cast<int>(stride);
    a.span = static_cast<int>(std::max({fx + 4, fy + 4, fz + 4, fr >= 0 ? fr + 2 : int64_t{0}}));
    return true;
}

// Device bytes at slot_w, free while inputs are staged.
size_t raw_staging_bytes(const lp_context* c) {
    return static_cast<size_t>(c->L.B) * static_cast<size_t>(c->L.P) * 24 * sizeof(float);
}

}  // namespace

namespace lp {

// The per-frame checks lp_batch_stage makes, without staging anything.
int validate_frame(const lp_context* c, const lp_segmenter_config& cfg, const lp_cloud& cl) {
    if (cl.num_points < 0 || cl.num_points > c->L.N)
        return fail(LP_ERR_CAPACITY, "frame has more points than max_points_per_frame");
    if (cl.num_points > 0 && (!cl.x || !cl.y || !cl.z)) return fail(LP_ERR_INVALID_ARGUMENT, "null coordinates");
    SegConst sc{};
    std::string why;
    if (!derive_seg(cfg, c->L, sc, why)) return fail(LP_ERR_INVALID_ARGUMENT, why);
    const bool soa = cl.xyz_stride == 4 && (cl.ring == nullptr || cl.ring_stride == 2);
    if (cl.num_points > 0 && !soa) {
        AosLayout a;
        if (!aos_layout(cl, a, why)) return fail(LP_ERR_INVALID_ARGUMENT, why);
        if (raw_staging_bytes(c) < static_cast<size_t>(a.stride))
            return fail(LP_ERR_CAPACITY, "point record larger than the staging area");
    }
    return LP_OK;
}

}  // namespace lp

namespace {

// Copies a cloud into the packed inputs (x, y, z, ring at offset `off`), or
// into the noise filter's raw inputs when `raw_inputs`.
int stage_cloud(lp_context* c, const lp_cloud& cl, int32_t off, uint8_t* raw_dev, bool raw_inputs = false) {
    DevBuf d = c->d;
    if (raw_inputs) {
        d.x = d.rx;
        d.y = d.ry;
        d.z = d.rz;
        d.ring = d.rring;
    }
    const int n = cl.num_points;
    if (n == 0) return LP_OK;
    const hipMemcpyKind kind = cl.on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const bool soa = cl.xyz_stride == 4 && (cl.ring == nullptr || cl.ring_stride == 2);
    if (soa) {
        LP_HIP(hipMemcpyAsync(d.x + off, cl.x, n * 4, kind, c->stream));
        LP_HIP(hipMemcpyAsync(d.y + off, cl.y, n * 4, kind, c->stream));
        LP_HIP(hipMemcpyAsync(d.z + off, cl.z, n * 4, kind, c->stream));
        if (cl.ring) LP_HIP(hipMemcpyAsync(d.ring + off, cl.ring, n * 2, kind, c->stream));
        return LP_OK;
    }
    // AoS: all fields inside one record of xyz_stride bytes, staged through
    // the device area at raw_dev in chunks of whole records.
    AosLayout a;
    std::string why;
    if (!aos_layout(cl, a, why)) return fail(LP_ERR_INVALID_ARGUMENT, why);
    const size_t stride = static_cast<size_t>(a.stride);
    const size_t chunk = raw_staging_bytes(c) / stride;
    if (chunk == 0) return fail(LP_ERR_CAPACITY, "point record larger than the staging area");
    for (size_t first = 0; first < static_cast<size_t>(n); first += chunk) {
        const int cnt = static_cast<int>(std::min(chunk, static_cast<size_t>(n) - first));
        // The last record is read up to its last field only: the caller's
        // array may end there.
        const size_t bytes = (static_cast<size_t>(cnt) - 1) * stride + static_cast<size_t>(a.span);
        LP_HIP(hipMemcpyAsync(raw_dev, a.lo + first * stride, bytes, kind, c->stream));
        const int o = off + static_cast<int>(first);
        hipLaunchKernelGGL(k_deinterleave, dim3((cnt + 255) / 256), dim3(256), 0, c->stream, raw_dev, cnt, a.stride,
                           a.fx, a.fy, a.fz, a.fr, d.x + o, d.y + o, d.z + o, d.ring + o);
        LP_HIP(hipGetLastError());
    }
    return LP_OK;
}

int run_pipeline(lp_context* c, int B, bool segment_only) {
    c->L.B_active = B;
    if (c->staged_filtered && !segment_only) {
        if (c->noise_kdtree)
            launch_noise_kdtree(c, B, true);
        else
            launch_noise(c, B, true);
    }
    launch_segment_front(c, B);
    launch_jcp(c, B);
    if (!segment_only) {
        launch_cluster(c, B, true);
        launch_hull(c, B);
        if (c->boxes_on || c->vehicles_on) launch_boxes(c, B);
        if (c->vehicles_on) launch_vehicles(c, B);
    }
    LP_HIP(hipGetLastError());
    return LP_OK;
}

// NoiseRemover::filter constants (noise_remover.cpp:44-48).
NoiseConst noise_const(const lp_noise_config& cfg) {
    NoiseConst nc{};
    nc.scaling = std::pow(static_cast<double>(cfg.radius_multiplier_m_per_m), 2.0);
    nc.min_sqr = cfg.min_search_radius_m * cfg.min_search_radius_m;
    nc.k = cfg.min_neighbours;
    return nc;
}

int status_of(uint32_t err) {
    if (err & kErrVoxelOverflow) return fail(LP_ERR_VOXEL_OVERFLOW, "more occupied voxels than max_voxels");
    if (err & kErrRansacSingle)
        return fail(LP_ERR_RANSAC_SINGLE_CANDIDATE, "exactly one RANSAC candidate (reference never terminates)");
    if (err & kErrNoiseInput)
        return fail(LP_ERR_INVALID_ARGUMENT,
                    "non-finite coordinates: the KD-tree search (std::nth_element) has no order for them");
    if (err & kErrNoiseCapacity)
        return fail(LP_ERR_CAPACITY, "KD-tree noise search: more pending search entries than the device holds");
    if (err & kErrHullFuzzyOrder)
        return fail(LP_ERR_INVALID_ARGUMENT,
                    "hull points with distinct x closer than DBL_EPSILON: the reference's sort comparator "
                    "(polygonizer.cpp:52-59) is not an ordering there");
    if (err & (kErrCellCapacity | kErrQueueCapacity | kErrHullCapacity | kErrRngExhausted | kErrBoxCapacity))
        return fail(LP_ERR_CAPACITY, "a device capacity was exceeded (cell, JCP queue, hull, box or RNG stream)");
    return LP_OK;
}

}  // namespace

extern "C" {

void lp_segmenter_config_default(lp_segmenter_config* c) {
    c->elevation_up_deg = 2.0F;
    c->elevation_down_deg = -24.8F;
    c->image_width = 2048;
    c->image_height = 64;
    c->assume_unorganized_cloud = 0;
    c->grid_radial_spacing_m = 2.0F;
    c->grid_slice_resolution_deg = 1.0F;
    c->ground_height_threshold_m = 0.2F;
    c->road_maximum_slope_m_per_m = 0.2F;
    c->min_distance_m = 2.0F;
    c->max_distance_m = 100.0F;
    c->sensor_height_m = 1.73F;
    c->kernel_threshold_distance_m = 1.0F;
    c->amplification_factor = 5.0F;
    c->z_min_m = -3.0F;
    c->z_max_m = 4.0F;
}

void lp_clusterer_config_default(lp_clusterer_config* c) {
    c->voxel_grid_range_resolution_m = 0.4F;
    c->voxel_grid_azimuth_resolution_deg = 1.0F;
    c->voxel_grid_elevation_resolution_deg = 1.5F;
    c->min_cluster_size = 3;
}

void lp_clusterer_config_shipped(lp_clusterer_config* c) {
    lp_clusterer_config_default(c);
    c->voxel_grid_elevation_resolution_deg = 3.0F;
}

void lp_capacity_default(lp_capacity* cap) {
    cap->max_frames = 1;
    cap->max_points_per_frame = 200000;  // Segmenter::MAX_CLOUD_SIZE, segmenter.hpp:151
    cap->max_image_height = 64;
    cap->max_image_width = 2048;
    cap->max_voxels = 200000;            // clusterer.cpp:36
    cap->device = 0;
    cap->streams = 1;
    cap->hash_load_factor = 0.5F;
}

const char* lp_last_error(void) { return g_last_error.c_str(); }

namespace {

int create_lane(const lp_capacity* cap, lp_context** out);

}  // namespace

int lp_context_create(const lp_capacity* cap, lp_context** out) {
    if (!cap || !out) return fail(LP_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (cap->streams < 1 || cap->streams > 8) return fail(LP_ERR_INVALID_ARGUMENT, "streams must be 1..8");
    const int S = std::min(cap->streams, std::max(cap->max_frames, 1));
    lp_capacity lc = *cap;
    lc.streams = 1;
    lc.max_frames = (cap->max_frames + S - 1) / S;
    lp_context* c = nullptr;
    int st = create_lane(&lc, &c);
    if (st != LP_OK) return st;
    for (int k = 1; k < S; ++k) {
        lp_context* lane = nullptr;
        st = create_lane(&lc, &lane);
        hipEvent_t ev = nullptr;
        if (st == LP_OK) {
            const hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
            if (e != hipSuccess) {
                lp_context_destroy(lane);
                st = fail(LP_ERR_HIP, std::string("lane event: ") + hipGetErrorString(e));
            }
        }
        if (st != LP_OK) {
            lp_context_destroy(c);
            return st;
        }
        c->lanes.push_back(lane);
        c->lane_done.push_back(ev);
    }
    *out = c;
    return LP_OK;
}

namespace {

// One lane: buffers, stream and graph cache for cap->max_frames frames.
int create_lane(const lp_capacity* cap, lp_context** out) {
    *out = nullptr;
    if (cap->max_frames < 1 || cap->max_points_per_frame < 1 || cap->max_image_height < 1 ||
        cap->max_image_width < 1 || cap->max_image_height > 1024 || cap->max_voxels < 1)
        return fail(LP_ERR_INVALID_ARGUMENT, "invalid capacity");
    if (!(cap->hash_load_factor >= 0.5F && cap->hash_load_factor <= 0.9F))
        return fail(LP_ERR_INVALID_ARGUMENT, "hash_load_factor must be in [0.5, 0.9]");
    int ndev = 0;
    LP_HIP(hipGetDeviceCount(&ndev));
    if (cap->device < 0 || cap->device >= ndev) return fail(LP_ERR_INVALID_ARGUMENT, "no such HIP device");
    LP_HIP(hipSetDevice(cap->device));
    auto* c = new lp_context();
    c->device = cap->device;
    c->cap = *cap;
    Layout& L = c->L;
    L.B = cap->max_frames;
    L.N = cap->max_points_per_frame;
    L.H = cap->max_image_height;
    L.W = cap->max_image_width;
    L.P = L.H * L.W;
    L.B_active = 0;
    L.idx_bits = std::max(1, ceil_log2(L.N));
    if (L.idx_bits > 24) {
        delete c;
        return fail(LP_ERR_INVALID_ARGUMENT, "max_points_per_frame above 2^24");
    }
    L.cell_stride = 32768;
    L.cand_stride = 4 * 4096;
    // Every point could occupy its own cell or voxel: the table holds N keys
    // at the chosen worst-case load.
    L.hash_cap = 1 << ceil_log2(static_cast<int64_t>(std::ceil(L.N / static_cast<double>(cap->hash_load_factor))));
    L.blk_stride = (std::max({L.N, L.P, L.hash_cap}) + 2047) / 2048 + 1;
    L.box_stride = L.N / 3 + 1024;
    const size_t bytes = layout_bytes(L, nullptr, nullptr);
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&c->arena, bytes);
    if (e != hipSuccess) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(LP_ERR_HIP, std::string("context allocation: ") + hipGetErrorString(e));
    }
    c->arena_bytes = bytes;
    layout_bytes(L, &c->d, static_cast<char*>(c->arena));
    launch_mt19937(c->d.mt_raw, kMtRawCount, c->stream);
    launch_hash_init(c);
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        lp_context_destroy(c);
        return fail(LP_ERR_HIP, std::string("context init: ") + hipGetErrorString(e));
    }
    *out = c;
    return LP_OK;
}

}  // namespace

int lp_context_destroy(lp_context* c) {
    if (!c) return LP_OK;
    for (auto* lane : c->lanes) lp_context_destroy(lane);
    for (auto e : c->lane_done) (void)hipEventDestroy(e);
    (void)hipSetDevice(c->device);
    for (auto& g : c->graphs) {
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphDestroy(g.graph);
    }
    for (auto e : c->events) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->arena) (void)hipFree(c->arena);
    if (c->clouds_mem) (void)hipFree(c->clouds_mem);
    if (c->kd_mem) (void)hipFree(c->kd_mem);
    if (c->image_dev) (void)hipFree(c->image_dev);
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return LP_OK;
}

void* lp_context_stream(lp_context* c) { return c ? c->stream : nullptr; }

namespace {

int stage_batch(lp_context* c, const lp_segmenter_config* seg_cfgs, int32_t num_seg_cfgs,
                const lp_clusterer_config* ccfg, int32_t num_frames, const lp_cloud* clouds, bool filter) {
    if (!c || !seg_cfgs || !ccfg || !clouds) return fail(LP_ERR_INVALID_ARGUMENT, "null argument");
    if (num_frames < 1 || num_frames > c->L.B) return fail(LP_ERR_CAPACITY, "batch larger than max_frames");
    if (num_seg_cfgs != 1 && num_seg_cfgs != num_frames)
        return fail(LP_ERR_INVALID_ARGUMENT, "num_seg_cfgs must be 1 or num_frames");
    LP_HIP(hipSetDevice(c->device));
    c->frames.assign(num_frames, FrameDesc{});
    int32_t off = 0;
    for (int f = 0; f < num_frames; ++f) {
        const lp_cloud& cl = clouds[f];
        if (cl.num_points < 0 || cl.num_points > c->L.N)
            return fail(LP_ERR_CAPACITY, "frame has more points than max_points_per_frame");
        if (cl.num_points > 0 && (!cl.x || !cl.y || !cl.z)) return fail(LP_ERR_INVALID_ARGUMENT, "null coordinates");
        FrameDesc& fd = c->frames[f];
        std::string why;
        if (!derive_seg(seg_cfgs[num_seg_cfgs == 1 ? 0 : f], c->L, fd.sc, why))
            return fail(LP_ERR_INVALID_ARGUMENT, why);
        fd.pt_off = off;
        fd.n = cl.num_points;
        fd.n_raw = cl.num_points;
        fd.pix_off = f * c->L.P;
        fd.has_ring = cl.ring != nullptr ? 1 : 0;
        off += cl.num_points;
    }
    // Inputs: SoA copies straight in; AoS records through a device staging
    // area (the slot-weight arrays are free at this point of the pipeline).
    // With the noise filter on they land in its raw inputs and the filter
    // writes the kept points where the segmenter reads them.
    uint8_t* raw = reinterpret_cast<uint8_t*>(c->d.slot_w);
    for (int f = 0; f < num_frames; ++f) {
        const int st = stage_cloud(c, clouds[f], c->frames[f].pt_off, raw, filter);
        if (st != LP_OK) return st;
    }
    derive_clu(*ccfg, c->cap.max_voxels, c->cc);
    if (filter) LP_HIP(hipMemcpyAsync(c->d.nc, &c->nc, sizeof(NoiseConst), hipMemcpyHostToDevice, c->stream));
    c->staged_filtered = filter;
    LP_HIP(hipMemcpyAsync(c->d.frames, c->frames.data(), sizeof(FrameDesc) * num_frames, hipMemcpyHostToDevice,
                          c->stream));
    LP_HIP(hipMemcpyAsync(c->d.cc, &c->cc, sizeof(CluConst), hipMemcpyHostToDevice, c->stream));
    c->num_frames = num_frames;
    c->staged_points = off;
    return LP_OK;
}

}  // namespace

int lp_batch_stage(lp_context* c, const lp_segmenter_config* seg_cfgs, int32_t num_seg_cfgs,
                   const lp_clusterer_config* ccfg, int32_t num_frames, const lp_cloud* clouds) {
    if (!c) return fail(LP_ERR_INVALID_ARGUMENT, "null context");
    if (c->lanes.empty()) return stage_batch(c, seg_cfgs, num_seg_cfgs, ccfg, num_frames, clouds, c->noise_on);
    const int S = 1 + static_cast<int>(c->lanes.size());
    if (!seg_cfgs || !clouds) return fail(LP_ERR_INVALID_ARGUMENT, "null argument");
    if (num_frames < 1 || num_frames > c->L.B * S) return fail(LP_ERR_CAPACITY, "batch larger than max_frames");
    if (num_seg_cfgs != 1 && num_seg_cfgs != num_frames)
        return fail(LP_ERR_INVALID_ARGUMENT, "num_seg_cfgs must be 1 or num_frames");
    // Contiguous shares, as even as possible, lane 0 first.
    const int g = (num_frames + S - 1) / S;
    c->lane_frames.assign(S, 0);
    for (int k = 0; k < S; ++k) {
        lp_context* lane = k == 0 ? c : c->lanes[k - 1];
        const int lo = k * g;
        const int cnt = std::max(0, std::min(g, num_frames - lo));
        if (cnt == 0) {
            lane->num_frames = 0;
            continue;
        }
        const int st = stage_batch(lane, num_seg_cfgs == 1 ? seg_cfgs : seg_cfgs + lo, num_seg_cfgs == 1 ? 1 : cnt,
                                   ccfg, cnt, clouds + lo, lane->noise_on);
        if (st != LP_OK) return st;
        c->lane_frames[k] = cnt;
    }
    return LP_OK;
}

namespace {

int run_lane(lp_context* c, int32_t use_graph) {
    LP_HIP(hipSetDevice(c->device));
    const int B = c->num_frames;
    if (!use_graph) return run_pipeline(c, B, false);
    const int options = (c->staged_filtered ? 1 : 0) | (c->boxes_on ? 2 : 0) | (c->clouds_on ? 4 : 0) |
                        (c->vehicles_on ? 8 : 0) | (c->staged_filtered && c->noise_kdtree ? 16 : 0) |
                        (c->persistent ? 32 : 0) | (c->point_wg << 8) | (c->mark_tile << 20);
    lp_context::GraphEntry* ge = nullptr;
    for (auto& g : c->graphs)
        if (g.frames == B && g.options == options) ge = &g;
    if (!ge) {
        LP_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int st = run_pipeline(c, B, false);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (st != LP_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return fail(LP_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        hipGraphExec_t ex = nullptr;
        const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            (void)hipGraphDestroy(g);
            return fail(LP_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
        }
        c->graphs.push_back({B, options, g, ex});
        ge = &c->graphs.back();
    }
    c->L.B_active = B;
    LP_HIP(hipGraphLaunch(ge->exec, c->stream));
    return LP_OK;
}

}  // namespace

int lp_batch_run(lp_context* c, int32_t use_graph) {
    if (!c || c->num_frames < 1) return fail(LP_ERR_INVALID_ARGUMENT, "nothing staged");
    int st = run_lane(c, use_graph);
    for (size_t k = 0; st == LP_OK && k < c->lanes.size(); ++k) {
        lp_context* lane = c->lanes[k];
        if (lane->num_frames < 1) continue;
        st = run_lane(lane, use_graph);
        if (st != LP_OK) break;
        LP_HIP(hipEventRecord(c->lane_done[k], lane->stream));
        LP_HIP(hipStreamWaitEvent(c->stream, c->lane_done[k], 0));
    }
    return st;
}

int lp_synchronize(lp_context* c) {
    if (!c) return fail(LP_ERR_INVALID_ARGUMENT, "null context");
    LP_HIP(hipStreamSynchronize(c->stream));
    return LP_OK;
}

int lp_query(lp_context* c) {
    if (!c) return fail(LP_ERR_INVALID_ARGUMENT, "null context");
    const hipError_t e = hipStreamQuery(c->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    LP_HIP(e);
    return LP_OK;
}

namespace {

int fetch_lane(lp_context* c, const lp_frame_outputs* outputs, lp_frame_result* results) {
    LP_HIP(hipSetDevice(c->device));
    const int B = c->num_frames;
    const Layout& L = c->L;
    std::vector<uint32_t> fs(static_cast<size_t>(B) * kFsCount);
    LP_HIP(hipMemcpyAsync(fs.data(), c->d.fs, fs.size() * 4, hipMemcpyDeviceToHost, c->stream));
    LP_HIP(hipStreamSynchronize(c->stream));
    int overall = LP_OK;
    for (int f = 0; f < B; ++f) {
        const uint32_t* s = fs.data() + static_cast<size_t>(f) * kFsCount;
        const FrameDesc& fd = c->frames[f];
        lp_frame_result r{};
        r.status = status_of(s[kFsErr] | s[kFsNoiseErr]);
        // Points the segmenter saw: the noise filter's kept count when it ran.
        const int n = c->staged_filtered ? static_cast<int>(s[kFsKept]) : fd.n;
        r.num_points = n;
        r.num_noise = c->staged_filtered ? static_cast<int32_t>(s[kFsNoise]) : 0;
        r.num_noise_searched = c->staged_filtered ? static_cast<int32_t>(s[kFsNoiseHard]) : 0;
        const bool boxes = c->boxes_on || c->vehicles_on;
        r.num_boxes = boxes ? static_cast<int32_t>(s[kFsBoxes]) : 0;
        r.num_obstacles = static_cast<int32_t>(s[kFsObstacles]);
        r.num_voxels = static_cast<int32_t>(s[kFsVoxels]);
        r.num_clusters = static_cast<int32_t>(s[kFsClusters]);
        r.num_hull_vertices = static_cast<int32_t>(s[kFsHullTotal]);
        r.jcp_queue = static_cast<int32_t>(s[kFsQueue]);
        if (c->clouds_on) {
            r.num_ground = static_cast<int32_t>(s[kFsGround]);
            r.num_unknown = n - r.num_ground - r.num_obstacles;
        }
        r.num_vehicles = c->vehicles_on ? static_cast<int32_t>(s[kFsVehicles]) : 0;
        r.num_hull_inexact = static_cast<int32_t>(s[kFsHullInexact]);
        r.num_hull_rounded = static_cast<int32_t>(s[kFsHullRounded]);
        r.ransac.candidates = s[kFsCands];
        r.ransac.best_inliers = s[kFsBest];
        r.ransac.applied = static_cast<int32_t>(s[kFsApplied]);
        uint32_t pl[4] = {s[kFsPlaneA], s[kFsPlaneB], s[kFsPlaneC], s[kFsPlaneD]};
        if (!r.ransac.applied) {
            const float dflt[4] = {0.0F, 0.0F, 1.0F, 0.0F};
            std::memcpy(pl, dflt, sizeof(pl));
        }
        std::memcpy(r.ransac.plane, pl, sizeof(pl));
        if (results) results[f] = r;
        if (r.status != LP_OK && overall == LP_OK) overall = r.status;
        if (!outputs) continue;
        const lp_frame_outputs& o = outputs[f];
        const size_t ob = static_cast<size_t>(f) * L.N;
        if (c->staged_filtered && o.noise_labels && fd.n_raw)
            LP_HIP(hipMemcpyAsync(o.noise_labels, c->d.noise + ob, fd.n_raw, hipMemcpyDeviceToHost, c->stream));
        if (o.seg_labels && n)
            LP_HIP(hipMemcpyAsync(o.seg_labels, c->d.labels + fd.pt_off, n * 4, hipMemcpyDeviceToHost, c->stream));
        if (r.status != LP_OK) continue;
        const int no = r.num_obstacles, K = r.num_clusters, V = r.num_hull_vertices;
        if (o.obstacle_index && no)
            LP_HIP(hipMemcpyAsync(o.obstacle_index, c->d.obs_src + ob, no * 4, hipMemcpyDeviceToHost, c->stream));
        if (o.cluster_labels && no)
            LP_HIP(hipMemcpyAsync(o.cluster_labels, c->d.clabel + ob, no * 4, hipMemcpyDeviceToHost, c->stream));
        if (o.hull_offsets) {
            if (K) LP_HIP(hipMemcpyAsync(o.hull_offsets, c->d.hull_off + ob, K * 4, hipMemcpyDeviceToHost, c->stream));
            o.hull_offsets[K] = V;
        }
        if (o.hull_indices && V) {
            if (V > o.hull_capacity) {
                if (results) results[f].status = fail(LP_ERR_CAPACITY, "hull_capacity too small");
                overall = LP_ERR_CAPACITY;
                continue;
            }
            LP_HIP(hipMemcpyAsync(o.hull_indices, c->d.hull_local + ob, V * 4, hipMemcpyDeviceToHost, c->stream));
        }
        if (boxes && o.boxes && K)
            LP_HIP(hipMemcpyAsync(o.boxes, c->d.boxes + static_cast<size_t>(f) * L.box_stride,
                                  K * sizeof(lp_bounding_box), hipMemcpyDeviceToHost, c->stream));
        if (c->vehicles_on && o.vehicle_class && K)
            LP_HIP(hipMemcpyAsync(o.vehicle_class, c->d.vclass + static_cast<size_t>(f) * L.box_stride,
                                  K * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        if (c->clouds_on && o.ground_index && r.num_ground)
            LP_HIP(hipMemcpyAsync(o.ground_index, c->d.split_idx + ob, r.num_ground * 4, hipMemcpyDeviceToHost,
                                  c->stream));
        if (c->clouds_on && o.unknown_index && r.num_unknown)
            LP_HIP(hipMemcpyAsync(o.unknown_index, c->d.split_idx + ob + r.num_ground, r.num_unknown * 4,
                                  hipMemcpyDeviceToHost, c->stream));
    }
    LP_HIP(hipStreamSynchronize(c->stream));
    return overall;
}

}  // namespace

int lp_batch_fetch(lp_context* c, const lp_frame_outputs* outputs, lp_frame_result* results) {
    if (!c || c->num_frames < 1) return fail(LP_ERR_INVALID_ARGUMENT, "nothing staged");
    if (c->lanes.empty()) return fetch_lane(c, outputs, results);
    int overall = LP_OK, off = 0;
    for (size_t k = 0; k < c->lane_frames.size(); ++k) {
        const int cnt = c->lane_frames[k];
        if (cnt == 0) continue;
        lp_context* lane = k == 0 ? c : c->lanes[k - 1];
        const int st = fetch_lane(lane, outputs ? outputs + off : nullptr, results ? results + off : nullptr);
        if (st != LP_OK && overall == LP_OK) overall = st;
        off += cnt;
    }
    return overall;
}

int lp_process_frames(lp_context* c, const lp_segmenter_config* seg_cfgs, int32_t num_seg_cfgs,
                      const lp_clusterer_config* ccfg, int32_t num_frames, const lp_cloud* clouds,
                      const lp_frame_outputs* outputs, lp_frame_result* results) {
    int st = lp_batch_stage(c, seg_cfgs, num_seg_cfgs, ccfg, num_frames, clouds);
    if (st != LP_OK) return st;
    st = lp_batch_run(c, 0);
    if (st != LP_OK) return st;
    return lp_batch_fetch(c, outputs, results);
}

int lp_segment(lp_context* c, const lp_segmenter_config* cfg, const lp_cloud* cloud, uint32_t* labels,
               lp_ransac_result* ransac) {
    if (!c || !cfg || !cloud || (!labels && cloud->num_points > 0))
        return fail(LP_ERR_INVALID_ARGUMENT, "null argument");
    lp_clusterer_config cc;
    lp_clusterer_config_shipped(&cc);
    int st = stage_batch(c, cfg, 1, &cc, 1, cloud, false);
    if (st != LP_OK) return st;
    st = run_pipeline(c, 1, true);
    if (st != LP_OK) return st;
    uint32_t fs[kFsCount];
    LP_HIP(hipMemcpyAsync(fs, c->d.fs, sizeof(fs), hipMemcpyDeviceToHost, c->stream));
    if (cloud->num_points)
        LP_HIP(hipMemcpyAsync(labels, c->d.labels, cloud->num_points * 4, hipMemcpyDeviceToHost, c->stream));
    LP_HIP(hipStreamSynchronize(c->stream));
    st = status_of(fs[kFsErr] & ~(kErrVoxelOverflow | kErrHullCapacity));
    if (ransac) {
        ransac->candidates = fs[kFsCands];
        ransac->best_inliers = fs[kFsBest];
        ransac->applied = static_cast<int32_t>(fs[kFsApplied]);
        if (ransac->applied) {
            std::memcpy(ransac->plane, &fs[kFsPlaneA], 16);
        } else {
            const float dflt[4] = {0.0F, 0.0F, 1.0F, 0.0F};
            std::memcpy(ransac->plane, dflt, 16);
        }
    }
    return st;
}

int lp_cluster(lp_context* c, const lp_clusterer_config* cfg, const lp_cloud* cloud, int32_t* labels) {
    if (!c || !cfg || !cloud || (!labels && cloud->num_points > 0))
        return fail(LP_ERR_INVALID_ARGUMENT, "null argument");
    if (cloud->num_points < 0 || cloud->num_points > c->L.N)
        return fail(LP_ERR_CAPACITY, "cloud has more points than max_points_per_frame");
    LP_HIP(hipSetDevice(c->device));
    FrameDesc fd{};
    fd.pt_off = 0;
    fd.n = cloud->num_points;
    fd.n_raw = cloud->num_points;
    fd.pix_off = 0;
    fd.has_ring = 0;  // no polar grid, no image: the clear stage only resets the voxel state
    c->frames.assign(1, fd);
    lp_cloud xyz = *cloud;
    xyz.ring = nullptr;
    int st = stage_cloud(c, xyz, 0, reinterpret_cast<uint8_t*>(c->d.slot_w));
    if (st != LP_OK) return st;
    derive_clu(*cfg, c->cap.max_voxels, c->cc);
    LP_HIP(hipMemcpyAsync(c->d.frames, c->frames.data(), sizeof(FrameDesc), hipMemcpyHostToDevice, c->stream));
    LP_HIP(hipMemcpyAsync(c->d.cc, &c->cc, sizeof(CluConst), hipMemcpyHostToDevice, c->stream));
    c->num_frames = 1;
    c->L.B_active = 1;
    launch_clear(c, 1);
    launch_cluster(c, 1, false);
    LP_HIP(hipGetLastError());
    uint32_t fs[kFsCount];
    LP_HIP(hipMemcpyAsync(fs, c->d.fs, sizeof(fs), hipMemcpyDeviceToHost, c->stream));
    if (cloud->num_points)
        LP_HIP(hipMemcpyAsync(labels, c->d.clabel, cloud->num_points * 4, hipMemcpyDeviceToHost, c->stream));
    LP_HIP(hipStreamSynchronize(c->stream));
    c->num_frames = 0;  // the staged frame is not a full-pipeline batch
    return status_of(fs[kFsErr]);
}

void lp_noise_config_default(lp_noise_config* cfg) {
    cfg->radius_multiplier_m_per_m = 0.02F;
    cfg->min_search_radius_m = 0.1F;
    cfg->min_neighbours = 4U;
}

int lp_set_noise_filter(lp_context* c, const lp_noise_config* cfg) {
    if (!c) return fail(LP_ERR_INVALID_ARGUMENT, "null context");
    if (!cfg) {
        c->noise_on = false;
        for (auto* lane : c->lanes) lane->noise_on = false;
        return LP_OK;
    }
    if (c->noise_kdtree && cfg->min_neighbours > kKdMaxNeighbours)
        return fail(LP_ERR_INVALID_ARGUMENT, "the KD-tree noise search supports min_neighbours up to 7");
    c->nc = noise_const(*cfg);
    c->noise_on = true;
    for (auto* lane : c->lanes) lp_set_noise_filter(lane, cfg);
    return LP_OK;
}

int lp_set_persistent(lp_context* c, int32_t enable) {
    if (!c) return fail(LP_ERR_INVALID_ARGUMENT, "null context");
    c->persistent = enable != 0;
    for (auto* lane : c->lanes) lane->persistent = c->persistent;
    return LP_OK;
}

int lp_set_launch_shape(lp_context* c, int32_t point_workgroup, int32_t mark_tile_width) {
    if (!c) return fail(LP_ERR_INVALID_ARGUMENT, "null context");
    auto ok = [](int32_t v) { return v == 64 || v == 128 || v == 256; };
    if (!ok(point_workgroup) || !ok(mark_tile_width))
        return fail(LP_ERR_INVALID_ARGUMENT, "workgroup size and tile width must be 64, 128 or 256");
    c->point_wg = point_workgroup;
    c->mark_tile = mark_tile_width;
    for (auto* lane : c->lanes) {
        lane->point_wg = point_workgroup;
        lane->mark_tile = mark_tile_width;
    }
    return LP_OK;
}

int lp_set_noise_search(lp_context* c, int32_t mode) {
    if (!c) return fail(LP_ERR_INVALID_ARGUMENT, "null context");
    if (mode != LP_NOISE_RULE && mode != LP_NOISE_KDTREE) return fail(LP_ERR_INVALID_ARGUMENT, "unknown noise search");
    // The KD-tree kernels (lp_noise_ref.hip) stay compiled but unreachable:
    // their first device run left the GPU faulted, and no cause has been
    // shown yet.
    if (mode == LP_NOISE_KDTREE && kKdSearchDisabled)
        return fail(LP_ERR_INVALID_ARGUMENT, "the KD-tree noise search is switched off in this build");
    if (mode == LP_NOISE_KDTREE) {
        if (c->noise_on && c->nc.k > static_cast<uint32_t>(kKdMaxNeighbours))
            return fail(LP_ERR_INVALID_ARGUMENT, "the KD-tree noise search supports min_neighbours up to 7");
        if (!c->kd_mem) {
            LP_HIP(hipSetDevice(c->device));
            const size_t cnt = static_cast<size_t>(c->L.B) * c->L.N;
            const size_t ent = cnt * kKdEntriesPerPoint;
            const size_t bytes = cnt * (sizeof(float4) + 2 * sizeof(int32_t) + sizeof(uint2) + 4 * sizeof(uint32_t) +
                                        kKdRecInts * sizeof(int32_t)) + 2 * ent * sizeof(uint64_t);
            LP_HIP(hipMalloc(&c->kd_mem, bytes));
            char* p = static_cast<char*>(c->kd_mem);
            auto take = [&](size_t b) { char* r = p; p += b; return r; };
            c->d.kd_csr = reinterpret_cast<uint64_t*>(take(ent * sizeof(uint64_t)));
            c->d.kd_live = reinterpret_cast<uint64_t*>(take(ent * sizeof(uint64_t)));
            c->d.kd = reinterpret_cast<float4*>(take(cnt * sizeof(float4)));
            c->d.kd_q = reinterpret_cast<uint2*>(take(cnt * sizeof(uint2)));
            c->d.kd_tl = reinterpret_cast<int32_t*>(take(cnt * sizeof(int32_t)));
            c->d.kd_tr = reinterpret_cast<int32_t*>(take(cnt * sizeof(int32_t)));
            c->d.kd_lcnt = reinterpret_cast<uint32_t*>(take(cnt * sizeof(uint32_t)));
            c->d.kd_loff = reinterpret_cast<uint32_t*>(take(cnt * sizeof(uint32_t)));
            c->d.kd_dflag = reinterpret_cast<uint32_t*>(take(cnt * sizeof(uint32_t)));
            c->d.kd_dpos = reinterpret_cast<uint32_t*>(take(cnt * sizeof(uint32_t)));
            c->d.kd_drec = reinterpret_cast<int32_t*>(take(cnt * kKdRecInts * sizeof(int32_t)));
        }
    }
    c->noise_kdtree = mode == LP_NOISE_KDTREE;
    for (auto* lane : c->lanes) {
        const int st = lp_set_noise_search(lane, mode);
        if (st != LP_OK) return st;
    }
    return LP_OK;
}

int lp_set_oriented_boxes(lp_context* c, int32_t enable) {
    if (!c) return fail(LP_ERR_INVALID_ARGUMENT, "null context");
    c->boxes_on = enable != 0;
    for (auto* lane : c->lanes) lane->boxes_on = c->boxes_on;
    return LP_OK;
}

int lp_set_label_clouds(lp_context* c, int32_t enable) {
    if (!c) return fail(LP_ERR_INVALID_ARGUMENT, "null context");
    if (enable && !c->clouds_mem) {
        LP_HIP(hipSetDevice(c->device));
        const size_t count = static_cast<size_t>(c->L.B) * c->L.N;
        LP_HIP(hipMalloc(&c->clouds_mem, 2 * count * sizeof(uint32_t)));
        c->d.gnd_flag = static_cast<uint32_t*>(c->clouds_mem);
        c->d.split_idx = reinterpret_cast<int32_t*>(c->d.gnd_flag + count);
    }
    c->clouds_on = enable != 0;
    for (auto* lane : c->lanes) {
        const int st = lp_set_label_clouds(lane, enable);
        if (st != LP_OK) return st;
    }
    return LP_OK;
}

int lp_set_vehicle_boxes(lp_context* c, int32_t enable) {
    if (!c) return fail(LP_ERR_INVALID_ARGUMENT, "null context");
    c->vehicles_on = enable != 0;
    for (auto* lane : c->lanes) lane->vehicles_on = c->vehicles_on;
    return LP_OK;
}

int lp_remove_noise(lp_context* c, const lp_noise_config* cfg, const lp_cloud* cloud, uint8_t* labels) {
    if (!c || !cfg || !cloud || (!labels && cloud->num_points > 0))
        return fail(LP_ERR_INVALID_ARGUMENT, "null argument");
    if (cloud->num_points < 0 || cloud->num_points > c->L.N)
        return fail(LP_ERR_CAPACITY, "cloud has more points than max_points_per_frame");
    if (cloud->num_points > 0 && (!cloud->x || !cloud->y || !cloud->z))
        return fail(LP_ERR_INVALID_ARGUMENT, "null coordinates");
    if (c->noise_kdtree && cfg->min_neighbours > kKdMaxNeighbours)
        return fail(LP_ERR_INVALID_ARGUMENT, "the KD-tree noise search supports min_neighbours up to 7");
    if (cloud->num_points == 0) return LP_OK;
    LP_HIP(hipSetDevice(c->device));
    FrameDesc fd{};
    fd.n = cloud->num_points;
    fd.n_raw = cloud->num_points;
    c->frames.assign(1, fd);
    lp_cloud xyz = *cloud;
    xyz.ring = nullptr;
    int st = stage_cloud(c, xyz, 0, reinterpret_cast<uint8_t*>(c->d.slot_w), true);
    if (st != LP_OK) return st;
    const NoiseConst nc = noise_const(*cfg);
    LP_HIP(hipMemcpyAsync(c->d.frames, c->frames.data(), sizeof(FrameDesc), hipMemcpyHostToDevice, c->stream));
    LP_HIP(hipMemcpyAsync(c->d.nc, &nc, sizeof(NoiseConst), hipMemcpyHostToDevice, c->stream));
    c->num_frames = 0;
    c->L.B_active = 1;
    if (c->noise_kdtree) {
        launch_noise_kdtree(c, 1, false);
    } else {
        launch_noise(c, 1, false);
    }
    LP_HIP(hipGetLastError());
    LP_HIP(hipMemcpyAsync(labels, c->d.noise, cloud->num_points, hipMemcpyDeviceToHost, c->stream));
    uint32_t err = 0;
    LP_HIP(hipMemcpyAsync(&err, c->d.fs + kFsNoiseErr, sizeof(err), hipMemcpyDeviceToHost, c->stream));
    LP_HIP(hipStreamSynchronize(c->stream));
    if (c->noise_kdtree) {
        const int st2 = status_of(err);
        if (st2 != LP_OK) return st2;
    }
    return LP_OK;
}

namespace {

// Stages num_polygons point lists (x, y packed by offsets) as frame 0's hull
// coordinates and runs `launch` over them; boxes[k] receives list k's box.
int polygon_boxes(lp_context* c, const double* x, const double* y, const int32_t* offsets, int32_t num_polygons,
                  lp_bounding_box* boxes, void (*launch)(lp_context*, int)) {
    if (!c || !offsets || num_polygons < 0 || (num_polygons > 0 && !boxes))
        return fail(LP_ERR_INVALID_ARGUMENT, "null argument");
    if (num_polygons == 0) return LP_OK;
    const int32_t total = offsets[num_polygons];
    if (offsets[0] != 0 || total < 0 || total > c->L.N || num_polygons > c->L.box_stride || num_polygons > c->L.N)
        return fail(LP_ERR_CAPACITY, "polygons exceed the context capacity");
    for (int k = 0; k < num_polygons; ++k)
        if (offsets[k + 1] < offsets[k]) return fail(LP_ERR_INVALID_ARGUMENT, "offsets must be non-decreasing");
    if (total > 0 && (!x || !y)) return fail(LP_ERR_INVALID_ARGUMENT, "null coordinates");
    LP_HIP(hipSetDevice(c->device));
    const size_t BN = static_cast<size_t>(c->L.B) * c->L.N;
    std::vector<uint32_t> fs(kFsCount, 0u);
    fs[kFsClusters] = static_cast<uint32_t>(num_polygons);
    fs[kFsHullTotal] = static_cast<uint32_t>(total);
    LP_HIP(hipMemcpyAsync(c->d.fs, fs.data(), kFsCount * 4, hipMemcpyHostToDevice, c->stream));
    LP_HIP(hipMemcpyAsync(c->d.hull_off, offsets, num_polygons * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    if (total) {
        LP_HIP(hipMemcpyAsync(c->d.hcx + BN, x, total * sizeof(double), hipMemcpyHostToDevice, c->stream));
        LP_HIP(hipMemcpyAsync(c->d.hcy + BN, y, total * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    c->num_frames = 0;
    c->L.B_active = 1;
    launch(c, num_polygons);
    LP_HIP(hipGetLastError());
    LP_HIP(hipMemcpyAsync(fs.data(), c->d.fs, kFsCount * 4, hipMemcpyDeviceToHost, c->stream));
    LP_HIP(hipMemcpyAsync(boxes, c->d.boxes, num_polygons * sizeof(lp_bounding_box), hipMemcpyDeviceToHost,
                          c->stream));
    LP_HIP(hipStreamSynchronize(c->stream));
    if (fs[kFsErr] & kErrBoxCapacity) return fail(LP_ERR_CAPACITY, "a polygon exceeds the box kernel's capacity");
    return LP_OK;
}

}  // namespace

int lp_oriented_boxes(lp_context* c, const double* x, const double* y, const int32_t* offsets, int32_t num_polygons,
                      lp_bounding_box* boxes) {
    return polygon_boxes(c, x, y, offsets, num_polygons, boxes,
                         [](lp_context* ctx, int) { launch_boxes(ctx, 1); });
}

int lp_pca_boxes(lp_context* c, const double* x, const double* y, const int32_t* offsets, int32_t num_polygons,
                 lp_bounding_box* boxes) {
    return polygon_boxes(c, x, y, offsets, num_polygons, boxes,
                         [](lp_context* ctx, int k) { launch_pca_boxes(ctx, k); });
}

int lp_add_ring_info(lp_context* c, const float* x, const float* y, int32_t n, uint16_t* ring) {
    if (!c || n < 0 || (n > 0 && (!x || !y || !ring))) return fail(LP_ERR_INVALID_ARGUMENT, "null argument");
    if (n > c->L.N) return fail(LP_ERR_CAPACITY, "cloud has more points than max_points_per_frame");
    if (n == 0) return LP_OK;
    LP_HIP(hipSetDevice(c->device));
    FrameDesc fd{};
    fd.n = n;
    c->frames.assign(1, fd);
    c->num_frames = 0;
    LP_HIP(hipMemcpyAsync(c->d.frames, &fd, sizeof(FrameDesc), hipMemcpyHostToDevice, c->stream));
    LP_HIP(hipMemcpyAsync(c->d.x, x, n * 4, hipMemcpyHostToDevice, c->stream));
    LP_HIP(hipMemcpyAsync(c->d.y, y, n * 4, hipMemcpyHostToDevice, c->stream));
    launch_ring_info(c, n);
    LP_HIP(hipGetLastError());
    LP_HIP(hipMemcpyAsync(ring, c->d.ring, n * 2, hipMemcpyDeviceToHost, c->stream));
    LP_HIP(hipStreamSynchronize(c->stream));
    return LP_OK;
}

int lp_convex_hulls(lp_context* c, const double* x, const double* y, const int32_t* offsets, int32_t num_lists,
                    int32_t* hull_offsets, int32_t* hull_indices, int32_t hull_capacity) {
    if (!c || !offsets || !hull_offsets || num_lists < 0) return fail(LP_ERR_INVALID_ARGUMENT, "null argument");
    if (num_lists == 0) {
        hull_offsets[0] = 0;
        return LP_OK;
    }
    const int32_t total = offsets[num_lists];
    if (offsets[0] != 0 || total < 0 || total > c->L.N || num_lists > c->L.N)
        return fail(LP_ERR_CAPACITY, "point lists exceed max_points_per_frame");
    for (int k = 0; k < num_lists; ++k)
        if (offsets[k + 1] < offsets[k]) return fail(LP_ERR_INVALID_ARGUMENT, "offsets must be non-decreasing");
    if (total > 0 && (!x || !y)) return fail(LP_ERR_INVALID_ARGUMENT, "null coordinates");
    LP_HIP(hipSetDevice(c->device));
    c->L.B_active = 1;
    LP_HIP(hipMemcpyAsync(c->d.hc_choff, offsets, (num_lists + 1) * sizeof(int32_t), hipMemcpyHostToDevice,
                          c->stream));
    if (total) {
        LP_HIP(hipMemcpyAsync(c->d.hcx, x, total * sizeof(double), hipMemcpyHostToDevice, c->stream));
        LP_HIP(hipMemcpyAsync(c->d.hcy, y, total * sizeof(double), hipMemcpyHostToDevice, c->stream));
        // (x, y) pairs kept for the canonical index among identical points.
        LP_HIP(hipMemcpy2DAsync(c->d.dsort, 16, x, 8, 8, total, hipMemcpyHostToDevice, c->stream));
        LP_HIP(hipMemcpy2DAsync(reinterpret_cast<char*>(c->d.dsort) + 8, 16, y, 8, 8, total, hipMemcpyHostToDevice,
                                c->stream));
    }
    launch_hull_lists(c, num_lists, total);
    LP_HIP(hipGetLastError());
    uint32_t fs[kFsCount];
    LP_HIP(hipMemcpyAsync(fs, c->d.fs, sizeof(fs), hipMemcpyDeviceToHost, c->stream));
    LP_HIP(hipStreamSynchronize(c->stream));
    const int32_t V = static_cast<int32_t>(fs[kFsHullTotal]);
    if (fs[kFsErr] & kErrHullCapacity) return fail(LP_ERR_CAPACITY, "hull candidates exceed the device capacity");
    if (V > hull_capacity || (V > 0 && !hull_indices)) return fail(LP_ERR_CAPACITY, "hull_capacity too small");
    LP_HIP(hipMemcpyAsync(hull_offsets, c->d.hull_off, num_lists * sizeof(int32_t), hipMemcpyDeviceToHost,
                          c->stream));
    if (V) LP_HIP(hipMemcpyAsync(hull_indices, c->d.hull_local, V * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    LP_HIP(hipStreamSynchronize(c->stream));
    hull_offsets[num_lists] = V;
    c->num_frames = 0;
    return LP_OK;
}

int lp_debug_fetch(lp_context* c, int32_t frame, int32_t which, void* dst, int32_t capacity) {
    if (!c || frame < 0 || frame >= static_cast<int32_t>(c->frames.size()) || !dst)
        return fail(LP_ERR_INVALID_ARGUMENT, "bad debug request");
    LP_HIP(hipSetDevice(c->device));
    LP_HIP(hipStreamSynchronize(c->stream));
    const FrameDesc& fd = c->frames[frame];
    const Layout& L = c->L;
    const int32_t P = fd.sc.H * fd.sc.W;
    const size_t po = static_cast<size_t>(frame) * L.P;
    uint32_t fs[kFsCount];
    LP_HIP(hipMemcpy(fs, c->d.fs + static_cast<size_t>(frame) * kFsCount, sizeof(fs), hipMemcpyDeviceToHost));
    const int32_t nq = static_cast<int32_t>(fs[kFsQueue]);
    switch (which) {
        case LP_DEBUG_ELEVATION_MAP: {
            const int32_t n = fd.sc.cells;
            if (n <= capacity)
                LP_HIP(hipMemcpy(dst, c->d.elev + static_cast<size_t>(frame) * L.cell_stride, n * 4,
                                 hipMemcpyDeviceToHost));
            return n;
        }
        case LP_DEBUG_MAPPING:
            if (P <= capacity) LP_HIP(hipMemcpy(dst, c->d.mapping + po, P * 4, hipMemcpyDeviceToHost));
            return P;
        case LP_DEBUG_STATE_RECM:
            if (P <= capacity) LP_HIP(hipMemcpy(dst, c->d.state + po, P, hipMemcpyDeviceToHost));
            return P;
        case LP_DEBUG_JCP_QUEUE:
            if (nq <= capacity && nq > 0) LP_HIP(hipMemcpy(dst, c->d.qlist + po, nq * 4, hipMemcpyDeviceToHost));
            return nq;
        case LP_DEBUG_STATE_FINAL: {
            if (P > capacity) return P;
            std::vector<uint8_t> st(P), fin(std::max(nq, 1));
            std::vector<int32_t> qm(P);
            LP_HIP(hipMemcpy(st.data(), c->d.state + po, P, hipMemcpyDeviceToHost));
            LP_HIP(hipMemcpy(qm.data(), c->d.qmap + po, P * 4, hipMemcpyDeviceToHost));
            if (nq > 0) LP_HIP(hipMemcpy(fin.data(), c->d.fin + po, nq, hipMemcpyDeviceToHost));
            uint8_t* o = static_cast<uint8_t*>(dst);
            for (int32_t p = 0; p < P; ++p) o[p] = qm[p] >= 0 ? fin[qm[p]] : st[p];
            return P;
        }
        default:
            return fail(LP_ERR_INVALID_ARGUMENT, "unknown debug buffer");
    }
}

int lp_frame_image(lp_context* c, int32_t frame, uint8_t* bgr, int32_t capacity) {
    if (!c || frame < 0 || !bgr) return fail(LP_ERR_INVALID_ARGUMENT, "bad image request");
    // The lane that ran the frame: lane k holds the k-th contiguous share.
    lp_context* lane = c;
    int32_t f = frame;
    if (!c->lanes.empty()) {
        size_t k = 0;
        while (k < c->lane_frames.size() && f >= c->lane_frames[k]) f -= c->lane_frames[k++];
        if (k == c->lane_frames.size()) return fail(LP_ERR_INVALID_ARGUMENT, "frame not in the last batch");
        lane = k == 0 ? c : c->lanes[k - 1];
    }
    if (f >= lane->num_frames || f >= static_cast<int32_t>(lane->frames.size()))
        return fail(LP_ERR_INVALID_ARGUMENT, "frame not in the last batch");
    const SegConst& sc = lane->frames[f].sc;
    const int32_t pixels = sc.H * sc.W;
    const int32_t bytes = 3 * pixels;
    if (bytes > capacity) return bytes;
    LP_HIP(hipSetDevice(lane->device));
    if (!lane->image_dev) LP_HIP(hipMalloc(&lane->image_dev, 3 * static_cast<size_t>(lane->L.P)));
    launch_label_image(lane, f, pixels, lane->image_dev);
    LP_HIP(hipGetLastError());
    LP_HIP(hipMemcpyAsync(bgr, lane->image_dev, bytes, hipMemcpyDeviceToHost, lane->stream));
    LP_HIP(hipStreamSynchronize(lane->stream));
    return bytes;
}

int lp_stage_names(lp_context* c, char* buf, int32_t buf_len) {
    if (!c || !buf || buf_len <= 0) return fail(LP_ERR_INVALID_ARGUMENT, "null argument");
    std::string s;
    for (size_t i = 0; i < c->stage_names.size(); ++i) {
        if (i) s += ",";
        s += c->stage_names[i];
    }
    std::snprintf(buf, static_cast<size_t>(buf_len), "%s", s.c_str());
    return static_cast<int>(c->stage_names.size());
}

int lp_profile_run(lp_context* c, int32_t runs, float* stage_ms, int32_t max_stages, int32_t* num_stages) {
    if (!c || c->num_frames < 1 || runs < 1) return fail(LP_ERR_INVALID_ARGUMENT, "nothing staged");
    LP_HIP(hipSetDevice(c->device));
    std::vector<float> acc;
    for (int r = 0; r < runs; ++r) {
        for (auto e : c->events) (void)hipEventDestroy(e);
        c->events.clear();
        c->stage_names.clear();
        c->profiling = true;
        const int st = run_pipeline(c, c->num_frames, false);
        stage_mark(c, "end");
        c->profiling = false;
        if (st != LP_OK) return st;
        LP_HIP(hipStreamSynchronize(c->stream));
        acc.resize(c->events.size() - 1, 0.0f);
        for (size_t i = 0; i + 1 < c->events.size(); ++i) {
            float ms = 0.0f;
            LP_HIP(hipEventElapsedTime(&ms, c->events[i], c->events[i + 1]));
            acc[i] += ms;
        }
    }
    const int ns = static_cast<int>(acc.size());
    for (int i = 0; i < ns && i < max_stages; ++i) stage_ms[i] = acc[i] / runs;
    if (num_stages) *num_stages = ns;
    c->stage_names.pop_back();
    return LP_OK;
}

// PointCloud2 field dispatch of Processor::run (src/processor/src/
// processor.cpp:494-549) and the per-layout conversions convert<PointT>
// (processor.cpp:42-179): the reference copies every point into a
// pcl::PointCloud; here the message buffer itself becomes the lp_cloud
// (byte-strided views), so the points go to the GPU straight from it.
static bool field_names_are(const lp_point_field* f, int32_t n, std::initializer_list<const char*> want) {
    if (n != static_cast<int32_t>(want.size())) return false;
    int32_t k = 0;
    for (const char* w : want) {
        if (!f[k].name || std::strcmp(f[k].name, w) != 0) return false;
        ++k;
    }
    return true;
}

int lp_cloud_from_pointcloud2(const uint8_t* data, uint32_t width, uint32_t height, uint32_t point_step,
                              uint32_t row_step, const lp_point_field* fields, int32_t num_fields, lp_cloud* cloud) {
    if (!fields || !cloud || num_fields < 0) return fail(LP_ERR_INVALID_ARGUMENT, "null argument");
    int ring_field = -1;
    if (field_names_are(fields, num_fields, {"x", "y", "z"}) ||
        field_names_are(fields, num_fields, {"x", "y", "z", "intensity"})) {
        ring_field = -1;
    } else if (field_names_are(fields, num_fields, {"x", "y", "z", "ring"})) {
        ring_field = 3;
    } else if (field_names_are(fields, num_fields, {"x", "y", "z", "intensity", "ring"})) {
        ring_field = 4;
    } else {
        return fail(LP_ERR_INVALID_ARGUMENT,
                    "Unsupported point format. Acceptable formats: pcl::PointXYZ, pcl::PointXYZI, "
                    "pcl::PointXYZR, pcl::PointXYZIR");
    }
    const uint64_t n = static_cast<uint64_t>(width) * height;
    if (n > 0x7fffffffull) return fail(LP_ERR_CAPACITY, "more than 2^31 - 1 points");
    if (height > 1 && static_cast<uint64_t>(row_step) != static_cast<uint64_t>(width) * point_step)
        return fail(LP_ERR_INVALID_ARGUMENT, "rows are not packed (row_step != width * point_step)");
    if (point_step > 0x7fffffffu) return fail(LP_ERR_INVALID_ARGUMENT, "point_step too large");
    for (int k = 0; k < 3; ++k)
        if (static_cast<uint64_t>(fields[k].offset) + 4 > point_step)
            return fail(LP_ERR_INVALID_ARGUMENT, "a coordinate field lies outside the point");
    if (ring_field >= 0 && static_cast<uint64_t>(fields[ring_field].offset) + 2 > point_step)
        return fail(LP_ERR_INVALID_ARGUMENT, "the ring field lies outside the point");
    if (n > 0 && !data) return fail(LP_ERR_INVALID_ARGUMENT, "null data");
    lp_cloud c{};
    c.num_points = static_cast<int32_t>(n);
    c.xyz_stride = static_cast<int32_t>(point_step);
    c.ring_stride = static_cast<int32_t>(point_step);
    c.on_device = 0;
    c.x = data ? data + fields[0].offset : nullptr;
    c.y = data ? data + fields[1].offset : nullptr;
    c.z = data ? data + fields[2].offset : nullptr;
    c.ring = (ring_field >= 0 && data) ? data + fields[ring_field].offset : nullptr;
    *cloud = c;
    return LP_OK;
}

}  // extern "C"
