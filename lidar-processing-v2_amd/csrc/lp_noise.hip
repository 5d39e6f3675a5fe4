// lp_noise.hip -- dynamic radius outlier removal (DROR) on the GPU.
//
// NoiseRemover::filter (lidar_processing_lib/src/noise_remover.cpp:38-68):
// a point is NOISE when fewer than min_neighbours points, itself included,
// satisfy dist_sqr(point, other) <= r^2, with
//   r^2 = max((float)(pow((double)multiplier, 2) * (double)(x*x + y*y)), min_r*min_r)
//   dist_sqr = (t0-p0)^2 + ((t1-p1)^2 + ((t2-p2)^2 + 0))   (kdtree.hpp:131-143)
// all in float except the widened product.
//
// Layout: the frame's points are bucketed into 1 m grid cells through an
// open-addressing hash of cell keys (one slot per occupied cell, load <= 0.5),
// counted, scanned and scattered into a cell-sorted float4 array.  One thread
// per point (in cell-sorted order, so a wavefront reads neighbouring cells)
// visits its own cell and then the cube of cells within ceil(r) + 1 and stops
// at the min_neighbours-th hit.
#include "lp_device.h"
#include "lp_internal.h"
#include "lp_kernels.h"

namespace lp {

using namespace lpd;

namespace {

constexpr int kNoiseMaxExtent = 4;        // cells each way; larger radii scan the whole frame
constexpr float kNoiseCellSlack = 0.015625f;

// 1 m cells; x and y clamped to [-512, 511] m and z to [-1024, 1023] m so the
// key packs into 31 bits (never -1, the empty marker).  Clamping is monotone
// and never moves two cells further apart, so the cube search still covers
// every cell that can hold a point within the radius.
__device__ __forceinline__ void noise_cell(float x, float y, float z, int& cx, int& cy, int& cz) {
    cx = static_cast<int>(fminf(fmaxf(floorf(x), -512.0f), 511.0f)) + 512;
    cy = static_cast<int>(fminf(fmaxf(floorf(y), -512.0f), 511.0f)) + 512;
    cz = static_cast<int>(fminf(fmaxf(floorf(z), -1024.0f), 1023.0f)) + 1024;
}

__device__ __forceinline__ int32_t noise_key(int cx, int cy, int cz) { return cx | (cy << 10) | (cz << 20); }

__device__ __forceinline__ uint32_t noise_slot(int32_t key, uint32_t mask) {
    uint32_t h = static_cast<uint32_t>(key) * 0x9E3779B1u;
    h ^= h >> 16;
    return h & mask;
}

__device__ __forceinline__ int32_t noise_find(const int32_t* hkey, uint32_t mask, int32_t key) {
    uint32_t h = noise_slot(key, mask);
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const int32_t k = hkey[h];
        if (k == key) return static_cast<int32_t>(h);
        if (k == -1) return -1;
        h = (h + 1) & mask;
    }
    return -1;
}

// noise_remover.cpp:46-59; std::max(a, b) keeps a unless a < b (a NaN stays).
__device__ __forceinline__ float noise_radius_sqr(const NoiseConst& nc, float x, float y) {
    const double range_sqr = static_cast<double>(x * x + y * y);
    const float r2 = static_cast<float>(nc.scaling * range_sqr);
    return (r2 < nc.min_sqr) ? nc.min_sqr : r2;
}

__device__ __forceinline__ float noise_dist(const float4& t, const float4& p) {
    const float d0 = t.x - p.x, d1 = t.y - p.y, d2 = t.z - p.z;
    return d0 * d0 + (d1 * d1 + (d2 * d2 + 0.0f));
}

__global__ void k_noise_clear(DevBuf d, Layout L) {
    const int f = blockIdx.y;
    const int stride = gridDim.x * blockDim.x;
    const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    for (int h = t0; h < L.hash_cap; h += stride) {
        const size_t o = (size_t)f * L.hash_cap + h;
        d.hkey[o] = -1;
        d.hmin[o] = 0;
        d.hcount[o] = 0;
    }
    if (t0 == 0) {
        d.fs[f * kFsCount + kFsKept] = 0;
        d.fs[f * kFsCount + kFsNoise] = 0;
    }
}

__global__ void k_noise_insert(DevBuf d, Layout L) {
    const int f = blockIdx.y;
    const FrameDesc& fd = d.frames[f];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fd.n_raw) return;
    const int gi = fd.pt_off + i;
    int cx, cy, cz;
    noise_cell(d.rx[gi], d.ry[gi], d.rz[gi], cx, cy, cz);
    const int32_t key = noise_key(cx, cy, cz);
    const uint32_t mask = static_cast<uint32_t>(L.hash_cap) - 1;
    const size_t ho = (size_t)f * L.hash_cap;
    // Lanes with the same cell (consecutive points mostly share one) insert
    // once: the lowest such lane probes and adds the group's count.
    const int lane = lane_id();
    uint64_t pending = __ballot(1);
    int32_t slot = 0;
    while (pending) {
        const int leader = __ffsll(static_cast<unsigned long long>(pending)) - 1;
        const int32_t lkey = __shfl(key, leader, kWave);
        const bool mine = ((pending >> lane) & 1ull) && key == lkey;
        const uint64_t grp = __ballot(mine);
        uint32_t h = 0;
        if (lane == leader) {
            h = noise_slot(key, mask);
            for (uint32_t probe = 0; probe <= mask; ++probe) {
                const int32_t prev = atomicCAS(&d.hkey[ho + h], -1, key);
                if (prev == -1 || prev == key) break;
                h = (h + 1) & mask;
            }
            atomicAdd(&d.hcount[ho + h], static_cast<uint32_t>(__popcll(grp)));
        }
        h = __shfl(h, leader, kWave);
        if (mine) slot = static_cast<int32_t>(h);
        pending &= ~grp;
    }
    d.vslot[(size_t)f * L.N + i] = slot;
}

// hdense holds the exclusive scan of hcount (cell starts) on entry.
__global__ void k_noise_scatter(DevBuf d, Layout L) {
    const int f = blockIdx.y;
    const FrameDesc& fd = d.frames[f];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fd.n_raw) return;
    const int gi = fd.pt_off + i;
    const size_t ho = (size_t)f * L.hash_cap;
    const int32_t h = d.vslot[(size_t)f * L.N + i];
    const int pos = d.hdense[ho + h] + static_cast<int>(wave_atomic_inc(reinterpret_cast<uint32_t*>(&d.hmin[ho + h])));
    const size_t po = (size_t)f * L.N;
    d.dsort[po + pos] = make_float4(d.rx[gi], d.ry[gi], d.rz[gi], 0.0f);
    d.didx[po + pos] = i;
}

// Squared distance from the point to the box of cell offset (dx, dy, dz)
// relative to its own cell, per axis from the point's position inside its
// 1 m cell.  Cells whose box is beyond the radius hold no neighbour.
__device__ __forceinline__ float cell_gap_sqr(float fx, float fy, float fz, int dx, int dy, int dz) {
    const float gx = dx > 0 ? (static_cast<float>(dx) - fx) : (dx < 0 ? (fx - static_cast<float>(dx + 1)) : 0.0f);
    const float gy = dy > 0 ? (static_cast<float>(dy) - fy) : (dy < 0 ? (fy - static_cast<float>(dy + 1)) : 0.0f);
    const float gz = dz > 0 ? (static_cast<float>(dz) - fz) : (dz < 0 ? (fz - static_cast<float>(dz + 1)) : 0.0f);
    return gx * gx + gy * gy + gz * gz;
}

__global__ __launch_bounds__(256) void k_noise_count(DevBuf d, Layout L) {
    const int f = blockIdx.y;
    const FrameDesc& fd = d.frames[f];
    const int n = fd.n_raw;
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= n) return;
    const NoiseConst nc = *d.nc;
    const size_t po = (size_t)f * L.N;
    const size_t ho = (size_t)f * L.hash_cap;
    const uint32_t mask = static_cast<uint32_t>(L.hash_cap) - 1;
    const float4 t = d.dsort[po + pos];
    const int i = d.didx[po + pos];
    const float r2 = noise_radius_sqr(nc, t.x, t.y);
    const uint32_t k = nc.k;
    uint32_t cnt = 0;
    if (k > 0 && r2 >= 0.0f) {
        const float r = sqrtf(r2);
        if (r < static_cast<float>(kNoiseMaxExtent - 1)) {
            const int e = static_cast<int>(floorf(r + kNoiseCellSlack)) + 1;
            int cx, cy, cz;
            noise_cell(t.x, t.y, t.z, cx, cy, cz);
            // Position inside the own cell (exact for in-range coordinates;
            // clamped cells give values outside [0, 1), which only make the
            // cull below more permissive).  The cull keeps every cell within
            // (r + slack): the float test admits a few ulps beyond r.
            const float fx = t.x - floorf(t.x), fy = t.y - floorf(t.y), fz = t.z - floorf(t.z);
            const float reach = (r + kNoiseCellSlack) * (r + kNoiseCellSlack);
            bool done = false;
            // Shells of growing Chebyshev radius: own cell, its 26
            // neighbours, then the outer shells, nearest first.
            for (int sh = 0; sh <= e && !done; ++sh) {
                for (int dz = -sh; dz <= sh && !done; ++dz)
                    for (int dy = -sh; dy <= sh && !done; ++dy)
                        for (int dx = -sh; dx <= sh && !done; ++dx) {
                            if (max(abs(dx), max(abs(dy), abs(dz))) != sh) continue;
                            if (sh > 0 && cell_gap_sqr(fx, fy, fz, dx, dy, dz) > reach) continue;
                            const int ax = cx + dx, ay = cy + dy, az = cz + dz;
                            if (ax < 0 || ax > 1023 || ay < 0 || ay > 1023 || az < 0 || az > 2047) continue;
                            const int32_t h = noise_find(d.hkey + ho, mask, noise_key(ax, ay, az));
                            if (h < 0) continue;
                            const int s0 = d.hdense[ho + h];
                            const int s1 = s0 + static_cast<int>(d.hcount[ho + h]);
                            // Four points per step, loads issued together;
                            // counting past the k-th hit leaves the label alone.
                            int s = s0;
                            for (; s + 4 <= s1 && !done; s += 4) {
                                const float4 a = d.dsort[po + s], b = d.dsort[po + s + 1];
                                const float4 c = d.dsort[po + s + 2], q = d.dsort[po + s + 3];
                                cnt += (noise_dist(t, a) <= r2 ? 1u : 0u) + (noise_dist(t, b) <= r2 ? 1u : 0u) +
                                       (noise_dist(t, c) <= r2 ? 1u : 0u) + (noise_dist(t, q) <= r2 ? 1u : 0u);
                                done = cnt >= k;
                            }
                            for (; s < s1 && !done; ++s) {
                                cnt += noise_dist(t, d.dsort[po + s]) <= r2 ? 1u : 0u;
                                done = cnt >= k;
                            }
                        }
            }
        } else {
            // Radius beyond the cell cube (range > ~150 m) or infinite: scan all.
            for (int s = 0; s < n; ++s)
                if (noise_dist(t, d.dsort[po + s]) <= r2 && ++cnt >= k) break;
        }
    }
    const bool noise = cnt < k;
    d.noise[po + i] = noise ? 1 : 0;
    d.obs_flag[po + i] = noise ? 0u : 1u;  // keep flag for the compaction scan
    if (noise) atomicAdd(&d.fs[f * kFsCount + kFsNoise], 1u);
}

// Stable compaction of the kept points into the segmenter's inputs; the
// frame's point count becomes the kept count.
__global__ void k_noise_compact(DevBuf d, Layout L) {
    const int f = blockIdx.y;
    FrameDesc& fd = d.frames[f];
    const int n = fd.n_raw;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t po = (size_t)f * L.N;
    if (blockIdx.x == 0 && threadIdx.x == 0) fd.n = static_cast<int32_t>(d.fs[f * kFsCount + kFsKept]);
    if (i >= n || !d.obs_flag[po + i]) return;
    const int gi = fd.pt_off + i;
    const int go = fd.pt_off + static_cast<int>(reinterpret_cast<const uint32_t*>(d.obs_src)[po + i]);
    d.x[go] = d.rx[gi];
    d.y[go] = d.ry[gi];
    d.z[go] = d.rz[gi];
    if (fd.has_ring) d.ring[go] = d.rring[gi];
}

}  // namespace

void launch_noise(lp_context* c, int B, bool compact) {
    DevBuf& d = c->d;
    const Layout& L = c->L;
    hipStream_t st = c->stream;
    const int nb = (L.N + 255) / 256;
    stage_mark(c, "noise");
    hipLaunchKernelGGL(k_noise_clear, dim3(256, B), dim3(256), 0, st, d, L);
    hipLaunchKernelGGL(k_noise_insert, dim3(nb, B), dim3(256), 0, st, d, L);
    launch_scan(c, B, d.hcount, reinterpret_cast<uint32_t*>(d.hdense), L.hash_cap, kCountHash, 0, -1);
    hipLaunchKernelGGL(k_noise_scatter, dim3(nb, B), dim3(256), 0, st, d, L);
    stage_mark(c, "noise_count");
    hipLaunchKernelGGL(k_noise_count, dim3(nb, B), dim3(256), 0, st, d, L);
    stage_mark(c, "noise_compact");
    if (compact) {
        launch_scan(c, B, d.obs_flag, reinterpret_cast<uint32_t*>(d.obs_src), L.N, kCountRawPoints, 0, kFsKept);
        hipLaunchKernelGGL(k_noise_compact, dim3(nb, B), dim3(256), 0, st, d, L);
    }
}

}  // namespace lp
