// lp_noise.hip -- dynamic radius outlier removal (DROR) on the GPU.
//
// NoiseRemover::filter (lidar_processing_lib/src/noise_remover.cpp:38-68):
// a point is NOISE when fewer than min_neighbours points, itself included,
// satisfy dist_sqr(point, other) <= r^2, with
//   r^2 = max((float)(pow((double)multiplier, 2) * (double)(x*x + y*y)), min_r*min_r)
//   dist_sqr = (t0-p0)^2 + ((t1-p1)^2 + ((t2-p2)^2 + 0))   (kdtree.hpp:131-143)
// all in float except the widened product.
//
// A first pass, one thread per point in input order, counts the points
// within the radius among its 8 input-order neighbours on each side: scan
// order keeps a point's azimuth neighbours of the same beam next to it, so
// this decides almost every VALID point of a surface with 17 coalesced
// distance tests (any min_neighbours points within the radius prove VALID).
// The rest are listed (the open list) and go through the grid.
//
// Layout: the frame's points are bucketed into 1 m grid cells through an
// open-addressing hash of cell keys (one slot per occupied cell, load at most
// the context's hash load factor), counted, scanned and scattered into a
// cell-sorted float4 array (w holds the cell's hash slot).  The same pass
// that runs the input-order test inserts the point and takes its place in
// the cell from the cell count's atomic add, so the scatter needs no
// atomics.  The grid count runs in two passes.  The first, one
// thread per open point, checks the point's own cell outward from its own
// position (points next to each other in scan order are mostly neighbours in
// space), so almost every open point of a dense surface reaches
// min_neighbours after a few distance tests; a point whose ball lies inside
// its cell is decided there too.  The rest (rain, isolated returns, points
// near a cell face) are listed and the second pass, one wavefront per listed
// point, visits the cube of cells within ceil(r) + 1 nearest first: the slow
// searches run side by side instead of holding up wavefronts of easy points.
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

__device__ __forceinline__ int32_t noise_find(const HashSlot* hs, uint32_t mask, int32_t key) {
    uint32_t h = noise_slot(key, mask);
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const int32_t k = hs[h].key;
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

// The hash slots are clean between uses: every user
// resets the slots it occupied when it is done (k_hash_reset), so only the
// column bits and the counters are cleared here.
__global__ void k_noise_clear(DevBuf d, Layout L) {
    const int f = blockIdx.y;
    const int stride = gridDim.x * blockDim.x;
    const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    for (int q = t0; q < kNoiseOccWords; q += stride) d.nocc[(size_t)f * kNoiseOccWords + q] = 0;
    if (t0 == 0) {
        d.fs[f * kFsCount + kFsKept] = 0;
        d.fs[f * kFsCount + kFsNoise] = 0;
        d.fs[f * kFsCount + kFsNoiseOpen] = 0;
        d.fs[f * kFsCount + kFsNoiseHard] = 0;
        d.fs[f * kFsCount + kFsNoiseCells] = 0;
        d.fs[f * kFsCount + kFsNoiseErr] = 0;
    }
}

constexpr int kNearWindow = 8;  // input-order neighbours tested on each side

constexpr int kPrepRows = 2;  // points per thread: consecutive runs of 256, their atomics in flight together
constexpr int kPrepTile = 256 * kPrepRows;

// Two jobs on the same coordinates, kPrepRows points per thread (rows of 256
// consecutive points, so the lanes of a wavefront hold consecutive points):
//  * input-order pass: noise[i] = 0 (VALID) when at least min_neighbours of
//    the points i - 8 .. i + 8 (i itself included) lie within the radius,
//    else 2, and the point joins the open list for the grid passes;
//  * grid insert: the point's cell gets a hash slot (vslot) and the point a
//    position inside its cell (nloc): the cell's count before the point's run
//    plus its rank in the run, so the scatter needs no atomics.
// The rows' atomics are issued side by side: the kernel waits on atomic
// round trips, and two points per thread halve the waits per point.
__global__ __launch_bounds__(256) void k_noise_prep(DevBuf d, Layout L) {
    __shared__ float4 tile_pts[kPrepTile + 2 * kNearWindow];
    __shared__ uint32_t rsv[256 / kWave + 1];
    int tile, f;
    xcd_tile_frame(tile, f);
    const FrameDesc& fd = d.frames[f];
    const int n = fd.n_raw;
    const int base = tile * kPrepTile;
    if (base >= n) return;
    for (int q = threadIdx.x; q < kPrepTile + 2 * kNearWindow; q += blockDim.x) {
        const int j = min(max(base - kNearWindow + q, 0), n - 1);
        const int gj = fd.pt_off + j;
        tile_pts[q] = make_float4(d.rx[gj], d.ry[gj], d.rz[gj], 0.0f);
    }
    __syncthreads();
    const NoiseConst nc = *d.nc;
    const size_t po = (size_t)f * L.N;
    const uint32_t mask = static_cast<uint32_t>(L.hash_cap) - 1;
    const size_t ho = (size_t)f * L.hash_cap;
    const int lane = lane_id();
    int i[kPrepRows], cxs[kPrepRows], cys[kPrepRows];
    bool open[kPrepRows];
    int32_t key[kPrepRows];
#pragma unroll
    for (int s = 0; s < kPrepRows; ++s) {
        const int e = s * 256 + static_cast<int>(threadIdx.x);
        i[s] = base + e;
        const bool in_frame = i[s] < n;
        const float4 t = tile_pts[e + kNearWindow];
        const float r2 = noise_radius_sqr(nc, t.x, t.y);
        const int lo = max(i[s] - kNearWindow, 0), hi = min(i[s] + kNearWindow, n - 1);
        uint32_t cnt = 0;
#pragma unroll
        for (int o = -kNearWindow; o <= kNearWindow; ++o) {
            const int j = i[s] + o;
            const bool in = j >= lo && j <= hi;
            cnt += (in && noise_dist(t, tile_pts[e + kNearWindow + o]) <= r2) ? 1u : 0u;
        }
        open[s] = in_frame && cnt < nc.k;
        if (in_frame) d.noise[po + i[s]] = open[s] ? 2 : 0;
        int cz;
        noise_cell(t.x, t.y, t.z, cxs[s], cys[s], cz);
        // Lanes past the frame's end take key -2: never a cell key (keys
        // are non-negative), they form their own trailing run and insert
        // nothing.
        key[s] = in_frame ? noise_key(cxs[s], cys[s], cz) : -2;
    }
    uint32_t q[kPrepRows];
    block_reserve<kPrepRows>(&d.fs[f * kFsCount + kFsNoiseOpen], open, q, rsv);
#pragma unroll
    for (int s = 0; s < kPrepRows; ++s)
        if (open[s]) d.nopen[po + q[s]] = i[s];

    // A run of lanes with the same cell (consecutive points mostly share one)
    // inserts once: its lowest lane probes and adds the run's count; all
    // runs of both rows at once.
    WaveRun run[kPrepRows];
    bool lead[kPrepRows];
    uint32_t h[kPrepRows];
    int32_t prev[kPrepRows];
#pragma unroll
    for (int s = 0; s < kPrepRows; ++s) {
        run[s] = wave_runs(key[s]);
        lead[s] = lane == run[s].leader && key[s] >= 0;
        h[s] = noise_slot(key[s], mask);
        prev[s] = lead[s] ? atomicCAS(&d.hslot[ho + h[s]].key, -1, key[s]) : key[s];
    }
#pragma unroll
    for (int s = 0; s < kPrepRows; ++s) {
        // Collisions (another cell's key in the probed slot) probe on.
        for (uint32_t probe = 1; probe <= mask && prev[s] != -1 && prev[s] != key[s]; ++probe) {
            h[s] = (h[s] + 1) & mask;
            prev[s] = atomicCAS(&d.hslot[ho + h[s]].key, -1, key[s]);
        }
    }
    bool fresh[kPrepRows];
    uint32_t before[kPrepRows];
#pragma unroll
    for (int s = 0; s < kPrepRows; ++s) {
        fresh[s] = lead[s] && prev[s] == -1;
        before[s] = 0;
        if (lead[s]) {
            before[s] = atomicAdd(&d.hslot[ho + h[s]].count, run[s].count);
            atomicOr(&d.nocc[(size_t)f * kNoiseOccWords + cys[s] * 32 + (cxs[s] >> 5)], 1u << (cxs[s] & 31));
        }
    }
    // Every occupied cell gets a dense id (in any order): the cell counts are
    // scanned over the ids, not over the whole table.
    uint32_t id[kPrepRows];
    block_reserve<kPrepRows>(&d.fs[f * kFsCount + kFsNoiseCells], fresh, id, rsv);
#pragma unroll
    for (int s = 0; s < kPrepRows; ++s) {
        if (fresh[s]) {
            d.ncell_slot[(size_t)f * L.N + id[s]] = static_cast<int32_t>(h[s]);
            d.hslot[ho + h[s]].dense = static_cast<int32_t>(id[s]);
        }
        const uint32_t slot = __shfl(h[s], run[s].leader, kWave);
        const uint32_t place = __shfl(before[s], run[s].leader, kWave) + run[s].rank;
        if (i[s] < n) {
            d.vslot[po + i[s]] = static_cast<int32_t>(slot);
            d.nloc[po + i[s]] = static_cast<int32_t>(place);
        }
    }
}

// Each slot's aux holds its cell's first cell-sorted position on entry (the
// exclusive scan of the cell counts by dense id); each point goes to that
// start plus its place in the cell.
__global__ void k_noise_scatter(DevBuf d, Layout L) {
    int tile, f;
    xcd_tile_frame(tile, f);
    const FrameDesc& fd = d.frames[f];
    const int i = tile * blockDim.x + threadIdx.x;
    if (i >= fd.n_raw) return;
    const int gi = fd.pt_off + i;
    const size_t po = (size_t)f * L.N;
    const size_t ho = (size_t)f * L.hash_cap;
    const int32_t h = d.vslot[po + i];
    const int loc = d.nloc[po + i];
    const int pos = d.hslot[ho + h].aux + loc;  // the cell's start plus the point's place
    d.dsort[po + pos] = make_float4(d.rx[gi], d.ry[gi], d.rz[gi], __int_as_float(h));
    d.spos[po + i] = pos;
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

// Writes the label of point i and counts NOISE labels per wavefront.
__device__ __forceinline__ void noise_label(const DevBuf& d, int f, size_t po, int i, bool noise) {
    d.noise[po + i] = noise ? 1 : 0;
    const uint64_t m = __ballot(noise);
    if (m && lane_id() == __ffsll(static_cast<unsigned long long>(m)) - 1)
        atomicAdd(&d.fs[f * kFsCount + kFsNoise], static_cast<uint32_t>(__popcll(m)));
}

constexpr int kOwnCellBudget = 64;  // distance tests of the own-cell pass

// Own-cell test of point i at cell-sorted position pos (see k_noise_count).
__device__ __forceinline__ void noise_own_cell(const DevBuf& d, const NoiseConst& nc, int f, size_t po, size_t ho,
                                               int i, int pos) {
    const float4 t = d.dsort[po + pos];
    const float r2 = noise_radius_sqr(nc, t.x, t.y);
    const uint32_t k = nc.k;
    int state;  // 0 VALID, 1 NOISE, 2 listed
    uint32_t carry = 0;  // listed: own-cell count and how far the scan got
    if (k == 0) {
        state = 0;
    } else if (!(r2 >= 0.0f)) {
        state = 1;  // a NaN radius admits no point
    } else {
        const int h = __float_as_int(t.w);
        const HashSlot hv = d.hslot[ho + h];
        const int s0 = hv.aux;
        const int s1 = s0 + static_cast<int>(hv.count);
        uint32_t cnt = 0;
        int budget = kOwnCellBudget;
        int up = pos, dn = pos - 1;
        while (cnt < k && budget > 0 && (up < s1 || dn >= s0)) {
            // Alternate after and before the point (its scan-order neighbours).
            const bool fwd = up < s1 && (dn < s0 || ((budget & 1) == 0));
            const int s = fwd ? up++ : dn--;
            cnt += noise_dist(t, d.dsort[po + s]) <= r2 ? 1u : 0u;
            --budget;
        }
        if (cnt >= k) {
            state = 0;
        } else {
            state = 2;
            carry = cnt | static_cast<uint32_t>(up - pos) << 8 | static_cast<uint32_t>(pos - 1 - dn) << 16;
            const float r = sqrtf(r2);
            int cx, cy, cz;
            noise_cell(t.x, t.y, t.z, cx, cy, cz);
            const bool whole = up >= s1 && dn < s0;
            const bool interior = cx > 0 && cx < 1023 && cy > 0 && cy < 1023 && cz > 0 && cz < 2047;
            if (whole && interior && r < static_cast<float>(kNoiseMaxExtent - 1)) {
                const float fx = t.x - floorf(t.x), fy = t.y - floorf(t.y), fz = t.z - floorf(t.z);
                const float reach = (r + kNoiseCellSlack) * (r + kNoiseCellSlack);
                if (cell_gap_sqr(fx, fy, fz, -1, 0, 0) > reach && cell_gap_sqr(fx, fy, fz, 1, 0, 0) > reach &&
                    cell_gap_sqr(fx, fy, fz, 0, -1, 0) > reach && cell_gap_sqr(fx, fy, fz, 0, 1, 0) > reach &&
                    cell_gap_sqr(fx, fy, fz, 0, 0, -1) > reach && cell_gap_sqr(fx, fy, fz, 0, 0, 1) > reach)
                    state = 1;
            }
        }
    }
    if (state == 2) {
        const uint32_t q = wave_atomic_inc(&d.fs[f * kFsCount + kFsNoiseHard]);
        d.vslot[po + q] = pos;  // the listed point: position, input index and carry, loaded side by side
        d.nloc[po + q] = i;
        d.obs_src[po + q] = static_cast<int32_t>(carry);
    } else {
        noise_label(d, f, po, i, state == 1);
    }
}

// Own-cell pass over the open list (a grid-stride loop).  A point is VALID
// once min_neighbours points of its cell are within its radius (the count
// includes itself, and the search order does not change the outcome), NOISE
// when its whole cell has been checked and no other cell comes within its
// radius, and listed for the neighbour-cell pass otherwise.
__global__ __launch_bounds__(256) void k_noise_count(DevBuf d, Layout L) {
    int tile, f;
    xcd_tile_frame(tile, f);
    const int nopen = static_cast<int>(d.fs[f * kFsCount + kFsNoiseOpen]);
    const NoiseConst nc = *d.nc;
    const size_t po = (size_t)f * L.N;
    const size_t ho = (size_t)f * L.hash_cap;
    for (int q = tile * blockDim.x + threadIdx.x; q < nopen; q += gridDim.x * blockDim.x) {
        const int i = d.nopen[po + q];
        const int pos = d.spos[po + i];
        noise_own_cell(d, nc, f, po, ho, i, pos);
    }
}

// Neighbour-cell pass over the listed points, one wavefront per point (a
// grid-stride loop over the list).  The lanes scan cell ranges 64 points at a
// time and count hits with a ballot, so a point next to a large cell costs
// a few steps instead of one lane walking the cell.  The own cell continues
// from where the first pass left off ([dn, up) was checked); then the cube of
// floor(r + 1/64) + 1 cells each way, shell by shell (Chebyshev radius), the
// lanes testing 64 cells at once: box distance within the radius, column
// occupied (one bit per 1 m (x, y) column), cell present in the hash.  The
// search stops at the min_neighbours-th hit.  Radii of 3 m and more (ranges
// beyond ~150 m) scan the whole frame.
__device__ __forceinline__ uint32_t wave_count_range(const DevBuf& d, size_t po, const float4& t, float r2, int a,
                                                     int b, uint32_t cnt, uint32_t k) {
    const int lane = lane_id();
    for (int base = a; base < b && cnt < k; base += kWave) {
        const int s = base + lane;
        const bool hit = s < b && noise_dist(t, d.dsort[po + s]) <= r2;
        cnt += static_cast<uint32_t>(__popcll(__ballot(hit)));
    }
    return cnt;
}

__global__ __launch_bounds__(256) void k_noise_search(DevBuf d, Layout L) {
    int tile, f;
    xcd_tile_frame(tile, f);
    const FrameDesc& fd = d.frames[f];
    const int n = fd.n_raw;
    const int nhard = static_cast<int>(d.fs[f * kFsCount + kFsNoiseHard]);
    const int lane = lane_id();
    const int waves = gridDim.x * (blockDim.x / kWave);
    const NoiseConst nc = *d.nc;
    const size_t po = (size_t)f * L.N;
    const size_t ho = (size_t)f * L.hash_cap;
    const uint32_t mask = static_cast<uint32_t>(L.hash_cap) - 1;
    const uint32_t k = nc.k;
    const uint32_t* occ = d.nocc + (size_t)f * kNoiseOccWords;
    uint32_t noise_here = 0;
    for (int q = tile * (blockDim.x / kWave) + threadIdx.x / kWave; q < nhard; q += waves) {
        const int pos = d.vslot[po + q];
        const int i = d.nloc[po + q];
        const uint32_t carry = static_cast<uint32_t>(d.obs_src[po + q]);
        const float4 t = d.dsort[po + pos];
        const float r2 = noise_radius_sqr(nc, t.x, t.y);
        const float r = sqrtf(r2);
        uint32_t cnt;
        if (r < static_cast<float>(kNoiseMaxExtent - 1)) {
            cnt = carry & 0xffu;
            const int h0 = __float_as_int(t.w);
            const HashSlot hv = d.hslot[ho + h0];
            const int s0 = hv.aux;
            const int s1 = s0 + static_cast<int>(hv.count);
            const int up = pos + static_cast<int>((carry >> 8) & 0xffu);
            const int dn = pos - 1 - static_cast<int>((carry >> 16) & 0xffu);
            cnt = wave_count_range(d, po, t, r2, up, s1, cnt, k);
            cnt = wave_count_range(d, po, t, r2, s0, dn + 1, cnt, k);
            const int e = static_cast<int>(floorf(r + kNoiseCellSlack)) + 1;
            int cx, cy, cz;
            noise_cell(t.x, t.y, t.z, cx, cy, cz);
            // Position inside the own cell (exact for in-range coordinates;
            // clamped cells give values outside [0, 1), which only make the
            // cull more permissive).  The cull keeps every cell within
            // (r + slack): the float test admits a few ulps beyond r.
            const float fx = t.x - floorf(t.x), fy = t.y - floorf(t.y), fz = t.z - floorf(t.z);
            const float reach = (r + kNoiseCellSlack) * (r + kNoiseCellSlack);
            for (int sh = 1; sh <= e && cnt < k; ++sh) {
                const int wd = 2 * sh + 1;
                const int cube = wd * wd * wd;
                for (int c0 = 0; c0 < cube && cnt < k; c0 += kWave) {
                    const int ci = c0 + lane;
                    const int dz = ci / (wd * wd) - sh, dy = (ci / wd) % wd - sh, dx = ci % wd - sh;
                    const int ax = cx + dx, ay = cy + dy, az = cz + dz;
                    int32_t h = -1;
                    if (ci < cube && max(abs(dx), max(abs(dy), abs(dz))) == sh &&
                        cell_gap_sqr(fx, fy, fz, dx, dy, dz) <= reach && ax >= 0 && ax <= 1023 && ay >= 0 &&
                        ay <= 1023 && az >= 0 && az <= 2047 && ((occ[ay * 32 + (ax >> 5)] >> (ax & 31)) & 1u))
                        h = noise_find(d.hslot + ho, mask, noise_key(ax, ay, az));
                    // The points of all cells found in this batch, 64 at a
                    // time: point g belongs to the lane whose prefix range
                    // holds it (binary search over the lanes' inclusive sums).
                    HashSlot hv{};
                    if (h >= 0) hv = d.hslot[ho + h];
                    const int c_n = h >= 0 ? static_cast<int>(hv.count) : 0;
                    const int c_a = h >= 0 ? hv.aux : 0;
                    const int incl = wave_inclusive_sum(c_n);
                    const int found = __shfl(incl, kWave - 1, kWave);
                    const int start = c_a - (incl - c_n);  // point g of this lane's cell: start + g
                    for (int base = 0; base < found && cnt < k; base += kWave) {
                        const int g = base + lane;
                        int lo = 0, hi = kWave - 1;
#pragma unroll
                        for (int it = 0; it < 6; ++it) {
                            const int mid = (lo + hi) >> 1;
                            if (__shfl(incl, mid, kWave) > g)
                                hi = mid;
                            else
                                lo = mid + 1;
                        }
                        const int sp = __shfl(start, lo, kWave) + g;
                        const bool hit = g < found && noise_dist(t, d.dsort[po + sp]) <= r2;
                        cnt += static_cast<uint32_t>(__popcll(__ballot(hit)));
                    }
                }
            }
        } else {
            // Radius beyond the cell cube (range > ~150 m) or infinite.
            cnt = wave_count_range(d, po, t, r2, 0, n, 0u, k);
        }
        if (lane == 0) {
            const bool noise = cnt < k;
            d.noise[po + i] = noise ? 1 : 0;
            noise_here += noise ? 1u : 0u;
        }
    }
    if (lane == 0 && noise_here) atomicAdd(&d.fs[f * kFsCount + kFsNoise], noise_here);
}

// Stable compaction of the kept points (label 0) into the segmenter's
// inputs, in two passes over tiles of 2048 points (eight rows of 256, each
// read and written with coalesced accesses): the kept count of every tile,
// then each tile's compaction after the counts of the tiles before it, which
// moves x, y, z, ring to their places.  The keep flags are the label bytes;
// the last tile writes the kept count, which becomes the frame's point count.
constexpr int kKeepThreads = 256;
constexpr int kKeepRows = 8;
constexpr int kKeepTile = kKeepThreads * kKeepRows;

__device__ __forceinline__ uint32_t keep_bits(const uint8_t* lab, int base, int n) {
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < kKeepRows; ++k) {
        const int i = base + k * kKeepThreads + static_cast<int>(threadIdx.x);
        m |= (i < n && lab[i] == 0) ? (1u << k) : 0u;
    }
    return m;
}

__global__ __launch_bounds__(kKeepThreads) void k_keep_count(DevBuf d, Layout L) {
    __shared__ uint32_t lds[kKeepThreads / kWave + 1];
    const int f = blockIdx.y;
    const int n = d.frames[f].n_raw;
    const int base = blockIdx.x * kKeepTile;
    if (base >= n) return;
    const uint32_t m = keep_bits(d.noise + (size_t)f * L.N, base, n);
    uint32_t total;
    block_exclusive_sum<kKeepThreads>(static_cast<uint32_t>(__popc(m)), lds, &total);
    if (threadIdx.x == 0) d.blk[(size_t)f * L.blk_stride + blockIdx.x] = total;
}

__global__ __launch_bounds__(kKeepThreads) void k_keep_move(DevBuf d, Layout L) {
    __shared__ uint32_t lds[kKeepRows * kKeepThreads / kWave + 1];
    const int f = blockIdx.y;
    FrameDesc& fd = d.frames[f];
    const int n = fd.n_raw;
    const int tile = blockIdx.x;
    const int base = tile * kKeepTile;
    if (base >= n) {
        if (tile == 0 && threadIdx.x == 0) {
            d.fs[f * kFsCount + kFsKept] = 0;
            fd.n = 0;
        }
        return;
    }
    const uint32_t* blk = d.blk + (size_t)f * L.blk_stride;
    uint32_t pre = 0;
    for (int i = threadIdx.x; i < tile; i += kKeepThreads) pre += blk[i];
    uint32_t pre_total;
    block_exclusive_sum<kKeepThreads>(pre, lds, &pre_total);
    const uint32_t m = keep_bits(d.noise + (size_t)f * L.N, base, n);
    uint32_t pos[kKeepRows];
    const uint32_t total = tile_compact<kKeepThreads, kKeepRows>(m, pos, lds);
    const int off = fd.pt_off;
    const bool ring = fd.has_ring != 0;
#pragma unroll
    for (int k = 0; k < kKeepRows; ++k) {
        if ((m >> k) & 1u) {
            const int gi = off + base + k * kKeepThreads + static_cast<int>(threadIdx.x);
            const int go = off + static_cast<int>(pre_total + pos[k]);
            d.x[go] = d.rx[gi];
            d.y[go] = d.ry[gi];
            d.z[go] = d.rz[gi];
            if (ring) d.ring[go] = d.rring[gi];
        }
    }
    if (threadIdx.x == 0 && base + kKeepTile >= n) {
        const uint32_t kept = pre_total + total;
        d.fs[f * kFsCount + kFsKept] = kept;
        fd.n = static_cast<int32_t>(kept);
    }
}

}  // namespace

void launch_noise(lp_context* c, int B, bool compact) {
    DevBuf& d = c->d;
    const Layout& L = c->L;
    hipStream_t st = c->stream;
    const int nb = (L.N + 255) / 256;
    stage_mark(c, "noise");
    hipLaunchKernelGGL(k_noise_clear, dim3(256, B), dim3(256), 0, st, d, L);
    hipLaunchKernelGGL(k_noise_prep, dim3((L.N + kPrepTile - 1) / kPrepTile, B), dim3(256), 0, st, d, L);
    // Cell starts (exclusive scan of the counts by dense id) straight into
    // the slots' aux fields.
    launch_scan_gather_into(c, B, &d.hslot[0].count, L.hash_cap, 4, d.ncell_slot,
                            reinterpret_cast<uint32_t*>(&d.hslot[0].aux), L.N, kCountSlot, kFsNoiseCells);
    hipLaunchKernelGGL(k_noise_scatter, dim3(nb, B), dim3(256), 0, st, d, L);
    stage_mark(c, "noise_count");
    // The open list is a few percent of the points: a quarter of the point
    // grid, striding over the list.
    hipLaunchKernelGGL(k_noise_count, dim3((nb + 3) / 4, B), dim3(256), 0, st, d, L);
    stage_mark(c, "noise_search");
    // About 8192 wavefronts in all (8 per SIMD: the search is latency-bound),
    // at least 128 and at most 1024 per frame.
    const int sb = std::min(256, std::max(32, 2048 / B));
    hipLaunchKernelGGL(k_noise_search, dim3(sb, B), dim3(256), 0, st, d, L);
    launch_hash_reset(c, B, d.ncell_slot, kFsNoiseCells);
    stage_mark(c, "noise_compact");
    if (compact) launch_noise_compact(c, B);
}

void launch_noise_compact(lp_context* c, int B) {
    const int nt = (c->L.N + kKeepTile - 1) / kKeepTile;  // at most blk_stride tiles
    hipLaunchKernelGGL(k_keep_count, dim3(nt, B), dim3(kKeepThreads), 0, c->stream, c->d, c->L);
    hipLaunchKernelGGL(k_keep_move, dim3(nt, B), dim3(kKeepThreads), 0, c->stream, c->d, c->L);
}

}  // namespace lp
