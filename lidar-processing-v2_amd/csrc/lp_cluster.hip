// lp_cluster.hip -- curved-voxel clustering on gfx950.
//
// Reference: lidar_processing_lib/src/clusterer.cpp:55-239.  The reference
// hashes (range, azimuth, elevation) voxel keys into a linear-probing map
// (static_unordered_map.hpp) and labels voxels by BFS from each unlabelled
// point in input order over the 26-neighbourhood (range and elevation
// bounded, azimuth wrapping by one), numbering components in the order of
// their first point; removeSmallClusters then drops components with fewer
// than min_cluster_size points and renumbers the rest in order.  The result
// is therefore: label = rank, among surviving components, of the
// component's minimum point index.
//
// Here: voxel keys go into an open-addressing HBM hash set by atomicCAS
// (one dense id per occupied voxel, per-voxel point count and minimum point
// index by atomics on the voxel's own slot), components come from a
// lock-free union-find over dense ids held in LDS (one workgroup per frame;
// a global-memory union-find takes over above kCcLds voxels), and labels
// from a scan over "first point of a surviving component" flags.
#include "lp_kernels.h"
#include "lp_queue.h"

namespace lp {

using namespace lpd;

constexpr float kTwoPiF = 6.28318548202514648438f;   // Clusterer::TWO_M_PIf
constexpr float kHalfPiF = 1.57079637050628662109f;  // M_PI_2f
constexpr int kCcLds = 32768;                        // voxels per frame for the LDS union-find

// Obstacle split: input-order compaction of OBSTACLE points
// (processor.cpp:562-579), in two passes over tiles of 2048 points (eight
// rows of 256, coalesced) straight from the labels: the obstacle count of
// every tile, then each tile's compaction after the counts of the tiles
// before it, which moves the obstacles' index and x, y, z.  cl_flag receives
// every point's exclusive obstacle count (the split into ground and unknown
// clouds reads it); the last tile writes the obstacle count.
constexpr int kObsThreads = 256;
constexpr int kObsRows = 8;
constexpr int kObsTile = kObsThreads * kObsRows;

__device__ __forceinline__ uint32_t obstacle_bits(const uint32_t* lab, int base, int n) {
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < kObsRows; ++k) {
        const int i = base + k * kObsThreads + static_cast<int>(threadIdx.x);
        m |= (i < n && lab[i] == LP_LABEL_OBSTACLE) ? (1u << k) : 0u;
    }
    return m;
}

__global__ __launch_bounds__(kObsThreads) void k_obs_count(DevBuf d, Layout L) {
    __shared__ uint32_t lds[kObsThreads / kWave + 1];
    const int f = blockIdx.y;
    const FrameDesc& fd = d.frames[f];
    const int n = fd.n;
    const int base = blockIdx.x * kObsTile;
    if (base >= n) return;
    const uint32_t m = obstacle_bits(d.labels + fd.pt_off, base, n);
    uint32_t total;
    block_exclusive_sum<kObsThreads>(static_cast<uint32_t>(__popc(m)), lds, &total);
    if (threadIdx.x == 0) d.blk[(size_t)f * L.blk_stride + blockIdx.x] = total;
}

__global__ __launch_bounds__(kObsThreads) void k_obs_gather(DevBuf d, Layout L) {
    __shared__ uint32_t lds[kObsRows * kObsThreads / kWave + 1];
    const int f = blockIdx.y;
    const FrameDesc& fd = d.frames[f];
    const int n = fd.n;
    const int tile = blockIdx.x;
    const int base = tile * kObsTile;
    if (base >= n) {
        if (tile == 0 && threadIdx.x == 0) d.fs[f * kFsCount + kFsObstacles] = 0;
        return;
    }
    const uint32_t* blk = d.blk + (size_t)f * L.blk_stride;
    uint32_t pre = 0;
    for (int i = threadIdx.x; i < tile; i += kObsThreads) pre += blk[i];
    uint32_t pre_total;
    block_exclusive_sum<kObsThreads>(pre, lds, &pre_total);
    const uint32_t m = obstacle_bits(d.labels + fd.pt_off, base, n);
    uint32_t pos[kObsRows];
    const uint32_t total = tile_compact<kObsThreads, kObsRows>(m, pos, lds);
    const size_t ob = (size_t)f * L.N;
#pragma unroll
    for (int k = 0; k < kObsRows; ++k) {
        const int i = base + k * kObsThreads + static_cast<int>(threadIdx.x);
        const uint32_t o = pre_total + pos[k];
        if (i < n) d.cl_flag[ob + i] = o;
        if ((m >> k) & 1u) {
            const int gi = fd.pt_off + i;
            d.obs_src[ob + o] = i;
            d.ox[ob + o] = d.x[gi];
            d.oy[ob + o] = d.y[gi];
            d.oz[ob + o] = d.z[gi];
        }
    }
    if (threadIdx.x == 0 && base + kObsTile >= n) d.fs[f * kFsCount + kFsObstacles] = pre_total + total;
}

// Ground and unsegmented clouds of the same split (processor.cpp:565-577):
// with gnd_flag scanned into obs_flag and cl_flag still the obstacle scan,
// an UNKNOWN point's position among the unknowns is its index less the
// ground and obstacle points before it.  split_idx holds the ground indices
// followed by the unknown ones.
__global__ void k_label_split(DevBuf d, Layout L) {
    const int f = blockIdx.y;
    const FrameDesc& fd = d.frames[f];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fd.n) return;
    const uint32_t lab = d.labels[fd.pt_off + i];
    if (lab == LP_LABEL_OBSTACLE) return;
    const size_t ob = (size_t)f * L.N;
    const uint32_t g = d.obs_flag[ob + i];
    if (lab == LP_LABEL_GROUND) {
        d.split_idx[ob + g] = i;
    } else {
        const uint32_t ng = d.fs[f * kFsCount + kFsGround];
        d.split_idx[ob + ng + (i - g - d.cl_flag[ob + i])] = i;
    }
}

// Standalone Clusterer::cluster input: the staged cloud is the obstacle cloud.
__global__ void k_cluster_input(DevBuf d, Layout L) {
    const int f = blockIdx.y;
    const FrameDesc& fd = d.frames[f];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) d.fs[f * kFsCount + kFsObstacles] = static_cast<uint32_t>(fd.n);
    if (i >= fd.n) return;
    const size_t ob = (size_t)f * L.N;
    const int gi = fd.pt_off + i;
    d.obs_src[ob + i] = i;
    d.ox[ob + i] = d.x[gi];
    d.oy[ob + i] = d.y[gi];
    d.oz[ob + i] = d.z[gi];
}

__device__ __forceinline__ uint32_t max_bits(float v) {
    // Non-negative floats order like their bit patterns; -0 counts as +0
    // (only ceilf(max / res) is used, which is the same for both).
    const uint32_t u = f2u(v);
    return u == 0x80000000u ? 0u : u;
}

// cartesianToSpherical (clusterer.cpp:55-100).
__device__ __forceinline__ void spherical_tile(const DevBuf& d, const Layout& L, int f, int tile) {
    const int nobs = static_cast<int>(d.fs[f * kFsCount + kFsObstacles]);
    const int k = tile * blockDim.x + threadIdx.x;
    if (tile * blockDim.x >= static_cast<unsigned>(nobs)) return;
    const size_t ob = (size_t)f * L.N;
    uint32_t mr = 0, ma = 0, me = 0;
    if (k < nobs) {
        const float x = d.ox[ob + k], y = d.oy[ob + k], z = d.oz[ob + k];
        float az = fdlibm_atan2f(y, x);
        az = (az < 0) ? (az + kTwoPiF) : az;
        const float dsq = x * x + y * y;
        const float dist = sqrtf(dsq);
        const float range = sqrtf(dsq + z * z);
        const float el = fdlibm_atanf(z / dist) + kHalfPiF;
        d.sr[ob + k] = range;
        d.sa[ob + k] = az;
        d.se[ob + k] = el;
        mr = max_bits(range);
        ma = max_bits(az);
        me = max_bits(el);
    }
    // The frame maxima: wavefront maxima (DPP inside 16-lane rows, then two
    // shuffles), the workgroup's in LDS, one atomic per workgroup and value
    // (per wavefront, the per-frame counters were a contention point).
    __shared__ uint32_t red[3][1024 / kWave];
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        if (o < 16) {
            mr = max(mr, lane_xor(mr, o));
            ma = max(ma, lane_xor(ma, o));
            me = max(me, lane_xor(me, o));
        } else {
            mr = max(mr, static_cast<uint32_t>(__shfl_xor(mr, o, kWave)));
            ma = max(ma, static_cast<uint32_t>(__shfl_xor(ma, o, kWave)));
            me = max(me, static_cast<uint32_t>(__shfl_xor(me, o, kWave)));
        }
    }
    const int w = static_cast<int>(threadIdx.x) / kWave;
    if (lane_id() == 0) {
        red[0][w] = mr;
        red[1][w] = ma;
        red[2][w] = me;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < static_cast<int>(blockDim.x) / kWave; ++q) {
            mr = max(mr, red[0][q]);
            ma = max(ma, red[1][q]);
            me = max(me, red[2][q]);
        }
        atomicMax(&d.fs[f * kFsCount + kFsMaxR], mr);
        atomicMax(&d.fs[f * kFsCount + kFsMaxA], ma);
        atomicMax(&d.fs[f * kFsCount + kFsMaxE], me);
    }
    __syncthreads();  // red is reused by the workgroup's next tile
}

__global__ void k_spherical(DevBuf d, Layout L) { spherical_tile(d, L, blockIdx.y, blockIdx.x); }

__global__ __launch_bounds__(256) void k_spherical_q(DevBuf d, Layout L, int B) {
    wq_run(d.wq + kWqSpherical * 8, B, blockDim.x, [&](int f) { return d.fs[f * kFsCount + kFsObstacles]; },
           [&](int f, int t) { spherical_tile(d, L, f, t); });
}

struct Dims {
    int32_t nr, na, ne;
};

// num_range / num_azimuth / num_elevation (clusterer.cpp:94-99).
__device__ __forceinline__ Dims dims_of(const DevBuf& d, int f, const CluConst& cc) {
    const uint32_t* fs = d.fs + f * kFsCount;
    Dims dm;
    dm.nr = static_cast<int32_t>(ceilf(u2f(fs[kFsMaxR]) / cc.res_r) + 1);
    dm.na = static_cast<int32_t>(ceilf(u2f(fs[kFsMaxA]) / cc.res_az) + 1);
    dm.ne = static_cast<int32_t>(ceilf(u2f(fs[kFsMaxE]) / cc.res_el) + 1);
    return dm;
}

__device__ __forceinline__ uint32_t hash_slot(int32_t key, uint32_t mask) {
    uint32_t h = static_cast<uint32_t>(key) * 0x9E3779B1u;
    h ^= h >> 15;
    return h & mask;
}

// Voxel keys (clusterer.cpp:102-120, flatVoxelIndex clusterer.hpp:135-141)
// inserted into the per-frame hash set.
constexpr int kVoxRows = 2;  // obstacles per thread: rows of blockDim.x, their atomics in flight together

// Obstacles tile * kVoxRows * blockDim.x + s * blockDim.x + threadIdx.x for
// the rows s: lanes of a wavefront hold consecutive obstacles.
__device__ __forceinline__ void voxel_insert_tile(const DevBuf& d, const Layout& L, int f, int tile) {
    const int nobs = static_cast<int>(d.fs[f * kFsCount + kFsObstacles]);
    const int base = tile * kVoxRows * static_cast<int>(blockDim.x);
    if (base >= nobs) return;
    const CluConst cc = *d.cc;
    const Dims dm = dims_of(d, f, cc);
    const size_t ob = (size_t)f * L.N;
    const uint32_t mask = static_cast<uint32_t>(L.hash_cap) - 1;
    const size_t ho = (size_t)f * L.hash_cap;
    const int lane = lane_id();
    int k[kVoxRows];
    int32_t key[kVoxRows];
#pragma unroll
    for (int s = 0; s < kVoxRows; ++s) {
        k[s] = base + s * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
        key[s] = -2;  // past the obstacles: never a voxel key (keys are non-negative), inserts nothing
        if (k[s] < nobs) {
            const int32_t ri = static_cast<int32_t>(d.sr[ob + k[s]] / cc.res_r);
            const int32_t ai = static_cast<int32_t>(d.sa[ob + k[s]] / cc.res_az);
            const int32_t ei = static_cast<int32_t>(d.se[ob + k[s]] / cc.res_el);
            key[s] = dm.nr * (dm.na * ei + ai) + ri;
        }
    }
    // A run of lanes with the same voxel inserts once: its lowest lane (which
    // holds the run's smallest point index, lanes follow point order) probes
    // and records the run's count and minimum index; all runs of both rows
    // at once.
    WaveRun run[kVoxRows];
    bool lead[kVoxRows];
    uint32_t h[kVoxRows];
    int32_t prev[kVoxRows];
#pragma unroll
    for (int s = 0; s < kVoxRows; ++s) {
        run[s] = wave_runs(key[s]);
        lead[s] = lane == run[s].leader && key[s] >= 0;
        h[s] = hash_slot(key[s], mask);
        prev[s] = lead[s] ? atomicCAS(&d.hslot[ho + h[s]].key, -1, key[s]) : key[s];
    }
#pragma unroll
    for (int s = 0; s < kVoxRows; ++s) {
        for (uint32_t probe = 1; probe <= mask && prev[s] != -1 && prev[s] != key[s]; ++probe) {
            h[s] = (h[s] + 1) & mask;
            prev[s] = atomicCAS(&d.hslot[ho + h[s]].key, -1, key[s]);
        }
    }
    bool fresh[kVoxRows];
#pragma unroll
    for (int s = 0; s < kVoxRows; ++s) {
        fresh[s] = lead[s] && prev[s] == -1;
        if (lead[s]) {
            atomicMin(&d.hslot[ho + h[s]].aux, k[s]);
            atomicAdd(&d.hslot[ho + h[s]].count, run[s].count);
        }
    }
    // Dense ids of the new voxels, in any order.
    uint32_t pos[kVoxRows];
    __shared__ uint32_t rsv[1024 / kWave + 1];
    block_reserve<kVoxRows>(&d.fs[f * kFsCount + kFsVoxels], fresh, pos, rsv);
#pragma unroll
    for (int s = 0; s < kVoxRows; ++s) {
        if (fresh[s]) {
            if (pos[s] < static_cast<uint32_t>(L.N)) {
                d.vlist[ob + pos[s]] = static_cast<int32_t>(h[s]);
                d.vpar[ob + pos[s]] = static_cast<int32_t>(pos[s]);
                d.vmin[ob + pos[s]] = 0x7fffffff;
                d.vcnt[ob + pos[s]] = 0;
            }
            d.hslot[ho + h[s]].dense = static_cast<int32_t>(pos[s]);
        }
        const int32_t slot = static_cast<int32_t>(__shfl(h[s], run[s].leader, kWave));
        if (k[s] < nobs) d.vslot[ob + k[s]] = slot;
    }
}

__global__ void k_voxel_insert(DevBuf d, Layout L) {
    int tile, f;
    xcd_tile_frame(tile, f);
    voxel_insert_tile(d, L, f, tile);
}

__global__ __launch_bounds__(256) void k_voxel_insert_q(DevBuf d, Layout L, int B) {
    wq_run(d.wq + kWqVoxelInsert * 8, B, kVoxRows * blockDim.x,
           [&](int f) { return d.fs[f * kFsCount + kFsObstacles]; },
           [&](int f, int t) { voxel_insert_tile(d, L, f, t); });
}


// Half of the 26-neighbourhood (the other half is covered from the
// neighbour's side): offsets (dr, da, de) with (de, da, dr) > (0, 0, 0).
__constant__ static const int8_t kHalfNb[13][3] = {
    {1, 0, 0},  {-1, 1, 0}, {0, 1, 0},  {1, 1, 0},  {-1, -1, 1}, {0, -1, 1}, {1, -1, 1},
    {-1, 0, 1}, {0, 0, 1},  {1, 0, 1},  {-1, 1, 1}, {0, 1, 1},   {1, 1, 1}};

// ---- union-find in LDS ----------------------------------------------------
__device__ __forceinline__ int32_t lds_find(volatile int32_t* par, int32_t x) {
    for (;;) {
        const int32_t p = par[x];
        if (p == x) return x;
        const int32_t gp = par[p];
        if (gp == p) return p;
        par[x] = gp;  // path halving; x is not a root, so no CAS targets it
        x = gp;
    }
}

__device__ __forceinline__ void lds_union(int32_t* par, int32_t a, int32_t b) {
    for (;;) {
        a = lds_find(par, a);
        b = lds_find(par, b);
        if (a == b) return;
        if (a > b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(par + b, b, a) == b) return;  // hook the larger root under the smaller
    }
}

// Forward neighbours of every voxel, looked up in the hash by the whole grid
// (the lookups are independent; the union-find below then only reads them).
__global__ void k_voxel_neighbours(DevBuf d, Layout L) {
    int tile, f;
    xcd_tile_frame(tile, f);
    // Voxel capacity (the reference's StaticUnorderedMap throws past it).
    if (tile == 0 && threadIdx.x == 0 && static_cast<int32_t>(d.fs[f * kFsCount + kFsVoxels]) > d.cc->max_voxels)
        atomicOr(&d.fs[f * kFsCount + kFsErr], kErrVoxelOverflow);
    const uint32_t U = min(d.fs[f * kFsCount + kFsVoxels], static_cast<uint32_t>(L.N));
    const uint32_t v = tile * blockDim.x + threadIdx.x;
    if (v >= U) return;
    const CluConst cc = *d.cc;
    const Dims dm = dims_of(d, f, cc);
    const size_t ob = (size_t)f * L.N;
    const size_t ho = (size_t)f * L.hash_cap;
    const uint32_t mask = static_cast<uint32_t>(L.hash_cap) - 1;
    const int32_t key = d.hslot[ho + d.vlist[ob + v]].key;
    int32_t* nb = d.vnb + ob * 13;
    // Neighbour q of the voxel (clusterer.cpp:150-176 bounds: range and
    // elevation in [0, num), azimuth wraps by one step) is its dense id, or -1
    // if that voxel is empty.  All 13 first probes are issued before any is
    // waited on (collisions, rare at the table's load, probe on afterwards).
    const int32_t ri = key % dm.nr;
    const int32_t rest = key / dm.nr;
    const int32_t ai = rest % dm.na;
    const int32_t ei = rest / dm.na;
    int32_t nk[13];
    uint32_t h[13];
    HashSlot rec[13];
#pragma unroll
    for (int q = 0; q < 13; ++q) {
        const int32_t rr = ri + kHalfNb[q][0];
        const int32_t ee = ei + kHalfNb[q][2];
        int32_t aa = ai + kHalfNb[q][1];
        if (aa < 0)
            aa += dm.na;
        else if (aa >= dm.na)
            aa -= dm.na;
        nk[q] = (rr < 0 || rr >= dm.nr || ee < 0 || ee >= dm.ne) ? -1 : dm.nr * (dm.na * ee + aa) + rr;
        h[q] = hash_slot(nk[q], mask);
        rec[q] = HashSlot{-1, 0u, -1, 0};
        if (nk[q] >= 0) rec[q] = d.hslot[ho + h[q]];
    }
#pragma unroll
    for (int q = 0; q < 13; ++q) {
        for (uint32_t probe = 1; probe <= mask && nk[q] >= 0 && rec[q].key != nk[q] && rec[q].key != -1; ++probe) {
            h[q] = (h[q] + 1) & mask;
            rec[q] = d.hslot[ho + h[q]];
        }
        nb[(size_t)q * L.N + v] = (nk[q] >= 0 && rec[q].key == nk[q]) ? rec[q].dense : -1;
    }
}

__global__ __launch_bounds__(1024) void k_voxel_cc(DevBuf d, Layout L) {
    __shared__ int32_t par[kCcLds];
    const int f = blockIdx.x;
    const int tid = threadIdx.x;
    const uint32_t U = min(d.fs[f * kFsCount + kFsVoxels], static_cast<uint32_t>(L.N));
    if (U > static_cast<uint32_t>(kCcLds)) {
        if (tid == 0) d.fs[f * kFsCount + kFsCcGlobal] = 1;
        return;
    }
    const size_t ob = (size_t)f * L.N;
    const int32_t* nb = d.vnb + ob * 13;
    for (uint32_t v = tid; v < U; v += 1024) par[v] = static_cast<int32_t>(v);
    __syncthreads();
    for (uint32_t v = tid; v < U; v += 1024) {
        int32_t u[13];
#pragma unroll
        for (int q = 0; q < 13; ++q) u[q] = nb[(size_t)q * L.N + v];
#pragma unroll
        for (int q = 0; q < 13; ++q)
            if (u[q] >= 0) lds_union(par, static_cast<int32_t>(v), u[q]);
    }
    __syncthreads();
    for (uint32_t v = tid; v < U; v += 1024) d.vpar[ob + v] = lds_find(par, static_cast<int32_t>(v));
}

// ---- union-find in global memory (frames with more than kCcLds voxels) ----
__device__ __forceinline__ int32_t load_parent(const int32_t* p, int32_t x) {
    return __hip_atomic_load(p + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int32_t g_find(int32_t* parent, int32_t x) {
    int32_t p = load_parent(parent, x);
    while (p != x) {
        const int32_t gp = load_parent(parent, p);
        if (gp != p) atomicCAS(parent + x, p, gp);
        x = p;
        p = gp;
    }
    return x;
}

__global__ void k_voxel_cc_global(DevBuf d, Layout L) {
    const int f = blockIdx.y;
    if (!d.fs[f * kFsCount + kFsCcGlobal]) return;
    const uint32_t U = min(d.fs[f * kFsCount + kFsVoxels], static_cast<uint32_t>(L.N));
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= U) return;
    const size_t ob = (size_t)f * L.N;
    int32_t* par = d.vpar + ob;
    const int32_t* nb = d.vnb + ob * 13;
    for (int q = 0; q < 13; ++q) {
        const int32_t u = nb[(size_t)q * L.N + v];
        if (u < 0) continue;
        int32_t a = static_cast<int32_t>(v), b = u;
        for (;;) {
            a = g_find(par, a);
            b = g_find(par, b);
            if (a == b) break;
            if (a > b) {
                const int32_t t = a;
                a = b;
                b = t;
            }
            if (atomicCAS(par + b, b, a) == b) break;
        }
    }
}

// Component minimum point index and point count (removeSmallClusters counts
// points, clusterer.cpp:214-220): one atomic per voxel.  After the
// global-memory union-find every voxel also compresses its own parent to the
// root (all unions are done, so concurrent finds only shorten paths within
// one tree); the LDS union-find leaves the parents compressed already.
__global__ void k_voxel_stats(DevBuf d, Layout L) {
    const int f = blockIdx.y;
    const uint32_t U = min(d.fs[f * kFsCount + kFsVoxels], static_cast<uint32_t>(L.N));
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= U) return;
    const size_t ob = (size_t)f * L.N;
    const size_t ho = (size_t)f * L.hash_cap;
    int32_t r = d.vpar[ob + v];
    if (d.fs[f * kFsCount + kFsCcGlobal]) {
        r = g_find(d.vpar + ob, static_cast<int32_t>(v));
        d.vpar[ob + v] = r;
    }
    const int32_t h = d.vlist[ob + v];
    const HashSlot hv = d.hslot[ho + h];
    // Voxels numbered one after the other mostly share a component: each run
    // of equal roots in the wavefront reduces its minimum and count first
    // (a segmented reduction down the run) and its leader adds them, so a
    // large component's root takes one atomic pair per run, not per voxel.
    const WaveRun run = wave_runs(r);
    const int lane = lane_id();
    const int end = run.leader + static_cast<int>(run.count);  // one past the run's last lane
    int32_t mn = hv.aux;
    uint32_t cnt = hv.count;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int32_t omn = __shfl_down(mn, o, kWave);
        const uint32_t ocnt = __shfl_down(cnt, o, kWave);
        if (lane + o < end) {
            mn = min(mn, omn);
            cnt += ocnt;
        }
    }
    if (lane == run.leader) {
        atomicMin(&d.vmin[ob + r], mn);
        atomicAdd(&d.vcnt[ob + r], cnt);
    }
}

__device__ __forceinline__ int32_t root_of_point(const DevBuf& d, size_t ob, size_t ho, int k) {
    return d.vpar[ob + d.hslot[ho + d.vslot[ob + k]].dense];
}

// Clusters in the order of their first point (the reference numbers them in
// input order): a component is a cluster when it has at least min_size
// points, and its first point is its minimum obstacle index.  Two passes over
// tiles of 2048 obstacles (eight rows of 256): the first points of every
// tile, then each tile's compaction after the counts of the tiles before it,
// which numbers the clusters, writes their root, size and an empty cursor,
// and gives each root its cluster label (vlab).  The last tile writes the
// cluster count.
constexpr int kCompThreads = 256;
constexpr int kCompRows = 8;
constexpr int kCompTile = kCompThreads * kCompRows;

__device__ __forceinline__ uint32_t first_point_bits(const DevBuf& d, size_t ob, size_t ho, int base, int nobs,
                                                     uint32_t min_size, int32_t (&root)[kCompRows]) {
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < kCompRows; ++k) {
        const int q = base + k * kCompThreads + static_cast<int>(threadIdx.x);
        root[k] = -1;
        if (q < nobs) {
            const int32_t r = root_of_point(d, ob, ho, q);
            root[k] = r;
            if (d.vmin[ob + r] == q && d.vcnt[ob + r] >= min_size) m |= 1u << k;
        }
    }
    return m;
}

__global__ __launch_bounds__(kCompThreads) void k_component_count(DevBuf d, Layout L) {
    __shared__ uint32_t lds[kCompThreads / kWave + 1];
    const int f = blockIdx.y;
    const int nobs = static_cast<int>(d.fs[f * kFsCount + kFsObstacles]);
    const int base = blockIdx.x * kCompTile;
    if (base >= nobs) return;
    const size_t ob = (size_t)f * L.N;
    const size_t ho = (size_t)f * L.hash_cap;
    int32_t root[kCompRows];
    const uint32_t m = first_point_bits(d, ob, ho, base, nobs, d.cc->min_size, root);
    uint32_t total;
    block_exclusive_sum<kCompThreads>(static_cast<uint32_t>(__popc(m)), lds, &total);
    if (threadIdx.x == 0) d.blk[(size_t)f * L.blk_stride + blockIdx.x] = total;
}

__global__ __launch_bounds__(kCompThreads) void k_component_number(DevBuf d, Layout L) {
    __shared__ uint32_t lds[kCompRows * kCompThreads / kWave + 1];
    const int f = blockIdx.y;
    const int nobs = static_cast<int>(d.fs[f * kFsCount + kFsObstacles]);
    const int tile = blockIdx.x;
    const int base = tile * kCompTile;
    if (base >= nobs) {
        if (tile == 0 && threadIdx.x == 0) d.fs[f * kFsCount + kFsClusters] = 0;
        return;
    }
    const uint32_t* blk = d.blk + (size_t)f * L.blk_stride;
    uint32_t pre = 0;
    for (int i = threadIdx.x; i < tile; i += kCompThreads) pre += blk[i];
    uint32_t pre_total;
    block_exclusive_sum<kCompThreads>(pre, lds, &pre_total);
    const size_t ob = (size_t)f * L.N;
    const size_t ho = (size_t)f * L.hash_cap;
    int32_t root[kCompRows];
    const uint32_t m = first_point_bits(d, ob, ho, base, nobs, d.cc->min_size, root);
    uint32_t pos[kCompRows];
    const uint32_t total = tile_compact<kCompThreads, kCompRows>(m, pos, lds);
#pragma unroll
    for (int k = 0; k < kCompRows; ++k) {
        if ((m >> k) & 1u) {
            const uint32_t lab = pre_total + pos[k];
            const int32_t r = root[k];
            d.cl_root[ob + lab] = r;
            d.cl_size[ob + lab] = d.vcnt[ob + r];
            d.cl_cursor[ob + lab] = 0;
            d.hc_count[ob + lab] = 0;  // the round-0 hull input count (first buffer)
            d.vlab[ob + r] = static_cast<int32_t>(lab);
        }
    }
    if (threadIdx.x == 0 && base + kCompTile >= nobs) d.fs[f * kFsCount + kFsClusters] = pre_total + total;
}

__global__ void k_component_label(DevBuf d, Layout L) {
    const int f = blockIdx.y;
    const int nobs = static_cast<int>(d.fs[f * kFsCount + kFsObstacles]);
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nobs) return;
    const size_t ob = (size_t)f * L.N;
    const size_t ho = (size_t)f * L.hash_cap;
    const int32_t r = root_of_point(d, ob, ho, k);
    d.clabel[ob + k] = d.vcnt[ob + r] >= d.cc->min_size ? d.vlab[ob + r] : -1;
}

// True when the obstacle at input index i (pixel pix of a frame W pixels
// wide) has, in the two range-image rows above it in its column, obstacles
// of its own cluster with exactly its x and y.  Vertical surfaces repeat one
// (x, y) down a column; such a point is never a hull vertex while a copy of
// it stays in the hull input, and down every column run of copies the top
// two never pass this test, so each run keeps two copies (a cluster of one
// repeated point keeps the two copies the reference returns for it).
__device__ __forceinline__ bool column_repeat(const DevBuf& d, const Layout& L, const FrameDesc& fd, size_t ob,
                                              int i, float x, float y, int32_t lab) {
    const int pix = d.pt_pix[fd.pt_off + i];
    const int W = fd.sc.W;
    if (pix < 2 * W) return false;
    const unsigned long long imask = (1ull << L.idx_bits) - 1;
#pragma unroll
    for (int dr = 1; dr <= 2; ++dr) {
        const unsigned long long key = d.pix_key[fd.pix_off + pix - dr * W];
        if (key == ~0ull) return false;
        const int j = static_cast<int>(key & imask);
        const int gj = fd.pt_off + j;
        if (d.labels[gj] != LP_LABEL_OBSTACLE || d.x[gj] != x || d.y[gj] != y) return false;
        if (d.clabel[ob + d.cl_flag[ob + j]] != lab) return false;
    }
    return true;
}

// Members of each cluster (unordered within the cluster) and the round-0
// hull input: the cluster's points as PointXY doubles (processor.cpp:647),
// less the column repeats (their own count, in the first hull buffer).
__global__ void k_cluster_scatter(DevBuf d, Layout L, int from_split) {
    const int f = blockIdx.y;
    const int nobs = static_cast<int>(d.fs[f * kFsCount + kFsObstacles]);
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nobs) return;
    const size_t ob = (size_t)f * L.N;
    const int32_t lab = d.clabel[ob + k];
    if (lab < 0) return;
    const uint32_t off = d.cl_off[ob + lab];
    const uint32_t pos = wave_atomic_inc(&d.cl_cursor[ob + lab]);
    d.cl_members[ob + off + pos] = k;
    const float x = d.ox[ob + k], y = d.oy[ob + k];
    if (from_split && column_repeat(d, L, d.frames[f], ob, d.obs_src[ob + k], x, y, lab)) return;
    const uint32_t hp = off + wave_atomic_inc(&d.hc_count[ob + lab]);
    d.hcx[ob + hp] = static_cast<double>(x);
    d.hcy[ob + hp] = static_cast<double>(y);
    d.hci[ob + hp] = k;
}

void launch_cluster(lp_context* c, int B, bool from_split) {
    DevBuf& d = c->d;
    const Layout& L = c->L;
    hipStream_t st = c->stream;
    const int nb = (L.N + 255) / 256;
    stage_mark(c, "split");
    if (from_split) {
        const int nt = (L.N + kObsTile - 1) / kObsTile;  // at most blk_stride tiles
        hipLaunchKernelGGL(k_obs_count, dim3(nt, B), dim3(kObsThreads), 0, st, d, L);
        hipLaunchKernelGGL(k_obs_gather, dim3(nt, B), dim3(kObsThreads), 0, st, d, L);
        if (c->clouds_on) {
            launch_scan(c, B, d.gnd_flag, d.obs_flag, L.N, kCountPoints, 0, kFsGround);
            hipLaunchKernelGGL(k_label_split, dim3(nb, B), dim3(256), 0, st, d, L);
        }
    } else {
        hipLaunchKernelGGL(k_cluster_input, dim3(nb, B), dim3(256), 0, st, d, L);
    }
    stage_mark(c, "voxels");
    const int wg = c->point_wg;
    if (c->persistent) {
        hipLaunchKernelGGL(k_spherical_q, dim3(kPersistentWorkgroups), dim3(wg), 0, st, d, L, B);
        hipLaunchKernelGGL(k_voxel_insert_q, dim3(kPersistentWorkgroups), dim3(wg), 0, st, d, L, B);
    } else {
        hipLaunchKernelGGL(k_spherical, dim3((L.N + wg - 1) / wg, B), dim3(wg), 0, st, d, L);
        hipLaunchKernelGGL(k_voxel_insert, dim3((L.N + kVoxRows * wg - 1) / (kVoxRows * wg), B), dim3(wg), 0, st, d,
                           L);
    }
    stage_mark(c, "components");
    hipLaunchKernelGGL(k_voxel_neighbours, dim3(nb, B), dim3(256), 0, st, d, L);
    hipLaunchKernelGGL(k_voxel_cc, dim3(B), dim3(1024), 0, st, d, L);
    hipLaunchKernelGGL(k_voxel_cc_global, dim3(nb, B), dim3(256), 0, st, d, L);
    hipLaunchKernelGGL(k_voxel_stats, dim3(nb, B), dim3(256), 0, st, d, L);
    const int ct = (L.N + kCompTile - 1) / kCompTile;  // at most blk_stride tiles
    hipLaunchKernelGGL(k_component_count, dim3(ct, B), dim3(kCompThreads), 0, st, d, L);
    hipLaunchKernelGGL(k_component_number, dim3(ct, B), dim3(kCompThreads), 0, st, d, L);
    hipLaunchKernelGGL(k_component_label, dim3(nb, B), dim3(256), 0, st, d, L);
    launch_scan_small(c, B, d.cl_size, d.cl_off, L.N, kCountSlot, kFsClusters, -1);
    hipLaunchKernelGGL(k_cluster_scatter, dim3(nb, B), dim3(256), 0, st, d, L, from_split ? 1 : 0);
    launch_hash_reset(c, B, d.vlist, kFsVoxels);
}

}  // namespace lp
