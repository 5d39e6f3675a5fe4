// lp_segment.hip -- polar grid, range-image winner, RECM and RANSAC on gfx950.
//
// Reference semantics (paths relative to the reference root):
//   constructPolarGrid   lidar_processing_lib/src/segmenter.cpp:103-204
//   RECM elevation map   lidar_processing_lib/src/segmenter.cpp:206-267
//   obstacle test        lidar_processing_lib/src/segmenter.cpp:269-283
//   RANSAC correction    lidar_processing_lib/src/segmenter.cpp:321-479
//   image transfer       lidar_processing_lib/src/segmenter.cpp:289-318
//
// Layout: points are CSR-packed per frame (FrameDesc::pt_off); per-frame
// arrays of cells / pixels / queue entries use fixed frame strides.  Grids
// are (tiles, frames) and sized for the batch capacity so one launch
// sequence (and one captured graph) serves every batch of a shape bucket.
#include "lp_kernels.h"
#include "lp_queue.h"

namespace lp {

using namespace lpd;

constexpr float kTwoPi = 6.28318548202514648438f;  // (float)(2*M_PI), segmenter.hpp:136
constexpr uint32_t kEmptyZ = 0x7fc00000u;           // NaN marks an empty cell

// ---------------------------------------------------------------------------
// Per-batch clear (fill kernels instead of memsets so the whole sequence can
// be captured in a graph).  The voxel hash is clean on entry (its previous
// users reset the slots they occupied).
__global__ void k_clear(DevBuf d, Layout L) {
    const int f = blockIdx.y;
    const FrameDesc fd = d.frames[f];
    const int cells = fd.sc.cells;
    const int P = fd.sc.H * fd.sc.W;
    const int stride = gridDim.x * blockDim.x;
    const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (f == 0 && t0 < kWqCounters) d.wq[t0] = 0;
    for (int c = t0; c < cells; c += stride) d.cell_count[(size_t)f * L.cell_stride + c] = 0;
    for (int p = t0; p < P; p += stride) d.pix_key[fd.pix_off + p] = ~0ull;
    for (int k = t0; k < kRansacIters; k += stride) d.hyp_count[f * kRansacIters + k] = 0;
    if (t0 < kFsKept) d.fs[f * kFsCount + t0] = 0;  // the noise filter's counts run ahead of this clear
}

// ---------------------------------------------------------------------------
// One thread per point: polar cell, range-image pixel, cell histogram and the
// pixel winner.  The winner of a pixel is the first point, in (cell, input)
// order, with the smallest x*x+y*y (segmenter.cpp:291-318, strict `<`).
// Equal depth implies equal radial cell, so the order reduces to
// (depth, slice, index): packed into one 64-bit key and resolved with a
// single atomicMin.
__device__ __forceinline__ void project_tile(const DevBuf& d, const Layout& L, int f, int tile) {
    const FrameDesc& fd = d.frames[f];
    const int i = tile * blockDim.x + threadIdx.x;
    if (i >= fd.n) return;
    const SegConst& s = fd.sc;
    const int gi = fd.pt_off + i;
    const float x = d.x[gi], y = d.y[gi], z = d.z[gi];
    int cell = -1, pix = -1;
    const bool finite = isfinite(x) && isfinite(y) && isfinite(z);
    if (finite && !(z < s.zlo || z > s.zhi)) {
        const float dd = x * x + y * y;
        const float dist = sqrtf(dd);
        const int radial = static_cast<int>(dist / s.spacing);
        if (!(dist < s.min_d || dist > s.max_d || radial >= s.rings)) {
            float az = atan2_approx(y, x);
            az = (az < 0) ? (az + kTwoPi) : az;
            const int slice = min(static_cast<int>(az / s.slice_res), s.slices - 1);
            int row = -1;
            if (s.organized && fd.has_ring) {
                const int h = d.ring[gi];
                if (h < s.H) row = h;
            } else {
                const float elev = fdlibm_atanf(z / dist);
                const int hs = static_cast<int>((elev - s.down_rad) / s.per_pixel);
                if (hs >= 0 && hs < s.H) row = hs;
            }
            if (row >= 0) {
                const int col = static_cast<uint16_t>((s.w_minus_1 * az) / kTwoPi);
                cell = slice * s.rings + radial;
                pix = row * s.W + col;
                d.pt_loc[gi] = static_cast<int32_t>(wave_atomic_inc(&d.cell_count[(size_t)f * L.cell_stride + cell]));
                const unsigned long long key = (static_cast<unsigned long long>(f2u(dd)) << 32) |
                                               (static_cast<unsigned long long>(slice) << L.idx_bits) |
                                               static_cast<unsigned long long>(i);
                atomicMin(&d.pix_key[fd.pix_off + pix], key);
            }
        }
    }
    d.pt_cell[gi] = cell;
    d.pt_pix[gi] = pix;
    d.pt_xyzc[gi] = make_float4(x, y, z, __int_as_float(cell));
}

__global__ void k_project(DevBuf d, Layout L) { project_tile(d, L, blockIdx.y, blockIdx.x); }

__global__ __launch_bounds__(256) void k_project_q(DevBuf d, Layout L, int B) {
    wq_run(d.wq + kWqProject * 8, B, blockDim.x, [&](int f) { return d.frames[f].n; },
           [&](int f, int t) { project_tile(d, L, f, t); });
}

// Exclusive scan of the cell histogram, one 1024-thread workgroup per frame.
// Cell starts (exclusive scan of the counts), the list of non-empty cells
// with their starts and counts for k_cell_sort, and the empty marker of every
// empty cell's minimum.
__global__ __launch_bounds__(1024) void k_cell_scan(DevBuf d, Layout L) {
    __shared__ uint64_t lds[1024 / kWave + 1];
    const int f = blockIdx.x;
    const int cells = d.frames[f].sc.cells;
    const size_t co = (size_t)f * L.cell_stride;
    const uint32_t* cnt = d.cell_count + co;
    uint32_t* st = d.cell_start + (size_t)f * (L.cell_stride + 1);
    uint4* list = d.cell_list + co;
    float* zmin = d.zmin + co;
    uint32_t carry = 0, used = 0;
    for (int base = 0; base < cells; base += 1024 * 8) {
        uint32_t v[8];
        uint32_t local = 0, nz = 0;
        const int b = base + threadIdx.x * 8;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            v[k] = (b + k < cells) ? cnt[b + k] : 0u;
            local += v[k];
            nz += v[k] ? 1u : 0u;
        }
        // Both sums in one block scan: the point count in the high word, the
        // non-empty cell count in the low one (each frame total < 2^32).
        uint64_t both_total;
        const uint64_t both = block_exclusive_sum<1024>((static_cast<uint64_t>(local) << 32) | nz, lds, &both_total);
        const uint32_t total = static_cast<uint32_t>(both_total >> 32), nz_total = static_cast<uint32_t>(both_total);
        uint32_t ex = static_cast<uint32_t>(both >> 32) + carry;
        uint32_t lx = static_cast<uint32_t>(both) + used;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (b + k < cells) {
                st[b + k] = ex;
                if (v[k])
                    list[lx++] = make_uint4(static_cast<uint32_t>(b + k), ex, v[k], 0u);
                else
                    zmin[b + k] = u2f(kEmptyZ);
            }
            ex += v[k];
        }
        carry += total;
        used += nz_total;
    }
    if (threadIdx.x == 0) {
        st[cells] = carry;
        d.fs[f * kFsCount + kFsCellsUsed] = used;
    }
}

// Scatter point indices into their cells (unordered; sorted per cell next),
// each at the place the histogram's atomic add gave it, with its z key beside
// it for the robust minimum.
__global__ void k_cell_fill(DevBuf d, Layout L) {
    const int f = blockIdx.y;
    const FrameDesc& fd = d.frames[f];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fd.n) return;
    const int gi = fd.pt_off + i;
    const int cell = d.pt_cell[gi];
    if (cell < 0) return;
    const uint32_t pos = d.cell_start[(size_t)f * (L.cell_stride + 1) + cell] + static_cast<uint32_t>(d.pt_loc[gi]);
    d.cell_items[(size_t)f * L.N + pos] = i;
    d.cell_zkey[(size_t)f * L.N + pos] = ord_f32(d.z[gi]);
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Bitonic sort of m (power of two) u32 keys in LDS by NT cooperating
// threads (tid in [0, NT)); SYNC() orders the stages.
template <int NT, typename Sync>
__device__ __forceinline__ void bitonic_u32(uint32_t* a, int m, int tid, Sync sync) {
    for (int k = 2; k <= m; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = tid; t < (m >> 1); t += NT) {
                const int lo = ((t & ~(j - 1)) << 1) | (t & (j - 1));
                const int hi = lo + j;
                const uint32_t x = a[lo], y = a[hi];
                const bool up = (lo & k) == 0;
                if ((x > y) == up) {
                    a[lo] = y;
                    a[hi] = x;
                }
            }
            sync();
        }
    }
}

// Robust minimum of a sorted cell (segmenter.cpp:249-260): s[0], replaced by
// s[i] for the largest i in [1, n/2] with s[i] - s[i-1] > 0.5f.  `keys` hold
// ord_f32(z) sorted ascending.  Returns via max-reduction over candidate i.
__device__ __forceinline__ float unord_f32(uint32_t k) {
    return u2f((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

template <int NT, typename Sync, typename MaxReduce>
__device__ __forceinline__ float robust_min(const uint32_t* keys, int n, int tid, Sync sync, MaxReduce maxred) {
    int best = 0;
    for (int i = 1 + tid; i <= n / 2; i += NT) {
        const float a = unord_f32(keys[i]);
        const float b = unord_f32(keys[i - 1]);
        if (a - b > 0.5F) best = i;  // i grows with the loop, keep the last hit
    }
    best = maxred(best);
    return unord_f32(keys[best]);
}

// Sorted cell of up to 64*S points held in registers: the point indices
// ascending for the RANSAC cells, and z for the robust minimum.
template <int S>
__device__ __forceinline__ void sort_cell_regs(int32_t* items, const uint32_t* zkeys, int n, int radial,
                                               float* zmin_cell) {
    const int lane = lane_id();
    uint32_t key[S];
    if (radial < kRansacRadialBins) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int e = s * kWave + lane;
            key[s] = e < n ? static_cast<uint32_t>(items[e]) : 0xffffffffu;
        }
        reg_sort_u32<S>(key);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int e = s * kWave + lane;
            if (e < n) items[e] = static_cast<int32_t>(key[s]);
        }
    }
    if (radial >= 1) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int e = s * kWave + lane;
            key[s] = e < n ? zkeys[e] : 0xffffffffu;
        }
        reg_sort_u32<S>(key);
        // Robust minimum (segmenter.cpp:249-260): the largest i in [1, n/2]
        // with s[i] - s[i-1] > 0.5f, else 0.
        int best = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int e = s * kWave + lane;
            const uint32_t up = __shfl_up(key[s], 1, kWave);
            const uint32_t wrap = __shfl(key[s > 0 ? s - 1 : 0], kWave - 1, kWave);
            const uint32_t prev = lane > 0 ? up : wrap;
            if (e >= 1 && e <= n / 2 && unord_f32(key[s]) - unord_f32(prev) > 0.5F) best = e;
        }
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) best = max(best, __shfl_xor(best, o, kWave));
        uint32_t kb = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const uint32_t v = __shfl(key[s], best % kWave, kWave);
            if (best / kWave == s) kb = v;
        }
        if (lane == 0) *zmin_cell = unord_f32(kb);
    }
}

// Ascending bitonic sort of 16 keys, one per lane of each 16-lane row (all
// exchanges are DPP moves inside the row).
__device__ __forceinline__ uint32_t row_sort16_u32(uint32_t key) {
    const int r = lane_id() & 15;
#pragma unroll
    for (int k = 2; k <= 16; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint32_t other = lane_xor(key, j);
            const bool asc = (r & k) == 0;
            const bool lower = (r & j) == 0;
            key = (lower == asc) ? min(key, other) : max(key, other);
        }
    }
    return key;
}

// sort_cell_regs for a cell of n <= 16 points held by one 16-lane row
// (element e = lane & 15); every lane of the row calls it.
__device__ __forceinline__ void sort_cell_row(int32_t* items, const uint32_t* zkeys, int n, int radial,
                                              float* zmin_cell) {
    const int e = lane_id() & 15;
    if (radial < kRansacRadialBins) {
        uint32_t key = e < n ? static_cast<uint32_t>(items[e]) : 0xffffffffu;
        key = row_sort16_u32(key);
        if (e < n) items[e] = static_cast<int32_t>(key);
    }
    if (radial >= 1) {
        uint32_t key = e < n ? zkeys[e] : 0xffffffffu;
        key = row_sort16_u32(key);
        // Robust minimum (segmenter.cpp:249-260): the largest i in [1, n/2]
        // with s[i] - s[i-1] > 0.5f, else 0.
        const uint32_t prev = __shfl_up(key, 1, 16);
        int best = (e >= 1 && e <= n / 2 && unord_f32(key) - unord_f32(prev) > 0.5F) ? e : 0;
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) best = max(best, static_cast<int>(lane_xor(static_cast<uint32_t>(best), o)));
        const uint32_t kb = __shfl(key, (lane_id() & ~15) + best, kWave);
        if (e == 0) *zmin_cell = unord_f32(kb);
    }
}

// One wavefront per non-empty cell (a grid-stride loop over the cell list):
// sorts the cell's point indices ascending (so the cell lists follow the
// reference's input order) for the RANSAC cells, and sorts z for the robust
// minimum.  A wavefront takes four consecutive cells of the list (each entry
// carries the cell's start and count): those of up to 16 points are sorted
// by one 16-lane row each (DPP bitonic), the others one after the other by
// the whole wavefront.  Cells up to kSmallCell (512) points sort in
// registers; larger ones are queued for k_cell_sort_big.
__global__ __launch_bounds__(256) void k_cell_sort(DevBuf d, Layout L) {
    static_assert(kSmallCell == 8 * kWave, "the register sorts cover cells up to kSmallCell points");
    int tile, f;
    xcd_tile_frame(tile, f);
    const FrameDesc& fd = d.frames[f];
    const SegConst& s = fd.sc;
    const int w = threadIdx.x / kWave;
    const int lane = lane_id();
    const size_t co = (size_t)f * L.cell_stride;
    const int used = static_cast<int>(d.fs[f * kFsCount + kFsCellsUsed]);
    const int waves = gridDim.x * 4;
    float* zmin = d.zmin + co;
    // Row g of the wavefront takes list entry q4 + g; the entry of its next
    // iteration is loaded one iteration ahead.
    const int g = lane >> 4;
    const int step = waves * 4;
    int q4 = (tile * 4 + w) * 4;
    const uint4 none = make_uint4(0xffffffffu, 0u, 0u, 0u);
    uint4 next = q4 + g < used ? d.cell_list[co + q4 + g] : none;
    for (; q4 < used; q4 += step) {
        const uint4 e = next;
        next = q4 + step + g < used ? d.cell_list[co + q4 + step + g] : none;
        const int cell_g = static_cast<int>(e.x);  // -1: no entry
        const uint32_t start_g = e.y;
        const uint32_t n_g = e.z;
        if (cell_g >= 0 && n_g <= 16u) {
            sort_cell_row(d.cell_items + (size_t)f * L.N + start_g, d.cell_zkey + (size_t)f * L.N + start_g,
                          static_cast<int>(n_g), cell_g % s.rings, zmin + cell_g);
        }
        // Rows whose cell is larger: the whole wavefront, one cell at a time.
        const uint64_t large = __ballot(cell_g >= 0 && n_g > 16u && (lane & 15) == 0);
        for (int gg = 0; gg < 4; ++gg) {
            if (!((large >> (gg * 16)) & 1ull)) continue;
            const int cell = __shfl(cell_g, gg * 16, kWave);
            const uint32_t n = static_cast<uint32_t>(__shfl(static_cast<int>(n_g), gg * 16, kWave));
            if (n > static_cast<uint32_t>(kSmallCell)) {
                if (lane == 0) {
                    const uint32_t k = atomicAdd(&d.fs[f * kFsCount + kFsNumBig], 1u);
                    if (k < static_cast<uint32_t>(L.cell_stride)) d.big_cells[co + k] = cell;
                }
                continue;
            }
            const uint32_t start = static_cast<uint32_t>(__shfl(static_cast<int>(start_g), gg * 16, kWave));
            int32_t* items = d.cell_items + (size_t)f * L.N + start;
            const uint32_t* zk = d.cell_zkey + (size_t)f * L.N + start;
            const int radial = cell % s.rings;
            if (n <= static_cast<uint32_t>(kWave)) {
                sort_cell_regs<1>(items, zk, static_cast<int>(n), radial, zmin + cell);
                continue;
            }
            if (n <= static_cast<uint32_t>(2 * kWave)) {
                sort_cell_regs<2>(items, zk, static_cast<int>(n), radial, zmin + cell);
                continue;
            }
            if (n <= static_cast<uint32_t>(4 * kWave)) {
                sort_cell_regs<4>(items, zk, static_cast<int>(n), radial, zmin + cell);
                continue;
            }
            sort_cell_regs<8>(items, zk, static_cast<int>(n), radial, zmin + cell);  // n <= kSmallCell
        }
    }
}

// Cells with more than kSmallCell points: one 1024-thread workgroup each.
__global__ __launch_bounds__(1024) void k_cell_sort_big(DevBuf d, Layout L) {
    __shared__ uint32_t a[kBigCell];
    __shared__ int red[1024 / kWave];
    const int f = blockIdx.y;
    const FrameDesc& fd = d.frames[f];
    const SegConst& s = fd.sc;
    const size_t co = (size_t)f * L.cell_stride;
    const uint32_t nbig = min(d.fs[f * kFsCount + kFsNumBig], static_cast<uint32_t>(L.cell_stride));
    const int tid = threadIdx.x;
    auto sync = [] { __syncthreads(); };
    for (uint32_t b = blockIdx.x; b < nbig; b += gridDim.x) {
        const int cell = d.big_cells[co + b];
        const uint32_t n = d.cell_count[co + cell];
        if (n > static_cast<uint32_t>(kBigCell)) {
            if (tid == 0) atomicOr(&d.fs[f * kFsCount + kFsErr], kErrCellCapacity);
            continue;
        }
        const uint32_t start = d.cell_start[(size_t)f * (L.cell_stride + 1) + cell];
        int32_t* items = d.cell_items + (size_t)f * L.N + start;
        int m = 1;
        while (m < static_cast<int>(n)) m <<= 1;
        const int radial = cell % s.rings;
        if (radial < kRansacRadialBins) {
            for (int t = tid; t < m; t += 1024) a[t] = t < static_cast<int>(n) ? static_cast<uint32_t>(items[t]) : 0xffffffffu;
            __syncthreads();
            bitonic_u32<1024>(a, m, tid, sync);
            for (int t = tid; t < static_cast<int>(n); t += 1024) items[t] = static_cast<int32_t>(a[t]);
            __syncthreads();
        }
        if (radial >= 1) {
            for (int t = tid; t < m; t += 1024)
                a[t] = t < static_cast<int>(n) ? d.cell_zkey[(size_t)f * L.N + start + t] : 0xffffffffu;
            __syncthreads();
            bitonic_u32<1024>(a, m, tid, sync);
            const float r = robust_min<1024>(a, static_cast<int>(n), tid, sync, [&](int v) {
#pragma unroll
                for (int o = kWave / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
                if (lane_id() == 0) red[tid / kWave] = v;
                __syncthreads();
                int m2 = 0;
                for (int i = 0; i < 1024 / kWave; ++i) m2 = max(m2, red[i]);
                __syncthreads();
                return m2;
            });
            if (tid == 0) d.zmin[co + cell] = r;
        }
        __syncthreads();
    }
}

// RECM recurrence along each azimuth slice (segmenter.cpp:215-267).  The
// 49 float additions must happen in order (a log-step min-plus scan would
// re-associate them), so each slice is one sequential lane.
// One wavefront per 64 slices.  A slice's cells are contiguous (cell =
// slice * rings + r), so the wavefront's 64 x rings minima are one
// contiguous block: it is loaded into LDS with coalesced reads, each lane runs
// its slice's recurrence there, and the elevations go back the same way.
// Grids with more than kRecmRings rings read and write global memory
// directly (eight rings' minima loaded ahead of their eight steps).
constexpr int kRecmRings = 64;

__global__ __launch_bounds__(kWave) void k_recm(DevBuf d, Layout L) {
    __shared__ float blk[kWave * kRecmRings];
    const int f = blockIdx.y;
    const SegConst& s = d.frames[f].sc;
    const int s0 = blockIdx.x * kWave;
    if (s0 >= s.slices) return;
    const int lane = lane_id();
    const int sl = s0 + lane;
    const int R = s.rings;
    const size_t co = (size_t)f * L.cell_stride;
    float* E = d.elev + co;
    const float* zm = d.zmin + co;
    auto step = [&](float prev, float z) {
        if (f2u(z) == kEmptyZ) return prev + s.dz;
        const float b = prev + s.dz;
        return (b < z) ? b : z;  // std::min(curr_z_min, prev + dz)
    };
    if (R <= kRecmRings) {
        const int ns = min(kWave, s.slices - s0);
        const int total = ns * R;
        const float* src = zm + (size_t)s0 * R;
        for (int i = lane; i < total; i += kWave) blk[i] = src[i];
        __syncthreads();
        if (lane < ns) {
            float* row = blk + lane * R;
            float prev = s.e0;
            row[0] = prev;
            for (int r = 1; r < R; ++r) {
                prev = step(prev, row[r]);
                row[r] = prev;
            }
        }
        __syncthreads();
        float* dst = E + (size_t)s0 * R;
        for (int i = lane; i < total; i += kWave) dst[i] = blk[i];
        return;
    }
    if (sl >= s.slices) return;
    const int off = sl * R;
    E[off] = s.e0;
    float prev = s.e0;
    for (int r0 = 1; r0 < R; r0 += 8) {
        float zz[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) zz[k] = (r0 + k < R) ? zm[off + r0 + k] : 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (r0 + k >= R) break;
            prev = step(prev, zz[k]);
            E[off + r0 + k] = prev;
        }
    }
}

// RANSAC candidate count per near cell, one wave per cell (a grid-stride
// loop over the frame's slices x 4 near cells), in the reference's iteration
// order: k = slice*4 + radial (segmenter.cpp:328-351).
__global__ __launch_bounds__(256) void k_cand_count(DevBuf d, Layout L) {
    const int f = blockIdx.y;
    const FrameDesc& fd = d.frames[f];
    const SegConst& s = fd.sc;
    const int w = threadIdx.x / kWave;
    const int lane = lane_id();
    for (int k = blockIdx.x * 4 + w; k < s.slices * kRansacRadialBins; k += gridDim.x * 4) {
    const int cell = (k / kRansacRadialBins) * s.rings + (k % kRansacRadialBins);
    const size_t co = (size_t)f * L.cell_stride;
    const uint32_t n = d.cell_count[co + cell];
    const uint32_t start = d.cell_start[(size_t)f * (L.cell_stride + 1) + cell];
    const uint32_t* zk = d.cell_zkey + (size_t)f * L.N + start;  // the cell's z in any order
    const float e = d.elev[co + cell];
    uint32_t c = 0;
    for (uint32_t t = lane; t < n; t += kWave) {
        const float z = unord_f32(zk[t]);
        c += (fabsf(e - z) < s.two_thr) ? 1u : 0u;
    }
    c = wave_sum(c);
    if (lane == 0) d.cand_count[(size_t)f * L.cand_stride + k] = c;
    }
}

__global__ __launch_bounds__(1024) void k_cand_scan(DevBuf d, Layout L) {
    __shared__ uint32_t lds[1024 / kWave + 1];
    const int f = blockIdx.x;
    const SegConst& s = d.frames[f].sc;
    const int cnt = s.slices * kRansacRadialBins;
    const uint32_t* in = d.cand_count + (size_t)f * L.cand_stride;
    uint32_t* out = d.cand_start + (size_t)f * L.cand_stride;
    uint32_t carry = 0;
    for (int base = 0; base < cnt; base += 1024 * 4) {
        uint32_t v[4];
        uint32_t local = 0;
        const int b = base + threadIdx.x * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            v[q] = (b + q < cnt) ? in[b + q] : 0u;
            local += v[q];
        }
        uint32_t total;
        uint32_t ex = block_exclusive_sum<1024>(local, lds, &total) + carry;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (b + q < cnt) out[b + q] = ex;
            ex += v[q];
        }
        carry += total;
    }
    if (threadIdx.x == 0) d.fs[f * kFsCount + kFsCands] = carry;
}

// Candidate compaction: each wave writes its cell's candidates in input
// order at the cell's scanned offset (grid-stride over the near cells).
__global__ __launch_bounds__(256) void k_cand_fill(DevBuf d, Layout L) {
    const int f = blockIdx.y;
    const FrameDesc& fd = d.frames[f];
    const SegConst& s = fd.sc;
    const int w = threadIdx.x / kWave;
    const int lane = lane_id();
    for (int k = blockIdx.x * 4 + w; k < s.slices * kRansacRadialBins; k += gridDim.x * 4) {
    const int cell = (k / kRansacRadialBins) * s.rings + (k % kRansacRadialBins);
    const size_t co = (size_t)f * L.cell_stride;
    const uint32_t n = d.cell_count[co + cell];
    const uint32_t start = d.cell_start[(size_t)f * (L.cell_stride + 1) + cell];
    const int32_t* items = d.cell_items + (size_t)f * L.N + start;
    const float e = d.elev[co + cell];
    uint32_t pos = d.cand_start[(size_t)f * L.cand_stride + k];
    const size_t ob = (size_t)f * L.N;
    for (uint32_t t0 = 0; t0 < n; t0 += kWave) {
        const uint32_t t = t0 + lane;
        bool pred = false;
        float x = 0, y = 0, z = 0;
        if (t < n) {
            const float4 pc = d.pt_xyzc[fd.pt_off + items[t]];  // one gather per point
            x = pc.x;
            y = pc.y;
            z = pc.z;
            pred = fabsf(e - z) < s.two_thr;
        }
        const uint64_t mask = __ballot(pred);
        if (pred) {
            const uint32_t o = pos + lanes_below(mask);
            d.cx[ob + o] = x;
            d.cy[ob + o] = y;
            d.cz[ob + o] = z;
        }
        pos += static_cast<uint32_t>(__popcll(mask));
    }
    }
}

// mt19937 with seed 42 (the reference's std::mt19937 gen{42},
// segmenter.cpp:369): the raw output stream does not depend on the frame, so
// it is generated once per context.  Standard MT19937 recurrence
// (Matsumoto & Nishimura 1998).
__global__ void k_mt19937(uint32_t* out, int count) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t mt[624];
    mt[0] = 42u;
    for (int i = 1; i < 624; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + static_cast<uint32_t>(i);
    int idx = 624;
    for (int k = 0; k < count; ++k) {
        if (idx >= 624) {
            for (int i = 0; i < 624; ++i) {
                const uint32_t yv = (mt[i] & 0x80000000u) | (mt[(i + 1) % 624] & 0x7fffffffu);
                mt[i] = mt[(i + 397) % 624] ^ (yv >> 1) ^ ((yv & 1u) ? 0x9908b0dfu : 0u);
            }
            idx = 0;
        }
        uint32_t yv = mt[idx++];
        yv ^= yv >> 11;
        yv ^= (yv << 7) & 0x9d2c5680u;
        yv ^= (yv << 15) & 0xefc60000u;
        yv ^= yv >> 18;
        out[k] = yv;
    }
}

// Hypotheses: lane 0 maps the raw stream to indices with libstdc++'s
// uniform_int_distribution<uint32_t> for a 32-bit engine (Lemire's
// nearly-divisionless method, bits/uniform_int_dist.h), redrawing p3 while
// it equals p2 (segmenter.cpp:379-386); lanes 0..59 then build the planes
// (segmenter.cpp:389-418).
__global__ __launch_bounds__(64) void k_ransac_hyp(DevBuf d, Layout L) {
    __shared__ uint32_t idx2[kRansacIters], idx3[kRansacIters];
    // 60 hypotheses draw about 120 values (a few more after rejections or
    // repeated indices): the first 256 are staged in LDS, later ones are read
    // from global memory.
    constexpr int kStaged = 256;
    __shared__ uint32_t raw[kStaged];
    __shared__ uint32_t ok;
    const int f = blockIdx.x;
    const FrameDesc& fd = d.frames[f];
    const SegConst& s = fd.sc;
    const uint32_t n = d.fs[f * kFsCount + kFsCands];
    const int lane = threadIdx.x;
    if (n == 0) return;  // no candidates: the reference returns (segmenter.cpp:353-356)
    if (n == 1) {
        if (lane == 0) atomicOr(&d.fs[f * kFsCount + kFsErr], kErrRansacSingle);
        return;
    }
#pragma unroll
    for (int i = lane; i < kStaged; i += 64) raw[i] = d.mt_raw[i];
    __syncthreads();
    if (lane == 0) {
        int pos = 0;
        bool good = true;
        auto draw = [&]() -> uint32_t {
            if (pos >= kMtRawCount) {
                good = false;
                return 0;
            }
            uint64_t prod = static_cast<uint64_t>(pos < kStaged ? raw[pos] : d.mt_raw[pos]) * n;
            ++pos;
            uint32_t low = static_cast<uint32_t>(prod);
            if (low < n) {
                const uint32_t thresh = (0u - n) % n;
                while (low < thresh && good) {
                    if (pos >= kMtRawCount) {
                        good = false;
                        break;
                    }
                    prod = static_cast<uint64_t>(pos < kStaged ? raw[pos] : d.mt_raw[pos]) * n;
                    ++pos;
                    low = static_cast<uint32_t>(prod);
                }
            }
            return static_cast<uint32_t>(prod >> 32);
        };
        for (int it = 0; it < kRansacIters && good; ++it) {
            const uint32_t a = draw();
            uint32_t b = draw();
            while (b == a && good) b = draw();
            idx2[it] = a;
            idx3[it] = b;
        }
        ok = good ? 1u : 0u;
        if (!good) atomicOr(&d.fs[f * kFsCount + kFsErr], kErrRngExhausted);
    }
    __syncthreads();
    if (!ok || lane >= kRansacIters) return;
    const size_t ob = (size_t)f * L.N;
    const float p1x = 0.0f, p1y = 0.0f, p1z = s.neg_h;
    const uint32_t i2 = idx2[lane], i3 = idx3[lane];
    const float p2x = d.cx[ob + i2], p2y = d.cy[ob + i2], p2z = d.cz[ob + i2];
    const float p3x = d.cx[ob + i3], p3y = d.cy[ob + i3], p3z = d.cz[ob + i3];
    float nx = ((p2y - p1y) * (p3z - p1z)) - ((p2z - p1z) * (p3y - p1y));
    float ny = ((p2z - p1z) * (p3x - p1x)) - ((p2x - p1x) * (p3z - p1z));
    float nz = ((p2x - p1x) * (p3y - p1y)) - ((p2y - p1y) * (p3x - p1x));
    const float den2 = (nx * nx) + (ny * ny) + (nz * nz);
    uint32_t valid = 0;
    float pd = 0.0f;
    if (!(den2 < 1.0e-5F)) {
        const float norm = 1.0F / sqrtf(den2);
        nz *= norm;
        if (!(fabsf(nz) < s.cos_max)) {
            nx *= norm;
            ny *= norm;
            pd = (nx * p1x) + (ny * p1y) + (nz * p1z);
            valid = 1;
        }
    }
    float* h = d.hyp + ((size_t)f * kRansacIters + lane) * 4;
    h[0] = nx;
    h[1] = ny;
    h[2] = nz;
    h[3] = pd;
    d.hyp_valid[f * kRansacIters + lane] = valid;
}

// Inlier counts of all 60 hypotheses at once: one candidate per thread,
// ballot + popcount per hypothesis, LDS totals, one atomic per block and
// hypothesis (segmenter.cpp:420-431).
constexpr int kScoreRows = 4;  // candidates per thread: rows of 256, each plane read once for all of them

// Every candidate scores all hypotheses (segmenter.cpp:410-430: |a x + b y +
// c z - d| < thr, the same float expression).  The planes are read as
// wavefront-uniform values (scalar loads, no LDS); each wavefront counts a
// plane's hits with ballots over its rows, lane `it` keeps plane it's count,
// and the block's counts meet in LDS before one atomic per plane.
__global__ __launch_bounds__(256) void k_ransac_count(DevBuf d, Layout L) {
    __shared__ uint32_t acc[kRansacIters];
    static_assert(kRansacIters <= kWave, "one lane per hypothesis holds its count");
    const int f = blockIdx.y;
    const SegConst& s = d.frames[f].sc;
    const uint32_t n = d.fs[f * kFsCount + kFsCands];
    const uint32_t base = blockIdx.x * (256u * kScoreRows);
    if (n < 2 || base >= n) return;
    if ((d.fs[f * kFsCount + kFsErr] & (kErrRansacSingle | kErrRngExhausted)) != 0) return;
    const int tid = threadIdx.x;
    const int lane = lane_id();
    if (tid < kRansacIters) acc[tid] = 0;
    const size_t ob = (size_t)f * L.N;
    float x[kScoreRows], y[kScoreRows], z[kScoreRows];
    bool act[kScoreRows];
#pragma unroll
    for (int r = 0; r < kScoreRows; ++r) {
        const uint32_t t = base + r * 256u + tid;
        act[r] = t < n;
        x[r] = act[r] ? d.cx[ob + t] : 0.0f;
        y[r] = act[r] ? d.cy[ob + t] : 0.0f;
        z[r] = act[r] ? d.cz[ob + t] : 0.0f;
    }
    const float4* hyp = reinterpret_cast<const float4*>(d.hyp) + (size_t)f * kRansacIters;
    const uint32_t* valid = d.hyp_valid + f * kRansacIters;
    const float thr = s.thr;
    uint32_t mine = 0;
    for (int it = 0; it < kRansacIters; ++it) {
        if (!valid[it]) continue;
        const float4 p = hyp[it];
        uint32_t c = 0;
#pragma unroll
        for (int r = 0; r < kScoreRows; ++r) {
            const float od = fabsf((p.x * x[r]) + (p.y * y[r]) + (p.z * z[r]) - p.w);
            c += static_cast<uint32_t>(__popcll(__ballot(act[r] && od < thr)));
        }
        mine += lane == it ? c : 0u;
    }
    __syncthreads();
    if (lane < kRansacIters && mine) atomicAdd(&acc[lane], mine);
    __syncthreads();
    if (tid < kRansacIters && valid[tid] && acc[tid]) atomicAdd(&d.hyp_count[f * kRansacIters + tid], acc[tid]);
}

// Best hypothesis: strict `>` in iteration order, then the sign flip
// (segmenter.cpp:433-453).
__global__ void k_ransac_best(DevBuf d, Layout L) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= L.B_active) return;
    uint32_t* fs = d.fs + f * kFsCount;
    if (fs[kFsCands] < 2 || (fs[kFsErr] & (kErrRansacSingle | kErrRngExhausted))) return;
    uint32_t best = 0;
    float a = 0.0f, b = 0.0f, c = 1.0f, dd = 0.0f;
    for (int it = 0; it < kRansacIters; ++it) {
        if (!d.hyp_valid[f * kRansacIters + it]) continue;
        const uint32_t cnt = d.hyp_count[f * kRansacIters + it];
        if (cnt > best) {
            best = cnt;
            const float* h = d.hyp + ((size_t)f * kRansacIters + it) * 4;
            a = h[0];
            b = h[1];
            c = h[2];
            dd = h[3];
        }
    }
    if (best > 0 && c < 0) {
        a = -a;
        b = -b;
        c = -c;
        dd = -dd;
    }
    fs[kFsBest] = best;
    fs[kFsApplied] = best > 0 ? 1u : 0u;
    fs[kFsPlaneA] = f2u(a);
    fs[kFsPlaneB] = f2u(b);
    fs[kFsPlaneC] = f2u(c);
    fs[kFsPlaneD] = f2u(dd);
}

// Pixel states after RECM + RANSAC: the winner's label (segmenter.cpp:
// 278-281 and 466-475), its index (cloud_mapping_indices_) and xyz.
__global__ void k_pixel_label(DevBuf d, Layout L) {
    const int f = blockIdx.y;
    const FrameDesc& fd = d.frames[f];
    const SegConst& s = fd.sc;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= s.H * s.W) return;
    const unsigned long long key = d.pix_key[fd.pix_off + p];
    const size_t po = fd.pix_off + p;
    if (key == ~0ull) {
        d.state[po] = kEmpty;
        d.mapping[po] = -1;
        d.pxyz[po] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));  // w: the winner, -1 for none
        return;
    }
    const int i = static_cast<int>(key & ((1ull << L.idx_bits) - 1));
    const float4 pc = d.pt_xyzc[fd.pt_off + i];  // one gather: x, y, z and the cell
    const float x = pc.x, y = pc.y, z = pc.z;
    const int cell = __float_as_int(pc.w);
    const float e = d.elev[(size_t)f * L.cell_stride + cell];
    uint8_t lab = (z >= e + s.thr) ? kObstacle : kGround;
    const uint32_t* fs = d.fs + f * kFsCount;
    if (fs[kFsApplied] && (cell % s.rings) < kRansacRadialBins) {
        const float a = u2f(fs[kFsPlaneA]), b = u2f(fs[kFsPlaneB]), c = u2f(fs[kFsPlaneC]), dd = u2f(fs[kFsPlaneD]);
        const float sd = (a * x) + (b * y) + (c * z) - dd;
        if (sd < s.thr) lab = kGround;
    }
    d.state[po] = lab;
    d.mapping[po] = i;
    d.pxyz[po] = make_float4(x, y, z, __int_as_float(i));
}

// ---------------------------------------------------------------------------
void launch_segment_front(lp_context* c, int B) {
    DevBuf& d = c->d;
    const Layout& L = c->L;
    hipStream_t st = c->stream;
    const int nb = (L.N + 255) / 256;
    launch_clear(c, B);
    stage_mark(c, "project");
    const int wg = c->point_wg;
    if (c->persistent)
        hipLaunchKernelGGL(k_project_q, dim3(kPersistentWorkgroups), dim3(wg), 0, st, d, L, B);
    else
        hipLaunchKernelGGL(k_project, dim3((L.N + wg - 1) / wg, B), dim3(wg), 0, st, d, L);
    stage_mark(c, "cells");
    hipLaunchKernelGGL(k_cell_scan, dim3(B), dim3(1024), 0, st, d, L);
    hipLaunchKernelGGL(k_cell_fill, dim3(nb, B), dim3(256), 0, st, d, L);
    // About 8192 wavefronts in all, at least 64 and at most 4096 per frame.
    hipLaunchKernelGGL(k_cell_sort, dim3(std::min(1024, std::max(16, 2048 / B)), B), dim3(256), 0, st, d, L);
    // Cells beyond kSmallCell points are rare: about 1024 workgroups in all.
    hipLaunchKernelGGL(k_cell_sort_big, dim3(std::min(64, std::max(4, 1024 / B)), B), dim3(1024), 0, st, d, L);
    stage_mark(c, "recm");
    hipLaunchKernelGGL(k_recm, dim3((L.cand_stride / kRansacRadialBins + kWave - 1) / kWave, B), dim3(kWave), 0, st, d,
                       L);
    stage_mark(c, "ransac");
    // 128 workgroups (512 wavefronts) per frame walk the slices x 4 near cells.
    hipLaunchKernelGGL(k_cand_count, dim3(128, B), dim3(256), 0, st, d, L);
    hipLaunchKernelGGL(k_cand_scan, dim3(B), dim3(1024), 0, st, d, L);
    hipLaunchKernelGGL(k_cand_fill, dim3(128, B), dim3(256), 0, st, d, L);
    hipLaunchKernelGGL(k_ransac_hyp, dim3(B), dim3(64), 0, st, d, L);
    hipLaunchKernelGGL(k_ransac_count, dim3((L.N + 256 * kScoreRows - 1) / (256 * kScoreRows), B), dim3(256), 0, st,
                       d, L);
    hipLaunchKernelGGL(k_ransac_best, dim3((B + 63) / 64), dim3(64), 0, st, d, L);
    stage_mark(c, "pixels");
    hipLaunchKernelGGL(k_pixel_label, dim3((L.P + 255) / 256, B), dim3(256), 0, st, d, L);
}

void launch_clear(lp_context* c, int B) {
    stage_mark(c, "clear");
    hipLaunchKernelGGL(k_clear, dim3(256, B), dim3(256), 0, c->stream, c->d, c->L);
}

// Hash slots back to their clean state: key -1 (free), count 0, aux
// kHashMinClean.  k_hash_reset takes the slots a stage occupied (slots[0 ..
// fs[count_slot]) per frame); k_hash_init clears every slot (context
// creation, or a frame whose list is incomplete).
__global__ void k_hash_reset(DevBuf d, Layout L, const int32_t* slots, int count_slot) {
    const int f = blockIdx.y;
    const uint32_t cnt = d.fs[f * kFsCount + count_slot];
    const size_t ho = (size_t)f * L.hash_cap;
    const int stride = gridDim.x * blockDim.x;
    const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (cnt > static_cast<uint32_t>(L.N)) {
        for (int h = t0; h < L.hash_cap; h += stride) {
            d.hslot[ho + h] = HashSlot{-1, 0u, 0, kHashMinClean};
        }
        return;
    }
    for (int v = t0; v < static_cast<int>(cnt); v += stride) {
        const int32_t h = slots[(size_t)f * L.N + v];
        d.hslot[ho + h] = HashSlot{-1, 0u, 0, kHashMinClean};
    }
}

__global__ void k_hash_init(DevBuf d, Layout L) {
    const size_t total = (size_t)L.B * L.hash_cap;
    for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < total; o += (size_t)gridDim.x * blockDim.x) {
        d.hslot[o] = HashSlot{-1, 0u, 0, kHashMinClean};
    }
}

void launch_hash_reset(lp_context* c, int B, const int32_t* slots, int count_slot) {
    hipLaunchKernelGGL(k_hash_reset, dim3(64, B), dim3(256), 0, c->stream, c->d, c->L, slots, count_slot);
}

void launch_hash_init(lp_context* c) {
    hipLaunchKernelGGL(k_hash_init, dim3(1024), dim3(256), 0, c->stream, c->d, c->L);
}

void launch_mt19937(uint32_t* out, int count, hipStream_t st) {
    hipLaunchKernelGGL(k_mt19937, dim3(1), dim3(1), 0, st, out, count);
}

}  // namespace lp
