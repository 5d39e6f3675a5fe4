// lp_noise_ref.hip -- NoiseRemover::filter with the reference object's KD-tree
// search, bit for bit.
//
// Reference (paths relative to the reference repository root):
//   lidar_processing_lib/src/noise_remover.cpp:38-68        filter: one radius_search_k_nearest per point
//   lidar_processing_lib/include/lidar_processing_lib/kdtree.hpp:170-213   rebuild (std::nth_element per range)
//   lidar_processing_lib/include/lidar_processing_lib/kdtree.hpp:339-397   radius_search_k_nearest
//
// radius_search_k_nearest stops at the k-th neighbour without clearing its
// member stack: the entries it did not visit stay, and the next query pushes
// its root on top of them.  A query whose own traversal finds fewer than k
// points (it would be NOISE by the radius rule) then goes on through the
// earlier queries' leftover subtrees with its own target and radius, and can
// count its neighbours again.  The device reproduces that sequence for one
// NoiseRemover per frame (an empty stack at the first query):
//
// 1. The tree.  Node storage is the input in order, then std::nth_element on
//    every range [b, e) at mid = b + (e-b)/2 by axis depth % 3.  The range
//    shape depends on n only; which point lands in which slot depends on
//    libstdc++'s introselect (median-of-three pivot, Hoare partition,
//    insertion sort below four elements, heap selection at the depth
//    limit), because ties on the split axis are common.  The kernels below
//    restate that algorithm: the Hoare partition of a range is computed in
//    parallel (the j-th element >= pivot from the left is swapped with the
//    j-th element <= pivot from the right while the first precedes the
//    second; the cut is the (m+1)-th left element or the m-th right one,
//    whichever comes first), ranges of up to kKdSmall points build their
//    whole subtree in LDS, and short ranges run the sequential form.
// 2. The own traversal of every query, in parallel.  Its hits are met in
//    pre-order (left before right), so the state at the k-th hit u is the
//    set of pending right children of u's ancestors: a bit per depth.  A
//    query with fewer than k hits records them (their slots).
// 3. The stale stack, one workgroup per frame, over the short-of-k queries
//    in input order.  The stack is the concatenation of the earlier
//    queries' pending entries.  Walking an entry's subtree with the short
//    query's target finds exactly its hits inside that slot range, so an
//    entry's count is a range test against the recorded slots; the entry
//    where the count reaches k is walked by pre-order to the hit that
//    stops the search, whose pending right children replace the entry.
#include <cfloat>
#include <climits>

#include "lp_kernels.h"

namespace lp {

using namespace lpd;

namespace {

constexpr int kKdSmall = 2048;          // subtree built in LDS (float4 records)
constexpr int kKdTiny = 32;             // ranges selected by one thread
constexpr int kKdBigThreads = 1024;
constexpr int kKdSmallThreads = 256;

__device__ __forceinline__ float kd_key(const float4& r, int axis) {
    return axis == 0 ? r.x : (axis == 1 ? r.y : r.z);
}

__device__ __forceinline__ void kd_swap(float4* A, int i, int j) {
    const float4 t = A[i];
    A[i] = A[j];
    A[j] = t;
}

// Range of node r (0 <= r < 2^level, left to right) of the tree over slots
// [0, n): the bits of r choose left (0) or right (1) from the top.
__device__ __forceinline__ void kd_level_range(int n, int level, int r, int& b, int& e) {
    b = 0;
    e = n;
    for (int d = level - 1; d >= 0 && b < e; --d) {
        const int m = b + (e - b) / 2;
        if ((r >> d) & 1)
            b = m + 1;
        else
            e = m;
    }
}

// ---- sequential libstdc++ selection (bits/stl_algo.h, bits/stl_heap.h) ----
__device__ void kd_move_median_to_first(float4* A, int result, int a, int b, int c, int ax) {
    const float ka = kd_key(A[a], ax), kb = kd_key(A[b], ax), kc = kd_key(A[c], ax);
    int m;
    if (ka < kb) {
        if (kb < kc)
            m = b;
        else if (ka < kc)
            m = c;
        else
            m = a;
    } else if (ka < kc) {
        m = a;
    } else if (kb < kc) {
        m = c;
    } else {
        m = b;
    }
    kd_swap(A, result, m);
}

// The scans stay inside [pivot, end) even where the median-of-three guard
// would be needed (it always holds for finite keys, which are checked).
__device__ int kd_unguarded_partition(float4* A, int first, int last, int pivot, int ax) {
    const float p = kd_key(A[pivot], ax);
    const int end = last;
    while (true) {
        while (first < end - 1 && kd_key(A[first], ax) < p) ++first;
        --last;
        while (last > pivot && p < kd_key(A[last], ax)) --last;
        if (!(first < last)) return first;
        kd_swap(A, first, last);
        ++first;
    }
}

__device__ void kd_push_heap(float4* A, int first, int hole, int top, float4 v, int ax) {
    int parent = (hole - 1) / 2;
    const float kv = kd_key(v, ax);
    while (hole > top && kd_key(A[first + parent], ax) < kv) {
        A[first + hole] = A[first + parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    A[first + hole] = v;
}

__device__ void kd_adjust_heap(float4* A, int first, int hole, int len, float4 v, int ax) {
    const int top = hole;
    int second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        if (kd_key(A[first + second], ax) < kd_key(A[first + second - 1], ax)) second--;
        A[first + hole] = A[first + second];
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        A[first + hole] = A[first + second - 1];
        hole = second - 1;
    }
    kd_push_heap(A, first, hole, top, v, ax);
}

__device__ void kd_heap_select(float4* A, int first, int middle, int last, int ax) {
    const int len = middle - first;
    if (len >= 2) {
        for (int parent = (len - 2) / 2;; --parent) {
            kd_adjust_heap(A, first, parent, len, A[first + parent], ax);
            if (parent == 0) break;
        }
    }
    for (int i = middle; i < last; ++i)
        if (kd_key(A[i], ax) < kd_key(A[first], ax)) {
            const float4 v = A[i];
            A[i] = A[first];
            kd_adjust_heap(A, first, 0, len, v, ax);
        }
}

__device__ void kd_insertion_sort(float4* A, int first, int last, int ax) {
    if (first == last) return;
    for (int i = first + 1; i != last; ++i) {
        const float4 v = A[i];
        const float kv = kd_key(v, ax);
        if (kv < kd_key(A[first], ax)) {
            for (int k = i; k > first; --k) A[k] = A[k - 1];
            A[first] = v;
        } else {
            int j = i, nx = i - 1;
            while (kv < kd_key(A[nx], ax)) {
                A[j] = A[nx];
                j = nx;
                --nx;
            }
            A[j] = v;
        }
    }
}

// std::nth_element(A + first, A + nth, A + last) by axis `ax`, one thread.
__device__ void kd_nth_seq(float4* A, int first, int nth, int last, int ax) {
    if (first == last || nth == last) return;
    int depth = 2 * (31 - __clz(last - first));
    while (last - first > 3) {
        if (depth == 0) {
            kd_heap_select(A, first, nth + 1, last, ax);
            kd_swap(A, first, nth);
            return;
        }
        --depth;
        kd_move_median_to_first(A, first, first + 1, first + (last - first) / 2, last - 1, ax);
        const int cut = kd_unguarded_partition(A, first + 1, last, first, ax);
        if (cut <= nth)
            first = cut;
        else
            last = cut;
    }
    kd_insertion_sort(A, first, last, ax);
}

// ---- parallel selection by a group of G threads -----------------------------
// Group policies: a whole workgroup (barriers, block scan) or one wavefront.
template <int G>
struct BlockGroup {
    uint64_t* lds;  // G/64 + 1 entries
    int* sh;        // 8 shared ints
    __device__ int rank() const { return threadIdx.x; }
    __device__ void sync() const { __syncthreads(); }
    // Inclusive prefix (and the group total) of a 64-bit value.
    __device__ uint64_t incl(uint64_t v, uint64_t* total) const {
        return block_exclusive_sum<G>(v, lds, total) + v;
    }
    __device__ int sum(int v) const {
        uint64_t t;
        block_exclusive_sum<G>(static_cast<uint64_t>(static_cast<uint32_t>(v)), lds, &t);
        return static_cast<int>(t);
    }
};

struct WaveGroup {
    int* sh;  // 8 ints private to this wavefront
    __device__ int rank() const { return lane_id(); }
    __device__ void sync() const { wave_sync(); }
    __device__ uint64_t incl(uint64_t v, uint64_t* total) const {
        const uint64_t r = wave_inclusive_sum(v);
        *total = __shfl(r, kWave - 1, kWave);
        return r;
    }
    __device__ int sum(int v) const { return wave_sum(v); }
};

// The Hoare partition of [F, Lst) around the key P, as the sequential scans
// perform it (left scan stops at keys >= P, right scan at keys <= P).  Lp /
// Rp receive the left / right stop positions in scan order (index 0 = the
// first stop); returns the cut.
template <int G, class Grp>
__device__ int kd_partition_group(const Grp& g, float4* A, int F, int Lst, float P, int ax, int* Lp, int* Rp) {
    const int t = g.rank();
    int cl = 0, cr = 0;
    for (int p = F + t; p < Lst; p += G) {
        const float k = kd_key(A[p], ax);
        cl += (k < P) ? 0 : 1;
        cr += (P < k) ? 0 : 1;
    }
    const int totL = g.sum(cl);
    const int totR = g.sum(cr);
    int carryL = 0, carryR = 0, matched = 0;
    for (int base = F; base < Lst; base += G) {
        const int p = base + t;
        uint32_t fl = 0, fr = 0;
        if (p < Lst) {
            const float k = kd_key(A[p], ax);
            fl = (k < P) ? 0u : 1u;
            fr = (P < k) ? 0u : 1u;
        }
        uint64_t tot;
        const uint64_t inc = g.incl((static_cast<uint64_t>(fl) << 32) | fr, &tot);
        const int rankL = carryL + static_cast<int>(inc >> 32);        // 1-based when fl
        const int inclR = carryR + static_cast<int>(inc & 0xffffffffu);
        if (fl) {
            Lp[rankL - 1] = p;
            if (totR - inclR >= rankL) ++matched;   // the rankL-th right stop lies beyond p
        }
        if (fr) Rp[totR - inclR] = p;               // rank from the right, 0-based
        carryL += static_cast<int>(tot >> 32);
        carryR += static_cast<int>(tot & 0xffffffffu);
    }
    const int m = g.sum(matched);
    g.sync();
    for (int j = t; j < m; j += G) kd_swap(A, Lp[j], Rp[j]);
    int cut = INT_MAX;
    if (m < totL) cut = Lp[m];
    if (m >= 1) cut = min(cut, Rp[m - 1]);
    if (cut == INT_MAX) cut = Lst;  // no stop at all: only for keys without an order
    g.sync();
    return cut;
}

// std::nth_element by a group: the median-of-three and the final steps on
// rank 0, the partitions in parallel.
template <int G, class Grp>
__device__ void kd_nth_group(const Grp& g, float4* A, int first, int nth, int last, int ax, int* Lp, int* Rp) {
    if (first == last || nth == last) return;
    int depth = 2 * (31 - __clz(last - first));
    while (last - first > 3) {
        if (depth == 0) {
            if (g.rank() == 0) {
                kd_heap_select(A, first, nth + 1, last, ax);
                kd_swap(A, first, nth);
            }
            g.sync();
            return;
        }
        --depth;
        if (g.rank() == 0) kd_move_median_to_first(A, first, first + 1, first + (last - first) / 2, last - 1, ax);
        g.sync();
        const float P = kd_key(A[first], ax);
        const int cut = kd_partition_group<G>(g, A, first + 1, last, P, ax, Lp + first, Rp + first);
        if (cut <= nth)
            first = cut;
        else
            last = cut;
    }
    if (g.rank() == 0) kd_insertion_sort(A, first, last, ax);
    g.sync();
}

// Level below which every range of frame f fits kKdSmall points.
__device__ __forceinline__ int kd_small_level(int n) {
    int l = 0;
    while ((n >> l) > kKdSmall) ++l;
    return l;
}

}  // namespace

// Slot i of the tree of frame f starts as input point i (nodes_ filled in
// order, kdtree.hpp:182-186); labels and keep flags start VALID.
__global__ void k_kd_init(DevBuf d, Layout L) {
    int tile, f;
    xcd_tile_frame(tile, f);
    const FrameDesc& fd = d.frames[f];
    const int i = tile * blockDim.x + threadIdx.x;
    const size_t po = (size_t)f * L.N;
    if (tile == 0 && threadIdx.x == 0) {
        d.fs[f * kFsCount + kFsKept] = 0;
        d.fs[f * kFsCount + kFsNoise] = 0;
        d.fs[f * kFsCount + kFsNoiseHard] = 0;
        d.fs[f * kFsCount + kFsNoiseCells] = 0;
        d.fs[f * kFsCount + kFsNoiseErr] = 0;
    }
    if (i >= fd.n_raw) return;
    const int gi = fd.pt_off + i;
    const float x = d.rx[gi], y = d.ry[gi], z = d.rz[gi];
    d.kd[po + i] = make_float4(x, y, z, __int_as_float(i));
    d.noise[po + i] = 0;
    d.obs_flag[po + i] = 1u;
}

// One workgroup per range of `level` that is still larger than kKdSmall
// points: the parallel selection in global memory.
__global__ __launch_bounds__(kKdBigThreads) void k_kd_split_big(DevBuf d, Layout L, int level) {
    __shared__ uint64_t lds[kKdBigThreads / kWave + 1];
    __shared__ int sh[8];
    const int f = blockIdx.y;
    const int n = d.frames[f].n_raw;
    if (level >= kd_small_level(n) || blockIdx.x >= (1u << level)) return;
    int b, e;
    kd_level_range(n, level, blockIdx.x, b, e);
    if (b >= e) return;
    const size_t po = (size_t)f * L.N;
    BlockGroup<kKdBigThreads> g{lds, sh};
    kd_nth_group<kKdBigThreads>(g, d.kd + po, b, b + (e - b) / 2, e, level % 3, d.kd_tl + po, d.kd_tr + po);
}

// One workgroup per range of the frame's first small level: the range goes
// to LDS and every deeper level is selected there (ranges above kKdTiny by a
// wavefront each, shorter ones by a thread each), then it is written back.
__global__ __launch_bounds__(kKdSmallThreads) void k_kd_split_small(DevBuf d, Layout L) {
    __shared__ float4 S[kKdSmall];
    __shared__ int Lt[kKdSmall];
    __shared__ int Rt[kKdSmall];
    __shared__ int sh[kKdSmallThreads / kWave][8];
    const int f = blockIdx.y;
    const int n = d.frames[f].n_raw;
    const int level = kd_small_level(n);
    if (blockIdx.x >= (1u << level)) return;
    int b0, e0;
    kd_level_range(n, level, blockIdx.x, b0, e0);
    const int m = e0 - b0;
    if (m <= 0) return;
    const size_t po = (size_t)f * L.N;
    float4* G = d.kd + po + b0;
    for (int i = threadIdx.x; i < m; i += kKdSmallThreads) S[i] = G[i];
    __syncthreads();
    const int w = threadIdx.x / kWave;
    constexpr int kWaves = kKdSmallThreads / kWave;
    WaveGroup wg{sh[w]};
    for (int s = 0;; ++s) {
        const int ranges = 1 << s;
        const int ax = (level + s) % 3;
        // Largest range of this sub-level: floor(m / 2^s).
        const int biggest = m >> s;
        if (biggest > kKdTiny) {
            for (int r = w; r < ranges; r += kWaves) {
                int b, e;
                kd_level_range(m, s, r, b, e);
                if (e - b > kKdTiny) {
                    kd_nth_group<kWave>(wg, S, b, b + (e - b) / 2, e, ax, Lt, Rt);
                } else if (b < e && lane_id() == 0) {
                    kd_nth_seq(S, b, b + (e - b) / 2, e, ax);
                }
            }
        } else {
            for (int r = threadIdx.x; r < ranges; r += kKdSmallThreads) {
                int b, e;
                kd_level_range(m, s, r, b, e);
                if (b < e) kd_nth_seq(S, b, b + (e - b) / 2, e, ax);
            }
        }
        __syncthreads();
        if (biggest <= 1) break;
    }
    for (int i = threadIdx.x; i < m; i += kKdSmallThreads) G[i] = S[i];
}

namespace {

constexpr int kKdMaxK = static_cast<int>(kKdMaxNeighbours);
constexpr int kKdResolveThreads = 1024;

// Pending entry: slot range [b, e) of a subtree root and its depth.
__device__ __forceinline__ uint64_t kd_entry(int b, int e, int dep) {
    return static_cast<uint64_t>(b) | (static_cast<uint64_t>(e) << 25) | (static_cast<uint64_t>(dep) << 50);
}
__device__ __forceinline__ void kd_entry_split(uint64_t v, int& b, int& e, int& dep) {
    b = static_cast<int>(v & 0x1ffffffu);
    e = static_cast<int>((v >> 25) & 0x1ffffffu);
    dep = static_cast<int>(v >> 50);
}

// noise_remover.cpp:46-59: the squared range summed in float, widened to
// double, scaled and rounded back; std::max keeps the first unless smaller.
__device__ __forceinline__ float kd_radius_sqr(const NoiseConst& nc, float x, float y) {
    const double range_sqr = static_cast<double>(x * x + y * y);
    const float r2 = static_cast<float>(nc.scaling * range_sqr);
    return (r2 < nc.min_sqr) ? nc.min_sqr : r2;
}

// kdtree.hpp:131-143 dist_sqr(target, node): (t0-p0)^2 + ((t1-p1)^2 + ((t2-p2)^2 + 0)).
__device__ __forceinline__ float kd_dist(float tx, float ty, float tz, const float4& p) {
    const float d0 = tx - p.x, d1 = ty - p.y, d2 = tz - p.z;
    return d0 * d0 + (d1 * d1 + (d2 * d2 + 0.0f));
}

// The own traversal of one query from an empty stack (kdtree.hpp:351-383).
// The stack is kept as the current subtree plus one bit per depth whose
// right child is pending; a pop resumes at the deepest such right child.
// Returns the hit count (capped at k); hits[] receives up to kKdMaxK hit
// slots in visiting order; at the k-th hit *uk is its slot and *mask the
// pending depths (the leftover the reference keeps).
__device__ int kd_own_search(const float4* T, int n, float tx, float ty, float tz, float prox, int k, int* hits,
                             int* uk, uint32_t* mask_out) {
    int b = 0, e = n, dep = 0, last = 0, hc = 0;
    uint32_t mask = 0;
    *uk = -1;
    while (true) {
        if (b >= e) {
            if (mask == 0) break;
            const int dd = 31 - __clz(mask);
            mask &= ~(1u << dd);
            // The pending right child at depth dd belongs to the ancestor of
            // the last visited node at that depth.
            int ab = 0, ae = n;
            for (int q = 0; q < dd; ++q) {
                const int m = ab + (ae - ab) / 2;
                if (last < m)
                    ae = m;
                else
                    ab = m + 1;
            }
            b = ab + (ae - ab) / 2 + 1;
            e = ae;
            dep = dd + 1;
            continue;
        }
        const int mid = b + (e - b) / 2;
        last = mid;
        const float4 nd = T[mid];
        if (kd_dist(tx, ty, tz, nd) <= prox) {
            if (hc < kKdMaxK) hits[hc] = mid;
            ++hc;
            if (hc >= k) {
                *uk = mid;
                break;
            }
        }
        const int ax = dep % 3;
        const float delta = kd_key(nd, ax) - (ax == 0 ? tx : (ax == 1 ? ty : tz));
        if (delta * delta <= prox) {
            if (mid + 1 < e) mask |= 1u << dep;
            e = mid;
        } else if (delta > 0) {
            e = mid;
        } else {
            b = mid + 1;
        }
        ++dep;
    }
    *mask_out = mask;
    return hc;
}

__device__ __forceinline__ bool kd_frame_failed(const DevBuf& d, int f) {
    return d.fs[f * kFsCount + kFsNoiseErr] != 0;
}

}  // namespace

// Own traversal of every query: leftover counts of the queries that reach k
// neighbours, a flag for the others.
__global__ void k_kd_query(DevBuf d, Layout L) {
    int tile, f;
    xcd_tile_frame(tile, f);
    const FrameDesc& fd = d.frames[f];
    const int n = fd.n_raw;
    const int i = tile * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t po = (size_t)f * L.N;
    const NoiseConst nc = *d.nc;
    const int gi = fd.pt_off + i;
    const float tx = d.rx[gi], ty = d.ry[gi], tz = d.rz[gi];
    if (!(isfinite(tx) && isfinite(ty) && isfinite(tz))) {
        // std::nth_element has no order for such keys: the frame fails (the
        // tree kernels stay inside their ranges whatever the keys).
        atomicOr(&d.fs[f * kFsCount + kFsNoiseErr], kErrNoiseInput);
        d.kd_lcnt[po + i] = 0;
        d.kd_dflag[po + i] = 0;
        return;
    }
    int hits[kKdMaxK];
    int uk;
    uint32_t mask;
    const int hc = kd_own_search(d.kd + po, n, tx, ty, tz, kd_radius_sqr(nc, tx, ty), static_cast<int>(nc.k), hits,
                                 &uk, &mask);
    const bool short_of_k = hc < static_cast<int>(nc.k);
    d.kd_q[po + i] = make_uint2(static_cast<uint32_t>(uk), mask);
    d.kd_lcnt[po + i] = short_of_k ? 0u : static_cast<uint32_t>(__popc(mask));
    d.kd_dflag[po + i] = short_of_k ? 1u : 0u;
}

// Leftover entries (ascending depth: the bottom of the stack first) of the
// queries that reached k, in query order; the hit slots of the others, in
// their order.
__global__ void k_kd_expand(DevBuf d, Layout L) {
    int tile, f;
    xcd_tile_frame(tile, f);
    const FrameDesc& fd = d.frames[f];
    const int n = fd.n_raw;
    const int i = tile * blockDim.x + threadIdx.x;
    if (i >= n || kd_frame_failed(d, f)) return;
    const size_t po = (size_t)f * L.N;
    const size_t cap = (size_t)L.N * kKdEntriesPerPoint;
    if (d.kd_dflag[po + i] == 0) {
        const uint2 q = d.kd_q[po + i];
        uint32_t mask = q.y;
        if (!mask) return;
        const uint32_t o = d.kd_loff[po + i];
        if (o + __popc(mask) > cap) {
            atomicOr(&d.fs[f * kFsCount + kFsNoiseErr], kErrNoiseCapacity);
            return;
        }
        uint64_t* out = d.kd_csr + po * kKdEntriesPerPoint + o;
        const int uk = static_cast<int>(q.x);
        int b = 0, e = n, j = 0;
        for (int dep = 0; mask && dep < 32; ++dep) {
            const int m = b + (e - b) / 2;
            if (mask & (1u << dep)) {
                out[j++] = kd_entry(m + 1, e, dep + 1);
                mask &= ~(1u << dep);
            }
            if (uk < m)
                e = m;
            else
                b = m + 1;
        }
        return;
    }
    const NoiseConst nc = *d.nc;
    const int gi = fd.pt_off + i;
    const float tx = d.rx[gi], ty = d.ry[gi], tz = d.rz[gi];
    int hits[kKdMaxK];
    int uk;
    uint32_t mask;
    const int hc = kd_own_search(d.kd + po, n, tx, ty, tz, kd_radius_sqr(nc, tx, ty), static_cast<int>(nc.k), hits,
                                 &uk, &mask);
    int32_t* rec = d.kd_drec + (po + d.kd_dpos[po + i]) * kKdRecInts;
    rec[0] = i;
    rec[1] = hc;
    for (int h = 0; h < kKdMaxK - 1; ++h) rec[2 + h] = h < hc ? hits[h] : -1;
}

namespace {

// Position of slot u in the pre-order of the subtree [b, e): u1 before u2.
__device__ __forceinline__ bool kd_pre_before(int b, int e, int u1, int u2) {
    while (b < e) {
        const int m = b + (e - b) / 2;
        if (u1 == m) return true;
        if (u2 == m) return false;
        const bool l1 = u1 < m, l2 = u2 < m;
        if (l1 != l2) return l1;
        if (l1)
            e = m;
        else
            b = m + 1;
    }
    return u1 < u2;
}

}  // namespace

// The stale stack of one frame (one workgroup per frame).  The short-of-k
// queries are taken in input order.  Before query q, the stack holds what
// the earlier short queries left plus the leftover entries of every query
// between the previous short query and q (a contiguous range of the
// leftover array, the newest part of the stack).  From the top, an entry
// counts the recorded hits whose slots lie in its range; the search stops
// inside the entry where the count reaches k, and that entry is replaced by
// the pending right children of the ancestors of the stopping hit.
__global__ __launch_bounds__(kKdResolveThreads) void k_kd_resolve(DevBuf d, Layout L) {
    __shared__ uint64_t lds[kKdResolveThreads / kWave + 1];
    __shared__ int brk_pos, brk_rem;
    __shared__ uint64_t brk_entry;
    __shared__ int rec_sh[kKdRecInts];
    __shared__ int rem_count;
    const int f = blockIdx.x;
    const FrameDesc& fd = d.frames[f];
    const int tid = threadIdx.x;
    if (kd_frame_failed(d, f)) return;
    const size_t po = (size_t)f * L.N;
    const size_t cap = (size_t)L.N * kKdEntriesPerPoint;
    const int nd = static_cast<int>(d.fs[f * kFsCount + kFsNoiseHard]);
    const uint64_t* csr = d.kd_csr + po * kKdEntriesPerPoint;
    uint64_t* live = d.kd_live + po * kKdEntriesPerPoint;
    const float4* T = d.kd + po;
    const NoiseConst nc = *d.nc;
    const int k = static_cast<int>(nc.k);
    size_t top = 0;        // live entries
    uint32_t pushed = 0;   // leftover entries of the queries before the current one
    uint32_t noise_count = 0;
    bool overflow = false;
    for (int j = 0; j < nd && !overflow; ++j) {
        if (tid < kKdRecInts) rec_sh[tid] = d.kd_drec[(po + j) * kKdRecInts + tid];
        __syncthreads();
        const int q = rec_sh[0];
        const int hc = rec_sh[1];
        int hs[kKdMaxK - 1];
#pragma unroll
        for (int h = 0; h < kKdMaxK - 1; ++h) hs[h] = rec_sh[2 + h];
        const uint32_t cs = pushed, ce = d.kd_loff[po + q];
        pushed = ce;
        int need = k - hc;
        bool rescued = false;
        // Region 0: the new leftover range [cs, ce); region 1: the live stack.
        for (int region = 0; region < 2 && !rescued && need > 0; ++region) {
            const uint64_t* src = region == 0 ? csr : live;
            const size_t lo = region == 0 ? cs : 0;
            size_t pos = region == 0 ? ce : top;
            while (pos > lo) {
                const size_t p = pos - 1 - tid;
                int h = 0;
                uint64_t ent = 0;
                if (tid < static_cast<int>(pos - lo) && p >= lo) {
                    ent = src[p];
                    int b, e, dep;
                    kd_entry_split(ent, b, e, dep);
#pragma unroll
                    for (int x = 0; x < kKdMaxK - 1; ++x) h += (x < hc && hs[x] >= b && hs[x] < e) ? 1 : 0;
                }
                uint64_t total;
                const uint64_t ex = block_exclusive_sum<kKdResolveThreads>(static_cast<uint64_t>(h), lds, &total);
                if (static_cast<int>(total) < need) {
                    need -= static_cast<int>(total);
                    pos = pos > lo + kKdResolveThreads ? pos - kKdResolveThreads : lo;
                    continue;
                }
                if (h > 0 && static_cast<int>(ex) < need && static_cast<int>(ex) + h >= need) {
                    brk_pos = static_cast<int>(p);
                    brk_rem = need - static_cast<int>(ex);
                    brk_entry = ent;
                }
                __syncthreads();
                const size_t bp = static_cast<size_t>(brk_pos);
                // The entries below the breaking one stay; from the leftover
                // range they move onto the live stack.
                if (region == 0) {
                    const size_t cnt = bp - cs;
                    if (top + cnt + 32 > cap) {
                        overflow = true;
                    } else {
                        for (size_t t = tid; t < cnt; t += kKdResolveThreads) live[top + t] = csr[cs + t];
                        top += cnt;
                    }
                } else {
                    top = bp;
                }
                __syncthreads();
                if (!overflow && tid < kWave) {
                    int b, e, dep;
                    kd_entry_split(brk_entry, b, e, dep);
                    // The stopping hit: the brk_rem-th recorded hit of this
                    // subtree in pre-order.
                    int u = -1;
                    {
                        int in[kKdMaxK - 1];
                        int m = 0;
#pragma unroll
                        for (int x = 0; x < kKdMaxK - 1; ++x)
                            if (x < hc && hs[x] >= b && hs[x] < e) in[m++] = hs[x];
                        for (int a = 1; a < m; ++a) {
                            const int v = in[a];
                            int c = a - 1;
                            while (c >= 0 && kd_pre_before(b, e, v, in[c])) {
                                in[c + 1] = in[c];
                                --c;
                            }
                            in[c + 1] = v;
                        }
                        u = in[brk_rem - 1];
                    }
                    // Lane l looks at the node l steps below the entry on the
                    // path to u: its right child is pending when u lies in
                    // its left subtree and both children were pushed.
                    const int gi = fd.pt_off + q;
                    const float tx = d.rx[gi], ty = d.ry[gi], tz = d.rz[gi];
                    const float prox = kd_radius_sqr(nc, tx, ty);
                    int ab = b, ae = e;
                    const int l = tid;
                    bool valid = true;
                    for (int s2 = 0; s2 < l; ++s2) {
                        const int m = ab + (ae - ab) / 2;
                        if (m == u) {
                            valid = false;
                            break;
                        }
                        if (u < m)
                            ae = m;
                        else
                            ab = m + 1;
                    }
                    bool push = false;
                    uint64_t out = 0;
                    if (valid && ab < ae) {
                        const int m = ab + (ae - ab) / 2;
                        if (m != u && u < m && m + 1 < ae) {
                            const int adep = dep + l;
                            const int ax = adep % 3;
                            const float4 nd2 = T[m];
                            const float delta = kd_key(nd2, ax) - (ax == 0 ? tx : (ax == 1 ? ty : tz));
                            if (delta * delta <= prox) {
                                push = true;
                                out = kd_entry(m + 1, ae, adep + 1);
                            }
                        }
                    }
                    const uint64_t bal = __ballot(push);
                    if (push) live[top + lanes_below(bal)] = out;
                    if (tid == 0) rem_count = static_cast<int>(__popcll(bal));
                }
                __syncthreads();
                if (!overflow) top += static_cast<size_t>(rem_count);
                rescued = true;
                break;
            }
            if (!rescued && region == 1) top = 0;
        }
        if (!rescued) {
            if (tid == 0) {
                d.noise[po + q] = 1;
                d.obs_flag[po + q] = 0u;
            }
            ++noise_count;
        }
        __syncthreads();
    }
    if (tid == 0) {
        d.fs[f * kFsCount + kFsNoise] = noise_count;
        if (overflow) atomicOr(&d.fs[f * kFsCount + kFsNoiseErr], kErrNoiseCapacity);
    }
}

int kd_small_level_host(int n) {
    int l = 0;
    while ((n >> l) > kKdSmall) ++l;
    return l;
}

void launch_noise_kdtree(lp_context* c, int B, bool compact) {
    DevBuf& d = c->d;
    const Layout& L = c->L;
    hipStream_t st = c->stream;
    const int nb = (L.N + 255) / 256;
    stage_mark(c, "noise");
    hipLaunchKernelGGL(k_kd_init, dim3(nb, B), dim3(256), 0, st, d, L);
    const int small = kd_small_level_host(L.N);
    for (int l = 0; l < small; ++l)
        hipLaunchKernelGGL(k_kd_split_big, dim3(1u << l, B), dim3(kKdBigThreads), 0, st, d, L, l);
    hipLaunchKernelGGL(k_kd_split_small, dim3(1u << small, B), dim3(kKdSmallThreads), 0, st, d, L);
    stage_mark(c, "noise_count");
    hipLaunchKernelGGL(k_kd_query, dim3(nb, B), dim3(256), 0, st, d, L);
    launch_scan(c, B, d.kd_lcnt, d.kd_loff, L.N, kCountRawPoints, 0, -1);
    launch_scan(c, B, d.kd_dflag, d.kd_dpos, L.N, kCountRawPoints, 0, kFsNoiseHard);
    hipLaunchKernelGGL(k_kd_expand, dim3(nb, B), dim3(256), 0, st, d, L);
    stage_mark(c, "noise_search");
    hipLaunchKernelGGL(k_kd_resolve, dim3(B), dim3(kKdResolveThreads), 0, st, d, L);
    stage_mark(c, "noise_compact");
    if (compact) launch_noise_compact(c, B);
}

}  // namespace lp
