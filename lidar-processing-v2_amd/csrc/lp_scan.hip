// lp_scan.hip -- reduce-then-scan exclusive prefix sums over per-frame
// arrays whose lengths are only known on the device (point counts, queue
// lengths, obstacle counts), so the launch shape stays fixed for graphs.
#include "lp_kernels.h"

namespace lp {

using namespace lpd;

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ int scan_count(const DevBuf& d, const Layout& L, int f, ScanCount kind, int slot) {
    const FrameDesc& fd = d.frames[f];
    if (kind == kCountPoints) return fd.n;
    if (kind == kCountPixels) return fd.sc.H * fd.sc.W;
    if (kind == kCountHash) return L.hash_cap;
    if (kind == kCountRawPoints) return fd.n_raw;
    return static_cast<int>(d.fs[f * kFsCount + slot]);
}

// Element i of frame f: in[f * in_stride + i], or with a gather index
// in[(f * in_stride + gather[f * stride + i]) * elem], elem words apart
// (a field of a record array).
struct ScanIn {
    const uint32_t* in;
    size_t in_stride;
    const int32_t* gather;
    int elem;
};

__device__ __forceinline__ uint32_t scan_value(const ScanIn& si, size_t stride, int f, int i) {
    if (!si.gather) return si.in[(size_t)f * si.in_stride + i];
    return si.in[((size_t)f * si.in_stride + si.gather[(size_t)f * stride + i]) * si.elem];
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(DevBuf d, Layout L, ScanIn si, size_t stride,
                                                              ScanCount kind, int slot) {
    __shared__ uint32_t lds[kScanThreads / kWave + 1];
    const int f = blockIdx.y;
    const int n = scan_count(d, L, f, kind, slot);
    const int base = blockIdx.x * kScanTile;
    if (base >= n) return;
    uint32_t s = 0;
    const int b = base + threadIdx.x * kScanItems;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (b + k < n) s += scan_value(si, stride, f, b + k);
    uint32_t total;
    block_exclusive_sum<kScanThreads>(s, lds, &total);
    if (threadIdx.x == 0) d.blk[(size_t)f * L.blk_stride + blockIdx.x] = total;
}

// Down pass: each tile adds the totals of the tiles before it (read straight
// from the reduce pass, at most blk_stride of them) to its own exclusive
// scan; the last tile writes the frame total.  Two launches per scan.
// With out_gathered the result for element i goes where its input came from,
// into the record field `out` points at (out is that field of record 0).
__global__ __launch_bounds__(kScanThreads) void k_scan_down(DevBuf d, Layout L, ScanIn si, uint32_t* out,
                                                            size_t stride, ScanCount kind, int slot, int total_slot,
                                                            bool out_gathered) {
    __shared__ uint32_t lds[kScanThreads / kWave + 1];
    const int f = blockIdx.y;
    const int n = scan_count(d, L, f, kind, slot);
    const int tile = blockIdx.x;
    const int base = tile * kScanTile;
    if (base >= n) {
        if (tile == 0 && threadIdx.x == 0 && total_slot >= 0) d.fs[f * kFsCount + total_slot] = 0;
        return;
    }
    const uint32_t* blk = d.blk + (size_t)f * L.blk_stride;
    uint32_t pre = 0;
    for (int i = threadIdx.x; i < tile; i += kScanThreads) pre += blk[i];
    uint32_t pre_total;
    block_exclusive_sum<kScanThreads>(pre, lds, &pre_total);
    uint32_t* o = out + (size_t)f * stride;
    uint32_t v[kScanItems];
    uint32_t s = 0;
    const int b = base + threadIdx.x * kScanItems;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (b + k < n) ? scan_value(si, stride, f, b + k) : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t ex = block_exclusive_sum<kScanThreads>(s, lds, &total) + pre_total;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (b + k < n) {
            if (out_gathered)
                out[((size_t)f * si.in_stride + si.gather[(size_t)f * stride + b + k]) * si.elem] = ex;
            else
                o[b + k] = ex;
        }
        ex += v[k];
    }
    if (threadIdx.x == 0 && total_slot >= 0 && base + kScanTile >= n)
        d.fs[f * kFsCount + total_slot] = pre_total + total;
}

// Scans whose counts are cluster-sized (a few hundred to a few thousand):
// one 1024-thread workgroup per frame walks the frame in tiles of 4096 with
// a running carry, one launch instead of a reduce and a down pass.  In-place
// use (in == out) is safe: a tile is loaded before the block sum's barriers.
constexpr int kSmallThreads = 1024;
constexpr int kSmallItems = 4;

__global__ __launch_bounds__(kSmallThreads) void k_scan_small(DevBuf d, Layout L, ScanIn si, uint32_t* out,
                                                              size_t stride, ScanCount kind, int slot,
                                                              int total_slot) {
    __shared__ uint32_t lds[kSmallThreads / kWave + 1];
    const int f = blockIdx.x;
    const int n = scan_count(d, L, f, kind, slot);
    uint32_t* o = out + (size_t)f * stride;
    uint32_t carry = 0;
    for (int base = 0; base < n; base += kSmallThreads * kSmallItems) {
        uint32_t v[kSmallItems];
        uint32_t s = 0;
        const int b = base + threadIdx.x * kSmallItems;
#pragma unroll
        for (int k = 0; k < kSmallItems; ++k) {
            v[k] = (b + k < n) ? scan_value(si, stride, f, b + k) : 0u;
            s += v[k];
        }
        uint32_t total;
        uint32_t ex = block_exclusive_sum<kSmallThreads>(s, lds, &total) + carry;
#pragma unroll
        for (int k = 0; k < kSmallItems; ++k) {
            if (b + k < n) o[b + k] = ex;
            ex += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0 && total_slot >= 0) d.fs[f * kFsCount + total_slot] = carry;
}

void launch_scan_small(lp_context* c, int B, const uint32_t* in, uint32_t* out, size_t stride, ScanCount kind,
                       int count_slot, int total_slot) {
    hipLaunchKernelGGL(k_scan_small, dim3(B), dim3(kSmallThreads), 0, c->stream, c->d, c->L,
                       ScanIn{in, stride, nullptr, 1}, out, stride, kind, count_slot, total_slot);
}

static void launch_scan_in(lp_context* c, int B, ScanIn si, uint32_t* out, size_t stride, ScanCount kind,
                           int count_slot, int total_slot, bool out_gathered = false) {
    const Layout& L = c->L;
    const int maxn = static_cast<int>(stride);
    const int nb = (maxn + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb, B), dim3(kScanThreads), 0, c->stream, c->d, L, si, stride, kind,
                       count_slot);
    hipLaunchKernelGGL(k_scan_down, dim3(nb, B), dim3(kScanThreads), 0, c->stream, c->d, L, si, out, stride, kind,
                       count_slot, total_slot, out_gathered);
}

void launch_scan(lp_context* c, int B, const uint32_t* in, uint32_t* out, size_t stride, ScanCount kind,
                 int count_slot, int total_slot) {
    launch_scan_in(c, B, ScanIn{in, stride, nullptr, 1}, out, stride, kind, count_slot, total_slot);
}

void launch_scan_gather(lp_context* c, int B, const uint32_t* in, size_t in_stride, int elem, const int32_t* gather,
                        uint32_t* out, size_t stride, ScanCount kind, int count_slot, int total_slot) {
    launch_scan_in(c, B, ScanIn{in, in_stride, gather, elem}, out, stride, kind, count_slot, total_slot);
}

void launch_scan_gather_into(lp_context* c, int B, const uint32_t* in, size_t in_stride, int elem,
                             const int32_t* gather, uint32_t* field, size_t stride, ScanCount kind, int count_slot) {
    launch_scan_in(c, B, ScanIn{in, in_stride, gather, elem}, field, stride, kind, count_slot, -1, true);
}

}  // namespace lp
