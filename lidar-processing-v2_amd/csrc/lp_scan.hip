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

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(DevBuf d, Layout L, const uint32_t* in, size_t stride,
                                                              ScanCount kind, int slot) {
    __shared__ uint32_t lds[kScanThreads / kWave + 1];
    const int f = blockIdx.y;
    const int n = scan_count(d, L, f, kind, slot);
    const int base = blockIdx.x * kScanTile;
    if (base >= n) return;
    const uint32_t* a = in + (size_t)f * stride;
    uint32_t s = 0;
    const int b = base + threadIdx.x * kScanItems;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (b + k < n) s += a[b + k];
    uint32_t total;
    block_exclusive_sum<kScanThreads>(s, lds, &total);
    if (threadIdx.x == 0) d.blk[(size_t)f * L.blk_stride + blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_blocks(DevBuf d, Layout L, ScanCount kind, int slot, int total_slot) {
    __shared__ uint32_t lds[1024 / kWave + 1];
    const int f = blockIdx.x;
    const int n = scan_count(d, L, f, kind, slot);
    const int nb = (n + kScanTile - 1) / kScanTile;
    uint32_t* blk = d.blk + (size_t)f * L.blk_stride;
    uint32_t carry = 0;
    for (int base = 0; base < nb; base += 1024) {
        const int i = base + threadIdx.x;
        const uint32_t v = i < nb ? blk[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_sum<1024>(v, lds, &total);
        if (i < nb) blk[i] = ex + carry;
        carry += total;
    }
    if (threadIdx.x == 0 && total_slot >= 0) d.fs[f * kFsCount + total_slot] = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_down(DevBuf d, Layout L, const uint32_t* in, uint32_t* out,
                                                            size_t stride, ScanCount kind, int slot) {
    __shared__ uint32_t lds[kScanThreads / kWave + 1];
    const int f = blockIdx.y;
    const int n = scan_count(d, L, f, kind, slot);
    const int base = blockIdx.x * kScanTile;
    if (base >= n) return;
    const uint32_t* a = in + (size_t)f * stride;
    uint32_t* o = out + (size_t)f * stride;
    uint32_t v[kScanItems];
    uint32_t s = 0;
    const int b = base + threadIdx.x * kScanItems;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (b + k < n) ? a[b + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t ex = block_exclusive_sum<kScanThreads>(s, lds, &total) + d.blk[(size_t)f * L.blk_stride + blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (b + k < n) o[b + k] = ex;
        ex += v[k];
    }
}

void launch_scan(lp_context* c, int B, const uint32_t* in, uint32_t* out, size_t stride, ScanCount kind,
                 int count_slot, int total_slot) {
    const Layout& L = c->L;
    const int maxn = static_cast<int>(stride);
    const int nb = (maxn + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb, B), dim3(kScanThreads), 0, c->stream, c->d, L, in, stride, kind,
                       count_slot);
    hipLaunchKernelGGL(k_scan_blocks, dim3(B), dim3(1024), 0, c->stream, c->d, L, kind, count_slot, total_slot);
    hipLaunchKernelGGL(k_scan_down, dim3(nb, B), dim3(kScanThreads), 0, c->stream, c->d, L, in, out, stride, kind,
                       count_slot);
}

}  // namespace lp
