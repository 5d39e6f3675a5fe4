// lp_ring.hip -- ring indices for scan-ordered clouds.
//
// Reference: Dataloader::addRingInfo (src/dataloader/src/dataloader.cpp:68-137):
// walking the cloud in order, the ring starts at 63 and drops by one at every
// step from the fourth azimuth quadrant to the first while it is above 0.
// Closed form: ring_i = 63 - min(63, number of such steps at indices <= i),
// so one flag kernel and a prefix sum replace the sequential walk.
#include <mutex>

#include "lp_kernels.h"

namespace lp {

using namespace lpd;

__device__ __forceinline__ int quadrant(float x, float y) {
    float az = fdlibm_atan2f(y, x);
    az = (az < 0) ? (az + 2.0F * 3.14159274101257324219f) : az;  // 2.0F * M_PIf
    if (az < 1.57079637050628662109f) return 0;                  // M_PI_2f
    if (az < 3.14159274101257324219f) return 1;                  // M_PIf
    if (az < 1.5F * 3.14159274101257324219f) return 2;           // 1.5F * M_PIf
    return 3;
}

// flag[i] = 1 when point i is in the first quadrant and point i-1 in the
// fourth (the previous quadrant starts as the first, dataloader.cpp:86).
__global__ void k_ring_flags(const float* x, const float* y, int n, uint32_t* flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t f = 0;
    if (i > 0 && quadrant(x[i], y[i]) == 0 && quadrant(x[i - 1], y[i - 1]) == 3) f = 1;
    flag[i] = f;
}

// scan holds the exclusive prefix of the flags.
__global__ void k_ring_assign(const uint32_t* flag, const uint32_t* scan, int n, uint16_t* ring) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = scan[i] + flag[i];
    ring[i] = static_cast<uint16_t>(63 - min(63u, c));
}

void launch_ring_info(lp_context* c, int n) {
    DevBuf& d = c->d;
    const int nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_ring_flags, dim3(nb), dim3(256), 0, c->stream, d.x, d.y, n, d.obs_flag);
    launch_scan(c, 1, d.obs_flag, d.cl_flag, c->L.N, kCountPoints, 0, -1);
    hipLaunchKernelGGL(k_ring_assign, dim3(nb), dim3(256), 0, c->stream, d.obs_flag, d.cl_flag, n, d.ring);
}

// ---------------------------------------------------------------------------
// Allocation-free form (lp_add_ring_info_noalloc): the caller's device arrays
// in and out, and the per-tile transition counts in a file-scope device
// array instead of context memory.  Three launches: tile counts, one
// workgroup scanning them in place, then each tile's rings from its own
// block scan plus its tile offset.
constexpr int kRingTile = 1024;                       // points per tile (256 threads x 4)
constexpr int kRingMaxTiles = 16384;
constexpr int kRingMaxPoints = kRingTile * kRingMaxTiles;
constexpr int kRingMaxDevices = 64;

// Per-tile transition counts, then their exclusive prefix.  Every call on a
// device shares this array; each call writes all entries it reads.
static __device__ __align__(256) uint32_t scratch[kRingMaxTiles];

__device__ __forceinline__ uint32_t ring_flag(const float* x, const float* y, int i) {
    return (i > 0 && quadrant(x[i], y[i]) == 0 && quadrant(x[i - 1], y[i - 1]) == 3) ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_ring_tile_counts(const float* x, const float* y, int n, uint32_t* tiles) {
    __shared__ uint32_t lds[256 / kWave + 1];
    const int b = blockIdx.x * kRingTile + threadIdx.x * 4;
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (b + k < n) c += ring_flag(x, y, b + k);
    uint32_t total;
    block_exclusive_sum<256>(c, lds, &total);
    if (threadIdx.x == 0) tiles[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_ring_tile_scan(uint32_t* tiles, int num_tiles) {
    __shared__ uint32_t lds[1024 / kWave + 1];
    constexpr int kPer = kRingMaxTiles / 1024;
    const int b = threadIdx.x * kPer;
    uint32_t v[kPer];
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        v[k] = (b + k < num_tiles) ? tiles[b + k] : 0u;
        local += v[k];
    }
    uint32_t total;
    uint32_t ex = block_exclusive_sum<1024>(local, lds, &total);
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        if (b + k < num_tiles) tiles[b + k] = ex;
        ex += v[k];
    }
}

__global__ __launch_bounds__(256) void k_ring_tile_assign(const float* x, const float* y, int n,
                                                          const uint32_t* tiles, uint16_t* ring) {
    __shared__ uint32_t lds[256 / kWave + 1];
    const int b = blockIdx.x * kRingTile + threadIdx.x * 4;
    uint32_t f[4];
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        f[k] = (b + k < n) ? ring_flag(x, y, b + k) : 0u;
        c += f[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_sum<256>(c, lds, &total) + tiles[blockIdx.x];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        run += f[k];
        if (b + k < n) ring[b + k] = static_cast<uint16_t>(63 - min(63u, run));
    }
}

namespace {

std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kRingMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy of the array).
int scratch_for_current_device(uint32_t** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kRingMaxDevices)
        return set_last_error(LP_ERR_HIP, "hipGetDevice failed or device ordinal too large");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return set_last_error(LP_ERR_HIP, "hipGetSymbolAddress(scratch) failed");
        g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch_ptr[dev];
    return LP_OK;
}

bool is_device_pointer(const void* p) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

}  // namespace

}  // namespace lp

extern "C" int lp_add_ring_info_noalloc(const float* x, const float* y, int32_t n, uint16_t* ring, void* stream) {
    using namespace lp;
    if (n < 0 || (n > 0 && (!x || !y || !ring))) return set_last_error(LP_ERR_INVALID_ARGUMENT, "null argument");
    if (n > kRingMaxPoints)
        return set_last_error(LP_ERR_CAPACITY, "lp_add_ring_info_noalloc: more than 16777216 points");
    if (n == 0) return LP_OK;
    if (!is_device_pointer(x) || !is_device_pointer(y) || !is_device_pointer(ring))
        return set_last_error(LP_ERR_INVALID_ARGUMENT, "lp_add_ring_info_noalloc: x, y and ring must be device memory");
    uint32_t* tiles = nullptr;
    const int st = scratch_for_current_device(&tiles);
    if (st != LP_OK) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int nt = (n + kRingTile - 1) / kRingTile;
    hipLaunchKernelGGL(k_ring_tile_counts, dim3(nt), dim3(256), 0, s, x, y, n, tiles);
    hipLaunchKernelGGL(k_ring_tile_scan, dim3(1), dim3(1024), 0, s, tiles, nt);
    hipLaunchKernelGGL(k_ring_tile_assign, dim3(nt), dim3(256), 0, s, x, y, n, tiles, ring);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_last_error(LP_ERR_HIP, hipGetErrorString(e));
    return LP_OK;
}
