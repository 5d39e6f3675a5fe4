// lp_ring.hip -- ring indices for scan-ordered clouds.
//
// Reference: Dataloader::addRingInfo (src/dataloader/src/dataloader.cpp:68-137):
// walking the cloud in order, the ring starts at 63 and drops by one at every
// step from the fourth azimuth quadrant to the first while it is above 0.
// Closed form: ring_i = 63 - min(63, number of such steps at indices <= i),
// so one flag kernel and a prefix sum replace the sequential walk.
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

}  // namespace lp
