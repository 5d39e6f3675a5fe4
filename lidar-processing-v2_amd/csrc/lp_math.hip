// lp_math.hip -- evaluation entry point for the libm restatements, so tests
// can compare the device code (and the same source compiled for the host)
// with the host C library the reference links.
#include <string>

#include "lp_kernels.h"

namespace lp {

using namespace lpd;

__host__ __device__ __forceinline__ float eval_math(int which, float a, float b) {
    switch (which) {
        case LP_MATH_EXPF: return glibc_expf(a);
        case LP_MATH_ATANF: return fdlibm_atanf(a);
        case LP_MATH_ATAN2F: return fdlibm_atan2f(a, b);
        default: return atan2_approx(a, b);
    }
}

__global__ void k_math(int which, const float* a, const float* b, float* out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = eval_math(which, a[i], b ? b[i] : 0.0f);
}

// Every float with bit pattern in [lo, lo + count): device result bits that
// differ from `ref` (host results for the same range) are counted.
__global__ void k_math_range(int which, uint32_t lo, uint32_t count, const uint32_t* ref, unsigned long long* bad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float x = u2f(lo + i);
    if (f2u(eval_math(which, x, 1.0f)) != ref[i]) atomicAdd(bad, 1ull);
}

}  // namespace lp

using namespace lp;

extern "C" int lp_math_eval(int32_t which, const float* a, const float* b, float* out, int32_t n,
                            int32_t on_device) {
    if (which < 0 || which > LP_MATH_ATAN2_APPROX || n < 0 || (n > 0 && (!a || !out))) return LP_ERR_INVALID_ARGUMENT;
    if (!on_device) {
        for (int32_t i = 0; i < n; ++i) out[i] = eval_math(which, a[i], b ? b[i] : 0.0f);
        return LP_OK;
    }
    if (n == 0) return LP_OK;
    float *da = nullptr, *db = nullptr, *dout = nullptr;
    const size_t bytes = static_cast<size_t>(n) * 4;
    hipError_t e = hipMalloc(&da, bytes);
    if (e == hipSuccess && b) e = hipMalloc(&db, bytes);
    if (e == hipSuccess) e = hipMalloc(&dout, bytes);
    if (e == hipSuccess) e = hipMemcpy(da, a, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess && b) e = hipMemcpy(db, b, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_math, dim3((n + 255) / 256), dim3(256), 0, nullptr, which, da, db, dout, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dout);
    return e == hipSuccess ? LP_OK : LP_ERR_HIP;
}

extern "C" int lp_math_check_range(int32_t which, uint32_t lo, uint32_t count, const uint32_t* ref_bits,
                                   uint64_t* mismatches) {
    if (which < 0 || which > LP_MATH_ATANF || !ref_bits || !mismatches) return LP_ERR_INVALID_ARGUMENT;
    *mismatches = 0;
    if (count == 0) return LP_OK;
    uint32_t* dref = nullptr;
    unsigned long long* dbad = nullptr;
    const size_t bytes = static_cast<size_t>(count) * 4;
    hipError_t e = hipMalloc(&dref, bytes);
    if (e == hipSuccess) e = hipMalloc(&dbad, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemcpy(dref, ref_bits, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(dbad, 0, sizeof(unsigned long long));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_math_range, dim3((count + 255) / 256), dim3(256), 0, nullptr, which, lo, count, dref,
                           dbad);
        e = hipGetLastError();
    }
    unsigned long long bad = 0;
    if (e == hipSuccess) e = hipMemcpy(&bad, dbad, sizeof(bad), hipMemcpyDeviceToHost);
    (void)hipFree(dref);
    (void)hipFree(dbad);
    *mismatches = bad;
    return e == hipSuccess ? LP_OK : LP_ERR_HIP;
}
