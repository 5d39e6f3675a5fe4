// lp_device.h -- device math and wavefront helpers for gfx950.
//
// Everything here must round exactly like the host code the reference
// runs: the library is compiled with -ffp-contract=off and IEEE division
// and square root, and the libm functions the reference calls are restated
// from the published algorithms glibc 2.35 uses on x86-64:
//   expf   -- the exp2f-table method of ARM Optimized Routines (Szabolcs
//             Nagy, 2017): N=32 table of 2^(i/32), degree-3 polynomial.
//             Equal to host expf for every float in [-20, 20] (exhaustive).
//   atanf  -- Sun fdlibm s_atanf.c (float): 4-way argument reduction and an
//             11-term odd/even split polynomial.  The constants are the
//             decimal literals of that source (the compiler uses those, not
//             the hex values in its comments).  Equal to host atanf for all
//             2^32 inputs (exhaustive).
//   atan2f -- Sun fdlibm e_atan2f.c on top of that atanf.  Equal to host
//             atan2f on 4e8 random and structured pairs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lpd {

#define LPD_HD __host__ __device__ __forceinline__

LPD_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
LPD_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

// ---- expf (exp2f-table method) -------------------------------------------
// tab[i] = bits(round_to_double(2^(i/32))) - (i << 47)
#define LPD_EXP2_TAB                                                                       \
    {0x3ff0000000000000ULL, 0x3fefd9b0d3158574ULL, 0x3fefb5586cf9890fULL, 0x3fef9301d0125b51ULL, \
     0x3fef72b83c7d517bULL, 0x3fef54873168b9aaULL, 0x3fef387a6e756238ULL, 0x3fef1e9df51fdee1ULL, \
     0x3fef06fe0a31b715ULL, 0x3feef1a7373aa9cbULL, 0x3feedea64c123422ULL, 0x3feece086061892dULL, \
     0x3feebfdad5362a27ULL, 0x3feeb42b569d4f82ULL, 0x3feeab07dd485429ULL, 0x3feea47eb03a5585ULL, \
     0x3feea09e667f3bcdULL, 0x3fee9f75e8ec5f74ULL, 0x3feea11473eb0187ULL, 0x3feea589994cce13ULL, \
     0x3feeace5422aa0dbULL, 0x3feeb737b0cdc5e5ULL, 0x3feec49182a3f090ULL, 0x3feed503b23e255dULL, \
     0x3feee89f995ad3adULL, 0x3feeff76f2fb5e47ULL, 0x3fef199bdd85529cULL, 0x3fef3720dcef9069ULL, \
     0x3fef5818dcfba487ULL, 0x3fef7c97337b9b5fULL, 0x3fefa4afa2a490daULL, 0x3fefd0765b6e4540ULL}
__device__ __constant__ static const uint64_t kExp2TabDev[32] = LPD_EXP2_TAB;
static const uint64_t kExp2TabHost[32] = LPD_EXP2_TAB;

LPD_HD float glibc_expf(float x) {
    const uint32_t abstop = (f2u(x) >> 20) & 0x7ff;
    if (abstop >= (0x42b00000u >> 20)) {  // |x| >= 88
        if (f2u(x) == 0xff800000u) return 0.0f;
        if (abstop >= (0x7f800000u >> 20)) return x + x;
        if (x > 0x1.62e42ep6f) return u2f(0x7f800000u);
        if (x < -0x1.9fe368p6f) return 0.0f;
    }
    const double xd = static_cast<double>(x);
    const double inv_ln2_n = 0x1.71547652b82fep+5;
    const double shift = 0x1.8p+52;
    double z = inv_ln2_n * xd;
    double kd = z + shift;
    const uint64_t ki = __builtin_bit_cast(uint64_t, kd);
    kd -= shift;
    const double r = z - kd;
#ifdef __HIP_DEVICE_COMPILE__
    uint64_t t = kExp2TabDev[ki % 32];
#else
    uint64_t t = kExp2TabHost[ki % 32];
#endif
    t += ki << 47;
    const double s = __builtin_bit_cast(double, t);
    z = 0x1.c6af84b912394p-20 * r + 0x1.ebfce50fac4f3p-13;
    const double r2 = r * r;
    double y = 0x1.62e42ff0c52d6p-6 * r + 1.0;
    y = z * r2 + y;
    y = y * s;
    return static_cast<float>(y);
}

// ---- atanf / atan2f (fdlibm) ----------------------------------------------
LPD_HD float fdlibm_atanf(float x) {
    const float atanhi0 = 4.6364760399e-01f, atanhi1 = 7.8539812565e-01f, atanhi2 = 9.8279368877e-01f,
                atanhi3 = 1.5707962513e+00f;
    const float atanlo0 = 5.0121582440e-09f, atanlo1 = 3.7748947079e-08f, atanlo2 = 3.4473217170e-08f,
                atanlo3 = 7.5497894159e-08f;
    const float aT0 = 3.3333334327e-01f, aT1 = -2.0000000298e-01f, aT2 = 1.4285714924e-01f,
                aT3 = -1.1111110449e-01f, aT4 = 9.0908870101e-02f, aT5 = -7.6918758452e-02f,
                aT6 = 6.6610731184e-02f, aT7 = -5.8335702866e-02f, aT8 = 4.9768779427e-02f,
                aT9 = -3.6531571299e-02f, aT10 = 1.6285819933e-02f;
    const int32_t hx = static_cast<int32_t>(f2u(x));
    const int32_t ix = hx & 0x7fffffff;
    int id;
    if (ix >= 0x4c000000) {  // |x| >= 2^25
        if (ix > 0x7f800000) return x + x;
        return hx > 0 ? atanhi3 + atanlo3 : -atanhi3 - atanlo3;
    }
    if (ix < 0x3ee00000) {  // |x| < 0.4375
        if (ix < 0x31000000) return x;  // |x| < 2^-29
        id = -1;
    } else {
        x = fabsf(x);
        if (ix < 0x3f980000) {
            if (ix < 0x3f300000) {
                id = 0;
                x = (2.0f * x - 1.0f) / (2.0f + x);
            } else {
                id = 1;
                x = (x - 1.0f) / (x + 1.0f);
            }
        } else {
            if (ix < 0x401c0000) {
                id = 2;
                x = (x - 1.5f) / (1.0f + 1.5f * x);
            } else {
                id = 3;
                x = -1.0f / x;
            }
        }
    }
    float z = x * x;
    const float w = z * z;
    const float s1 = z * (aT0 + w * (aT2 + w * (aT4 + w * (aT6 + w * (aT8 + w * aT10)))));
    const float s2 = w * (aT1 + w * (aT3 + w * (aT5 + w * (aT7 + w * aT9))));
    if (id < 0) return x - x * (s1 + s2);
    const float hi = id == 0 ? atanhi0 : id == 1 ? atanhi1 : id == 2 ? atanhi2 : atanhi3;
    const float lo = id == 0 ? atanlo0 : id == 1 ? atanlo1 : id == 2 ? atanlo2 : atanlo3;
    z = hi - ((x * (s1 + s2) - lo) - x);
    return hx < 0 ? -z : z;
}

LPD_HD float fdlibm_atan2f(float y, float x) {
    const float pi_o_4 = 7.8539818525e-01f, pi_o_2 = 1.5707963705e+00f, pi = 3.1415927410e+00f,
                pi_lo = -8.7422776573e-08f, tiny = 1.0e-30f;
    const int32_t hx = static_cast<int32_t>(f2u(x)), ix = hx & 0x7fffffff;
    const int32_t hy = static_cast<int32_t>(f2u(y)), iy = hy & 0x7fffffff;
    if (ix > 0x7f800000 || iy > 0x7f800000) return x + y;
    if (hx == 0x3f800000) return fdlibm_atanf(y);
    const int m = ((hy >> 31) & 1) | ((hx >> 30) & 2);
    if (iy == 0) {
        if (m <= 1) return y;
        return m == 2 ? pi + tiny : -pi - tiny;
    }
    if (ix == 0) return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
    if (ix == 0x7f800000) {
        if (iy == 0x7f800000) {
            switch (m) {
                case 0: return pi_o_4 + tiny;
                case 1: return -pi_o_4 - tiny;
                case 2: return 3.0f * pi_o_4 + tiny;
                default: return -3.0f * pi_o_4 - tiny;
            }
        }
        switch (m) {
            case 0: return 0.0f;
            case 1: return -0.0f;
            case 2: return pi + tiny;
            default: return -pi - tiny;
        }
    }
    if (iy == 0x7f800000) return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
    const int k = (iy - ix) >> 23;
    float z;
    if (k > 60)
        z = pi_o_2 + 0.5f * pi_lo;
    else if (hx < 0 && k < -60)
        z = 0.0f;
    else
        z = fdlibm_atanf(fabsf(y / x));
    switch (m) {
        case 0: return z;
        case 1: return -z;
        case 2: return pi - (z - pi_lo);
        default: return (z - pi_lo) - pi;
    }
}

// ---- atan2Approx (lidar_processing_lib/include/lidar_processing_lib/common.hpp:33-62)
LPD_HD float atan2_approx(float y, float x) {
    const float ax = fabsf(x);
    const float ay = fabsf(y);
    const float mx = (ay < ax) ? ax : ay;  // std::max(ay, ax)
    const float mn = (ax < ay) ? ax : ay;  // std::min(ay, ax)
    const float a = mn / mx;
    const float s = a * a;
    const float c = s * a;
    const float q = s * s;
    float r = 0.024840285F * q + 0.18681418F;
    const float t = -0.094097948F * q - 0.33213072F;
    r = r * s + t;
    r = r * c + a;
    if (ay > ax) r = 1.57079637F - r;
    if (x < 0) r = 3.14159274F - r;
    if (y < 0) r = -r;
    return r;
}

// Order-preserving maps of float / double bit patterns to unsigned integers.
__device__ __forceinline__ uint32_t ord_f32(float f) {
    const uint32_t u = f2u(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ uint64_t ord_f64(double d) {
    const uint64_t u = __builtin_bit_cast(uint64_t, d);
    return (u & 0x8000000000000000ULL) ? ~u : (u | 0x8000000000000000ULL);
}

// ---- wavefront (64 lanes) helpers --------------------------------------
constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return static_cast<int>(__lane_id()); }

// Orders LDS accesses between the lanes of one wavefront.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Number of set bits of `mask` in lanes below this one.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Workgroups go to the 8 XCDs round-robin by linear workgroup id.  For a
// grid of gridDim.x tiles x gridDim.y frames with gridDim.y a multiple of 8,
// this maps the workgroup to a (tile, frame) such that every tile of frame f
// runs on XCD f % 8, so a frame's arrays stay in one XCD's L2 (other grids
// keep the plain mapping).  The mapping is a bijection either way.
__device__ __forceinline__ void xcd_tile_frame(int& tile, int& frame) {
    const int T = static_cast<int>(gridDim.x), B = static_cast<int>(gridDim.y);
    if (B % 8 != 0) {
        tile = static_cast<int>(blockIdx.x);
        frame = static_cast<int>(blockIdx.y);
        return;
    }
    const int p = static_cast<int>(blockIdx.x) + static_cast<int>(blockIdx.y) * T;
    const int q = p >> 3;
    frame = (p & 7) + 8 * (q / T);
    tile = q % T;
}

// Runs of equal `key` among the active lanes of the wavefront (consecutive
// active lanes with one key; points handled by neighbouring lanes mostly
// share a cell, voxel or cluster).  leader: the run's lowest lane; count: the
// run's lanes; rank: active lanes of the run below this one.  A key may form
// several runs when its lanes are not adjacent.
struct WaveRun {
    int leader;
    uint32_t count;
    uint32_t rank;
};

template <typename K>
__device__ __forceinline__ WaveRun wave_runs(K key) {
    const int lane = lane_id();
    const uint64_t active = __ballot(1);
    const uint64_t below = active & ((1ull << lane) - 1ull);
    const int prev = below ? 63 - __clzll(static_cast<long long>(below)) : lane;
    const K pk = __shfl(key, prev, kWave);
    const bool start = !below || pk != key;
    const uint64_t starts = __ballot(start);
    const uint64_t upto = starts & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
    WaveRun r;
    r.leader = 63 - __clzll(static_cast<long long>(upto));
    const uint64_t after = starts & ~(r.leader == 63 ? ~0ull : ((2ull << r.leader) - 1ull));
    const int next = after ? __ffsll(static_cast<unsigned long long>(after)) - 1 : kWave;
    const uint64_t from_leader = ~((1ull << r.leader) - 1ull);
    const uint64_t to_next = next == kWave ? ~0ull : ((1ull << next) - 1ull);
    r.count = static_cast<uint32_t>(__popcll(active & from_leader & to_next));
    r.rank = static_cast<uint32_t>(__popcll(active & from_leader & ((1ull << lane) - 1ull)));
    return r;
}

// atomicAdd(ctr, 1) for every active lane, with each run of lanes that target
// the same counter combined into one atomic (all runs at once).  Returns what
// atomicAdd would have: the counter's value before this lane's increment,
// the lanes of a run taking consecutive values in lane order.
__device__ __forceinline__ uint32_t wave_atomic_inc(uint32_t* ctr) {
    const WaveRun r = wave_runs(reinterpret_cast<unsigned long long>(ctr));
    uint32_t base = 0;
    if (lane_id() == r.leader) base = atomicAdd(ctr, r.count);
    return __shfl(base, r.leader, kWave) + r.rank;
}

// Bitonic sort (ascending) of 64*S u32 keys held in registers, element
// e = s * 64 + lane in slot s: partners within a lane take min / max of two
// registers, others exchange through lane shuffles.
// Lane (lane ^ j) of v.  Partners inside a 16-lane row move by DPP (no LDS
// round trip): quad permutes for j = 1, 2, the rotation by 8 for j = 8, and
// for j = 4 the rotations by 4 and 12 (lane i reads lane (i - 4) mod 16 and
// (i + 4) mod 16), picked by lane bit 2.  Larger j go through a shuffle.
__device__ __forceinline__ uint32_t lane_xor(uint32_t v, int j) {
    const int x = static_cast<int>(v);
    switch (j) {
        case 1:
            return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(x, x, 0xB1, 0xf, 0xf, false));
        case 2:
            return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(x, x, 0x4E, 0xf, 0xf, false));
        case 4: {
            const int a = __builtin_amdgcn_update_dpp(x, x, 0x124, 0xf, 0xf, false);
            const int b = __builtin_amdgcn_update_dpp(x, x, 0x12C, 0xf, 0xf, false);
            return static_cast<uint32_t>((lane_id() & 4) ? a : b);
        }
        case 8:
            return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(x, x, 0x128, 0xf, 0xf, false));
        default:
            return __shfl_xor(v, j, kWave);
    }
}

template <int S>
__device__ __forceinline__ void reg_sort_u32(uint32_t (&key)[S]) {
    const int lane = lane_id();
#pragma unroll
    for (int k = 2; k <= kWave * S; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= kWave) {
                const int js = j / kWave;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    if (s & js) continue;
                    const int t = s | js;
                    const bool asc = ((s * kWave) & k) == 0;
                    const uint32_t lo = min(key[s], key[t]), hi = max(key[s], key[t]);
                    key[s] = asc ? lo : hi;
                    key[t] = asc ? hi : lo;
                }
            } else {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const uint32_t other = lane_xor(key[s], j);
                    const bool asc = ((s * kWave + lane) & k) == 0;
                    const bool lower = (lane & j) == 0;
                    key[s] = (lower == asc) ? min(key[s], other) : max(key[s], other);
                }
            }
        }
    }
}

template <typename T>
__device__ __forceinline__ T wave_inclusive_sum(T v) {
    const int l = lane_id();
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const T t = __shfl_up(v, o, kWave);
        if (l >= o) v += t;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Exclusive prefix sum over a workgroup of NT threads (NT multiple of 64).
// `lds` must hold NT/64 + 1 entries.  Returns this thread's exclusive prefix
// and writes the total to *total.
template <int NT, typename T>
__device__ __forceinline__ T block_exclusive_sum(T v, T* lds, T* total) {
    const int tid = threadIdx.x;
    const int w = tid / kWave;
    const T inc = wave_inclusive_sum(v);
    if (lane_id() == kWave - 1) lds[w] = inc;
    __syncthreads();
    if (tid == 0) {
        T run = 0;
        for (int i = 0; i < NT / kWave; ++i) {
            const T t = lds[i];
            lds[i] = run;
            run += t;
        }
        lds[NT / kWave] = run;
    }
    __syncthreads();
    const T res = inc - v + lds[w];
    *total = lds[NT / kWave];
    __syncthreads();
    return res;
}

// Places in a counter for R rows of flags: one atomic add by the lowest
// flagged lane for the whole wavefront; each flagged lane's place is the
// counter before the add plus the flagged (row, lane) pairs before it, rows
// in order.
template <int R>
__device__ __forceinline__ void wave_reserve(uint32_t* ctr, const bool (&flag)[R], uint32_t (&out)[R]) {
    uint64_t m[R];
    uint64_t any = 0;
    uint32_t total = 0;
#pragma unroll
    for (int s = 0; s < R; ++s) {
        m[s] = __ballot(flag[s]);
        any |= m[s];
        total += static_cast<uint32_t>(__popcll(m[s]));
    }
    uint32_t base = 0;
    if (total) {
        const int first = static_cast<int>(__ffsll(static_cast<unsigned long long>(any))) - 1;
        if (lane_id() == first) base = atomicAdd(ctr, total);
        base = __shfl(base, first, kWave);
    }
#pragma unroll
    for (int s = 0; s < R; ++s) {
        out[s] = base + lanes_below(m[s]);
        base += static_cast<uint32_t>(__popcll(m[s]));
    }
}

// wave_reserve for a whole workgroup: one atomic add per workgroup
// (per-frame counters hit by every wavefront are a contention point).  All
// threads of the workgroup call it; `lds` holds blockDim.x / 64 + 1 words.
// Places are in (wavefront, row, lane) order.
template <int R>
__device__ __forceinline__ void block_reserve(uint32_t* ctr, const bool (&flag)[R], uint32_t (&out)[R],
                                              uint32_t* lds) {
    const int NW = static_cast<int>(blockDim.x) / kWave;
    const int w = static_cast<int>(threadIdx.x) / kWave;
    uint64_t m[R];
    uint32_t total = 0;
#pragma unroll
    for (int s = 0; s < R; ++s) {
        m[s] = __ballot(flag[s]);
        total += static_cast<uint32_t>(__popcll(m[s]));
    }
    if (lane_id() == 0) lds[w] = total;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int q = 0; q < NW; ++q) {
            const uint32_t t = lds[q];
            lds[q] = run;
            run += t;
        }
        lds[NW] = run ? atomicAdd(ctr, run) : 0u;
    }
    __syncthreads();
    uint32_t base = lds[NW] + lds[w];
#pragma unroll
    for (int s = 0; s < R; ++s) {
        out[s] = base + lanes_below(m[s]);
        base += static_cast<uint32_t>(__popcll(m[s]));
    }
    __syncthreads();
}

// Stream compaction inside one tile of NT * R elements held row by row
// (element k * NT + threadIdx.x of the tile is the thread's element of row
// k, so every row is read and written with coalesced accesses).  bits has
// bit k set when the thread's element of row k is kept; pos[k] receives the
// number of kept elements before it in tile order.  Returns the tile's kept
// count.  lds holds R * NT / 64 + 1 words.
template <int NT, int R>
__device__ __forceinline__ uint32_t tile_compact(uint32_t bits, uint32_t (&pos)[R], uint32_t* lds) {
    constexpr int W = NT / kWave;
    static_assert(R * W <= kWave, "one wavefront scans the row totals");
    const int w = static_cast<int>(threadIdx.x) / kWave;
    uint64_t bal[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        bal[k] = __ballot((bits >> k) & 1u);
        if (lane_id() == 0) lds[k * W + w] = static_cast<uint32_t>(__popcll(bal[k]));
    }
    __syncthreads();
    if (threadIdx.x < kWave) {
        const uint32_t v = threadIdx.x < R * W ? lds[threadIdx.x] : 0u;
        const uint32_t inc = wave_inclusive_sum(v);
        if (threadIdx.x < R * W) lds[threadIdx.x] = inc - v;
        if (threadIdx.x == kWave - 1) lds[R * W] = inc;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < R; ++k) pos[k] = lds[k * W + w] + lanes_below(bal[k]);
    const uint32_t total = lds[R * W];
    __syncthreads();
    return total;
}

}  // namespace lpd
