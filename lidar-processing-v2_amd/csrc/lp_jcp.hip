// lp_jcp.hip -- Jump-Convolution-Process on the range image, exact.
//
// Reference: lidar_processing_lib/src/segmenter.cpp:481-638 (JCP) and
// :640-669 (populateLabels).  The reference dilates the obstacle mask with a
// 5x5 rectangle (out-of-image pixels never contribute), queues every ground
// pixel next to an obstacle in raster order, and then visits the queue
// sequentially.  Each visit reads the CURRENT label of its 24 neighbours and
// writes 24 weight / label slots that are member arrays reset only once per
// frame: a slot whose neighbour falls outside the image is skipped without
// being rewritten, so it keeps the value of the last earlier queue pixel for
// which that slot was inside the image ("stale" slots).
//
// Parallel form used here (exactly equivalent):
//   * every dependency of a queue pixel p is on raster-earlier queue pixels;
//   * all of them lie in earlier image rows, except the two same-row
//     neighbours (0,-1) and (0,-2) -- a stale slot's writer and the writer's
//     neighbour are always in earlier rows or static;
//   * so rows are solved in order, and inside a row the state of pixel j is
//     T_j(state(j-2), state(j-1)) over 3 states {ground, obstacle, pending}:
//     an associative scan of 9-entry transfer maps solves the row at once.
// The weights w_i = u_i / sum do not depend on labels and are precomputed
// for all queue pixels in parallel (k_jcp_slots); the per-row solve only
// re-runs the ordered float sums that the reference performs.
#include "lp_kernels.h"

namespace lp {

using namespace lpd;

__constant__ static const int8_t kOffR[24] = {-2, -2, -2, -2, -2, -1, -1, -1, -1, -1, 0, 0,
                                              0,  0,  1,  1,  1,  1,  1,  2,  2,  2,  2, 2};
__constant__ static const int8_t kOffC[24] = {-2, -1, 0, 1, 2, -2, -1, 0, 1, 2, -2, -1,
                                              1,  2,  -2, -1, 0, 1, 2, -2, -1, 0, 1, 2};

// Slot label sources: >= 0 a queue index (dynamic), else a static label.
constexpr int32_t kSrcUnknown = -1;
constexpr int32_t kSrcGround = -2;
constexpr int32_t kSrcObstacle = -3;

// Internal solve states.
constexpr uint32_t kSG = 0, kSO = 1, kSU = 2;

// Queue membership: a ground pixel with an obstacle pixel in its clipped 5x5
// window (cv::dilate of the red channel, then the raster scan of
// segmenter.cpp:491-514).
// Queue flags: a ground pixel with an obstacle pixel in its 5x5 window
// clipped to the image (the dilated red channel of segmenter.cpp:484-513).
// Each workgroup takes 256 columns of one row and stages rows r-2..r+2 of
// columns c0-2..c0+257 in LDS (out-of-image pixels read as empty).
__global__ __launch_bounds__(256) void k_jcp_mark(DevBuf d, Layout L) {
    constexpr int kSpan = 256 + 4;
    __shared__ uint8_t win[5][kSpan];
    const int f = blockIdx.y;
    const FrameDesc& fd = d.frames[f];
    const int H = fd.sc.H, W = fd.sc.W;
    const int tpr = (W + 255) / 256;
    const int r = static_cast<int>(blockIdx.x) / tpr, c0 = (static_cast<int>(blockIdx.x) % tpr) * 256;
    if (r >= H) return;
    const uint8_t* st = d.state + fd.pix_off;
    for (int q = threadIdx.x; q < 5 * kSpan; q += blockDim.x) {
        const int rr = r + q / kSpan - 2, cc = c0 - 2 + q % kSpan;
        win[q / kSpan][q % kSpan] = (rr >= 0 && rr < H && cc >= 0 && cc < W) ? st[rr * W + cc] : kEmpty;
    }
    __syncthreads();
    const int c = c0 + static_cast<int>(threadIdx.x);
    if (c >= W) return;
    uint32_t q = 0;
    if (win[2][threadIdx.x + 2] == kGround) {
#pragma unroll
        for (int dr = 0; dr < 5; ++dr)
#pragma unroll
            for (int dc = 0; dc < 5; ++dc) q |= win[dr][threadIdx.x + dc] == kObstacle ? 1u : 0u;
    }
    d.qflag[(size_t)f * L.P + r * W + c] = q;
}

// qmap holds the exclusive scan of qflag on entry.
__global__ void k_jcp_fill(DevBuf d, Layout L) {
    const int f = blockIdx.y;
    const FrameDesc& fd = d.frames[f];
    const int P = fd.sc.H * fd.sc.W;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const size_t o = (size_t)f * L.P + p;
    const int pos = d.qmap[o];
    if (d.qflag[o]) {
        d.qlist[(size_t)f * L.P + pos] = p;
        d.qmap[o] = pos;
    } else {
        d.qmap[o] = -1;
    }
}

__device__ __forceinline__ int lower_bound_i32(const int32_t* a, int n, int v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

struct SlotVal {
    float u;
    int32_t src;
    bool near;
};

// Slot value computed by evaluator pixel `pe` (core point `core`) for its
// neighbour pixel np (segmenter.cpp:558-606).
__device__ __forceinline__ SlotVal eval_slot(const DevBuf& d, size_t po, size_t qo, const SegConst& s, int pe,
                                             float4 core, int np) {
    SlotVal v{0.0f, kSrcUnknown, false};
    if (d.mapping[po + np] == -1) return v;
    const float4 nb = d.pxyz[po + np];
    const float dx = core.x - nb.x;
    const float dy = core.y - nb.y;
    const float dz = core.z - nb.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    if (d2 > s.kthr2) return v;
    v.u = glibc_expf(s.neg_amp * sqrtf(d2));
    v.near = true;
    const int qm = d.qmap[qo + np];
    if (qm >= 0) {
        v.src = (np < pe) ? qm : kSrcUnknown;  // later queue pixels are still pending
    } else {
        const uint8_t st = d.state[po + np];
        v.src = st == kGround ? kSrcGround : (st == kObstacle ? kSrcObstacle : kSrcUnknown);
    }
    return v;
}

// Per queue pixel: the 24 slot weights w_i = u_i / sum and label sources,
// including stale slots resolved through their writer.
__global__ void k_jcp_slots(DevBuf d, Layout L) {
    int tile, f;
    xcd_tile_frame(tile, f);
    const FrameDesc& fd = d.frames[f];
    const SegConst& s = fd.sc;
    const int nq = static_cast<int>(d.fs[f * kFsCount + kFsQueue]);
    const int j = tile * blockDim.x + threadIdx.x;
    if (j >= nq) return;
    const int H = s.H, W = s.W;
    const size_t po = fd.pix_off;
    const size_t qo = (size_t)f * L.P;
    const int32_t* ql = d.qlist + qo;
    const int p = ql[j];
    const int r = p / W, c = p % W;
    const float4 core = d.pxyz[po + p];
    float u[24];
    int32_t src[24];
    float sum = 0.0f;
#pragma unroll
    for (int i = 0; i < 24; ++i) {
        const int dr = kOffR[i], dc = kOffC[i];
        const int nr = r + dr, nc = c + dc;
        if (nr >= 0 && nr < H && nc >= 0 && nc < W) {
            const SlotVal v = eval_slot(d, po, qo, s, p, core, nr * W + nc);
            u[i] = v.u;
            src[i] = v.src;
            if (v.near) sum += v.u;
        } else {
            // Stale slot: find the last earlier queue pixel for which this
            // slot was inside the image.
            const int rlo = max(0, -dr), rhi = min(H, H - dr);
            const int clo = max(0, -dc), chi = min(W, W - dc);
            int khi = j - 1;
            if (rhi < H) khi = min(khi, lower_bound_i32(ql, nq, rhi * W) - 1);
            const int klo = rlo > 0 ? lower_bound_i32(ql, nq, rlo * W) : 0;
            int k = khi;
            for (; k >= klo; --k) {
                const int qc = ql[k] % W;
                if (qc >= clo && qc < chi) break;
            }
            if (k >= klo) {
                const int q = ql[k];
                const int qr = q / W, qc = q % W;
                const SlotVal v = eval_slot(d, po, qo, s, q, d.pxyz[po + q], (qr + dr) * W + (qc + dc));
                u[i] = v.u;
                src[i] = v.src;
            } else {
                u[i] = 0.0f;  // slots start at zero / UNKNOWN (segmenter.cpp:523-525)
                src[i] = kSrcUnknown;
            }
        }
    }
    const bool ok = fabsf(sum) > 1.1920928955078125e-07f;  // FLT_EPSILON
    float* wo = d.slot_w + ((size_t)f * L.P + j) * 24;
    int32_t* so = d.slot_src + ((size_t)f * L.P + j) * 24;
#pragma unroll
    for (int i = 0; i < 24; ++i) {
        wo[i] = ok ? u[i] / sum : 0.0f;
        so[i] = src[i];
    }
    d.qsum_ok[(size_t)f * L.P + j] = ok ? 1 : 0;
}

// Transfer map over the nine pair states ps = a*3 + b: entry ps (0..8) is
// byte ps of (lo, hi, x), i.e. entries 0-3, 4-7 and 8 (the upper bytes of x
// are don't-care values in 0..8).
struct TMap {
    uint32_t lo, hi, x;
};
constexpr uint32_t kIdLo = 0x03020100u, kIdHi = 0x07060504u, kIdX = 0x08u;

// Entries f[sel_k] for the four selector bytes sel_k (each 0..8) of sel:
// the byte permute takes entries 0..7 from (hi:lo), and the bytes whose
// selector is 8 (bit 3 set) are replaced by entry 8 (f8, in every byte).
// The byte mask is a permute too: selector 12 gives 0x00 and 13 gives 0xff.
__device__ __forceinline__ uint32_t tmap_pick(const TMap& f, uint32_t f8, uint32_t sel) {
    const uint32_t v = __builtin_amdgcn_perm(f.hi, f.lo, sel);
    const uint32_t m = __builtin_amdgcn_perm(0u, 0u, ((sel >> 3) & 0x01010101u) | 0x0c0c0c0cu);
    return (f8 & m) | (v & ~m);
}

// f after g.
__device__ __forceinline__ TMap tmap_compose(const TMap& f, const TMap& g) {
    const uint32_t f8 = __builtin_amdgcn_perm(0u, f.x, 0u);
    return TMap{tmap_pick(f, f8, g.lo), tmap_pick(f, f8, g.hi), tmap_pick(f, f8, g.x)};
}

__device__ __forceinline__ uint32_t tmap_get(const TMap& m, uint32_t ps) {
    const uint32_t w = ps < 4 ? m.lo : (ps < 8 ? m.hi : m.x);
    return (w >> (8 * (ps & 3))) & 0xffu;
}

// The map of the lane `kCtrl` names (DPP row shift or row broadcast); lanes
// outside kRowMask or without a source lane read the identity.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ TMap tmap_dpp(const TMap& m) {
    return TMap{static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(kIdLo), static_cast<int>(m.lo),
                                                                  kCtrl, kRowMask, 0xf, false)),
                static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(kIdHi), static_cast<int>(m.hi),
                                                                  kCtrl, kRowMask, 0xf, false)),
                static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(kIdX), static_cast<int>(m.x),
                                                                  kCtrl, kRowMask, 0xf, false))};
}

// Lane `src` (wavefront-uniform) of a map.
__device__ __forceinline__ TMap tmap_readlane(const TMap& m, int src) {
    return TMap{static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(m.lo), src)),
                static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(m.hi), src)),
                static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(m.x), src))};
}

// Inclusive scan of the wavefront's maps, m := m_lane o ... o m_0: row
// shifts by 1, 2, 4 and 8 inside each 16-lane row, then lane 15 of rows 0 / 2
// into rows 1 / 3 and lane 31 into rows 2 and 3.
__device__ __forceinline__ TMap tmap_wave_scan(TMap m) {
    m = tmap_compose(m, tmap_dpp<0x111, 0xf>(m));
    m = tmap_compose(m, tmap_dpp<0x112, 0xf>(m));
    m = tmap_compose(m, tmap_dpp<0x114, 0xf>(m));
    m = tmap_compose(m, tmap_dpp<0x118, 0xf>(m));
    m = tmap_compose(m, tmap_dpp<0x142, 0xa>(m));
    m = tmap_compose(m, tmap_dpp<0x143, 0xc>(m));
    return m;
}
constexpr int kSolveThreads = 512;
constexpr int kSolveWaves = kSolveThreads / kWave;
constexpr int kFinLds = 65536;
// LDS label table of the solve: entries 0, 1, 2 are the static labels of
// sources kSrcObstacle, kSrcGround, kSrcUnknown; queue pixel j is entry j + 3.
constexpr int kFinBase = 3;
static_assert(kSrcObstacle + kFinBase == 0 && kSrcGround + kFinBase == 1 && kSrcUnknown + kFinBase == 2,
              "static sources map to the first three label entries");

// wg += w for a GROUND slot, wo += w for an OBSTACLE slot (segmenter.cpp:
// 613-626).  Written as two selects: every weight is +0 or positive and the
// sums start at +0, so adding +0.0f to the other sum leaves it bit-identical,
// and the select form keeps both sums in registers.
__device__ __forceinline__ void add_slot(uint32_t l, float w, float& wg, float& wo) {
    wg += (l == kSG) ? w : 0.0f;
    wo += (l == kSO) ? w : 0.0f;
}

// Slot weights, label sources and the sum test of one queue pixel, as
// written by k_jcp_slots.  They do not depend on solved labels, so the solve
// loads the next chunk's while it solves the current one.
struct SlotData {
    float4 w[6];
    int4 s[6];
    uint32_t ok;
};

__device__ __forceinline__ void load_slots(const DevBuf& d, size_t qo, int j, SlotData& sd) {
    const float4* w4 = reinterpret_cast<const float4*>(d.slot_w + (qo + j) * 24);
    const int4* s4 = reinterpret_cast<const int4*>(d.slot_src + (qo + j) * 24);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        sd.w[q] = w4[q];
        sd.s[q] = s4[q];
    }
    sd.ok = d.qsum_ok[qo + j];
}

// Transfer map of queue pixel j (pair state (state(j-2), state(j-1)) ->
// (state(j-1), state(j))).  Labels of earlier rows come from `fin`.
__device__ __forceinline__ TMap pixel_map(const SlotData& sd, int j, int rs, const uint8_t* fin) {
    const bool ok = sd.ok != 0;
    float wv[24];
    int32_t src[24];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const float4 a = sd.w[q];
        const int4 b = sd.s[q];
        wv[4 * q] = a.x;
        wv[4 * q + 1] = a.y;
        wv[4 * q + 2] = a.z;
        wv[4 * q + 3] = a.w;
        src[4 * q] = b.x;
        src[4 * q + 1] = b.y;
        src[4 * q + 2] = b.z;
        src[4 * q + 3] = b.w;
    }
    // Static and earlier-row labels.  A source in this row can only be the
    // queue pixel before (j-1) or two before (j-2), through slots (0,-2) and
    // (0,-1) (indices 10, 11): stale slots always read earlier rows.
    // `fin` holds the three static labels below the queue states (entry
    // sv + kFinBase), so every slot is one unconditional LDS read; sources in
    // this row (sv >= rs) read as pending here and come from the pair state.
    uint32_t lab[24];
#pragma unroll
    for (int i = 0; i < 24; ++i) {
        const int32_t sv = src[i];
        const uint32_t fv = fin[sv + kFinBase];
        lab[i] = sv < rs ? fv : kSU;
    }
    // Only sources inside this row come from the scan's pair state; a stale
    // column slot at column 0 or 1 can name the last queue pixels of the
    // previous row, whose labels are already final in `fin`.
    const int sel10 = src[10] < rs ? 0 : (src[10] == j - 2 ? 1 : (src[10] == j - 1 ? 2 : 0));  // 1: j-2, 2: j-1
    const bool sel11 = src[11] >= rs && src[11] == j - 1;
    if (!ok) {
        // Every pair state (a, b) goes to (b, UNKNOWN).
        return TMap{(0u * 3 + kSU) | ((1u * 3 + kSU) << 8) | ((2u * 3 + kSU) << 16) | ((0u * 3 + kSU) << 24),
                    (1u * 3 + kSU) | ((2u * 3 + kSU) << 8) | ((0u * 3 + kSU) << 16) | ((1u * 3 + kSU) << 24),
                    2u * 3 + kSU};
    }
    // Ordered sums of segmenter.cpp:613-626: slots 0..9 are fixed, 10 and 11
    // may depend on the row's previous states, 12..23 are fixed again.  The
    // ground sum only sees which of slots 10, 11 are GROUND and the obstacle
    // sum which are OBSTACLE, so four chains of each (q bit 0: slot 10
    // counts, bit 1: slot 11 counts), each in the reference's order, cover
    // all nine pair states; the chains are independent and interleave.
    // Per-slot addends (the weight or +0) of the two sums, formed once.
    float ag[24], ao[24];
#pragma unroll
    for (int i = 0; i < 24; ++i) {
        ag[i] = lab[i] == kSG ? wv[i] : 0.0f;
        ao[i] = lab[i] == kSO ? wv[i] : 0.0f;
    }
    float wg0 = 0.0f, wo0 = 0.0f;
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        wg0 += ag[i];
        wo0 += ao[i];
    }
    float tg[4], to[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float g = wg0, o = wo0;
        g += (q & 1) ? wv[10] : 0.0f;
        o += (q & 1) ? wv[10] : 0.0f;
        g += (q & 2) ? wv[11] : 0.0f;
        o += (q & 2) ? wv[11] : 0.0f;
#pragma unroll
        for (int i = 12; i < 24; ++i) {
            g += ag[i];
            o += ao[i];
        }
        tg[q] = g;
        to[q] = o;
    }
    uint32_t e[3] = {0u, 0u, 0u};
#pragma unroll
    for (int ps = 0; ps < 9; ++ps) {
        const uint32_t a = ps / 3, b = ps % 3;
        const uint32_t l10 = sel10 == 1 ? a : (sel10 == 2 ? b : lab[10]);
        const uint32_t l11 = sel11 ? b : lab[11];
        const float g01 = l10 == kSG ? tg[1] : tg[0], g23 = l10 == kSG ? tg[3] : tg[2];
        const float o01 = l10 == kSO ? to[1] : to[0], o23 = l10 == kSO ? to[3] : to[2];
        const float g = l11 == kSG ? g23 : g01;
        const float o = l11 == kSO ? o23 : o01;
        const uint32_t t = (o > g) ? kSO : kSG;
        e[ps / 4] |= (b * 3 + t) << (8 * (ps % 4));
    }
    return TMap{e[0], e[1], e[2]};
}

// Row-ordered exact solve, one workgroup per frame: per row, transfer maps
// of its queue pixels, a wavefront DPP scan of the maps, one LDS step across
// the eight wavefronts (two per SIMD, so a row of up to 512 queue pixels is
// one step and the SIMDs interleave two wavefronts' map evaluation).
__global__ __launch_bounds__(kSolveThreads) void k_jcp_solve(DevBuf d, Layout L) {
    __shared__ uint8_t fin[kFinBase + kFinLds];
    __shared__ int row_start[1025];
    __shared__ TMap wtot[kSolveWaves];
    const int f = blockIdx.x;
    const FrameDesc& fd = d.frames[f];
    const SegConst& s = fd.sc;
    const int nq = static_cast<int>(d.fs[f * kFsCount + kFsQueue]);
    const int tid = threadIdx.x;
    const int w = tid / kWave;
    const int lane = lane_id();
    if (nq > kFinLds || s.H > 1024) {
        if (tid == 0) atomicOr(&d.fs[f * kFsCount + kFsErr], kErrQueueCapacity);
        return;
    }
    if (nq == 0) return;
    const size_t qo = (size_t)f * L.P;
    const int32_t* ql = d.qlist + qo;
    for (int r = tid; r <= s.H; r += kSolveThreads) row_start[r] = lower_bound_i32(ql, nq, r * s.W);
    if (tid < kFinBase) fin[tid] = static_cast<uint8_t>(tid == 0 ? kSO : (tid == 1 ? kSG : kSU));
    // Chunk bases run rs, rs + T, ... below re, then the next non-empty row
    // starts at re: the first chunk starts at 0 and each next base is known.
    // Slot loads are unconditional (index clamped to the last queue pixel):
    // with the same number of loads in flight on every path the compiler
    // waits for the current chunk's data only, not for the prefetch.
    SlotData cur, nxt;
    load_slots(d, qo, min(tid, nq - 1), cur);
    __syncthreads();
    for (int r = 0; r < s.H; ++r) {
        const int rs = row_start[r], re = row_start[r + 1];
        if (rs == re) continue;
        uint32_t carry = kSU * 3 + kSU;
        for (int base = rs; base < re; base += kSolveThreads) {
            const int j = base + tid;
            const int jn = ((base + kSolveThreads < re) ? base + kSolveThreads : re) + tid;
            load_slots(d, qo, min(jn, nq - 1), nxt);
            TMap m = j < re ? pixel_map(cur, j, rs, fin) : TMap{kIdLo, kIdHi, kIdX};
            m = tmap_wave_scan(m);  // m := m_j o ... o m_(first of this wavefront)
            if (lane == kWave - 1) wtot[w] = m;
            __syncthreads();
            // Prefixes of the wavefront totals: lane q scans wtot[q] o ... o
            // wtot[0] with DPP row shifts (the totals sit in lanes 0..7 of
            // row 0); pre is the prefix of the earlier wavefronts, chain
            // that of all of them.
            TMap t = lane < kSolveWaves ? wtot[lane] : TMap{kIdLo, kIdHi, kIdX};
            t = tmap_compose(t, tmap_dpp<0x111, 0xf>(t));
            t = tmap_compose(t, tmap_dpp<0x112, 0xf>(t));
            t = tmap_compose(t, tmap_dpp<0x114, 0xf>(t));
            static_assert(kSolveWaves <= 8, "three row-shift steps scan at most 8 wavefront totals");
            const TMap chain = tmap_readlane(t, kSolveWaves - 1);
            if (w > 0) m = tmap_compose(m, tmap_readlane(t, w - 1));
            const uint32_t ps = tmap_get(m, carry);
            if (j < re) fin[kFinBase + j] = static_cast<uint8_t>(ps % 3);
            carry = tmap_get(chain, carry);
            cur = nxt;  // the prefetch has had the whole chunk to land
            __syncthreads();
        }
    }
    for (int j = tid; j < nq; j += kSolveThreads) {
        const uint32_t v = fin[kFinBase + j];
        d.fin[qo + j] = v == kSG ? kGround : (v == kSO ? kObstacle : kPending);
    }
}

// Labels per input point (populateLabels, segmenter.cpp:640-669): only the
// winner of a pixel gets its pixel's final label; PENDING stays UNKNOWN.
// Also flags obstacle points for the input-order split of processor.cpp:562-579.
__global__ void k_seg_labels(DevBuf d, Layout L) {
    const int f = blockIdx.y;
    const FrameDesc& fd = d.frames[f];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fd.n) return;
    const int gi = fd.pt_off + i;
    uint32_t lab = LP_LABEL_UNKNOWN;
    const int pix = d.pt_pix[gi];
    if (pix >= 0) {
        const unsigned long long key = d.pix_key[fd.pix_off + pix];
        if (static_cast<int>(key & ((1ull << L.idx_bits) - 1)) == i) {
            const int qm = d.qmap[(size_t)f * L.P + pix];
            const uint8_t st = qm >= 0 ? d.fin[(size_t)f * L.P + qm] : d.state[fd.pix_off + pix];
            lab = st == kGround ? LP_LABEL_GROUND : (st == kObstacle ? LP_LABEL_OBSTACLE : LP_LABEL_UNKNOWN);
        }
    }
    d.labels[gi] = lab;
    d.obs_flag[(size_t)f * L.N + i] = lab == LP_LABEL_OBSTACLE ? 1u : 0u;
    if (d.gnd_flag) d.gnd_flag[(size_t)f * L.N + i] = lab == LP_LABEL_GROUND ? 1u : 0u;
}

// Segmenter::image after segment() (segmenter.hpp:163-166), as BGR bytes.
// The winners paint ground green and obstacles red (segmenter.cpp:288-318);
// the red channel is dilated with a 5x5 rectangle whose out-of-image part
// never contributes (:484-489), which turns empty pixels near an obstacle
// red and ground pixels near one yellow; yellow pixels with a point become
// the blue JCP queue (:491-513) and take the JCP colour where the weights
// resolve (:626-635).  A yellow pixel without a point (white, :510) cannot
// occur: only a winner paints green.
__global__ void k_label_image(DevBuf d, Layout L, int f, uint8_t* img) {
    const FrameDesc& fd = d.frames[f];
    const int H = fd.sc.H, W = fd.sc.W;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= H * W) return;
    const uint8_t* st = d.state + fd.pix_off;
    const int qm = d.qmap[(size_t)f * L.P + p];
    uint8_t b = 0, g = 0, r = 0;
    const uint8_t s = qm >= 0 ? d.fin[(size_t)f * L.P + qm] : st[p];
    if (s == kGround) {
        g = 255;
    } else if (s == kObstacle) {
        r = 255;
    } else if (s == kPending) {
        b = 255;
    } else {
        const int row = p / W, col = p % W;
        for (int dr = -2; dr <= 2 && !r; ++dr) {
            const int rr = row + dr;
            if (rr < 0 || rr >= H) continue;
            for (int dc = -2; dc <= 2; ++dc) {
                const int cc = col + dc;
                if (cc >= 0 && cc < W && st[rr * W + cc] == kObstacle) {
                    r = 255;
                    break;
                }
            }
        }
    }
    img[3 * p + 0] = b;
    img[3 * p + 1] = g;
    img[3 * p + 2] = r;
}

void launch_jcp(lp_context* c, int B) {
    DevBuf& d = c->d;
    const Layout& L = c->L;
    hipStream_t st = c->stream;
    const int pb = (L.P + 255) / 256;
    stage_mark(c, "jcp_queue");
    hipLaunchKernelGGL(k_jcp_mark, dim3(L.H * ((L.W + 255) / 256), B), dim3(256), 0, st, d, L);
    launch_scan(c, B, d.qflag, reinterpret_cast<uint32_t*>(d.qmap), L.P, kCountPixels, 0, kFsQueue);
    hipLaunchKernelGGL(k_jcp_fill, dim3(pb, B), dim3(256), 0, st, d, L);
    stage_mark(c, "jcp_slots");
    hipLaunchKernelGGL(k_jcp_slots, dim3(pb, B), dim3(256), 0, st, d, L);
    stage_mark(c, "jcp_solve");
    hipLaunchKernelGGL(k_jcp_solve, dim3(B), dim3(kSolveThreads), 0, st, d, L);
    stage_mark(c, "labels");
    hipLaunchKernelGGL(k_seg_labels, dim3((L.N + 255) / 256, B), dim3(256), 0, st, d, L);
}

void launch_label_image(lp_context* c, int f, int pixels, uint8_t* img) {
    hipLaunchKernelGGL(k_label_image, dim3((pixels + 255) / 256), dim3(256), 0, c->stream, c->d, c->L, f, img);
}

}  // namespace lp
