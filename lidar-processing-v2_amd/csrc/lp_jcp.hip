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
#include "lp_queue.h"

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
// Each workgroup takes T columns of one row (T = 64, 128 or 256, the
// tile width the launch chooses) and stages rows r-2..r+2 of columns
// c0-2..c0+T+1 in LDS (out-of-image pixels read as empty).
template <int T>
__global__ __launch_bounds__(T) void k_jcp_mark(DevBuf d, Layout L) {
    constexpr int kSpan = T + 4;
    __shared__ uint8_t win[5][kSpan];
    const int f = blockIdx.y;
    const FrameDesc& fd = d.frames[f];
    const int H = fd.sc.H, W = fd.sc.W;
    const int tpr = (W + T - 1) / T;
    const int r = static_cast<int>(blockIdx.x) / tpr, c0 = (static_cast<int>(blockIdx.x) % tpr) * T;
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
    d.qflag[(size_t)f * L.P + r * W + c] = static_cast<uint8_t>(q);
}

// The queue in raster order, from the flags in two passes over tiles of 2048
// pixels (eight rows of 256, coalesced): the queue count of every tile, then
// each tile's compaction after the counts of the tiles before it.  qlist[j]
// is the pixel of queue entry j, qmap[p] the queue index of pixel p (-1 when
// not queued), qrow[r] the first queue index at or after row r's first pixel
// (kept for the row solve); the last tile writes the queue length.
constexpr int kQueueThreads = 256;
constexpr int kQueueRows = 8;
constexpr int kQueueTile = kQueueThreads * kQueueRows;

__device__ __forceinline__ uint32_t queue_bits(const uint8_t* qf, int base, int P) {
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < kQueueRows; ++k) {
        const int p = base + k * kQueueThreads + static_cast<int>(threadIdx.x);
        m |= (p < P && qf[p] != 0) ? (1u << k) : 0u;
    }
    return m;
}

__global__ __launch_bounds__(kQueueThreads) void k_jcp_count(DevBuf d, Layout L) {
    __shared__ uint32_t lds[kQueueThreads / kWave + 1];
    const int f = blockIdx.y;
    const FrameDesc& fd = d.frames[f];
    const int P = fd.sc.H * fd.sc.W;
    const int base = blockIdx.x * kQueueTile;
    if (base >= P) return;
    const uint32_t m = queue_bits(d.qflag + (size_t)f * L.P, base, P);
    uint32_t total;
    block_exclusive_sum<kQueueThreads>(static_cast<uint32_t>(__popc(m)), lds, &total);
    if (threadIdx.x == 0) d.blk[(size_t)f * L.blk_stride + blockIdx.x] = total;
}

__global__ __launch_bounds__(kQueueThreads) void k_jcp_fill(DevBuf d, Layout L) {
    __shared__ uint32_t lds[kQueueRows * kQueueThreads / kWave + 1];
    const int f = blockIdx.y;
    const FrameDesc& fd = d.frames[f];
    const int W = fd.sc.W;
    const int P = fd.sc.H * W;
    const int tile = blockIdx.x;
    const int base = tile * kQueueTile;
    if (base >= P) {
        if (tile == 0 && threadIdx.x == 0) d.fs[f * kFsCount + kFsQueue] = 0;
        return;
    }
    const uint32_t* blk = d.blk + (size_t)f * L.blk_stride;
    uint32_t pre = 0;
    for (int i = threadIdx.x; i < tile; i += kQueueThreads) pre += blk[i];
    uint32_t pre_total;
    block_exclusive_sum<kQueueThreads>(pre, lds, &pre_total);
    const size_t po = (size_t)f * L.P;
    const uint32_t m = queue_bits(d.qflag + po, base, P);
    uint32_t pos[kQueueRows];
    const uint32_t total = tile_compact<kQueueThreads, kQueueRows>(m, pos, lds);
#pragma unroll
    for (int k = 0; k < kQueueRows; ++k) {
        const int p = base + k * kQueueThreads + static_cast<int>(threadIdx.x);
        if (p >= P) continue;
        const int j = static_cast<int>(pre_total + pos[k]);
        if (p % W == 0) d.qrow[(size_t)f * (L.H + 1) + p / W] = j;
        if ((m >> k) & 1u) {
            d.qlist[po + j] = p;
            d.qmap[po + p] = j;
        } else {
            d.qmap[po + p] = -1;
        }
    }
    if (threadIdx.x == 0 && base + kQueueTile >= P) d.fs[f * kFsCount + kFsQueue] = pre_total + total;
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
    const float4 nb = d.pxyz[po + np];  // w holds the pixel's winner, -1 when empty
    if (__float_as_int(nb.w) == -1) return v;
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
    const int32_t* qrow = d.qrow + (size_t)f * (L.H + 1);
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
            // First queue index at or after a row's first pixel: qrow.
            if (rhi < H) khi = min(khi, qrow[rhi] - 1);
            const int klo = rlo > 0 ? qrow[rlo] : 0;
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
    float w[24];
#pragma unroll
    for (int i = 0; i < 24; ++i) w[i] = ok ? u[i] / sum : 0.0f;
    // Head slots 0..11 (rows -2, -1 and the two left neighbours): sources
    // and weights as they are.  Tail slots 12..23 (right neighbours and rows
    // +1, +2) name a neighbour raster-after their writer (the pixel itself
    // or, stale, an earlier queue pixel), so their source is static or a
    // pending queue pixel and their addends are known here: the ground
    // addend (w or +0) and the obstacle addend of each tail slot.
    float* wo = d.slot_w + ((size_t)f * L.P + j) * 24;
    int32_t* so = d.slot_src + ((size_t)f * L.P + j) * 24;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        wo[k] = src[12 + k] == kSrcGround ? w[12 + k] : 0.0f;
        wo[12 + k] = src[12 + k] == kSrcObstacle ? w[12 + k] : 0.0f;
        so[k] = src[k];
        so[12 + k] = __float_as_int(w[k]);
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

// Slot data of one queue pixel, as written by k_jcp_slots: w holds the
// tail slots' ground addends (entries 0..11) and obstacle addends (12..23),
// s the head slots' sources (0..11) and weight bit patterns (12..23), ok the
// sum test.  None of it depends on solved labels, so the solve loads the next
// chunk's while it solves the current one.
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

// Byte vectors over the nine pair states ps = a*3 + b, laid out like a TMap
// (bytes 0-3 of lo, 4-7 of hi, byte 0 of x): 3a, 3b and b.
constexpr uint32_t kA3Lo = 0x03000000u, kA3Hi = 0x06060303u, kA3X = 0x06u;
constexpr uint32_t kB3Lo = 0x00060300u, kB3Hi = 0x03000603u, kB3X = 0x06u;
constexpr uint32_t kB1Lo = 0x00020100u, kB1Hi = 0x01000201u, kB1X = 0x02u;

__device__ __forceinline__ uint32_t obstacle_wins(float o, float g) { return o > g ? kSO : kSG; }

// Transfer map of queue pixel j (pair state (state(j-2), state(j-1)) ->
// (state(j-1), state(j))).  Labels of earlier rows come from `fin`.
__device__ __forceinline__ TMap pixel_map(const SlotData& sd, int j, int rs, const uint8_t* fin) {
    float tw[24], hw[12];
    int32_t hs[12];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        tw[4 * q] = sd.w[q].x;
        tw[4 * q + 1] = sd.w[q].y;
        tw[4 * q + 2] = sd.w[q].z;
        tw[4 * q + 3] = sd.w[q].w;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        hs[4 * q] = sd.s[q].x;
        hs[4 * q + 1] = sd.s[q].y;
        hs[4 * q + 2] = sd.s[q].z;
        hs[4 * q + 3] = sd.s[q].w;
        hw[4 * q] = __int_as_float(sd.s[q + 3].x);
        hw[4 * q + 1] = __int_as_float(sd.s[q + 3].y);
        hw[4 * q + 2] = __int_as_float(sd.s[q + 3].z);
        hw[4 * q + 3] = __int_as_float(sd.s[q + 3].w);
    }
    // Ordered sums of segmenter.cpp:613-626 (wg += w for a GROUND slot, wo +=
    // w for an OBSTACLE slot), written as selects of w or +0: every weight
    // is +0 or positive and the sums start at +0, so adding +0 leaves a sum
    // bit-identical.  Slots 0..9 (rows -2, -1) name
    // static labels or queue pixels of earlier rows, all final in `fin`
    // (which holds the three static labels below the queue states, entry
    // sv + kFinBase).
    float wg0 = 0.0f, wo0 = 0.0f;
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        const uint32_t l = fin[hs[i] + kFinBase];
        wg0 += l == kSG ? hw[i] : 0.0f;
        wo0 += l == kSO ? hw[i] : 0.0f;
    }
    // Slots (0,-2), (0,-1) (10, 11): a source in this row is the queue pixel
    // before (j-1) or two before (j-2) and comes from the pair state; other
    // sources are static or earlier rows (a stale slot at column 0 or 1 can
    // name the last queue pixels of the previous row).
    const int32_t s10 = hs[10], s11 = hs[11];
    const uint32_t lab10 = s10 < rs ? fin[s10 + kFinBase] : kSU;
    const uint32_t lab11 = s11 < rs ? fin[s11 + kFinBase] : kSU;
    const int sel10 = s10 < rs ? 0 : (s10 == j - 2 ? 1 : (s10 == j - 1 ? 2 : 0));  // 1: j-2, 2: j-1
    const bool sel11 = s11 >= rs && s11 == j - 1;
    // Each sum with slot 10 / slot 11 counted or not (q bit 0 / bit 1): the
    // ground sum only sees which of them are GROUND and the obstacle sum which
    // are OBSTACLE, so four values of each cover all nine pair states.
    float g0 = wg0, g1 = wg0 + hw[10], g2 = wg0 + hw[11];
    float g3 = g1 + hw[11];
    float o0 = wo0, o1 = wo0 + hw[10], o2 = wo0 + hw[11];
    float o3 = o1 + hw[11];
    // The tail (slots 12..23) in slot order, its addends precomputed.
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        const float a = tw[k];
        g0 += a;
        g1 += a;
        g2 += a;
        g3 += a;
    }
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        const float a = tw[12 + k];
        o0 += a;
        o1 += a;
        o2 += a;
        o3 += a;
    }
    // Decision table over the label pairs (l10, l11), byte l10 * 3 + l11:
    // ground sum g[(l10 == G) | (l11 == G) << 1] against obstacle sum
    // o[(l10 == O) | (l11 == O) << 1].  A pixel whose weight sum fails the
    // test goes to PENDING for every pair state.
    TMap T{obstacle_wins(o0, g3) | obstacle_wins(o2, g1) << 8 | obstacle_wins(o0, g1) << 16 |
               obstacle_wins(o1, g2) << 24,
           obstacle_wins(o3, g0) | obstacle_wins(o1, g0) << 8 | obstacle_wins(o0, g2) << 16 |
               obstacle_wins(o2, g0) << 24,
           obstacle_wins(o0, g0)};
    if (sd.ok == 0) T = TMap{kSU * 0x01010101u, kSU * 0x01010101u, kSU};
    // Per pair state (a, b): l10 is a, b or static, l11 is b or static; the
    // entry is (b, t(l10, l11)).
    const uint32_t r10 = lab10 * 3u * 0x01010101u, r11 = lab11 * 0x01010101u;
    const uint32_t ilo = (sel10 == 1 ? kA3Lo : (sel10 == 2 ? kB3Lo : r10)) + (sel11 ? kB1Lo : r11);
    const uint32_t ihi = (sel10 == 1 ? kA3Hi : (sel10 == 2 ? kB3Hi : r10)) + (sel11 ? kB1Hi : r11);
    const uint32_t ix = (sel10 == 1 ? kA3X : (sel10 == 2 ? kB3X : r10)) + (sel11 ? kB1X : r11);
    const uint32_t t8 = T.x * 0x01010101u;
    return TMap{tmap_pick(T, t8, ilo) + kB3Lo, tmap_pick(T, t8, ihi) + kB3Hi, tmap_pick(T, t8, ix) + kB3X};
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
    for (int r = tid; r <= s.H; r += kSolveThreads) row_start[r] = r < s.H ? d.qrow[(size_t)f * (L.H + 1) + r] : nq;
    if (tid < kFinBase) fin[tid] = static_cast<uint8_t>(tid == 0 ? kSO : (tid == 1 ? kSG : kSU));
    __syncthreads();
    auto row_at = [&](int r) { return __builtin_amdgcn_readfirstlane(row_start[r]); };
    // Chunks run in queue order: a row's chunks start at rs, rs + T, ... below
    // its end re, and the next chunk after the row's last starts at re, the
    // first queue pixel of the next non-empty row (row_start[H] = nq).
    int r = 0;
    while (row_at(r + 1) == 0) ++r;
    int base = 0;
    uint32_t carry = kSU * 3 + kSU;
    // One chunk with the slot data `cur`, prefetching the next chunk's into
    // `nxt` (its index is clamped to the last queue pixel, so every path has
    // the same loads in flight and the compiler waits for the current chunk's
    // data only).  Returns false after the last chunk.
    auto chunk = [&](const SlotData& cur, SlotData& nxt) -> bool {
        const int rs = row_at(r), re = row_at(r + 1);
        const int j = base + tid;
        const int nb = base + kSolveThreads < re ? base + kSolveThreads : re;
        load_slots(d, qo, min(nb + tid, nq - 1), nxt);
        TMap m = j < re ? pixel_map(cur, j, rs, fin) : TMap{kIdLo, kIdHi, kIdX};
        m = tmap_wave_scan(m);  // m := m_j o ... o m_(first of this wavefront)
        if (lane == kWave - 1) wtot[w] = m;
        __syncthreads();
        // Prefixes of the wavefront totals: lane q scans wtot[q] o ... o
        // wtot[0] with DPP row shifts (the totals sit in lanes 0..7 of row
        // 0); pre is the prefix of the earlier wavefronts, chain that of all
        // of them.
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
        __syncthreads();
        if (nb < re) {
            base = nb;
            return true;
        }
        if (nb >= nq) return false;
        ++r;
        while (row_at(r + 1) == nb) ++r;
        base = nb;
        carry = kSU * 3 + kSU;
        return true;
    };
    SlotData sa, sb;
    load_slots(d, qo, min(tid, nq - 1), sa);
    while (chunk(sa, sb) && chunk(sb, sa)) {
    }
    for (int j = tid; j < nq; j += kSolveThreads) {
        const uint32_t v = fin[kFinBase + j];
        d.fin[qo + j] = v == kSG ? kGround : (v == kSO ? kObstacle : kPending);
    }
}

// Labels per input point (populateLabels, segmenter.cpp:640-669): only the
// winner of a pixel gets its pixel's final label; PENDING stays UNKNOWN.
__device__ __forceinline__ void seg_labels_tile(const DevBuf& d, const Layout& L, int f, int tile) {
    const FrameDesc& fd = d.frames[f];
    const int i = tile * blockDim.x + threadIdx.x;
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
    if (d.gnd_flag) d.gnd_flag[(size_t)f * L.N + i] = lab == LP_LABEL_GROUND ? 1u : 0u;
}

__global__ void k_seg_labels(DevBuf d, Layout L) { seg_labels_tile(d, L, blockIdx.y, blockIdx.x); }

__global__ __launch_bounds__(256) void k_seg_labels_q(DevBuf d, Layout L, int B) {
    wq_run(d.wq + kWqSegLabels * 8, B, blockDim.x, [&](int f) { return d.frames[f].n; },
           [&](int f, int t) { seg_labels_tile(d, L, f, t); });
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
    const int mt = c->mark_tile;
    if (mt == 64)
        hipLaunchKernelGGL(k_jcp_mark<64>, dim3(L.H * ((L.W + 63) / 64), B), dim3(64), 0, st, d, L);
    else if (mt == 128)
        hipLaunchKernelGGL(k_jcp_mark<128>, dim3(L.H * ((L.W + 127) / 128), B), dim3(128), 0, st, d, L);
    else
        hipLaunchKernelGGL(k_jcp_mark<256>, dim3(L.H * ((L.W + 255) / 256), B), dim3(256), 0, st, d, L);
    const int qt = (L.P + kQueueTile - 1) / kQueueTile;  // at most blk_stride tiles
    hipLaunchKernelGGL(k_jcp_count, dim3(qt, B), dim3(kQueueThreads), 0, st, d, L);
    hipLaunchKernelGGL(k_jcp_fill, dim3(qt, B), dim3(kQueueThreads), 0, st, d, L);
    stage_mark(c, "jcp_slots");
    hipLaunchKernelGGL(k_jcp_slots, dim3(pb, B), dim3(256), 0, st, d, L);
    stage_mark(c, "jcp_solve");
    hipLaunchKernelGGL(k_jcp_solve, dim3(B), dim3(kSolveThreads), 0, st, d, L);
    stage_mark(c, "labels");
    const int wg = c->point_wg;
    if (c->persistent)
        hipLaunchKernelGGL(k_seg_labels_q, dim3(kPersistentWorkgroups), dim3(wg), 0, st, d, L, B);
    else
        hipLaunchKernelGGL(k_seg_labels, dim3((L.N + wg - 1) / wg, B), dim3(wg), 0, st, d, L);
}

void launch_label_image(lp_context* c, int f, int pixels, uint8_t* img) {
    hipLaunchKernelGGL(k_label_image, dim3((pixels + 255) / 256), dim3(256), 0, c->stream, c->d, c->L, f, img);
}

}  // namespace lp
