// lp_obb.hip -- rotating-calipers oriented bounding boxes of hull polygons.
//
// Reference: lidar_processing_lib/src/polygonizer.cpp:93-161
// (findAntipodalPairsOfConvexHull, Shamos' walk with areaOfTriangle,
// polygonizer.hpp:228-231) and :163-278 (boundingBoxRotatingCalipers).  For
// every antipodal pair (i, j), for index in {i, j}, for offset in {-1, +1},
// the edge (index, index+offset) of length >= 1e-6 defines a frame in which
// the polygon's extents give a box; the first candidate, in that order, whose
// double area is below the running minimum replaces the box.  The running
// minimum is the box's float `area` member, so near-ties are decided against
// the float-rounded value exactly as the reference does.
//
// One workgroup per polygon.  The first wavefront walks the antipodal pairs
// (the outer step is sequential; each inner run of j advances is tested on
// 64 lanes at once) while the second sums the centroid in the reference's
// order.  The candidates (i, -1) and (i - 1, +1) share one edge with opposite
// directions, which negates every rotated coordinate exactly and leaves the
// area bit-identical, so the workgroup evaluates one area per edge, O(n)
// each; the first wavefront then applies the sequential selection rule over
// the candidates in the reference's order (64 at a time, with a ballot for
// those below the running minimum) and recomputes the chosen candidate's
// extents in its own direction for the corners.
#include "lp_device.h"
#include "lp_internal.h"
#include "lp_kernels.h"

namespace lp {

using namespace lpd;

namespace {

constexpr int kBoxThreads = 256;
constexpr int kBoxVerts = 1024;       // polygon vertices held in LDS (larger: capacity error)
constexpr int kBoxPairs = 2 * kBoxVerts;  // antipodal pairs (at most 3n/2)

// areaOfTriangle (polygonizer.hpp:228-231), same operand order.
__device__ __forceinline__ double tri_area_xy(double ax, double ay, double bx, double by, double cx, double cy) {
    return fabs((ax * (by - cy) + bx * (cy - ay) + cx * (ay - by)) * 0.5);
}

__device__ __forceinline__ int wrap(int k, int n) { return k >= n ? k - n : k; }

// Number of consecutive steps j -> j+1 -> ... the reference's loop
// `while (area(i, i+1, j+1) > area(i, i+1, j)) ++j` takes from j, counted
// up to `cap` (one wavefront tests 64 steps at a time).
__device__ int advance_run(const double* X, const double* Y, int n, int i, int j, int cap) {
    const int lane = lane_id();
    const int i1 = wrap(i + 1, n);
    const double xi = X[i], yi = Y[i], xi1 = X[i1], yi1 = Y[i1];
    int t = 0;
    while (t < cap) {
        const int s = t + lane;
        const int a = (j + s) % n, b = (j + s + 1) % n;
        const bool go = tri_area_xy(xi, yi, xi1, yi1, X[b], Y[b]) > tri_area_xy(xi, yi, xi1, yi1, X[a], Y[a]);
        const uint64_t stop = __ballot(!go);
        if (stop) return min(cap, t + __ffsll(static_cast<unsigned long long>(stop)) - 1);
        t += kWave;
    }
    return cap;
}

// Antipodal pairs of polygon (X, Y, n >= 3) in the reference's order
// (polygonizer.cpp:93-161), by one wavefront; returns the count, or -1 when
// the walk exceeds 4n + 8 steps or kBoxPairs.
__device__ int antipodal_walk(const double* X, const double* Y, int n, int2* pairs) {
    const int lane = lane_id();
    const int i0 = n - 1;
    const int limit = 4 * n + 8;
    int guard = 0, count = 0;
    int i = 0, j = 1;
    {
        // Initial advance: every step counts against the guard.
        const int T = advance_run(X, Y, n, i, j, limit - guard + 1);
        if (guard + T > limit) return -1;
        guard += T;
        j = (j + T) % n;
    }
    const int j0 = j;
    while (i != j0) {
        i = wrap(i + 1, n);
        if (count >= kBoxPairs) return -1;
        if (lane == 0) pairs[count] = make_int2(i, j);
        ++count;
        // Inner advance, step t = 1..T: j + t; stop (returning the pairs so
        // far) when i == j0 and j + t == i0, fail past the guard.
        const int T = advance_run(X, Y, n, i, j, limit - guard + 1);
        int t_term = T + 1;
        if (i == j0) {
            const int d = ((i0 - j) % n + n) % n;  // first step t >= 1 with j + t == i0 (mod n)
            const int t_first = d > 0 ? d : n;
            if (t_first <= T) t_term = t_first;
        }
        const int t_guard = limit - guard + 1;  // first step that trips the guard
        if (t_term <= T && t_term <= t_guard) {
            if (count + t_term - 1 > kBoxPairs) return -1;
            for (int t = 1 + lane; t < t_term; t += kWave) pairs[count + t - 1] = make_int2(i, (j + t) % n);
            return count + t_term - 1;
        }
        if (t_guard <= T) return -1;
        if (count + T > kBoxPairs) return -1;
        for (int t = 1 + lane; t <= T; t += kWave) pairs[count + t - 1] = make_int2(i, (j + t) % n);
        count += T;
        guard += T;
        j = (j + T) % n;
        // areaOfTriangle(j, i+1, j+1) == areaOfTriangle(i, i+1, j)
        const int i1 = wrap(i + 1, n), j1 = wrap(j + 1, n);
        if (tri_area_xy(X[j], Y[j], X[i1], Y[i1], X[j1], Y[j1]) == tri_area_xy(X[i], Y[i], X[i1], Y[i1], X[j], Y[j])) {
            if (count >= kBoxPairs) return -1;
            if (lane == 0) pairs[count] = (i == j0 && j == i0) ? make_int2(i1, j) : make_int2(i, j1);
            ++count;
        }
        if (++guard > limit) return -1;
    }
    return count;
}

// Extents of the polygon rotated into the frame of unit direction (ux, uy)
// about (cx, cy), as the reference computes them (std::min / std::max order).
struct Extent {
    double mnx, mxx, mny, mxy;
};

__device__ __forceinline__ void extent_add(Extent& e, double rx, double ry) {
    e.mnx = (rx < e.mnx) ? rx : e.mnx;  // std::min(min_x, x)
    e.mxx = (e.mxx < rx) ? rx : e.mxx;  // std::max(max_x, x)
    e.mny = (ry < e.mny) ? ry : e.mny;
    e.mxy = (e.mxy < ry) ? ry : e.mxy;
}

__device__ __forceinline__ Extent extent_init() {
    return Extent{1.7976931348623157e308, -1.7976931348623157e308, 1.7976931348623157e308,
                  -1.7976931348623157e308};
}

// Candidate edge (idx, idx + off) of the reference: unit direction, or
// false when shorter than 1e-6.
__device__ __forceinline__ bool cand_dir(const double* X, const double* Y, int n, int idx, int off, double& ux,
                                         double& uy) {
    int nb = idx + off;
    if (nb < 0)
        nb += n;
    else if (nb >= n)
        nb -= n;
    const double ex = X[nb] - X[idx];
    const double ey = Y[nb] - Y[idx];
    const double len = sqrt(ex * ex + ey * ey);
    if (len < 1.0e-6) return false;
    ux = ex / len;
    uy = ey / len;
    return true;
}

// Polygon k of frame f: vertices VX/VY[ob + off[k] .. off[k+1]) with the
// offsets' end given by `total`.
__global__ __launch_bounds__(kBoxThreads) void k_boxes(DevBuf d, Layout L, const double* VX, const double* VY,
                                                      const uint32_t* off) {
    __shared__ double X[kBoxVerts];
    __shared__ double Y[kBoxVerts];
    __shared__ double area[kBoxVerts];  // per edge e (e -> e+1); NaN when too short
    __shared__ int2 pairs[kBoxPairs];
    __shared__ int np_s;
    __shared__ double cen_s[2];
    int tile, f;
    xcd_tile_frame(tile, f);
    const int tid = threadIdx.x;
    const int w = tid / kWave;
    const int lane = lane_id();
    const uint32_t* fs = d.fs + f * kFsCount;
    const int K = static_cast<int>(fs[kFsClusters]);
    const uint32_t total = fs[kFsHullTotal];
    const size_t ob = (size_t)f * L.N;
    if (K > L.box_stride) {
        if (tile == 0 && tid == 0) atomicOr(&d.fs[f * kFsCount + kFsErr], kErrBoxCapacity);
        return;
    }
    for (int k = tile; k < K; k += gridDim.x) {
        const uint32_t v0 = off[ob + k];
        const uint32_t v1 = k + 1 < K ? off[ob + k + 1] : total;
        const int n = static_cast<int>(v1 - v0);
        lp_bounding_box box;
        for (int q = 0; q < 8; ++q) box.corners[q] = 0.0;
        box.area = 0.0f;
        box.angle_rad = 0.0f;
        box.is_valid = 0;
        box.reserved = 0;
        if (n > kBoxVerts) {
            if (tid == 0) atomicOr(&d.fs[f * kFsCount + kFsErr], kErrBoxCapacity);
        } else if (n >= 3) {
            for (int q = tid; q < n; q += kBoxThreads) {
                X[q] = VX[ob + v0 + q];
                Y[q] = VY[ob + v0 + q];
            }
            __syncthreads();
            if (w == 0) {
                const int np = antipodal_walk(X, Y, n, pairs);
                if (lane == 0) np_s = np;
            } else if (w == 1 && lane == 0) {
                double cx = 0.0, cy = 0.0;
                for (int q = 0; q < n; ++q) {
                    cx += X[q];
                    cy += Y[q];
                }
                cen_s[0] = cx / n;
                cen_s[1] = cy / n;
            }
            __syncthreads();
            const int np = np_s;
            const double cx = cen_s[0], cy = cen_s[1];
            if (np < 0) {
                if (tid == 0) atomicOr(&d.fs[f * kFsCount + kFsErr], kErrBoxCapacity);
            } else {
                // Area of every edge in its forward direction.
                for (int e = tid; e < n; e += kBoxThreads) {
                    double ux, uy;
                    double a = __builtin_nan("");
                    if (cand_dir(X, Y, n, e, 1, ux, uy)) {
                        Extent ex = extent_init();
                        for (int q = 0; q < n; ++q) {
                            const double tx = X[q] - cx;
                            const double ty = Y[q] - cy;
                            extent_add(ex, tx * ux + ty * uy, -tx * uy + ty * ux);
                        }
                        a = (ex.mxx - ex.mnx) * (ex.mxy - ex.mny);
                    }
                    area[e] = a;
                }
                __syncthreads();
                if (w == 0) {
                    // Candidate c: pair c/4, vertex (c & 2 ? j : i), offset (c & 1 ? +1 : -1);
                    // a candidate replaces the box when its area is below
                    // the float running minimum (polygonizer.cpp:257).
                    float best = __builtin_inff();  // float(DBL_MAX)
                    int sel = -1;
                    const int nc = 4 * np;
                    for (int base = 0; base < nc; base += kWave) {
                        const int c = base + lane;
                        double a = __builtin_nan("");
                        if (c < nc) {
                            const int2 pr = pairs[c >> 2];
                            const int idx = (c & 2) ? pr.y : pr.x;
                            a = area[(c & 1) ? idx : (idx == 0 ? n - 1 : idx - 1)];
                        }
                        uint64_t beat = __ballot(a < static_cast<double>(best));
                        while (beat) {
                            const int q = __ffsll(static_cast<unsigned long long>(beat)) - 1;
                            beat &= beat - 1;
                            const double aq = __shfl(a, q, kWave);
                            if (aq < static_cast<double>(best)) {
                                best = static_cast<float>(aq);
                                sel = base + q;
                            }
                        }
                    }
                    if (sel >= 0) {
                        // The chosen candidate in its own direction.
                        const int2 pr = pairs[sel >> 2];
                        const int idx = (sel & 2) ? pr.y : pr.x;
                        double ux = 0.0, uy = 0.0;
                        cand_dir(X, Y, n, idx, (sel & 1) ? 1 : -1, ux, uy);
                        // Lanes take every 64th vertex; the reduction keeps,
                        // among equal values, the one of the lowest vertex
                        // index, which is what the sequential std::min /
                        // std::max keep (equal values differ at most in the
                        // sign of a zero).
                        Extent ex = extent_init();
                        int qa = n, qb = n, qc = n, qd = n;
                        for (int q = lane; q < n; q += kWave) {
                            const double tx = X[q] - cx;
                            const double ty = Y[q] - cy;
                            const double rx = tx * ux + ty * uy, ry = -tx * uy + ty * ux;
                            if (rx < ex.mnx) { ex.mnx = rx; qa = q; }
                            if (ex.mxx < rx) { ex.mxx = rx; qb = q; }
                            if (ry < ex.mny) { ex.mny = ry; qc = q; }
                            if (ex.mxy < ry) { ex.mxy = ry; qd = q; }
                        }
#pragma unroll
                        for (int o = kWave / 2; o > 0; o >>= 1) {
                            const double a0 = __shfl_xor(ex.mnx, o, kWave), a1 = __shfl_xor(ex.mxx, o, kWave);
                            const double a2 = __shfl_xor(ex.mny, o, kWave), a3 = __shfl_xor(ex.mxy, o, kWave);
                            const int b0 = __shfl_xor(qa, o, kWave), b1 = __shfl_xor(qb, o, kWave);
                            const int b2 = __shfl_xor(qc, o, kWave), b3 = __shfl_xor(qd, o, kWave);
                            if (a0 < ex.mnx || (a0 == ex.mnx && b0 < qa)) { ex.mnx = a0; qa = b0; }
                            if (ex.mxx < a1 || (a1 == ex.mxx && b1 < qb)) { ex.mxx = a1; qb = b1; }
                            if (a2 < ex.mny || (a2 == ex.mny && b2 < qc)) { ex.mny = a2; qc = b2; }
                            if (ex.mxy < a3 || (a3 == ex.mxy && b3 < qd)) { ex.mxy = a3; qd = b3; }
                        }
                        box.area = best;
                        box.corners[0] = ex.mnx * ux - ex.mny * uy + cx;
                        box.corners[1] = ex.mnx * uy + ex.mny * ux + cy;
                        box.corners[2] = ex.mxx * ux - ex.mny * uy + cx;
                        box.corners[3] = ex.mxx * uy + ex.mny * ux + cy;
                        box.corners[4] = ex.mxx * ux - ex.mxy * uy + cx;
                        box.corners[5] = ex.mxx * uy + ex.mxy * ux + cy;
                        box.corners[6] = ex.mnx * ux - ex.mxy * uy + cx;
                        box.corners[7] = ex.mnx * uy + ex.mxy * ux + cy;
                        box.angle_rad = atan2_approx(static_cast<float>(uy), static_cast<float>(ux));
                        box.is_valid = 1;
                    }
                }
            }
        }
        if (tid == 0) {
            d.boxes[(size_t)f * L.box_stride + k] = box;
            if (box.is_valid) atomicAdd(&d.fs[f * kFsCount + kFsBoxes], 1u);
        }
        __syncthreads();
    }
}

// Hull vertex coordinates (PointXY doubles of the obstacle points, in hull
// order) into the second hull candidate buffer, free once the hulls are done.
__global__ void k_box_gather(DevBuf d, Layout L) {
    const int f = blockIdx.y;
    const uint32_t* fs = d.fs + f * kFsCount;
    const int K = static_cast<int>(fs[kFsClusters]);
    const uint32_t total = fs[kFsHullTotal];
    const size_t ob = (size_t)f * L.N;
    const size_t BN = (size_t)L.B * L.N;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < total; v += gridDim.x * blockDim.x) {
        int lo = 0, hi = K;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (d.hull_off[ob + mid] <= v)
                lo = mid + 1;
            else
                hi = mid;
        }
        const int c = lo - 1;
        const uint32_t i = v - d.hull_off[ob + c];
        const int32_t idx = d.hull_idx[ob + d.cl_off[ob + c] + i];
        d.hcx[BN + ob + v] = static_cast<double>(d.ox[ob + idx]);
        d.hcy[BN + ob + v] = static_cast<double>(d.oy[ob + idx]);
    }
}

}  // namespace

void launch_boxes(lp_context* c, int B, bool from_hulls) {
    DevBuf& d = c->d;
    const Layout& L = c->L;
    hipStream_t st = c->stream;
    const size_t BN = (size_t)L.B * L.N;
    stage_mark(c, "boxes");
    if (from_hulls) hipLaunchKernelGGL(k_box_gather, dim3(256, B), dim3(256), 0, st, d, L);
    hipLaunchKernelGGL(k_boxes, dim3(128, B), dim3(kBoxThreads), 0, st, d, L, d.hcx + BN, d.hcy + BN, d.hull_off);
}

}  // namespace lp
