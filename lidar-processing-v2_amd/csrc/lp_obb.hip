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
// Polygons of up to 64 vertices (nearly all hulls) take one wavefront each,
// larger ones a workgroup.  One lane walks the antipodal pairs and the
// centroid is summed in the reference's order.  The candidates (i, -1) and
// (i - 1, +1) share one edge with opposite directions, which negates every
// rotated coordinate exactly and leaves the area bit-identical, so one area
// is evaluated per edge, O(n) each, a lane per edge; one wavefront then
// applies the sequential selection rule over the candidates in the
// reference's order (64 at a time, through a prefix minimum of the areas) and
// recomputes the chosen candidate's extents in its own direction for the
// corners.
#include <algorithm>

#include "lp_device.h"
#include "lp_internal.h"
#include "lp_kernels.h"

namespace lp {

using namespace lpd;

namespace {

constexpr int kBoxThreads = 256;
// Polygons of up to kBoxVerts vertices are held in LDS (larger: capacity
// error).  Antipodal pairs: at most 3n/2, room for 2n.
constexpr int kBoxVerts = 1024;

// areaOfTriangle (polygonizer.hpp:228-231), same operand order.
__device__ __forceinline__ double tri_area_xy(double ax, double ay, double bx, double by, double cx, double cy) {
    return fabs((ax * (by - cy) + bx * (cy - ay) + cx * (ay - by)) * 0.5);
}

__device__ __forceinline__ int wrap(int k, int n) { return k >= n ? k - n : k; }

// Antipodal pairs of polygon (n >= 3 vertices) in the reference's order
// (polygonizer.cpp:93-161), walked as one sequential chain with the areas of
// the running pair carried from step to step.  vert(stream, k, x, y) fetches
// vertex k; stream 0 serves i and stream 1 serves j, and each stream's
// vertex indices advance by one (mod n) from call to call.  emit(c, i, j)
// records pair c.  Returns the count, or -1 when the walk exceeds 4n + 8
// steps or `cap` pairs (the reference's loop need not end on degenerate
// input).
template <typename Vert, typename Emit>
__device__ __forceinline__ int antipodal_walk(int n, int cap, Vert vert, Emit emit) {
    const int i0 = n - 1;
    const int limit = 4 * n + 8;
    int guard = 0, count = 0;
    int i = 0, j = 1;
    double xi, yi, xi1, yi1, xj, yj, xn, yn;
    vert(0, 0, xi, yi);
    vert(0, 1, xi1, yi1);
    xj = xi1;
    yj = yi1;
    int jn = wrap(2, n);
    vert(1, jn, xn, yn);
    double a_cur = tri_area_xy(xi, yi, xi1, yi1, xj, yj);
    // while (area(i, i+1, j+1) > area(i, i+1, j)) ++j
    for (;;) {
        const double a_nx = tri_area_xy(xi, yi, xi1, yi1, xn, yn);
        if (!(a_nx > a_cur)) break;
        j = jn;
        xj = xn;
        yj = yn;
        a_cur = a_nx;
        jn = wrap(j + 1, n);
        vert(1, jn, xn, yn);
        if (++guard > limit) return -1;
    }
    const int j0 = j;
    while (i != j0) {
        i = wrap(i + 1, n);
        xi = xi1;
        yi = yi1;
        const int i1 = wrap(i + 1, n);
        vert(0, i1, xi1, yi1);
        if (count >= cap) return -1;
        emit(count++, i, j);
        a_cur = tri_area_xy(xi, yi, xi1, yi1, xj, yj);
        for (;;) {
            const double a_nx = tri_area_xy(xi, yi, xi1, yi1, xn, yn);
            if (!(a_nx > a_cur)) break;
            j = jn;
            xj = xn;
            yj = yn;
            a_cur = a_nx;
            jn = wrap(j + 1, n);
            vert(1, jn, xn, yn);
            if (i == j0 && j == i0) return count;
            if (count >= cap) return -1;
            emit(count++, i, j);
            if (++guard > limit) return -1;
        }
        // areaOfTriangle(j, i+1, j+1) == areaOfTriangle(i, i+1, j)
        if (tri_area_xy(xj, yj, xi1, yi1, xn, yn) == a_cur) {
            if (count >= cap) return -1;
            if (i == j0 && j == i0)
                emit(count++, i1, j);
            else
                emit(count++, i, jn);
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

// Area of the box in the frame of edge e (e -> e+1), or NaN when the edge is
// shorter than 1e-6.  The candidates (e, +1) and (e + 1, -1) share it.
// vert(q, x, y) fetches vertex q for the extents (the same for every lane).
template <typename Vert>
__device__ __forceinline__ double edge_box_area(const double* X, const double* Y, int n, int e, double cx, double cy,
                                                Vert vert) {
    double ux, uy;
    if (!cand_dir(X, Y, n, e, 1, ux, uy)) return __builtin_nan("");
    Extent ex = extent_init();
    // Unrolled so the vertex fetches of two steps are in flight together;
    // the order of the updates stays.
#pragma unroll 2
    for (int q = 0; q < n; ++q) {
        double xq, yq;
        vert(q, xq, yq);
        const double tx = xq - cx;
        const double ty = yq - cy;
        extent_add(ex, tx * ux + ty * uy, -tx * uy + ty * ux);
    }
    return (ex.mxx - ex.mnx) * (ex.mxy - ex.mny);
}

// The reference's sequential choice among the 4 * np candidates of the
// antipodal pairs, run by one whole wavefront (every lane returns the same
// box).  Candidate c: pair c/4, vertex (c & 2 ? j : i), offset (c & 1 ? +1 :
// -1); a candidate replaces the box when its area is below the float running
// minimum (polygonizer.cpp:257).  `area` holds the area of every edge.
__device__ void choose_box(const double* X, const double* Y, int n, const double* area, const int2* pairs, int np,
                           double cx, double cy, lp_bounding_box& box) {
    const int lane = lane_id();
    // The running minimum after candidate c is float(min of the areas up to
    // c): float rounding is monotone, so a candidate that does not replace
    // the box lies between the smallest area so far and its float value and
    // rounds to the same float.  So candidate c replaces the box exactly when
    // its area is below float(min of the areas before it), which a prefix
    // minimum over the lanes gives for 64 candidates at once; the box is the
    // last replacing candidate.  NaN areas (short edges, no candidate) never
    // replace and do not enter the minimum.
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
        double incl = (a == a) ? a : __builtin_inf();
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const double u = __shfl_up(incl, o, kWave);
            if (lane >= o) incl = fmin(incl, u);
        }
        double excl = __shfl_up(incl, 1, kWave);
        if (lane == 0) excl = __builtin_inf();
        const float before = fminf(best, static_cast<float>(excl));
        const uint64_t rep = __ballot(a < static_cast<double>(before));
        if (rep) sel = base + 63 - __clzll(static_cast<long long>(rep));
        best = fminf(best, static_cast<float>(__shfl(incl, kWave - 1, kWave)));
    }
    if (sel < 0) return;
    // The chosen candidate in its own direction.
    const int2 pr = pairs[sel >> 2];
    const int idx = (sel & 2) ? pr.y : pr.x;
    double ux = 0.0, uy = 0.0;
    cand_dir(X, Y, n, idx, (sel & 1) ? 1 : -1, ux, uy);
    // Lanes take every 64th vertex; the reduction keeps, among equal values,
    // the one of the lowest vertex index, which is what the sequential
    // std::min / std::max keep (equal values differ at most in the sign of a
    // zero).
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

__device__ __forceinline__ lp_bounding_box empty_box() {
    lp_bounding_box box;
    for (int q = 0; q < 8; ++q) box.corners[q] = 0.0;
    box.area = 0.0f;
    box.angle_rad = 0.0f;
    box.is_valid = 0;
    box.reserved = 0;
    return box;
}

// Polygons of up to one vertex per lane: one wavefront each, four per
// workgroup (a grid-stride loop over the frame's polygons), the vertices in
// the wavefront's own LDS slice.  Polygon k of frame f: vertices
// VX/VY[ob + off[k] .. off[k+1]) with the offsets' end given by `total`.
// Larger polygons are listed for k_boxes.
constexpr int kBoxWaves = kBoxThreads / kWave;

__global__ __launch_bounds__(kBoxThreads) void k_boxes_wave(DevBuf d, Layout L, const double* VX, const double* VY,
                                                            const uint32_t* off) {
    __shared__ double Xs[kBoxWaves][kWave];
    __shared__ double Ys[kBoxWaves][kWave];
    __shared__ double areas[kBoxWaves][kWave];
    __shared__ int2 pairs_s[kBoxWaves][2 * kWave];
    int tile, f;
    xcd_tile_frame(tile, f);
    const int w = threadIdx.x / kWave;
    const int lane = lane_id();
    double* X = Xs[w];
    double* Y = Ys[w];
    double* area = areas[w];
    int2* pairs = pairs_s[w];
    const uint32_t* fs = d.fs + f * kFsCount;
    const int K = static_cast<int>(fs[kFsClusters]);
    const uint32_t total = fs[kFsHullTotal];
    const size_t ob = (size_t)f * L.N;
    if (K > L.box_stride) {
        if (tile == 0 && threadIdx.x == 0) atomicOr(&d.fs[f * kFsCount + kFsErr], kErrBoxCapacity);
        return;
    }
    for (int k = tile * kBoxWaves + w; k < K; k += gridDim.x * kBoxWaves) {
        const uint32_t v0 = off[ob + k];
        const uint32_t v1 = k + 1 < K ? off[ob + k + 1] : total;
        const int n = static_cast<int>(v1 - v0);
        if (n > kWave) {
            if (lane == 0) {
                const uint32_t q = atomicAdd(&d.fs[f * kFsCount + kFsBoxLarge], 1u);
                d.box_large[(size_t)f * L.box_stride + q] = k;
            }
            continue;
        }
        lp_bounding_box box = empty_box();
        if (n >= 3) {
            if (lane < n) {
                X[lane] = VX[ob + v0 + lane];
                Y[lane] = VY[ob + v0 + lane];
            }
            wave_sync();
            int np = 0;
            if (lane == 0)
                np = antipodal_walk(
                    n, 2 * kWave, [&](int, int q, double& vx, double& vy) { vx = X[q]; vy = Y[q]; },
                    [&](int c, int a, int b) { pairs[c] = make_int2(a, b); });
            np = __shfl(np, 0, kWave);
            wave_sync();
            // Centroid: the vertices summed in order, as the reference does.
            double cx = 0.0, cy = 0.0;
            for (int q = 0; q < n; ++q) {
                cx += X[q];
                cy += Y[q];
            }
            cx /= n;
            cy /= n;
            if (np < 0) {
                if (lane == 0) atomicOr(&d.fs[f * kFsCount + kFsErr], kErrBoxCapacity);
            } else {
                if (lane < n)
                    area[lane] = edge_box_area(X, Y, n, lane, cx, cy, [&](int q, double& vx, double& vy) {
                        vx = X[q];
                        vy = Y[q];
                    });
                wave_sync();
                choose_box(X, Y, n, area, pairs, np, cx, cy, box);
            }
        }
        if (lane == 0) {
            d.boxes[(size_t)f * L.box_stride + k] = box;
            if (box.is_valid) atomicAdd(&d.fs[f * kFsCount + kFsBoxes], 1u);
        }
        wave_sync();
    }
}

// Polygons of more than one vertex per lane, one workgroup each (up to
// kBoxVerts vertices; larger: capacity error): workgroup `tile` takes the
// entries tile, tile + gridDim.x, ... of the frame's list written by
// k_boxes_wave.
__global__ __launch_bounds__(kBoxThreads) void k_boxes(DevBuf d, Layout L, const double* VX, const double* VY,
                                                      const uint32_t* off) {
    __shared__ double X[kBoxVerts];
    __shared__ double Y[kBoxVerts];
    __shared__ double area[kBoxVerts];  // per edge e (e -> e+1); NaN when too short
    __shared__ int2 pairs[2 * kBoxVerts];
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
    const int m = static_cast<int>(fs[kFsBoxLarge]);
    const size_t ob = (size_t)f * L.N;
    if (K > L.box_stride) return;  // flagged by k_boxes_wave
    for (int j = tile; j < m; j += gridDim.x) {
        const int k = d.box_large[(size_t)f * L.box_stride + j];
        const uint32_t v0 = off[ob + k];
        const uint32_t v1 = k + 1 < K ? off[ob + k + 1] : total;
        const int n = static_cast<int>(v1 - v0);
        lp_bounding_box box = empty_box();
        if (n > kBoxVerts) {
            if (tid == 0) atomicOr(&d.fs[f * kFsCount + kFsErr], kErrBoxCapacity);
        } else {
            for (int q = tid; q < n; q += kBoxThreads) {
                X[q] = VX[ob + v0 + q];
                Y[q] = VY[ob + v0 + q];
            }
            __syncthreads();
            // The first wavefront walks the antipodal pairs while the
            // others sum the centroid (each on its own, in the
            // reference's order) and evaluate the edge areas.
            if (w == 0) {
                if (lane == 0)
                    np_s = antipodal_walk(
                        n, 2 * kBoxVerts, [&](int, int q, double& x, double& y) { x = X[q]; y = Y[q]; },
                        [&](int c, int a, int b) { pairs[c] = make_int2(a, b); });
            } else {
                double cx = 0.0, cy = 0.0;
                for (int q = 0; q < n; ++q) {
                    cx += X[q];
                    cy += Y[q];
                }
                cx /= n;
                cy /= n;
                if (tid == kWave) {
                    cen_s[0] = cx;
                    cen_s[1] = cy;
                }
                for (int e = tid - kWave; e < n; e += kBoxThreads - kWave)
                    area[e] = edge_box_area(X, Y, n, e, cx, cy, [&](int q, double& vx, double& vy) {
                        vx = X[q];
                        vy = Y[q];
                    });
            }
            __syncthreads();
            const int np = np_s;
            if (np < 0) {
                if (tid == 0) atomicOr(&d.fs[f * kFsCount + kFsErr], kErrBoxCapacity);
            } else if (w == 0) {
                choose_box(X, Y, n, area, pairs, np, cen_s[0], cen_s[1], box);
            }
        }
        if (tid == 0) {
            d.boxes[(size_t)f * L.box_stride + k] = box;
            if (box.is_valid) atomicAdd(&d.fs[f * kFsCount + kFsBoxes], 1u);
        }
        __syncthreads();
    }
}

// Vehicle size classes of the processor's polygon simplification
// (src/processor/src/processor.hpp:78-192): the five base dimension ranges
// widened by the length / width / height tolerances, their volume and area
// bounds, and the loosest height and volume bounds over all five.
struct VehicleConst {
    double dims[5][6];    // min/max length, width, height
    double bounds[5][4];  // min/max volume, min/max area
    double min_h, max_h, min_vol, max_vol;
};

VehicleConst vehicle_const() {
    static const double base[5][6] = {{4.3, 4.6, 1.6, 1.9, 1.4, 1.5},
                                      {4.6, 5.0, 1.6, 1.9, 1.4, 1.5},
                                      {4.6, 5.2, 1.7, 2.1, 1.7, 1.8},
                                      {5.2, 5.8, 1.9, 2.2, 1.8, 2.0},
                                      {4.9, 5.2, 1.7, 2.1, 1.7, 1.8}};
    const double tol[3] = {0.8, 0.5, 0.5};
    VehicleConst v{};
    v.min_h = 1.7976931348623157e308;
    v.max_h = -1.7976931348623157e308;
    v.min_vol = 1.7976931348623157e308;
    v.max_vol = -1.7976931348623157e308;
    for (int i = 0; i < 5; ++i) {
        for (int k = 0; k < 3; ++k) {
            v.dims[i][2 * k] = base[i][2 * k] - tol[k];
            v.dims[i][2 * k + 1] = base[i][2 * k + 1] + tol[k];
        }
        const double* a = v.dims[i];
        v.bounds[i][0] = a[0] * a[2] * a[4];
        v.bounds[i][1] = a[1] * a[3] * a[5];
        v.bounds[i][2] = a[0] * a[2];
        v.bounds[i][3] = a[1] * a[3];
        if (a[4] < v.min_h) v.min_h = a[4];
        if (a[5] > v.max_h) v.max_h = a[5];
        if (v.bounds[i][0] < v.min_vol) v.min_vol = v.bounds[i][0];
        if (v.bounds[i][1] > v.max_vol) v.max_vol = v.bounds[i][1];
    }
    return v;
}

__device__ __forceinline__ double dist_xy(double ax, double ay, double bx, double by) {
    const double dx = ax - bx;
    const double dy = ay - by;
    return sqrt(dx * dx + dy * dy);
}

// Processor::run's polygon simplification test (processor.cpp:680-757), one
// wavefront per cluster: the cluster's z range over its points, then on the
// first lane the shoelace area of the hull in hull order (polygonArea,
// polygonizer.hpp:185-198), the volume and box tests and the first vehicle
// class whose strict bounds all hold.  The box is the cluster hull's
// rotating-calipers box of the preceding stage.
__global__ __launch_bounds__(256) void k_vehicles(DevBuf d, Layout L, VehicleConst vc) {
    const int f = blockIdx.y;
    const uint32_t* fs = d.fs + f * kFsCount;
    const int K = min(static_cast<int>(fs[kFsClusters]), L.box_stride);
    const size_t ob = (size_t)f * L.N;
    const size_t BN = (size_t)L.B * L.N;
    const int lane = lane_id();
    const int waves = gridDim.x * (blockDim.x / kWave);
    for (int c = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; c < K; c += waves) {
        const uint32_t size = d.cl_size[ob + c];
        const uint32_t base = d.cl_off[ob + c];
        float zlo = 3.402823466e38f, zhi = -3.402823466e38f;
        for (uint32_t j = lane; j < size; j += kWave) {
            const float z = d.oz[ob + d.cl_members[ob + base + j]];
            zlo = z < zlo ? z : zlo;
            zhi = zhi < z ? z : zhi;
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            const float a = __shfl_xor(zlo, o, kWave), b = __shfl_xor(zhi, o, kWave);
            zlo = a < zlo ? a : zlo;
            zhi = zhi < b ? b : zhi;
        }
        if (lane == 0) {
            int32_t cls = -1;
            const int nv = static_cast<int>(d.hull_count[ob + c]);
            const double height = static_cast<double>(zhi) - static_cast<double>(zlo);
            if (nv >= 3 && size > 150u && height > vc.min_h && height < vc.max_h) {
                const double* X = d.hcx + BN + ob + d.hull_off[ob + c];
                const double* Y = d.hcy + BN + ob + d.hull_off[ob + c];
                double area = 0.0;
                for (int i = 0; i < nv - 1; ++i) area += X[i] * Y[i + 1] - X[i + 1] * Y[i];
                area += X[nv - 1] * Y[0] - X[0] * Y[nv - 1];
                area = fabs(area) * 0.5;
                const double volume = area * height;
                const lp_bounding_box& box = d.boxes[(size_t)f * L.box_stride + c];
                if (volume > vc.min_vol && volume < vc.max_vol && box.is_valid &&
                    area / static_cast<double>(box.area) > 0.4) {
                    const double* q = box.corners;
                    const double e1 = dist_xy(q[0], q[1], q[2], q[3]);
                    const double e2 = dist_xy(q[2], q[3], q[4], q[5]);
                    const double len = e1 < e2 ? e2 : e1;  // std::max
                    const double wid = e2 < e1 ? e2 : e1;  // std::min
                    for (int i = 0; i < 5; ++i) {
                        const double* v = vc.dims[i];
                        const double* b = vc.bounds[i];
                        if (len > v[0] && len < v[1] && wid > v[2] && wid < v[3] && height > v[4] && height < v[5] &&
                            volume > b[0] && volume < b[1] && area > b[2] && area < b[3]) {
                            cls = i;
                            break;
                        }
                    }
                }
            }
            d.vclass[(size_t)f * L.box_stride + c] = cls;
            if (cls >= 0) atomicAdd(&d.fs[f * kFsCount + kFsVehicles], 1u);
        }
    }
}

// One Jacobi rotation (c, s) of the 2x2 SVD sweep: the rotation that
// diagonalises the symmetric 2x2 [[x, y], [y, z]] (Golub & Van Loan's
// symmetric Schur step, with the sign choice t = 1 / (tau +- sqrt(tau^2 + 1))).
__device__ __forceinline__ void jacobi_rot(double x, double y, double z, double& c, double& s) {
    const double deno = 2.0 * fabs(y);
    if (deno < 2.2250738585072014e-308) {
        c = 1.0;
        s = 0.0;
        return;
    }
    const double tau = (x - z) / deno;
    const double w = sqrt(tau * tau + 1.0);
    const double t = tau > 0.0 ? 1.0 / (tau + w) : 1.0 / (tau - w);
    const double sign_t = t > 0.0 ? 1.0 : -1.0;
    const double n = 1.0 / sqrt(t * t + 1.0);
    s = -sign_t * (y / fabs(y)) * fabs(t) * n;
    c = n;
}

// Right singular vectors of the 2x2 covariance by two-sided Jacobi sweeps
// (the real 2x2 step: a rotation making the block symmetric, then the
// symmetric Schur rotation), largest singular value first.  v = [v00 v01;
// v10 v11], columns are the vectors.  False for a non-finite matrix.
__device__ bool svd2_v(const double (&a)[2][2], double (&v)[2][2]) {
    double scale = fmax(fmax(fabs(a[0][0]), fabs(a[0][1])), fmax(fabs(a[1][0]), fabs(a[1][1])));
    if (!isfinite(scale)) return false;
    if (scale == 0.0) scale = 1.0;
    double w[2][2] = {{a[0][0] / scale, a[0][1] / scale}, {a[1][0] / scale, a[1][1] / scale}};
    v[0][0] = 1.0;
    v[0][1] = 0.0;
    v[1][0] = 0.0;
    v[1][1] = 1.0;
    const double precision = 2.0 * 2.220446049250313e-16;
    double max_diag = fmax(fabs(w[0][0]), fabs(w[1][1]));
    // p = 1, q = 0 throughout (the only off-diagonal pair of a 2x2).
    for (int sweep = 0; sweep < 64; ++sweep) {
        const double thr = fmax(2.2250738585072014e-308, precision * max_diag);
        if (!(fabs(w[1][0]) > thr || fabs(w[0][1]) > thr)) break;
        // Block m = [[w11, w10], [w01, w00]] in (p, q) order.
        const double m00 = w[1][1], m01 = w[1][0], m10 = w[0][1], m11 = w[0][0];
        double c1 = 1.0, s1 = 0.0;
        const double tt = m00 + m11, dd = m10 - m01;
        if (!(fabs(dd) < 2.2250738585072014e-308)) {
            const double u = tt / dd;
            const double tmp = sqrt(1.0 + u * u);
            s1 = 1.0 / tmp;
            c1 = u / tmp;
        }
        // m <- rot1 applied on the left (rows 0, 1 of m); the symmetric
        // rotation reads m00, m01 and m11 of the result.
        double n00 = m00, n01 = m01, n11 = m11;
        if (!(c1 == 1.0 && s1 == 0.0)) {
            n00 = c1 * m00 + s1 * m10;
            n01 = c1 * m01 + s1 * m11;
            n11 = -s1 * m01 + c1 * m11;
        }
        double cr, sr;
        jacobi_rot(n00, n01, n11, cr, sr);
        // left = rot1 * right^T, right^T = (cr, -sr).
        const double cl = c1 * cr - s1 * (-sr);
        const double sl = c1 * (-sr) + s1 * cr;
        // w.applyOnTheLeft(p=1, q=0, left): x = row 1, y = row 0.
        if (!(cl == 1.0 && sl == 0.0)) {
            for (int k = 0; k < 2; ++k) {
                const double xi = w[1][k], yi = w[0][k];
                w[1][k] = cl * xi + sl * yi;
                w[0][k] = -sl * xi + cl * yi;
            }
        }
        // w and v .applyOnTheRight(p=1, q=0, right): x = column 1, y = column 0,
        // rotated by right^T = (cr, -sr).
        if (!(cr == 1.0 && -sr == 0.0)) {
            for (int k = 0; k < 2; ++k) {
                const double xi = w[k][1], yi = w[k][0];
                w[k][1] = cr * xi + (-sr) * yi;
                w[k][0] = -(-sr) * xi + cr * yi;
                const double vx = v[k][1], vy = v[k][0];
                v[k][1] = cr * vx + (-sr) * vy;
                v[k][0] = -(-sr) * vx + cr * vy;
            }
        }
        max_diag = fmax(max_diag, fmax(fabs(w[1][1]), fabs(w[0][0])));
    }
    const double sv0 = fabs(w[0][0]) * scale, sv1 = fabs(w[1][1]) * scale;
    if (sv0 != 0.0 || sv1 != 0.0) {
        if (sv1 > sv0) {
            const double t0 = v[0][0], t1 = v[1][0];
            v[0][0] = v[0][1];
            v[1][0] = v[1][1];
            v[0][1] = t0;
            v[1][1] = t1;
        }
    }
    return true;
}

// Polygonizer::boundingBoxPrincipalComponentAnalysis (polygonizer.cpp:280-360),
// one wavefront per list.  Every lane evaluates the mean and the covariance
// in the same sequential order (so all lanes hold identical values with no
// exchange), the lanes share the rotation of the centred points and the
// extents, and lane 0 writes the box.
__global__ __launch_bounds__(256) void k_pca_boxes(DevBuf d, Layout L, const double* VX, const double* VY,
                                                  const uint32_t* off) {
    const uint32_t* fs = d.fs;
    const int K = static_cast<int>(fs[kFsClusters]);
    const uint32_t total = fs[kFsHullTotal];
    const int lane = lane_id();
    const int waves = gridDim.x * (blockDim.x / kWave);
    for (int k = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; k < K; k += waves) {
        const uint32_t beg = off[k];
        const uint32_t end = k + 1 < K ? off[k + 1] : total;
        const int n = static_cast<int>(end - beg);
        const double* X = VX + beg;
        const double* Y = VY + beg;
        lp_bounding_box box{};
        double v[2][2] = {{1.0, 0.0}, {0.0, 1.0}};
        bool ok = n >= 3;
        double mx = 0.0, my = 0.0;
        if (ok) {
            double sx = X[0], sy = Y[0];
            for (int i = 1; i < n; ++i) {
                sx += X[i];
                sy += Y[i];
            }
            mx = sx / static_cast<double>(n);
            my = sy / static_cast<double>(n);
            double c00 = 0.0, c01 = 0.0, c11 = 0.0;
            for (int i = 0; i < n; ++i) {
                const double cx = X[i] - mx, cy = Y[i] - my;
                c00 += cx * cx;
                c01 += cx * cy;
                c11 += cy * cy;
            }
            const double dn = static_cast<double>(static_cast<uint32_t>(n - 1));
            const double a[2][2] = {{c00 / dn, c01 / dn}, {c01 / dn, c11 / dn}};
            ok = svd2_v(a, v);
        }
        if (ok) {
            Extent e = extent_init();
            for (int i = lane; i < n; i += kWave) {
                const double cx = X[i] - mx, cy = Y[i] - my;
                extent_add(e, cx * v[0][0] + cy * v[1][0], cx * v[0][1] + cy * v[1][1]);
            }
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                const double a0 = __shfl_xor(e.mnx, o, kWave), a1 = __shfl_xor(e.mxx, o, kWave);
                const double a2 = __shfl_xor(e.mny, o, kWave), a3 = __shfl_xor(e.mxy, o, kWave);
                e.mnx = a0 < e.mnx ? a0 : e.mnx;
                e.mxx = e.mxx < a1 ? a1 : e.mxx;
                e.mny = a2 < e.mny ? a2 : e.mny;
                e.mxy = e.mxy < a3 ? a3 : e.mxy;
            }
            const double cxs[4] = {e.mnx, e.mxx, e.mxx, e.mnx};
            const double cys[4] = {e.mny, e.mny, e.mxy, e.mxy};
            for (int j = 0; j < 4; ++j) {
                box.corners[2 * j] = (cxs[j] * v[0][0] + cys[j] * v[0][1]) + mx;
                box.corners[2 * j + 1] = (cxs[j] * v[1][0] + cys[j] * v[1][1]) + my;
            }
            box.area = static_cast<float>((e.mxx - e.mnx) * (e.mxy - e.mny));
            box.angle_rad = static_cast<float>(atan2(v[1][0], v[0][0]));
            box.is_valid = 1;
        }
        if (lane == 0) d.boxes[k] = box;
    }
}

}  // namespace

void launch_vehicles(lp_context* c, int B) {
    stage_mark(c, "vehicles");
    hipLaunchKernelGGL(k_vehicles, dim3(64, B), dim3(256), 0, c->stream, c->d, c->L, vehicle_const());
}

void launch_pca_boxes(lp_context* c, int num_polygons) {
    const size_t BN = (size_t)c->L.B * c->L.N;
    const int blocks = std::max(1, std::min(1024, (num_polygons + 3) / 4));
    hipLaunchKernelGGL(k_pca_boxes, dim3(blocks), dim3(256), 0, c->stream, c->d, c->L, c->d.hcx + BN,
                       c->d.hcy + BN, c->d.hull_off);
}

void launch_boxes(lp_context* c, int B) {
    DevBuf& d = c->d;
    const Layout& L = c->L;
    hipStream_t st = c->stream;
    const size_t BN = (size_t)L.B * L.N;
    stage_mark(c, "boxes");
    // Four polygons per workgroup; about 128 wavefronts per frame on large
    // batches, more for few frames (standalone polygon lists).
    const int tiles = std::min(1024, std::max(32, 4096 / B));
    hipLaunchKernelGGL(k_boxes_wave, dim3(tiles, B), dim3(kBoxThreads), 0, st, d, L, d.hcx + BN, d.hcy + BN,
                       d.hull_off);
    // Enough workgroups that each takes at most one of a frame's few large
    // polygons, whose walk is the longest chain of the stage.
    hipLaunchKernelGGL(k_boxes, dim3(std::min(256, std::max(4, 1024 / B)), B), dim3(kBoxThreads), 0, st, d, L,
                       d.hcx + BN, d.hcy + BN, d.hull_off);
}

}  // namespace lp
