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
// One wavefront per polygon: lane 0 walks the antipodal pairs (sequential by
// nature, O(n)) into LDS; the lanes then evaluate 64 candidates at a time
// (O(n) each) and lane 0 applies the sequential selection rule to the few
// that beat the running minimum.
#include "lp_device.h"
#include "lp_internal.h"
#include "lp_kernels.h"

namespace lp {

using namespace lpd;

namespace {

constexpr int kBoxWaves = 4;
constexpr int kBoxPairs = 512;   // antipodal pairs held per polygon (about 1.5 per vertex)
constexpr int kBoxVerts = 256;   // polygons up to this size are staged in LDS

struct Cand {
    double area, ux, uy, mnx, mxx, mny, mxy;
    int valid;
};

// areaOfTriangle (polygonizer.hpp:228-231), same operand order.
__device__ __forceinline__ double tri_area_xy(double ax, double ay, double bx, double by, double cx, double cy) {
    return fabs((ax * (by - cy) + bx * (cy - ay) + cx * (ay - by)) * 0.5);
}

// Antipodal pairs of polygon (X, Y, n >= 3) in the reference's order; returns
// the count, or -1 when the walk exceeds 4n + 8 steps or the LDS capacity.
// Vertices i, i+1, j and j+1 stay in registers: each step loads one vertex.
__device__ int antipodal_walk(const double* X, const double* Y, int n, int2* pairs) {
    int count = 0;
    auto push = [&](int a, int b) -> bool {
        if (count >= kBoxPairs) return false;
        pairs[count++] = make_int2(a, b);
        return true;
    };
    auto nx = [n](int k) { return k + 1 == n ? 0 : k + 1; };
    const int i0 = n - 1;
    int i = 0, j = 1;
    double xi = X[0], yi = Y[0], xi1 = X[1], yi1 = Y[1];
    double xj = X[1], yj = Y[1], xj1 = X[2], yj1 = Y[2];
    auto adv_i = [&]() {
        i = nx(i);
        xi = xi1;
        yi = yi1;
        const int q = nx(i);
        xi1 = X[q];
        yi1 = Y[q];
    };
    auto adv_j = [&]() {
        j = nx(j);
        xj = xj1;
        yj = yj1;
        const int q = nx(j);
        xj1 = X[q];
        yj1 = Y[q];
    };
    const int limit = 4 * n + 8;
    int guard = 0;
    // areaOfTriangle(i, i+1, j+1) > areaOfTriangle(i, i+1, j)
    while (tri_area_xy(xi, yi, xi1, yi1, xj1, yj1) > tri_area_xy(xi, yi, xi1, yi1, xj, yj)) {
        adv_j();
        if (++guard > limit) return -1;
    }
    const int j0 = j;
    while (i != j0) {
        adv_i();
        if (!push(i, j)) return -1;
        while (tri_area_xy(xi, yi, xi1, yi1, xj1, yj1) > tri_area_xy(xi, yi, xi1, yi1, xj, yj)) {
            adv_j();
            if (!(i == j0 && j == i0)) {
                if (!push(i, j)) return -1;
            } else {
                return count;
            }
            if (++guard > limit) return -1;
        }
        // areaOfTriangle(j, i+1, j+1) == areaOfTriangle(i, i+1, j)
        if (tri_area_xy(xj, yj, xi1, yi1, xj1, yj1) == tri_area_xy(xi, yi, xi1, yi1, xj, yj)) {
            if (!(i == j0 && j == i0)) {
                if (!push(i, nx(j))) return -1;
            } else {
                if (!push(nx(i), j)) return -1;
            }
        }
        if (++guard > limit) return -1;
    }
    return count;
}

// Polygon k of frame f: vertices VX/VY[vo + off[k] .. off[k+1]) with the
// offsets' end given by `total`.
__global__ __launch_bounds__(kBoxWaves * kWave) void k_boxes(DevBuf d, Layout L, const double* VX, const double* VY,
                                                             const uint32_t* off) {
    __shared__ int2 pairs_l[kBoxWaves][kBoxPairs];
    __shared__ Cand cand_l[kBoxWaves][kWave];
    __shared__ int np_l[kBoxWaves];
    __shared__ double cen_l[kBoxWaves][2];
    __shared__ double vx_l[kBoxWaves][kBoxVerts];
    __shared__ double vy_l[kBoxWaves][kBoxVerts];
    const int f = blockIdx.y;
    const int w = threadIdx.x / kWave;
    const int lane = lane_id();
    const uint32_t* fs = d.fs + f * kFsCount;
    const int K = static_cast<int>(fs[kFsClusters]);
    const uint32_t total = fs[kFsHullTotal];
    const size_t ob = (size_t)f * L.N;
    if (K > L.box_stride) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&d.fs[f * kFsCount + kFsErr], kErrBoxCapacity);
        return;
    }
    int2* pairs = pairs_l[w];
    Cand* cl = cand_l[w];
    for (int k = blockIdx.x * kBoxWaves + w; k < K; k += gridDim.x * kBoxWaves) {
        const uint32_t v0 = off[ob + k];
        const uint32_t v1 = k + 1 < K ? off[ob + k + 1] : total;
        const int n = static_cast<int>(v1 - v0);
        const double* X = VX + ob + v0;
        const double* Y = VY + ob + v0;
        if (n <= kBoxVerts) {
            for (int q = lane; q < n; q += kWave) {
                vx_l[w][q] = X[q];
                vy_l[w][q] = Y[q];
            }
            wave_sync();
            X = vx_l[w];
            Y = vy_l[w];
        }
        lp_bounding_box box;
        for (int q = 0; q < 8; ++q) box.corners[q] = 0.0;
        box.area = 0.0f;
        box.angle_rad = 0.0f;
        box.is_valid = 0;
        box.reserved = 0;
        if (n >= 3) {
            if (lane == 0) {
                np_l[w] = antipodal_walk(X, Y, n, pairs);
                double cx = 0.0, cy = 0.0;
                for (int q = 0; q < n; ++q) {
                    cx += X[q];
                    cy += Y[q];
                }
                cen_l[w][0] = cx / n;
                cen_l[w][1] = cy / n;
            }
            wave_sync();
            const int np = np_l[w];
            const double cx = cen_l[w][0], cy = cen_l[w][1];
            if (np < 0) {
                if (lane == 0) atomicOr(&d.fs[f * kFsCount + kFsErr], kErrBoxCapacity);
            } else {
                float best = __builtin_inff();  // float(DBL_MAX)
                Cand sel{};
                sel.valid = 0;
                const int nc = 4 * np;
                for (int base = 0; base < nc; base += kWave) {
                    const int c = base + lane;
                    Cand r{};
                    r.valid = 0;
                    if (c < nc) {
                        const int2 pr = pairs[c >> 2];
                        const int idx = (c & 2) ? pr.y : pr.x;
                        int nb = idx + ((c & 1) ? 1 : -1);
                        if (nb < 0)
                            nb += n;
                        else if (nb >= n)
                            nb -= n;
                        const double ex = X[nb] - X[idx];
                        const double ey = Y[nb] - Y[idx];
                        const double len = sqrt(ex * ex + ey * ey);
                        if (!(len < 1.0e-6)) {
                            const double ux = ex / len;
                            const double uy = ey / len;
                            double mnx = 1.7976931348623157e308, mxx = -1.7976931348623157e308;
                            double mny = 1.7976931348623157e308, mxy = -1.7976931348623157e308;
                            for (int q = 0; q < n; ++q) {
                                const double tx = X[q] - cx;
                                const double ty = Y[q] - cy;
                                const double rx = tx * ux + ty * uy;
                                const double ry = -tx * uy + ty * ux;
                                mnx = (rx < mnx) ? rx : mnx;  // std::min(min_x, x)
                                mxx = (mxx < rx) ? rx : mxx;  // std::max(max_x, x)
                                mny = (ry < mny) ? ry : mny;
                                mxy = (mxy < ry) ? ry : mxy;
                            }
                            r.area = (mxx - mnx) * (mxy - mny);
                            r.ux = ux;
                            r.uy = uy;
                            r.mnx = mnx;
                            r.mxx = mxx;
                            r.mny = mny;
                            r.mxy = mxy;
                            r.valid = 1;
                        }
                    }
                    cl[lane] = r;
                    wave_sync();
                    // Only candidates below the round's starting minimum can
                    // replace the box (the minimum only falls within a round);
                    // lane 0 applies the rule to them in candidate order.
                    uint64_t beat = __ballot(r.valid && r.area < static_cast<double>(best));
                    if (lane == 0) {
                        while (beat) {
                            const int q = __ffsll(static_cast<unsigned long long>(beat)) - 1;
                            beat &= beat - 1;
                            const Cand& e = cl[q];
                            if (e.area < static_cast<double>(best)) {
                                best = static_cast<float>(e.area);
                                sel = e;
                            }
                        }
                    }
                    best = __shfl(best, 0, kWave);
                    wave_sync();
                }
                if (lane == 0 && sel.valid) {
                    const double ux = sel.ux, uy = sel.uy;
                    box.area = best;
                    box.corners[0] = sel.mnx * ux - sel.mny * uy + cx;
                    box.corners[1] = sel.mnx * uy + sel.mny * ux + cy;
                    box.corners[2] = sel.mxx * ux - sel.mny * uy + cx;
                    box.corners[3] = sel.mxx * uy + sel.mny * ux + cy;
                    box.corners[4] = sel.mxx * ux - sel.mxy * uy + cx;
                    box.corners[5] = sel.mxx * uy + sel.mxy * ux + cy;
                    box.corners[6] = sel.mnx * ux - sel.mxy * uy + cx;
                    box.corners[7] = sel.mnx * uy + sel.mxy * ux + cy;
                    box.angle_rad = atan2_approx(static_cast<float>(uy), static_cast<float>(ux));
                    box.is_valid = 1;
                }
            }
        }
        if (lane == 0) {
            d.boxes[(size_t)f * L.box_stride + k] = box;
            if (box.is_valid) atomicAdd(&d.fs[f * kFsCount + kFsBoxes], 1u);
        }
        wave_sync();
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
    hipLaunchKernelGGL(k_boxes, dim3(64, B), dim3(kBoxWaves * kWave), 0, st, d, L, d.hcx + BN, d.hcy + BN,
                       d.hull_off);
}

}  // namespace lp
