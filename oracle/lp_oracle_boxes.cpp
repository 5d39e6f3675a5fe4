/*
 * lp_oracle_boxes.cpp -- CPU restatement of the principal-component box and
 * of the processor's vehicle test.
 *
 * TEST INFRASTRUCTURE ONLY (part of the oracle library; see lp_oracle.h).
 *
 * PCA box: lidar_processing_lib/src/polygonizer.cpp:280-362
 * (boundingBoxPrincipalComponentAnalysis).  The reference evaluates it with
 * Eigen 3.4 (find_package(Eigen3 3.4), lidar_processing_lib/CMakeLists.txt:28),
 * which is absent here:
 *   - colwise().mean() and the covariance product centred^T * centred are
 *     restated as sequential sums over the points; Eigen may vectorise or
 *     block these sums in another order, so the last bits of the mean and
 *     covariance are not pinned to the reference;
 *   - JacobiSVD<MatrixXd>::compute(cov, ComputeThinV) on the 2x2 covariance
 *     is restated from its published algorithm: scale by the largest
 *     absolute entry, sweep the (p, q) = (1, 0) pair with the real 2x2 step
 *     (a rotation that symmetrises the block, then the symmetric Jacobi
 *     rotation of Golub & Van Loan 8.4.2) while an off-diagonal entry exceeds
 *     max(DBL_MIN, 2 eps * max |diagonal|), then order the singular values
 *     (and the columns of V) largest first.
 * The box itself (rotation of the centred points into V's frame, extents,
 * corners (min,min), (max,min), (max,max), (min,max) mapped back by V^T plus
 * the mean, float area, float atan2(V10, V00)) follows polygonizer.cpp:337-360.
 *
 * Vehicle test: src/processor/src/processor.cpp:672-760 with the tables of
 * src/processor/src/processor.hpp:60-192 (base dimensions, tolerances, the
 * adjusted dimensions, per-type volume / area bounds and the loosest height /
 * volume bounds) and polygonArea / distance (polygonizer.hpp:185-214).
 */
#include "lp_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>

namespace {

struct VehicleTables {
    double dims[5][6];
    double bounds[5][4];
    double min_h, max_h, min_vol, max_vol;
};

// processor.hpp:78-192, evaluated in the same order.
VehicleTables make_tables() {
    static const double base[5][6] = {
        {4.3, 4.6, 1.6, 1.9, 1.4, 1.5},  // compact
        {4.6, 5.0, 1.6, 1.9, 1.4, 1.5},  // sedan
        {4.6, 5.2, 1.7, 2.1, 1.7, 1.8},  // SUV
        {5.2, 5.8, 1.9, 2.2, 1.8, 2.0},  // truck
        {4.9, 5.2, 1.7, 2.1, 1.7, 1.8},  // minivan
    };
    const double length_tol = 0.8, width_tol = 0.5, height_tol = 0.5;
    VehicleTables t{};
    for (int i = 0; i < 5; ++i) {
        t.dims[i][0] = base[i][0] - length_tol;
        t.dims[i][1] = base[i][1] + length_tol;
        t.dims[i][2] = base[i][2] - width_tol;
        t.dims[i][3] = base[i][3] + width_tol;
        t.dims[i][4] = base[i][4] - height_tol;
        t.dims[i][5] = base[i][5] + height_tol;
    }
    for (int i = 0; i < 5; ++i) {
        const double* d = t.dims[i];
        t.bounds[i][0] = d[0] * d[2] * d[4];
        t.bounds[i][1] = d[1] * d[3] * d[5];
        t.bounds[i][2] = d[0] * d[2];
        t.bounds[i][3] = d[1] * d[3];
    }
    t.min_h = std::numeric_limits<double>::max();
    t.max_h = std::numeric_limits<double>::lowest();
    t.min_vol = std::numeric_limits<double>::max();
    t.max_vol = std::numeric_limits<double>::lowest();
    for (int i = 0; i < 5; ++i) {
        if (t.dims[i][4] < t.min_h) t.min_h = t.dims[i][4];
        if (t.dims[i][5] > t.max_h) t.max_h = t.dims[i][5];
        if (t.bounds[i][0] < t.min_vol) t.min_vol = t.bounds[i][0];
        if (t.bounds[i][1] > t.max_vol) t.max_vol = t.bounds[i][1];
    }
    return t;
}

const VehicleTables& tables() {
    static const VehicleTables t = make_tables();
    return t;
}

double distance_xy(double ax, double ay, double bx, double by) {
    const double dx = ax - bx;
    const double dy = ay - by;
    return std::sqrt(dx * dx + dy * dy);
}

// The symmetric Jacobi rotation (c, s) that zeroes y in [[x, y], [y, z]]:
// tau = (x - z) / 2|y|, t = 1 / (tau + sqrt(tau^2 + 1)) for tau > 0 else
// 1 / (tau - sqrt(tau^2 + 1)), c = 1 / sqrt(t^2 + 1), s = -sign(t) (y/|y|) |t| c.
void symmetric_rotation(double x, double y, double z, double& c, double& s) {
    const double deno = 2.0 * std::fabs(y);
    if (deno < std::numeric_limits<double>::min()) {
        c = 1.0;
        s = 0.0;
        return;
    }
    const double tau = (x - z) / deno;
    const double w = std::sqrt(tau * tau + 1.0);
    const double t = tau > 0.0 ? 1.0 / (tau + w) : 1.0 / (tau - w);
    const double sign_t = t > 0.0 ? 1.0 : -1.0;
    const double n = 1.0 / std::sqrt(t * t + 1.0);
    s = -sign_t * (y / std::fabs(y)) * std::fabs(t) * n;
    c = n;
}

// Rows x, y of a 2x2 (or columns, with the caller's indexing) rotated by
// (c, s): x' = c x + s y, y' = -s x + c y; the identity leaves them alone.
void rotate_pair(double& x0, double& x1, double& y0, double& y1, double c, double s) {
    if (c == 1.0 && s == 0.0) return;
    const double a0 = x0, a1 = x1, b0 = y0, b1 = y1;
    x0 = c * a0 + s * b0;
    x1 = c * a1 + s * b1;
    y0 = -s * a0 + c * b0;
    y1 = -s * a1 + c * b1;
}

bool svd2_v(const double a[2][2], double v[2][2]) {
    double scale = std::max(std::max(std::fabs(a[0][0]), std::fabs(a[0][1])),
                            std::max(std::fabs(a[1][0]), std::fabs(a[1][1])));
    if (!std::isfinite(scale)) return false;
    if (scale == 0.0) scale = 1.0;
    double w[2][2] = {{a[0][0] / scale, a[0][1] / scale}, {a[1][0] / scale, a[1][1] / scale}};
    v[0][0] = 1.0;
    v[0][1] = 0.0;
    v[1][0] = 0.0;
    v[1][1] = 1.0;
    const double precision = 2.0 * std::numeric_limits<double>::epsilon();
    const double tiny = std::numeric_limits<double>::min();
    double max_diag = std::max(std::fabs(w[0][0]), std::fabs(w[1][1]));
    bool finished = false;
    while (!finished) {
        finished = true;
        const int p = 1, q = 0;
        const double threshold = std::max(tiny, precision * max_diag);
        if (std::fabs(w[p][q]) > threshold || std::fabs(w[q][p]) > threshold) {
            finished = false;
            // The block [[w_pp, w_pq], [w_qp, w_qq]].
            double m00 = w[p][p], m01 = w[p][q], m10 = w[q][p], m11 = w[q][q];
            // Symmetrising rotation.
            double c1 = 1.0, s1 = 0.0;
            const double t = m00 + m11;
            const double d = m10 - m01;
            if (!(std::fabs(d) < tiny)) {
                const double u = t / d;
                const double tmp = std::sqrt(1.0 + u * u);
                s1 = 1.0 / tmp;
                c1 = u / tmp;
            }
            rotate_pair(m00, m01, m10, m11, c1, s1);
            double cr, sr;
            symmetric_rotation(m00, m01, m11, cr, sr);
            // Left rotation = symmetrising rotation times the transpose of
            // the right one; the right one acts on columns as its transpose.
            const double cl = c1 * cr - s1 * (-sr);
            const double sl = c1 * (-sr) + s1 * cr;
            rotate_pair(w[p][0], w[p][1], w[q][0], w[q][1], cl, sl);
            rotate_pair(w[0][p], w[1][p], w[0][q], w[1][q], cr, -sr);
            rotate_pair(v[0][p], v[1][p], v[0][q], v[1][q], cr, -sr);
            max_diag = std::max(max_diag, std::max(std::fabs(w[p][p]), std::fabs(w[q][q])));
        }
    }
    const double sv0 = std::fabs(w[0][0]) * scale, sv1 = std::fabs(w[1][1]) * scale;
    if (!(sv0 == 0.0 && sv1 == 0.0) && sv1 > sv0) {
        std::swap(v[0][0], v[0][1]);
        std::swap(v[1][0], v[1][1]);
    }
    return true;
}

}  // namespace

extern "C" {

int lpo_svd2_v(const double* a, double* v) {
    const double m[2][2] = {{a[0], a[1]}, {a[2], a[3]}};
    double r[2][2];
    if (!svd2_v(m, r)) return LPO_ERR_INVALID;
    v[0] = r[0][0];
    v[1] = r[0][1];
    v[2] = r[1][0];
    v[3] = r[1][1];
    return LPO_OK;
}

int lpo_pca_box(const double* x, const double* y, int32_t n, lpo_bounding_box* out) {
    for (int k = 0; k < 8; ++k) out->corners[k] = 0.0;
    out->area = 0.0f;
    out->angle_rad = 0.0f;
    out->is_valid = 0;
    if (n < 3) return LPO_OK;
    double sx = x[0], sy = y[0];
    for (int32_t i = 1; i < n; ++i) {
        sx += x[i];
        sy += y[i];
    }
    const double mx = sx / static_cast<double>(n);
    const double my = sy / static_cast<double>(n);
    double c00 = 0.0, c01 = 0.0, c11 = 0.0;
    for (int32_t i = 0; i < n; ++i) {
        const double cx = x[i] - mx, cy = y[i] - my;
        c00 += cx * cx;
        c01 += cx * cy;
        c11 += cy * cy;
    }
    const double dn = static_cast<double>(static_cast<uint32_t>(n) - 1u);
    const double cov[2][2] = {{c00 / dn, c01 / dn}, {c01 / dn, c11 / dn}};
    double v[2][2];
    if (!svd2_v(cov, v)) return LPO_OK;
    double mnx = std::numeric_limits<double>::max(), mxx = std::numeric_limits<double>::lowest();
    double mny = std::numeric_limits<double>::max(), mxy = std::numeric_limits<double>::lowest();
    for (int32_t i = 0; i < n; ++i) {
        const double cx = x[i] - mx, cy = y[i] - my;
        const double rx = cx * v[0][0] + cy * v[1][0];
        const double ry = cx * v[0][1] + cy * v[1][1];
        mnx = std::min(mnx, rx);
        mxx = std::max(mxx, rx);
        mny = std::min(mny, ry);
        mxy = std::max(mxy, ry);
    }
    const double px[4] = {mnx, mxx, mxx, mnx};
    const double py[4] = {mny, mny, mxy, mxy};
    for (int j = 0; j < 4; ++j) {
        out->corners[2 * j] = (px[j] * v[0][0] + py[j] * v[0][1]) + mx;
        out->corners[2 * j + 1] = (px[j] * v[1][0] + py[j] * v[1][1]) + my;
    }
    out->area = static_cast<float>((mxx - mnx) * (mxy - mny));
    out->angle_rad = static_cast<float>(std::atan2(v[1][0], v[0][0]));
    out->is_valid = 1;
    return LPO_OK;
}

void lpo_vehicle_tables(double* dims, double* bounds, double* loose) {
    const VehicleTables& t = tables();
    for (int i = 0; i < 5; ++i) {
        for (int k = 0; k < 6; ++k) dims[6 * i + k] = t.dims[i][k];
        for (int k = 0; k < 4; ++k) bounds[4 * i + k] = t.bounds[i][k];
    }
    loose[0] = t.min_h;
    loose[1] = t.max_h;
    loose[2] = t.min_vol;
    loose[3] = t.max_vol;
}

int32_t lpo_vehicle_class(const double* hx, const double* hy, int32_t nv, int32_t cluster_size, float z_min,
                          float z_max, const lpo_bounding_box* box) {
    const VehicleTables& t = tables();
    const double height = static_cast<double>(z_max) - static_cast<double>(z_min);
    if (!(nv >= 3 && cluster_size > 150 && height > t.min_h && height < t.max_h)) return -1;
    double area = 0.0;
    for (int32_t i = 0; i < nv - 1; ++i) area += (hx[i] * hy[i + 1]) - (hx[i + 1] * hy[i]);
    area += (hx[nv - 1] * hy[0]) - (hx[0] * hy[nv - 1]);
    area = std::fabs(area) * 0.5;
    const double volume = area * height;
    if (!(volume > t.min_vol && volume < t.max_vol)) return -1;
    if (!box->is_valid) return -1;
    const double iou = area / box->area;
    if (!(iou > 0.4)) return -1;
    const double* c = box->corners;
    const double e1 = distance_xy(c[0], c[1], c[2], c[3]);
    const double e2 = distance_xy(c[2], c[3], c[4], c[5]);
    const double length = std::max(e1, e2);
    const double width = std::min(e1, e2);
    for (int i = 0; i < 5; ++i) {
        const double* d = t.dims[i];
        const double* b = t.bounds[i];
        if (length > d[0] && length < d[1] && width > d[2] && width < d[3] && height > d[4] && height < d[5] &&
            volume > b[0] && volume < b[1] && area > b[2] && area < b[3])
            return i;
    }
    return -1;
}

}  // extern "C"
