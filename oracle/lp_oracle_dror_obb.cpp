/*
 * lp_oracle_dror_obb.cpp -- CPU restatement of the DROR noise filter and the
 * rotating-calipers oriented bounding box.
 *
 * TEST INFRASTRUCTURE ONLY (part of the oracle library; see lp_oracle.h).
 *
 * Noise filter (paths relative to the reference repository root):
 *   lidar_processing_lib/src/noise_remover.cpp:38-68            NoiseRemover::filter
 *   lidar_processing_lib/include/lidar_processing_lib/noise_remover.hpp:41-54  configuration
 *   lidar_processing_lib/include/lidar_processing_lib/kdtree.hpp:131-143   dist_sqr
 *   lidar_processing_lib/include/lidar_processing_lib/kdtree.hpp:170-213   rebuild
 *   lidar_processing_lib/include/lidar_processing_lib/kdtree.hpp:339-397   radius_search_k_nearest
 *
 * Two forms are restated:
 *   lpo_noise_definition      the filter's stated rule: a point is NOISE when
 *                             fewer than min_neighbours points (itself
 *                             included) lie within its search radius;
 *   lpo_noise_remover_filter  a persistent NoiseRemover object with the
 *                             KD-tree's search stack kept between queries and
 *                             between frames: radius_search_k_nearest leaves
 *                             the stack non-empty when it stops at the k-th
 *                             neighbour (kdtree.hpp:351-371, the `break`) and
 *                             the next query pushes its root on top of those
 *                             entries, so a query whose own traversal ends
 *                             with fewer than k neighbours goes on through the
 *                             earlier queries' unvisited subtrees and can count
 *                             the same point again.
 *
 * Oriented box: lidar_processing_lib/src/polygonizer.cpp:93-278
 * (findAntipodalPairsOfConvexHull, boundingBoxRotatingCalipers) with
 * areaOfTriangle (polygonizer.hpp:228-231) and atan2Approx (common.hpp:33-62).
 */
#include "lp_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

namespace {

// Search radius squared of one point (noise_remover.cpp:46-59): the range is
// summed in float and widened to double; the product with the squared
// multiplier is rounded back to float.
inline float dror_radius_sqr(const lpo_noise_config& cfg, float x, float y) {
    const double scaling = std::pow(static_cast<double>(cfg.radius_multiplier_m_per_m), 2.0);
    const float min_sqr = cfg.min_search_radius_m * cfg.min_search_radius_m;
    const double range_sqr = (x * x) + (y * y);
    return std::max(static_cast<float>(scaling * range_sqr), min_sqr);
}

// dist_sqr (kdtree.hpp:131-143): (t0-p0)^2 + ((t1-p1)^2 + ((t2-p2)^2 + 0)).
inline float dist_sqr3(const float* t, const float* p) {
    const float d0 = t[0] - p[0], d1 = t[1] - p[1], d2 = t[2] - p[2];
    return d0 * d0 + (d1 * d1 + (d2 * d2 + 0.0f));
}

// ---------------------------------------------------------------- KD tree
struct KdNode {
    float p[3];
    uint32_t index;
    int32_t left;   // slot of the left child, -1 = none
    int32_t right;
};

struct Range {
    int32_t begin, end;
    int32_t* ref;   // where the chosen slot is stored (root or a child field)
    uint32_t depth;
};

struct SearchEntry {
    int32_t slot;   // -1 = null node
    uint32_t axis;
};

}  // namespace

// The reference object: nodes_ is reserved for 200000 entries once
// (noise_remover.cpp:32-36, kdtree.hpp:170-180), so node addresses are stable
// slots and the entries a stale stack holds refer to whatever the slot holds
// at the time they are popped.
struct lpo_noise_remover {
    std::vector<KdNode> nodes;      // fixed capacity, slots beyond the frame keep old contents
    std::vector<SearchEntry> stack; // radius_search_stack_, kept between calls
    int32_t root = -1;
    int32_t capacity = 200000;

    lpo_noise_remover() : nodes(200000, KdNode{{0, 0, 0}, 0, -1, -1}) {}

    void rebuild(const float* x, const float* y, const float* z, int32_t n) {
        root = -1;
        if (n == 0) return;
        for (int32_t i = 0; i < n; ++i) nodes[i] = KdNode{{x[i], y[i], z[i]}, static_cast<uint32_t>(i), -1, -1};
        std::vector<Range> rs;
        rs.push_back({0, n, &root, 0u});
        while (!rs.empty()) {
            const Range r = rs.back();
            rs.pop_back();
            if (r.begin >= r.end) continue;
            const uint32_t axis = r.depth % 3u;
            const int32_t mid = r.begin + (r.end - r.begin) / 2;
            std::nth_element(nodes.begin() + r.begin, nodes.begin() + mid, nodes.begin() + r.end,
                             [axis](const KdNode& a, const KdNode& b) { return a.p[axis] < b.p[axis]; });
            *r.ref = mid;
            if (mid > r.begin) rs.push_back({r.begin, mid, &nodes[mid].left, r.depth + 1});
            if (mid + 1 < r.end) rs.push_back({mid + 1, r.end, &nodes[mid].right, r.depth + 1});
        }
    }

    // radius_search_k_nearest (kdtree.hpp:339-397) with sort_ = false; returns
    // the neighbour count it reports.
    uint32_t search(const float* t, float prox, uint32_t k) {
        uint32_t found = 0;
        if (root < 0) return 0;
        stack.push_back({root, 0u});
        while (!stack.empty()) {
            const SearchEntry e = stack.back();
            stack.pop_back();
            if (e.slot < 0) continue;
            const KdNode& nd = nodes[e.slot];
            const float d = dist_sqr3(t, nd.p);
            if (d <= prox) {
                ++found;
                if (found >= k) break;
            }
            const uint32_t next = (e.axis + 1u) % 3u;
            const float delta = nd.p[e.axis] - t[e.axis];
            if (delta * delta <= prox) {
                stack.push_back({nd.right, next});
                stack.push_back({nd.left, next});
            } else {
                stack.push_back({delta > 0 ? nd.left : nd.right, next});
            }
        }
        return found;
    }
};

extern "C" {

void lpo_noise_config_default(lpo_noise_config* c) {
    c->radius_multiplier_m_per_m = 0.02f;
    c->min_search_radius_m = 0.1f;
    c->min_neighbours = 4u;
}

float lpo_noise_radius_sqr(const lpo_noise_config* cfg, float x, float y) { return dror_radius_sqr(*cfg, x, y); }

int lpo_noise_definition(const lpo_noise_config* cfg, const float* x, const float* y, const float* z, int32_t n,
                         uint8_t* labels, uint32_t* counts) {
    if (!cfg || n < 0) return LPO_ERR_INVALID;
    if (n == 0) return LPO_OK;
    // Uniform 1 m grid over the finite points' bounding box; a query with
    // radius r scans the cells within floor(r + 1/64) + 1 of its own, which
    // holds every point whose float distance passes the test (the float test
    // admits at most a few ulps beyond r).  Only points within 1e6 m of the
    // origin on every axis enter the grid; a query uses it when its target
    // lies 128 m inside that box and its radius is under 64 m (so every
    // possible neighbour is in the grid), and compares against every point
    // otherwise (non-finite or far targets, huge radii).
    auto finite = [&](int32_t i) {
        return std::fabs(x[i]) <= 1e6f && std::fabs(y[i]) <= 1e6f && std::fabs(z[i]) <= 1e6f;
    };
    auto inner = [&](int32_t i) {
        return std::fabs(x[i]) <= 1e6f - 128.0f && std::fabs(y[i]) <= 1e6f - 128.0f &&
               std::fabs(z[i]) <= 1e6f - 128.0f;
    };
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    bool any = false;
    const float* c3[3] = {x, y, z};
    for (int32_t i = 0; i < n; ++i) {
        if (!finite(i)) continue;
        for (int a = 0; a < 3; ++a) {
            lo[a] = any ? std::min(lo[a], c3[a][i]) : c3[a][i];
            hi[a] = any ? std::max(hi[a], c3[a][i]) : c3[a][i];
        }
        any = true;
    }
    int64_t dim[3];
    for (int a = 0; a < 3; ++a) dim[a] = static_cast<int64_t>(std::floor(hi[a]) - std::floor(lo[a])) + 1;
    auto cell_of = [&](int a, float v) { return static_cast<int64_t>(std::floor(v) - std::floor(lo[a])); };
    std::vector<int32_t> order;
    std::vector<int64_t> skey;
    {
        std::vector<std::pair<int64_t, int32_t>> kv;
        for (int32_t i = 0; i < n; ++i)
            if (finite(i)) kv.push_back({(cell_of(2, z[i]) * dim[1] + cell_of(1, y[i])) * dim[0] + cell_of(0, x[i]), i});
        std::sort(kv.begin(), kv.end());
        for (const auto& e : kv) {
            skey.push_back(e.first);
            order.push_back(e.second);
        }
    }
    for (int32_t i = 0; i < n; ++i) {
        const float t[3] = {x[i], y[i], z[i]};
        const float r2 = dror_radius_sqr(*cfg, x[i], y[i]);
        uint32_t cnt = 0;
        const bool grid = inner(i) && r2 >= 0.0f && std::sqrt(static_cast<double>(r2)) < 64.0;
        if (!grid) {
            for (int32_t j = 0; j < n; ++j) {
                const float p[3] = {x[j], y[j], z[j]};
                if (dist_sqr3(t, p) <= r2) ++cnt;
            }
        } else {
            const int64_t e = static_cast<int64_t>(std::floor(std::sqrt(static_cast<double>(r2)) + 1.0 / 64)) + 1;
            const int64_t cx = cell_of(0, x[i]), cy = cell_of(1, y[i]), cz = cell_of(2, z[i]);
            for (int64_t kz = std::max<int64_t>(cz - e, 0); kz <= std::min<int64_t>(cz + e, dim[2] - 1); ++kz)
                for (int64_t ky = std::max<int64_t>(cy - e, 0); ky <= std::min<int64_t>(cy + e, dim[1] - 1); ++ky) {
                    const int64_t row = (kz * dim[1] + ky) * dim[0];
                    const int64_t k0 = row + std::max<int64_t>(cx - e, 0);
                    const int64_t k1 = row + std::min<int64_t>(cx + e, dim[0] - 1);
                    auto it = std::lower_bound(skey.begin(), skey.end(), k0);
                    for (; it != skey.end() && *it <= k1; ++it) {
                        const int32_t j = order[it - skey.begin()];
                        const float p[3] = {x[j], y[j], z[j]};
                        if (dist_sqr3(t, p) <= r2) ++cnt;
                    }
                }
        }
        if (counts) counts[i] = cnt;
        labels[i] = cnt < cfg->min_neighbours ? 1 : 0;
    }
    return LPO_OK;
}

lpo_noise_remover* lpo_noise_remover_create(void) { return new lpo_noise_remover(); }
void lpo_noise_remover_destroy(lpo_noise_remover* p) { delete p; }

int lpo_noise_remover_filter(lpo_noise_remover* p, const lpo_noise_config* cfg, const float* x, const float* y,
                             const float* z, int32_t n, uint8_t* labels) {
    if (!p || !cfg || n < 0) return LPO_ERR_INVALID;
    // More points than the reserved 200000 nodes would move the node storage
    // under the stale stack entries (undefined in the reference).
    if (n > p->capacity) return LPO_ERR_INVALID;
    for (int32_t i = 0; i < n; ++i) labels[i] = 0;
    p->rebuild(x, y, z, n);
    for (int32_t i = 0; i < n; ++i) {
        const float t[3] = {x[i], y[i], z[i]};
        const float r2 = dror_radius_sqr(*cfg, x[i], y[i]);
        if (p->search(t, r2, cfg->min_neighbours) < cfg->min_neighbours) labels[i] = 1;
    }
    return LPO_OK;
}

int32_t lpo_noise_remover_stack_depth(const lpo_noise_remover* p) { return static_cast<int32_t>(p->stack.size()); }

// ------------------------------------------------------------ oriented box

static inline double area_of_triangle(const double* x, const double* y, int32_t a, int32_t b, int32_t c) {
    return std::fabs((x[a] * (y[b] - y[c]) + x[b] * (y[c] - y[a]) + x[c] * (y[a] - y[b])) * 0.5);
}

int32_t lpo_antipodal_pairs(const double* x, const double* y, int32_t n, int32_t* pairs, int32_t cap) {
    // polygonizer.cpp:93-161.  Returns the pair count (pairs may be truncated
    // to cap), or -1 when the walk exceeds 4n+8 steps.
    int32_t count = 0;
    auto push = [&](int32_t a, int32_t b) {
        if (count < cap) {
            pairs[2 * count] = a;
            pairs[2 * count + 1] = b;
        }
        ++count;
    };
    if (n < 2) return 0;
    const int32_t i0 = n - 1;
    int32_t i = 0, j = 1;
    auto nx = [n](int32_t k) { return k + 1 == n ? 0 : k + 1; };
    int64_t guard = 0;
    const int64_t limit = 4 * static_cast<int64_t>(n) + 8;
    while (area_of_triangle(x, y, i, nx(i), nx(j)) > area_of_triangle(x, y, i, nx(i), j)) {
        j = nx(j);
        if (++guard > limit) return -1;
    }
    const int32_t j0 = j;
    while (i != j0) {
        i = nx(i);
        push(i, j);
        while (area_of_triangle(x, y, i, nx(i), nx(j)) > area_of_triangle(x, y, i, nx(i), j)) {
            j = nx(j);
            if (!(i == j0 && j == i0)) {
                push(i, j);
            } else {
                return count;
            }
            if (++guard > limit) return -1;
        }
        if (area_of_triangle(x, y, j, nx(i), nx(j)) == area_of_triangle(x, y, i, nx(i), j)) {
            if (!(i == j0 && j == i0)) {
                push(i, nx(j));
            } else {
                push(nx(i), j);
            }
        }
        if (++guard > limit) return -1;
    }
    return count;
}

int lpo_obb_rotating_calipers(const double* x, const double* y, int32_t n, lpo_bounding_box* out) {
    // polygonizer.cpp:163-278.  The box's area is a float member compared
    // against each double candidate area (strict <), so the float rounding of
    // the running minimum decides near-ties exactly as in the reference.
    for (int k = 0; k < 8; ++k) out->corners[k] = 0.0;
    out->area = 0.0f;
    out->angle_rad = 0.0f;
    out->is_valid = 0;
    if (n < 3) return LPO_OK;
    std::vector<int32_t> pairs(2 * (4 * static_cast<size_t>(n) + 16));
    const int32_t np = lpo_antipodal_pairs(x, y, n, pairs.data(), static_cast<int32_t>(pairs.size() / 2));
    if (np < 0) return LPO_ERR_INVALID;
    if (np > static_cast<int32_t>(pairs.size() / 2)) return LPO_ERR_INVALID;
    double cx = 0.0, cy = 0.0;
    for (int32_t k = 0; k < n; ++k) {
        cx += x[k];
        cy += y[k];
    }
    cx /= n;
    cy /= n;
    float best = std::numeric_limits<float>::infinity();  // float(DBL_MAX)
    for (int32_t q = 0; q < np; ++q) {
        const int32_t pr[2] = {pairs[2 * q], pairs[2 * q + 1]};
        for (int32_t idx : pr) {
            for (int32_t off : {-1, 1}) {
                int32_t nb = idx + off;
                if (nb < 0) nb += n;
                else if (nb >= n) nb -= n;
                const double ex = x[nb] - x[idx];
                const double ey = y[nb] - y[idx];
                const double len = std::sqrt(ex * ex + ey * ey);
                if (len < 1.0e-6) continue;
                const double ux = ex / len;
                const double uy = ey / len;
                double mnx = std::numeric_limits<double>::max(), mxx = std::numeric_limits<double>::lowest();
                double mny = std::numeric_limits<double>::max(), mxy = std::numeric_limits<double>::lowest();
                for (int32_t k = 0; k < n; ++k) {
                    const double tx = x[k] - cx;
                    const double ty = y[k] - cy;
                    const double rx = tx * ux + ty * uy;
                    const double ry = -tx * uy + ty * ux;
                    mnx = std::min(mnx, rx);
                    mxx = std::max(mxx, rx);
                    mny = std::min(mny, ry);
                    mxy = std::max(mxy, ry);
                }
                const double area = (mxx - mnx) * (mxy - mny);
                if (area < best) {
                    best = static_cast<float>(area);
                    out->area = best;
                    out->corners[0] = mnx * ux - mny * uy + cx;
                    out->corners[1] = mnx * uy + mny * ux + cy;
                    out->corners[2] = mxx * ux - mny * uy + cx;
                    out->corners[3] = mxx * uy + mny * ux + cy;
                    out->corners[4] = mxx * ux - mxy * uy + cx;
                    out->corners[5] = mxx * uy + mxy * ux + cy;
                    out->corners[6] = mnx * ux - mxy * uy + cx;
                    out->corners[7] = mnx * uy + mxy * ux + cy;
                    out->angle_rad = lpo_atan2_approx(static_cast<float>(uy), static_cast<float>(ux));
                    out->is_valid = 1;
                }
            }
        }
    }
    return LPO_OK;
}

}  // extern "C"
