/*
 * lp_oracle.cpp -- CPU restatement of the lidar_processing_lib hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See lp_oracle.h for scope.
 * Built like the reference Release build: g++ -O3, no -march, no fast-math
 * (lidar_processing_lib/CMakeLists.txt:11-21), so no FMA contraction happens
 * and every float expression rounds exactly as the reference's does.
 * Every expression that feeds a comparison keeps the reference's operand
 * order; comments cite the reference line each block restates.
 */
#include "lp_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <limits>
#include <numeric>
#include <random>
#include <set>
#include <vector>

namespace {

constexpr float kDegToRad = static_cast<float>(M_PI / 180.0);  // segmenter.hpp:135
constexpr float kTwoPi = static_cast<float>(2.0 * M_PI);        // segmenter.hpp:136
constexpr float kInvalidZ = std::numeric_limits<float>::max();  // segmenter.hpp:139

enum : uint32_t { UNKNOWN = 0, GROUND = 1, OBSTACLE = 2 };  // segmenter.hpp:69-74

// BGR colours of the label image (segmenter.hpp:144-148).
struct Bgr {
    uint8_t b, g, r;
    bool operator==(const Bgr& o) const { return b == o.b && g == o.g && r == o.r; }
};
constexpr Bgr kObstacle{0, 0, 255};
constexpr Bgr kGround{0, 255, 0};
constexpr Bgr kIntersectionOrUnknown{0, 255, 255};
constexpr Bgr kIntersection{255, 0, 0};
constexpr Bgr kUnknownColour{255, 255, 255};

// common.hpp:33-62
inline float atan2Approx(const float y, const float x) {
    const float ax = std::fabs(x);
    const float ay = std::fabs(y);
    const float mx = std::max(ay, ax);
    const float mn = std::min(ay, ax);
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

struct GridPoint {  // SegmenterPoint, segmenter.hpp:76-85
    float x, y, z;
    uint32_t label;
    uint16_t height_index;
    uint16_t width_index;
    int32_t cloud_index;
};

struct Seg {
    lpo_seg_config cfg;
    float slice_res_rad;
    int32_t rings;
    int32_t slices;
    std::vector<std::vector<GridPoint>> grid;
    std::vector<float> elevation;
    std::vector<int32_t> mapping;
    std::vector<float> depth;
    std::vector<Bgr> image;
    std::vector<float> cell_z;
    struct RP { float x, y, z; };
    std::vector<RP> ransac_points;
    // Eigen::Vector<float,24> members and std::array<Label,24>, reset once per
    // frame (segmenter.cpp:523-525) and NOT per queue pixel.
    float u[24];
    float w[24];
    uint32_t m[24];

    explicit Seg(const lpo_seg_config& c) : cfg(c) {
        // segmenter.cpp:38-71
        slice_res_rad = cfg.grid_slice_resolution_deg * kDegToRad;
        rings = static_cast<int32_t>(cfg.max_distance_m / cfg.grid_radial_spacing_m);
        slices = static_cast<int32_t>(kTwoPi / slice_res_rad);
        grid.resize(static_cast<size_t>(rings) * slices);
        for (auto& cell : grid) cell.reserve(400);
        cell_z.reserve(1000);
        elevation.assign(static_cast<size_t>(rings) * slices, kInvalidZ);
        const size_t P = static_cast<size_t>(cfg.image_height) * cfg.image_width;
        mapping.assign(P, -1);
        depth.assign(P, std::numeric_limits<float>::max());
        image.assign(P, Bgr{0, 0, 0});
        ransac_points.reserve(200000);
    }

    // segmenter.cpp:73-85
    void reset() {
        std::fill(image.begin(), image.end(), Bgr{0, 0, 0});
        std::fill(elevation.begin(), elevation.end(), kInvalidZ);
        std::fill(mapping.begin(), mapping.end(), -1);
        std::fill(depth.begin(), depth.end(), std::numeric_limits<float>::max());
        for (auto& cell : grid) cell.clear();
    }

    // segmenter.cpp:103-204
    void polarGrid(const float* X, const float* Y, const float* Z, const uint16_t* ring, int32_t n) {
        const float up = cfg.elevation_up_deg * kDegToRad;
        const float down = cfg.elevation_down_deg * kDegToRad;
        const float fov = up - down;
        const float per_pixel = fov / cfg.image_height;
        for (auto& cell : grid) cell.clear();
        if (n == 0) return;
        const float zlo = -cfg.sensor_height_m + cfg.z_min_m;
        const float zhi = -cfg.sensor_height_m + cfg.z_max_m;
        for (int32_t i = 0; i < n; ++i) {
            const float px = X[i], py = Y[i], pz = Z[i];
            if (pz < zlo || pz > zhi) continue;
            const float d = std::sqrt(px * px + py * py);
            const auto radial = static_cast<int32_t>(d / cfg.grid_radial_spacing_m);
            if (d < cfg.min_distance_m || d > cfg.max_distance_m || radial >= rings) continue;
            float az = atan2Approx(py, px);
            az = (az < 0) ? (az + kTwoPi) : az;
            const auto slice = std::min(static_cast<int32_t>(az / slice_res_rad), slices - 1);
            uint16_t h;
            if (ring != nullptr && !cfg.assume_unorganized_cloud) {
                h = ring[i];
                if (h >= cfg.image_height) continue;
            } else {
                const float elev = std::atan(pz / d);
                const auto hs = static_cast<int32_t>((elev - down) / per_pixel);
                if (hs < 0 || hs >= cfg.image_height) continue;
                h = static_cast<uint16_t>(hs);
            }
            const auto wi = static_cast<uint16_t>((cfg.image_width - 1) * az / kTwoPi);
            const auto cell = static_cast<uint32_t>(slice * rings + radial);
            grid[cell].push_back({px, py, pz, GROUND, h, wi, i});
        }
    }

    // segmenter.cpp:206-319
    int recm(lpo_seg_debug* dbg) {
        const float dz = std::min(cfg.grid_radial_spacing_m * std::tan(cfg.road_maximum_slope_m_per_m),
                                  cfg.ground_height_threshold_m - std::numeric_limits<float>::epsilon());
        for (int32_t s = 0; s < slices; ++s) {
            const int32_t off = s * rings;
            elevation[off] = -cfg.sensor_height_m + cfg.ground_height_threshold_m;
            float prev = elevation[off];
            for (int32_t r = 1; r < rings; ++r) {
                const int32_t c = off + r;
                const auto& cell = grid[c];
                if (cell.empty()) {
                    elevation[c] = prev + dz;
                    prev = elevation[c];
                    continue;
                }
                cell_z.clear();
                for (const auto& p : cell) cell_z.push_back(p.z);
                std::sort(cell_z.begin(), cell_z.end());
                float zmin = cell_z[0];
                for (int32_t i = static_cast<int32_t>(cell_z.size()) / 2; i >= 1; --i) {
                    if (cell_z[i] - cell_z[i - 1] > 0.5F) {
                        zmin = cell_z[i];
                        break;
                    }
                }
                elevation[c] = std::min(zmin, prev + dz);
                prev = elevation[c];
            }
        }
        for (uint32_t c = 0; c < grid.size(); ++c) {
            const float e = elevation[c];
            for (auto& p : grid[c])
                if (p.z >= e + cfg.ground_height_threshold_m) p.label = OBSTACLE;
        }
        const int st = ransac(dbg);
        if (st != LPO_OK) return st;
        for (const auto& cell : grid) {
            for (const auto& p : cell) {
                const int32_t h = p.height_index, wi = p.width_index;
                const int32_t pix = h * cfg.image_width + wi;
                const float dep = (p.x * p.x) + (p.y * p.y);
                if (dep < depth[pix]) {
                    if (p.label == GROUND) {
                        image[pix] = kGround;
                        depth[pix] = dep;
                        mapping[pix] = p.cloud_index;
                    } else if (p.label == OBSTACLE) {
                        image[pix] = kObstacle;
                        depth[pix] = dep;
                        mapping[pix] = p.cloud_index;
                    }
                }
            }
        }
        return LPO_OK;
    }

    // segmenter.cpp:321-479
    int ransac(lpo_seg_debug* dbg) {
        constexpr int32_t kMaxRadial = 4;
        constexpr uint32_t kIters = 60;
        ransac_points.clear();
        for (int32_t s = 0; s < slices; ++s) {
            const int32_t off = s * rings;
            for (int32_t r = 0; r < kMaxRadial; ++r) {
                const int32_t c = off + r;
                const float zmin = elevation[c];
                for (const auto& p : grid[c])
                    if (std::fabs(zmin - p.z) < 2.0F * cfg.ground_height_threshold_m)
                        ransac_points.push_back({p.x, p.y, p.z});
            }
        }
        if (dbg) {
            dbg->ransac_candidates = static_cast<uint32_t>(ransac_points.size());
            dbg->ransac_best_inliers = 0;
            dbg->ransac_applied = 0;
            dbg->ransac_plane[0] = 0.0F;
            dbg->ransac_plane[1] = 0.0F;
            dbg->ransac_plane[2] = 1.0F;
            dbg->ransac_plane[3] = 0.0F;
        }
        if (ransac_points.empty()) return LPO_OK;
        // With exactly one candidate the reference's `while (p3 == p2)` redraw
        // loop (segmenter.cpp:383-386) never ends; report it instead.
        if (ransac_points.size() == 1) return LPO_ERR_RANSAC_SINGLE;
        const RP p1{0.0F, 0.0F, -cfg.sensor_height_m};
        const float cos_max = std::cos(std::tan(cfg.road_maximum_slope_m_per_m));
        float a = 0.0F, b = 0.0F, c = 1.0F, d = 0.0F;
        std::mt19937 gen{42};
        std::uniform_int_distribution<uint32_t> dist{0, static_cast<uint32_t>(ransac_points.size()) - 1U};
        uint32_t best = 0U;
        for (uint32_t it = 0; it < kIters; ++it) {
            const uint32_t i2 = dist(gen);
            const auto& p2 = ransac_points[i2];
            uint32_t i3 = dist(gen);
            while (i3 == i2) i3 = dist(gen);
            const auto& p3 = ransac_points[i3];
            if (dbg) {
                dbg->ransac_draws[2 * it] = i2;
                dbg->ransac_draws[2 * it + 1] = i3;
            }
            float nx = ((p2.y - p1.y) * (p3.z - p1.z)) - ((p2.z - p1.z) * (p3.y - p1.y));
            float ny = ((p2.z - p1.z) * (p3.x - p1.x)) - ((p2.x - p1.x) * (p3.z - p1.z));
            float nz = ((p2.x - p1.x) * (p3.y - p1.y)) - ((p2.y - p1.y) * (p3.x - p1.x));
            const float den2 = (nx * nx) + (ny * ny) + (nz * nz);
            if (den2 < 1.0e-5F) continue;
            const float norm = 1.0F / std::sqrt(den2);
            nz *= norm;
            if (std::fabs(nz) < cos_max) continue;
            nx *= norm;
            ny *= norm;
            const float pd = (nx * p1.x) + (ny * p1.y) + (nz * p1.z);
            uint32_t count = 0U;
            for (const auto& q : ransac_points) {
                const float od = std::fabs((nx * q.x) + (ny * q.y) + (nz * q.z) - pd);
                if (od < cfg.ground_height_threshold_m) ++count;
            }
            if (count > best) {
                best = count;
                a = nx;
                b = ny;
                c = nz;
                d = pd;
            }
        }
        if (best > 0) {
            if (c < 0) {
                a = -a;
                b = -b;
                c = -c;
                d = -d;
            }
            for (int32_t s = 0; s < slices; ++s) {
                const int32_t off = s * rings;
                for (int32_t r = 0; r < kMaxRadial; ++r) {
                    for (auto& p : grid[off + r]) {
                        const float sd = (a * p.x) + (b * p.y) + (c * p.z) - d;
                        if (sd < cfg.ground_height_threshold_m) p.label = GROUND;
                    }
                }
            }
        }
        if (dbg) {
            dbg->ransac_best_inliers = best;
            dbg->ransac_applied = best > 0;
            dbg->ransac_plane[0] = a;
            dbg->ransac_plane[1] = b;
            dbg->ransac_plane[2] = c;
            dbg->ransac_plane[3] = d;
        }
        return LPO_OK;
    }

    // segmenter.cpp:481-638
    void jcp(const float* X, const float* Y, const float* Z, lpo_seg_debug* dbg) {
        const int32_t H = cfg.image_height, W = cfg.image_width;
        // cv::split / cv::dilate(red, 5x5 MORPH_RECT, default anchor and
        // border: out-of-image pixels never contribute) / cv::merge.
        std::vector<uint8_t> red(static_cast<size_t>(H) * W), dil(static_cast<size_t>(H) * W);
        for (size_t p = 0; p < red.size(); ++p) red[p] = image[p].r;
        for (int32_t r = 0; r < H; ++r)
            for (int32_t c = 0; c < W; ++c) {
                uint8_t mx = 0;
                for (int32_t dr = -2; dr <= 2; ++dr)
                    for (int32_t dc = -2; dc <= 2; ++dc) {
                        const int32_t rr = r + dr, cc = c + dc;
                        if (rr < 0 || rr >= H || cc < 0 || cc >= W) continue;
                        mx = std::max(mx, red[static_cast<size_t>(rr) * W + cc]);
                    }
                dil[static_cast<size_t>(r) * W + c] = mx;
            }
        for (size_t p = 0; p < red.size(); ++p) image[p].r = dil[p];

        std::deque<int32_t> queue;  // lidar_processing_lib::Queue<Index> FIFO
        for (int32_t r = 0; r < H; ++r)
            for (int32_t c = 0; c < W; ++c) {
                const int32_t p = r * W + c;
                if (image[p] == kIntersectionOrUnknown) {
                    if (mapping[p] != -1) {
                        queue.push_back(p);
                        image[p] = kIntersection;
                    } else {
                        image[p] = kUnknownColour;
                    }
                }
            }
        if (dbg && dbg->queue) {
            int32_t k = 0;
            for (int32_t p : queue) dbg->queue[k++] = p;
        }
        if (dbg) dbg->queue_len = static_cast<int32_t>(queue.size());

        for (int i = 0; i < 24; ++i) {
            m[i] = UNKNOWN;
            u[i] = 0.0F;
            w[i] = 0.0F;
        }
        static constexpr int32_t kOff[24][2] = {
            {-2, -2}, {-2, -1}, {-2, 0}, {-2, 1}, {-2, 2}, {-1, -2}, {-1, -1}, {-1, 0},
            {-1, 1},  {-1, 2},  {0, -2}, {0, -1}, {0, 1},  {0, 2},   {1, -2},  {1, -1},
            {1, 0},   {1, 1},   {1, 2},  {2, -2}, {2, -1}, {2, 0},   {2, 1},   {2, 2}};
        const float thr2 = cfg.kernel_threshold_distance_m * cfg.kernel_threshold_distance_m;
        while (!queue.empty()) {
            const int32_t p = queue.front();
            queue.pop_front();
            const int32_t r = p / W, c = p % W;
            const int32_t ci = mapping[p];
            const float cx = X[ci], cy = Y[ci], cz = Z[ci];
            float sum = 0.0F;
            for (int i = 0; i < 24; ++i) {
                const int32_t nr = r + kOff[i][0], nc = c + kOff[i][1];
                if (nr < 0 || nr >= H || nc < 0 || nc >= W) continue;  // stale slot kept
                const int32_t np = nr * W + nc;
                const int32_t ni = mapping[np];
                if (ni == -1) {
                    u[i] = 0.0F;
                    m[i] = UNKNOWN;
                    continue;
                }
                const float dx = cx - X[ni];
                const float dy = cy - Y[ni];
                const float dzz = cz - Z[ni];
                const float d2 = dx * dx + dy * dy + dzz * dzz;
                if (d2 > thr2) {
                    u[i] = 0.0F;
                    m[i] = UNKNOWN;
                    continue;
                }
                const float dist = std::sqrt(d2);
                u[i] = std::exp(-cfg.amplification_factor * dist);
                sum += u[i];
                const Bgr& col = image[np];
                if (col == kGround)
                    m[i] = GROUND;
                else if (col == kObstacle)
                    m[i] = OBSTACLE;
                else
                    m[i] = UNKNOWN;
            }
            if (std::fabs(sum) > std::numeric_limits<float>::epsilon()) {
                for (int i = 0; i < 24; ++i) w[i] = u[i] / sum;  // Eigen quotient
                float wo = 0.0F, wg = 0.0F;
                for (int i = 0; i < 24; ++i) {
                    if (m[i] == GROUND)
                        wg += w[i];
                    else if (m[i] == OBSTACLE)
                        wo += w[i];
                }
                image[p] = (wo > wg) ? kObstacle : kGround;
            }
        }
    }

    // segmenter.cpp:640-669
    void populate(uint32_t* labels) {
        const int32_t P = cfg.image_height * cfg.image_width;
        for (int32_t p = 0; p < P; ++p) {
            const int32_t ci = mapping[p];
            if (ci == -1) continue;
            if (image[p] == kGround)
                labels[ci] = GROUND;
            else if (image[p] == kObstacle)
                labels[ci] = OBSTACLE;
        }
    }

    int segment(const float* X, const float* Y, const float* Z, const uint16_t* ring, int32_t n,
                uint32_t* labels, lpo_seg_debug* dbg) {
        for (int32_t i = 0; i < n; ++i) labels[i] = UNKNOWN;  // labels.assign
        reset();
        polarGrid(X, Y, Z, ring, n);
        const int st = recm(dbg);
        if (st != LPO_OK) return st;
        const int32_t P = cfg.image_height * cfg.image_width;
        if (dbg && dbg->elevation_map)
            std::memcpy(dbg->elevation_map, elevation.data(), elevation.size() * sizeof(float));
        if (dbg && dbg->mapping) std::memcpy(dbg->mapping, mapping.data(), P * sizeof(int32_t));
        if (dbg && dbg->state_recm)
            for (int32_t p = 0; p < P; ++p)
                dbg->state_recm[p] = image[p] == kGround ? 1 : image[p] == kObstacle ? 2 : 0;
        jcp(X, Y, Z, dbg);
        populate(labels);
        if (dbg && dbg->image)
            for (int32_t p = 0; p < P; ++p) {
                dbg->image[3 * p + 0] = image[p].b;
                dbg->image[3 * p + 1] = image[p].g;
                dbg->image[3 * p + 2] = image[p].r;
            }
        if (dbg && dbg->state_final)
            for (int32_t p = 0; p < P; ++p) {
                if (mapping[p] == -1)
                    dbg->state_final[p] = 0;
                else
                    dbg->state_final[p] = image[p] == kGround ? 1 : image[p] == kObstacle ? 2 : 3;
            }
        return LPO_OK;
    }
};

// Open addressing with linear probing, identity hash, power-of-two buckets,
// rehash above load 0.75, and a hard element capacity whose overflow the
// reference reports with std::overflow_error
// (static_unordered_map.hpp:52-97, 292-321).
class VoxelMap {
  public:
    VoxelMap(uint32_t buckets, uint32_t max_elements) : max_elements_(max_elements) {
        uint32_t b = 1;
        while (b < buckets) b <<= 1;
        keys_.assign(b, kEmpty);
        vals_.assign(b, 0);
    }
    void clear() {
        std::fill(keys_.begin(), keys_.end(), kEmpty);
        count_ = 0;
    }
    // Returns false on capacity overflow.
    bool insert(int32_t key, int32_t value) {
        if (static_cast<float>(count_) / static_cast<float>(keys_.size()) > 0.75F) rehash();
        const size_t mask = keys_.size() - 1;
        size_t h = static_cast<uint32_t>(key) & mask;
        for (;;) {
            if (keys_[h] == kEmpty) {
                if (count_ >= max_elements_) return false;
                keys_[h] = key;
                vals_[h] = value;
                ++count_;
                return true;
            }
            if (keys_[h] == key) {
                vals_[h] = value;
                return true;
            }
            h = (h + 1) & mask;
        }
    }
    int32_t* find(int32_t key) {
        const size_t mask = keys_.size() - 1;
        size_t h = static_cast<uint32_t>(key) & mask;
        for (;;) {
            if (keys_[h] == kEmpty) return nullptr;
            if (keys_[h] == key) return &vals_[h];
            h = (h + 1) & mask;
        }
    }

  private:
    static constexpr int64_t kEmpty = std::numeric_limits<int64_t>::min();
    void rehash() {
        std::vector<int64_t> ok(keys_.size() * 2, kEmpty);
        std::vector<int32_t> ov(keys_.size() * 2, 0);
        const size_t mask = ok.size() - 1;
        for (size_t i = 0; i < keys_.size(); ++i) {
            if (keys_[i] == kEmpty) continue;
            size_t h = static_cast<uint32_t>(keys_[i]) & mask;
            while (ok[h] != kEmpty) h = (h + 1) & mask;
            ok[h] = keys_[i];
            ov[h] = vals_[i];
        }
        keys_.swap(ok);
        vals_.swap(ov);
    }
    std::vector<int64_t> keys_;
    std::vector<int32_t> vals_;
    uint32_t count_ = 0;
    uint32_t max_elements_;
};

struct Clu {
    lpo_cluster_config cfg;
    float res_r, res_az, res_el;
    int32_t num_range = 0, num_az = 0, num_el = 0;
    struct Sph { float range, az, el; };
    struct Key { int32_t r, a, e; };
    std::vector<Sph> sph;
    std::vector<int32_t> vidx;
    std::vector<Key> vkeys;
    VoxelMap map;
    std::deque<Key> queue;
    std::set<int32_t> visited;
    std::vector<uint32_t> counts;
    std::vector<int32_t> cache;

    Clu(const lpo_cluster_config& c, int32_t max_voxels) : cfg(c), map(262144, static_cast<uint32_t>(max_voxels)) {
        // clusterer.hpp:178-185, clusterer.cpp:27-41
        res_r = cfg.voxel_grid_range_resolution_m;
        res_az = cfg.voxel_grid_azimuth_resolution_deg * kDegToRad;
        res_el = cfg.voxel_grid_elevation_resolution_deg * kDegToRad;
    }
    int32_t flat(int32_t r, int32_t a, int32_t e) const { return num_range * (num_az * e + a) + r; }

    // clusterer.cpp:55-100
    void spherical(const float* X, const float* Y, const float* Z, int32_t n) {
        sph.clear();
        if (n == 0) return;
        float mr = std::numeric_limits<float>::lowest();
        float ma = std::numeric_limits<float>::lowest();
        float me = std::numeric_limits<float>::lowest();
        for (int32_t i = 0; i < n; ++i) {
            float az = std::atan2(Y[i], X[i]);
            az = (az < 0) ? (az + kTwoPi) : az;
            ma = std::max(ma, az);
            const float dsq = X[i] * X[i] + Y[i] * Y[i];
            const float d = std::sqrt(dsq);
            const float range = std::sqrt(dsq + Z[i] * Z[i]);
            mr = std::max(mr, range);
            const float el = std::atan(Z[i] / d) + static_cast<float>(M_PI_2);
            me = std::max(me, el);
            sph.push_back({range, az, el});
        }
        num_range = static_cast<int32_t>(std::ceil(mr / res_r) + 1);
        num_az = static_cast<int32_t>(std::ceil(ma / res_az) + 1);
        num_el = static_cast<int32_t>(std::ceil(me / res_el) + 1);
    }

    // clusterer.cpp:102-120
    bool build() {
        map.clear();
        vidx.clear();
        vkeys.clear();
        for (const auto& p : sph) {
            const auto ri = static_cast<int32_t>(p.range / res_r);
            const auto ai = static_cast<int32_t>(p.az / res_az);
            const auto ei = static_cast<int32_t>(p.el / res_el);
            const int32_t v = flat(ri, ai, ei);
            if (!map.insert(v, -1)) return false;
            vidx.push_back(v);
            vkeys.push_back({ri, ai, ei});
        }
        return true;
    }

    // clusterer.cpp:122-193
    void components(int32_t* labels) {
        static constexpr int32_t kOff[26][3] = {
            {-1, -1, -1}, {-1, 0, -1}, {-1, 1, -1}, {0, -1, -1}, {0, 0, -1}, {0, 1, -1}, {1, -1, -1},
            {1, 0, -1},   {1, 1, -1},  {-1, -1, 0}, {-1, 0, 0},  {-1, 1, 0}, {0, -1, 0}, {0, 1, 0},
            {1, -1, 0},   {1, 0, 0},   {1, 1, 0},   {-1, -1, 1}, {-1, 0, 1}, {-1, 1, 1}, {0, -1, 1},
            {0, 0, 1},    {0, 1, 1},   {1, -1, 1},  {1, 0, 1},   {1, 1, 1}};
        const int32_t n = static_cast<int32_t>(sph.size());
        for (int32_t i = 0; i < n; ++i) labels[i] = -1;
        int32_t label = 0;
        for (int32_t i = 0; i < n; ++i) {
            if (*map.find(vidx[i]) != -1) continue;
            visited.clear();
            visited.insert(vidx[i]);
            queue.push_back(vkeys[i]);
            while (!queue.empty()) {
                const Key k = queue.front();
                queue.pop_front();
                *map.find(flat(k.r, k.a, k.e)) = label;
                for (const auto& o : kOff) {
                    const int32_t rr = k.r + o[0];
                    const int32_t ee = k.e + o[2];
                    if (rr < 0 || rr >= num_range || ee < 0 || ee >= num_el) continue;
                    int32_t aa = k.a + o[1];
                    if (aa < 0)
                        aa += num_az;
                    else if (aa >= num_az)
                        aa -= num_az;
                    const int32_t nv = flat(rr, aa, ee);
                    if (map.find(nv) != nullptr && visited.find(nv) == visited.end()) {
                        queue.push_back({rr, aa, ee});
                        visited.insert(nv);
                    }
                }
            }
            ++label;
        }
        for (int32_t i = 0; i < n; ++i) labels[i] = *map.find(vidx[i]);
    }

    // clusterer.cpp:195-239
    void removeSmall(int32_t* labels, int32_t n) {
        if (n == 0) return;
        const int32_t mx = *std::max_element(labels, labels + n);
        if (mx == -1) return;
        counts.assign(static_cast<size_t>(mx) + 1, 0U);
        for (int32_t i = 0; i < n; ++i)
            if (labels[i] != -1) ++counts[labels[i]];
        cache.assign(static_cast<size_t>(mx) + 1, -1);
        int32_t cur = 0;
        for (int32_t l = 0; l <= mx; ++l)
            if (counts[l] >= cfg.min_cluster_size) cache[l] = cur++;
        for (int32_t i = 0; i < n; ++i) labels[i] = cache[labels[i]];
    }

    int cluster(const float* X, const float* Y, const float* Z, int32_t n, int32_t* labels) {
        spherical(X, Y, Z, n);
        if (!build()) return LPO_ERR_VOXEL_OVERFLOW;
        components(labels);
        removeSmall(labels, n);
        return LPO_OK;
    }
};

// polygonizer.cpp:33-91 (PointXY of doubles).
int32_t convexHull(const double* X, const double* Y, int32_t n, std::vector<int32_t>& sorted,
                   int32_t* out) {
    if (n < 3) {
        for (int32_t i = 0; i < n; ++i) out[i] = i;
        return n;
    }
    sorted.resize(n);
    std::iota(sorted.begin(), sorted.end(), 0);
    std::sort(sorted.begin(), sorted.end(), [X, Y](const int32_t i, const int32_t j) {
        return X[i] < X[j] || (std::fabs(X[i] - X[j]) < std::numeric_limits<double>::epsilon() && Y[i] < Y[j]);
    });
    auto ccw_or_colinear = [X, Y](int32_t a, int32_t b, int32_t c) {
        return (X[b] - X[a]) * (Y[c] - Y[a]) - (Y[b] - Y[a]) * (X[c] - X[a]) <= 0;
    };
    int32_t k = 0;
    for (int32_t i = 0; i < n; ++i) {
        while (k > 1 && ccw_or_colinear(out[k - 2], out[k - 1], sorted[i])) --k;
        out[k++] = sorted[i];
    }
    for (int32_t i = n - 2, t = k + 1; i >= 0; --i) {
        while (k >= t && ccw_or_colinear(out[k - 2], out[k - 1], sorted[i])) --k;
        out[k++] = sorted[i];
    }
    return k - 1;
}

}  // namespace

extern "C" {

void lpo_seg_config_default(lpo_seg_config* c) {
    c->elevation_up_deg = 2.0F;
    c->elevation_down_deg = -24.8F;
    c->image_width = 2048;
    c->image_height = 64;
    c->assume_unorganized_cloud = 0;
    c->grid_radial_spacing_m = 2.0F;
    c->grid_slice_resolution_deg = 1.0F;
    c->ground_height_threshold_m = 0.2F;
    c->road_maximum_slope_m_per_m = 0.2F;
    c->min_distance_m = 2.0F;
    c->max_distance_m = 100.0F;
    c->sensor_height_m = 1.73F;
    c->kernel_threshold_distance_m = 1.0F;
    c->amplification_factor = 5.0F;
    c->z_min_m = -3.0F;
    c->z_max_m = 4.0F;
}

// processor.param.yaml:32-35 (the shipped values; the header default for the
// elevation resolution is 1.5 deg, clusterer.hpp:65).
void lpo_cluster_config_default(lpo_cluster_config* c) {
    c->voxel_grid_range_resolution_m = 0.4F;
    c->voxel_grid_azimuth_resolution_deg = 1.0F;
    c->voxel_grid_elevation_resolution_deg = 3.0F;
    c->min_cluster_size = 3;
}

struct lpo_processor {
    Seg seg;
    Clu clu;
    std::vector<float> ox, oy, oz;
    std::vector<double> gx, gy, px, py;
    std::vector<int32_t> sorted, out;
    std::vector<float> kx, ky, kz;
    std::vector<uint16_t> kring;
    lpo_processor(const lpo_seg_config& s, const lpo_cluster_config& c, int32_t max_voxels)
        : seg(s), clu(c, max_voxels) {}
};

lpo_processor* lpo_processor_create(const lpo_seg_config* scfg, const lpo_cluster_config* ccfg,
                                    int32_t max_voxels) {
    if (scfg == nullptr || ccfg == nullptr || scfg->image_width <= 0 || scfg->image_height <= 0) return nullptr;
    auto* p = new lpo_processor(*scfg, *ccfg, max_voxels);
    if (p->seg.rings < 4) {
        delete p;
        return nullptr;
    }
    return p;
}

void lpo_processor_destroy(lpo_processor* p) { delete p; }

int lpo_processor_segment(lpo_processor* p, const float* x, const float* y, const float* z,
                          const uint16_t* ring, int32_t n, uint32_t* labels, lpo_seg_debug* dbg) {
    if (p == nullptr || n < 0) return LPO_ERR_INVALID;
    return p->seg.segment(x, y, z, ring, n, labels, dbg);
}

int lpo_processor_cluster(lpo_processor* p, const float* x, const float* y, const float* z, int32_t n,
                          int32_t* labels) {
    if (p == nullptr || n < 0) return LPO_ERR_INVALID;
    return p->clu.cluster(x, y, z, n, labels);
}

static int frame_impl(lpo_processor* p, const float* x, const float* y, const float* z, const uint16_t* ring,
                      int32_t n, uint32_t* seg_labels, int32_t* obstacle_count, int32_t* cluster_labels,
                      int32_t* num_clusters, int32_t* hull_offsets, int32_t hull_offsets_cap, int32_t* hull_indices,
                      int32_t hull_cap, lpo_bounding_box* boxes) {
    if (p == nullptr || n < 0) return LPO_ERR_INVALID;
    int st = p->seg.segment(x, y, z, ring, n, seg_labels, nullptr);
    if (st != LPO_OK) return st;
    // processor.cpp:562-579: obstacle cloud in input order.
    auto& ox = p->ox;
    auto& oy = p->oy;
    auto& oz = p->oz;
    ox.clear();
    oy.clear();
    oz.clear();
    for (int32_t i = 0; i < n; ++i)
        if (seg_labels[i] == OBSTACLE) {
            ox.push_back(x[i]);
            oy.push_back(y[i]);
            oz.push_back(z[i]);
        }
    const int32_t no = static_cast<int32_t>(ox.size());
    *obstacle_count = no;
    st = p->clu.cluster(ox.data(), oy.data(), oz.data(), no, cluster_labels);
    if (st != LPO_OK) return st;
    // processor.cpp:612-670: per label, an O(N_obs) gather then the hull.
    int32_t mx = -1;
    for (int32_t i = 0; i < no; ++i) mx = std::max(mx, cluster_labels[i]);
    *num_clusters = mx + 1;
    if (mx + 2 > hull_offsets_cap) return LPO_ERR_INVALID;
    int32_t total = 0;
    hull_offsets[0] = 0;
    for (int32_t l = 0; l <= mx; ++l) {
        p->gx.clear();
        p->gy.clear();
        for (int32_t i = 0; i < no; ++i)
            if (cluster_labels[i] == l) {
                p->gx.push_back(ox[i]);
                p->gy.push_back(oy[i]);
            }
        p->out.resize(2 * p->gx.size() + 1);
        const int32_t k = convexHull(p->gx.data(), p->gy.data(), static_cast<int32_t>(p->gx.size()),
                                     p->sorted, p->out.data());
        if (total + k > hull_cap) return LPO_ERR_INVALID;
        std::memcpy(hull_indices + total, p->out.data(), sizeof(int32_t) * k);
        total += k;
        hull_offsets[l + 1] = total;
        if (boxes) {
            // processor.cpp:663-704: the polygon of the hull vertices, then
            // boundingBoxRotatingCalipers on it.
            p->px.clear();
            p->py.clear();
            for (int32_t v = 0; v < k; ++v) {
                p->px.push_back(p->gx[p->out[v]]);
                p->py.push_back(p->gy[p->out[v]]);
            }
            const int bs = lpo_obb_rotating_calipers(p->px.data(), p->py.data(), k, &boxes[l]);
            if (bs != LPO_OK) return bs;
        }
    }
    return LPO_OK;
}

int lpo_processor_frame(lpo_processor* p, const float* x, const float* y, const float* z,
                        const uint16_t* ring, int32_t n, uint32_t* seg_labels, int32_t* obstacle_count,
                        int32_t* cluster_labels, int32_t* num_clusters, int32_t* hull_offsets,
                        int32_t hull_offsets_cap, int32_t* hull_indices, int32_t hull_cap) {
    return frame_impl(p, x, y, z, ring, n, seg_labels, obstacle_count, cluster_labels, num_clusters, hull_offsets,
                      hull_offsets_cap, hull_indices, hull_cap, nullptr);
}

int lpo_full_frame(lpo_processor* p, lpo_noise_remover* nr, const lpo_noise_config* ncfg, const float* x,
                   const float* y, const float* z, const uint16_t* ring, int32_t n, uint8_t* noise_labels,
                   int32_t* kept, uint32_t* seg_labels, int32_t* obstacle_count, int32_t* cluster_labels,
                   int32_t* num_clusters, int32_t* hull_offsets, int32_t hull_offsets_cap, int32_t* hull_indices,
                   int32_t hull_cap, lpo_bounding_box* boxes) {
    if (p == nullptr || ncfg == nullptr || n < 0) return LPO_ERR_INVALID;
    const int st = nr ? lpo_noise_remover_filter(nr, ncfg, x, y, z, n, noise_labels)
                      : lpo_noise_definition(ncfg, x, y, z, n, noise_labels, nullptr);
    if (st != LPO_OK) return st;
    p->kx.clear();
    p->ky.clear();
    p->kz.clear();
    p->kring.clear();
    for (int32_t i = 0; i < n; ++i)
        if (noise_labels[i] == 0) {
            p->kx.push_back(x[i]);
            p->ky.push_back(y[i]);
            p->kz.push_back(z[i]);
            if (ring) p->kring.push_back(ring[i]);
        }
    *kept = static_cast<int32_t>(p->kx.size());
    return frame_impl(p, p->kx.data(), p->ky.data(), p->kz.data(), ring ? p->kring.data() : nullptr, *kept,
                      seg_labels, obstacle_count, cluster_labels, num_clusters, hull_offsets, hull_offsets_cap,
                      hull_indices, hull_cap, boxes);
}

int lpo_segment(const lpo_seg_config* cfg, const float* x, const float* y, const float* z,
                const uint16_t* ring, int32_t n, uint32_t* labels, lpo_seg_debug* dbg) {
    lpo_cluster_config cc;
    lpo_cluster_config_default(&cc);
    lpo_processor* p = lpo_processor_create(cfg, &cc, 200000);
    if (p == nullptr) return LPO_ERR_INVALID;
    const int st = lpo_processor_segment(p, x, y, z, ring, n, labels, dbg);
    lpo_processor_destroy(p);
    return st;
}

int lpo_cluster(const lpo_cluster_config* cfg, const float* x, const float* y, const float* z,
                int32_t n, int32_t max_voxels, int32_t* labels) {
    if (cfg == nullptr || n < 0) return LPO_ERR_INVALID;
    Clu c(*cfg, max_voxels);
    return c.cluster(x, y, z, n, labels);
}

int32_t lpo_convex_hull(const double* x, const double* y, int32_t n, int32_t* out) {
    std::vector<int32_t> sorted;
    return convexHull(x, y, n, sorted, out);
}

int lpo_process_frame(const lpo_seg_config* scfg, const lpo_cluster_config* ccfg, const float* x,
                      const float* y, const float* z, const uint16_t* ring, int32_t n,
                      int32_t max_voxels, uint32_t* seg_labels, int32_t* obstacle_count,
                      int32_t* cluster_labels, int32_t* num_clusters, int32_t* hull_offsets,
                      int32_t hull_offsets_cap, int32_t* hull_indices, int32_t hull_cap) {
    lpo_processor* p = lpo_processor_create(scfg, ccfg, max_voxels);
    if (p == nullptr) return LPO_ERR_INVALID;
    const int st = lpo_processor_frame(p, x, y, z, ring, n, seg_labels, obstacle_count, cluster_labels,
                                       num_clusters, hull_offsets, hull_offsets_cap, hull_indices, hull_cap);
    lpo_processor_destroy(p);
    return st;
}

// dataloader.cpp:68-137
void lpo_add_ring_info(const float* x, const float* y, int32_t n, uint16_t* ring) {
    if (n <= 0) return;
    int prev_q = 0;
    uint16_t r = 63;
    for (int32_t i = 0; i < n; ++i) {
        float az = std::atan2(y[i], x[i]);
        az = (az < 0) ? (az + 2.0F * static_cast<float>(M_PI)) : az;
        int q;
        if (az < static_cast<float>(M_PI_2))
            q = 0;
        else if (az < static_cast<float>(M_PI))
            q = 1;
        else if (az < 1.5F * static_cast<float>(M_PI))
            q = 2;
        else
            q = 3;
        if (q == 0 && prev_q == 3 && r > 0U) --r;
        prev_q = q;
        ring[i] = r;
    }
}

float lpo_atan2_approx(float y, float x) { return atan2Approx(y, x); }

void lpo_std_uniform_draws(uint32_t n, int32_t count, uint32_t* out) {
    std::mt19937 gen{42};
    std::uniform_int_distribution<uint32_t> dist{0, n - 1U};
    for (int32_t i = 0; i < count; ++i) out[i] = dist(gen);
}

void lpo_mt19937_raw(int32_t count, uint32_t* out) {
    std::mt19937 gen{42};
    for (int32_t i = 0; i < count; ++i) out[i] = static_cast<uint32_t>(gen());
}

// Host libm over float bit patterns [lo, lo+count): which 0 expf, 1 atanf.
void lpo_host_math_range(int32_t which, uint32_t lo, uint32_t count, uint32_t* out_bits) {
    for (uint32_t i = 0; i < count; ++i) {
        uint32_t b = lo + i;
        float x;
        std::memcpy(&x, &b, 4);
        const float r = which == 0 ? std::exp(x) : std::atan(x);
        std::memcpy(&out_bits[i], &r, 4);
    }
}

float lpo_host_expf(float x) { return std::exp(x); }
float lpo_host_atanf(float x) { return std::atan(x); }
float lpo_host_atan2f(float y, float x) { return std::atan2(y, x); }

}  // extern "C"
