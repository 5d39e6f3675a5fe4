// lidar_processing_lib_amd.hpp -- the reference's Segmenter / Clusterer /
// Polygonizer interface (namespace lidar_processing_lib) implemented on the
// MI355X C ABI of include/lp_amd.h.
//
// Same names, configuration fields, defaults and call signatures as
//   lidar_processing_lib/include/lidar_processing_lib/segmenter.hpp:69-169
//   lidar_processing_lib/include/lidar_processing_lib/clusterer.hpp:52-94
//   lidar_processing_lib/include/lidar_processing_lib/polygonizer.hpp:62-106
// so the processor's calls (src/processor/src/processor.cpp:556, 599, 663)
// compile unchanged against a pcl::PointCloud<PointT>: the templates take any
// cloud type exposing `points` (a contiguous vector of structs with float
// x, y, z and, for organized clouds, an integral `ring`).  Errors throw the
// reference's exception types: capacity overflow of the voxel map ->
// std::overflow_error (static_unordered_map.hpp:79-81); anything else ->
// std::runtime_error.
#pragma once

#include <algorithm>
#include <cstddef>
#include <array>
#include <cstdint>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "lp_amd.h"

namespace lidar_processing_lib {

enum class Label : std::uint32_t { UNKNOWN = 0, GROUND, OBSTACLE };

using ClusterLabel = std::int32_t;

struct SegmenterConfiguration {
    float elevation_up_deg = 2.0F;
    float elevation_down_deg = -24.8F;
    std::int32_t image_width = 2048;
    std::int32_t image_height = 64;
    bool assume_unorganized_cloud = false;
    float grid_radial_spacing_m = 2.0F;
    float grid_slice_resolution_deg = 1.0F;
    float ground_height_threshold_m = 0.2F;
    float road_maximum_slope_m_per_m = 0.2F;
    float min_distance_m = 2.0F;
    float max_distance_m = 100.0F;
    float sensor_height_m = 1.73F;
    float kernel_threshold_distance_m = 1.0F;
    float amplification_factor = 5.0F;
    float z_min_m = -3.0F;
    float z_max_m = 4.0F;
    bool display_recm_with_low_confidence_points = false;  // accepted, ignored (OpenCV window)
};

struct ClustererConfiguration {
    float voxel_grid_range_resolution_m = 0.4F;
    float voxel_grid_azimuth_resolution_deg = 1.0F;
    float voxel_grid_elevation_resolution_deg = 1.5F;
    std::uint32_t min_cluster_size = 3;
};

struct PointXY {
    double x;
    double y;
};

// BoundingBox (polygonizer.hpp:75-81).
struct BoundingBox {
    std::array<PointXY, 4> corners;
    float area;
    float angle_rad;
    bool is_valid;
};

// NoiseRemoverConfiguration / NoiseRemoverLabel (noise_remover.hpp:35-54).
enum class NoiseRemoverLabel : std::uint8_t { VALID = 0, NOISE = 1 };

struct NoiseRemoverConfiguration {
    float radius_multiplier_m_per_m = 0.02F;
    float min_search_radius_m = 0.1F;
    std::uint32_t min_neighbours = 4U;
};

namespace detail {

template <typename P, typename = void>
struct has_ring : std::false_type {};
template <typename P>
struct has_ring<P, std::void_t<decltype(std::declval<P>().ring)>> : std::true_type {};

[[noreturn]] inline void raise(int code) {
    const std::string msg = std::string("lidar_processing_lib (MI355X): ") + lp_last_error();
    if (code == LP_ERR_VOXEL_OVERFLOW) throw std::overflow_error(msg);
    throw std::runtime_error(msg);
}

inline void check(int code) {
    if (code != LP_OK) raise(code);
}

// One HIP context per object, like the reference's per-object scratch.  The
// reference reserves 200 000 points (segmenter.cpp:62, clusterer.cpp:33-38,
// noise_remover.cpp:32-36) as a hint and keeps working past it, so a call
// that needs more points or a larger image than the context holds recreates
// it with room for that call (device memory is allocated only then).
class Context {
  public:
    explicit Context(std::int32_t max_points = 200000, std::int32_t max_h = 128, std::int32_t max_w = 2048) {
        create(max_points, max_h, max_w);
    }
    ~Context() { lp_context_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    lp_context* get() const noexcept { return ctx_; }

    void ensure(std::size_t points, std::int32_t h = 0, std::int32_t w = 0) {
        if (points <= static_cast<std::size_t>(points_) && h <= h_ && w <= w_) return;
        if (points > static_cast<std::size_t>(INT32_MAX)) throw std::length_error("cloud too large");
        const auto n = std::max(points_, static_cast<std::int32_t>(points));
        const auto hh = std::max(h_, h);
        const auto ww = std::max(w_, w);
        lp_context_destroy(ctx_);
        ctx_ = nullptr;
        create(n, hh, ww);
    }

  private:
    void create(std::int32_t max_points, std::int32_t max_h, std::int32_t max_w) {
        lp_capacity cap;
        lp_capacity_default(&cap);
        cap.max_points_per_frame = max_points;
        cap.max_image_height = max_h;
        cap.max_image_width = max_w;
        check(lp_context_create(&cap, &ctx_));
        points_ = max_points;
        h_ = max_h;
        w_ = max_w;
    }

    lp_context* ctx_ = nullptr;
    std::int32_t points_ = 0;
    std::int32_t h_ = 0;
    std::int32_t w_ = 0;
};

template <typename CloudT>
lp_cloud make_cloud(const CloudT& cloud, bool use_ring) {
    using PointT = typename std::decay_t<decltype(cloud.points)>::value_type;
    lp_cloud c{};
    c.num_points = static_cast<std::int32_t>(cloud.points.size());
    c.xyz_stride = static_cast<std::int32_t>(sizeof(PointT));
    c.ring_stride = static_cast<std::int32_t>(sizeof(PointT));
    c.on_device = 0;
    if (c.num_points == 0) return c;
    const PointT& p0 = cloud.points[0];
    c.x = &p0.x;
    c.y = &p0.y;
    c.z = &p0.z;
    if constexpr (has_ring<PointT>::value) {
        static_assert(sizeof(p0.ring) == 2, "ring must be a 16-bit field");
        c.ring = use_ring ? &p0.ring : nullptr;
    } else {
        c.ring = nullptr;
    }
    return c;
}

}  // namespace detail

// Zero-copy view of a sensor_msgs::msg::PointCloud2 (or any type with the
// same members: fields[i].name / .offset, data, width, height, point_step,
// row_step) as segmenter input, with Processor::run's field dispatch
// (processor.cpp:494-549): x, y, z [, intensity] [, ring] in that order or
// std::runtime_error("Unsupported point format ...").  Rows padded beyond
// width * point_step are packed into the view's own buffer first.
class CloudView {
  public:
    template <typename MsgT>
    explicit CloudView(const MsgT& msg) {
        std::vector<lp_point_field> f;
        names_.reserve(msg.fields.size());
        for (const auto& fld : msg.fields) names_.push_back(std::string(fld.name));
        for (std::size_t k = 0; k < msg.fields.size(); ++k)
            f.push_back({names_[k].c_str(), static_cast<std::uint32_t>(msg.fields[k].offset)});
        const std::uint8_t* data = msg.data.data();
        std::uint32_t row_step = msg.row_step;
        const std::size_t packed_row = static_cast<std::size_t>(msg.width) * msg.point_step;
        if (msg.height > 1 && msg.row_step != packed_row) {
            packed_.resize(packed_row * msg.height);
            for (std::uint32_t r = 0; r < msg.height; ++r)
                std::copy(data + static_cast<std::size_t>(r) * msg.row_step,
                          data + static_cast<std::size_t>(r) * msg.row_step + packed_row,
                          packed_.begin() + static_cast<std::ptrdiff_t>(r * packed_row));
            data = packed_.data();
            row_step = static_cast<std::uint32_t>(packed_row);
        }
        detail::check(lp_cloud_from_pointcloud2(data, msg.width, msg.height, msg.point_step, row_step, f.data(),
                                                static_cast<std::int32_t>(f.size()), &cloud_));
    }

    const lp_cloud& cloud() const noexcept { return cloud_; }
    std::size_t size() const noexcept { return static_cast<std::size_t>(cloud_.num_points); }

  private:
    lp_cloud cloud_{};
    std::vector<std::string> names_;
    std::vector<std::uint8_t> packed_;
};

// The label image of Segmenter::image() (a CV_8UC3 cv::Mat in the reference,
// segmenter.hpp:163-166): rows x cols BGR pixels, row-major.
struct LabelImage {
    std::int32_t rows = 0;
    std::int32_t cols = 0;
    std::vector<std::uint8_t> bgr;
    const std::uint8_t* at(std::int32_t row, std::int32_t col) const { return bgr.data() + 3 * (row * cols + col); }
};

class Segmenter {
  public:
    Segmenter() : ctx_(new detail::Context()) {}

    void config(const SegmenterConfiguration& config) { config_ = config; }

    const SegmenterConfiguration& config() const noexcept { return config_; }

    template <typename CloudT>
    void segment(const CloudT& cloud, std::vector<Label>& labels) {
        lp_segmenter_config c = to_c();
        lp_cloud in = detail::make_cloud(cloud, !config_.assume_unorganized_cloud);
        labels.assign(cloud.points.size(), Label::UNKNOWN);
        ctx_->ensure(cloud.points.size(), config_.image_height, config_.image_width);
        detail::check(lp_segment(ctx_->get(), &c, &in, reinterpret_cast<std::uint32_t*>(labels.data()), &ransac_));
        image_valid_ = false;
    }

    // A PointCloud2 message as it arrives (Processor::run without the
    // conversion to a pcl::PointCloud).  A ring field is used unless the
    // configuration assumes an unorganized cloud.
    void segment(const CloudView& view, std::vector<Label>& labels) {
        lp_segmenter_config c = to_c();
        lp_cloud in = view.cloud();
        if (config_.assume_unorganized_cloud) in.ring = nullptr;
        labels.assign(view.size(), Label::UNKNOWN);
        ctx_->ensure(view.size(), config_.image_height, config_.image_width);
        detail::check(lp_segment(ctx_->get(), &c, &in, reinterpret_cast<std::uint32_t*>(labels.data()), &ransac_));
        image_valid_ = false;
    }

    // The label image of the last segment() call, built on the device when
    // first asked for.
    const LabelImage& image() const {
        if (!image_valid_) {
            image_.rows = config_.image_height;
            image_.cols = config_.image_width;
            image_.bgr.resize(3 * static_cast<std::size_t>(image_.rows) * image_.cols);
            const int got = lp_frame_image(ctx_->get(), 0, image_.bgr.data(), static_cast<std::int32_t>(image_.bgr.size()));
            if (got < 0) detail::check(got);
            image_valid_ = true;
        }
        return image_;
    }

    // The RANSAC plane the reference keeps in locals (segmenter.cpp:363-374).
    const lp_ransac_result& ransac() const noexcept { return ransac_; }

  private:
    lp_segmenter_config to_c() const {
        lp_segmenter_config c;
        c.elevation_up_deg = config_.elevation_up_deg;
        c.elevation_down_deg = config_.elevation_down_deg;
        c.image_width = config_.image_width;
        c.image_height = config_.image_height;
        c.assume_unorganized_cloud = config_.assume_unorganized_cloud ? 1 : 0;
        c.grid_radial_spacing_m = config_.grid_radial_spacing_m;
        c.grid_slice_resolution_deg = config_.grid_slice_resolution_deg;
        c.ground_height_threshold_m = config_.ground_height_threshold_m;
        c.road_maximum_slope_m_per_m = config_.road_maximum_slope_m_per_m;
        c.min_distance_m = config_.min_distance_m;
        c.max_distance_m = config_.max_distance_m;
        c.sensor_height_m = config_.sensor_height_m;
        c.kernel_threshold_distance_m = config_.kernel_threshold_distance_m;
        c.amplification_factor = config_.amplification_factor;
        c.z_min_m = config_.z_min_m;
        c.z_max_m = config_.z_max_m;
        return c;
    }

    SegmenterConfiguration config_;
    std::unique_ptr<detail::Context> ctx_;
    lp_ransac_result ransac_{};
    mutable LabelImage image_;
    mutable bool image_valid_ = false;
};

class Clusterer {
  public:
    static constexpr std::int32_t INVALID_LABEL = -1;

    Clusterer() : ctx_(new detail::Context()) {}

    void config(const ClustererConfiguration& config) { config_ = config; }
    const ClustererConfiguration& config() const noexcept { return config_; }

    template <typename CloudT>
    void cluster(const CloudT& cloud, std::vector<ClusterLabel>& labels) {
        lp_clusterer_config c;
        c.voxel_grid_range_resolution_m = config_.voxel_grid_range_resolution_m;
        c.voxel_grid_azimuth_resolution_deg = config_.voxel_grid_azimuth_resolution_deg;
        c.voxel_grid_elevation_resolution_deg = config_.voxel_grid_elevation_resolution_deg;
        c.min_cluster_size = config_.min_cluster_size;
        lp_cloud in = detail::make_cloud(cloud, false);
        labels.assign(cloud.points.size(), INVALID_LABEL);
        ctx_->ensure(cloud.points.size());
        detail::check(lp_cluster(ctx_->get(), &c, &in, labels.data()));
    }

  private:
    ClustererConfiguration config_;
    std::unique_ptr<detail::Context> ctx_;
};

class Polygonizer {
  public:
    Polygonizer() : ctx_(new detail::Context()) {}

    // Counter-clockwise hull indices into `points` (polygonizer.cpp:33-91).
    template <typename PointT>
    void convexHull(const std::vector<PointT>& points, std::vector<std::int32_t>& indices) {
        std::vector<std::vector<std::int32_t>> out;
        convexHulls(std::vector<const std::vector<PointT>*>{&points}, out);
        indices = std::move(out[0]);
    }

    // Many point lists in one call (one launch sequence for all clusters).
    template <typename PointT>
    void convexHulls(const std::vector<const std::vector<PointT>*>& lists,
                     std::vector<std::vector<std::int32_t>>& indices) {
        std::vector<double> x, y;
        std::vector<std::int32_t> offsets(1, 0);
        for (const auto* l : lists) {
            for (const auto& p : *l) {
                x.push_back(static_cast<double>(p.x));
                y.push_back(static_cast<double>(p.y));
            }
            offsets.push_back(static_cast<std::int32_t>(x.size()));
        }
        std::vector<std::int32_t> hoff(lists.size() + 1);
        std::vector<std::int32_t> hidx(x.size() + 4);
        ctx_->ensure(x.size());
        detail::check(lp_convex_hulls(ctx_->get(), x.data(), y.data(), offsets.data(),
                                      static_cast<std::int32_t>(lists.size()), hoff.data(), hidx.data(),
                                      static_cast<std::int32_t>(hidx.size())));
        indices.resize(lists.size());
        for (std::size_t k = 0; k < lists.size(); ++k)
            indices[k].assign(hidx.begin() + hoff[k], hidx.begin() + hoff[k + 1]);
    }

    // Oriented box of a convex polygon, vertices counter-clockwise
    // (polygonizer.cpp:163-278).
    BoundingBox boundingBoxRotatingCalipers(const std::vector<PointXY>& convex_hull_points) {
        std::vector<BoundingBox> out;
        boundingBoxesRotatingCalipers(std::vector<const std::vector<PointXY>*>{&convex_hull_points}, out);
        return out[0];
    }

    // Many polygons in one call.
    void boundingBoxesRotatingCalipers(const std::vector<const std::vector<PointXY>*>& polygons,
                                       std::vector<BoundingBox>& boxes) {
        boxes_of(polygons, boxes, lp_oriented_boxes);
    }

    // Principal-component box of a point list (polygonizer.cpp:280-362).
    BoundingBox boundingBoxPrincipalComponentAnalysis(const std::vector<PointXY>& convex_hull_points) {
        std::vector<BoundingBox> out;
        boxes_of(std::vector<const std::vector<PointXY>*>{&convex_hull_points}, out, lp_pca_boxes);
        return out[0];
    }

    void boundingBoxesPrincipalComponentAnalysis(const std::vector<const std::vector<PointXY>*>& polygons,
                                                 std::vector<BoundingBox>& boxes) {
        boxes_of(polygons, boxes, lp_pca_boxes);
    }

  private:
    using BoxFn = int (*)(lp_context*, const double*, const double*, const std::int32_t*, std::int32_t,
                          lp_bounding_box*);

    void boxes_of(const std::vector<const std::vector<PointXY>*>& polygons, std::vector<BoundingBox>& boxes,
                  BoxFn fn) {
        std::vector<double> x, y;
        std::vector<std::int32_t> offsets(1, 0);
        for (const auto* poly : polygons) {
            for (const auto& p : *poly) {
                x.push_back(p.x);
                y.push_back(p.y);
            }
            offsets.push_back(static_cast<std::int32_t>(x.size()));
        }
        std::vector<lp_bounding_box> raw(polygons.size());
        ctx_->ensure(std::max(x.size(), polygons.size()));
        detail::check(fn(ctx_->get(), x.data(), y.data(), offsets.data(), static_cast<std::int32_t>(polygons.size()),
                         raw.data()));
        boxes.resize(polygons.size());
        for (std::size_t k = 0; k < polygons.size(); ++k) {
            for (int c = 0; c < 4; ++c) boxes[k].corners[c] = {raw[k].corners[2 * c], raw[k].corners[2 * c + 1]};
            boxes[k].area = raw[k].area;
            boxes[k].angle_rad = raw[k].angle_rad;
            boxes[k].is_valid = raw[k].is_valid != 0;
        }
    }

    std::unique_ptr<detail::Context> ctx_;
};

// NoiseRemover::filter (noise_remover.cpp:38-68) over an array of xyz points
// (KDTree<float, 3>::PointT, i.e. std::array<float, 3>).
class NoiseRemover {
  public:
    using PointT = std::array<float, 3>;

    NoiseRemover() : ctx_(new detail::Context()) {}

    void reserve(std::uint32_t max_pts) { ctx_->ensure(max_pts); }

    void config(const NoiseRemoverConfiguration& config) { config_ = config; }
    const NoiseRemoverConfiguration& config() const noexcept { return config_; }

    void filter(const std::vector<PointT>& points, std::vector<NoiseRemoverLabel>& labels) {
        labels.assign(points.size(), NoiseRemoverLabel::VALID);
        if (points.empty()) return;
        reserve(static_cast<std::uint32_t>(points.size()));
        lp_noise_config c{config_.radius_multiplier_m_per_m, config_.min_search_radius_m, config_.min_neighbours};
        lp_cloud cloud{};
        cloud.x = &points[0][0];
        cloud.y = &points[0][1];
        cloud.z = &points[0][2];
        cloud.xyz_stride = static_cast<std::int32_t>(sizeof(PointT));
        cloud.num_points = static_cast<std::int32_t>(points.size());
        static_assert(sizeof(NoiseRemoverLabel) == 1, "labels are bytes");
        detail::check(lp_remove_noise(ctx_->get(), &c, &cloud, reinterpret_cast<std::uint8_t*>(labels.data())));
    }

  private:
    NoiseRemoverConfiguration config_;
    std::unique_ptr<detail::Context> ctx_;
};

}  // namespace lidar_processing_lib
