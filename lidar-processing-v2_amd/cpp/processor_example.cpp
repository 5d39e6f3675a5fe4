// processor_example.cpp -- the reference processor's per-frame calls
// (src/processor/src/processor.cpp:556-670) written against the drop-in
// lidar_processing_lib classes of lidar_processing_lib_amd.hpp.
//
// Usage: processor_example <in.bin> <out.bin> [full | msg | image]
//   in.bin : int32 n, then n records {float x, y, z, pad, intensity; uint16 ring; 10 bytes padding}
//            (the 32-byte pcl::PointXYZIR layout)
//   out.bin: uint32 seg_labels[n], int32 n_obs, int32 cluster_labels[n_obs],
//            int32 k, int32 hull_offsets[k+1], int32 hull_indices[...]
//   full   : the noise filter runs first (NoiseRemover::filter, VALID points
//            kept in input order; seg_labels then covers the n_kept points) and
//            every hull polygon gets a rotating-calipers box; out.bin then
//            starts with int32 n_kept, uint8 noise[n] and ends with k boxes of
//            {double corners[8]; float area; float angle; int32 valid}.
//   msg    : the records are handed to the segmenter as a PointCloud2 message
//            (fields x, y, z, intensity, ring; CloudView, no conversion to a
//            pcl::PointCloud); the outputs are those of the default mode.
//   image  : the default mode's outputs followed by the segmenter's label
//            image (Segmenter::image, published by processor.cpp:804-816):
//            int32 rows, int32 cols, rows*cols BGR pixels.
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "lidar_processing_lib_amd.hpp"

namespace {

struct PointXYZIR {  // pcl::PointXYZIR (point_types.hpp:83-89): 16-byte aligned
    float x, y, z, pad;
    float intensity;
    std::uint16_t ring;
    std::uint8_t tail[10];
};
static_assert(sizeof(PointXYZIR) == 32, "PointXYZIR must be 32 bytes");

struct PointXYZ {
    float x, y, z;
};

template <typename P>
struct Cloud {  // the part of pcl::PointCloud<P> the library uses
    std::vector<P> points;
    std::size_t size() const { return points.size(); }
};

}  // namespace

int main(int argc, char** argv) {
    const std::string mode = argc == 4 ? argv[3] : "";
    if (argc != 3 && !(argc == 4 && (mode == "full" || mode == "msg" || mode == "image"))) {
        std::fprintf(stderr, "usage: %s in.bin out.bin [full | msg | image]\n", argv[0]);
        return 2;
    }
    const bool full = mode == "full";
    const bool as_msg = mode == "msg";
    std::FILE* fi = std::fopen(argv[1], "rb");
    if (!fi) return 2;
    std::int32_t n_in = 0;
    if (std::fread(&n_in, 4, 1, fi) != 1) return 2;
    Cloud<PointXYZIR> input;
    input.points.resize(n_in);
    if (std::fread(input.points.data(), sizeof(PointXYZIR), n_in, fi) != static_cast<std::size_t>(n_in)) return 2;
    std::fclose(fi);

    // Noise filter ahead of the segmenter: keep VALID points in input order.
    std::vector<lidar_processing_lib::NoiseRemoverLabel> noise;
    Cloud<PointXYZIR> cloud;
    if (full) {
        lidar_processing_lib::NoiseRemover remover;
        std::vector<lidar_processing_lib::NoiseRemover::PointT> pts(n_in);
        for (std::int32_t i = 0; i < n_in; ++i) pts[i] = {input.points[i].x, input.points[i].y, input.points[i].z};
        remover.filter(pts, noise);
        for (std::int32_t i = 0; i < n_in; ++i)
            if (noise[i] == lidar_processing_lib::NoiseRemoverLabel::VALID) cloud.points.push_back(input.points[i]);
    } else {
        cloud.points = input.points;
    }
    const std::int32_t n = static_cast<std::int32_t>(cloud.points.size());

    lidar_processing_lib::Segmenter segmenter;
    lidar_processing_lib::Clusterer clusterer;
    lidar_processing_lib::Polygonizer polygonizer;
    lidar_processing_lib::ClustererConfiguration ccfg;
    ccfg.voxel_grid_elevation_resolution_deg = 3.0F;  // processor.param.yaml:34
    clusterer.config(ccfg);

    std::vector<lidar_processing_lib::Label> labels;
    if (as_msg) {
        // The message a PointCloud2 subscriber receives (processor.cpp:494).
        struct Field {
            std::string name;
            std::uint32_t offset;
        };
        struct Message {
            std::vector<Field> fields;
            std::vector<std::uint8_t> data;
            std::uint32_t width, height, point_step, row_step;
        } msg;
        msg.fields = {{"x", 0}, {"y", 4}, {"z", 8}, {"intensity", 16}, {"ring", 20}};
        msg.width = static_cast<std::uint32_t>(n);
        msg.height = 1;
        msg.point_step = sizeof(PointXYZIR);
        msg.row_step = msg.width * msg.point_step;
        const auto* raw = reinterpret_cast<const std::uint8_t*>(cloud.points.data());
        msg.data.assign(raw, raw + static_cast<std::size_t>(n) * sizeof(PointXYZIR));
        segmenter.segment(lidar_processing_lib::CloudView(msg), labels);
    } else {
        segmenter.segment(cloud, labels);  // processor.cpp:556
    }

    Cloud<PointXYZ> obstacles;  // processor.cpp:562-579
    for (std::int32_t i = 0; i < n; ++i)
        if (labels[i] == lidar_processing_lib::Label::OBSTACLE)
            obstacles.points.push_back({cloud.points[i].x, cloud.points[i].y, cloud.points[i].z});

    std::vector<lidar_processing_lib::ClusterLabel> clabels;
    clusterer.cluster(obstacles, clabels);  // processor.cpp:599

    std::int32_t max_label = -1;
    for (auto l : clabels) max_label = l > max_label ? l : max_label;
    std::vector<std::vector<lidar_processing_lib::PointXY>> groups(max_label + 1);
    for (std::size_t i = 0; i < clabels.size(); ++i)  // processor.cpp:627-658
        if (clabels[i] >= 0)
            groups[clabels[i]].push_back({static_cast<double>(obstacles.points[i].x),
                                          static_cast<double>(obstacles.points[i].y)});
    std::vector<const std::vector<lidar_processing_lib::PointXY>*> lists;
    for (auto& g : groups) lists.push_back(&g);
    std::vector<std::vector<std::int32_t>> hulls;
    polygonizer.convexHulls(lists, hulls);  // processor.cpp:663, all clusters at once

    // processor.cpp:665-704: the hull polygon of each cluster, then its box.
    std::vector<lidar_processing_lib::BoundingBox> boxes;
    if (full) {
        std::vector<std::vector<lidar_processing_lib::PointXY>> polys(hulls.size());
        std::vector<const std::vector<lidar_processing_lib::PointXY>*> plist;
        for (std::size_t c = 0; c < hulls.size(); ++c) {
            for (auto v : hulls[c]) polys[c].push_back(groups[c][v]);
            plist.push_back(&polys[c]);
        }
        polygonizer.boundingBoxesRotatingCalipers(plist, boxes);
    }

    std::FILE* fo = std::fopen(argv[2], "wb");
    if (!fo) return 2;
    if (full) {
        std::fwrite(&n, 4, 1, fo);
        std::fwrite(noise.data(), 1, noise.size(), fo);
    }
    std::fwrite(labels.data(), 4, labels.size(), fo);
    const std::int32_t no = static_cast<std::int32_t>(clabels.size());
    std::fwrite(&no, 4, 1, fo);
    std::fwrite(clabels.data(), 4, clabels.size(), fo);
    const std::int32_t k = max_label + 1;
    std::fwrite(&k, 4, 1, fo);
    std::int32_t off = 0;
    std::fwrite(&off, 4, 1, fo);
    for (auto& h : hulls) {
        off += static_cast<std::int32_t>(h.size());
        std::fwrite(&off, 4, 1, fo);
    }
    for (auto& h : hulls) std::fwrite(h.data(), 4, h.size(), fo);
    for (const auto& b : boxes) {
        for (const auto& c : b.corners) {
            std::fwrite(&c.x, 8, 1, fo);
            std::fwrite(&c.y, 8, 1, fo);
        }
        const std::int32_t valid = b.is_valid ? 1 : 0;
        std::fwrite(&b.area, 4, 1, fo);
        std::fwrite(&b.angle_rad, 4, 1, fo);
        std::fwrite(&valid, 4, 1, fo);
    }
    if (mode == "image") {
        const lidar_processing_lib::LabelImage& img = segmenter.image();
        std::fwrite(&img.rows, 4, 1, fo);
        std::fwrite(&img.cols, 4, 1, fo);
        std::fwrite(img.bgr.data(), 1, img.bgr.size(), fo);
    }
    std::fclose(fo);
    std::printf("points %d obstacles %d clusters %d\n", n, no, k);
    return 0;
}
