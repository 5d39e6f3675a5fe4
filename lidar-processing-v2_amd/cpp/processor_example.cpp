// processor_example.cpp -- the reference processor's per-frame calls
// (src/processor/src/processor.cpp:556-670) written against the drop-in
// lidar_processing_lib classes of lidar_processing_lib_amd.hpp.
//
// Usage: processor_example <in.bin> <out.bin>
//   in.bin : int32 n, then n records {float x, y, z, pad, intensity; uint16 ring; 10 bytes padding}
//            (the 32-byte pcl::PointXYZIR layout)
//   out.bin: uint32 seg_labels[n], int32 n_obs, int32 cluster_labels[n_obs],
//            int32 k, int32 hull_offsets[k+1], int32 hull_indices[...]
#include <cstdio>
#include <cstdint>
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
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
        return 2;
    }
    std::FILE* fi = std::fopen(argv[1], "rb");
    if (!fi) return 2;
    std::int32_t n = 0;
    if (std::fread(&n, 4, 1, fi) != 1) return 2;
    Cloud<PointXYZIR> cloud;
    cloud.points.resize(n);
    if (std::fread(cloud.points.data(), sizeof(PointXYZIR), n, fi) != static_cast<std::size_t>(n)) return 2;
    std::fclose(fi);

    lidar_processing_lib::Segmenter segmenter;
    lidar_processing_lib::Clusterer clusterer;
    lidar_processing_lib::Polygonizer polygonizer;
    lidar_processing_lib::ClustererConfiguration ccfg;
    ccfg.voxel_grid_elevation_resolution_deg = 3.0F;  // processor.param.yaml:34
    clusterer.config(ccfg);

    std::vector<lidar_processing_lib::Label> labels;
    segmenter.segment(cloud, labels);  // processor.cpp:556

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

    std::FILE* fo = std::fopen(argv[2], "wb");
    if (!fo) return 2;
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
    std::fclose(fo);
    std::printf("points %d obstacles %d clusters %d\n", n, no, k);
    return 0;
}
