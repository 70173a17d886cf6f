// Drives the C++ drop-in facade the way the reference's example does
// (main.cpp:6-47): train -> save -> load into a second detector -> match.
// Usage: facade_main model.bin scene.bin out.model
// Files: int32 n, then n*3 float points, then (model only) n*3 float normals.
#include <ppf/ppf.h>

#include <cstdio>
#include <fstream>
#include <iostream>

static ppf::PointCloud read_cloud(const char *path, bool normals) {
    std::ifstream f(path, std::ios::binary);
    int           n = 0;
    f.read(reinterpret_cast<char *>(&n), 4);
    ppf::PointCloud c;
    c.point.resize(n);
    f.read(reinterpret_cast<char *>(c.point.data()), 12 * (size_t)n);
    if (normals) {
        c.normal.resize(n);
        f.read(reinterpret_cast<char *>(c.normal.data()), 12 * (size_t)n);
    }
    return c;
}

int main(int argc, char **argv) {
    if (argc < 4)
        return 2;
    ppf::PointCloud model = read_cloud(argv[1], true);
    ppf::PointCloud scene = read_cloud(argv[2], false);
    scene.viewPoint       = ppf::Vec3f(620, 100, 500);
    {
        ppf::Detector detector;
        detector.trainModel(model, 0.04f);
        detector.save(argv[3]);
    }
    // the reference's exceptions come through unchanged
    try {
        ppf::Detector empty;
        std::vector<ppf::Mat4f> p;
        std::vector<float>      s;
        empty.matchScene(scene, p, s);
        return 3;
    } catch (const std::runtime_error &e) {
        std::cout << "expected: " << e.what() << "\n";
    }
    try {
        ppf::TrainParam bad(0.0f);
        return 4;
    } catch (const std::range_error &e) {
        std::cout << "expected: " << e.what() << "\n";
    }
    ppf::Detector detector;
    detector.load(argv[3]);
    std::vector<ppf::Mat4f> pose;
    std::vector<float>      score;
    ppf::MatchResult        result;
    detector.matchScene(scene, pose, score, 0.04f, 0.1f, 0.2f, ppf::MatchParam(3), &result);
    std::cout << "matches " << pose.size() << " sampled " << result.sampledScene.size()
              << " keypoints " << result.keyPoint.size() << "\n";
    for (size_t i = 0; i < pose.size(); i++) {
        std::cout << "score " << score[i] << " pose";
        for (int k = 0; k < 16; k++)
            std::cout << " " << pose[i].data()[k];
        std::cout << "\n";
    }
    return 0;
}
