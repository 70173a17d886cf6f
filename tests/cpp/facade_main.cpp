// Drives the C++ drop-in facade the way the reference's example does
// (main.cpp:6-47): readPLY -> train -> save -> load into a second detector
// -> match -> transformPointCloud / writePLY of the results, then
// sampleMesh on a mesh file.
// Usage: facade_main model.ply scene.ply out.model outdir mesh.ply
#include <ppf/ppf.h>

#include <cstdio>
#include <iostream>

int main(int argc, char **argv) {
    if (argc < 6)
        return 2;
    ppf::PointCloud model, scene;
    if (!ppf::readPLY(argv[1], model) || !ppf::readPLY(argv[2], scene))
        return 5;
    scene.viewPoint = ppf::Vec3f(620, 100, 500);
    {
        ppf::Timer    t("train model");
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
    const std::string dir = argv[4];
    for (size_t i = 0; i < pose.size(); i++) {
        std::cout << "score " << score[i] << " pose";
        for (int k = 0; k < 16; k++)
            std::cout << " " << pose[i].data()[k];
        std::cout << "\n";
        auto pc = ppf::transformPointCloud(model, pose[i]);
        if (!ppf::writePLY(dir + "/out" + std::to_string(i) + ".ply", pc))
            return 6;
    }
    if (!ppf::writePLY(dir + "/sampledScene.ply", result.sampledScene) ||
        !ppf::writePLY(dir + "/sampledKeypoint.ply", result.keyPoint, true))
        return 7;
    // a missing file: false and the reason on stdout
    ppf::PointCloud none;
    if (ppf::readPLY(dir + "/missing.ply", none))
        return 8;
    ppf::PointCloud mesh;
    if (!ppf::readPLY(argv[5], mesh))
        return 9;
    ppf::PointCloud samples = ppf::sampleMesh(mesh, 0.1f, 42);
    std::cout << "mesh faces " << mesh.face.size() << " samples " << samples.size() << "\n";
    auto valid = ppf::removeNan(samples, true);
    auto box   = ppf::computeBoundingBox(samples, valid);
    std::cout << "valid " << valid.size() << " box " << box.min[0] << " " << box.max[0] << "\n";
    return 0;
}
