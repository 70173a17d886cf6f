// ppf/ppf.h -- C++ drop-in for SurfaceMan/surface_match's public API
// (ppf.h:10-62, type.h:8-181) over the C ABI of libppf_amd.so (ppf_c.h).
//
// Same namespace, class names, member names, signatures, defaults and
// exception types as the reference.  With Eigen available the point and
// pose types are Eigen's (Vector3f, Matrix4f); without it, layout-identical
// PODs stand in, so client code that only fills and reads PointClouds
// compiles against either.
#pragma once

#include "../ppf_c.h"

#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#if defined(PPF_WITH_EIGEN) || (defined(__has_include) && __has_include(<Eigen/Core>) && !defined(PPF_NO_EIGEN))
#include <Eigen/Core>
#define PPF_EIGEN_TYPES 1
#endif

namespace ppf {

#ifdef PPF_EIGEN_TYPES
using Vec3f = Eigen::Vector3f;
using Vec3i = Eigen::Vector3i;
using Mat4f = Eigen::Matrix4f;
#else
// 12-byte point, the memory layout of Eigen::Vector3f
struct Vec3f {
    float v[3] = {0, 0, 0};
    Vec3f() = default;
    Vec3f(float x, float y, float z) : v{x, y, z} {}
    float &operator[](int i) { return v[i]; }
    float  operator[](int i) const { return v[i]; }
    float &x() { return v[0]; }
    float &y() { return v[1]; }
    float &z() { return v[2]; }
    float  x() const { return v[0]; }
    float  y() const { return v[1]; }
    float  z() const { return v[2]; }
    float *data() { return v; }
    const float *data() const { return v; }
};
struct Vec3i {
    int v[3] = {0, 0, 0};
    Vec3i() = default;
    Vec3i(int a, int b, int c) : v{a, b, c} {}
    int       &operator[](int i) { return v[i]; }
    int        operator[](int i) const { return v[i]; }
    int       *data() { return v; }
    const int *data() const { return v; }
};
// 4x4 column-major, the memory layout of Eigen::Matrix4f
struct Mat4f {
    float m[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    float &operator()(int r, int c) { return m[c * 4 + r]; }
    float  operator()(int r, int c) const { return m[c * 4 + r]; }
    float *data() { return m; }
    const float *data() const { return m; }
};
#endif
static_assert(sizeof(Vec3f) == 12, "point layout must be 3 packed floats");

namespace detail {
inline void throw_status(int rc) {
    if (rc == PPF_OK)
        return;
    const char *msg = ppf_last_error();
    if (rc == PPF_ERR_RANGE)
        throw std::range_error(msg);
    throw std::runtime_error(msg);
}
}  // namespace detail

// type.h:8-20
struct BoundingBox {
    Vec3f min{0, 0, 0};
    Vec3f max{0, 0, 0};
    BoundingBox() = default;
    BoundingBox(Vec3f mn, Vec3f mx) : min(mn), max(mx) {}
    Vec3f size() const { return Vec3f(max[0] - min[0], max[1] - min[1], max[2] - min[2]); }
    Vec3f center() const {
        return Vec3f((min[0] + max[0]) / 2.f, (min[1] + max[1]) / 2.f, (min[2] + max[2]) / 2.f);
    }
    float diameter() const {
        Vec3f s = size();
        return std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    }
};

// type.h:22-33
struct PointCloud {
    std::vector<Vec3f> point;
    std::vector<Vec3f> normal;
    std::vector<Vec3i> face;  // a model with faces is sampled by trainModel (ppf.cpp:63-70)
    BoundingBox        box;
    Vec3f              viewPoint = Vec3f(NAN, NAN, NAN);

    bool        hasNormal() const { return !point.empty() && point.size() == normal.size(); }
    std::size_t size() const { return point.size(); }
    bool        empty() const { return point.empty(); }
};

// type.h:39-74
struct TrainParam {
    float featDistanceStepRel;
    int   featAngleResolution;
    float poseRefRelSamplingDistance;
    int   knnNormal;
    bool  smoothNormal;

    explicit TrainParam(float featDistanceStepRel_ = 0.04f, int featAngleResolution_ = 30,
                        float poseRefRelSamplingDistance_ = 0.01f, int knnNormal_ = 10,
                        bool smoothNormal_ = true)
        : featDistanceStepRel(featDistanceStepRel_), featAngleResolution(featAngleResolution_),
          poseRefRelSamplingDistance(poseRefRelSamplingDistance_), knnNormal(knnNormal_),
          smoothNormal(smoothNormal_) {
        ppf_train_param p = c();
        detail::throw_status(ppf_train_param_check(&p));
    }
    ppf_train_param c() const {
        return ppf_train_param{featDistanceStepRel, featAngleResolution, poseRefRelSamplingDistance,
                               knnNormal, smoothNormal ? 1 : 0};
    }
};

// type.h:80-176
struct MatchParam {
    int   numMatches;
    int   knnNormal;
    bool  smoothNormal;
    bool  invertNormal;
    float maxOverlapDistRel;
    float maxOverlapDistAbs;
    bool  sparsePoseRefinement;
    bool  densePoseRefinement;
    int   poseRefNumSteps;
    float poseRefDistThresholdRel;
    float poseRefDistThresholdAbs;
    float poseRefScoringDistRel;
    float poseRefScoringDistAbs;

    explicit MatchParam(int numMatches_ = 1, int knnNormal_ = 10, bool smoothNormal_ = true,
                        bool invertNormal_ = false, float maxOverlapDistRel_ = 0.5f,
                        float maxOverlapDistAbs_ = 0, bool sparsePoseRefinement_ = true,
                        bool densePoseRefinement_ = true, int poseRefNumSteps_ = 5,
                        float poseRefDistThresholdRel_ = 0.1f, float poseRefDistThresholdAbs_ = 0,
                        float poseRefScoringDistRel_ = 0.01f, float poseRefScoringDistAbs_ = 0)
        : numMatches(numMatches_), knnNormal(knnNormal_), smoothNormal(smoothNormal_),
          invertNormal(invertNormal_), maxOverlapDistRel(maxOverlapDistRel_),
          maxOverlapDistAbs(maxOverlapDistAbs_), sparsePoseRefinement(sparsePoseRefinement_),
          densePoseRefinement(densePoseRefinement_), poseRefNumSteps(poseRefNumSteps_),
          poseRefDistThresholdRel(poseRefDistThresholdRel_),
          poseRefDistThresholdAbs(poseRefDistThresholdAbs_),
          poseRefScoringDistRel(poseRefScoringDistRel_),
          poseRefScoringDistAbs(poseRefScoringDistAbs_) {
        ppf_match_param p = c();
        detail::throw_status(ppf_match_param_check(&p));
    }
    ppf_match_param c() const {
        return ppf_match_param{numMatches, knnNormal, smoothNormal ? 1 : 0, invertNormal ? 1 : 0,
                               maxOverlapDistRel, maxOverlapDistAbs, sparsePoseRefinement ? 1 : 0,
                               densePoseRefinement ? 1 : 0, poseRefNumSteps,
                               poseRefDistThresholdRel, poseRefDistThresholdAbs,
                               poseRefScoringDistRel, poseRefScoringDistAbs};
    }
};

// type.h:178-181
struct MatchResult {
    PointCloud sampledScene;
    PointCloud keyPoint;
};

// ppf.h:10-62
class Detector {
public:
    Detector() {
        ppf_detector h = nullptr;
        detail::throw_status(ppf_detector_create(&h));
        h_.reset(h);
    }
    ~Detector() = default;

    // key of the mesh sampler's Philox stream (trainModel of a model with faces)
    uint64_t meshSeed = 0x5EED;

    void trainModel(const PointCloud &model, float samplingDistanceRel = 0.04f,
                    TrainParam param = TrainParam()) {
        float           box[6];
        ppf_train_param p = param.c();
        pack_box(model, box);
        if (!model.face.empty()) {
            // the mesh sampler's uniforms are seeded (the reference draws its
            // seed from std::random_device)
            static_assert(sizeof(Vec3i) == 3 * sizeof(int), "faces are packed int triples");
            detail::throw_status(ppf_train_model_mesh(
                h_.get(), model.point.empty() ? nullptr : model.point[0].data(),
                model.hasNormal() ? model.normal[0].data() : nullptr, model.point.size(),
                model.face[0].data(), model.face.size(), box, model.viewPoint.data(),
                samplingDistanceRel, &p, meshSeed));
            return;
        }
        detail::throw_status(ppf_train_model(
            h_.get(), model.point.empty() ? nullptr : model.point[0].data(),
            model.hasNormal() ? model.normal[0].data() : nullptr, model.point.size(), box,
            model.viewPoint.data(), samplingDistanceRel, &p));
    }

    void matchScene(const PointCloud &scene, std::vector<Mat4f> &pose, std::vector<float> &score,
                    float samplingDistanceRel = 0.04f, float keyPointFraction = 0.2f,
                    float minScore = 0.2f, MatchParam param = MatchParam(),
                    MatchResult *matchResult = nullptr) {
        float            box[6];
        ppf_match_param  p   = param.c();
        size_t           cap = param.numMatches > 0 ? (size_t)param.numMatches : 1, n = 0;
        std::vector<float> poses(16 * cap), scores(cap);
        pack_box(scene, box);
        ppf_match_result mr{};
        std::vector<float> sp, sn, kp, kn;
        if (matchResult) {
            size_t c = scene.point.size();
            sp.resize(3 * c), sn.resize(3 * c), kp.resize(3 * c), kn.resize(3 * c);
            mr = ppf_match_result{sp.data(), sn.data(), c, 0, kp.data(), kn.data(), c, 0};
        }
        detail::throw_status(ppf_match_scene(
            h_.get(), scene.point.empty() ? nullptr : scene.point[0].data(),
            scene.hasNormal() ? scene.normal[0].data() : nullptr, scene.point.size(), box,
            scene.viewPoint.data(), samplingDistanceRel, keyPointFraction, minScore, &p,
            poses.data(), scores.data(), cap, &n, matchResult ? &mr : nullptr, nullptr));
        pose.resize(n);
        score.resize(n);
        for (size_t i = 0; i < n; i++) {
            std::copy(poses.begin() + 16 * i, poses.begin() + 16 * (i + 1), pose[i].data());
            score[i] = scores[i];
        }
        if (matchResult) {
            unpack(sp, sn, mr.nSampled, matchResult->sampledScene);
            unpack(kp, kn, mr.nKey, matchResult->keyPoint);
        }
    }

    void save(const std::string &filename) const {
        detail::throw_status(ppf_save(h_.get(), filename.c_str()));
    }
    void load(const std::string &filename) {
        detail::throw_status(ppf_load(h_.get(), filename.c_str()));
    }

private:
    struct Deleter {
        void operator()(ppf_detector d) const { ppf_detector_destroy(d); }
    };
    std::unique_ptr<ppf_detector_s, Deleter> h_;

    static void pack_box(const PointCloud &c, float box[6]) {
        for (int a = 0; a < 3; a++) {
            box[a]     = c.box.min[a];
            box[a + 3] = c.box.max[a];
        }
    }
    static void unpack(const std::vector<float> &p, const std::vector<float> &n, size_t count,
                       PointCloud &out) {
        out.point.resize(count);
        out.normal.resize(count);
        for (size_t i = 0; i < count; i++) {
            out.point[i]  = Vec3f(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
            out.normal[i] = Vec3f(n[3 * i], n[3 * i + 1], n[3 * i + 2]);
        }
    }
};

}  // namespace ppf
