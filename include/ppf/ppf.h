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

#include <chrono>
#include <cmath>
#include <cstdint>
#include <iostream>
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

static_assert(sizeof(Vec3i) == 3 * sizeof(int), "face layout must be 3 packed ints");

// ---- filePLY.h:7-9 ----------------------------------------------------
// readPLY: points, normals (when the file has nx) and ear-clipped faces; on
// failure prints the reason and returns false, as the reference does
inline bool readPLY(const std::string &filename, PointCloud &mesh) {
    ppf_ply h = nullptr;
    if (ppf_ply_read(filename.c_str(), &h) != PPF_OK) {
        std::cout << ppf_last_error() << std::endl;
        return false;
    }
    size_t nv = 0, nn = 0, nf = 0;
    ppf_ply_info(h, &nv, &nn, &nf);
    mesh.point.resize(nv);
    mesh.normal.resize(nn);
    mesh.face.resize(nf);
    ppf_ply_copy(h, nv ? mesh.point[0].data() : nullptr, nn ? mesh.normal[0].data() : nullptr,
                 nf ? reinterpret_cast<int *>(mesh.face.data()) : nullptr);
    ppf_ply_free(h);
    return true;
}

inline bool writePLY(const std::string &filename, const PointCloud &mesh,
                     bool write_ascii = false) {
    const int rc = ppf_ply_write(
        filename.c_str(), mesh.point.empty() ? nullptr : mesh.point[0].data(),
        mesh.hasNormal() ? mesh.normal[0].data() : nullptr, mesh.point.size(),
        mesh.face.empty() ? nullptr : reinterpret_cast<const int *>(mesh.face.data()),
        mesh.face.size(), write_ascii ? 1 : 0);
    if (rc != PPF_OK) {
        std::cout << ppf_last_error() << std::endl;
        return false;
    }
    return true;
}

// ---- util.h:8-21 ------------------------------------------------------
// sampleMesh (util.cpp:20-84) on the GPU; the uniforms come from a Philox
// stream keyed by seed where the reference draws from std::random_device
inline PointCloud sampleMesh(const PointCloud &pc, float radius, uint64_t seed = 0x5EED) {
    PointCloud result;
    if (pc.face.empty())
        return result;
    const float *pts   = pc.point.empty() ? nullptr : pc.point[0].data();
    const int   *faces = reinterpret_cast<const int *>(pc.face.data());
    size_t       n     = 0;
    detail::throw_status(ppf_sample_mesh(pts, pc.point.size(), faces, pc.face.size(), radius,
                                         seed, nullptr, nullptr, 0, &n));
    result.point.resize(n);
    result.normal.resize(n);
    if (n)
        detail::throw_status(ppf_sample_mesh(pts, pc.point.size(), faces, pc.face.size(), radius,
                                             seed, result.point[0].data(),
                                             result.normal[0].data(), n, &n));
    result.box = pc.box;
    return result;
}

namespace detail {
inline bool finite3(const Vec3f &v) {
    return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]);
}
// Eigen's normalized(): v / |v|, v itself when |v|^2 is 0
inline Vec3f normalized(const Vec3f &v) {
    const float s = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    if (!(s > 0.0f))
        return v;
    const float l = std::sqrt(s);
    return Vec3f(v[0] / l, v[1] / l, v[2] / l);
}
}  // namespace detail

// removeNan (util.cpp:86-105): indices of the points (and normals, when
// asked and present) with all coordinates finite
inline std::vector<int> removeNan(const PointCloud &pc, bool checkNormal = false) {
    checkNormal = checkNormal && pc.hasNormal();
    std::vector<int> result;
    result.reserve(pc.size());
    for (size_t i = 0; i < pc.size(); i++)
        if (detail::finite3(pc.point[i]) && (!checkNormal || detail::finite3(pc.normal[i])))
            result.push_back((int)i);
    return result;
}

// computeBoundingBox (util.cpp:178-207), over validIndices when given
inline BoundingBox computeBoundingBox(const PointCloud &pc,
                                      const std::vector<int> &validIndices = {}) {
    Vec3f mn = validIndices.empty() ? pc.point[0] : pc.point[validIndices[0]];
    Vec3f mx = mn;
    const size_t m = validIndices.empty() ? pc.size() : validIndices.size();
    for (int dim = 0; dim < 3; dim++)
        for (size_t t = 0; t < m; t++) {
            const Vec3f &p = pc.point[validIndices.empty() ? t : (size_t)validIndices[t]];
            if (p[dim] > mx[dim])
                mx[dim] = p[dim];
            else if (p[dim] < mn[dim])
                mn[dim] = p[dim];
        }
    return BoundingBox(mn, mx);
}

// extraIndices (util.cpp:141-162): the listed points, normals normalized
inline PointCloud extraIndices(const PointCloud &pc, const std::vector<std::size_t> &indices) {
    PointCloud result;
    const bool hasNormal = pc.hasNormal();
    result.point.resize(indices.size());
    if (hasNormal)
        result.normal.resize(indices.size());
    for (size_t i = 0; i < indices.size(); i++) {
        result.point[i] = pc.point[indices[i]];
        if (hasNormal)
            result.normal[i] = detail::normalized(pc.normal[indices[i]]);
    }
    result.box = pc.box;
    if (result.box.diameter() == 0)
        result.box = computeBoundingBox(result);
    return result;
}

// normalizeNormal (util.cpp:164-176)
inline void normalizeNormal(PointCloud &pc, bool invert = false) {
    for (auto &n : pc.normal) {
        const Vec3f u = detail::normalized(n);
        n             = invert ? Vec3f(-u[0], -u[1], -u[2]) : u;
    }
}

// transformPointCloud (util.cpp:209-232): R p + t and R n, faces kept, box
// recomputed
inline PointCloud transformPointCloud(const PointCloud &pc, const Mat4f &pose,
                                      bool useNormal = true) {
    const bool doNormal = pc.hasNormal() && useNormal;
    PointCloud result;
    result.point.resize(pc.size());
    if (doNormal)
        result.normal.resize(pc.size());
    auto rot = [&](const Vec3f &v) {
        return Vec3f((pose(0, 0) * v[0] + pose(0, 1) * v[1]) + pose(0, 2) * v[2],
                     (pose(1, 0) * v[0] + pose(1, 1) * v[1]) + pose(1, 2) * v[2],
                     (pose(2, 0) * v[0] + pose(2, 1) * v[1]) + pose(2, 2) * v[2]);
    };
    for (size_t i = 0; i < pc.size(); i++) {
        const Vec3f r   = rot(pc.point[i]);
        result.point[i] = Vec3f(r[0] + pose(0, 3), r[1] + pose(1, 3), r[2] + pose(2, 3));
        if (doNormal)
            result.normal[i] = rot(pc.normal[i]);
    }
    result.face = pc.face;
    if (!result.point.empty())
        result.box = computeBoundingBox(result);
    return result;
}

// ---- helper.h ---------------------------------------------------------
// prints "<content> cost(ms): <ms>" when released or destroyed
class Timer {
public:
    explicit Timer(std::string content)
        : content_(std::move(content)), start_(std::chrono::steady_clock::now()) {}
    ~Timer() {
        if (!released_)
            release();
    }
    void release() {
        released_ = true;
        const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(
            std::chrono::steady_clock::now() - start_);
        std::cout << content_ << " cost(ms): " << ms.count() << std::endl;
    }

private:
    std::string                           content_;
    std::chrono::steady_clock::time_point start_;
    bool                                  released_ = false;
};

}  // namespace ppf
