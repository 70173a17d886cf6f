// detector.hip -- host orchestration of the PPF path and the C ABI
// (include/ppf_c.h).
//
// Detector::trainModel (ppf.cpp:36-147) and Detector::matchScene
// (ppf.cpp:149-510) are restated stage by stage; each stage runs on the GPU
// (grid.hip, table.hip, vote.hip, cluster.hip, icp.hip) and the host keeps
// only the bookkeeping: parameter derivation, small read-backs, and the
// ordered accept/reject walk over refined hypotheses.  Model files keep the
// reference's layout (ppf.cpp:512-547, serialize.cpp:5-186).
#include "ppf_dev.h"
#include "../../include/ppf_c.h"


#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <fstream>
#include <functional>
#include <memory>
#include <mutex>
#include <numeric>
#include <shared_mutex>
#include <string>
#include <thread>
#include <vector>

using namespace ppfd;

namespace {

thread_local std::string g_lastError;

const int   kVersion = 100;      // ppf.cpp:21
const int   kMagic   = 0x7F27F;  // ppf.cpp:22

struct HostCloud {
    std::vector<float> p, n;  // n x 3 each (n may be empty)
    float boxMin[3] = {0, 0, 0}, boxMax[3] = {0, 0, 0};
    float vp[3] = {NAN, NAN, NAN};
    int   size() const { return (int)(p.size() / 3); }
};

struct HostModel {
    float           samplingDistanceRel = 0.04f;
    ppf_train_param param{};
    HostCloud       sampled, reSampled;
    std::vector<uint32_t> keys, refs;
    std::vector<int>      offs;
    std::vector<float>    alphas;
};

float box_diameter(const float *lo, const float *hi) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}

void normalize3(float *v) {
    float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0]    = v[0] / l;
    v[1]    = v[1] / l;
    v[2]    = v[2] / l;
}

bool finite3(const float *v) {
    return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]);
}

// computeBoundingBox (util.cpp:178-207) over the valid indices
void bbox_host(const float *pts, const std::vector<int> &valid, float lo[3], float hi[3]) {
    const float *p0 = pts + 3 * (valid.empty() ? 0 : valid[0]);
    for (int a = 0; a < 3; a++)
        lo[a] = hi[a] = p0[a];
    for (int i : valid) {
        const float *p = pts + 3 * (size_t)i;
        for (int a = 0; a < 3; a++) {
            if (p[a] > hi[a])
                hi[a] = p[a];
            else if (p[a] < lo[a])
                lo[a] = p[a];
        }
    }
}

// removeNan (util.cpp:86-105)
std::vector<int> remove_nan(const float *pts, const float *nrm, int n) {
    std::vector<int> v;
    v.reserve(n);
    for (int i = 0; i < n; i++) {
        if (!finite3(pts + 3 * (size_t)i))
            continue;
        if (nrm && !finite3(nrm + 3 * (size_t)i))
            continue;
        v.push_back(i);
    }
    return v;
}

__global__ void split_kernel(const float *aos, int n, float *x, float *y, float *z) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    x[i] = aos[3 * (size_t)i];
    y[i] = aos[3 * (size_t)i + 1];
    z[i] = aos[3 * (size_t)i + 2];
}
__global__ void fill_nan_kernel(float *a, float *b, float *c, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        a[i] = b[i] = c[i] = NAN;
}
// normalizeNormal (util.cpp:163-176)
__global__ void normalize_kernel(float *x, float *y, float *z, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float a = x[i], b = y[i], c = z[i];
    float l = sqrtf(a * a + b * b + c * c);
    x[i]    = a / l;
    y[i]    = b / l;
    z[i]    = c / l;
}
__global__ void gather_aos_kernel(const int *idx, int m, const float *x, const float *y,
                                  const float *z, float *out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m)
        return;
    int i          = idx[t];
    out[3 * t]     = x[i];
    out[3 * t + 1] = y[i];
    out[3 * t + 2] = z[i];
}

// Cloud held on the device: originals (as uploaded) and working copies that
// the pipeline mutates (normal estimation moves points to centroids).
struct DevCloud {
    int    n = 0;
    bool   hasNormal = false;
    DevBuf aos, aosN, ox, oy, oz, onx, ony, onz, px, py, pz, nx, ny, nz, valid;
    int    nValid = 0;
    float  boxMin[3], boxMax[3], vp[3];
    // bounding box of the valid points (min xyz, max xyz): every working
    // position stays inside it (normal estimation moves points to
    // centroids of valid points), so grids over them need no device box
    float  ext[6];

    // the caller's AoS arrays to HBM, then removeNan + computeBoundingBox
    // (util.cpp:86-105,178-207) on the device: SoA originals, the ascending
    // valid indices and the box of the valid points (ppf.cpp:170-176)
    void upload(const float *pts, const float *nrm, int count, const float *box6,
                const float *viewPoint, hipStream_t s, IngestScratch &is) {
        n         = count;
        hasNormal = nrm != nullptr;
        int    m  = std::max(n, 1);
        float *A  = aos.as<float>((size_t)m * 3);
        float *B  = hasNormal ? aosN.as<float>((size_t)m * 3) : nullptr;
        PPF_HIP_CHECK(hipMemcpyAsync(A, pts, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
        if (hasNormal)
            PPF_HIP_CHECK(hipMemcpyAsync(B, nrm, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
        float *X = ox.as<float>(m), *Y = oy.as<float>(m), *Z = oz.as<float>(m);
        float *NX = onx.as<float>(m), *NY = ony.as<float>(m), *NZ = onz.as<float>(m);
        px.as<float>(m);
        py.as<float>(m);
        pz.as<float>(m);
        nx.as<float>(m);
        ny.as<float>(m);
        nz.as<float>(m);
        nValid = ingest_cloud(is, A, B, n, X, Y, Z, hasNormal ? NX : nullptr,
                              hasNormal ? NY : nullptr, hasNormal ? NZ : nullptr,
                              valid.as<int>(m), ext, s);
        if (box6 && box_diameter(box6, box6 + 3) != 0.0f) {
            std::memcpy(boxMin, box6, sizeof boxMin);
            std::memcpy(boxMax, box6 + 3, sizeof boxMax);
        } else {
            std::memcpy(boxMin, ext, sizeof boxMin);
            std::memcpy(boxMax, ext + 3, sizeof boxMax);
        }
        for (int a = 0; a < 3; a++)
            vp[a] = viewPoint ? viewPoint[a] : NAN;
    }

    // working copies := originals
    void reset(hipStream_t s) {
        size_t b = sizeof(float) * n;
        PPF_HIP_CHECK(hipMemcpyAsync(px.get<float>(), ox.get<float>(), b, hipMemcpyDeviceToDevice, s));
        PPF_HIP_CHECK(hipMemcpyAsync(py.get<float>(), oy.get<float>(), b, hipMemcpyDeviceToDevice, s));
        PPF_HIP_CHECK(hipMemcpyAsync(pz.get<float>(), oz.get<float>(), b, hipMemcpyDeviceToDevice, s));
        if (hasNormal) {
            PPF_HIP_CHECK(hipMemcpyAsync(nx.get<float>(), onx.get<float>(), b, hipMemcpyDeviceToDevice, s));
            PPF_HIP_CHECK(hipMemcpyAsync(ny.get<float>(), ony.get<float>(), b, hipMemcpyDeviceToDevice, s));
            PPF_HIP_CHECK(hipMemcpyAsync(nz.get<float>(), onz.get<float>(), b, hipMemcpyDeviceToDevice, s));
            if (n > 0)
                normalize_kernel<<<div_up(n, 256), 256, 0, s>>>(nx.get<float>(), ny.get<float>(),
                                                                 nz.get<float>(), n);
        } else if (n > 0) {
            fill_nan_kernel<<<div_up(n, 256), 256, 0, s>>>(nx.get<float>(), ny.get<float>(),
                                                            nz.get<float>(), n);
        }
    }
};

}  // namespace

struct ppf_scene_s {
    DevCloud   cloud;
    std::mutex busy;  // a match moves the scene's working points: one at a time
};

struct ppf_icp_target_s {
    hipStream_t stream = nullptr;
    DevBuf      x, y, z, nx, ny, nz, aos;
    Grid        grid, full;
    Workspace   ws;
    DevBuf      sx, sy, sz, snx, sny, snz, saos;
    DevBuf      views, poseTarget;  // ppf_icp_refine_multi: every target's grids, each pose's target
    int         n = 0;
    hipEvent_t  ev[2] = {nullptr, nullptr};
    ~ppf_icp_target_s() {
        for (auto &e : ev)
            if (e)
                (void)hipEventDestroy(e);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
};

// Everything one match call writes: a stream, scratch, index buffers,
// events, the host-array scene copy and the helper lanes.  A detector keeps
// a pool of these; each call takes one for its duration, so calls on one
// detector run concurrently (the reference's matchScene only reads the
// trained model, ppf.cpp:149-510).
struct MatchCtx {
    hipStream_t   stream = nullptr;
    Workspace     ws;
    IngestScratch ingest;
    DevCloud      tmpScene;
    DevBuf        idxA, idxB, idxC, idxD;  // sampled / keypoints / dense / scratch
    hipEvent_t    ev[13] = {};
    // per-scene state of a batched match (grown on demand, reused)
    struct SceneWork {
        DevBuf samp, kp, dense;
        Grid   gVote, gSparse, gFull, gDense;
        int    ns = 0, nk = 0, nd = 0;
    };
    std::vector<std::unique_ptr<SceneWork>> sceneWork;
    DevBuf batchKpScene, batchKpId, batchKpDesc, batchGrids, batchHypScene;
    Workspace afterWs;  // the lanes' after-pass work of a batched match
    // grids of every scene of a batched match, built together
    MultiGrid batchVote, batchSparse, batchFull, batchDense, batchNormal;
    // the per-scene stages of a batched match run on up to kLanes host
    // threads, each with its own stream (default priority) and workspace,
    // so the small kernels and read-backs of different scenes overlap
    static constexpr int kLanes = 4;
    struct Lane {
        hipStream_t stream = nullptr;
        hipEvent_t  ev[3]  = {nullptr, nullptr, nullptr};  // timed lane work; end of a pass
        Workspace   ws;
        ~Lane() {
            for (auto e : ev)
                if (e)
                    (void)hipEventDestroy(e);
            if (stream)
                (void)hipStreamDestroy(stream);
        }
    };
    std::vector<std::unique_ptr<Lane>> lanes;

    MatchCtx() {
        PPF_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        for (auto &e : ev)
            PPF_HIP_CHECK(hipEventCreate(&e));
    }
    ~MatchCtx() {
        lanes.clear();
        for (auto &e : ev)
            if (e)
                (void)hipEventDestroy(e);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
};

// The trained model and the launch shapes (read by every call, written by
// train / load / set_tuning / autotune under the exclusive lock) and the
// pool of call contexts.
struct ppf_detector_s {
    bool         trained = false;
    HostModel    hm;
    DevModel     dm;
    TrainScratch ts;
    VoteConfig   vcfg;
    IcpConfig    icfg;
    int          normCell4 = 4;  // scene normals grid cell = sample step * normCell4 / 4
    int          refTail = 1;
    std::shared_mutex                      model;   // shared: readers, unique: writers
    std::mutex                             poolMu;  // guards the two lists below
    std::vector<std::unique_ptr<MatchCtx>> ctxs;
    std::vector<MatchCtx *>                freeCtx;
};

namespace {

// a call context from the detector's pool for the duration of one call
struct Lease {
    ppf_detector_s &D;
    MatchCtx       *X = nullptr;
    explicit Lease(ppf_detector_s &d) : D(d) {
        std::lock_guard<std::mutex> g(D.poolMu);
        if (D.freeCtx.empty()) {
            D.ctxs.push_back(std::make_unique<MatchCtx>());
            X = D.ctxs.back().get();
        } else {
            X = D.freeCtx.back();
            D.freeCtx.pop_back();
        }
    }
    Lease(const Lease &) = delete;
    Lease &operator=(const Lease &) = delete;
    ~Lease() {
        std::lock_guard<std::mutex> g(D.poolMu);
        D.freeCtx.push_back(X);
    }
    MatchCtx &operator*() const { return *X; }
};
using ReadLock  = std::shared_lock<std::shared_mutex>;
using WriteLock = std::unique_lock<std::shared_mutex>;

// cells of the refinement's target grids, in units of the sample step
// (sparse stage) and of the resample step (dense stage): measured on the
// 1M-point scene (sparse 1x 1.09 ms, 2x 0.89, 3x 0.94; dense 1x 5.5 ms,
// 2x 1.97, 3x 1.74, 4x 1.82, 6x 2.28).  No cell changes a result.
constexpr float kSparseCell = 2.0f;
constexpr float kDenseCell  = 3.0f;

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess)
        return 0.0f;
    return ms;
}

// TrainParam constructor checks (type.cpp:71-89)
void check_train_param(const ppf_train_param &p) {
    if (p.featDistanceStepRel > 1 || p.featDistanceStepRel <= 0)
        throw std::range_error("Invalid Input: featDistanceStepRel range mismatch in TrainParam()");
    if (p.featAngleResolution <= 0)
        throw std::range_error("Invalid Input: featAngleResolution range mismatch in TrainParam()");
    if (p.poseRefRelSamplingDistance > 1 || p.poseRefRelSamplingDistance <= 0)
        throw std::range_error(
            "Invalid Input: poseRefRelSamplingDistance range mismatch in TrainParam()");
}

// MatchParam constructor checks (type.cpp:91-127), including the reference's
// guard on !densePoseRefinement
void check_match_param(const ppf_match_param &p) {
    if (p.numMatches < 1)
        throw std::range_error("Invalid Input: numMatches range mismatch in MatchParam()");
    if ((p.maxOverlapDistRel < 0) && (p.maxOverlapDistAbs < 0))
        throw std::range_error("Invalid Input: maxOverlapDist range mismatch in MatchParam()");
    if (!p.densePoseRefinement) {
        if (p.poseRefNumSteps < 1)
            throw std::range_error("Invalid Input: poseRefNumSteps range mismatch in MatchParam()");
        if ((p.poseRefDistThresholdRel < 0) && (p.poseRefDistThresholdAbs < 0))
            throw std::range_error(
                "Invalid Input: poseRefDistThreshold range mismatch in MatchParam()");
        if ((p.poseRefScoringDistRel < 0) && (p.poseRefScoringDistAbs < 0))
            throw std::range_error(
                "Invalid Input: poseRefScoringDist range mismatch in MatchParam()");
    }
}

void download_aos(const int *idxDev, int m, const DevCloud &c, bool normals, std::vector<float> &out,
                  DevBuf &scratch, hipStream_t s) {
    out.resize((size_t)m * 3);
    if (m == 0)
        return;
    float *d = scratch.as<float>((size_t)m * 3);
    if (normals)
        gather_aos_kernel<<<div_up(m, 256), 256, 0, s>>>(idxDev, m, c.nx.get<float>(),
                                                          c.ny.get<float>(), c.nz.get<float>(), d);
    else
        gather_aos_kernel<<<div_up(m, 256), 256, 0, s>>>(idxDev, m, c.px.get<float>(),
                                                          c.py.get<float>(), c.pz.get<float>(), d);
    PPF_HIP_CHECK(hipMemcpyAsync(out.data(), d, sizeof(float) * 3 * m, hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
}

void train_impl(ppf_detector_s &D, MatchCtx &X, const float *pts, const float *nrm, int n, const float *box6,
                const float *vp, float rel, const ppf_train_param &param) {
    // ppf.cpp:39-44
    if (rel > 1 || rel < 0)
        throw std::range_error("Invalid Input: samplingDistanceRel range mismatch in trainModel");
    if (n <= 0 || !pts)
        throw std::runtime_error("Invalid Input: empty model in trainModel");
    check_train_param(param);
    hipStream_t s = X.stream;
    DevCloud   &c = X.tmpScene;
    c.upload(pts, nrm, n, box6, vp, s, X.ingest);
    c.reset(s);
    float Dm           = box_diameter(c.boxMin, c.boxMax);             // ppf.cpp:54
    float sampleStep   = Dm * rel;                                     // :56
    float reSampleStep = Dm * param.poseRefRelSamplingDistance;        // :57
    float distanceStep = Dm * param.featDistanceStepRel;               // :58
    float angleStep    = ppfc::kTwoPi / (float)param.featAngleResolution;  // :59
    int  *cnt          = X.ws.smallCnt.as<int>(8);
    int  *s1 = X.idxA.as<int>(std::max(c.nValid, 1)), *s2 = X.idxB.as<int>(std::max(c.nValid, 1));
    int   n1 = voxel_sample(X.ws, c.px.get<float>(), c.py.get<float>(), c.pz.get<float>(),
                            c.valid.get<int>(), c.nValid, sampleStep, c.boxMin, s1, cnt, s);
    int   n2 = voxel_sample(X.ws, c.px.get<float>(), c.py.get<float>(), c.pz.get<float>(),
                            c.valid.get<int>(), c.nValid, reSampleStep, c.boxMin, s2, cnt, s);
    if (!c.hasNormal) {
        float cellHint = sampleStep * 0.5f;
        estimate_normals(X.ws, n, c.px.get<float>(), c.py.get<float>(), c.pz.get<float>(),
                         c.nx.get<float>(), c.ny.get<float>(), c.nz.get<float>(),
                         c.valid.get<int>(), c.nValid, s1, n1, param.knnNormal, param.smoothNormal,
                         0, c.vp, cellHint, s);
        estimate_normals(X.ws, n, c.px.get<float>(), c.py.get<float>(), c.pz.get<float>(),
                         c.nx.get<float>(), c.ny.get<float>(), c.nz.get<float>(),
                         c.valid.get<int>(), c.nValid, s2, n2, param.knnNormal, param.smoothNormal,
                         0, c.vp, cellHint, s);
    }
    HostModel hm;
    hm.samplingDistanceRel = rel;
    hm.param               = param;
    DevBuf scratch;
    download_aos(s1, n1, c, false, hm.sampled.p, scratch, s);
    download_aos(s1, n1, c, true, hm.sampled.n, scratch, s);
    download_aos(s2, n2, c, false, hm.reSampled.p, scratch, s);
    download_aos(s2, n2, c, true, hm.reSampled.n, scratch, s);
    // extraIndices normalises the gathered normals (util.cpp:153)
    for (int i = 0; i < n1; i++)
        normalize3(&hm.sampled.n[3 * (size_t)i]);
    for (int i = 0; i < n2; i++)
        normalize3(&hm.reSampled.n[3 * (size_t)i]);
    for (HostCloud *hc : {&hm.sampled, &hm.reSampled}) {
        std::memcpy(hc->boxMin, c.boxMin, sizeof c.boxMin);
        std::memcpy(hc->boxMax, c.boxMax, sizeof c.boxMax);
    }
    // table on the device from the sampled model (uploaded as SoA)
    DevModel &dm = D.dm;
    model_upload(dm, n1, hm.sampled.p.data(), hm.sampled.n.data(), n2, hm.reSampled.p.data(),
                 hm.reSampled.n.data(), {}, std::vector<int>(1, 0), {}, {},
                 param.featAngleResolution, Dm, param.featDistanceStepRel, s);
    train_table(D.ts, n1, dm.sx.get<float>(), dm.sy.get<float>(), dm.sz.get<float>(),
                dm.snx.get<float>(), dm.sny.get<float>(), dm.snz.get<float>(), angleStep,
                distanceStep, D.refTail, hm.keys, hm.offs, hm.refs, hm.alphas, s);
    model_upload(dm, n1, hm.sampled.p.data(), hm.sampled.n.data(), n2, hm.reSampled.p.data(),
                 hm.reSampled.n.data(), hm.keys, hm.offs, hm.refs, hm.alphas,
                 param.featAngleResolution, Dm, param.featDistanceStepRel, s);
    D.hm      = std::move(hm);
    D.trained = true;
}

void upload_model(ppf_detector_s &D, MatchCtx &X) {
    HostModel &hm = D.hm;
    float      Dm = box_diameter(hm.sampled.boxMin, hm.sampled.boxMax);  // ppf.cpp:183
    model_upload(D.dm, hm.sampled.size(), hm.sampled.p.data(), hm.sampled.n.data(),
                 hm.reSampled.size(), hm.reSampled.p.data(), hm.reSampled.n.data(), hm.keys,
                 hm.offs, hm.refs, hm.alphas, hm.param.featAngleResolution, Dm,
                 hm.param.featDistanceStepRel, X.stream);
}

struct MatchOut {
    std::vector<std::pair<float, std::vector<float>>> results;  // score, pose 3x4
};

void write_result_cloud(MatchCtx &X, DevCloud &c, const int *idxDev, int m, float *pOut,
                        float *nOut, size_t cap, size_t &countOut) {
    countOut = (size_t)m;
    size_t w = std::min((size_t)m, cap);
    if (w == 0)
        return;
    std::vector<float> tmp;
    DevBuf             scratch;
    if (pOut) {
        download_aos(idxDev, (int)w, c, false, tmp, scratch, X.stream);
        std::memcpy(pOut, tmp.data(), sizeof(float) * 3 * w);
    }
    if (nOut) {
        download_aos(idxDev, (int)w, c, true, tmp, scratch, X.stream);
        for (size_t i = 0; i < w; i++)
            normalize3(&tmp[3 * i]);
        std::memcpy(nOut, tmp.data(), sizeof(float) * 3 * w);
    }
}

// the detector's lane l (created on first use)
MatchCtx::Lane &lane_of(MatchCtx &X, int l) {
    while ((int)X.lanes.size() <= l) {
        auto ln = std::make_unique<MatchCtx::Lane>();
        PPF_HIP_CHECK(hipStreamCreateWithFlags(&ln->stream, hipStreamNonBlocking));
        for (auto &e : ln->ev)
            PPF_HIP_CHECK(hipEventCreate(&e));
        X.lanes.push_back(std::move(ln));
    }
    return *X.lanes[l];
}

// the dense prep waits for the vote's pair kernel
constexpr bool PREP_AFTER_PAIRS = true;

// ends a helper's wait for the pair kernel however the vote ends (no pair
// kernel queued: 2), without overwriting the vote's own 1
struct PairRelease {
    std::atomic<int> &f;
    void operator()() {
        int z = 0;
        f.compare_exchange_strong(z, 2, std::memory_order_acq_rel);
    }
    ~PairRelease() { (*this)(); }
};

// joins a helper thread however the scope is left
struct JoinGuard {
    std::thread &t;
    ~JoinGuard() {
        if (t.joinable())
            t.join();
    }
};

void match_impl(ppf_detector_s &D, MatchCtx &X, DevCloud &c, float rel, float kpFrac, float minScore,
                const ppf_match_param &p, MatchOut &out, ppf_match_result *mr,
                ppf_match_stats *st) {
    // ppf.cpp:155-168
    if (!D.trained)
        throw std::runtime_error("No trained model in matchScene");
    if (rel > 1 || rel < 0)
        throw std::range_error("Invalid Input: samplingDistanceRel range mismatch in matchScene");
    if (kpFrac > 1 || kpFrac < 0)
        throw std::range_error("Invalid Input: keyPointFraction range mismatch in matchScene");
    if (minScore > 1 || minScore < 0)
        throw std::range_error("Invalid Input: minScore range mismatch in matchScene");
    if (c.n <= 0)
        throw std::runtime_error("Invalid Input: empty scene in matchScene");
    check_match_param(p);
    hipStream_t s   = X.stream;
    HostModel  &hm  = D.hm;
    DevModel   &dm  = D.dm;
    ppf_match_stats stl{};
    if (!st)
        st = &stl;
    std::memset(st, 0, sizeof *st);
    PPF_HIP_CHECK(hipEventRecord(X.ev[0], s));
    c.reset(s);
    // ppf.cpp:183-193
    const float Dm           = box_diameter(hm.sampled.boxMin, hm.sampled.boxMax);
    const float angleStep    = ppfc::kTwoPi / (float)hm.param.featAngleResolution;
    const int   refNum       = hm.sampled.size();
    const float reSampleStep = Dm * hm.param.poseRefRelSamplingDistance;
    const float sampleStep   = Dm * rel;  // :200
    const float normCell     = sampleStep * (float)D.normCell4 * 0.25f;
    // the dense samples' normals (queries 16x denser) measured best with a
    // cell of one sample step whatever the scene pass uses
    const float denseNormCell = sampleStep;
    // :223-240
    const float voteThreshold = (float)refNum * minScore;
    float maxOverlapDist = 0;
    if (p.maxOverlapDistRel > 0) maxOverlapDist = Dm * p.maxOverlapDistRel;
    if (p.maxOverlapDistAbs > 0) maxOverlapDist = p.maxOverlapDistAbs;
    float poseRefDist = 0;
    if (p.poseRefDistThresholdRel > 0) poseRefDist = Dm * p.poseRefDistThresholdRel;
    if (p.poseRefDistThresholdAbs > 0) poseRefDist = p.poseRefDistThresholdAbs;
    float scoringDist = 0;
    if (p.poseRefScoringDistRel > 0) scoringDist = Dm * p.poseRefScoringDistRel;
    if (p.poseRefScoringDistAbs > 0) scoringDist = p.poseRefScoringDistAbs;

    float *PX = c.px.get<float>(), *PY = c.py.get<float>(), *PZ = c.pz.get<float>();
    float *NX = c.nx.get<float>(), *NY = c.ny.get<float>(), *NZ = c.nz.get<float>();
    int   *cnt = X.ws.smallCnt.as<int>(8);
    int   *samp = X.idxA.as<int>(std::max(c.nValid, 1));
    // scene sample1 (ppf.cpp:199-203)
    int ns = voxel_sample(X.ws, PX, PY, PZ, c.valid.get<int>(), c.nValid, sampleStep, c.boxMin,
                          samp, cnt, s);
    PPF_HIP_CHECK(hipEventRecord(X.ev[1], s));
    // normals (ppf.cpp:204-211)
    if (!c.hasNormal)
        estimate_normals(X.ws, c.n, PX, PY, PZ, NX, NY, NZ, c.valid.get<int>(), c.nValid, samp, ns,
                         p.knnNormal, p.smoothNormal, p.invertNormal, c.vp, normCell, s, c.ext);
    PPF_HIP_CHECK(hipEventRecord(X.ev[2], s));
    // key points (ppf.cpp:212-216)
    const float keySampleStep = std::sqrt(1.0f / kpFrac) * sampleStep;
    int *kp = X.idxB.as<int>(std::max(ns, 1));
    int  nk = voxel_sample(X.ws, PX, PY, PZ, samp, ns, keySampleStep, c.boxMin, kp, cnt, s);
    PPF_HIP_CHECK(hipEventRecord(X.ev[3], s));
    if (mr) {  // ppf.cpp:242-245
        write_result_cloud(X, c, kp, nk, mr->keyPoints, mr->keyNormals, mr->keyCap, mr->nKey);
        write_result_cloud(X, c, samp, ns, mr->sampledPoints, mr->sampledNormals, mr->sampledCap,
                           mr->nSampled);
    }
    // vote (ppf.cpp:247-411)
    grid_build(X.ws.gVote, PX, PY, PZ, NX, NY, NZ, samp, ns, Dm / (float)D.vcfg.cellDiv, s, c.ext,
               true);
    // dense prep (ppf.cpp:427-434) on a lane, overlapping the vote and the
    // clustering: it samples the scene again and estimates normals only for
    // points without one (so never a sampled point, which all have one by
    // now) and writes only those points, while the vote reads its grid's
    // own copies and the key points, all sampled points.  The refinement's
    // grids (their own buffers, over the final positions) follow on the
    // same lane.
    int        *dense = X.idxC.as<int>(std::max(c.nValid, 1));
    int         nd    = 0;
    float       prepMs = 0.0f;
    int         np    = 0;
    std::thread prep;
    std::exception_ptr prepErr;
    JoinGuard   prepJoin{prep};
    // released after the vote, or on the way out if anything throws first
    // (before the join above)
    PairRelease pairRelease{X.ws.pairQueued};
    const bool laneGrids = p.sparsePoseRefinement != 0;
    if (p.densePoseRefinement || laneGrids) {
        MatchCtx::Lane &ln = lane_of(X, 0);
        PPF_HIP_CHECK(hipEventRecord(X.ev[12], s));
        PPF_HIP_CHECK(hipStreamWaitEvent(ln.stream, X.ev[12], 0));
        X.ws.pairQueued.store(0, std::memory_order_relaxed);
        int dev = 0;
        PPF_HIP_CHECK(hipGetDevice(&dev));
        prep = std::thread([&, dev] {
            try {
                PPF_HIP_CHECK(hipSetDevice(dev));
                // start after the pair kernel (the VALU-bound stage), beside
                // the LDS-bound hit voting and the clustering
                if (PREP_AFTER_PAIRS) {
                    int q;
                    while ((q = X.ws.pairQueued.load(std::memory_order_acquire)) == 0)
                        std::this_thread::yield();
                    if (q == 1)
                        PPF_HIP_CHECK(hipStreamWaitEvent(ln.stream, X.ws.voteEv[2], 0));
                }
                PPF_HIP_CHECK(hipEventRecord(ln.ev[0], ln.stream));
                if (p.densePoseRefinement) {
                    nd = voxel_sample(ln.ws, PX, PY, PZ, c.valid.get<int>(), c.nValid,
                                      reSampleStep, c.boxMin, dense, ln.ws.smallCnt.as<int>(8),
                                      ln.stream);
                    if (!c.hasNormal)
                        estimate_normals(ln.ws, c.n, PX, PY, PZ, NX, NY, NZ, c.valid.get<int>(),
                                         c.nValid, dense, nd, p.knnNormal, 0, p.invertNormal,
                                         c.vp, denseNormCell, ln.stream, c.ext);
                }
                if (laneGrids) {
                    grid_build(X.ws.gSparse, PX, PY, PZ, NX, NY, NZ, samp, ns,
                               sampleStep * kSparseCell, ln.stream, c.ext, true);
                    grid_build(X.ws.gFull, PX, PY, PZ, nullptr, nullptr, nullptr,
                               c.valid.get<int>(), c.nValid,
                               std::max(scoringDist * 2.0f, sampleStep * 0.25f), ln.stream, c.ext,
                               true);
                    if (p.densePoseRefinement)
                        grid_build(X.ws.gDense, PX, PY, PZ, NX, NY, NZ, dense, nd,
                                   reSampleStep * kDenseCell, ln.stream, c.ext, true);
                }
                PPF_HIP_CHECK(hipEventRecord(ln.ev[1], ln.stream));
                PPF_HIP_CHECK(hipStreamSynchronize(ln.stream));
                PPF_HIP_CHECK(hipEventElapsedTime(&prepMs, ln.ev[0], ln.ev[1]));
            } catch (...) {
                prepErr = std::current_exception();
            }
        });
    }
    VoteStats vs{};
    PPF_HIP_CHECK(hipEventRecord(X.ev[4], s));
    np = vote_and_poses(X.ws, dm, X.ws.gVote.view, PX, PY, PZ, NX, NY, NZ, kp, nk,
                        voteThreshold, D.vcfg, &vs, nullptr, 0, s);
    pairRelease();
    PPF_HIP_CHECK(hipEventRecord(X.ev[5], s));
    // cluster (ppf.cpp:413-420)
    std::vector<Hypo> hyp;
    float center[3];
    for (int a = 0; a < 3; a++)
        center[a] = (hm.sampled.boxMin[a] + hm.sampled.boxMax[a]) / 2.0f;
    int nClusters = 0;
    cluster_poses(X.ws, X.ws.poses.get<PoseRec>(), np, 0.1f * Dm, angleStep, center,
                  maxOverlapDist, hyp, &nClusters, s);
    PPF_HIP_CHECK(hipEventRecord(X.ev[6], s));
    // the dense prep has finished before the refinement reads the scene
    if (prep.joinable())
        prep.join();
    if (prepErr)
        std::rethrow_exception(prepErr);
    PPF_HIP_CHECK(hipEventRecord(X.ev[7], s));
    // icp (ppf.cpp:436-495)
    std::vector<int> cand;
    for (int i = 0; i < (int)hyp.size(); i++)
        if (!(hyp[i].votes < voteThreshold))
            cand.push_back(i);
    std::vector<float> poseNow(hyp.size() * 12), scoreNow(hyp.size());
    std::vector<char>  alive(hyp.size(), 0);
    for (int i : cand) {
        std::memcpy(&poseNow[12 * (size_t)i], hyp[i].pose, sizeof(float) * 12);
        scoreNow[i] = hyp[i].votes;
        alive[i]    = 1;
    }
    int nSparse = 0, nDenseIcp = 0;
    bool needGrids = p.sparsePoseRefinement && !cand.empty();
    if (needGrids) {
        // gSparse and gFull were built on the lane (laneGrids)
        IcpCriteria sc{5, poseRefDist, sampleStep * 0.5f, sampleStep, 0.05f};  // ppf.cpp:423
        std::vector<float>        init(cand.size() * 12);
        std::vector<IcpResultDev> res(cand.size());
        for (size_t k = 0; k < cand.size(); k++)
            std::memcpy(&init[12 * k], hyp[cand[k]].pose, sizeof(float) * 12);
        PPF_HIP_CHECK(hipEventRecord(X.ev[9], s));
        icp_batch(X.ws, refNum, dm.sx.get<float>(), dm.sy.get<float>(), dm.sz.get<float>(),
                  dm.snx.get<float>(), dm.sny.get<float>(), dm.snz.get<float>(),
                  X.ws.gSparse.view, X.ws.gFull.view, init.data(), (int)cand.size(), sc,
                  scoringDist, D.icfg, res.data(), s);
        PPF_HIP_CHECK(hipEventRecord(X.ev[10], s));
        nSparse = (int)cand.size();
        std::vector<int> survivors;
        for (size_t k = 0; k < cand.size(); k++) {
            int i = cand[k];
            if (!res[k].converged) {
                alive[i] = 0;
                continue;
            }
            std::memcpy(&poseNow[12 * (size_t)i], res[k].pose, sizeof(float) * 12);
            float score = (float)res[k].inliers / (float)refNum;
            if (score > 1.0f)
                score = 1.0f;
            scoreNow[i] = score;
            if (score < minScore) {
                alive[i] = 0;
                continue;
            }
            survivors.push_back(i);
        }
        if (p.densePoseRefinement && !survivors.empty()) {
            // gDense was built on the lane
            IcpCriteria dc{p.poseRefNumSteps, poseRefDist, reSampleStep * 0.5f, sampleStep, 0.05f};
            std::vector<float>        init2(survivors.size() * 12);
            std::vector<IcpResultDev> res2(survivors.size());
            for (size_t k = 0; k < survivors.size(); k++)
                std::memcpy(&init2[12 * k], &poseNow[12 * (size_t)survivors[k]], sizeof(float) * 12);
            IcpConfig dcfg = D.icfg;
            PPF_HIP_CHECK(hipEventRecord(X.ev[11], s));
            icp_batch(X.ws, hm.reSampled.size(), dm.rx.get<float>(), dm.ry.get<float>(),
                      dm.rz.get<float>(), dm.rnx.get<float>(), dm.rny.get<float>(),
                      dm.rnz.get<float>(), X.ws.gDense.view, X.ws.gFull.view, init2.data(),
                      (int)survivors.size(), dc, scoringDist, dcfg, res2.data(), s);
            PPF_HIP_CHECK(hipEventRecord(X.ev[12], s));
            nDenseIcp = (int)survivors.size();
            for (size_t k = 0; k < survivors.size(); k++) {
                int i = survivors[k];
                if (!res2[k].converged) {
                    alive[i] = 0;
                    continue;
                }
                std::memcpy(&poseNow[12 * (size_t)i], res2[k].pose, sizeof(float) * 12);
                float score = (float)res2[k].inliers / (float)hm.reSampled.size();
                if (score > 1.0f)
                    score = 1.0f;
                scoreNow[i] = score;
            }
        }
    }
    // ordered accept walk (ppf.cpp:439-495)
    for (int i : cand) {
        if (!alive[i])
            continue;
        float score = scoreNow[i];
        if ((p.sparsePoseRefinement || p.densePoseRefinement) && score < minScore)
            continue;
        out.results.emplace_back(score, std::vector<float>(&poseNow[12 * (size_t)i],
                                                           &poseNow[12 * (size_t)i] + 12));
        if ((int)out.results.size() >= p.numMatches)
            break;
    }
    // ppf.cpp:498-499 (stable on ties)
    std::stable_sort(out.results.begin(), out.results.end(),
                     [](const auto &a, const auto &b) { return a.first > b.first; });
    PPF_HIP_CHECK(hipEventRecord(X.ev[8], s));
    PPF_HIP_CHECK(hipEventSynchronize(X.ev[8]));
    st->nScene          = c.n;
    st->nValid          = c.nValid;
    st->nSampled        = ns;
    st->nKeypoints      = nk;
    st->nVotedKeypoints = vs.voted;
    st->nPoses          = np;
    st->nClusters       = nClusters;
    st->nHypotheses     = (int)hyp.size();
    st->nSparseIcp      = nSparse;
    st->nDenseIcp       = nDenseIcp;
    st->nResults        = (int)out.results.size();
    st->nDense          = nd;
    st->nPairs          = vs.pairs;
    st->nKeyHits        = vs.hits;
    st->nVotes          = vs.votes;
    st->msSample        = elapsed(X.ev[0], X.ev[1]);
    st->msNormals       = elapsed(X.ev[1], X.ev[2]);
    st->msKeypoints     = elapsed(X.ev[2], X.ev[3]);
    st->msVote          = elapsed(X.ev[3], X.ev[5]);
    st->msVoteKernel    = vs.kernelMs;
    st->msPairKernel    = vs.pairMs;
    st->msHitsKernel    = vs.hitsMs;
    st->nVoteFallback   = vs.overflow;
    st->msCluster       = elapsed(X.ev[5], X.ev[6]);
    st->msIcpPrepare    = prepMs;  // overlapped with the vote and the clustering
    st->msIcp           = elapsed(X.ev[7], X.ev[8]);
    st->msTotal         = elapsed(X.ev[0], X.ev[8]);
    st->msIcpSparse     = nSparse ? elapsed(X.ev[9], X.ev[10]) : 0.0f;
    st->msIcpDense      = nDenseIcp ? elapsed(X.ev[11], X.ev[12]) : 0.0f;
}

// Two passes of the lanes over the scenes, in the background of the
// calling thread: f(i, workspace, stream) for every scene, then, once
// open() is called, g(i, ...) for every scene.  join() waits for both and
// rethrows a lane's error; leaving the scope without open() skips g.
struct LanePasses {
    using Fn = std::function<void(int, Workspace &, hipStream_t)>;
    using AfterFn = std::function<void(Workspace &, hipStream_t)>;
    MatchCtx                       &X;
    int                             n;
    Fn                              f, g;
    AfterFn                         afterF;  // once every lane's f work is done (may be empty)
    std::atomic<int>                next1{0}, next2{0}, gate{0};  // gate: 0 wait, 1 go, -1 skip
    std::atomic<int>                doneF{0};
    int                             nLanes = 0;
    std::vector<std::thread>        th;
    std::vector<std::exception_ptr> err;

    LanePasses(MatchCtx &X_, int n_, Fn f_, Fn g_, AfterFn a_ = {})
        : X(X_), n(n_), f(std::move(f_)), g(std::move(g_)), afterF(std::move(a_)) {
        PPF_HIP_CHECK(hipStreamSynchronize(X.stream));
        const int L = std::max(1, std::min(n, MatchCtx::kLanes));
        lane_of(X, L - 1);
        nLanes = L;
        int dev = 0;
        PPF_HIP_CHECK(hipGetDevice(&dev));
        err.resize(L);
        try {
            start(L, dev);
        } catch (...) {
            // a thread that could not be started: the started ones skip the
            // second pass and are joined before the error leaves
            int z = 0;
            gate.compare_exchange_strong(z, -1);
            for (auto &t : th)
                if (t.joinable())
                    t.join();
            throw;
        }
    }
    void start(int L, int dev) {
        for (int l = 0; l < L; l++)
            th.emplace_back([this, l, dev] {
                try {
                    PPF_HIP_CHECK(hipSetDevice(dev));
                    MatchCtx::Lane &ln = *X.lanes[l];
                    for (int i = next1++; i < n; i = next1++)
                        f(i, ln.ws, ln.stream);
                    if (afterF) {
                        // the last lane to finish queuing its f work waits
                        // on every lane's stream, then queues afterF on its
                        // own, with the context's own workspace for it (so
                        // its buffers stay allocated whichever lane runs it)
                        PPF_HIP_CHECK(hipEventRecord(ln.ev[2], ln.stream));
                        if (doneF.fetch_add(1, std::memory_order_acq_rel) + 1 == nLanes) {
                            for (int o = 0; o < nLanes; o++)
                                PPF_HIP_CHECK(hipStreamWaitEvent(ln.stream, X.lanes[o]->ev[2], 0));
                            afterF(X.afterWs, ln.stream);
                        }
                    }
                    int q;
                    while ((q = gate.load(std::memory_order_acquire)) == 0)
                        std::this_thread::yield();
                    if (q == 1)
                        for (int i = next2++; i < n; i = next2++)
                            g(i, ln.ws, ln.stream);
                    PPF_HIP_CHECK(hipStreamSynchronize(ln.stream));
                } catch (...) {
                    err[l] = std::current_exception();
                }
            });
    }
    void open() { gate.store(1, std::memory_order_release); }
    void join() {
        int z = 0;
        gate.compare_exchange_strong(z, -1);
        for (auto &t : th)
            if (t.joinable())
                t.join();
        for (auto &e : err)
            if (e)
                std::rethrow_exception(e);
    }
    ~LanePasses() {
        int z = 0;
        gate.compare_exchange_strong(z, -1);
        for (auto &t : th)
            if (t.joinable())
                t.join();
    }
};

// Several scenes matched together (same parameters): the scenes' sampling,
// normals, key points and grids run as one launch sequence for all scenes
// (each scene in its own slice of the shared tables), the key points of all
// scenes are voted in one launch (each key point searches its own scene's
// grid), the poses are clustered per scene, and the hypotheses of all
// scenes are refined in one batched refinement per stage (each against its
// own scene).  Every per-scene result equals match_impl's on that scene.
void match_batch_impl(ppf_detector_s &D, MatchCtx &X, const std::vector<DevCloud *> &clouds, float rel,
                      float kpFrac, float minScore, const ppf_match_param &p,
                      std::vector<MatchOut> &outs, ppf_match_stats *st) {
    if (!D.trained)
        throw std::runtime_error("No trained model in matchScene");
    if (rel > 1 || rel < 0)
        throw std::range_error("Invalid Input: samplingDistanceRel range mismatch in matchScene");
    if (kpFrac > 1 || kpFrac < 0)
        throw std::range_error("Invalid Input: keyPointFraction range mismatch in matchScene");
    if (minScore > 1 || minScore < 0)
        throw std::range_error("Invalid Input: minScore range mismatch in matchScene");
    for (DevCloud *c : clouds)
        if (c->n <= 0)
            throw std::runtime_error("Invalid Input: empty scene in matchScene");
    check_match_param(p);
    const int   n  = (int)clouds.size();
    hipStream_t s  = X.stream;
    HostModel  &hm = D.hm;
    DevModel   &dm = D.dm;
    ppf_match_stats stl{};
    if (!st)
        st = &stl;
    std::memset(st, 0, sizeof *st);
    outs.assign(n, MatchOut{});
    while ((int)X.sceneWork.size() < n)
        X.sceneWork.emplace_back(new MatchCtx::SceneWork());
    // ppf.cpp:183-240 (the same derived parameters for every scene)
    const float Dm           = box_diameter(hm.sampled.boxMin, hm.sampled.boxMax);
    const float angleStep    = ppfc::kTwoPi / (float)hm.param.featAngleResolution;
    const int   refNum       = hm.sampled.size();
    const float reSampleStep = Dm * hm.param.poseRefRelSamplingDistance;
    const float sampleStep   = Dm * rel;
    const float normCell     = sampleStep * (float)D.normCell4 * 0.25f;
    // the dense samples' normals (queries 16x denser) measured best with a
    // cell of one sample step whatever the scene pass uses
    const float denseNormCell = sampleStep;
    const float voteThreshold = (float)refNum * minScore;
    float maxOverlapDist = 0;
    if (p.maxOverlapDistRel > 0) maxOverlapDist = Dm * p.maxOverlapDistRel;
    if (p.maxOverlapDistAbs > 0) maxOverlapDist = p.maxOverlapDistAbs;
    float poseRefDist = 0;
    if (p.poseRefDistThresholdRel > 0) poseRefDist = Dm * p.poseRefDistThresholdRel;
    if (p.poseRefDistThresholdAbs > 0) poseRefDist = p.poseRefDistThresholdAbs;
    float scoringDist = 0;
    if (p.poseRefScoringDistRel > 0) scoringDist = Dm * p.poseRefScoringDistRel;
    if (p.poseRefScoringDistAbs > 0) scoringDist = p.poseRefScoringDistAbs;
    const float keySampleStep = std::sqrt(1.0f / kpFrac) * sampleStep;

    PPF_HIP_CHECK(hipEventRecord(X.ev[0], s));
    // ---- every scene's sampled points in one launch sequence (ppf.cpp:199-203)
    std::vector<VoxSceneIn> vin(n);
    std::vector<int>        cnts;
    for (int i = 0; i < n; i++) {
        DevCloud &c  = *clouds[i];
        auto     &sw = *X.sceneWork[i];
        c.reset(s);
        vin[i] = VoxSceneIn{c.px.get<float>(), c.py.get<float>(), c.pz.get<float>(),
                            c.valid.get<int>(), c.nValid, {c.boxMin[0], c.boxMin[1], c.boxMin[2]},
                            sw.samp.as<int>(std::max(c.nValid, 1))};
    }
    voxel_sample_multi(X.ws, vin, sampleStep, cnts, s);
    for (int i = 0; i < n; i++)
        X.sceneWork[i]->ns = cnts[i];
    // ---- the normals of every scene's sampled points (ppf.cpp:204-211)
    auto normalsOf = [&](auto query, float cell, int smooth, Workspace &ws, hipStream_t ls) {
        std::vector<NormSceneIn> ns;
        for (int i = 0; i < n; i++) {
            DevCloud &c = *clouds[i];
            if (c.hasNormal)
                continue;
            NormSceneIn in{c.px.get<float>(), c.py.get<float>(), c.pz.get<float>(),
                           c.nx.get<float>(), c.ny.get<float>(), c.nz.get<float>(), c.n,
                           c.valid.get<int>(), c.nValid, nullptr, 0, {}, {}};
            query(i, in);
            std::memcpy(in.vp, c.vp, sizeof in.vp);
            std::memcpy(in.box, c.ext, sizeof in.box);
            ns.push_back(in);
        }
        estimate_normals_multi(ws, X.batchNormal, ns, p.knnNormal, smooth, p.invertNormal, cell,
                               ls);
    };
    normalsOf(
        [&](int i, NormSceneIn &in) {
            in.query = X.sceneWork[i]->samp.get<int>();
            in.nq    = X.sceneWork[i]->ns;
        },
        normCell, p.smoothNormal, X.ws, s);
    // ---- every scene's key points and vote grid together (ppf.cpp:212-216)
    std::vector<GridSceneIn> gin(n);
    for (int i = 0; i < n; i++) {
        DevCloud &c  = *clouds[i];
        auto     &sw = *X.sceneWork[i];
        vin[i]       = VoxSceneIn{c.px.get<float>(), c.py.get<float>(), c.pz.get<float>(),
                            sw.samp.get<int>(), sw.ns, {c.boxMin[0], c.boxMin[1], c.boxMin[2]},
                            sw.kp.as<int>(std::max(sw.ns, 1))};
        gin[i]       = GridSceneIn{c.px.get<float>(), c.py.get<float>(), c.pz.get<float>(),
                             c.nx.get<float>(), c.ny.get<float>(), c.nz.get<float>(),
                             sw.samp.get<int>(), sw.ns, {}};
        std::memcpy(gin[i].box, c.ext, sizeof c.ext);
    }
    voxel_sample_multi(X.ws, vin, keySampleStep, cnts, s);
    for (int i = 0; i < n; i++)
        X.sceneWork[i]->nk = cnts[i];
    std::vector<GridView> voteViews(n);
    if (grid_build_multi(X.batchVote, gin, Dm / (float)D.vcfg.cellDiv, true, s)) {
        voteViews = X.batchVote.views;
    } else {
        for (int i = 0; i < n; i++) {
            DevCloud &c  = *clouds[i];
            auto     &sw = *X.sceneWork[i];
            grid_build(sw.gVote, c.px.get<float>(), c.py.get<float>(), c.pz.get<float>(),
                       c.nx.get<float>(), c.ny.get<float>(), c.nz.get<float>(), sw.samp.get<int>(),
                       sw.ns, Dm / (float)D.vcfg.cellDiv, s, c.ext, true);
            voteViews[i] = sw.gVote.view;
        }
    }
    std::vector<int> ko(n + 1, 0);
    float            minCell = FLT_MAX;
    for (int i = 0; i < n; i++) {
        minCell   = std::min(minCell, voteViews[i].cell);
        ko[i + 1] = ko[i] + X.sceneWork[i]->nk;
    }
    // every scene's key points gathered before the lanes start (a point
    // whose normal the sampled pass left undefined can be given one, and
    // moved, by its scene's dense pass)
    const int    K = ko[n];
    VoteStats    vs{};
    int          np = 0;
    float *kx = nullptr, *ky = nullptr, *kz = nullptr, *knx = nullptr, *kny = nullptr,
          *knz = nullptr;
    int   *kpScene = nullptr, *kpId = nullptr;
    GridView *gv = nullptr;
    std::vector<GridView> views(n);
    if (K > 0) {
        kx  = X.ws.kpX.as<float>(K);
        ky  = X.ws.kpY.as<float>(K);
        kz  = X.ws.kpZ.as<float>(K);
        knx = X.ws.kpNX.as<float>(K);
        kny = X.ws.kpNY.as<float>(K);
        knz = X.ws.kpNZ.as<float>(K);
        kpScene = X.batchKpScene.as<int>(K);
        kpId    = X.batchKpId.as<int>(K);
        std::vector<KpSceneIn> kin(n);
        for (int i = 0; i < n; i++) {
            DevCloud &c = *clouds[i];
            views[i]    = voteViews[i];
            kin[i]      = KpSceneIn{X.sceneWork[i]->kp.get<int>(), c.px.get<float>(),
                               c.py.get<float>(), c.pz.get<float>(), c.nx.get<float>(),
                               c.ny.get<float>(), c.nz.get<float>()};
        }
        gather_keypoints_multi(kin, ko, X.batchKpDesc, kx, ky, kz, knx, kny, knz, kpScene, kpId, s);
        gv = X.batchGrids.as<GridView>(n);
        PPF_HIP_CHECK(hipMemcpyAsync(gv, views.data(), sizeof(GridView) * n,
                                     hipMemcpyHostToDevice, s));
    }
    PPF_HIP_CHECK(hipEventRecord(X.ev[3], s));
    // ---- per scene, on the lanes while the vote and the clustering
    // (ppf.cpp:413-420) run: the refinement's dense samples with their
    // normals (ppf.cpp:427-434) and its grids over the final positions, as
    // the single-scene match prepares them beside its vote (they write only
    // points without a normal, and the vote reads the scene only through its
    // grids' copies and the key points gathered above)
    const bool laneGrids = p.sparsePoseRefinement != 0;
    float center[3];
    for (int a = 0; a < 3; a++)
        center[a] = (hm.sampled.boxMin[a] + hm.sampled.boxMax[a]) / 2.0f;
    std::vector<std::vector<Hypo>> hyp(n);
    std::vector<int>               ncs(n, 0);
    std::vector<int>               pOff(n + 1, 0);
    for (int i = 0; i < n; i++)
        X.sceneWork[i]->nd = 0;
    // the refinement grids of every scene over the final positions, queued
    // once every lane's normals are: together when they fit one build,
    // else scene by scene
    std::vector<GridView> sparseViews(n), fullViews(n), denseViews(n);
    auto refineGrids = [&](hipStream_t ls) {
        std::vector<GridSceneIn> gs(n), gf(n), gd(n);
        for (int i = 0; i < n; i++) {
            DevCloud &c  = *clouds[i];
            auto     &sw = *X.sceneWork[i];
            float *PX = c.px.get<float>(), *PY = c.py.get<float>(), *PZ = c.pz.get<float>();
            float *NX = c.nx.get<float>(), *NY = c.ny.get<float>(), *NZ = c.nz.get<float>();
            gs[i] = GridSceneIn{PX, PY, PZ, NX, NY, NZ, sw.samp.get<int>(), sw.ns, {}};
            gf[i] = GridSceneIn{PX, PY, PZ, nullptr, nullptr, nullptr, c.valid.get<int>(), c.nValid, {}};
            gd[i] = GridSceneIn{PX, PY, PZ, NX, NY, NZ, sw.dense.get<int>(), sw.nd, {}};
            std::memcpy(gs[i].box, c.ext, sizeof c.ext);
            std::memcpy(gf[i].box, c.ext, sizeof c.ext);
            std::memcpy(gd[i].box, c.ext, sizeof c.ext);
        }
        const float fullCell = std::max(scoringDist * 2.0f, sampleStep * 0.25f);
        if (grid_build_multi(X.batchSparse, gs, sampleStep * kSparseCell, true, ls))
            sparseViews = X.batchSparse.views;
        else
            for (int i = 0; i < n; i++) {
                grid_build(X.sceneWork[i]->gSparse, gs[i].X, gs[i].Y, gs[i].Z, gs[i].NX, gs[i].NY,
                           gs[i].NZ, gs[i].subset, gs[i].m, sampleStep * kSparseCell, ls,
                           clouds[i]->ext, true);
                sparseViews[i] = X.sceneWork[i]->gSparse.view;
            }
        // the scoring grid is read by radius tests alone: float4 copies only
        if (grid_build_multi(X.batchFull, gf, fullCell, true, ls, false))
            fullViews = X.batchFull.views;
        else
            for (int i = 0; i < n; i++) {
                grid_build(X.sceneWork[i]->gFull, gf[i].X, gf[i].Y, gf[i].Z, nullptr, nullptr,
                           nullptr, gf[i].subset, gf[i].m, fullCell, ls, clouds[i]->ext, true);
                fullViews[i] = X.sceneWork[i]->gFull.view;
            }
        if (!p.densePoseRefinement)
            return;
        if (grid_build_multi(X.batchDense, gd, reSampleStep * kDenseCell, true, ls))
            denseViews = X.batchDense.views;
        else
            for (int i = 0; i < n; i++) {
                grid_build(X.sceneWork[i]->gDense, gd[i].X, gd[i].Y, gd[i].Z, gd[i].NX, gd[i].NY,
                           gd[i].NZ, gd[i].subset, gd[i].m, reSampleStep * kDenseCell, ls,
                           clouds[i]->ext, true);
                denseViews[i] = X.sceneWork[i]->gDense.view;
            }
    };
    // on one lane: every scene's dense-sample normals, then the refinement
    // grids over the final positions
    auto laneWork = [&](Workspace &ws, hipStream_t ls) {
        if (p.densePoseRefinement) {
            // the dense samples (the positions of the sampled points are
            // final by now), then their normals
            std::vector<VoxSceneIn> din(n);
            std::vector<int>        dc;
            for (int i = 0; i < n; i++) {
                DevCloud &c = *clouds[i];
                din[i]      = VoxSceneIn{c.px.get<float>(), c.py.get<float>(), c.pz.get<float>(),
                                    c.valid.get<int>(), c.nValid,
                                    {c.boxMin[0], c.boxMin[1], c.boxMin[2]},
                                    X.sceneWork[i]->dense.as<int>(std::max(c.nValid, 1))};
            }
            voxel_sample_multi(ws, din, reSampleStep, dc, ls);
            for (int i = 0; i < n; i++)
                X.sceneWork[i]->nd = dc[i];
            normalsOf(
                [&](int i, NormSceneIn &in) {
                    in.query = X.sceneWork[i]->dense.get<int>();
                    in.nq    = X.sceneWork[i]->nd;
                },
                denseNormCell, 0, ws, ls);
        }
        if (laneGrids)
            refineGrids(ls);
    };
    LanePasses lanes(
        X, n, [](int, Workspace &, hipStream_t) {},
        [&](int i, Workspace &ws, hipStream_t ls) {
            // scenes with long pose lists: clustered one by one on the lanes
            // (the vote's pose buffer may have grown during the vote)
            cluster_poses(ws, X.ws.poses.get<PoseRec>() + pOff[i], pOff[i + 1] - pOff[i],
                          0.1f * Dm, angleStep, center, maxOverlapDist, hyp[i], &ncs[i], ls);
        },
        laneWork);
    // ---- every scene's key points in one vote
    if (K > 0) {
        VoteKeypoints in;
        in.g        = views[0];
        in.gridsDev = gv;
        in.kpScene  = kpScene;
        in.minCell  = minCell;
        in.kx = kx;
        in.ky = ky;
        in.kz = kz;
        in.knx = knx;
        in.kny = kny;
        in.knz = knz;
        in.kp  = kpId;
        in.nk  = K;
        in.nScenes = n;
        np = vote_keypoints(X.ws, dm, in, voteThreshold, D.vcfg, &vs, nullptr, 0, s);
        // poses are compacted in key point order: scene i's are those of its key points
        std::vector<int> offs(K + 1);
        PPF_HIP_CHECK(hipMemcpyAsync(offs.data(), X.ws.voteCnt.get<int>() + K + 1,
                                     sizeof(int) * (K + 1), hipMemcpyDeviceToHost, s));
        PPF_HIP_CHECK(hipStreamSynchronize(s));
        for (int i = 0; i <= n; i++)
            pOff[i] = offs[ko[i]];
    }
    PPF_HIP_CHECK(hipEventRecord(X.ev[5], s));
    // every scene's clustering in one pass over the vote's poses, beside the
    // lanes' refinement preparation, when every scene's list is short enough;
    // else scene by scene on the lanes after it
    if (cluster_multi_fits(pOff)) {
        cluster_poses_multi(X.ws, X.ws.poses.get<PoseRec>(), pOff, 0.1f * Dm, angleStep, center,
                            maxOverlapDist, hyp, ncs, s);
    } else {
        lanes.open();
    }
    lanes.join();
    int nClustersTot = 0;
    for (int i = 0; i < n; i++)
        nClustersTot += ncs[i];
    PPF_HIP_CHECK(hipEventRecord(X.ev[6], s));
    // ---- refinement (ppf.cpp:423-495), every scene's hypotheses together
    struct Cand {
        int scene, h;
    };
    std::vector<Cand>               cand;
    std::vector<std::vector<float>> poseNow(n), scoreNow(n);
    std::vector<std::vector<char>>  alive(n);
    for (int i = 0; i < n; i++) {
        poseNow[i].assign(hyp[i].size() * 12, 0.0f);
        scoreNow[i].assign(hyp[i].size(), 0.0f);
        alive[i].assign(hyp[i].size(), 0);
        for (int h = 0; h < (int)hyp[i].size(); h++)
            if (!(hyp[i][h].votes < voteThreshold)) {
                cand.push_back({i, h});
                std::memcpy(&poseNow[i][12 * (size_t)h], hyp[i][h].pose, sizeof(float) * 12);
                scoreNow[i][h] = hyp[i][h].votes;
                alive[i][h]    = 1;
            }
    }
    std::vector<char> sceneHasCand(n, 0);
    for (auto &c : cand)
        sceneHasCand[c.scene] = 1;
    int nSparse = 0, nDenseIcp = 0, nDenseTot = 0;
    if (p.sparsePoseRefinement && !cand.empty()) {
        std::vector<GridView> &vs1 = sparseViews, &vf = fullViews, &vd = denseViews;
        for (int i = 0; i < n; i++)
            if (p.densePoseRefinement && sceneHasCand[i])
                nDenseTot += X.sceneWork[i]->nd;
        GridView *gS = X.batchGrids.as<GridView>((size_t)3 * n);
        PPF_HIP_CHECK(hipMemcpyAsync(gS, vs1.data(), sizeof(GridView) * n, hipMemcpyHostToDevice, s));
        PPF_HIP_CHECK(hipMemcpyAsync(gS + n, vf.data(), sizeof(GridView) * n, hipMemcpyHostToDevice,
                                     s));
        if (p.densePoseRefinement)
            PPF_HIP_CHECK(hipMemcpyAsync(gS + 2 * n, vd.data(), sizeof(GridView) * n,
                                         hipMemcpyHostToDevice, s));
        std::vector<float> init(cand.size() * 12);
        std::vector<int>   hs(cand.size());
        for (size_t q = 0; q < cand.size(); q++) {
            std::memcpy(&init[12 * q], hyp[cand[q].scene][cand[q].h].pose, sizeof(float) * 12);
            hs[q] = cand[q].scene;
        }
        int *hsD = X.batchHypScene.as<int>(cand.size());
        PPF_HIP_CHECK(hipMemcpyAsync(hsD, hs.data(), sizeof(int) * hs.size(), hipMemcpyHostToDevice,
                                     s));
        const int first = cand[0].scene;
        IcpCriteria sc{5, poseRefDist, sampleStep * 0.5f, sampleStep, 0.05f};  // ppf.cpp:423
        std::vector<IcpResultDev> res(cand.size());
        icp_batch(X.ws, refNum, dm.sx.get<float>(), dm.sy.get<float>(), dm.sz.get<float>(),
                  dm.snx.get<float>(), dm.sny.get<float>(), dm.snz.get<float>(), vs1[first],
                  vf[first], init.data(), (int)cand.size(), sc, scoringDist, D.icfg, res.data(), s,
                  gS, gS + n, hsD);
        nSparse = (int)cand.size();
        std::vector<Cand> survivors;
        for (size_t q = 0; q < cand.size(); q++) {
            const int i = cand[q].scene, h = cand[q].h;
            if (!res[q].converged) {
                alive[i][h] = 0;
                continue;
            }
            std::memcpy(&poseNow[i][12 * (size_t)h], res[q].pose, sizeof(float) * 12);
            float score = (float)res[q].inliers / (float)refNum;
            if (score > 1.0f)
                score = 1.0f;
            scoreNow[i][h] = score;
            if (score < minScore) {
                alive[i][h] = 0;
                continue;
            }
            survivors.push_back(cand[q]);
        }
        if (p.densePoseRefinement && !survivors.empty()) {
            std::vector<float> init2(survivors.size() * 12);
            std::vector<int>   hs2(survivors.size());
            for (size_t q = 0; q < survivors.size(); q++) {
                std::memcpy(&init2[12 * q], &poseNow[survivors[q].scene][12 * (size_t)survivors[q].h],
                            sizeof(float) * 12);
                hs2[q] = survivors[q].scene;
            }
            int *hsD2 = X.batchHypScene.as<int>(survivors.size());
            PPF_HIP_CHECK(hipMemcpyAsync(hsD2, hs2.data(), sizeof(int) * hs2.size(),
                                         hipMemcpyHostToDevice, s));
            const int first2 = survivors[0].scene;
            IcpCriteria dc{p.poseRefNumSteps, poseRefDist, reSampleStep * 0.5f, sampleStep, 0.05f};
            std::vector<IcpResultDev> res2(survivors.size());
            icp_batch(X.ws, hm.reSampled.size(), dm.rx.get<float>(), dm.ry.get<float>(),
                      dm.rz.get<float>(), dm.rnx.get<float>(), dm.rny.get<float>(),
                      dm.rnz.get<float>(), vd[first2], vf[first2], init2.data(),
                      (int)survivors.size(), dc, scoringDist, D.icfg, res2.data(), s, gS + 2 * n,
                      gS + n, hsD2);
            nDenseIcp = (int)survivors.size();
            for (size_t q = 0; q < survivors.size(); q++) {
                const int i = survivors[q].scene, h = survivors[q].h;
                if (!res2[q].converged) {
                    alive[i][h] = 0;
                    continue;
                }
                std::memcpy(&poseNow[i][12 * (size_t)h], res2[q].pose, sizeof(float) * 12);
                float score = (float)res2[q].inliers / (float)hm.reSampled.size();
                if (score > 1.0f)
                    score = 1.0f;
                scoreNow[i][h] = score;
            }
        }
    }
    // ordered accept walk per scene (ppf.cpp:439-495), stable sort by score
    for (int i = 0; i < n; i++) {
        for (int h = 0; h < (int)hyp[i].size(); h++) {
            if (!alive[i][h])
                continue;
            float score = scoreNow[i][h];
            if ((p.sparsePoseRefinement || p.densePoseRefinement) && score < minScore)
                continue;
            outs[i].results.emplace_back(score, std::vector<float>(&poseNow[i][12 * (size_t)h],
                                                                   &poseNow[i][12 * (size_t)h] + 12));
            if ((int)outs[i].results.size() >= p.numMatches)
                break;
        }
        std::stable_sort(outs[i].results.begin(), outs[i].results.end(),
                         [](const auto &a, const auto &b) { return a.first > b.first; });
    }
    PPF_HIP_CHECK(hipEventRecord(X.ev[8], s));
    PPF_HIP_CHECK(hipEventSynchronize(X.ev[8]));
    for (int i = 0; i < n; i++) {
        st->nScene += clouds[i]->n;
        st->nValid += clouds[i]->nValid;
        st->nSampled += X.sceneWork[i]->ns;
        st->nResults += (int)outs[i].results.size();
    }
    st->nKeypoints      = K;
    st->nVotedKeypoints = vs.voted;
    st->nPoses          = np;
    st->nClusters       = nClustersTot;
    int nh = 0;
    for (auto &v : hyp)
        nh += (int)v.size();
    st->nHypotheses   = nh;
    st->nSparseIcp    = nSparse;
    st->nDenseIcp     = nDenseIcp;
    st->nDense        = nDenseTot;
    st->nPairs        = vs.pairs;
    st->nKeyHits      = vs.hits;
    st->nVotes        = vs.votes;
    st->msSample      = elapsed(X.ev[0], X.ev[3]);  // sampling + normals + key points
    st->msVote        = elapsed(X.ev[3], X.ev[5]);
    st->msVoteKernel  = vs.kernelMs;
    st->msPairKernel  = vs.pairMs;
    st->msHitsKernel  = vs.hitsMs;
    st->nVoteFallback = vs.overflow;
    st->msCluster     = elapsed(X.ev[5], X.ev[6]);
    st->msIcp         = elapsed(X.ev[6], X.ev[8]);
    st->msTotal       = elapsed(X.ev[0], X.ev[8]);
}

void write_poses(const MatchOut &o, float *poses16, float *scores, size_t cap, size_t *nOut) {
    size_t n = o.results.size();
    if (nOut)
        *nOut = n;
    for (size_t i = 0; i < n && i < cap; i++) {
        const float *P = o.results[i].second.data();
        if (poses16) {
            float *M = poses16 + 16 * i;  // column-major 4x4 (Eigen::Matrix4f)
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 4; c++)
                    M[c * 4 + r] = P[r * 4 + c];
            M[3] = M[7] = M[11] = 0.0f;
            M[15]               = 1.0f;
        }
        if (scores)
            scores[i] = o.results[i].first;
    }
}

// ---------------------------------------------------------------- model file

template <class T> void wr(std::ofstream &f, const T &v) {
    f.write(reinterpret_cast<const char *>(&v), sizeof(T));
}
template <class T> void rd(std::ifstream &f, T &v) {
    f.read(reinterpret_cast<char *>(&v), sizeof(T));
}

void write_cloud(std::ofstream &f, const HostCloud &c) {  // serialize.cpp:59-86
    uint32_t np = (uint32_t)c.size();
    wr(f, np);
    f.write(reinterpret_cast<const char *>(c.p.data()), sizeof(float) * c.p.size());
    uint32_t nn = (uint32_t)(c.n.size() / 3);
    wr(f, nn);
    f.write(reinterpret_cast<const char *>(c.n.data()), sizeof(float) * c.n.size());
    f.write(reinterpret_cast<const char *>(c.boxMin), sizeof(float) * 3);
    f.write(reinterpret_cast<const char *>(c.boxMax), sizeof(float) * 3);
    f.write(reinterpret_cast<const char *>(c.vp), sizeof(float) * 3);
}

// bytes of the file not read yet (0 once the stream has failed)
size_t remaining(std::ifstream &f, size_t fileSize) {
    if (!f.good())
        return 0;
    std::streamoff at = f.tellg();
    return at < 0 || (size_t)at > fileSize ? 0 : fileSize - (size_t)at;
}

// an element count read from the file, checked against what is left of it
// before anything is allocated for it (a corrupt count would otherwise ask
// for gigabytes or wrap a size computation)
uint32_t read_count(std::ifstream &f, size_t fileSize, size_t elemBytes, const char *path) {
    uint32_t n = 0;
    rd(f, n);
    if (!f.good() || (size_t)n > remaining(f, fileSize) / elemBytes)
        throw std::runtime_error(std::string("truncated model file:") + path);
    return n;
}

void read_cloud(std::ifstream &f, size_t fileSize, HostCloud &c, const char *path) {
    uint32_t np = read_count(f, fileSize, 12, path);
    c.p.resize((size_t)np * 3);
    f.read(reinterpret_cast<char *>(c.p.data()), sizeof(float) * c.p.size());
    uint32_t nn = read_count(f, fileSize, 12, path);
    c.n.resize((size_t)nn * 3);
    f.read(reinterpret_cast<char *>(c.n.data()), sizeof(float) * c.n.size());
    f.read(reinterpret_cast<char *>(c.boxMin), sizeof(float) * 3);
    f.read(reinterpret_cast<char *>(c.boxMax), sizeof(float) * 3);
    f.read(reinterpret_cast<char *>(c.vp), sizeof(float) * 3);
    // a trained model has a normal per point (trainModel estimates the
    // missing ones); the voting frames and the refinement read both arrays
    if (f.good() && nn != np)
        throw std::runtime_error(std::string("corrupt point cloud in:") + path);
}

void save_impl(const ppf_detector_s &D, const char *path) {
    // ppf.cpp:512-527 (the file is created before the model check, as there)
    std::ofstream f(path, std::ios::out | std::ios::binary);
    if (!f.is_open())
        throw std::runtime_error(std::string("failed to open file:") + path);
    if (!D.trained)
        throw std::runtime_error("No trained model in save");
    const HostModel &m = D.hm;
    wr(f, kMagic);
    wr(f, kVersion);
    wr(f, m.samplingDistanceRel);
    wr(f, m.param.featDistanceStepRel);  // serialize.cpp:180-186
    wr(f, m.param.featAngleResolution);
    wr(f, m.param.poseRefRelSamplingDistance);
    wr(f, m.param.knnNormal);
    uint8_t smooth = m.param.smoothNormal ? 1 : 0;
    wr(f, smooth);
    write_cloud(f, m.sampled);
    write_cloud(f, m.reSampled);
    uint32_t nk = (uint32_t)m.keys.size();  // serialize.cpp:164-170
    wr(f, nk);
    for (uint32_t k = 0; k < nk; k++) {
        wr(f, m.keys[k]);
        uint32_t len = (uint32_t)(m.offs[k + 1] - m.offs[k]);
        wr(f, len);
        f.write(reinterpret_cast<const char *>(&m.refs[m.offs[k]]), sizeof(uint32_t) * len);
        wr(f, len);
        f.write(reinterpret_cast<const char *>(&m.alphas[m.offs[k]]), sizeof(float) * len);
    }
    f.close();
}

void load_impl(ppf_detector_s &D, MatchCtx &X, const char *path) {
    std::ifstream f(path, std::ios::binary);  // ppf.cpp:529-547
    if (!f.is_open())
        throw std::runtime_error(std::string("failed to open file:") + path);
    D.trained = false;
    f.seekg(0, std::ios::end);
    const size_t fileSize = f.tellg() < 0 ? 0 : (size_t)f.tellg();
    f.seekg(0, std::ios::beg);
    int magic = 0, version = 0;
    rd(f, magic);
    if (magic != kMagic)
        throw std::runtime_error(std::string("unsupported file format:") + path);
    rd(f, version);  // not checked, as in the reference
    HostModel m;
    rd(f, m.samplingDistanceRel);
    rd(f, m.param.featDistanceStepRel);
    rd(f, m.param.featAngleResolution);
    rd(f, m.param.poseRefRelSamplingDistance);
    rd(f, m.param.knnNormal);
    uint8_t smooth = 0;
    rd(f, smooth);
    m.param.smoothNormal = smooth ? 1 : 0;
    if (!f.good())
        throw std::runtime_error(std::string("truncated model file:") + path);
    // the stored TrainParam must be one the constructor accepts: the angle
    // resolution sizes the vote accumulator and the steps the quantisation
    check_train_param(m.param);
    read_cloud(f, fileSize, m.sampled, path);
    read_cloud(f, fileSize, m.reSampled, path);
    if (m.sampled.size() == 0)
        throw std::runtime_error(std::string("corrupt point cloud in:") + path);
    // a bucket takes at least 12 bytes (key and two counts)
    uint32_t nk = read_count(f, fileSize, 12, path);
    // buckets keep the file's intra-bucket order (it decides which entries
    // take the bin tail rule); buckets are re-ordered by key for the CSR.
    std::vector<std::pair<uint32_t, std::pair<std::vector<uint32_t>, std::vector<float>>>> b(nk);
    const uint32_t nRef = (uint32_t)m.sampled.size();
    for (uint32_t k = 0; k < nk && f.good(); k++) {
        rd(f, b[k].first);
        uint32_t len = read_count(f, fileSize, 4, path);
        b[k].second.first.resize(len);
        f.read(reinterpret_cast<char *>(b[k].second.first.data()), sizeof(uint32_t) * len);
        uint32_t lenA = read_count(f, fileSize, 4, path);
        if (lenA != len)
            throw std::runtime_error(std::string("corrupt hash table in:") + path);
        b[k].second.second.resize(lenA);
        f.read(reinterpret_cast<char *>(b[k].second.second.data()), sizeof(float) * lenA);
        // every entry names a sampled model point (its frame and its row of
        // the vote accumulator)
        for (uint32_t r : b[k].second.first)
            if (r >= nRef)
                throw std::runtime_error(std::string("corrupt hash table in:") + path);
        // training never stores a NaN alpha (ppf.cpp:131-139); the voting
        // kernels rely on finite ones
        for (float al : b[k].second.second)
            if (!std::isfinite(al))
                throw std::runtime_error(std::string("corrupt hash table in:") + path);
    }
    if (!f.good())
        throw std::runtime_error(std::string("truncated model file:") + path);
    std::stable_sort(b.begin(), b.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    m.offs.assign(1, 0);
    for (size_t k = 0; k < b.size(); k++) {
        // a key stored twice keeps its first bucket, as the reference's
        // flat_hash_map::emplace does (serialize.cpp:150-159)
        if (k > 0 && b[k].first == b[k - 1].first)
            continue;
        auto &e = b[k];
        m.keys.push_back(e.first);
        m.refs.insert(m.refs.end(), e.second.first.begin(), e.second.first.end());
        m.alphas.insert(m.alphas.end(), e.second.second.begin(), e.second.second.end());
        m.offs.push_back((int)m.refs.size());
    }
    D.hm = std::move(m);
    upload_model(D, X);
    D.trained = true;
}

template <class F> int guarded(F &&f) {
    try {
        f();
        return PPF_OK;
    } catch (const std::range_error &e) {
        g_lastError = e.what();
        return PPF_ERR_RANGE;
    } catch (const HipError &e) {
        g_lastError = e.what();
        return PPF_ERR_DEVICE;
    } catch (const std::exception &e) {
        g_lastError = e.what();
        return PPF_ERR_RUNTIME;
    } catch (...) {
        g_lastError = "unknown error";
        return PPF_ERR_RUNTIME;
    }
}

#define PPF_REQUIRE(cond, msg)                 \
    do {                                       \
        if (!(cond)) {                         \
            g_lastError = msg;                 \
            return PPF_ERR_ARG;                \
        }                                      \
    } while (0)

// ---------------------------------------------------------------- small kernels for stage APIs

__global__ void radius_kernel(GridView g, const float *q, int nq, float r2, int *counts) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq)
        return;
    float qx = q[3 * t], qy = q[3 * t + 1], qz = q[3 * t + 2];
    float r  = sqrtf(r2) * 1.0001f + 1e-3f * g.cell;
    int   xlo = cell_coord(qx - r, g.ox, g.cell, g.dx), xhi = cell_coord(qx + r, g.ox, g.cell, g.dx);
    int   ylo = cell_coord(qy - r, g.oy, g.cell, g.dy), yhi = cell_coord(qy + r, g.oy, g.cell, g.dy);
    int   zlo = cell_coord(qz - r, g.oz, g.cell, g.dz), zhi = cell_coord(qz + r, g.oz, g.cell, g.dz);
    int   c = 0;
    for (int cz = zlo; cz <= zhi; cz++)
        for (int cy = ylo; cy <= yhi; cy++)
            for (int cx = xlo; cx <= xhi; cx++) {
                int s0, c0;
                if (!cell_lookup(g, cx, cy, cz, s0, c0))
                    continue;
                for (int k = s0; k < s0 + c0; k++)
                    if (ppfc::dist2(qx, qy, qz, g.x[k], g.y[k], g.z[k]) < r2)
                        c++;
            }
    counts[t] = c;
}

__global__ void radius_fill_kernel(GridView g, const float *q, int nq, float r2, const int *offs,
                                   int *ids, int cap) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq)
        return;
    float qx = q[3 * t], qy = q[3 * t + 1], qz = q[3 * t + 2];
    float r  = sqrtf(r2) * 1.0001f + 1e-3f * g.cell;
    int   xlo = cell_coord(qx - r, g.ox, g.cell, g.dx), xhi = cell_coord(qx + r, g.ox, g.cell, g.dx);
    int   ylo = cell_coord(qy - r, g.oy, g.cell, g.dy), yhi = cell_coord(qy + r, g.oy, g.cell, g.dy);
    int   zlo = cell_coord(qz - r, g.oz, g.cell, g.dz), zhi = cell_coord(qz + r, g.oz, g.cell, g.dz);
    int   o = offs[t];
    for (int cz = zlo; cz <= zhi; cz++)
        for (int cy = ylo; cy <= yhi; cy++)
            for (int cx = xlo; cx <= xhi; cx++) {
                int s0, c0;
                if (!cell_lookup(g, cx, cy, cz, s0, c0))
                    continue;
                for (int k = s0; k < s0 + c0; k++)
                    if (ppfc::dist2(qx, qy, qz, g.x[k], g.y[k], g.z[k]) < r2) {
                        if (o < cap)
                            ids[o] = g.id[k];
                        o++;
                    }
            }
}

__global__ void knn_kernel(GridView g, const float *q, int nq, int k, int *ids, float *d2) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq)
        return;
    KnnList<32> L;
    grid_knn<32>(g, q[3 * t], q[3 * t + 1], q[3 * t + 2], k, L);
    for (int i = 0; i < 32; i++) {
        if (i >= k)
            break;
        ids[(size_t)t * k + i] = i < L.cnt ? L.id[i] : -1;
        d2[(size_t)t * k + i]  = i < L.cnt ? L.d[i] : INFINITY;
    }
}

// ceilings of this GPU for the roofline: a streaming copy (16-byte loads and
// stores, grid-stride) and LDS atomic adds at pseudo-random addresses
__global__ void ceiling_copy_kernel(const uint4 *src, uint4 *dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
__global__ void __launch_bounds__(1024) ceiling_lds_kernel(int iters, uint32_t *sink) {
    __shared__ uint32_t acc[16384];
    for (int i = threadIdx.x; i < 16384; i += blockDim.x)
        acc[i] = 0u;
    __syncthreads();
    uint32_t x = (blockIdx.x * 1024u + threadIdx.x) * 2654435761u + 1u;
    for (int it = 0; it < iters; it++) {
        x = x * 1664525u + 1013904223u;  // Numerical Recipes LCG
        atomicAdd(&acc[x >> 18], 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int i = 0; i < 16384; i += 997)
            s += acc[i];
        atomicAdd(sink, s);
    }
}

__global__ void math_kernel(int op, const float *a, const float *b, int n, uint32_t *out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float x = a[i], y = b ? b[i] : 0.0f;
    float r = 0.0f;
    switch (op) {
        case 0: r = ppfc::atanf_c(x); break;
        case 1: r = ppfc::atan2f_c(x, y); break;
        case 2: r = x / y; break;
        case 3: r = sqrtf(x); break;
        case 4:
            out[i] = ppfc::murmur_key4(__float_as_uint(x), __float_as_uint(y), 0u, 0u);
            return;
        case 5: out[i] = ppfc::quant(x, y, false); return;
        case 6: out[i] = ppfc::quant(x, y, true); return;
        case 7: {
            double d = sqrt((double)x * (double)y);
            unsigned long long u = __double_as_longlong(d);
            out[i] = (uint32_t)(u >> 32) ^ (uint32_t)u;
            return;
        }
        default: break;
    }
    out[i] = __float_as_uint(r);
}

hipStream_t util_stream() {
    static hipStream_t s = nullptr;
    if (!s)
        PPF_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
}

struct TmpCloud {
    DevBuf x, y, z, aos;
    void up(const float *p, int n, hipStream_t s) {
        int m = std::max(n, 1);
        PPF_HIP_CHECK(hipMemcpyAsync(aos.as<float>((size_t)m * 3), p, sizeof(float) * 3 * n,
                                     hipMemcpyHostToDevice, s));
        split_kernel<<<div_up(m, 256), 256, 0, s>>>(aos.get<float>(), n, x.as<float>(m),
                                                     y.as<float>(m), z.as<float>(m));
    }
};

}  // namespace

// ---------------------------------------------------------------- allocation-free voxel sampling
//
// ppf_voxel_sample_noalloc's scratch: the voxel table, flags and block
// counts of up to kNoAllocMaxPoints points, in a fixed device array that
// each GPU holds from module load to process end.  Host code never passes
// the array's name to a kernel (in host code it is a stand-in, not a device
// address): the device address comes from hipGetSymbolAddress, looked up
// once per device with that device current, and only that address is used.

constexpr int    kNoAllocMaxPoints = 1 << 20;
constexpr size_t kNoAllocScratch   = 34619392;  // voxel_scratch_bytes(kNoAllocMaxPoints)
static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];

namespace {
constexpr int  kMaxDevices = 64;
std::mutex     g_scratchMu;
unsigned char *g_scratchPtr[kMaxDevices] = {};

unsigned char *scratch_ptr() {
    int dev = 0;
    PPF_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        throw HipError("ppf_voxel_sample_noalloc: device index beyond the library's table");
    std::lock_guard<std::mutex> g(g_scratchMu);
    if (!g_scratchPtr[dev]) {
        void *p = nullptr;
        PPF_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratchPtr[dev] = static_cast<unsigned char *>(p);
    }
    return g_scratchPtr[dev];
}
}  // namespace

// ======================================================================== C ABI

extern "C" {

PPF_API int ppf_train_param_default(ppf_train_param *p) {
    PPF_REQUIRE(p, "null parameter");
    *p = ppf_train_param{0.04f, 30, 0.01f, 10, 1};  // type.h:71-73
    return PPF_OK;
}

PPF_API int ppf_match_param_default(ppf_match_param *p) {
    PPF_REQUIRE(p, "null parameter");
    *p = ppf_match_param{1, 10, 1, 0, 0.5f, 0.0f, 1, 1, 5, 0.1f, 0.0f, 0.01f, 0.0f};  // type.h:170-175
    return PPF_OK;
}

PPF_API int ppf_train_param_check(const ppf_train_param *p) {
    PPF_REQUIRE(p, "null parameter");
    return guarded([&] { check_train_param(*p); });
}

PPF_API int ppf_match_param_check(const ppf_match_param *p) {
    PPF_REQUIRE(p, "null parameter");
    return guarded([&] { check_match_param(*p); });
}

PPF_API int ppf_detector_create(ppf_detector *out) {
    PPF_REQUIRE(out, "null output handle");
    *out = nullptr;
    return guarded([&] {
        int dev = 0;
        PPF_HIP_CHECK(hipGetDeviceCount(&dev));
        if (dev < 1)
            throw HipError("no HIP device available");
        *out = new ppf_detector_s();
    });
}

PPF_API void ppf_detector_destroy(ppf_detector d) { delete d; }

PPF_API int ppf_train_model(ppf_detector d, const float *points, const float *normals, size_t n,
                            const float *box6, const float *viewPoint, float rel,
                            const ppf_train_param *param) {
    PPF_REQUIRE(d, "null detector");
    return guarded([&] {
        ppf_train_param p;
        if (param)
            p = *param;
        else
            ppf_train_param_default(&p);
        WriteLock w(d->model);
        Lease     L(*d);
        train_impl(*d, *L, points, normals, (int)n, box6, viewPoint, rel, p);
    });
}

PPF_API int ppf_train_model_mesh(ppf_detector d, const float *points, const float *normals,
                                 size_t n, const int *faces, size_t nFaces, const float *box6,
                                 const float *viewPoint, float rel, const ppf_train_param *param,
                                 uint64_t seed) {
    PPF_REQUIRE(d && (faces || nFaces == 0), "null argument");
    return guarded([&] {
        ppf_train_param p;
        if (param)
            p = *param;
        else
            ppf_train_param_default(&p);
        WriteLock w(d->model);
        Lease     L(*d);
        if (nFaces == 0) {
            train_impl(*d, *L, points, normals, (int)n, box6, viewPoint, rel, p);
            return;
        }
        // ppf.cpp:39-52: the checks, then the box of the given model, whose
        // diameter sets the sampling radius reSampleStep / 4 (:57, :66)
        if (rel > 1 || rel < 0)
            throw std::range_error("Invalid Input: samplingDistanceRel range mismatch in trainModel");
        if (n == 0 || !points)
            throw std::runtime_error("Invalid Input: empty model in trainModel");
        check_train_param(p);
        float box[6];
        if (box6 && box_diameter(box6, box6 + 3) != 0.0f) {
            std::memcpy(box, box6, sizeof box);
        } else {
            std::vector<int> v = remove_nan(points, normals, (int)n);
            bbox_host(points, v, box, box + 3);
        }
        const float Dm           = box_diameter(box, box + 3);
        const float reSampleStep = Dm * p.poseRefRelSamplingDistance;
        std::vector<float> sp, sn;
        sample_mesh(points, (int)n, faces, (int)nFaces, (float)(reSampleStep / 4.), seed, sp, sn,
                    (*L).stream);
        if (sp.empty())
            throw std::runtime_error("Invalid Input: mesh too small to sample in trainModel");
        train_impl(*d, *L, sp.data(), sn.data(), (int)(sp.size() / 3), box, viewPoint, rel, p);
    });
}

PPF_API int ppf_match_scene(ppf_detector d, const float *points, const float *normals, size_t n,
                            const float *box6, const float *viewPoint, float rel, float kpFrac,
                            float minScore, const ppf_match_param *param, float *poses16,
                            float *scores, size_t cap, size_t *nOut, ppf_match_result *result,
                            ppf_match_stats *stats) {
    PPF_REQUIRE(d, "null detector");
    return guarded([&] {
        ReadLock r(d->model);
        if (!d->trained)
            throw std::runtime_error("No trained model in matchScene");
        ppf_match_param p;
        if (param)
            p = *param;
        else
            ppf_match_param_default(&p);
        if (n == 0 || !points)
            throw std::runtime_error("Invalid Input: empty scene in matchScene");
        Lease     L(*d);
        MatchCtx &X = *L;
        const auto t0 = std::chrono::steady_clock::now();
        X.tmpScene.upload(points, normals, (int)n, box6, viewPoint, X.stream, X.ingest);
        const float ingestMs =
            std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        MatchOut o;
        match_impl(*d, X, X.tmpScene, rel, kpFrac, minScore, p, o, result, stats);
        if (stats)
            stats->msIngest = ingestMs;
        write_poses(o, poses16, scores, cap, nOut);
    });
}

PPF_API int ppf_save(ppf_detector d, const char *path) {
    PPF_REQUIRE(d && path, "null argument");
    return guarded([&] {
        ReadLock r(d->model);
        save_impl(*d, path);
    });
}

PPF_API int ppf_load(ppf_detector d, const char *path) {
    PPF_REQUIRE(d && path, "null argument");
    return guarded([&] {
        WriteLock w(d->model);
        Lease     L(*d);
        load_impl(*d, *L, path);
    });
}

PPF_API const char *ppf_last_error(void) { return g_lastError.c_str(); }

PPF_API int ppf_scene_upload(ppf_detector d, const float *points, const float *normals, size_t n,
                             const float *box6, const float *viewPoint, ppf_scene *out) {
    PPF_REQUIRE(d && out, "null argument");
    *out = nullptr;
    return guarded([&] {
        auto  sc = std::make_unique<ppf_scene_s>();
        Lease L(*d);
        sc->cloud.upload(points, normals, (int)n, box6, viewPoint, (*L).stream, (*L).ingest);
        *out = sc.release();
    });
}

PPF_API void ppf_scene_free(ppf_scene s) { delete s; }

PPF_API int ppf_scene_download(ppf_scene sc, size_t *n, size_t *nValid, int *hasNormals,
                               float *points, float *normals, int *valid, float box6[6]) {
    PPF_REQUIRE(sc, "null scene");
    return guarded([&] {
        std::lock_guard<std::mutex> busy(sc->busy);
        const DevCloud &c = sc->cloud;
        if (n)
            *n = (size_t)c.n;
        if (nValid)
            *nValid = (size_t)c.nValid;
        if (hasNormals)
            *hasNormals = c.hasNormal ? 1 : 0;
        if (box6) {
            std::memcpy(box6, c.boxMin, sizeof c.boxMin);
            std::memcpy(box6 + 3, c.boxMax, sizeof c.boxMax);
        }
        if (c.n > 0 && points)
            PPF_HIP_CHECK(hipMemcpy(points, c.aos.get<float>(), sizeof(float) * 3 * c.n,
                                    hipMemcpyDeviceToHost));
        if (c.n > 0 && normals && c.hasNormal)
            PPF_HIP_CHECK(hipMemcpy(normals, c.aosN.get<float>(), sizeof(float) * 3 * c.n,
                                    hipMemcpyDeviceToHost));
        if (c.nValid > 0 && valid)
            PPF_HIP_CHECK(hipMemcpy(valid, c.valid.get<int>(), sizeof(int) * c.nValid,
                                    hipMemcpyDeviceToHost));
    });
}

PPF_API int ppf_match_scene_resident(ppf_detector d, ppf_scene sc, float rel, float kpFrac,
                                     float minScore, const ppf_match_param *param, float *poses16,
                                     float *scores, size_t cap, size_t *nOut,
                                     ppf_match_stats *stats) {
    PPF_REQUIRE(d && sc, "null argument");
    return guarded([&] {
        ppf_match_param p;
        if (param)
            p = *param;
        else
            ppf_match_param_default(&p);
        ReadLock                    r(d->model);
        std::lock_guard<std::mutex> busy(sc->busy);
        Lease                       L(*d);
        MatchOut                    o;
        match_impl(*d, *L, sc->cloud, rel, kpFrac, minScore, p, o, nullptr, stats);
        write_poses(o, poses16, scores, cap, nOut);
    });
}

static void pose16_to_12(const float *p16, float *p12) {
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 4; c++)
            p12[r * 4 + c] = p16[c * 4 + r];
}

PPF_API int ppf_scene_inliers(ppf_detector d, ppf_scene sc, const float *poses16, size_t nPoses,
                              int dense, float dist, int *inliers) {
    PPF_REQUIRE(d && sc && (poses16 || nPoses == 0) && (inliers || nPoses == 0),
                "null argument");
    return guarded([&] {
        ReadLock                    r(d->model);
        std::lock_guard<std::mutex> busy(sc->busy);
        Lease                       L(*d);
        MatchCtx                   &X = *L;
        if (!d->trained)
            throw std::runtime_error("No trained model in matchScene");
        DevCloud   &c  = sc->cloud;
        hipStream_t s  = X.stream;
        DevModel   &dm = d->dm;
        const float Dm = box_diameter(d->hm.sampled.boxMin, d->hm.sampled.boxMax);
        grid_build(X.ws.gFull, c.px.get<float>(), c.py.get<float>(), c.pz.get<float>(), nullptr,
                   nullptr, nullptr, c.valid.get<int>(), c.nValid,
                   std::max(dist * 2.0f, Dm * 1e-3f), s, nullptr, true);
        std::vector<float> P(12 * nPoses);
        for (size_t h = 0; h < nPoses; h++)
            pose16_to_12(poses16 + 16 * h, &P[12 * h]);
        if (dense)
            icp_inliers(X.ws, X.ws.gFull.view, dm.nReSampled, dm.rx.get<float>(),
                        dm.ry.get<float>(), dm.rz.get<float>(), P.data(), (int)nPoses, dist,
                        inliers, s);
        else
            icp_inliers(X.ws, X.ws.gFull.view, dm.nSampled, dm.sx.get<float>(),
                        dm.sy.get<float>(), dm.sz.get<float>(), P.data(), (int)nPoses, dist,
                        inliers, s);
    });
}

PPF_API int ppf_match_scenes_resident(ppf_detector d, const ppf_scene *scenes, size_t nScenes,
                                      float rel, float kpFrac, float minScore,
                                      const ppf_match_param *param, float *poses16, float *scores,
                                      size_t cap, size_t *nOut, ppf_match_stats *stats) {
    PPF_REQUIRE(d && (scenes || nScenes == 0) && nOut, "null argument");
    for (size_t i = 0; i < nScenes; i++)
        PPF_REQUIRE(scenes[i], "null scene");
    return guarded([&] {
        ppf_match_param p;
        if (param)
            p = *param;
        else
            ppf_match_param_default(&p);
        // a scene named more than once is matched once (its stages move its
        // points in place, so two lanes must never work on one scene) and
        // its result is written to every entry that names it
        std::vector<DevCloud *> clouds;
        std::vector<size_t>     slot(nScenes);
        for (size_t i = 0; i < nScenes; i++) {
            size_t j = 0;
            while (j < i && scenes[j] != scenes[i])
                j++;
            if (j == i) {
                slot[i] = clouds.size();
                clouds.push_back(&scenes[i]->cloud);
            } else {
                slot[i] = slot[j];
            }
        }
        // each distinct scene is held for the call, in address order (no
        // deadlock between calls naming the same scenes in other orders)
        std::vector<ppf_scene_s *> held;
        for (size_t i = 0; i < nScenes; i++)
            if (slot[i] == held.size())
                held.push_back(scenes[i]);
        std::vector<ppf_scene_s *> order(held);
        std::sort(order.begin(), order.end());
        std::vector<std::unique_lock<std::mutex>> busy;
        for (ppf_scene_s *h : order)
            busy.emplace_back(h->busy);
        ReadLock              r(d->model);
        Lease                 L(*d);
        std::vector<MatchOut> outs;
        if (!clouds.empty())
            match_batch_impl(*d, *L, clouds, rel, kpFrac, minScore, p, outs, stats);
        for (size_t i = 0; i < nScenes; i++)
            write_poses(outs[slot[i]], poses16 ? poses16 + 16 * cap * i : nullptr,
                        scores ? scores + cap * i : nullptr, cap, nOut + i);
    });
}

namespace {
ppf_tuning tuning_of(const ppf_detector_s &d) {
    ppf_tuning t;
    t.voteBlock    = d.vcfg.block;
    t.voteXcdRemap = d.vcfg.xcdRemap;
    t.voteCellDiv  = d.vcfg.cellDiv;
    t.icpBlock     = d.icfg.block;
    t.icpSpread    = d.icfg.spread;
    t.normCell4    = d.normCell4;
    t.voteSplit    = d.vcfg.split;
    t.voteHitCap   = d.vcfg.hitCap;
    t.voteTiles    = d.vcfg.tiles;
    t.voteOrder    = d.vcfg.order;
    return t;
}
}  // namespace

PPF_API int ppf_get_tuning(ppf_detector d, ppf_tuning *out) {
    PPF_REQUIRE(d && out, "null argument");
    ReadLock r(d->model);
    *out = tuning_of(*d);
    return PPF_OK;
}

PPF_API int ppf_set_tuning(ppf_detector d, const ppf_tuning *t) {
    PPF_REQUIRE(d && t, "null argument");
    PPF_REQUIRE(t->voteBlock == 256 || t->voteBlock == 512 || t->voteBlock == 1024,
                "vote block must be 256, 512 or 1024");
    PPF_REQUIRE(t->voteCellDiv >= 2 && t->voteCellDiv <= 4, "vote cell divisor must be 2, 3 or 4");
    PPF_REQUIRE(t->icpBlock == 64 || t->icpBlock == 128 || t->icpBlock == 256 ||
                    t->icpBlock == 512 || t->icpBlock == 1024,
                "icp block must be 64..1024, a power of two");
    PPF_REQUIRE(t->voteHitCap >= 1 && t->voteHitCap <= (1 << 20),
                "vote hit capacity must be 1 .. 2^20");
    PPF_REQUIRE(t->voteTiles >= 1 && t->voteTiles <= 16, "vote tiles must be 1 .. 16");
    PPF_REQUIRE(t->normCell4 >= 1 && t->normCell4 <= 16,
                "normals cell must be 1 .. 16 quarters of the sample step");
    WriteLock w(d->model);
    d->vcfg.block    = t->voteBlock;
    d->vcfg.xcdRemap = t->voteXcdRemap ? 1 : 0;
    d->vcfg.cellDiv  = t->voteCellDiv;
    d->icfg.block    = t->icpBlock;
    d->icfg.spread   = t->icpSpread ? 1 : 0;
    d->normCell4     = t->normCell4;
    d->vcfg.split    = t->voteSplit ? 1 : 0;
    d->vcfg.hitCap   = t->voteHitCap;
    d->vcfg.tiles    = t->voteTiles;
    d->vcfg.order    = t->voteOrder ? 1 : 0;
    return PPF_OK;
}

PPF_API int ppf_autotune(ppf_detector d, ppf_scene sc, float rel, float kpFrac, float minScore,
                         const ppf_match_param *param, ppf_tune_report *rep) {
    PPF_REQUIRE(d && sc && rep, "null argument");
    return guarded([&] {
        ppf_match_param p;
        if (param)
            p = *param;
        else
            ppf_match_param_default(&p);
        WriteLock                   w(d->model);
        std::lock_guard<std::mutex> busy(sc->busy);
        Lease                       L(*d);
        VoteConfig bestV = d->vcfg;
        IcpConfig  bestI = d->icfg;
        float      bestVoteMs = 1e30f, bestIcpMs = 1e30f;
        int        tried = 0;
        auto run = [&](float &voteMs, float &icpMs) {
            voteMs = icpMs = 1e30f;
            for (int rep2 = 0; rep2 < 2; rep2++) {
                MatchOut        o;
                ppf_match_stats st{};
                match_impl(*d, *L, sc->cloud, rel, kpFrac, minScore, p, o, nullptr, &st);
                voteMs = std::min(voteMs, st.msVoteKernel);
                icpMs  = std::min(icpMs, st.msIcp);
            }
        };
        // vote stage: one knob at a time, the others at their best so far
        auto tryVote = [&](int VoteConfig::*knob, std::initializer_list<int> values) {
            for (int v0 : values) {
                d->vcfg       = bestV;
                d->vcfg.*knob = v0;
                float v, i;
                run(v, i);
                tried++;
                if (v < bestVoteMs) {
                    bestVoteMs = v;
                    bestV      = d->vcfg;
                }
            }
        };
        tryVote(&VoteConfig::split, {1, 0});
        tryVote(&VoteConfig::block, {256, 512, 1024});
        tryVote(&VoteConfig::cellDiv, {2, 3, 4});
        tryVote(&VoteConfig::tiles, {1, 2, 3, 4});
        tryVote(&VoteConfig::order, {1, 0});
        tryVote(&VoteConfig::xcdRemap, {0, 1});
        d->vcfg = bestV;
        for (int spread : {1, 0})
            for (int block : {64, 128, 256, 512, 1024}) {
                d->icfg.spread = spread;
                d->icfg.block  = block;
                float v, i;
                run(v, i);
                tried++;
                if (i < bestIcpMs) {
                    bestIcpMs = i;
                    bestI     = d->icfg;
                }
            }
        d->icfg               = bestI;
        // normals: the grid cell
        float bestNormMs = 1e30f;
        auto  tryNormals = [&](int ppf_detector_s::*knob, std::initializer_list<int> values) {
            int best = d->*knob;
            for (int q : values) {
                d->*knob = q;
                float ms = 1e30f;
                for (int rep2 = 0; rep2 < 2; rep2++) {
                    MatchOut        o;
                    ppf_match_stats st{};
                    match_impl(*d, *L, sc->cloud, rel, kpFrac, minScore, p, o, nullptr, &st);
                    ms = std::min(ms, st.msNormals + st.msIcpPrepare);
                }
                tried++;
                if (ms < bestNormMs) {
                    bestNormMs = ms;
                    best       = q;
                }
            }
            d->*knob = best;
        };
        tryNormals(&ppf_detector_s::normCell4, {3, 4, 5, 6});
        rep->normalsMs        = bestNormMs;
        rep->best             = tuning_of(*d);
        rep->voteMs           = bestVoteMs;
        rep->icpMs            = bestIcpMs;
        rep->configsTried     = tried;
    });
}

PPF_API int ppf_model_get_info(ppf_detector d, ppf_model_info *info) {
    PPF_REQUIRE(d && info, "null argument");
    return guarded([&] {
        ReadLock r(d->model);
        if (!d->trained)
            throw std::runtime_error("No trained model");
        const HostModel &m     = d->hm;
        info->nSampled         = m.sampled.size();
        info->nReSampled       = m.reSampled.size();
        info->nKeys            = (int)m.keys.size();
        info->nEntries         = (int)m.refs.size();
        info->angleNum         = m.param.featAngleResolution;
        info->diameter         = box_diameter(m.sampled.boxMin, m.sampled.boxMax);
        info->samplingDistanceRel = m.samplingDistanceRel;
        std::memcpy(info->boxMin, m.sampled.boxMin, sizeof info->boxMin);
        std::memcpy(info->boxMax, m.sampled.boxMax, sizeof info->boxMax);
    });
}

PPF_API int ppf_model_get_sampled(ppf_detector d, float *pts, float *nrm, float *rePts,
                                  float *reNrm) {
    PPF_REQUIRE(d, "null detector");
    return guarded([&] {
        ReadLock r(d->model);
        if (!d->trained)
            throw std::runtime_error("No trained model");
        const HostModel &m = d->hm;
        if (pts) std::memcpy(pts, m.sampled.p.data(), sizeof(float) * m.sampled.p.size());
        if (nrm) std::memcpy(nrm, m.sampled.n.data(), sizeof(float) * m.sampled.n.size());
        if (rePts) std::memcpy(rePts, m.reSampled.p.data(), sizeof(float) * m.reSampled.p.size());
        if (reNrm) std::memcpy(reNrm, m.reSampled.n.data(), sizeof(float) * m.reSampled.n.size());
    });
}

PPF_API int ppf_model_get_table(ppf_detector d, uint32_t *keys, int *offs, uint32_t *refs,
                                float *alphas) {
    PPF_REQUIRE(d, "null detector");
    return guarded([&] {
        ReadLock r(d->model);
        if (!d->trained)
            throw std::runtime_error("No trained model");
        const HostModel &m = d->hm;
        if (keys) std::memcpy(keys, m.keys.data(), sizeof(uint32_t) * m.keys.size());
        if (offs) std::memcpy(offs, m.offs.data(), sizeof(int) * m.offs.size());
        if (refs) std::memcpy(refs, m.refs.data(), sizeof(uint32_t) * m.refs.size());
        if (alphas) std::memcpy(alphas, m.alphas.data(), sizeof(float) * m.alphas.size());
    });
}

PPF_API int ppf_pair_features(const float p1n1[6], const float *pts, const float *nrm, size_t n,
                              float angleStep, float distStep, int tail, uint32_t *keys,
                              float *alphas) {
    PPF_REQUIRE(p1n1 && (pts || n == 0) && (nrm || n == 0) && keys && alphas, "null argument");
    return guarded([&] {
        hipStream_t s = util_stream();
        TmpCloud    P, N;
        P.up(pts, (int)n, s);
        N.up(nrm, (int)n, s);
        DevBuf kd, ad;
        uint32_t *k = kd.as<uint32_t>(std::max<size_t>(n, 1));
        float    *a = ad.as<float>(std::max<size_t>(n, 1));
        pair_features(p1n1, (int)n, P.x.get<float>(), P.y.get<float>(), P.z.get<float>(),
                      N.x.get<float>(), N.y.get<float>(), N.z.get<float>(), angleStep, distStep,
                      tail, k, a, s);
        PPF_HIP_CHECK(hipMemcpyAsync(keys, k, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
        PPF_HIP_CHECK(hipMemcpyAsync(alphas, a, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        PPF_HIP_CHECK(hipStreamSynchronize(s));
    });
}

PPF_API int ppf_vote(ppf_detector d, const float *pts, const float *nrm, size_t n,
                     const int *keypoints, size_t nk, float minScore, int *cand3, int32_t *accDump,
                     size_t dumpCount, float *poses12, int *nPoses) {
    PPF_REQUIRE(d && pts && nrm && keypoints && cand3, "null argument");
    return guarded([&] {
        ReadLock  r(d->model);
        Lease     L(*d);
        MatchCtx &X = *L;
        if (!d->trained)
            throw std::runtime_error("No trained model in matchScene");
        hipStream_t s = X.stream;
        DevCloud   &c = X.tmpScene;
        c.upload(pts, nrm, (int)n, nullptr, nullptr, s, X.ingest);
        // the vote stage reads positions and normals as given
        size_t b = sizeof(float) * n;
        PPF_HIP_CHECK(hipMemcpyAsync(c.px.get<float>(), c.ox.get<float>(), b, hipMemcpyDeviceToDevice, s));
        PPF_HIP_CHECK(hipMemcpyAsync(c.py.get<float>(), c.oy.get<float>(), b, hipMemcpyDeviceToDevice, s));
        PPF_HIP_CHECK(hipMemcpyAsync(c.pz.get<float>(), c.oz.get<float>(), b, hipMemcpyDeviceToDevice, s));
        PPF_HIP_CHECK(hipMemcpyAsync(c.nx.get<float>(), c.onx.get<float>(), b, hipMemcpyDeviceToDevice, s));
        PPF_HIP_CHECK(hipMemcpyAsync(c.ny.get<float>(), c.ony.get<float>(), b, hipMemcpyDeviceToDevice, s));
        PPF_HIP_CHECK(hipMemcpyAsync(c.nz.get<float>(), c.onz.get<float>(), b, hipMemcpyDeviceToDevice, s));
        const HostModel &hm = d->hm;
        float Dm = box_diameter(hm.sampled.boxMin, hm.sampled.boxMax);
        // every point is a sampled point
        std::vector<int> all(n);
        std::iota(all.begin(), all.end(), 0);
        int *samp = X.idxA.as<int>(std::max<size_t>(n, 1));
        int *kp   = X.idxB.as<int>(std::max<size_t>(nk, 1));
        PPF_HIP_CHECK(hipMemcpyAsync(samp, all.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        PPF_HIP_CHECK(hipMemcpyAsync(kp, keypoints, sizeof(int) * nk, hipMemcpyHostToDevice, s));
        grid_build(X.ws.gVote, c.px.get<float>(), c.py.get<float>(), c.pz.get<float>(),
                   c.nx.get<float>(), c.ny.get<float>(), c.nz.get<float>(), samp, (int)n,
                   Dm / (float)d->vcfg.cellDiv, s, nullptr, true);
        DevBuf dump;
        int32_t *dd = nullptr;
        if (accDump && dumpCount) {
            dd = dump.as<int32_t>(dumpCount * (size_t)d->dm.nSampled * d->dm.angleNum);
            PPF_HIP_CHECK(hipMemsetAsync(dd, 0, sizeof(int32_t) * dumpCount * d->dm.nSampled * d->dm.angleNum, s));
        }
        float voteThreshold = (float)hm.sampled.size() * minScore;
        int   np = vote_and_poses(X.ws, d->dm, X.ws.gVote.view, c.px.get<float>(),
                                  c.py.get<float>(), c.pz.get<float>(), c.nx.get<float>(),
                                  c.ny.get<float>(), c.nz.get<float>(), kp, (int)nk, voteThreshold,
                                  d->vcfg, nullptr, dd, (int)dumpCount, s);
        std::vector<PoseRec> pr(np);
        PPF_HIP_CHECK(hipMemcpyAsync(pr.data(), X.ws.poses.get<PoseRec>(), sizeof(PoseRec) * np,
                                     hipMemcpyDeviceToHost, s));
        if (dd)
            PPF_HIP_CHECK(hipMemcpyAsync(accDump, dd,
                                         sizeof(int32_t) * dumpCount * d->dm.nSampled * d->dm.angleNum,
                                         hipMemcpyDeviceToHost, s));
        PPF_HIP_CHECK(hipStreamSynchronize(s));
        for (size_t k = 0; k < nk * 9; k++)
            cand3[k] = -1;
        for (int i = 0; i < np; i++) {
            int *c3 = cand3 + ((size_t)pr[i].keypoint * 3 + pr[i].rank) * 3;
            c3[0]   = pr[i].refId;
            c3[1]   = pr[i].angleId;
            c3[2]   = (int)pr[i].votes;
            if (poses12)
                std::memcpy(poses12 + 12 * (size_t)i, pr[i].pose, sizeof(float) * 12);
        }
        if (nPoses)
            *nPoses = np;
    });
}

PPF_API int ppf_sample_mesh(const float *pts, size_t n, const int *faces, size_t nFaces,
                            float radius, uint64_t seed, float *points, float *normals,
                            size_t cap, size_t *nOut) {
    PPF_REQUIRE(nOut && (pts || n == 0) && (faces || nFaces == 0), "null argument");
    return guarded([&] {
        hipStream_t s = util_stream();
        if (!points || !normals) {
            *nOut = (size_t)sample_mesh_count(pts, (int)n, faces, (int)nFaces, radius, s);
            return;
        }
        std::vector<float> sp, sn;
        sample_mesh(pts, (int)n, faces, (int)nFaces, radius, seed, sp, sn, s);
        *nOut = sp.size() / 3;
        if (cap < *nOut)
            return;
        std::memcpy(points, sp.data(), sizeof(float) * sp.size());
        std::memcpy(normals, sn.data(), sizeof(float) * sn.size());
    });
}

PPF_API int ppf_voxel_sample(const float *pts, size_t n, const int *subset, size_t nSubset,
                             float step, const float boxMin[3], int *out, size_t *nOut) {
    PPF_REQUIRE(pts && boxMin && out && nOut, "null argument");
    return guarded([&] {
        hipStream_t s = util_stream();
        TmpCloud    P;
        P.up(pts, (int)n, s);
        static Workspace *wsp = new Workspace();  // lives until exit, never destroyed
        Workspace        &ws  = *wsp;
        DevBuf sub, o, cnt;
        int    m  = subset ? (int)nSubset : (int)n;
        int   *sd = nullptr;
        if (subset) {
            sd = sub.as<int>(std::max(m, 1));
            PPF_HIP_CHECK(hipMemcpyAsync(sd, subset, sizeof(int) * m, hipMemcpyHostToDevice, s));
        }
        int *od = o.as<int>(std::max(m, 1));
        int  c  = voxel_sample(ws, P.x.get<float>(), P.y.get<float>(), P.z.get<float>(), sd, m,
                               step, boxMin, od, cnt.as<int>(4), s);
        PPF_HIP_CHECK(hipMemcpyAsync(out, od, sizeof(int) * c, hipMemcpyDeviceToHost, s));
        PPF_HIP_CHECK(hipStreamSynchronize(s));
        *nOut = (size_t)c;
    });
}

PPF_API int ppf_voxel_sample_noalloc(const float *x, const float *y, const float *z, size_t n,
                                     const int *subset, size_t nSubset, float step,
                                     const float boxMin[3], int *out, int *count, void *stream) {
    PPF_REQUIRE(x && y && z && boxMin && out && count && (subset || nSubset == 0),
                "null argument");
    return guarded([&] {
        const size_t m = subset ? nSubset : n;
        if (m > (size_t)kNoAllocMaxPoints)
            throw std::range_error(
                "ppf_voxel_sample_noalloc: at most 1048576 points (its fixed scratch); use "
                "ppf_voxel_sample for more");
        if (voxel_scratch_bytes((int)m) > kNoAllocScratch)
            throw std::range_error("ppf_voxel_sample_noalloc: scratch too small");
        voxel_sample_into(x, y, z, subset, (int)m, step, boxMin, out, count, scratch_ptr(),
                          static_cast<hipStream_t>(stream));
    });
}

PPF_API int ppf_radius_search(const float *pts, size_t n, const float *q, size_t nq, float r2,
                              float cell, int *counts, int *ids, size_t cap) {
    PPF_REQUIRE(pts && q && counts, "null argument");
    return guarded([&] {
        hipStream_t s = util_stream();
        TmpCloud    P, Q;
        P.up(pts, (int)n, s);
        Grid g;
        grid_build(g, P.x.get<float>(), P.y.get<float>(), P.z.get<float>(), nullptr, nullptr,
                   nullptr, nullptr, (int)n, cell, s);
        DevBuf qd, cd, od, idd, tmp;
        float *qq = qd.as<float>(std::max<size_t>(nq, 1) * 3);
        PPF_HIP_CHECK(hipMemcpyAsync(qq, q, sizeof(float) * 3 * nq, hipMemcpyHostToDevice, s));
        int *c = cd.as<int>(nq + 1), *o = od.as<int>(nq + 1);
        if (nq)
            radius_kernel<<<div_up(nq, 128), 128, 0, s>>>(g.view, qq, (int)nq, r2, c);
        PPF_HIP_CHECK(hipMemcpyAsync(counts, c, sizeof(int) * nq, hipMemcpyDeviceToHost, s));
        PPF_HIP_CHECK(hipStreamSynchronize(s));
        if (ids && cap) {
            std::vector<int> off(nq + 1, 0);
            for (size_t t = 0; t < nq; t++)
                off[t + 1] = off[t] + counts[t];
            PPF_HIP_CHECK(hipMemcpyAsync(o, off.data(), sizeof(int) * (nq + 1), hipMemcpyHostToDevice, s));
            int *idv = idd.as<int>(cap);
            if (nq)
                radius_fill_kernel<<<div_up(nq, 128), 128, 0, s>>>(g.view, qq, (int)nq, r2, o, idv,
                                                                    (int)cap);
            size_t w = std::min<size_t>(cap, (size_t)off[nq]);
            PPF_HIP_CHECK(hipMemcpyAsync(ids, idv, sizeof(int) * w, hipMemcpyDeviceToHost, s));
            PPF_HIP_CHECK(hipStreamSynchronize(s));
            // ascending ids per query, as the oracle reports them
            for (size_t t = 0; t < nq; t++) {
                size_t b = std::min<size_t>(off[t], w), e = std::min<size_t>(off[t + 1], w);
                std::sort(ids + b, ids + e);
            }
        }
    });
}

PPF_API int ppf_knn_search(const float *pts, size_t n, const float *q, size_t nq, int k,
                           float cell, int *ids, float *d2) {
    PPF_REQUIRE(pts && q && ids && d2 && k >= 1 && k <= 32, "bad argument");
    return guarded([&] {
        hipStream_t s = util_stream();
        TmpCloud    P;
        P.up(pts, (int)n, s);
        Grid g;
        grid_build(g, P.x.get<float>(), P.y.get<float>(), P.z.get<float>(), nullptr, nullptr,
                   nullptr, nullptr, (int)n, cell, s);
        DevBuf qd, idd, dd;
        float *qq = qd.as<float>(std::max<size_t>(nq, 1) * 3);
        PPF_HIP_CHECK(hipMemcpyAsync(qq, q, sizeof(float) * 3 * nq, hipMemcpyHostToDevice, s));
        int   *iv = idd.as<int>(std::max<size_t>(nq, 1) * k);
        float *dv = dd.as<float>(std::max<size_t>(nq, 1) * k);
        if (nq)
            knn_kernel<<<div_up(nq, 128), 128, 0, s>>>(g.view, qq, (int)nq, k, iv, dv);
        PPF_HIP_CHECK(hipMemcpyAsync(ids, iv, sizeof(int) * nq * k, hipMemcpyDeviceToHost, s));
        PPF_HIP_CHECK(hipMemcpyAsync(d2, dv, sizeof(float) * nq * k, hipMemcpyDeviceToHost, s));
        PPF_HIP_CHECK(hipStreamSynchronize(s));
    });
}

PPF_API int ppf_estimate_normals(float *pts, float *nrm, size_t n, const int *valid,
                                 size_t nValid, const int *query, size_t nq, int k, int smooth,
                                 int invert, const float viewPoint[3]) {
    PPF_REQUIRE(pts && nrm && valid && query, "null argument");
    return guarded([&] {
        hipStream_t s = util_stream();
        TmpCloud    P, N;
        P.up(pts, (int)n, s);
        N.up(nrm, (int)n, s);
        DevBuf vd, qd;
        int   *v  = vd.as<int>(std::max<size_t>(nValid, 1));
        int   *qq = qd.as<int>(std::max<size_t>(nq, 1));
        PPF_HIP_CHECK(hipMemcpyAsync(v, valid, sizeof(int) * nValid, hipMemcpyHostToDevice, s));
        PPF_HIP_CHECK(hipMemcpyAsync(qq, query, sizeof(int) * nq, hipMemcpyHostToDevice, s));
        static Workspace *wsp = new Workspace();  // lives until exit, never destroyed
        Workspace        &ws  = *wsp;
        // cell ~ typical kNN radius: twice the cube root of volume per point
        std::vector<int> vv(valid, valid + nValid);
        float lo[3], hi[3];
        bbox_host(pts, vv, lo, hi);
        double vol = 1.0;
        for (int a = 0; a < 3; a++)
            vol *= (double)(hi[a] - lo[a]) + 1e-3;
        float cell = (float)(2.0 * std::cbrt(vol / std::max<size_t>(nValid, 1)));
        estimate_normals(ws, (int)n, P.x.get<float>(), P.y.get<float>(), P.z.get<float>(),
                         N.x.get<float>(), N.y.get<float>(), N.z.get<float>(), v, (int)nValid, qq,
                         (int)nq, k, smooth, invert, viewPoint, cell, s);
        DevBuf outb;
        float *o = outb.as<float>(std::max<size_t>(n, 1) * 3);
        int    m = std::max<int>((int)n, 1);
        std::vector<int> all(n);
        std::iota(all.begin(), all.end(), 0);
        DevBuf ad;
        int *alld = ad.as<int>(m);
        PPF_HIP_CHECK(hipMemcpyAsync(alld, all.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        if (n) {
            gather_aos_kernel<<<div_up(n, 256), 256, 0, s>>>(alld, (int)n, P.x.get<float>(),
                                                              P.y.get<float>(), P.z.get<float>(), o);
            PPF_HIP_CHECK(hipMemcpyAsync(pts, o, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
            PPF_HIP_CHECK(hipStreamSynchronize(s));
            gather_aos_kernel<<<div_up(n, 256), 256, 0, s>>>(alld, (int)n, N.x.get<float>(),
                                                              N.y.get<float>(), N.z.get<float>(), o);
            PPF_HIP_CHECK(hipMemcpyAsync(nrm, o, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
        }
        PPF_HIP_CHECK(hipStreamSynchronize(s));
    });
}

PPF_API int ppf_estimate_normals_mls(float *pts, float *nrm, size_t n, const int *valid,
                                     size_t nValid, const int *query, size_t nq, float radius,
                                     int order, int invert, const float viewPoint[3],
                                     ppf_mls_stats *stats) {
    PPF_REQUIRE(pts && nrm && (valid || nValid == 0) && (query || nq == 0), "null argument");
    return guarded([&] {
        for (size_t t = 0; t < nq; t++)
            if (query[t] < 0 || (size_t)query[t] >= n)
                throw std::range_error("Invalid Input: MLS query index out of range");
        for (size_t t = 0; t < nValid; t++)
            if (valid[t] < 0 || (size_t)valid[t] >= n)
                throw std::range_error("Invalid Input: MLS valid index out of range");
        hipStream_t s = util_stream();
        static Workspace *wsp = new Workspace();  // lives until exit, never destroyed
        Workspace        &ws  = *wsp;
        TmpCloud P, O, N;
        P.up(pts, (int)n, s);
        O.up(pts, (int)n, s);
        N.up(nrm, (int)n, s);
        DevBuf vd, qd;
        int   *v  = vd.as<int>(std::max<size_t>(nValid, 1));
        int   *qq = qd.as<int>(std::max<size_t>(nq, 1));
        PPF_HIP_CHECK(hipMemcpyAsync(v, valid, sizeof(int) * nValid, hipMemcpyHostToDevice, s));
        PPF_HIP_CHECK(hipMemcpyAsync(qq, query, sizeof(int) * nq, hipMemcpyHostToDevice, s));
        hipEvent_t ev[2];
        PPF_HIP_CHECK(hipEventCreate(&ev[0]));
        PPF_HIP_CHECK(hipEventCreate(&ev[1]));
        PPF_HIP_CHECK(hipEventRecord(ev[0], s));
        MlsStats st;
        estimate_normals_mls(ws, (int)n, P.x.get<float>(), P.y.get<float>(), P.z.get<float>(),
                             O.x.get<float>(), O.y.get<float>(), O.z.get<float>(),
                             N.x.get<float>(), N.y.get<float>(), N.z.get<float>(), v, (int)nValid,
                             qq, (int)nq, radius, order, invert, viewPoint, &st, s);
        PPF_HIP_CHECK(hipEventRecord(ev[1], s));
        DevBuf           outb, ad;  // outlive the final synchronisation
        std::vector<int> all(n);
        if (n) {
            float *o    = outb.as<float>(n * 3);
            int   *alld = ad.as<int>(n);
            std::iota(all.begin(), all.end(), 0);
            PPF_HIP_CHECK(hipMemcpyAsync(alld, all.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
            gather_aos_kernel<<<div_up(n, 256), 256, 0, s>>>(alld, (int)n, O.x.get<float>(),
                                                              O.y.get<float>(), O.z.get<float>(), o);
            PPF_HIP_CHECK(hipMemcpyAsync(pts, o, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
            PPF_HIP_CHECK(hipStreamSynchronize(s));
            gather_aos_kernel<<<div_up(n, 256), 256, 0, s>>>(alld, (int)n, N.x.get<float>(),
                                                              N.y.get<float>(), N.z.get<float>(), o);
            PPF_HIP_CHECK(hipMemcpyAsync(nrm, o, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
        }
        PPF_HIP_CHECK(hipStreamSynchronize(s));
        float ms = 0.0f;
        PPF_HIP_CHECK(hipEventElapsedTime(&ms, ev[0], ev[1]));
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        if (stats) {
            stats->neighbours = st.neighbours;
            stats->queries    = st.queries;
            stats->deviceMs   = ms;
        }
    });
}

PPF_API int ppf_icp_target_create(const float *pts, const float *nrm, size_t n, float cell,
                                  ppf_icp_target *out) {
    PPF_REQUIRE(pts && nrm && out && n > 0, "bad argument");
    *out = nullptr;
    return guarded([&] {
        auto t = std::make_unique<ppf_icp_target_s>();
        PPF_HIP_CHECK(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
        for (auto &e : t->ev)
            PPF_HIP_CHECK(hipEventCreate(&e));
        hipStream_t s = t->stream;
        int         m = (int)n;
        float      *A = t->aos.as<float>((size_t)m * 3);
        PPF_HIP_CHECK(hipMemcpyAsync(A, pts, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
        split_kernel<<<div_up(m, 256), 256, 0, s>>>(A, m, t->x.as<float>(m), t->y.as<float>(m),
                                                     t->z.as<float>(m));
        PPF_HIP_CHECK(hipStreamSynchronize(s));
        PPF_HIP_CHECK(hipMemcpyAsync(A, nrm, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
        split_kernel<<<div_up(m, 256), 256, 0, s>>>(A, m, t->nx.as<float>(m), t->ny.as<float>(m),
                                                     t->nz.as<float>(m));
        if (!(cell > 0.0f)) {
            std::vector<int> v(n);
            std::iota(v.begin(), v.end(), 0);
            float lo[3], hi[3];
            bbox_host(pts, v, lo, hi);
            double vol = 1.0;
            for (int a = 0; a < 3; a++)
                vol *= (double)(hi[a] - lo[a]) + 1e-3;
            cell = (float)(2.0 * std::cbrt(vol / (double)n));
        }
        grid_build(t->grid, t->x.get<float>(), t->y.get<float>(), t->z.get<float>(),
                   t->nx.get<float>(), t->ny.get<float>(), t->nz.get<float>(), nullptr, m, cell, s,
                   nullptr, true);
        grid_build(t->full, t->x.get<float>(), t->y.get<float>(), t->z.get<float>(), nullptr,
                   nullptr, nullptr, nullptr, m, cell, s, nullptr, true);
        t->n = m;
        *out = t.release();
    });
}

PPF_API void ppf_icp_target_destroy(ppf_icp_target t) { delete t; }

static void upload_src(ppf_icp_target_s &t, const float *src, const float *nrm, int n) {
    hipStream_t s = t.stream;
    int         m = std::max(n, 1);
    float      *A = t.saos.as<float>((size_t)m * 3);
    PPF_HIP_CHECK(hipMemcpyAsync(A, src, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
    split_kernel<<<div_up(m, 256), 256, 0, s>>>(A, n, t.sx.as<float>(m), t.sy.as<float>(m),
                                                 t.sz.as<float>(m));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
    t.snx.as<float>(m);
    t.sny.as<float>(m);
    t.snz.as<float>(m);
    if (nrm) {
        PPF_HIP_CHECK(hipMemcpyAsync(A, nrm, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
        split_kernel<<<div_up(m, 256), 256, 0, s>>>(A, n, t.snx.get<float>(), t.sny.get<float>(),
                                                     t.snz.get<float>());
        PPF_HIP_CHECK(hipStreamSynchronize(s));
    }
}

PPF_API int ppf_icp_refine(ppf_icp_target t, const float *src, const float *srcNrm, size_t nSrc,
                           const float *initPoses16, size_t nPoses, const ppf_icp_criteria *c,
                           int mode, int batched, int block, float scoringDist,
                           ppf_icp_result *out, float *deviceMs) {
    PPF_REQUIRE(t && src && srcNrm && initPoses16 && c && out, "null argument");
    return guarded([&] {
        upload_src(*t, src, srcNrm, (int)nSrc);
        std::vector<float> init(12 * nPoses);
        for (size_t h = 0; h < nPoses; h++)
            pose16_to_12(initPoses16 + 16 * h, &init[12 * h]);
        IcpCriteria ic{c->iterations, c->rejectDist, c->mseMin, c->mseMax, c->tolerance};
        std::vector<IcpResultDev> res(nPoses);
        hipStream_t s = t->stream;
        PPF_HIP_CHECK(hipEventRecord(t->ev[0], s));
        if (batched) {
            IcpConfig cfg;
            cfg.mode   = mode;
            cfg.block  = block > 0 ? block : 256;
            cfg.spread = batched == 2 ? 0 : 1;
            icp_batch(t->ws, (int)nSrc, t->sx.get<float>(), t->sy.get<float>(), t->sz.get<float>(),
                      t->snx.get<float>(), t->sny.get<float>(), t->snz.get<float>(),
                      t->grid.view, t->full.view, init.data(), (int)nPoses, ic, scoringDist, cfg,
                      nullptr, s);
        } else {
            icp_plain(t->ws, (int)nSrc, t->sx.get<float>(), t->sy.get<float>(), t->sz.get<float>(),
                      t->snx.get<float>(), t->sny.get<float>(), t->snz.get<float>(),
                      t->grid.view, t->full.view, init.data(), (int)nPoses, ic, scoringDist, mode,
                      nullptr, s);
        }
        PPF_HIP_CHECK(hipEventRecord(t->ev[1], s));
        PPF_HIP_CHECK(hipMemcpyAsync(res.data(), t->ws.icpOut.get<IcpResultDev>(),
                                     sizeof(IcpResultDev) * nPoses, hipMemcpyDeviceToHost, s));
        PPF_HIP_CHECK(hipStreamSynchronize(s));
        if (deviceMs)
            *deviceMs = elapsed(t->ev[0], t->ev[1]);
        for (size_t h = 0; h < nPoses; h++) {
            ppf_icp_result &o = out[h];
            const float    *P = res[h].pose;
            for (int r = 0; r < 3; r++)
                for (int c2 = 0; c2 < 4; c2++)
                    o.pose16[c2 * 4 + r] = P[r * 4 + c2];
            o.pose16[3] = o.pose16[7] = o.pose16[11] = 0.0f;
            o.pose16[15]   = 1.0f;
            o.type         = res[h].type;
            o.mse          = res[h].mse;
            o.convergeRate = res[h].convergeRate;
            o.iterations   = res[h].iterations;
            o.converged    = res[h].converged;
            o.inliers      = res[h].inliers;
        }
    });
}

static void write_icp_result(const IcpResultDev &r, ppf_icp_result &o) {
    const float *P = r.pose;
    for (int row = 0; row < 3; row++)
        for (int col = 0; col < 4; col++)
            o.pose16[col * 4 + row] = P[row * 4 + col];
    o.pose16[3] = o.pose16[7] = o.pose16[11] = 0.0f;
    o.pose16[15]   = 1.0f;
    o.type         = r.type;
    o.mse          = r.mse;
    o.convergeRate = r.convergeRate;
    o.iterations   = r.iterations;
    o.converged    = r.converged;
    o.inliers      = r.inliers;
}

PPF_API int ppf_icp_refine_multi(const ppf_icp_target *targets, size_t nTargets,
                                 const int *poseTarget, const float *src, const float *srcNrm,
                                 size_t nSrc, const float *initPoses16, size_t nPoses,
                                 const ppf_icp_criteria *c, int mode, int batched, int block,
                                 float scoringDist, ppf_icp_result *out, float *deviceMs) {
    PPF_REQUIRE(targets && nTargets > 0 && poseTarget && src && srcNrm && initPoses16 && c && out,
                "null argument");
    for (size_t i = 0; i < nTargets; i++)
        PPF_REQUIRE(targets[i], "null target");
    for (size_t h = 0; h < nPoses; h++)
        PPF_REQUIRE(poseTarget[h] >= 0 && (size_t)poseTarget[h] < nTargets,
                    "pose target index out of range");
    if (nPoses == 0)
        return PPF_OK;
    return guarded([&] {
        // the first target's stream, scratch and source buffers serve the call
        ppf_icp_target_s &t0 = *targets[0];
        hipStream_t       s  = t0.stream;
        upload_src(t0, src, srcNrm, (int)nSrc);
        std::vector<float> init(12 * nPoses);
        for (size_t h = 0; h < nPoses; h++)
            pose16_to_12(initPoses16 + 16 * h, &init[12 * h]);
        IcpCriteria ic{c->iterations, c->rejectDist, c->mseMin, c->mseMax, c->tolerance};
        std::vector<IcpResultDev> res(nPoses);
        PPF_HIP_CHECK(hipEventRecord(t0.ev[0], s));
        if (batched) {
            std::vector<GridView> v(2 * nTargets);
            for (size_t i = 0; i < nTargets; i++) {
                v[i]            = targets[i]->grid.view;
                v[nTargets + i] = targets[i]->full.view;
            }
            GridView *gv = t0.views.as<GridView>(2 * nTargets);
            int      *pt = t0.poseTarget.as<int>(nPoses);
            PPF_HIP_CHECK(hipMemcpyAsync(gv, v.data(), sizeof(GridView) * v.size(),
                                         hipMemcpyHostToDevice, s));
            PPF_HIP_CHECK(hipMemcpyAsync(pt, poseTarget, sizeof(int) * nPoses,
                                         hipMemcpyHostToDevice, s));
            IcpConfig cfg;
            cfg.mode   = mode;
            cfg.block  = block > 0 ? block : 256;
            cfg.spread = batched == 2 ? 0 : 1;
            const ppf_icp_target_s &tf = *targets[poseTarget[0]];
            icp_batch(t0.ws, (int)nSrc, t0.sx.get<float>(), t0.sy.get<float>(), t0.sz.get<float>(),
                      t0.snx.get<float>(), t0.sny.get<float>(), t0.snz.get<float>(), tf.grid.view,
                      tf.full.view, init.data(), (int)nPoses, ic, scoringDist, cfg, nullptr, s, gv,
                      gv + nTargets, pt);
            PPF_HIP_CHECK(hipEventRecord(t0.ev[1], s));
            PPF_HIP_CHECK(hipMemcpyAsync(res.data(), t0.ws.icpOut.get<IcpResultDev>(),
                                         sizeof(IcpResultDev) * nPoses, hipMemcpyDeviceToHost, s));
            PPF_HIP_CHECK(hipStreamSynchronize(s));
        } else {
            // the conventional refiner, one pose at a time against its target
            for (size_t h = 0; h < nPoses; h++) {
                const ppf_icp_target_s &t = *targets[poseTarget[h]];
                icp_plain(t0.ws, (int)nSrc, t0.sx.get<float>(), t0.sy.get<float>(),
                          t0.sz.get<float>(), t0.snx.get<float>(), t0.sny.get<float>(),
                          t0.snz.get<float>(), t.grid.view, t.full.view, &init[12 * h], 1, ic,
                          scoringDist, mode, &res[h], s);
            }
            PPF_HIP_CHECK(hipEventRecord(t0.ev[1], s));
            PPF_HIP_CHECK(hipStreamSynchronize(s));
        }
        if (deviceMs)
            *deviceMs = elapsed(t0.ev[0], t0.ev[1]);
        for (size_t h = 0; h < nPoses; h++)
            write_icp_result(res[h], out[h]);
    });
}

PPF_API int ppf_icp_adjoint(ppf_icp_target t, const float *src, size_t nSrc, const float *poses16,
                            const double *xi6, size_t nPoses, float rejectDist, int center,
                            double *out, int *corr) {
    PPF_REQUIRE(t && src && poses16 && xi6 && out, "null argument");
    return guarded([&] {
        upload_src(*t, src, nullptr, (int)nSrc);
        std::vector<float> T0(12 * nPoses);
        for (size_t h = 0; h < nPoses; h++)
            pose16_to_12(poses16 + 16 * h, &T0[12 * h]);
        icp_adjoint(t->ws, (int)nSrc, t->sx.get<float>(), t->sy.get<float>(), t->sz.get<float>(),
                    t->grid.view, T0.data(), xi6, (int)nPoses, rejectDist, center, out, corr,
                    t->stream);
    });
}

PPF_API int ppf_device_math(int op, const float *a, const float *b, size_t n, uint32_t *out) {
    PPF_REQUIRE(a && out, "null argument");
    return guarded([&] {
        hipStream_t s = util_stream();
        DevBuf      ad, bd, od;
        float      *A = ad.as<float>(std::max<size_t>(n, 1));
        float      *B = b ? bd.as<float>(std::max<size_t>(n, 1)) : nullptr;
        uint32_t   *O = od.as<uint32_t>(std::max<size_t>(n, 1));
        PPF_HIP_CHECK(hipMemcpyAsync(A, a, sizeof(float) * n, hipMemcpyHostToDevice, s));
        if (b)
            PPF_HIP_CHECK(hipMemcpyAsync(B, b, sizeof(float) * n, hipMemcpyHostToDevice, s));
        if (n)
            math_kernel<<<div_up(n, 256), 256, 0, s>>>(op, A, B, (int)n, O);
        PPF_HIP_CHECK(hipMemcpyAsync(out, O, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
        PPF_HIP_CHECK(hipStreamSynchronize(s));
    });
}

struct ppf_ply_s {
    PlyCloud c;
};

PPF_API int ppf_ply_read(const char *path, ppf_ply *out) {
    PPF_REQUIRE(path && out, "null argument");
    *out = nullptr;
    return guarded([&] {
        auto p = std::make_unique<ppf_ply_s>();
        ply_read(path, p->c);
        *out = p.release();
    });
}

PPF_API int ppf_ply_info(ppf_ply p, size_t *nPoints, size_t *nNormals, size_t *nFaces) {
    PPF_REQUIRE(p, "null argument");
    if (nPoints)
        *nPoints = p->c.points.size() / 3;
    if (nNormals)
        *nNormals = p->c.normals.size() / 3;
    if (nFaces)
        *nFaces = p->c.faces.size() / 3;
    return PPF_OK;
}

PPF_API int ppf_ply_copy(ppf_ply p, float *points, float *normals, int *faces) {
    PPF_REQUIRE(p, "null argument");
    if (points && !p->c.points.empty())
        std::memcpy(points, p->c.points.data(), sizeof(float) * p->c.points.size());
    if (normals && !p->c.normals.empty())
        std::memcpy(normals, p->c.normals.data(), sizeof(float) * p->c.normals.size());
    if (faces && !p->c.faces.empty())
        std::memcpy(faces, p->c.faces.data(), sizeof(int) * p->c.faces.size());
    return PPF_OK;
}

PPF_API void ppf_ply_free(ppf_ply p) { delete p; }

PPF_API int ppf_ply_write(const char *path, const float *points, const float *normals, size_t n,
                          const int *faces, size_t nFaces, int ascii) {
    PPF_REQUIRE(path && (points || n == 0) && (faces || nFaces == 0), "null argument");
    return guarded([&] { ply_write(path, points, normals, n, faces, nFaces, ascii != 0); });
}

PPF_API int ppf_scene_upload_ply(ppf_detector d, const char *path, const float *viewPoint,
                                 ppf_scene *out) {
    PPF_REQUIRE(d && path && out, "null argument");
    *out = nullptr;
    return guarded([&] {
        PlyCloud c;
        ply_read(path, c);
        const size_t n  = c.points.size() / 3;
        const bool   hn = c.normals.size() == c.points.size();  // type.cpp:36-38
        auto         sc = std::make_unique<ppf_scene_s>();
        Lease        L(*d);
        sc->cloud.upload(c.points.data(), hn ? c.normals.data() : nullptr, (int)n, nullptr,
                         viewPoint, (*L).stream, (*L).ingest);
        *out = sc.release();
    });
}

PPF_API int ppf_measure_ceilings(float *copyGBps, float *ldsAtomicsPerSec) {
    PPF_REQUIRE(copyGBps && ldsAtomicsPerSec, "null argument");
    return guarded([&] {
        hipStream_t s = util_stream();
        hipEvent_t  e0, e1;
        PPF_HIP_CHECK(hipEventCreate(&e0));
        PPF_HIP_CHECK(hipEventCreate(&e1));
        const size_t bytes = size_t(1) << 30;
        DevBuf       a, b, sink;
        uint4       *A = a.as<uint4>(bytes / 16), *B = b.as<uint4>(bytes / 16);
        PPF_HIP_CHECK(hipMemsetAsync(A, 1, bytes, s));
        float best = 1e30f, ms = 0.0f;
        for (int rep = 0; rep < 4; rep++) {
            PPF_HIP_CHECK(hipEventRecord(e0, s));
            ceiling_copy_kernel<<<8192, 256, 0, s>>>(A, B, bytes / 16);
            PPF_HIP_CHECK(hipEventRecord(e1, s));
            PPF_HIP_CHECK(hipEventSynchronize(e1));
            PPF_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, ms);
        }
        *copyGBps = (float)(2.0 * (double)bytes / (best * 1e-3) / 1e9);
        uint32_t *S = sink.as<uint32_t>(1);
        PPF_HIP_CHECK(hipMemsetAsync(S, 0, 4, s));
        const int blocks = 4096, iters = 4096;
        best             = 1e30f;
        for (int rep = 0; rep < 3; rep++) {
            PPF_HIP_CHECK(hipEventRecord(e0, s));
            ceiling_lds_kernel<<<blocks, 1024, 0, s>>>(iters, S);
            PPF_HIP_CHECK(hipEventRecord(e1, s));
            PPF_HIP_CHECK(hipEventSynchronize(e1));
            PPF_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, ms);
        }
        *ldsAtomicsPerSec = (float)((double)blocks * 1024.0 * iters / (best * 1e-3));
        PPF_HIP_CHECK(hipEventDestroy(e0));
        PPF_HIP_CHECK(hipEventDestroy(e1));
    });
}

}  // extern "C"
