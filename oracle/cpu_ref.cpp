// cpu_ref.cpp -- the CPU baseline's timing mode (test and measurement
// infrastructure: bench.py's cpu_baseline leg loads it; nothing in the
// product links it).  Detector::matchScene (ppf.cpp:149-510) restated on the
// reference's own data structures, so its times stand for the reference's
// CPU path:
//   * the reference's kd-tree (kdtree/kdtree.h + nanoflann.hpp, included
//     from the reference tree by oracle/Makefile, never copied), built over
//     the scene with the caller's box and the valid indices (ppf.cpp:196-198);
//   * greedy Poisson sampling over the tree order vAcc_, OpenMP-parallel
//     as in samplePointCloud (util.cpp:107-139), and the tree reduced to the
//     sampled slots for the key points (ppf.cpp:212-216);
//   * estimateNormal (util.cpp:234-334): kNN through the tree, FP32
//     centroid and covariance, the eigenvector of the smallest |eigenvalue|,
//     the point moved to the centroid, optional smoothing, orientation;
//   * the vote loop (ppf.cpp:262-391): radiusSearch in the reduced tree,
//     the pair features with the scalar-tail rule on the last n mod 8
//     partners in traversal order (util.cpp:783-784), a std::unordered_map
//     from key to bucket (the reference uses gtl::flat_hash_map), the
//     accumulator firstprivate per OpenMP thread;
//   * clustering, ICP and scoring through the oracle's functions (its ICP
//     searches a uniform grid, not the kd-tree).
// Stage times use the reference's Timer labels (helper.h:9-32).  The
// arithmetic of every feature, bin, frame and pose is the oracle's (canonical
// FP32), so only the data structures, the sampling and the order of work
// are the reference's.  Build: oracle/Makefile `make ref` (-O3 -mavx
// -fopenmp -ffp-contract=off, as CMakeLists.txt:44 builds the reference).
#include <kdtree.h>

#include "ppf_oracle.h"

#include <omp.h>

#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace {

using Cloud  = std::vector<std::array<float, 3>>;
using KDTree = KDTreeVectorOfVectorsAdaptor<Cloud, float>;
using Clock  = std::chrono::steady_clock;

double ms_since(Clock::time_point t) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}

struct Feature {  // privateType.h:31-41
    std::vector<uint32_t> refInd;
    std::vector<float>    alphaAngle;
};

struct CpuModel {
    const orc_model                       *m;
    std::unordered_map<uint32_t, Feature> table;
    std::vector<float>                    frames;  // per sampled point, 3x4
};

// samplePointCloud (util.cpp:107-139): walk vAcc_ in parallel; a point not
// yet suppressed is kept and suppresses everything within sampleStep.  The
// keep flags are written by several threads, as in the reference.
std::vector<int> sample_tree(const KDTree &tree, const Cloud &pts, float sampleStep,
                             std::vector<int> *indicesOfIndices) {
    const auto                 &vacc = tree.index->vAcc_;
    const int                   size = (int)vacc.size();
    std::vector<unsigned char>  keep(pts.size(), 1);
    const float                 radius = sampleStep * sampleStep;
    std::vector<int>            result;
    if (indicesOfIndices)
        indicesOfIndices->assign(size, -1);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < size; i++) {
        const int index = vacc[i];
        if (index < 0 || !keep[index])
            continue;
#pragma omp critical
        { result.push_back(index); }
        if (indicesOfIndices)
            (*indicesOfIndices)[i] = index;
        std::vector<nanoflann::ResultItem<int, float>> indices;
        nanoflann::RadiusResultSet<float, int>         rs(radius, indices);
        tree.index->findNeighbors(rs, &pts[index][0], nanoflann::SearchParameters(0, false));
        for (auto &r : indices)
            keep[r.first] = 0;
    }
    return result;
}

// computeNormal (util.cpp:234-270) for point idx; returns its neighbours
void compute_normal(Cloud &pts, std::vector<std::array<float, 3>> &nrm, int idx,
                    const KDTree &tree, int k, std::vector<int> *neighbour) {
    auto &normal = nrm[idx];
    if (std::isfinite(normal[0]) && std::isfinite(normal[1]) && std::isfinite(normal[2]))
        return;
    std::vector<int>   ind(k);
    std::vector<float> d2(k);
    const size_t       got = tree.index->knnSearch(&pts[idx][0], k, ind.data(), d2.data());
    if (got < 3)
        return;
    ind.resize(got);
    float c[3] = {0, 0, 0};
    for (int j : ind)
        for (int a = 0; a < 3; a++)
            c[a] += pts[j][a];
    for (int a = 0; a < 3; a++)
        c[a] /= (float)got;
    float cov[9] = {0};
    for (int j : ind) {
        float d[3] = {pts[j][0] - c[0], pts[j][1] - c[1], pts[j][2] - c[2]};
        for (int r = 0; r < 3; r++)
            for (int q = 0; q < 3; q++)
                cov[r * 3 + q] += d[r] * d[q];
    }
    double A[9], V[9], ev[3];
    for (int i = 0; i < 9; i++)
        A[i] = cov[i];
    orc_jacobi_eig(3, A, V, ev);
    int m = 0;
    for (int i = 1; i < 3; i++)
        if (std::fabs(ev[i]) < std::fabs(ev[m]))
            m = i;
    for (int a = 0; a < 3; a++)
        normal[a] = (float)V[a * 3 + m];
    if (neighbour)
        *neighbour = ind;
    for (int a = 0; a < 3; a++)
        pts[idx][a] = c[a];
}

// estimateNormal (util.cpp:272-334)
void estimate_normals(Cloud &pts, std::vector<std::array<float, 3>> &nrm,
                      const std::vector<int> &indices, const KDTree &tree, int k, bool smooth,
                      bool invert, const float vp[3]) {
    if (k < 3)
        k = 3;
    const int size = (int)indices.size();
#pragma omp parallel for
    for (int i = 0; i < size; i++) {
        const int        idx = indices[i];
        std::vector<int> nb;
        compute_normal(pts, nrm, idx, tree, k, &nb);
        if (nb.empty() || !smooth)
            continue;
        for (int j : nb)
            compute_normal(pts, nrm, j, tree, k, nullptr);
        float s[3] = {0, 0, 0};
        auto &n0   = nrm[idx];
        for (int j : nb) {
            auto       &n   = nrm[j];
            const float dot = n[0] * n0[0] + n[1] * n0[1] + n[2] * n0[2];
            for (int a = 0; a < 3; a++)
                s[a] += dot > 0 ? n[a] : -n[a];
        }
        const float l = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        for (int a = 0; a < 3; a++)
            n0[a] = s[a] / l;
    }
    const bool hasVp = vp && std::isfinite(vp[0]) && std::isfinite(vp[1]) && std::isfinite(vp[2]);
#pragma omp parallel for
    for (int i = 0; i < size; i++) {
        const int idx = indices[i];
        auto     &n   = nrm[idx];
        const float d = hasVp ? n[0] * (vp[0] - pts[idx][0]) + n[1] * (vp[1] - pts[idx][1]) +
                                    n[2] * (vp[2] - pts[idx][2])
                              : n[2];
        if (invert ? d > 0.f : d < 0.f)
            for (int a = 0; a < 3; a++)
                n[a] = -n[a];
    }
}

void fill_pose(orc_pose_t &p) {
    float R[9] = {p.pose[0], p.pose[1], p.pose[2], p.pose[4], p.pose[5],
                  p.pose[6], p.pose[8], p.pose[9], p.pose[10]};
    orc_quat_from_rot(R, p.q);
    const float n = std::sqrt(p.q[1] * p.q[1] + p.q[2] * p.q[2] + p.q[3] * p.q[3]);
    p.angle       = 2.0f * orc_atan2f(n, std::fabs(p.q[0]));
}

}  // namespace

extern "C" {

// stage times (ms), Timer labels of ppf.cpp:171-436
enum {
    ST_REMOVE_NAN, ST_KDTREE, ST_SAMPLE1, ST_NORMAL, ST_SAMPLE2, ST_PPF, ST_CLUSTER,
    ST_ICP_PREPARE, ST_ICP, ST_TOTAL, ST_N
};
typedef struct {
    float ms[ST_N];
    int   nSampled, nKeypoints, nVotedKeypoints, nPoses, nResults, threads;
    long  nPairs, nVotes;
} cpuref_stats;

// the hash table of a trained oracle model as an unordered_map, and the
// sampled model's frames (training is not timed)
void *cpuref_model(const orc_model *m) {
    auto *cm = new CpuModel;
    cm->m    = m;
    const orc_table &t = m->table;
    cm->table.reserve((size_t)t.nKeys * 2);
    for (int k = 0; k < t.nKeys; k++) {
        Feature &f = cm->table[t.keys[k]];
        f.refInd.assign(t.refs + t.offs[k], t.refs + t.offs[k + 1]);
        f.alphaAngle.assign(t.alphas + t.offs[k], t.alphas + t.offs[k + 1]);
    }
    cm->frames.resize((size_t)m->nSampled * 12);
    for (int r = 0; r < m->nSampled; r++)
        orc_frame(m->sampledP + 3 * r, m->sampledN + 3 * r, &cm->frames[(size_t)r * 12]);
    return cm;
}

void cpuref_model_free(void *p) { delete static_cast<CpuModel *>(p); }

// matchScene on a scene without normals.  kpStride > 1 votes only every
// kpStride-th key point (a bounded sample of the vote stage; its time then
// stands for 1 / kpStride of the full stage).  Returns the result count.
int cpuref_match(void *model, int n, const float *pts, const float vp[3], float rel, float kpFrac,
                 float minScore, const orc_match_param *p, int threads, int kpStride,
                 float *poses12, float *scores, int cap, cpuref_stats *st) {
    const CpuModel  &cm = *static_cast<CpuModel *>(model);
    const orc_model &m  = *cm.m;
    std::memset(st, 0, sizeof *st);
    const int prev = omp_get_max_threads();
    omp_set_num_threads(threads);
    st->threads         = threads;
    const auto t0       = Clock::now();
    auto       tStage   = Clock::now();
    // [1] copy, removeNan, box (ppf.cpp:153,170-176)
    Cloud scene(n);
    for (int i = 0; i < n; i++)
        scene[i] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    std::vector<int> valid(n);
    valid.resize(orc_remove_nan(n, pts, nullptr, valid.data()));
    float box[6];
    orc_bounding_box(n, pts, valid.data(), (int)valid.size(), box);
    st->ms[ST_REMOVE_NAN] = (float)ms_since(tStage);
    // [2] derived parameters (ppf.cpp:183-240)
    const float D = std::sqrt((m.boxMax[0] - m.boxMin[0]) * (m.boxMax[0] - m.boxMin[0]) +
                              (m.boxMax[1] - m.boxMin[1]) * (m.boxMax[1] - m.boxMin[1]) +
                              (m.boxMax[2] - m.boxMin[2]) * (m.boxMax[2] - m.boxMin[2]));
    const int   angleNum     = m.param.featAngleResolution;
    const float angleStep    = 6.28318530717958647692f / (float)angleNum;
    const float distanceStep = D * m.param.featDistanceStepRel;
    const int   refNum       = m.nSampled;
    const float sqD          = D * D;
    const float reSampleStep = D * m.param.poseRefRelSamplingDistance;
    const float sampleStep   = D * rel;
    const float voteThreshold = (float)refNum * minScore;
    float maxOverlapDist = 0, poseRefDist = 0, scoringDist = 0;
    if (p->maxOverlapDistRel > 0) maxOverlapDist = D * p->maxOverlapDistRel;
    if (p->maxOverlapDistAbs > 0) maxOverlapDist = p->maxOverlapDistAbs;
    if (p->poseRefDistThresholdRel > 0) poseRefDist = D * p->poseRefDistThresholdRel;
    if (p->poseRefDistThresholdAbs > 0) poseRefDist = p->poseRefDistThresholdAbs;
    if (p->poseRefScoringDistRel > 0) scoringDist = D * p->poseRefScoringDistRel;
    if (p->poseRefScoringDistAbs > 0) scoringDist = p->poseRefScoringDistAbs;

    tStage = Clock::now();
    Cloud  boxCloud = {{box[0], box[1], box[2]}, {box[3], box[4], box[5]}};
    std::vector<int> vIdx(valid);
    KDTree tree(scene, 10, boxCloud, vIdx);
    tree.index->buildIndex();
    st->ms[ST_KDTREE] = (float)ms_since(tStage);

    tStage = Clock::now();
    std::vector<int> slotOfSample;
    std::vector<int> sampled = sample_tree(tree, scene, sampleStep, &slotOfSample);
    st->ms[ST_SAMPLE1] = (float)ms_since(tStage);
    st->nSampled       = (int)sampled.size();

    tStage = Clock::now();
    std::vector<std::array<float, 3>> normal(n, {NAN, NAN, NAN});
    estimate_normals(scene, normal, sampled, tree, p->knnNormal, p->smoothNormal != 0,
                     p->invertNormal != 0, vp);
    st->ms[ST_NORMAL] = (float)ms_since(tStage);

    tStage = Clock::now();
    tree.reduce(slotOfSample);
    const float      keySampleStep = std::sqrt(1.f / kpFrac) * sampleStep;
    std::vector<int> keypoint      = sample_tree(tree, scene, keySampleStep, nullptr);
    st->ms[ST_SAMPLE2] = (float)ms_since(tStage);
    st->nKeypoints     = (int)keypoint.size();

    // [3] vote (ppf.cpp:247-411)
    tStage                  = Clock::now();
    const int stride        = angleNum + 1;
    const int nk            = (int)keypoint.size();
    std::vector<orc_pose_t> slots((size_t)nk * 3);
    std::vector<int>        nSlot(nk, 0);
    long                    pairs = 0, votes = 0;
    int                     voted = 0;
    const int               step  = kpStride > 1 ? kpStride : 1;
#pragma omp parallel reduction(+ : pairs, votes, voted)
    {
        std::vector<int32_t> acc((size_t)refNum * stride);
        std::vector<nanoflann::ResultItem<int, float>> indices;
#pragma omp for schedule(dynamic, 4)
        for (int c = 0; c < nk; c += step) {
            const int    id = keypoint[c];
            const float *p1 = &scene[id][0], *n1 = &normal[id][0];
            if (!(std::isfinite(n1[0]) && std::isfinite(n1[1]) && std::isfinite(n1[2])))
                continue;
            indices.clear();
            const size_t searched = tree.index->radiusSearch(p1, sqD, indices,
                                                             nanoflann::SearchParameters(0, false));
            if ((float)searched < voteThreshold)
                continue;
            voted++;
            std::fill(acc.begin(), acc.end(), 0);
            float Ts[12];
            orc_frame(p1, n1, Ts);
            const int rows = (int)searched - 1;
            const int body = rows - rows % 8;  // util.cpp:783-784
            int       r    = 0;
            for (auto &it : indices) {
                if (it.first == id)
                    continue;
                const float   *p2 = &scene[it.first][0], *n2 = &normal[it.first][0];
                const uint32_t key = orc_ppf_key(p1, n1, p2, n2, angleStep, distanceStep,
                                                 r >= body, nullptr, nullptr);
                r++;
                pairs++;
                auto f = cm.table.find(key);
                if (f == cm.table.end())
                    continue;
                const float alphaS = orc_alpha(Ts, p2);
                if (std::isnan(alphaS))
                    continue;
                const Feature &ft   = f->second;
                const int      size = (int)ft.refInd.size(), vec = size - size % 8;
                for (int e = 0; e < size; e++) {
                    const int bin = orc_angle_bin(ft.alphaAngle[e], alphaS, angleNum, e >= vec);
                    int32_t  *a   = &acc[(size_t)ft.refInd[e] * stride];
                    a[0]++;
                    a[bin + 1]++;
                }
                votes += size;
            }
            orc_candidate cand[3];
            const int     nc = orc_peaks(acc.data(), refNum, angleNum, voteThreshold, cand);
            for (int x = 0; x < nc; x++) {
                orc_pose_t &pp = slots[(size_t)c * 3 + x];
                orc_pose(Ts, &cm.frames[(size_t)cand[x].refId * 12], cand[x].angleId, angleNum,
                         pp.pose);
                pp.votes = cand[x].vote;
                fill_pose(pp);
            }
            nSlot[c] = nc;
        }
    }
    std::vector<orc_pose_t> poseList;
    for (int c = 0; c < nk; c++)
        for (int x = 0; x < nSlot[c]; x++)
            poseList.push_back(slots[(size_t)c * 3 + x]);
    st->ms[ST_PPF]         = (float)ms_since(tStage);
    st->nPairs             = pairs;
    st->nVotes             = votes;
    st->nVotedKeypoints    = voted;
    st->nPoses             = (int)poseList.size();

    // [4] cluster (ppf.cpp:413-420)
    tStage = Clock::now();
    const int               np = (int)poseList.size();
    std::vector<int>        clOf(np ? np : 1);
    std::vector<orc_pose_t> clusters(np ? np : 1), nms(np ? np : 1);
    const int ncl = orc_cluster_poses(poseList.data(), np, 0.1f * D, angleStep, clOf.data(),
                                      clusters.data());
    float center[3];
    for (int a = 0; a < 3; a++)
        center[a] = (m.boxMin[a] + m.boxMax[a]) / 2.0f;
    const int nn = orc_nms_poses(clusters.data(), ncl, center, maxOverlapDist, nms.data());
    st->ms[ST_CLUSTER] = (float)ms_since(tStage);

    // [5] icp prepare (ppf.cpp:427-434): the tree restored, the scene
    // resampled at reSampleStep, normals for the new samples; grids for the
    // oracle's refinement
    tStage = Clock::now();
    tree.restore();
    std::vector<int> dense;
    if (p->densePoseRefinement) {
        dense = sample_tree(tree, scene, reSampleStep, nullptr);
        estimate_normals(scene, normal, dense, tree, p->knnNormal, false, p->invertNormal != 0,
                         vp);
    }
    std::vector<float> P((size_t)n * 3), N((size_t)n * 3);
    for (int i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) {
            P[(size_t)3 * i + a] = scene[i][a];
            N[(size_t)3 * i + a] = normal[i][a];
        }
    orc_grid gSparse, gDense, gFull;
    orc_grid_build(&gSparse, n, P.data(), sampled.data(), (int)sampled.size(), sampleStep);
    orc_grid_build(&gFull, n, P.data(), valid.data(), (int)valid.size(),
                   scoringDist > 0 ? scoringDist * 4.0f : sampleStep);
    if (p->densePoseRefinement)
        orc_grid_build(&gDense, n, P.data(), dense.data(), (int)dense.size(), reSampleStep * 2.0f);
    st->ms[ST_ICP_PREPARE] = (float)ms_since(tStage);

    // [6] icp + scoring + the ordered accept walk (ppf.cpp:436-507)
    tStage = Clock::now();
    const orc_icp_criteria sparseC = {5, poseRefDist, sampleStep * 0.5f, sampleStep, 0.05f};
    const orc_icp_criteria denseC  = {p->poseRefNumSteps, poseRefDist, reSampleStep * 0.5f,
                                      sampleStep, 0.05f};
    std::vector<std::pair<float, std::array<float, 12>>> res;
    for (int c = 0; c < nn; c++) {
        std::array<float, 12> pose;
        std::memcpy(pose.data(), nms[c].pose, sizeof(float) * 12);
        float score = nms[c].votes;
        if (score < voteThreshold)
            continue;
        if (p->sparsePoseRefinement) {
            orc_icp_result ir;
            orc_icp_regist(refNum, m.sampledP, m.sampledN, &gSparse, N.data(), pose.data(),
                           &sparseC, 0, &ir);
            if (!ir.converged)
                continue;
            std::memcpy(pose.data(), ir.pose, sizeof(float) * 12);
            score = std::min(1.0f, (float)orc_inliers(refNum, m.sampledP, pose.data(), &gFull,
                                                      scoringDist) / (float)refNum);
            if (score < minScore)
                continue;
            if (p->densePoseRefinement) {
                orc_icp_regist(m.nReSampled, m.reSampledP, m.reSampledN, &gDense, N.data(),
                               pose.data(), &denseC, 0, &ir);
                if (!ir.converged)
                    continue;
                std::memcpy(pose.data(), ir.pose, sizeof(float) * 12);
                score = std::min(1.0f, (float)orc_inliers(m.nReSampled, m.reSampledP,
                                                          pose.data(), &gFull, scoringDist) /
                                           (float)m.nReSampled);
                if (score < minScore)
                    continue;
            }
        }
        res.emplace_back(score, pose);
        if ((int)res.size() >= p->numMatches)
            break;
    }
    std::stable_sort(res.begin(), res.end(),
                     [](const auto &x, const auto &y) { return x.first > y.first; });
    st->ms[ST_ICP] = (float)ms_since(tStage);
    orc_grid_free(&gSparse);
    orc_grid_free(&gFull);
    if (p->densePoseRefinement)
        orc_grid_free(&gDense);
    const int nOut = std::min((int)res.size(), cap);
    for (int i = 0; i < nOut; i++) {
        if (poses12)
            std::memcpy(poses12 + 12 * i, res[i].second.data(), sizeof(float) * 12);
        if (scores)
            scores[i] = res[i].first;
    }
    st->nResults    = (int)res.size();
    st->ms[ST_TOTAL] = (float)ms_since(t0);
    omp_set_num_threads(prev);
    return (int)res.size();
}
}
