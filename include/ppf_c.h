/*
 * ppf_c.h -- C ABI of the MI355X-native PPF detection library
 * (libppf_amd.so).  It is the drop-in boundary for SurfaceMan/surface_match's
 * ppf::Detector (ppf.h:10-62): every entry point takes plain pointers and
 * sizes, reports errors through a status code plus ppf_last_error(), and is
 * synchronous at the API edge, like the reference's member functions.
 *
 * Point clouds cross the boundary as the reference's PointCloud does
 * (type.h:22-33): points and normals as n x 3 float arrays (the memory
 * layout of std::vector<Eigen::Vector3f>), a bounding box as 6 floats
 * {min.x,min.y,min.z,max.x,max.y,max.z} (diameter 0 or NULL = compute it),
 * and a view point as 3 floats (NULL or NaN = none).  Poses are returned as
 * 4x4 column-major floats, the layout of Eigen::Matrix4f.
 *
 * Status codes map onto the reference's exceptions: PPF_ERR_RANGE is
 * std::range_error, PPF_ERR_RUNTIME is std::runtime_error (same messages).
 */
#ifndef PPF_C_H
#define PPF_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define PPF_API __attribute__((visibility("default")))
#else
#define PPF_API
#endif

enum {
    PPF_OK          = 0,
    PPF_ERR_RANGE   = 1, /* std::range_error in the reference   */
    PPF_ERR_RUNTIME = 2, /* std::runtime_error in the reference */
    PPF_ERR_DEVICE  = 3, /* HIP runtime / device failure        */
    PPF_ERR_ARG     = 4  /* NULL handle or buffer               */
};

typedef struct ppf_detector_s *ppf_detector;
typedef struct ppf_scene_s    *ppf_scene;
typedef struct ppf_ply_s      *ppf_ply;

/* TrainParam, type.h:39-74 (smoothNormal is the reference's bool). */
typedef struct {
    float featDistanceStepRel;
    int   featAngleResolution;
    float poseRefRelSamplingDistance;
    int   knnNormal;
    int   smoothNormal;
} ppf_train_param;

/* MatchParam, type.h:80-176. */
typedef struct {
    int   numMatches;
    int   knnNormal;
    int   smoothNormal;
    int   invertNormal;
    float maxOverlapDistRel;
    float maxOverlapDistAbs;
    int   sparsePoseRefinement;
    int   densePoseRefinement;
    int   poseRefNumSteps;
    float poseRefDistThresholdRel;
    float poseRefDistThresholdAbs;
    float poseRefScoringDistRel;
    float poseRefScoringDistAbs;
} ppf_match_param;

/* MatchResult, type.h:178-181: the sampled scene and the key points.  The
 * caller supplies buffers of the given capacities (points); the library
 * writes min(count, cap) points and always reports the full count. */
typedef struct {
    float *sampledPoints, *sampledNormals;
    size_t sampledCap, nSampled;
    float *keyPoints, *keyNormals;
    size_t keyCap, nKey;
} ppf_match_result;

/* Work counters and stage timings of one match (no reference counterpart;
 * the stage names are the reference's Timer labels, ppf.cpp:196-436). */
typedef struct {
    int    nScene, nValid, nSampled, nKeypoints, nVotedKeypoints, nPoses, nClusters,
           nHypotheses, nSparseIcp, nDenseIcp, nResults, nDense;
    unsigned long long nPairs, nKeyHits, nVotes;
    float  msSample, msNormals, msKeypoints, msVote, msCluster, msIcpPrepare, msIcp, msTotal;
    float  msVoteKernel; /* device time of the vote kernel alone */
    float  msIcpSparse, msIcpDense; /* device time of the two batched refinements */
    float  msPairKernel;  /* of msVoteKernel: the pair/key-hit kernel (split vote path) */
    float  msHitsKernel;  /* of msVoteKernel: hit voting + peaks (split vote path) */
    int    nVoteFallback; /* key points voted by the fused kernel for lack of hit capacity */
    float  msIngest;      /* ppf_match_scene: host wall time of the scene's upload (copy to
                             HBM, removeNan, bounding box; ppf.cpp:153,170-176) */
} ppf_match_stats;

/* -- parameters (TrainParam/MatchParam constructors, type.cpp:71-127) -- */
PPF_API int ppf_train_param_default(ppf_train_param *p);
PPF_API int ppf_match_param_default(ppf_match_param *p);
PPF_API int ppf_train_param_check(const ppf_train_param *p);
PPF_API int ppf_match_param_check(const ppf_match_param *p);

/* -- Detector (ppf.h:10-62) -- */
PPF_API int  ppf_detector_create(ppf_detector *out);          /* Detector::Detector  ppf.cpp:26-31 */
PPF_API void ppf_detector_destroy(ppf_detector d);            /* Detector::~Detector ppf.cpp:33-34 */

/* Detector::trainModel, ppf.h:25-26 / ppf.cpp:36-147 */
PPF_API int ppf_train_model(ppf_detector d, const float *points, const float *normals, size_t n,
                            const float *box6, const float *viewPoint, float samplingDistanceRel,
                            const ppf_train_param *param);

/* trainModel for a model with faces (ppf.cpp:63-70): the mesh is sampled
 * (sampleMesh, util.cpp:20-84) at radius reSampleStep / 4 with area-weighted
 * points and the triangles' normals, and the sample trained on with the
 * given model's box.  faces: nFaces x 3 point indices.  The reference seeds
 * its sampler from std::random_device; here the uniforms come from
 * Philox4x32-10 keyed by seed, so a seed fixes the model.  nFaces == 0 is
 * ppf_train_model. */
PPF_API int ppf_train_model_mesh(ppf_detector d, const float *points, const float *normals,
                                 size_t n, const int *faces, size_t nFaces, const float *box6,
                                 const float *viewPoint, float rel, const ppf_train_param *param,
                                 uint64_t seed);

/* Detector::matchScene, ppf.h:39-42 / ppf.cpp:149-510.  Writes up to cap
 * poses (16 floats each, column-major 4x4) and scores; *nOut = number of
 * matches (<= numMatches).  result and stats may be NULL. */
PPF_API int ppf_match_scene(ppf_detector d, const float *points, const float *normals, size_t n,
                            const float *box6, const float *viewPoint, float samplingDistanceRel,
                            float keyPointFraction, float minScore, const ppf_match_param *param,
                            float *poses16, float *scores, size_t cap, size_t *nOut,
                            ppf_match_result *result, ppf_match_stats *stats);

/* Detector::save / load, ppf.h:49,57 / ppf.cpp:512-547 (model file layout
 * of serialize.cpp:5-186, unchanged). */
PPF_API int ppf_save(ppf_detector d, const char *path);
PPF_API int ppf_load(ppf_detector d, const char *path);

/* message of the last failed call on this thread (what() of the
 * reference's exception) */
PPF_API const char *ppf_last_error(void);

/* -- device-resident scenes (no reference counterpart: the reference has no
 * device; these keep a scene in HBM so repeated matches skip the upload) -- */
PPF_API int ppf_scene_upload(ppf_detector d, const float *points, const float *normals, size_t n,
                             const float *box6, const float *viewPoint, ppf_scene *out);
PPF_API void ppf_scene_free(ppf_scene s);
/* what an upload kept, as matchScene's first steps see it (ppf.cpp:153,
 * 170-176): the points and normals as given (n x 3 each; normals only when
 * the scene has them), the ascending indices of the valid points (finite
 * point and, with normals, finite normal: removeNan, util.cpp:86-105) and
 * the box {min xyz, max xyz} (the caller's, else computeBoundingBox of the
 * valid points, util.cpp:178-207).  Any output may be NULL; points / normals
 * need room for n, valid for *nValid points (sizes from a first call with
 * NULL buffers). */
PPF_API int ppf_scene_download(ppf_scene s, size_t *n, size_t *nValid, int *hasNormals,
                               float *points, float *normals, int *valid, float box6[6]);
PPF_API int ppf_match_scene_resident(ppf_detector d, ppf_scene s, float samplingDistanceRel,
                                     float keyPointFraction, float minScore,
                                     const ppf_match_param *param, float *poses16, float *scores,
                                     size_t cap, size_t *nOut, ppf_match_stats *stats);
/* inliner (util.cpp:628-641) as matchScene scores a refined pose
 * (ppf.cpp:456-459, 479-482): for each pose (column-major 4x4), the number
 * of sampled (dense = 0) or resampled (dense = 1) model points whose
 * transformed position has a valid point of the scene within dist (dist2 <
 * dist^2).  The scene is taken as the last match left it (points that got a
 * fresh normal sit at their neighbourhood centroid, estimateNormal's
 * write-back), so this scores poses exactly as that match did. */
PPF_API int ppf_scene_inliers(ppf_detector d, ppf_scene s, const float *poses16, size_t nPoses,
                              int dense, float dist, int *inliers);
/* -- several resident scenes in one pass (no reference counterpart: the
 * reference matches one scene per call).  Each scene is sampled and given
 * normals on its own; the key points of all scenes are voted in one launch
 * (each against its own scene) and the hypotheses of all scenes are refined
 * in one batched refinement per stage.  Scene i's results (at most cap) go
 * to poses16[16 * cap * i ..] and scores[cap * i ..], their count to
 * nOut[i]; they equal ppf_match_scene_resident's for that scene.  stats
 * (optional) sums the counts over the scenes and times the whole batch. -- */
PPF_API int ppf_match_scenes_resident(ppf_detector d, const ppf_scene *scenes, size_t nScenes,
                                      float samplingDistanceRel, float keyPointFraction,
                                      float minScore, const ppf_match_param *param,
                                      float *poses16, float *scores, size_t cap, size_t *nOut,
                                      ppf_match_stats *stats);

/* -- launch tuning (no reference counterpart).  ppf_autotune times the
 * vote stage over {split, fused} kernels, then workgroup sizes {256, 512,
 * 1024}, then the vote-grid cell {D/2, D/3, D/4}, then reference tiles
 * {1, 2, 3, 4}, then the key point order {Morton, plain}, then the XCD
 * mapping {round-robin, contiguous} (one knob at a time), then the batched ICP
 * over {spread, persistent} x {64, 128, 256, 512, 1024} threads per
 * workgroup, then the normals grid cell over {1/4, 1/2, 1} sample step, on
 * the given resident scene, and keeps the fastest in the detector.  None of
 * these changes any result. -- */
typedef struct {
    int voteBlock;    /* threads per vote workgroup: 256, 512 or 1024 */
    int voteXcdRemap; /* 1: consecutive key points stay on one XCD */
    int voteCellDiv;  /* vote grid cell = model diameter / voteCellDiv: 2, 3 or 4 */
    int icpBlock;     /* threads per ICP workgroup: 64 .. 1024, a power of two */
    int icpSpread;    /* 1: each hypothesis spread over workgroups, 0: persistent */
    int normCell4;    /* scene normals grid cell = sample step * normCell4 / 4: 1 .. 16 */
    int voteSplit;    /* 1: pair kernel + hit voting kernel, 0: one fused kernel */
    int voteHitCap;   /* key hits kept per key point on the split path (>= 1); key points
                         with more are voted by the fused kernel */
    int voteTiles;    /* at least this many reference tiles per key point (1 .. 16): smaller
                         LDS accumulators, more workgroups per CU */
    int voteOrder;    /* 1: workgroups take key points in (scene, Morton code of position)
                         order, 0: in key point order */
} ppf_tuning;
typedef struct {
    ppf_tuning best;
    float      voteMs, icpMs; /* device time of the vote kernel / both refinements */
    int        configsTried;
    float      normalsMs;     /* device time of both normal estimations */
} ppf_tune_report;
PPF_API int ppf_autotune(ppf_detector d, ppf_scene s, float samplingDistanceRel,
                         float keyPointFraction, float minScore, const ppf_match_param *param,
                         ppf_tune_report *report);
PPF_API int ppf_get_tuning(ppf_detector d, ppf_tuning *out);
PPF_API int ppf_set_tuning(ppf_detector d, const ppf_tuning *t);

/* -- trained-model introspection (the IMPL fields, privateType.h:56-66) -- */
typedef struct {
    int   nSampled, nReSampled, nKeys, nEntries, angleNum;
    float diameter, samplingDistanceRel;
    float boxMin[3], boxMax[3];
} ppf_model_info;
PPF_API int ppf_model_get_info(ppf_detector d, ppf_model_info *info);
/* sampled / resampled model (n x 3 each) and the hash table in CSR form:
 * keys[nKeys] ascending, offs[nKeys+1], refs[nEntries], alphas[nEntries];
 * any pointer may be NULL. */
PPF_API int ppf_model_get_sampled(ppf_detector d, float *pts, float *nrm, float *rePts,
                                  float *reNrm);
PPF_API int ppf_model_get_table(ppf_detector d, uint32_t *keys, int *offs, uint32_t *refs,
                                float *alphas);

/* -- stage entry points (each replaces one reference function; used by the
 * parity tests and by callers composing their own pipeline) -- */

/* computePPF + computeAlpha (util.cpp:779-872) of (p1,n1) against n
 * partners (n x 3 arrays); tail != 0 applies the scalar-tail formula
 * (util.cpp:712-730) to all partners. */
PPF_API int ppf_pair_features(const float p1n1[6], const float *pts, const float *nrm, size_t n,
                              float angleStep, float distStep, int tail, uint32_t *keys,
                              float *alphas);

/* the vote loop of matchScene (ppf.cpp:247-409) over given key points of a
 * given sampled scene (n x 3 points/normals, positions as the vote stage
 * sees them).  Writes candidates (refId, angleId, vote) as 3 ints per slot,
 * 3 slots per key point (-1 = none) and, if accDump != NULL, the
 * accumulator bins of the first dumpCount key points (refNum x angleNum
 * ints each). */
PPF_API int ppf_vote(ppf_detector d, const float *pts, const float *nrm, size_t n,
                     const int *keypoints, size_t nk, float minScore, int *cand3,
                     int32_t *accDump, size_t dumpCount, float *poses12, int *nPoses);

/* voxel-grid downsampling (replaces samplePointCloud, util.cpp:107-139) */
PPF_API int ppf_voxel_sample(const float *pts, size_t n, const int *subset, size_t nSubset,
                             float step, const float boxMin[3], int *out, size_t *nOut);
/* ppf_voxel_sample without any allocation (no reference counterpart: for
 * callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  x, y, z:
 * device arrays of n points; subset (device, nullable) selects nSubset of
 * them; out (device, room for every selected point) receives the sampled
 * ids in the order of ppf_voxel_sample, *count (device) their number.  The
 * work is queued on stream (a hipStream_t; NULL = the default stream) and
 * not waited for.  At most 1048576 selected points (PPF_ERR_RANGE beyond):
 * the voxel table lives in a fixed 33 MiB device array that every GPU the
 * library runs on holds until the process ends, shared by all calls on
 * that GPU -- calls that overlap on one GPU race. */
PPF_API int ppf_voxel_sample_noalloc(const float *x, const float *y, const float *z, size_t n,
                                     const int *subset, size_t nSubset, float step,
                                     const float boxMin[3], int *out, int *count, void *stream);
/* exact radius search: per query the ids with dist2 < r2, ascending
 * (RadiusResultSet, nanoflann.hpp:364-418); counts[nq], ids packed in query
 * order up to cap */
PPF_API int ppf_radius_search(const float *pts, size_t n, const float *q, size_t nq, float r2,
                              float cell, int *counts, int *ids, size_t cap);
/* k nearest by (dist2, id) (KNNResultSet, nanoflann.hpp:169-242) */
PPF_API int ppf_knn_search(const float *pts, size_t n, const float *q, size_t nq, int k,
                           float cell, int *ids, float *d2);
/* estimateNormal (util.cpp:272-334), snapshot semantics; pts/nrm in-out */
PPF_API int ppf_estimate_normals(float *pts, float *nrm, size_t n, const int *valid,
                                 size_t nValid, const int *query, size_t nq, int k, int smooth,
                                 int invert, const float viewPoint[3]);

/* estimateNormalMLS (privateUtil.h:13-14, util.cpp:336-457): for each
 * query, PCA of its neighbours within radius among the valid points, then
 * a weighted polynomial fit (order 1..3) of their heights in the PCA
 * plane's Darboux frame; the query moves onto the fitted surface and its
 * normal follows the fit's slopes, oriented towards viewPoint (or +z when
 * it is NULL or not finite; away when invert).  pts/nrm (n x 3) in-out;
 * neighbours are read from the positions before any point moves.  stats
 * may be NULL. */
typedef struct {
    long long neighbours; /* sum of the queries' neighbourhood sizes */
    int       queries;
    float     deviceMs;   /* search + sort + fit on the device */
} ppf_mls_stats;
PPF_API int ppf_estimate_normals_mls(float *pts, float *nrm, size_t n, const int *valid,
                                     size_t nValid, const int *query, size_t nq, float radius,
                                     int order, int invert, const float viewPoint[3],
                                     ppf_mls_stats *stats);

/* -- ICP on explicit hypotheses (ICP::regist with a prebuilt search
 * structure and several initial poses, icp.h:86-90 / icp.cpp:214-278) -- */
typedef struct ppf_icp_target_s *ppf_icp_target;
/* ConvergenceCriteria, icp.h:7-25 */
typedef struct {
    int   iterations;
    float rejectDist, mseMin, mseMax, tolerance;
} ppf_icp_criteria;
/* ConvergenceResult, icp.h:34-44, plus the inlier count of util.cpp:628-641 */
typedef struct {
    float pose16[16]; /* column-major 4x4 */
    int   type;       /* 0 ITER, 1 MSE, 2 CONVERGE_RATE, 3 NO_CORRESPONDS */
    float mse, convergeRate;
    int   iterations, converged, inliers;
} ppf_icp_result;
/* the target cloud with normals, gridded once (cell: grid edge, 0 = auto) */
PPF_API int  ppf_icp_target_create(const float *pts, const float *nrm, size_t n, float cell,
                                   ppf_icp_target *out);
PPF_API void ppf_icp_target_destroy(ppf_icp_target t);
/* refine nPoses initial poses of src (with normals).  mode 0 = the
 * reference's symmetric objective, 1 = point-to-plane.  batched == 1: all
 * poses in one replayed hipGraph, each pose cut into workgroups of `block`
 * source points over the whole chip; batched == 2: all poses in one
 * persistent launch, one workgroup of `block` threads per pose; batched == 0:
 * the conventional refiner, one pose at a time with a host stop check per
 * iteration.  block: 0 = 256.
 * deviceMs (optional) receives the refinement's device time. */
PPF_API int ppf_icp_refine(ppf_icp_target t, const float *src, const float *srcNrm, size_t nSrc,
                           const float *initPoses16, size_t nPoses, const ppf_icp_criteria *c,
                           int mode, int batched, int block, float scoringDist,
                           ppf_icp_result *out, float *deviceMs);
/* hypotheses of several scenes refined together (no reference counterpart:
 * the batching extension, for callers with the pose hypotheses of many
 * scenes).  Pose h is refined against targets[poseTarget[h]] exactly as
 * ppf_icp_refine refines it against that target alone; batched == 1 or 2:
 * every pose in one batched refinement (as ppf_icp_refine), batched == 0:
 * the conventional refiner, one pose at a time.  The call runs on the first
 * target's stream and scratch, so no other call may use that target
 * meanwhile.  Errors: PPF_ERR_ARG for a null argument or target, or a
 * poseTarget entry outside [0, nTargets). */
PPF_API int ppf_icp_refine_multi(const ppf_icp_target *targets, size_t nTargets,
                                 const int *poseTarget, const float *src, const float *srcNrm,
                                 size_t nSrc, const float *initPoses16, size_t nPoses,
                                 const ppf_icp_criteria *c, int mode, int batched, int block,
                                 float scoringDist, ppf_icp_result *out, float *deviceMs);
/* point-to-plane cost adjoint (no reference counterpart; its Jacobian rows
 * are icp.cpp:318-320's): for each pose T = U(xi) T0 with U(xi) = [R(w), v]
 * (xi = (w, v), R(w) = exp([w]x) by Rodrigues' formula), the refiner's
 * correspondences at T and out[36] per pose = {cost, count, c[3], g[6],
 * J^T J[21] (upper triangle, row-major), 0 x4}.  g = J^T r is the gradient
 * with respect to a perturbation d = (om, u) applied on the left of T as
 * [R(om), c - R(om) c + u], about the centre c: the origin (center 0) or
 * the winners' centroid (center 1); at xi = 0 it is the gradient in xi
 * when c = 0, and d cost / d xi = (Jl(w)^T (g_om + g_u x (v - c)), g_u)
 * in general (Jl = SO(3)'s left Jacobian).  corr (optional, nPoses x nSrc)
 * receives each source point's target id (-1 = none). */
PPF_API int ppf_icp_adjoint(ppf_icp_target t, const float *src, size_t nSrc, const float *poses16,
                            const double *xi6, size_t nPoses, float rejectDist, int center,
                            double *out, int *corr);

/* L0 scalar math on the device, for bitwise host/device checks:
 * op 0 atanf(a), 1 atan2f(a,b), 2 a/b, 3 sqrtf(a), 4 murmur of 4 uint words
 * (a,b bit patterns as words 0,1 and 2,3 = 0), 5 quant ceil(a/b), 6 quant trunc,
 * 7 double sqrt((double)a*b) rounded to float bits of the double's high word */
PPF_API int ppf_device_math(int op, const float *a, const float *b, size_t n, uint32_t *out);

/* -- PLY files (filePLY.h:6-10).  ppf_ply_read replaces readPLY
 * (filePLY.cpp:199-257): vertex x/y/z, vertex nx/ny/nz when nx is present,
 * and the polygons of face "vertex_indices" (else "vertex_index") split
 * into triangles by ear clipping (filePLY.cpp:45-107); ascii and binary
 * little / big endian, every scalar type.  The counts come from
 * ppf_ply_info, the data (n x 3 floats, faces x 3 ints) from ppf_ply_copy.
 * Failures return PPF_ERR_RUNTIME with readPLY's message (the reference
 * prints it and returns false). -- */
PPF_API int  ppf_ply_read(const char *path, ppf_ply *out);
PPF_API int  ppf_ply_info(ppf_ply p, size_t *nPoints, size_t *nNormals, size_t *nFaces);
PPF_API int  ppf_ply_copy(ppf_ply p, float *points, float *normals, int *faces);
PPF_API void ppf_ply_free(ppf_ply p);
/* writePLY (filePLY.cpp:259-333): binary little endian (ascii != 0: ascii),
 * "comment Created by Open3D", float x/y/z (and nx/ny/nz when normals is
 * given), triangles as uchar-counted uint lists. */
PPF_API int ppf_ply_write(const char *path, const float *points, const float *normals, size_t n,
                          const int *faces, size_t nFaces, int ascii);
/* a PLY file read and uploaded as a resident scene (no reference
 * counterpart: readPLY followed by ppf_scene_upload); normals are used when
 * the file has one per point (type.cpp:36-38) */
PPF_API int ppf_scene_upload_ply(ppf_detector d, const char *path, const float *viewPoint,
                                 ppf_scene *out);

/* sampleMesh (util.h:8 / util.cpp:20-84): int(area / radius^2) points,
 * area-weighted over the triangles (faces: nFaces x 3 point indices), each
 * with its triangle's unit normal.  The uniforms of point j come from
 * Philox4x32-10 at counter (j, 0, 0, 0) under key seed (the reference
 * seeds std::mt19937 from std::random_device).  *nOut = the number of
 * points, which does not depend on seed; with points or normals NULL (a
 * size query) or cap < *nOut nothing is written.  nFaces == 0 gives 0
 * points (the reference's empty cloud). */
PPF_API int ppf_sample_mesh(const float *pts, size_t n, const int *faces, size_t nFaces,
                            float radius, uint64_t seed, float *points, float *normals,
                            size_t cap, size_t *nOut);

/* measured ceilings of this GPU for the rooflines (no reference
 * counterpart): a 1 GiB device-to-device streaming copy (bytes read +
 * written per second, best of 4) and LDS atomic adds at pseudo-random
 * addresses over 4096 workgroups of 1024 threads (best of 3) */
PPF_API int ppf_measure_ceilings(float *copyGBps, float *ldsAtomicsPerSec);

#ifdef __cplusplus
}
#endif
#endif
