/*
 * ppf_oracle.h -- CPU oracle for the PPF detection path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a single-threaded (optionally OpenMP for
 * timing) C restatement of SurfaceMan/surface_match's PPF path
 * (ppf.cpp, util.cpp, icp.cpp).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links
 * or calls it.
 *
 * Semantics ("canonical mode"):
 *   - all FP32 arithmetic is IEEE single precision with no FMA contraction
 *     (built with -ffp-contract=off), in the operation order spelled out in
 *     each function's comment;
 *   - transcendental functions are this file's own (Cephes-method atanf),
 *     never libm, so that the device library can reproduce every bit;
 *   - order-dependent reference behaviour (kd-tree traversal order, OpenMP
 *     races) is replaced by the order-free definitions documented per
 *     function.
 *
 * Parity status per function is listed in the header comment of
 * ppf_oracle.c.
 */
#ifndef PPF_ORACLE_H
#define PPF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- L0 scalar math -------------------------------------------------- */
float    orc_atanf(float x);
float    orc_atan2f(float y, float x);
uint32_t orc_murmur3_x86_32(const uint8_t *data, uint32_t len, uint32_t seed);
uint32_t orc_hash_key4(const uint32_t q[4]);

/* ---- L1 pair features ------------------------------------------------- */
/* f[0..3] = f1,f2,f3,|d|; returns the hash key.  tail != 0 selects the
 * reference's scalar-tail formula (util.cpp:712-730 + :690-695). */
uint32_t orc_ppf_key(const float p1[3], const float n1[3], const float p2[3], const float n2[3],
                     float angleStep, float distStep, int tail, float f[4], uint32_t q[4]);
/* 3x4 row-major rigid frame taking p to the origin and n to +x. */
void  orc_frame(const float p[3], const float n[3], float rt[12]);
float orc_alpha(const float rt[12], const float p2[3]);
/* angle bin of a model alpha against a scene alpha (ppf.cpp:316-340). */
int orc_angle_bin(float alphaModel, float alphaScene, int angleNum, int tail);

/* ---- L3 model table ---------------------------------------------------- */
typedef struct {
    int       nKeys;
    int       nEntries;
    uint32_t *keys;   /* nKeys, ascending                                   */
    int      *offs;   /* nKeys+1                                            */
    uint32_t *refs;   /* nEntries, bucket-major, insertion order in bucket  */
    float    *alphas; /* nEntries                                           */
} orc_table;

/* trainModel's table stage (ppf.cpp:105-146), single-thread insertion
 * order (i ascending, j ascending).  refTail: apply the scalar-tail key
 * formula to partners j >= n - n%8 (util.cpp:783-784,822-824). */
int  orc_build_table(int n, const float *pts, const float *nrm, float angleStep, float distStep,
                     int refTail, orc_table *out);
void orc_free_table(orc_table *t);
int  orc_table_find(const orc_table *t, uint32_t key);

/* ---- L4 vote / peaks / poses ----------------------------------------- */
typedef struct {
    int   refId;
    int   angleId;
    float vote;
} orc_candidate;

/* Vote one keypoint against its neighbour list (self already excluded).
 * acc has refNum*(angleNum+1) ints, laid out as the reference's
 * accumulator (slot 0 = count, slots 1..angleNum = bins), ppf.cpp:301-346. */
void orc_vote_keypoint(const orc_table *t, int refNum, int angleNum, const float kp[3],
                       const float kn[3], int nNb, const float *nbPts, const float *nbNrm,
                       float angleStep, float distStep, int32_t *acc);
/* Peak extraction + 0.95 filter (ppf.cpp:363-397).  Returns count (<=3). */
int orc_peaks(const int32_t *acc, int refNum, int angleNum, float voteThreshold,
              orc_candidate out[3]);
/* Pose = inv(Ts) * Rx(alpha_j) * Tm (ppf.cpp:393-405), 3x4 row-major. */
void orc_pose(const float sceneRt[12], const float modelRt[12], int angleId, int angleNum,
              float out[12]);

/* ---- L2 spatial search (uniform grid, exact) ------------------------- */
typedef struct {
    int          n;
    const float *pts; /* n x 3 */
    float        org[3];
    float        cell;
    int          dim[3];
    int         *start; /* ncell+1 */
    int         *idx;   /* n, point ids sorted by cell then id */
} orc_grid;

int  orc_grid_build(orc_grid *g, int n, const float *pts, const int *subset, int nSubset,
                    float cell);
void orc_grid_free(orc_grid *g);
/* removeNan (util.cpp:86-105) and computeBoundingBox (util.cpp:178-207) */
int  orc_remove_nan(int n, const float *pts, const float *nrm, int *out);
void orc_bounding_box(int n, const float *pts, const int *idx, int m, float box6[6]);

/* nanoflann L2 metric for DIM=3: ((0+dx^2)+dy^2)+dz^2, dx = q - p. */
float orc_dist2(const float q[3], const float p[3]);
/* all ids with dist2 < r2 (nanoflann RadiusResultSet::addPoint), ascending id */
int orc_grid_radius(const orc_grid *g, const float q[3], float r2, int *out, int cap);
/* k smallest by (dist2, id) */
int orc_grid_knn(const orc_grid *g, const float q[3], int k, int *out, float *d2);

/* ---- sampling ------------------------------------------------------- */
/* voxel-grid downsampling: one point per voxel of edge `step`, the one
 * nearest the voxel centre (ties: lowest id); output ascending id. */
int orc_voxel_sample(int n, const float *pts, const int *subset, int nSubset, float step,
                     const float boxMin[3], int *out);
/* greedy Poisson-disk sampling in a given visiting order (util.cpp:107-139) */
int orc_greedy_sample(int n, const float *pts, const int *order, int nOrder, float step, int *out);

/* ---- mesh sampling (util.cpp:20-84) --------------------------------- */
/* Philox4x32-10 (Salmon et al., SC'11) in place on ctr under key */
void orc_philox4x32_10(uint32_t ctr[4], const uint32_t key[2]);
/* sampleMesh with point j's uniforms from Philox at counter (j,0,0,0),
 * key (seed low word, high word).  Returns the number of points and writes
 * them (with their triangles' normals) when cap allows; -1 for a face
 * index out of range, -2 for an area or radius out of range. */
int orc_sample_mesh(int nPts, const float *pts, int nFaces, const int *faces, float radius,
                    uint64_t seed, float *outP, float *outN, int cap);

/* ---- normals (util.cpp:234-334), snapshot semantics ------------------ */
int orc_estimate_normals(int n, float *pts, float *nrm, const int *valid, int nValid,
                         const int *query, int nQuery, int k, int smooth, int invert,
                         const float viewPoint[3]);

/* ---- MLS normals (util.cpp:336-457), snapshot semantics --------------- */
/* e^x, Cephes expf method (Cody-Waite reduction, degree-5 polynomial) */
float orc_expf(float x);
/* Eigen's unitOrthogonal() of a 3-vector */
void orc_unit_orthogonal(const float n[3], float o[3]);
/* estimateNormalMLS over `query` (indices into pts), neighbours within
 * radius among `valid`, polynomial order 1..3; pts/nrm in-out.  Returns
 * the total neighbourhood size, or -1 for an order or radius out of range. */
long orc_estimate_normals_mls(int n, float *pts, float *nrm, const int *valid, int nValid,
                              const int *query, int nQuery, float radius, int order, int invert,
                              const float vp[3]);

/* ---- small linear algebra ------------------------------------------- */
int  orc_jacobi_eig(int n, double *a, double *evec, double *eval);
void orc_quat_from_rot(const float R[9], float q[4]); /* q = (w,x,y,z) */
void orc_rot_from_quat(const float q[4], float R[9]);
float orc_rot_angle(const float R[9]);

/* ---- pose clustering (util.cpp:486-601) ------------------------------ */
typedef struct {
    float pose[12]; /* 3x4 row-major */
    float votes;
    float angle;    /* rotation angle of pose */
    float q[4];
} orc_pose_t;
int orc_cluster_poses(const orc_pose_t *in, int n, float distThr, float angleThr,
                      int *clusterOf, orc_pose_t *centersOut);
int orc_nms_poses(const orc_pose_t *in, int n, const float center[3], float thr,
                  orc_pose_t *out);

/* ---- ICP (icp.cpp) --------------------------------------------------- */
typedef struct {
    int   iterations;
    float rejectDist, mseMin, mseMax, tolerance;
} orc_icp_criteria;
typedef struct {
    float pose[12];
    int   type; /* 0 ITER 1 MSE 2 CONVERGE_RATE 3 NO_CORRESPONDS */
    float mse, convergeRate;
    int   iterations, converged;
} orc_icp_result;
int orc_icp_regist(int nSrc, const float *srcP, const float *srcN, const orc_grid *dst,
                   const float *dstN, const float initPose[12], const orc_icp_criteria *c,
                   int mode, orc_icp_result *res);
int orc_inliers(int nSrc, const float *srcP, const float pose[12], const orc_grid *g,
                float dist);
void orc_p2pl_system(int nSrc, const float *srcP, const double T[12], const int *corr,
                     const float *dstP, const float *dstN, const double c[3], const double xi[6],
                     int tukey, double out[28]);

/* ---- end-to-end (ppf.cpp:36-147,149-510) ----------------------------- */
typedef struct {
    float featDistanceStepRel;
    int   featAngleResolution;
    float poseRefRelSamplingDistance;
    int   knnNormal;
    int   smoothNormal;
} orc_train_param;
typedef struct {
    int   numMatches, knnNormal, smoothNormal, invertNormal;
    float maxOverlapDistRel, maxOverlapDistAbs;
    int   sparsePoseRefinement, densePoseRefinement, poseRefNumSteps;
    float poseRefDistThresholdRel, poseRefDistThresholdAbs;
    float poseRefScoringDistRel, poseRefScoringDistAbs;
} orc_match_param;

typedef struct {
    float           samplingDistanceRel;
    orc_train_param param;
    int             nSampled, nReSampled;
    float          *sampledP, *sampledN, *reSampledP, *reSampledN;
    float           boxMin[3], boxMax[3], viewPoint[3];
    orc_table       table;
} orc_model;

int  orc_train(int n, const float *pts, const float *nrm, const float *box6, const float vp[3],
               float rel, const orc_train_param *p, int refTail, orc_model *out);
void orc_free_model(orc_model *m);

typedef struct {
    int   nSampled, nKeypoints, nPoses, nClusters, nNms, nVotedKeypoints;
    long  nPairs, nKeyHits, nVotes;
    float timeMs[8];
} orc_match_stats;

int orc_match(const orc_model *m, int n, const float *pts, const float *nrm, const float *box6,
              const float vp[3], float rel, float kpFrac, float minScore,
              const orc_match_param *p, int cap, float *poses12, float *scores,
              orc_match_stats *st, int nThreads);

#ifdef __cplusplus
}
#endif
#endif
