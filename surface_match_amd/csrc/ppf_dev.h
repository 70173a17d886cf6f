// ppf_dev.h -- internal interface between the PPF host orchestration and
// the gfx950 kernels.
#pragma once

#include "dev_common.h"

#include <atomic>

namespace ppfd {

// A captured launch and the argument bytes it was captured with.
struct GraphCache {
    hipGraphExec_t exec = nullptr;
    int            H = 0, nSrc = 0, block = 0, spread = 0;
    unsigned char  args[1024];
    ~GraphCache() {
        if (exec)
            (void)hipGraphExecDestroy(exec);
    }
};

// Reusable device scratch of one detector (grown on demand, never shrunk).
struct Workspace {
    DevBuf voxKey, voxBest, voxFlag, voxIota, cubTmp, counter, voxMulti;
    DevBuf p0x, p0y, p0z, n0x, n0y, n0z, rawX, rawY, rawZ, cenX, cenY, cenZ, have, touch,
        touchList, nbq, cntq, outN, mark, qidx;
    Grid   gNormal, gVote, gSparse, gDense, gFull, gMls;
    // MLS normals: per-query neighbour counts, 64-bit offsets, ids (twice)
    DevBuf mlsCnt, mlsOff, mlsIds, mlsIds2, mlsP4, mlsKey, mlsKey2, mlsQ, mlsQ2;
    // cooperative kNN normals: each point's list (k ids) and count
    DevBuf knnIds, knnCnt;
    // vote
    DevBuf kpX, kpY, kpZ, kpNX, kpNY, kpNZ, cand, candTile, poses, voteStats, accDump, voteCnt,
        smallCnt, hitStart, hitLen, hitAlpha, hitBucket, hitCount, splits, kpKey, kpKey2, kpOrd,
        kpOrd2;
    // the bucket split points in `splits` were computed for this model
    // generation, tile size and tile count
    unsigned splitGen = ~0u;
    int      splitRefs = 0, splitTiles = 0;
    // clustering
    DevBuf clSortKey, clSortKey2, clSortVal, clSortVal2, clStatus, clPred, clPredOff,
        clCenters, clNms, clMembers;
    DevBuf clKey64, clKey64b, clSceneOff;  // clustering of several scenes at once
    // icp
    DevBuf icpPose, icpCur, icpState, icpHash, icpPartial, icpOut, icpInl, icpSt, icpArrive,
        icpPart;
    GraphCache icpGraph[4];  // captured batched refinements, replaced round-robin
    int        icpGraphNext = 0;
    // events bracketing the vote kernel (created on first use)
    hipEvent_t voteEv[3] = {nullptr, nullptr, nullptr};
    // set once the pair kernel and voteEv[2] after it are queued (split
    // path): lets a helper thread start work that should follow the pair
    // kernel on another stream
    std::atomic<int> pairQueued{0};
    ~Workspace() {
        for (auto &e : voteEv)
            if (e)
                (void)hipEventDestroy(e);
    }
};

// ---------------------------------------------------------------- model

// Device form of the trained model: the hash table as CSR buckets reached
// through an open-addressing key table, plus per-reference local frames.
struct DevModel {
    int      nSampled = 0, nReSampled = 0, angleNum = 30, nKeys = 0, nEntries = 0;
    float    diameter = 0, angleStep = 0, distStep = 0;
    uint32_t tmask = 0;
    DevBuf   sx, sy, sz, snx, sny, snz;      // sampled model (SoA)
    DevBuf   rx, ry, rz, rnx, rny, rnz;      // resampled model (SoA)
    DevBuf   frames;                         // nSampled x 12 (3x4 row-major)
    DevBuf   entries;                        // uint2 {ref, alpha bits}, bucket-major
    DevBuf   ventries;                       // uint2 {ref x angleNum, alpha bits}, same order,
                                             // each bucket padded to a multiple of 8 entries
    DevBuf   slots;                          // uint4 {key, start in ventries, len, bucket index}
    DevBuf   buckets;                        // uint2 {start in ventries, len} per bucket index
    DevBuf   cosSin;                         // 2 * angleNum floats
    int      refsSorted = 0;                 // every bucket's refs ascending
    // the most entries any bucket has for one reference: a key point with n
    // hits adds at most n * maxMult votes to any accumulator cell
    uint32_t maxMult = 1;
    unsigned generation = 0;                 // bumped by every upload
    // pair-stage pre-filters over quantized features (filterA == 0: none):
    // bitmaps of the (q4), (q3, q4) and (q1, q3, q4) projections of every
    // feature tuple whose key is in the table; angle indices 0 .. filterA - 2
    // and distance indices 0 .. filterD - 2 are values, the last index of
    // each stands for any other value and is always marked
    DevBuf   keyFilter;
    int      filterA = 0, filterD = 0;
};

// Records of the model pair stage before grouping (train on device).
struct TrainScratch {
    DevBuf keys, keys2, idx, idx2, refs, alphas, flags, cnt, tmp, ukeys, ucount, uoffs;
};

// Build the model table on device from the sampled model (SoA device
// arrays); returns host copies of keys/offsets/refs/alphas (bucket-major,
// buckets by ascending key, insertion order (i asc, j asc) inside).
void train_table(TrainScratch &ts, int n, const float *px, const float *py, const float *pz,
                 const float *nx, const float *ny, const float *nz, float angleStep,
                 float distStep, int refTail, std::vector<uint32_t> &keys, std::vector<int> &offs,
                 std::vector<uint32_t> &refs, std::vector<float> &alphas, hipStream_t s);

// Pair features for testing: key/alpha of (p1,n1) against partners.
void pair_features(const float *p1n1 /*6 floats host*/, int n, const float *px, const float *py,
                   const float *pz, const float *nx, const float *ny, const float *nz,
                   float angleStep, float distStep, int tail, uint32_t *keys, float *alphas,
                   hipStream_t s);

// Upload buckets + frames of a model already on the host.
void model_upload(DevModel &m, int nSampled, const float *sp, const float *sn, int nRe,
                  const float *rp, const float *rn, const std::vector<uint32_t> &keys,
                  const std::vector<int> &offs, const std::vector<uint32_t> &refs,
                  const std::vector<float> &alphas, int angleNum, float diameter,
                  float distStepRel, hipStream_t s);

// ---------------------------------------------------------------- ingest
// removeNan + computeBoundingBox (util.cpp:86-105,178-207) of a cloud whose
// AoS arrays are already on the device (ingest.hip).
struct IngestScratch {
    DevBuf    masks, blockCnt, blockOff, keys, devInfo;
    uint32_t *info = nullptr;  // pinned read-back of {count, box bits}
    IngestScratch() = default;
    IngestScratch(const IngestScratch &) = delete;
    IngestScratch &operator=(const IngestScratch &) = delete;
    ~IngestScratch();
};
// SoA originals (normals only when aosN != NULL), the ascending indices of
// the valid points into valid[], and the box of the valid points (zeros when
// none); returns the number of valid points.  Synchronises s.
int ingest_cloud(IngestScratch &is, const float *aosP, const float *aosN, int n, float *ox,
                 float *oy, float *oz, float *onx, float *ony, float *onz, int *valid,
                 float ext[6], hipStream_t s);

// ---------------------------------------------------------------- grid/sample/normals
void device_bbox(const float *X, const float *Y, const float *Z, const int *subset, int m,
                 float *bb6dev, hipStream_t s);
int  voxel_sample(Workspace &ws, const float *X, const float *Y, const float *Z,
                  const int *subset, int m, float step, const float bmin[3], int *out,
                  int *countDev, hipStream_t s);
// voxel_sample without allocation, memset or host synchronisation: scratch
// (device, at least voxel_scratch_bytes(m)) holds the voxel table, flags and
// block counts; the count goes to countDev.  Same output as voxel_sample.
size_t voxel_scratch_bytes(int m);
void   voxel_sample_into(const float *X, const float *Y, const float *Z, const int *subset, int m,
                         float step, const float bmin[3], int *out, int *countDev,
                         unsigned char *scratch, hipStream_t s);
// ---- several scenes in one launch sequence (the batched match): each
// scene's part computes exactly what the single-scene call computes on it

// One scene's input to voxel_sample_multi: its points, the subset sampled
// (nullptr: 0..m-1), the voxel origin and its output list.
struct VoxSceneIn {
    const float *X, *Y, *Z;
    const int   *subset;
    int          m;
    float        bmin[3];
    int         *out;
};
// voxel_sample of every scene at one step; counts[i] is scene i's output
// size.  One read-back for all scenes.
void voxel_sample_multi(Workspace &ws, const std::vector<VoxSceneIn> &sc, float step,
                        std::vector<int> &counts, hipStream_t s);

// One scene's input to grid_build_multi (grid_build's arguments).
struct GridSceneIn {
    const float *X, *Y, *Z, *NX, *NY, *NZ;  // NX nullptr: no normals
    const int   *subset;
    int          m;
    float        box[6];  // a box known to hold the points (grid_build's hostBox6)
    int          idOff;   // added to every id the grid stores (0: the scene's own indices)
};
// The grids of several scenes built by one counting sort: every scene's
// view indexes shared slot arrays; cells per scene as grid_build chooses
// them.  Returns false (nothing built) when a scene's grid would exceed the
// dense cell limit or all scenes' cells together would; the caller then
// builds them one by one.
struct MultiGrid {
    DevBuf keys, vals, rank, cnt, cstart, occ, x, y, z, id, nx, ny, nz, p4, n4, tmp, scanTmp, htab;
    std::vector<GridView> views;
    int                   total = 0;  // points of all scenes (the shared slot arrays' length)
};
// soa false: only the {x, y, z, id} float4 copies (aos must be true), for
// grids read by radius tests alone (the scoring grid); the views' x, y, z
// and id are then nullptr.
bool grid_build_multi(MultiGrid &mg, const std::vector<GridSceneIn> &sc, float cellHint,
                      bool aos, hipStream_t s, bool soa = true);

void estimate_normals(Workspace &ws, int n, float *PX, float *PY, float *PZ, float *NX,
                      float *NY, float *NZ, const int *valid, int nValid, const int *query,
                      int nq, int k, int smooth, int invert, const float vp[3], float cellHint,
                      hipStream_t s, const float *hostBox6 = nullptr);

// One scene's input to estimate_normals_multi (estimate_normals' arguments).
struct NormSceneIn {
    float     *PX, *PY, *PZ, *NX, *NY, *NZ;
    int        n;
    const int *valid;
    int        nValid;
    const int *query;
    int        nq;
    float      vp[3];   // viewpoint (a non-finite one: none)
    float      box[6];  // a box known to hold the points
};
// estimate_normals of several scenes with one kNN grid build (mg) and one
// launch sequence: the scenes' points side by side in one id space, each
// query searching its own scene's cells.  Falls back to estimate_normals
// scene by scene when the scenes' grids do not fit one build.
void estimate_normals_multi(Workspace &ws, MultiGrid &mg, const std::vector<NormSceneIn> &sc, int k,
                            int smooth, int invert, float cellHint, hipStream_t s);

// estimateNormalMLS (util.cpp:336-457) for the queries (indices into the
// n-point cloud PX/PY/PZ, which is not modified): moved points into OX/OY/OZ
// (which must hold copies of the inputs), normals updated in NX/NY/NZ.
// Neighbours come from the valid points.  Asynchronous on s after its one
// read-back of the neighbour total.
constexpr int MLS_MAX_ORDER = 3;
struct MlsStats {
    long long neighbours = 0;
    int       queries    = 0;
};
void estimate_normals_mls(Workspace &ws, int n, const float *PX, const float *PY, const float *PZ,
                          float *OX, float *OY, float *OZ, float *NX, float *NY, float *NZ,
                          const int *valid, int nValid, const int *query, int nq, float radius,
                          int order, int invert, const float vp[3], MlsStats *st, hipStream_t s);

// ---------------------------------------------------------------- vote

// One pose hypothesis from the vote stage (A11/A12).
struct PoseRec {
    float pose[12];
    float votes;
    float angle;
    float q[4];
    int   keypoint;
    int   rank;
    int   refId;
    int   angleId;
};

struct VoteStats {
    unsigned long long pairs, hits, votes, candidates;
    int   voted;      // key points that passed the neighbour-count threshold
    float kernelMs;   // device time of the vote stage kernels
    float pairMs;     // of which the pair kernel (split path)
    float hitsMs;     // of which hit voting and any fused fallback (split path)
    int   overflow;   // key points voted by the fused kernel for lack of hit capacity
};

struct VoteConfig {
    int block    = 256;   // threads per voting workgroup
    int xcdRemap = 0;     // keep spatially adjacent keypoints on one XCD
    int cellDiv  = 3;     // vote grid cell = model diameter / cellDiv (2..4)
    int split    = 1;     // 1: pair kernel + hit voting kernel; 0: fused kernel
    int hitCap   = 2048;  // hits kept per key point on the split path
    int tiles    = 1;     // at least this many reference tiles per key point
    int order    = 1;     // visit key points in (scene, Morton code of position) order
};

// Vote every keypoint (indices into the sampled-scene grid's clouds),
// extract peaks and build poses.  Returns the number of poses; the poses
// are in (keypoint, rank) order in ws.poses.
// Key points of one or several scenes for the vote: flat per-key-point
// position / normal / point id, and (several scenes) each key point's scene
// index into a device array of the scenes' vote grids.
struct VoteKeypoints {
    GridView        g;         // the grid when there is one scene
    const GridView *gridsDev;  // device array, one per scene (nullptr: g)
    const int      *kpScene;   // device, per key point (with gridsDev)
    float           minCell;   // smallest grid cell (host check of the cell bound)
    const float    *kx, *ky, *kz, *knx, *kny, *knz;
    const int      *kp;        // point id within its scene
    int             nk;
    int             nScenes = 1;
};
// position / normal of points kp[0..nk) of a cloud into flat arrays
void gather_keypoints(const int *kp, int nk, const float *PX, const float *PY, const float *PZ,
                      const float *NX, const float *NY, const float *NZ, float *kx, float *ky,
                      float *kz, float *knx, float *kny, float *knz, hipStream_t s);
// gather_keypoints of several scenes into one flat list: scene w's key
// points at [ko[w], ko[w + 1]), with their scene (kpScene) and point id
// within it (kpId).  desc: device scratch for the scenes' descriptors.
struct KpSceneIn {
    const int   *kp;
    const float *PX, *PY, *PZ, *NX, *NY, *NZ;
};
void gather_keypoints_multi(const std::vector<KpSceneIn> &sc, const std::vector<int> &ko,
                            DevBuf &desc, float *kx, float *ky, float *kz, float *knx, float *kny,
                            float *knz, int *kpScene, int *kpId, hipStream_t s);
int vote_keypoints(Workspace &ws, const DevModel &m, const VoteKeypoints &in, float voteThreshold,
                   const VoteConfig &cfg, VoteStats *stats, int32_t *accDump, int dumpCount,
                   hipStream_t s);

int vote_and_poses(Workspace &ws, const DevModel &m, const GridView &g, const float *PX,
                   const float *PY, const float *PZ, const float *NX, const float *NY,
                   const float *NZ, const int *keypoints, int nk, float voteThreshold,
                   const VoteConfig &cfg, VoteStats *stats, int32_t *accDump, int dumpCount,
                   hipStream_t s);
int vote_max_tile_refs(int angleNum, int block);

// ---------------------------------------------------------------- clustering
// Greedy pose clustering + Markley mean + overlap NMS (util.cpp:486-601).
// Returns the NMS'd hypotheses (votes desc) in host memory.
struct Hypo {
    float pose[12];
    float votes;
};
void cluster_poses(Workspace &ws, const PoseRec *posesDev, int np, float distThr, float angleThr,
                   const float center[3], float nmsThr, std::vector<Hypo> &out, int *nClusters,
                   hipStream_t s);
// cluster_poses for several scenes at once: scene i's poses are
// posesDev[pOff[i] .. pOff[i + 1]); out[i] and nClusters[i] equal what
// cluster_poses returns for them alone.  Only for scenes of at most
// CLUSTER_MULTI_MAX poses each (cluster_multi_fits).
bool cluster_multi_fits(const std::vector<int> &pOff);
void cluster_poses_multi(Workspace &ws, const PoseRec *posesDev, const std::vector<int> &pOff,
                         float distThr, float angleThr, const float center[3], float nmsThr,
                         std::vector<std::vector<Hypo>> &out, std::vector<int> &nClusters,
                         hipStream_t s);

// ---------------------------------------------------------------- ICP
struct IcpCriteria {
    int   iterations;
    float rejectDist, mseMin, mseMax, tolerance;
};
struct IcpResultDev {
    float pose[12];
    int   type;
    float mse, convergeRate;
    int   iterations, converged;
    int   inliers;
};
struct IcpConfig {
    int block  = 512;
    int mode   = 0;  // 0 symmetric (reference), 1 point-to-plane
    int graph  = 1;  // replay the launch sequence from a captured hipGraph
    int spread = 1;  // 1: hypotheses cut into chunks of `block` points over
                     // the whole chip; 0: one persistent workgroup each
};
// Refine H hypotheses (poses12 host) of src against dst in one persistent
// launch; scores by inliers against `full` within scoringDist.
void icp_batch(Workspace &ws, int nSrc, const float *sx, const float *sy, const float *sz,
               const float *snx, const float *sny, const float *snz, const GridView &dst,
               const GridView &full, const float *init12, int H, const IcpCriteria &c,
               float scoringDist, const IcpConfig &cfg, IcpResultDev *out, hipStream_t s,
               const GridView *dstG = nullptr, const GridView *fullG = nullptr,
               const int *hypScene = nullptr);
// (with dstG: hypothesis h refines against the device grids dstG[hypScene[h]]
// and fullG[hypScene[h]]; dst / full then only name the target kind)
// The conventional per-hypothesis GPU refiner (grid over source points per
// step, host stop check per iteration), kept for comparison with icp_batch.
void icp_plain(Workspace &ws, int nSrc, const float *sx, const float *sy, const float *sz,
               const float *snx, const float *sny, const float *snz, const GridView &dst,
               const GridView &full, const float *init12, int H, const IcpCriteria &c,
               float scoringDist, int mode, IcpResultDev *out, hipStream_t s);

// Point-to-plane cost, se(3) gradient and Gauss-Newton J^T J of H poses
// (T = U(xi) T0) against dst, with the refiner's correspondence rule.  out:
// H x 36 doubles {cost, count, centre[3], grad[6], JtJ[21], 0 x4}; corr
// (optional): H x nSrc dst ids of the winners (-1 = none).
// inliner (util.cpp:628-641) at given poses: per pose (H x 12 host, row-major
// 3x4) the count of src points whose transformed position has a point of
// `full` within dist (dist2 < dist^2) -- the refiner's scoring step alone.
void icp_inliers(Workspace &ws, const GridView &full, int nSrc, const float *sx, const float *sy,
                 const float *sz, const float *poses12Host, int H, float dist, int *inliersHost,
                 hipStream_t s);
void icp_adjoint(Workspace &ws, int nSrc, const float *sx, const float *sy, const float *sz,
                 const GridView &dst, const float *T0host, const double *xiHost, int H,
                 float rejectDist, int center, double *outHost, int *corrHost, hipStream_t s);

// ---------------------------------------------------------------- mesh

// sampleMesh (util.cpp:20-84) with Philox4x32-10 uniforms keyed by seed:
// points and normals (n x 3 each) of the sample
void sample_mesh(const float *pts, int nPts, const int *faces, int nFaces, float radius,
                 uint64_t seed, std::vector<float> &outP, std::vector<float> &outN,
                 hipStream_t s);
// the number of points sample_mesh returns (the same for every seed)
int sample_mesh_count(const float *pts, int nPts, const int *faces, int nFaces, float radius,
                      hipStream_t s);

// ---------------------------------------------------------------- PLY

// a PLY file's points and normals (n x 3 floats) and triangles (3 ints each)
struct PlyCloud {
    std::vector<float> points, normals;
    std::vector<int>   faces;
};
// readPLY / writePLY (filePLY.cpp:199-333); throw std::runtime_error with
// the reference's messages
void ply_read(const char *path, PlyCloud &out);
void ply_write(const char *path, const float *pts, const float *nrm, size_t n, const int *faces,
               size_t nFaces, bool ascii);

}  // namespace ppfd
