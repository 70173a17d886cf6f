// cluster.hip -- pose clustering on gfx950 (util.cpp:486-601).
//
// clusterPose walks the poses by votes (descending) and puts each one in
// the first existing cluster whose centre (its first pose) is within 0.1 D
// in translation and angleStep in rotation-angle magnitude, else opens a new
// cluster.  The outcome is fixed by one rule -- pose i joins the earliest
// centre among its matching predecessors, and is a centre iff none of them
// is -- which is resolved here in parallel rounds: a pose decides as soon as
// every matching predecessor before its first matching centre has decided.
// Each round is a pass of the workgroup over the undecided poses; the
// result equals the sequential loop's.  clusterPose2's overlap NMS has the
// same structure and is resolved the same way.
#include "ppf_dev.h"

#include <cstring>
#include <rocprim/rocprim.hpp>

#include <algorithm>

namespace ppfd {

enum : int { UNDECIDED = -2, CENTER = -1 };

struct SortedPoses {
    float *tx, *ty, *tz, *angle, *votes;
    int   *ord;  // sorted position -> input index
};

__global__ void vote_keys_kernel(const PoseRec *p, int n, uint32_t *keys, int *vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    keys[i] = 0xFFFFFFFFu - __float_as_uint(p[i].votes);  // votes > 0: descending
    vals[i] = i;
}

__global__ void gather_sorted_kernel(const PoseRec *p, const int *ord, int n, SortedPoses s) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const PoseRec &r = p[ord[i]];
    s.tx[i]    = r.pose[3];
    s.ty[i]    = r.pose[7];
    s.tz[i]    = r.pose[11];
    s.angle[i] = r.angle;
    s.votes[i] = r.votes;
}

// comparePose (util.cpp:493-497)
__device__ __forceinline__ bool pose_match(const SortedPoses &s, int i, int j, float dThr,
                                           float aThr) {
    float dx = s.tx[i] - s.tx[j], dy = s.ty[i] - s.ty[j], dz = s.tz[i] - s.tz[j];
    float d  = sqrtf(dx * dx + dy * dy + dz * dz);
    float ph = fabsf(s.angle[i] - s.angle[j]);
    return d < dThr && ph < aThr;
}

// Greedy clustering rounds.  status[i]: UNDECIDED, CENTER, or the sorted
// position of the centre i joined.  Candidate predecessors come from the
// translation grid (cell >= dThr), so only 27 cells are examined.
__global__ void __launch_bounds__(1024) cluster_rounds_kernel(SortedPoses s, int n, GridView g,
                                                              float dThr, float aThr, int *status,
                                                              int *roundsOut) {
    __shared__ int sChanged, sUndecided;
    int rounds = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        status[i] = UNDECIDED;
    __syncthreads();
    for (;;) {
        if (threadIdx.x == 0) {
            sChanged   = 0;
            sUndecided = 0;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            if (status[i] != UNDECIDED)
                continue;
            float qx = s.tx[i], qy = s.ty[i], qz = s.tz[i];
            int   minCenter = 0x7fffffff, minUnd = 0x7fffffff;
            int   xlo = cell_coord(qx - dThr, g.ox, g.cell, g.dx), xhi = cell_coord(qx + dThr, g.ox, g.cell, g.dx);
            int   ylo = cell_coord(qy - dThr, g.oy, g.cell, g.dy), yhi = cell_coord(qy + dThr, g.oy, g.cell, g.dy);
            int   zlo = cell_coord(qz - dThr, g.oz, g.cell, g.dz), zhi = cell_coord(qz + dThr, g.oz, g.cell, g.dz);
            for (int cz = zlo; cz <= zhi; cz++)
                for (int cy = ylo; cy <= yhi; cy++)
                    for (int cx = xlo; cx <= xhi; cx++) {
                        int st, ct;
                        if (!cell_lookup(g, cx, cy, cz, st, ct))
                            continue;
                        for (int t = st; t < st + ct; t++) {
                            int j = g.id[t];
                            if (j >= i || !pose_match(s, i, j, dThr, aThr))
                                continue;
                            int sj = status[j];
                            if (sj == CENTER)
                                minCenter = min(minCenter, j);
                            else if (sj == UNDECIDED)
                                minUnd = min(minUnd, j);
                        }
                    }
            if (minUnd < minCenter) {
                atomicAdd(&sUndecided, 1);
                continue;
            }
            status[i] = (minCenter == 0x7fffffff) ? CENTER : minCenter;
            sChanged  = 1;
        }
        __syncthreads();
        rounds++;
        int und = sUndecided, ch = sChanged;
        __syncthreads();
        if (und == 0 || !ch)
            break;
        // make this round's decisions visible to the next round
        __threadfence_block();
    }
    if (threadIdx.x == 0)
        *roundsOut = rounds;
}

// The same rule walked in order by one wave, for pose lists that fit in
// LDS: 64 poses at a time, each lane first looks for its pose's first
// matching centre among those opened before the block (centres keep their
// creation order); the lanes left over are then settled in order, each
// against the centres opened in the block so far (one ballot: the lowest
// matching lane is the earliest), opening a centre when none matches.  No
// rounds, so short lists finish in one pass instead of a chain of
// workgroup-wide rounds.
constexpr int CLUSTER_SEQ_MAX = 512;
// several scenes at once, one wave each: the walks run side by side, so
// longer lists pay (80 KB of LDS at the limit)
constexpr int CLUSTER_MULTI_MAX = 4096;

// poses [b, b + n) of the sorted list, status written at the same
// positions; centres kept in sC (5 words each, capacity n)
__device__ void cluster_seq_range(const SortedPoses &s, int b, int n, float dThr, float aThr,
                                  int *status, float *sC) {
    // centres in creation order: {x, y, z, angle}, then sorted positions
    float4   *cc = reinterpret_cast<float4 *>(sC);
    int      *ci = reinterpret_cast<int *>(sC + 4 * n);
    const int lane = threadIdx.x;
    int       nc   = 0;  // wave-uniform
    for (int b0 = 0; b0 < n; b0 += 64) {
        const int  i     = b + b0 + lane;
        const bool valid = b0 + lane < n;
        float      tx = 0.0f, ty = 0.0f, tz = 0.0f, ta = 0.0f;
        if (valid) {
            tx = s.tx[i];
            ty = s.ty[i];
            tz = s.tz[i];
            ta = s.angle[i];
        }
        // comparePose (util.cpp:493-497) with the later pose first, as
        // pose_match(s, i, centre)
        auto match = [&](float x, float y, float z, float a) {
            const float dx = tx - x, dy = ty - y, dz = tz - z;
            const float d  = sqrtf(dx * dx + dy * dy + dz * dz);
            return d < dThr && fabsf(ta - a) < aThr;
        };
        int       st  = UNDECIDED;
        const int nc0 = nc;
        if (valid) {
            // four centres per step, their loads issued together, tested in order
            int t = 0;
            for (; t + 3 < nc0; t += 4) {
                const float4 c0 = cc[t], c1 = cc[t + 1], c2 = cc[t + 2], c3 = cc[t + 3];
                const int    f  = match(c0.x, c0.y, c0.z, c0.w)   ? 0
                                  : match(c1.x, c1.y, c1.z, c1.w) ? 1
                                  : match(c2.x, c2.y, c2.z, c2.w) ? 2
                                  : match(c3.x, c3.y, c3.z, c3.w) ? 3
                                                                  : 4;
                if (f < 4) {
                    st = ci[t + f];
                    break;
                }
            }
            if (st == UNDECIDED)
                for (; t < nc0; t++) {
                    const float4 c0 = cc[t];
                    if (match(c0.x, c0.y, c0.z, c0.w)) {
                        st = ci[t];
                        break;
                    }
                }
        }
        bool               opened = false;
        unsigned long long rest   = __ballot(valid && st == UNDECIDED);
        while (rest) {
            const int u = __ffsll((long long)rest) - 1;
            rest &= rest - 1ull;
            const float ux = __shfl(tx, u, 64), uy = __shfl(ty, u, 64), uz = __shfl(tz, u, 64),
                        ua = __shfl(ta, u, 64);
            bool m = false;
            if (opened) {
                // the centre is this lane's pose, the later pose is u's
                const float dx = ux - tx, dy = uy - ty, dz = uz - tz;
                const float d  = sqrtf(dx * dx + dy * dy + dz * dz);
                m              = d < dThr && fabsf(ua - ta) < aThr;
            }
            const unsigned long long mm = __ballot(m);
            if (mm) {
                if (lane == u)
                    st = b + b0 + __ffsll((long long)mm) - 1;
            } else {
                if (lane == u) {
                    st     = CENTER;
                    opened = true;
                    cc[nc] = make_float4(tx, ty, tz, ta);
                    ci[nc] = i;
                }
                nc++;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (valid)
            status[i] = st;
    }
}

__global__ void __launch_bounds__(64) cluster_seq_kernel(SortedPoses s, int n, float dThr,
                                                         float aThr, int *status) {
    extern __shared__ float sC[];
    cluster_seq_range(s, 0, n, dThr, aThr, status, sC);
}

// one wave per scene: scene w's poses are sorted positions [off[w], off[w + 1])
__global__ void __launch_bounds__(64) cluster_seq_multi_kernel(SortedPoses s, const int *off,
                                                               float dThr, float aThr,
                                                               int *status) {
    extern __shared__ float sC[];
    const int b = off[blockIdx.x];
    cluster_seq_range(s, b, off[blockIdx.x + 1] - b, dThr, aThr, status, sC);
}

// cluster index of each centre = its rank among centres (creation order)
__global__ void center_rank_kernel(const int *status, int n, int *isCenter) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        isCenter[i] = status[i] == CENTER ? 1 : 0;
}
__global__ void member_cluster_kernel(const int *status, const int *rank, int n, uint32_t *cl,
                                      int *vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int c = status[i] == CENTER ? i : status[i];
    cl[i]   = (uint32_t)rank[c];
    vals[i] = i;
}

// avgClusters + avgQuaternionMarkley (util.cpp:553-601): one thread per
// cluster sums its members in sorted order (FP32 as the reference), then
// an FP64 Jacobi eigen-solve of sum(q q^T)/M.
__global__ void cluster_avg_kernel(const PoseRec *p, const int *ord, const int *members,
                                   const int *mOff, int nc, Hypo *out, float *angleOut) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc)
        return;
    float tx = 0, ty = 0, tz = 0, votes = 0;
    float A[16];
    for (int r = 0; r < 16; r++)
        A[r] = 0.0f;
    int b = mOff[c], e = mOff[c + 1];
    for (int t = b; t < e; t++) {
        const PoseRec &r = p[ord[members[t]]];
        tx    = tx + r.pose[3];
        ty    = ty + r.pose[7];
        tz    = tz + r.pose[11];
        votes = votes + r.votes;
        for (int x = 0; x < 4; x++)
            for (int y = 0; y < 4; y++)
                A[x * 4 + y] = A[x * 4 + y] + r.q[x] * r.q[y];
    }
    int    m = e - b;
    double Ad[16], V[16], W[4];
    for (int r = 0; r < 16; r++)
        Ad[r] = (double)(A[r] / (float)m);
    ppfc::jacobi_eig<4>(Ad, V, W);
    int mx = 0;
    for (int r = 1; r < 4; r++)
        if (W[r] > W[mx])
            mx = r;
    double qv[4] = {V[0 * 4 + mx], V[1 * 4 + mx], V[2 * 4 + mx], V[3 * 4 + mx]};
    double nn    = sqrt(qv[0] * qv[0] + qv[1] * qv[1] + qv[2] * qv[2] + qv[3] * qv[3]);
    double sg    = qv[0] < 0.0 ? -1.0 : 1.0;
    float  q[4];
    for (int r = 0; r < 4; r++)
        q[r] = (float)(sg * qv[r] / nn);
    float R[9];
    ppfc::rot_from_quat(q, R);
    Hypo &o = out[c];
    for (int r = 0; r < 3; r++) {
        o.pose[r * 4 + 0] = R[r * 3 + 0];
        o.pose[r * 4 + 1] = R[r * 3 + 1];
        o.pose[r * 4 + 2] = R[r * 3 + 2];
    }
    float fm    = (float)m;
    o.pose[3]   = tx / fm;
    o.pose[7]   = ty / fm;
    o.pose[11]  = tz / fm;
    o.votes     = votes;
    float q2[4];
    ppfc::quat_from_rot(R, q2);
    angleOut[c] = ppfc::angle_of_quat(q2);
}

// clusterPose2 (util.cpp:522-551) on the vote-sorted cluster list: pose j is
// absorbed by the earliest kept pose within the overlap distance.  C is
// small (hundreds), so predecessors are scanned directly.
__global__ void __launch_bounds__(1024) nms_kernel(const Hypo *h, const int *ord, int C,
                                                   const float *center, float thr, int *status,
                                                   float *tr, Hypo *out, int *nOut) {
    __shared__ int sChanged, sUndecided;
    float sq = thr * thr;
    for (int i = threadIdx.x; i < C; i += blockDim.x) {
        const float *P = h[ord[i]].pose;
        for (int r = 0; r < 3; r++)
            tr[3 * i + r] = P[r * 4 + 0] * center[0] + P[r * 4 + 1] * center[1] +
                            P[r * 4 + 2] * center[2] + P[r * 4 + 3];
        status[i] = UNDECIDED;
    }
    __syncthreads();
    for (;;) {
        if (threadIdx.x == 0) {
            sChanged   = 0;
            sUndecided = 0;
        }
        __syncthreads();
        for (int j = threadIdx.x; j < C; j += blockDim.x) {
            if (status[j] != UNDECIDED)
                continue;
            int minKept = 0x7fffffff, minUnd = 0x7fffffff;
            for (int i = 0; i < j; i++) {
                float dx = tr[3 * i] - tr[3 * j], dy = tr[3 * i + 1] - tr[3 * j + 1],
                      dz = tr[3 * i + 2] - tr[3 * j + 2];
                if (!(dx * dx + dy * dy + dz * dz < sq))
                    continue;
                int si = status[i];
                if (si == CENTER) {
                    minKept = i;
                    break;
                }
                if (si == UNDECIDED) {
                    minUnd = i;
                    break;
                }
            }
            if (minUnd < minKept) {
                atomicAdd(&sUndecided, 1);
                continue;
            }
            status[j] = (minKept == 0x7fffffff) ? CENTER : minKept;
            sChanged  = 1;
        }
        __syncthreads();
        int und = sUndecided, ch = sChanged;
        __syncthreads();
        if (und == 0 || !ch)
            break;
    }
    // kept poses in order, each with its members' votes added in ascending
    // member order (the reference's accumulation order); one thread per pose
    __shared__ int sScan[1024 / 64 + 1];
    int            base = 0;
    for (int c0 = 0; c0 < C; c0 += 1024) {
        const int  i      = c0 + (int)threadIdx.x;
        const bool center = i < C && status[i] == CENTER;
        int        total;
        int        pos = block_excl_scan<1024>(center ? 1 : 0, sScan, total);
        if (center) {
            Hypo o = h[ord[i]];
            for (int j = i + 1; j < C; j++)
                if (status[j] == i)
                    o.votes = o.votes + h[ord[j]].votes;
            out[base + pos] = o;
        }
        base += total;
    }
    if (threadIdx.x == 0)
        *nOut = base;
}

// The same rule walked in order by one wave: pose j takes the first kept
// pose within the overlap distance, else it is kept.  The wave walks 64
// poses at a time: each lane first looks for its pose's first kept pose
// among those kept before the block (those precede any kept in the
// block); the lanes left over are then settled in order, each against the
// poses of the block kept so far (one ballot), becoming kept when none is
// near.  Used while the kept list fits in LDS.
constexpr int NMS_SEQ_MAX = 4096;

// NT threads: all of them load the poses and the votes (their global loads
// overlap), the first wave walks, all of them add up the kept poses' votes
template <int NT>
__device__ void nms_seq_range(const Hypo *h, const int *ord, int C, const float *center,
                              float thr, int *status, Hypo *out, int *nOut, float *sKept) {
    // kept: {x, y, z, pose index bits}[C]; then every pose's centre x/y/z[C]
    float4     *kp = reinterpret_cast<float4 *>(sKept);
    float      *px = sKept + 4 * C, *py = sKept + 5 * C, *pz = sKept + 6 * C;
    const int   lane = threadIdx.x;
    const float sq   = thr * thr;
    const float c0 = center[0], c1 = center[1], c2 = center[2];
    __shared__ int sNk;
    for (int j = lane; j < C; j += NT) {
        const float *P = h[ord[j]].pose;
        px[j]          = P[0] * c0 + P[1] * c1 + P[2] * c2 + P[3];
        py[j]          = P[4] * c0 + P[5] * c1 + P[6] * c2 + P[7];
        pz[j]          = P[8] * c0 + P[9] * c1 + P[10] * c2 + P[11];
    }
    __syncthreads();
    int  nk    = 0;  // wave-uniform
    int *sStat = reinterpret_cast<int *>(px);  // px[j] is dead once its block is walked
    for (int b0 = 0; b0 < (lane < 64 ? C : 0); b0 += 64) {
        const int  j     = b0 + lane;
        const bool valid = j < C;
        float      tx = 0.0f, ty = 0.0f, tz = 0.0f;
        if (valid) {
            tx = px[j];
            ty = py[j];
            tz = pz[j];
        }
        // the first pose kept before this block within the distance
        int       st  = UNDECIDED;
        const int nk0 = nk;
        auto      near_kept = [&](const float4 &k) {
            const float dx = k.x - tx, dy = k.y - ty, dz = k.z - tz;
            return dx * dx + dy * dy + dz * dz < sq;
        };
        if (valid) {
            // four kept poses per step, their loads issued together, tested in order
            int t = 0;
            for (; t + 3 < nk0; t += 4) {
                const float4 k0 = kp[t], k1 = kp[t + 1], k2 = kp[t + 2], k3 = kp[t + 3];
                const float  w  = near_kept(k0)   ? k0.w
                                  : near_kept(k1) ? k1.w
                                  : near_kept(k2) ? k2.w
                                  : near_kept(k3) ? k3.w
                                                  : __int_as_float(UNDECIDED);
                if (__float_as_int(w) != UNDECIDED) {
                    st = __float_as_int(w);
                    break;
                }
            }
            if (st == UNDECIDED)
                for (; t < nk0; t++) {
                    const float4 k0 = kp[t];
                    if (near_kept(k0)) {
                        st = __float_as_int(k0.w);
                        break;
                    }
                }
        }
        // the rest in order, against the poses kept in this block so far
        bool               keptHere = false;
        unsigned long long rest     = __ballot(valid && st == UNDECIDED);
        while (rest) {
            const int u = __ffsll((long long)rest) - 1;
            rest &= rest - 1ull;
            const float ux = __shfl(tx, u, 64), uy = __shfl(ty, u, 64), uz = __shfl(tz, u, 64);
            bool        near = false;
            if (keptHere) {
                const float dx = tx - ux, dy = ty - uy, dz = tz - uz;
                near           = dx * dx + dy * dy + dz * dz < sq;
            }
            const unsigned long long m = __ballot(near);
            if (m) {
                if (lane == u)
                    st = b0 + __ffsll((long long)m) - 1;
            } else {
                if (lane == u) {
                    st       = CENTER;
                    keptHere = true;
                    kp[nk]   = make_float4(tx, ty, tz, __int_as_float(j));
                }
                nk++;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (valid) {
            status[j] = st;
            sStat[j]  = st;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (lane == 0)
        sNk = nk;
    __syncthreads();
    nk = sNk;
    // every pose's votes next to its status in LDS (py is dead too)
    float *sVotes = py;
    for (int j = lane; j < C; j += NT)
        sVotes[j] = h[ord[j]].votes;
    __syncthreads();
    // kept poses in order, each with its members' votes added in ascending
    // member order (the reference's accumulation order)
    for (int t = lane; t < nk; t += NT) {
        const int i = __float_as_int(kp[t].w);
        Hypo      o = h[ord[i]];
        int       j = i + 1;
        // four statuses per step, loads issued together, votes added in order
        for (; j + 3 < C; j += 4) {
            const int s0 = sStat[j], s1 = sStat[j + 1], s2 = sStat[j + 2], s3 = sStat[j + 3];
            if (s0 == i)
                o.votes = o.votes + sVotes[j];
            if (s1 == i)
                o.votes = o.votes + sVotes[j + 1];
            if (s2 == i)
                o.votes = o.votes + sVotes[j + 2];
            if (s3 == i)
                o.votes = o.votes + sVotes[j + 3];
        }
        for (; j < C; j++)
            if (sStat[j] == i)
                o.votes = o.votes + sVotes[j];
        out[t] = o;
    }
    if (lane == 0)
        *nOut = nk;
}

constexpr int NMS_THREADS = 256;

__global__ void __launch_bounds__(NMS_THREADS) nms_seq_kernel(const Hypo *h, const int *ord, int C,
                                                              const float *center, float thr,
                                                              int *status, Hypo *out, int *nOut) {
    extern __shared__ float sKept[];
    nms_seq_range<NMS_THREADS>(h, ord, C, center, thr, status, out, nOut, sKept);
}

// one wave per scene: scene w's clusters are the sorted positions
// [off[w], off[w + 1]) of ord; its kept poses go to out + off[w], their
// count to nOut[w]
__global__ void __launch_bounds__(NMS_THREADS) nms_seq_multi_kernel(const Hypo *h, const int *ord,
                                                                    const int *off,
                                                                    const float *center, float thr,
                                                                    int *status, Hypo *out,
                                                                    int *nOut) {
    extern __shared__ float sKept[];
    const int b = off[blockIdx.x];
    nms_seq_range<NMS_THREADS>(h, ord + b, off[blockIdx.x + 1] - b, center, thr, status + b,
                               out + b, nOut + blockIdx.x, sKept);
}

__global__ void hypo_keys_kernel(const Hypo *h, int n, uint32_t *keys, int *vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    keys[i] = 0xFFFFFFFFu - __float_as_uint(h[i].votes);
    vals[i] = i;
}

static void sort_pairs(Workspace &ws, uint32_t *k, uint32_t *k2, int *v, int *v2, int n,
                       hipStream_t s) {
    size_t tb = 0;
    PPF_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tb, k, k2, v, v2, n, 0, 32, s));
    PPF_HIP_CHECK(rocprim::radix_sort_pairs(ws.cubTmp.reserve(tb), tb, k, k2, v, v2, n, 0, 32, s));
}

void cluster_poses(Workspace &ws, const PoseRec *poses, int np, float dThr, float aThr,
                   const float center[3], float nmsThr, std::vector<Hypo> &out, int *nClusters,
                   hipStream_t s) {
    out.clear();
    if (nClusters)
        *nClusters = 0;
    if (np <= 0)
        return;
    // every scratch array is sized for np up front (nc <= np), so no buffer
    // is reallocated while kernels that use it may still be queued
    uint32_t *k1 = ws.clSortKey.as<uint32_t>(np), *k2 = ws.clSortKey2.as<uint32_t>(np);
    int      *v1 = ws.clSortVal.as<int>(np), *v2 = ws.clSortVal2.as<int>(np);
    uint32_t *k3 = ws.clPred.as<uint32_t>(np);
    float    *buf = ws.clCenters.as<float>((size_t)np * 6);
    int      *status = ws.clStatus.as<int>((size_t)np * 4 + 8);
    int      *members = ws.clMembers.as<int>((size_t)np * 2 + 2);
    int      *cntsBuf = ws.clPredOff.as<int>((size_t)np * 3 + 4);
    Hypo     *centers = ws.clNms.as<Hypo>((size_t)np * 2 + 2);
    vote_keys_kernel<<<div_up(np, 256), 256, 0, s>>>(poses, np, k1, v1);
    sort_pairs(ws, k1, k2, v1, v2, np, s);
    int *ord = v2;  // sorted position -> pose index (stable: ties keep (keypoint, rank) order)
    SortedPoses sp{buf, buf + np, buf + 2 * (size_t)np, buf + 3 * (size_t)np, buf + 4 * (size_t)np,
                   ord};
    gather_sorted_kernel<<<div_up(np, 256), 256, 0, s>>>(poses, ord, np, sp);
    // long lists: a translation grid (ids are sorted positions) for the rounds
    int *rank   = status + np;
    int *isC    = status + 2 * (size_t)np;
    int *misc   = status + 3 * (size_t)np;
    if (np <= CLUSTER_SEQ_MAX) {
        const size_t lds = (size_t)20 * np;
        PPF_HIP_CHECK(hipFuncSetAttribute((const void *)cluster_seq_kernel,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        cluster_seq_kernel<<<1, 64, lds, s>>>(sp, np, dThr, aThr, status);
    } else {
        grid_build(ws.gNormal, sp.tx, sp.ty, sp.tz, nullptr, nullptr, nullptr, nullptr, np, dThr,
                   s);
        cluster_rounds_kernel<<<1, 1024, 0, s>>>(sp, np, ws.gNormal.view, dThr, aThr, status,
                                                 misc);
    }
    center_rank_kernel<<<div_up(np, 256), 256, 0, s>>>(status, np, isC);
    size_t tb = 0;
    PPF_HIP_CHECK(rocprim::exclusive_scan(nullptr, tb, isC, rank, 0, np, rocprim::plus<int>(), s));
    PPF_HIP_CHECK(rocprim::exclusive_scan(ws.cubTmp.reserve(tb), tb, isC, rank, 0, np, rocprim::plus<int>(), s));
    int lastRank = 0, lastIs = 0;
    PPF_HIP_CHECK(hipMemcpyAsync(&lastRank, rank + np - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipMemcpyAsync(&lastIs, isC + np - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
    int nc = lastRank + lastIs;
    if (nClusters)
        *nClusters = nc;
    // members grouped by cluster, ascending sorted position inside
    member_cluster_kernel<<<div_up(np, 256), 256, 0, s>>>(status, rank, np, k1, v1);
    int endBit = 1;
    while ((1 << endBit) < nc + 1)
        endBit++;
    tb = 0;
    PPF_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tb, k1, k3, v1, members, np, 0, endBit, s));
    int *mOff    = members + np;
    PPF_HIP_CHECK(rocprim::radix_sort_pairs(ws.cubTmp.reserve(tb), tb, k1, k3, v1, members, np, 0, endBit, s));
    // cluster offsets: counts by run-length of k3 (clusters are 0..nc-1, all non-empty)
    int *cnts = cntsBuf;
    uint32_t *uk = reinterpret_cast<uint32_t *>(k1);
    int      *nr = misc + 1;
    tb = 0;
    PPF_HIP_CHECK(rocprim::run_length_encode(nullptr, tb, k3, np, uk, cnts, nr, s));
    PPF_HIP_CHECK(rocprim::run_length_encode(ws.cubTmp.reserve(tb), tb, k3, np, uk, cnts, nr, s));
    PPF_HIP_CHECK(hipMemsetAsync(cnts + nc, 0, sizeof(int), s));
    tb = 0;
    PPF_HIP_CHECK(rocprim::exclusive_scan(nullptr, tb, cnts, mOff, 0, nc + 1, rocprim::plus<int>(), s));
    PPF_HIP_CHECK(rocprim::exclusive_scan(ws.cubTmp.reserve(tb), tb, cnts, mOff, 0, nc + 1, rocprim::plus<int>(), s));
    // v2 still holds ord (the second sort wrote to `members`)
    float *angles  = buf + (size_t)np * 5;
    cluster_avg_kernel<<<div_up(nc, 64), 64, 0, s>>>(poses, ord, members, mOff, nc, centers,
                                                     angles);
    // NMS on clusters sorted by votes (stable on creation order)
    uint32_t *hk = k1;
    int      *hv = cntsBuf + nc + 2;
    uint32_t *hk2 = k3;
    int      *hv2 = status;  // reuse
    hypo_keys_kernel<<<div_up(nc, 256), 256, 0, s>>>(centers, nc, hk, hv);
    sort_pairs(ws, hk, hk2, hv, hv2, nc, s);
    float *cdev = reinterpret_cast<float *>(misc + 4);
    PPF_HIP_CHECK(hipMemcpyAsync(cdev, center, sizeof(float) * 3, hipMemcpyHostToDevice, s));
    int  *nmsStatus = rank;  // reuse (nc <= np)
    float *tr       = buf;   // reuse
    Hypo *kept      = centers + nc;
    int  *nOut      = misc + 2;
    if (nc <= NMS_SEQ_MAX) {
        const size_t lds = (size_t)28 * std::max(nc, 1);
        PPF_HIP_CHECK(hipFuncSetAttribute((const void *)nms_seq_kernel,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        nms_seq_kernel<<<1, NMS_THREADS, lds, s>>>(centers, hv2, nc, cdev, nmsThr, nmsStatus, kept,
                                                   nOut);
    } else {
        nms_kernel<<<1, 1024, 0, s>>>(centers, hv2, nc, cdev, nmsThr, nmsStatus, tr, kept, nOut);
    }
    // the kept count and every slot it may cover (at most nc) in one read-back
    int nk = 0;
    out.resize(std::max(nc, 1));
    PPF_HIP_CHECK(hipMemcpyAsync(&nk, nOut, sizeof(int), hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipMemcpyAsync(out.data(), kept, sizeof(Hypo) * nc, hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
    if (nk < 0 || nk > nc)
        throw std::logic_error("cluster_poses: kept count out of range");
    out.resize(nk);
    PPF_HIP_CHECK(hipGetLastError());
}

// ---- several scenes at once (the batched match): the same steps over the
// concatenated pose list, every sort keyed by (scene, key) so each scene's
// poses and clusters stay contiguous and in their own order, one wave per
// scene for the ordered walks

bool cluster_multi_fits(const std::vector<int> &pOff) {
    for (size_t w = 0; w + 1 < pOff.size(); w++)
        if (pOff[w + 1] - pOff[w] > CLUSTER_MULTI_MAX)
            return false;
    return true;
}

// scene w of item i: off[w] <= i < off[w + 1] (off non-decreasing)
__device__ __forceinline__ int scene_of(const int *off, int S, int i) {
    int lo = 0, hi = S;  // first w with off[w] > i, minus one
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= i)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo - 1;
}

__global__ void vote_keys_multi_kernel(const PoseRec *p, int n, const int *off, int S,
                                       unsigned long long *keys, int *vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    keys[i] = ((unsigned long long)scene_of(off, S, i) << 32) |
              (0xFFFFFFFFu - __float_as_uint(p[i].votes));
    vals[i] = i;
}

// each scene's first cluster index (centres are ranked scene by scene, in
// creation order), and the total at off[S]
__global__ void scene_cluster_off_kernel(const int *rank, const int *isC, const int *off, int S,
                                         int np, int *cOff) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w > S)
        return;
    const int total = np > 0 ? rank[np - 1] + isC[np - 1] : 0;
    cOff[w]         = off[w] < np ? rank[off[w]] : total;
}

__global__ void hypo_keys_multi_kernel(const Hypo *h, int n, const int *cOff, int S,
                                       unsigned long long *keys, int *vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    keys[i] = ((unsigned long long)scene_of(cOff, S, i) << 32) |
              (0xFFFFFFFFu - __float_as_uint(h[i].votes));
    vals[i] = i;
}

static int bits_for(long long v) {
    int b = 1;
    while ((1ll << b) < v)
        b++;
    return b;
}

void cluster_poses_multi(Workspace &ws, const PoseRec *poses, const std::vector<int> &pOff,
                         float dThr, float aThr, const float center[3], float nmsThr,
                         std::vector<std::vector<Hypo>> &out, std::vector<int> &nClusters,
                         hipStream_t s) {
    const int S  = (int)pOff.size() - 1;
    const int np = S > 0 ? pOff[S] : 0;
    out.assign(std::max(S, 0), {});
    nClusters.assign(std::max(S, 0), 0);
    if (np <= 0)
        return;
    int maxNp = 0;
    for (int w = 0; w < S; w++)
        maxNp = std::max(maxNp, pOff[w + 1] - pOff[w]);
    if (maxNp > CLUSTER_MULTI_MAX)
        throw std::logic_error("cluster_poses_multi: a scene holds more poses than its walk keeps");
    unsigned long long *ka = ws.clKey64.as<unsigned long long>(np),
                       *kb = ws.clKey64b.as<unsigned long long>(np);
    uint32_t *k1 = ws.clSortKey.as<uint32_t>(np), *k3 = ws.clPred.as<uint32_t>(np);
    int      *v1 = ws.clSortVal.as<int>(np), *v2 = ws.clSortVal2.as<int>(np);
    float    *buf     = ws.clCenters.as<float>((size_t)np * 6);
    int      *status  = ws.clStatus.as<int>((size_t)np * 4 + 8);
    int      *members = ws.clMembers.as<int>((size_t)np * 2 + 2);
    int      *cntsBuf = ws.clPredOff.as<int>((size_t)np * 3 + 4);
    Hypo     *centers = ws.clNms.as<Hypo>((size_t)np * 2 + 2);
    int      *sOff    = ws.clSceneOff.as<int>((size_t)3 * (S + 1) + 4);
    int      *cOff    = sOff + (S + 1);
    int      *nOut    = cOff + (S + 1);
    float    *cdev    = reinterpret_cast<float *>(nOut + S + 1);
    PPF_HIP_CHECK(hipMemcpyAsync(sOff, pOff.data(), sizeof(int) * (S + 1), hipMemcpyHostToDevice, s));
    PPF_HIP_CHECK(hipMemcpyAsync(cdev, center, sizeof(float) * 3, hipMemcpyHostToDevice, s));
    const int sceneBits = bits_for(S);
    vote_keys_multi_kernel<<<div_up(np, 256), 256, 0, s>>>(poses, np, sOff, S, ka, v1);
    size_t tb = 0;
    PPF_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tb, ka, kb, v1, v2, np, 0, 32 + sceneBits, s));
    PPF_HIP_CHECK(rocprim::radix_sort_pairs(ws.cubTmp.reserve(tb), tb, ka, kb, v1, v2, np, 0,
                                            32 + sceneBits, s));
    int *ord = v2;  // sorted position -> pose index
    SortedPoses sp{buf, buf + np, buf + 2 * (size_t)np, buf + 3 * (size_t)np, buf + 4 * (size_t)np,
                   ord};
    gather_sorted_kernel<<<div_up(np, 256), 256, 0, s>>>(poses, ord, np, sp);
    {
        const size_t lds = (size_t)20 * maxNp;
        PPF_HIP_CHECK(hipFuncSetAttribute((const void *)cluster_seq_multi_kernel,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        cluster_seq_multi_kernel<<<S, 64, lds, s>>>(sp, sOff, dThr, aThr, status);
    }
    int *rank = status + np;
    int *isC  = status + 2 * (size_t)np;
    int *misc = status + 3 * (size_t)np;
    center_rank_kernel<<<div_up(np, 256), 256, 0, s>>>(status, np, isC);
    tb = 0;
    PPF_HIP_CHECK(rocprim::exclusive_scan(nullptr, tb, isC, rank, 0, np, rocprim::plus<int>(), s));
    PPF_HIP_CHECK(rocprim::exclusive_scan(ws.cubTmp.reserve(tb), tb, isC, rank, 0, np,
                                          rocprim::plus<int>(), s));
    scene_cluster_off_kernel<<<div_up(S + 1, 256), 256, 0, s>>>(rank, isC, sOff, S, np, cOff);
    std::vector<int> co(S + 1);
    PPF_HIP_CHECK(hipMemcpyAsync(co.data(), cOff, sizeof(int) * (S + 1), hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
    const int nc = co[S];
    int       maxC = 0;
    for (int w = 0; w < S; w++) {
        nClusters[w] = co[w + 1] - co[w];
        maxC         = std::max(maxC, nClusters[w]);
    }
    // members grouped by cluster (clusters numbered scene by scene), ascending
    // sorted position inside
    member_cluster_kernel<<<div_up(np, 256), 256, 0, s>>>(status, rank, np, k1, v1);
    const int endBit = bits_for((long long)nc + 1);
    tb = 0;
    PPF_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tb, k1, k3, v1, members, np, 0, endBit, s));
    PPF_HIP_CHECK(rocprim::radix_sort_pairs(ws.cubTmp.reserve(tb), tb, k1, k3, v1, members, np, 0,
                                            endBit, s));
    int      *mOff = members + np;
    int      *cnts = cntsBuf;
    uint32_t *uk   = k1;
    int      *nr   = misc + 1;
    tb             = 0;
    PPF_HIP_CHECK(rocprim::run_length_encode(nullptr, tb, k3, np, uk, cnts, nr, s));
    PPF_HIP_CHECK(rocprim::run_length_encode(ws.cubTmp.reserve(tb), tb, k3, np, uk, cnts, nr, s));
    PPF_HIP_CHECK(hipMemsetAsync(cnts + nc, 0, sizeof(int), s));
    tb = 0;
    PPF_HIP_CHECK(rocprim::exclusive_scan(nullptr, tb, cnts, mOff, 0, nc + 1, rocprim::plus<int>(), s));
    PPF_HIP_CHECK(rocprim::exclusive_scan(ws.cubTmp.reserve(tb), tb, cnts, mOff, 0, nc + 1,
                                          rocprim::plus<int>(), s));
    float *angles = buf + (size_t)np * 5;
    cluster_avg_kernel<<<div_up(nc, 64), 64, 0, s>>>(poses, ord, members, mOff, nc, centers, angles);
    // each scene's clusters by votes (stable on creation order), then its NMS
    int *hv = cntsBuf + nc + 2, *hv2 = status;  // status is free again (nc <= np)
    hypo_keys_multi_kernel<<<div_up(nc, 256), 256, 0, s>>>(centers, nc, cOff, S, ka, hv);
    tb = 0;
    PPF_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tb, ka, kb, hv, hv2, nc, 0, 32 + sceneBits, s));
    PPF_HIP_CHECK(rocprim::radix_sort_pairs(ws.cubTmp.reserve(tb), tb, ka, kb, hv, hv2, nc, 0,
                                            32 + sceneBits, s));
    int  *nmsStatus = rank;  // reuse (nc <= np)
    Hypo *kept      = centers + nc;
    if (maxC <= NMS_SEQ_MAX) {
        const size_t lds = (size_t)28 * std::max(maxC, 1);
        PPF_HIP_CHECK(hipFuncSetAttribute((const void *)nms_seq_multi_kernel,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        nms_seq_multi_kernel<<<S, NMS_THREADS, lds, s>>>(centers, hv2, cOff, cdev, nmsThr, nmsStatus, kept,
                                                nOut);
    } else {
        float *tr = buf;  // reuse
        for (int w = 0; w < S; w++)
            nms_kernel<<<1, 1024, 0, s>>>(centers, hv2 + co[w], co[w + 1] - co[w], cdev, nmsThr,
                                          nmsStatus + co[w], tr, kept + co[w], nOut + w);
    }
    std::vector<int>  nk(S);
    std::vector<Hypo> all(std::max(nc, 1));
    PPF_HIP_CHECK(hipMemcpyAsync(nk.data(), nOut, sizeof(int) * S, hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipMemcpyAsync(all.data(), kept, sizeof(Hypo) * nc, hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
    PPF_HIP_CHECK(hipGetLastError());
    for (int w = 0; w < S; w++)
        out[w].assign(all.begin() + co[w], all.begin() + co[w] + nk[w]);
}

}  // namespace ppfd
