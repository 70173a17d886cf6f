// vote.hip -- scene voting, peak extraction and pose construction on gfx950.
//
// Reference: ppf.cpp:247-409.  Per keypoint the reference radius-searches
// the sampled scene (radius = model diameter), computes the PPF key and
// alpha of every neighbour, looks the key up in the model table and adds
// one vote per bucket entry to a refNum x (angleNum+1) int accumulator, then
// extracts up to three peaks.
//
// Here one workgroup owns one keypoint (and one tile of model references if
// the accumulator does not fit in LDS).  The accumulator lives in LDS
// without the reference's per-ref count slot (it always equals the sum of
// the ref's bins, so the peak stage recomputes it).  Neighbour candidates
// come from the hashed grid cells overlapping the vote sphere.  Each wave
// works on its own: it takes chunks of 64 candidates from a shared counter,
// compacts the in-sphere pairs into a per-wave LDS queue so the feature and
// key stage runs on full waves, queues the key hits, and votes 64 hits at a
// time by cutting their buckets into segments of 8 entries striped over its
// lanes (bucket-size skew does not idle lanes).  The last queued hits of all
// waves are voted by the whole workgroup.  Integer votes commute, so the
// accumulator does not depend on the order of any of this.
#include "ppf_dev.h"

#include <type_traits>

#include <cstring>
#include <rocprim/rocprim.hpp>

#include <algorithm>

namespace ppfd {

constexpr int VOTE_AXIS_CELLS = 11;  // cells per axis over the vote sphere (host check)
constexpr int VOTE_MAX_CELLS  = VOTE_AXIS_CELLS * VOTE_AXIS_CELLS * VOTE_AXIS_CELLS;
constexpr int SEG            = 8;  // bucket entries voted by one lane
constexpr int VOTE_STATIC_LDS = 55 * 1024;   // bound on the kernel's static LDS (1024 threads)
constexpr int VOTE_MAX_CHUNKS = 512;         // chunks of 64 candidates with a first-cell entry
constexpr int LDS_LIMIT       = 160 * 1024;  // gfx950 LDS per workgroup

// references per accumulator tile: the LDS left beside the kernels' own
// arrays, less the spare row that single-tile voting sends padding entries to
int vote_max_tile_refs(int angleNum, int block) {
    (void)block;
    return (LDS_LIMIT - VOTE_STATIC_LDS) / (angleNum * 4) - 1;
}

struct VoteArgs {
    GridView        g;      // sampled scene, sorted, with normals (one scene)
    const GridView *grids;  // per scene (several scenes; nullptr: g)
    const int      *kpScene;  // scene of each key point (several scenes)
    // key point k: position / normal, and its point id in its scene (self
    // exclusion)
    const float *kpX, *kpY, *kpZ, *kpNX, *kpNY, *kpNZ;
    const int   *kp;
    int          nk;
    float        D;  // model diameter (vote radius)
    const uint4 *slots;
    uint32_t     tmask;
    const uint2 *ventries;      // bucket entries {ref x angleNum, alpha}, buckets 64 B aligned
    int          refNum, angleNum, tileRefs, nTiles;
    float        angleStep, distStep, sMaxId, D2, voteThreshold;
    unsigned long long *cand;   // [nk * nTiles * 3]
    unsigned long long *stats;  // pairs, hits, votes (optional)
    int32_t     *accDump;       // [dumpCount * refNum * angleNum] (optional)
    int          dumpCount;
    int          xcdRemap;
    const int   *kpSel;         // key points to vote (nullptr: all nk)
    // split path: hits of key point k at [k * hitCap, k * hitCap + nHit[k])
    int          hitCap;
    uint32_t    *hitStart, *hitLen, *hitBucket;
    float       *hitAlpha;
    // with buckets sorted by reference: entry index where tile t of bucket b
    // starts, splits[b * (nTiles + 1) + t] (nullptr: tiles scan whole buckets)
    const uint32_t *splits;
    int         *nHit, *nbr;    // hit count (-1: overflow), neighbour count incl. self
    int         *overflow, *nOverflow;
    int          countPairs;    // fused kernel adds its pairs and hits to stats
    // quantized-feature pre-filters of the pair kernel (nullptr: none)
    const uint32_t *filter;
    int          fA, fD;
    float        maxBin;   // angleNum - 1
    uint32_t     maxMult;  // most entries of one reference in one bucket
};

__device__ __forceinline__ bool table_find(const uint4 *slots, uint32_t tmask, uint32_t key,
                                           uint32_t &start, uint32_t &len, uint32_t &bucket) {
    uint32_t h = key & tmask;
    for (;;) {
        uint4 s = slots[h];
        if (s.z == 0u)
            return false;
        if (s.x == key) {
            start  = s.y;
            len    = s.z;
            bucket = s.w;
            return true;
        }
        h = (h + 1u) & tmask;
    }
}

// candidate code: (vote << 32) | ((0xFFFFF - ref) << 11) | bin -- larger is
// better under the reference's (vote desc, ref asc) order.  The best three
// codes, descending; insertion is a branch-free min/max network, so the
// three stay in registers.
struct Top3 {
    unsigned long long a, b, c;
};
__device__ __forceinline__ Top3 top3_add(Top3 t, unsigned long long x) {
    const unsigned long long hi1 = max(t.a, x), lo1 = min(t.a, x);
    const unsigned long long hi2 = max(t.b, lo1), lo2 = min(t.b, lo1);
    return Top3{hi1, hi2, max(t.c, lo2)};
}

// Votes of one segment: the 8 bucket entries from entry index e0 (a
// multiple of 8: buckets start on 64-byte boundaries, so the segment is four
// aligned 16-byte loads) of which entries [e0 + vlo, e0 + vhi) are voted,
// into the accumulator tile whose first reference row starts at r0A
// (= r0 x angleNum); entries from tailStart on take the bucket-tail bin rule.
// The vote entries hold ref x angleNum, so the accumulator index is one add.
// Returns the votes cast.
struct SegRun {
    uint32_t e0, tailStart;
    int      vlo, vhi;
    float    offBody, offTail;
};

__device__ __forceinline__ void load_segment(const VoteArgs &a, uint32_t e0, uint2 (&v)[SEG]) {
    const uint4 *p = reinterpret_cast<const uint4 *>(a.ventries + e0);
#pragma unroll
    for (int q = 0; q < SEG / 2; q++) {
        const uint4 w = p[q];
        v[2 * q]      = make_uint2(w.x, w.y);
        v[2 * q + 1]  = make_uint2(w.z, w.w);
    }
}
// N consecutive entries from e (a multiple of N; N = 1, 2, 4 or 8)
template <int N>
__device__ __forceinline__ void load_entries(const VoteArgs &a, uint32_t e, uint2 (&v)[N]) {
    if (N == 1) {
        v[0] = a.ventries[e];
    } else {
        const uint4 *p = reinterpret_cast<const uint4 *>(a.ventries + e);
#pragma unroll
        for (int q = 0; q < N / 2; q++) {
            const uint4 w = p[q];
            v[2 * q]      = make_uint2(w.x, w.y);
            v[2 * q + 1]  = make_uint2(w.z, w.w);
        }
    }
}

//
// A segment is all body or all tail: buckets and their tails start on
// multiples of 8 entries.  With one reference tile (RANGE false) the run is
// [0, vhi) and the entries past vhi are the bucket's padding, whose
// reference is the accumulator's spare row (refNum x angleNum), so every
// entry is voted without a per-entry mask.
template <bool RANGE>
__device__ __forceinline__ int vote_run_loaded(const VoteArgs &a, uint32_t *acc, int r0A, int r1A,
                                               const SegRun &g, const uint2 (&v)[SEG]) {
    const float off = g.e0 >= g.tailStart ? g.offTail : g.offBody;
    if (!RANGE) {
#pragma unroll
        for (int kk = 0; kk < SEG; kk++) {
            const int bin = ppfc::bin_wrapped(__uint_as_float(v[kk].y), off, a.sMaxId, a.angleNum);
            atomicAdd(&acc[(int)v[kk].x + bin], 1u);
        }
        return g.vhi - g.vlo;
    }
    int nv = 0;
#pragma unroll
    for (int kk = 0; kk < SEG; kk++) {
        const int refA = (int)v[kk].x;
        if (kk >= g.vlo && kk < g.vhi && refA >= r0A && refA < r1A) {
            const int bin = ppfc::bin_wrapped(__uint_as_float(v[kk].y), off, a.sMaxId, a.angleNum);
            atomicAdd(&acc[refA - r0A + bin], 1u);
            nv++;
        }
    }
    return nv;
}

// The same for one reference tile with 16-bit counters, two per word: the
// vote adds 1 to the cell's half word.  The returned old words are OR-ed
// into ovf: a counter that ends above 0x8000 was incremented from 0x8000 at
// some point, and that increment returned a word with bit 15 or 31 set.  So
// when no returned word had either bit set, every counter stayed at or below
// 0x8000 and no half ever carried into its neighbour; otherwise (checked
// once, after the votes) the key point is voted again with 32-bit counters.
// A set bit from a counter that stopped at exactly 0x8000, or from the
// spare row, only sends the key point to the 32-bit kernel unnecessarily.
//
// RET false: the key point's hits cannot bring any real cell past 0xFFFF
// (n hits x maxMult <= 0xFFFF), so the adds return nothing; the spare row
// may wrap, which only moves counts between spare cells.
// Entry alphas and scene alphas are finite (training skips NaN alphas,
// ppf.cpp:131-139, loading rejects them; hits have finite scene alphas), so
// the bin is bin_wrapped_fast's.  The accumulator is the kernel's only LDS
// and starts at LDS address 0 (checked on the host before the first
// launch), so the adds go through an LDS pointer of known value 0 and the
// word index, masked to 18 bits (the accumulator is far smaller), is the
// whole address computation.
// With N < SEG the segment is shared by SEG / N lanes, each voting N of its
// entries (the caller counts the segment's votes once).
typedef __attribute__((address_space(3))) uint32_t lds_u32;
template <bool RET, int N>
__device__ __forceinline__ int vote_run_packed(const VoteArgs &a, const SegRun &g,
                                               const uint2 (&v)[N], uint32_t &ovf) {
    lds_u32    *A   = (lds_u32 *)(size_t)0;
    const float off = g.e0 >= g.tailStart ? g.offTail : g.offBody;
#pragma unroll
    for (int kk = 0; kk < N; kk++) {
        const uint32_t bin =
            (uint32_t)ppfc::bin_wrapped_fast(__uint_as_float(v[kk].y), off, a.sMaxId, a.maxBin);
        const uint32_t cell = v[kk].x + bin;
        const uint32_t inc  = 1u << ((cell & 1u) << 4);
        lds_u32       *w    = A + ((cell >> 1) & 0x3FFFFu);
        if (RET)
            ovf |= __hip_atomic_fetch_add(w, inc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else
            __hip_atomic_fetch_add(w, inc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    return g.vhi - g.vlo;
}

template <bool RANGE>
__device__ __forceinline__ int vote_run(const VoteArgs &a, uint32_t *acc, int r0A, int r1A,
                                        const SegRun &g) {
    uint2 v[SEG];
    load_segment(a, g.e0, v);
    return vote_run_loaded<RANGE>(a, acc, r0A, r1A, g, v);
}

// two segments with their entry loads issued together (the second only
// when it holds entries)
template <bool RANGE>
__device__ __forceinline__ int vote_run2(const VoteArgs &a, uint32_t *acc, int r0A, int r1A,
                                         const SegRun &g1, const SegRun &g2) {
    uint2 v1[SEG], v2[SEG];
    load_segment(a, g1.e0, v1);
    if (g2.vhi > g2.vlo)
        load_segment(a, g2.e0, v2);
    int nv = vote_run_loaded<RANGE>(a, acc, r0A, r1A, g1, v1);
    if (g2.vhi > g2.vlo)
        nv += vote_run_loaded<RANGE>(a, acc, r0A, r1A, g2, v2);
    return nv;
}

// orders this wave's LDS writes before its later LDS reads (a wave runs in
// lock step; this keeps the compiler from reordering across the point)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Vote one queue of hits (bucket start, length, scene alpha) held by this
// wave: each bucket is cut into segments of SEG entries and the wave's lanes
// take segments in turn.  The reference's tail rule covers the last len % 8
// entries, which start at a multiple of 8, so a segment is entirely body or
// entirely tail.  Lane-uniform control flow except inside the segment loop.
template <int QCAP>
__device__ __forceinline__ void wave_vote(const VoteArgs &a, uint32_t *acc, int r0, int r1,
                                          const uint32_t *qStart, const uint32_t *qLen,
                                          const float *qAlpha, int *qPre, int qn,
                                          unsigned long long &nVotes) {
    const int lane = threadIdx.x & 63;
    int       s0 = 0, s1 = 0;
    if (lane < qn)
        s0 = (int)((qLen[lane] + SEG - 1) / SEG);
    if (lane + 64 < qn)
        s1 = (int)((qLen[lane + 64] + SEG - 1) / SEG);
    int x0 = s0, x1 = s1;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y0 = __shfl_up(x0, o, 64), y1 = __shfl_up(x1, o, 64);
        if (lane >= o) {
            x0 += y0;
            x1 += y1;
        }
    }
    const int tot0     = __shfl(x0, 63, 64);
    const int totalSeg = tot0 + __shfl(x1, 63, 64);
    if (lane < qn)
        qPre[lane] = x0 - s0;
    if (lane + 64 < qn)
        qPre[lane + 64] = tot0 + x1 - s1;
    wave_sync();
    const int r0A = r0 * a.angleNum, r1A = r1 * a.angleNum;
    int       nv  = 0;
    for (int t = lane; t < totalSeg; t += 64) {
        int lo = 0, hi = qn - 1;  // queue entry owning segment t
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (qPre[mid] <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        const uint32_t st = qStart[lo], L = qLen[lo];
        SegRun         g;
        g.e0        = st + (uint32_t)(t - qPre[lo]) * SEG;
        g.vlo       = 0;
        g.vhi       = (int)min((uint32_t)SEG, st + L - g.e0);
        g.tailStart = st + L - (L & 7u);
        g.offBody   = ppfc::bin_offset_body(qAlpha[lo]);
        g.offTail   = ppfc::bin_offset_tail(qAlpha[lo]);
        nv += a.nTiles > 1 ? vote_run<true>(a, acc, r0A, r1A, g) : vote_run<false>(a, acc, r0A, r1A, g);
    }
    nVotes += (unsigned long long)nv;
    // the queue is refilled after this point
    wave_sync();
}

// ---- pieces shared by the kernels below

// block -> (key point, tile).  Workgroups are dealt round-robin to the 8
// XCDs (each with its own L2).  With xcdRemap, XCD x works through a
// contiguous run of key points, so neighbouring key points -- which read the
// same scene cells and similar buckets -- share an L2.  The grid is then
// padded to 8 * per * nTiles blocks.  Returns -1 for a padding block.
__device__ __forceinline__ int block_keypoint(const VoteArgs &a, int nTiles, int &tile) {
    const int b = blockIdx.x;
    tile        = b % nTiles;
    int kb      = b / nTiles;
    if (a.xcdRemap) {
        int per = (a.nk + 7) / 8;
        kb      = (kb % 8) * per + kb / 8;
    }
    if (kb >= a.nk)
        return -1;
    return a.kpSel ? a.kpSel[kb] : kb;
}

// Cells of the sampled-scene grid overlapping the vote sphere around q (any
// order: voting commutes), their candidate prefix and, for every chunk of 64
// candidates below VOTE_MAX_CHUNKS, the cell holding its first candidate.
// Returns the candidate count; the caller synchronises before use.
template <int BLOCK>
__device__ int vote_cells(const VoteArgs &a, const GridView &g, float qx, float qy, float qz,
                          int *sCellStart, int *sCellPre, int *sChunkCell, int *sScan,
                          int &nCells) {
    constexpr int CPT = (VOTE_MAX_CELLS + BLOCK - 1) / BLOCK;
    const int     tid = threadIdx.x;
    const float   R   = a.D * 1.0001f + 1e-3f * g.cell;  // the sphere's cells, with a margin
    int xlo = cell_coord(qx - R, g.ox, g.cell, g.dx), xhi = cell_coord(qx + R, g.ox, g.cell, g.dx);
    int ylo = cell_coord(qy - R, g.oy, g.cell, g.dy), yhi = cell_coord(qy + R, g.oy, g.cell, g.dy);
    int zlo = cell_coord(qz - R, g.oz, g.cell, g.dz), zhi = cell_coord(qz + R, g.oz, g.cell, g.dz);
    int ncx = xhi - xlo + 1, ncy = yhi - ylo + 1, ncz = zhi - zlo + 1;
    nCells  = ncx * ncy * ncz;  // <= VOTE_MAX_CELLS: the host bounds cells per axis
    int cStart[CPT], cCnt[CPT], mySum = 0;
#pragma unroll
    for (int j = 0; j < CPT; j++) {
        int ci    = tid * CPT + j;
        cStart[j] = 0;
        cCnt[j]   = 0;
        if (ci < nCells) {
            int cx = xlo + ci % ncx, cy = ylo + (ci / ncx) % ncy, cz = zlo + ci / (ncx * ncy);
            if (!coarse_empty(g, cx, cy, cz) && cell_min_d2(g, cx, cy, cz, qx, qy, qz) <= a.D2) {
                int s0, c0;
                if (cell_lookup(g, cx, cy, cz, s0, c0)) {
                    cStart[j] = s0;
                    cCnt[j]   = c0;
                }
            }
        }
        mySum += cCnt[j];
    }
    int total;
    int pre = block_excl_scan<BLOCK>(mySum, sScan, total);
#pragma unroll
    for (int j = 0; j < CPT; j++) {
        int ci = tid * CPT + j;
        if (ci < nCells) {
            sCellStart[ci] = cStart[j];
            sCellPre[ci]   = pre;
            // chunk starts 64 t inside [pre, pre + cnt) begin in this cell
            for (int t = (pre + 63) >> 6; t < min((pre + cCnt[j] + 63) >> 6, VOTE_MAX_CHUNKS); t++)
                sChunkCell[t] = ci;
        }
        pre += cCnt[j];
    }
    if (tid == 0)
        sCellPre[nCells] = total;
    return total;
}

// sorted-grid slot of candidate c (chunk t = c / 64)
__device__ __forceinline__ int candidate_slot(int t, int c, int nCells, const int *sCellStart,
                                              const int *sCellPre, const int *sChunkCell) {
    int lo;  // the cell holding candidate c: largest cell with sCellPre[cell] <= c
    if (t < VOTE_MAX_CHUNKS) {
        lo = sChunkCell[t];  // holds candidate 64 t; c is at most 63 further
        while (sCellPre[lo + 1] <= c)
            lo++;
    } else {
        lo     = 0;
        int hi = nCells - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (sCellPre[mid] <= c)
                lo = mid;
            else
                hi = mid - 1;
        }
    }
    return sCellStart[lo] + (c - sCellPre[lo]);
}

// peaks of a voted accumulator tile (ppf.cpp:363-391): the best three
// (vote, ref, bin) codes of the tile
// accumulator cell c: a 32-bit word, or with PACKED a 16-bit half word
template <bool PACKED>
__device__ __forceinline__ uint32_t acc_get(const uint32_t *acc, int c) {
    if (!PACKED)
        return acc[c];
    return (acc[c >> 1] >> ((c & 1) << 4)) & 0xFFFFu;
}

constexpr int PEAK_BINS = 30;  // featAngleResolution's default (type.h:71)

template <int BLOCK, bool PACKED = false>
__device__ void peaks_top3(const VoteArgs &a, const uint32_t *acc, int nrTile, int r0,
                           unsigned long long *sTop, unsigned long long *candOut) {
    const int   tid  = threadIdx.x;
    const float thre = a.voteThreshold / 2.0f;
    const int   A    = a.angleNum, maxI = A - 1;
    Top3        top{0ull, 0ull, 0ull};
    if (PACKED && A == PEAK_BINS) {
        // the default 30 bins, packed two to a word: a reference's bins are
        // the 15 words from r * 15 on, read once, unrolled
        for (int r = tid; r < nrTile; r += BLOCK) {
            const uint32_t *w   = acc + r * (PEAK_BINS / 2);
            uint32_t        wd[PEAK_BINS / 2];
#pragma unroll
            for (int q = 0; q < PEAK_BINS / 2; q++)
                wd[q] = w[q];
            int sum = 0, best = -1, j = 0;
#pragma unroll
            for (int q = 0; q < PEAK_BINS / 2; q++) {
                const int lo = (int)(wd[q] & 0xFFFFu), hi = (int)(wd[q] >> 16);
                sum += lo + hi;
                if (lo > best) {
                    best = lo;
                    j    = 2 * q;
                }
                if (hi > best) {
                    best = hi;
                    j    = 2 * q + 1;
                }
            }
            if ((float)sum < thre || (float)best < thre)
                continue;
            const int jl = j == 0 ? maxI : j - 1, jr = j == maxI ? 0 : j + 1;
            const int v  = best + (int)acc_get<true>(acc, r * A + jl) +
                          (int)acc_get<true>(acc, r * A + jr);
            unsigned long long code = ((unsigned long long)(uint32_t)v << 32) |
                                      ((unsigned long long)(0xFFFFFu - (uint32_t)(r0 + r)) << 11) |
                                      (uint32_t)j;
            top = top3_add(top, code);
        }
    }
    for (int r = (PACKED && A == PEAK_BINS) ? nrTile : tid; r < nrTile; r += BLOCK) {
        auto bins = [&](int x) { return acc_get<PACKED>(acc, r * A + x); };
        int sum = 0, j = 0, best = -1;
        for (int x = 0; x < A; x++) {
            int v = (int)bins(x);
            sum += v;
            if (v > best) {
                best = v;
                j    = x;
            }
        }
        if ((float)sum < thre || (float)best < thre)
            continue;
        int v = best + (int)((j == 0) ? bins(maxI) : bins(j - 1)) +
                (int)((j == maxI) ? bins(0) : bins(j + 1));
        unsigned long long code = ((unsigned long long)(uint32_t)v << 32) |
                                  ((unsigned long long)(0xFFFFFu - (uint32_t)(r0 + r)) << 11) |
                                  (uint32_t)j;
        top = top3_add(top, code);
    }
    // the wave's top-3: usually few lanes found a peak, so their lists are
    // gathered lane by lane (ascending lane, as readlane broadcasts); many
    // lanes merge by an xor butterfly.  Codes are distinct (disjoint refs),
    // so both give the same three codes.
    const unsigned long long withPeak = __ballot(top.a != 0ull);
    if (__popcll(withPeak) <= 8) {
        Top3               wt{0ull, 0ull, 0ull};
        unsigned long long m = withPeak;
        while (m) {
            const int l = __ffsll((long long)m) - 1;
            m &= m - 1ull;
            wt = top3_add(top3_add(top3_add(wt, __shfl(top.a, l, 64)), __shfl(top.b, l, 64)),
                          __shfl(top.c, l, 64));
        }
        top = wt;
    } else {
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long u0 = __shfl_xor(top.a, o, 64), u1 = __shfl_xor(top.b, o, 64),
                                     u2 = __shfl_xor(top.c, o, 64);
            top = top3_add(top3_add(top3_add(top, u0), u1), u2);
        }
    }
    // every lane now holds the wave's top-3
    if ((tid & 63) == 0) {
        sTop[(tid >> 6) * 3 + 0] = top.a;
        sTop[(tid >> 6) * 3 + 1] = top.b;
        sTop[(tid >> 6) * 3 + 2] = top.c;
    }
    __syncthreads();
    if (tid == 0) {
        Top3 best{0ull, 0ull, 0ull};
        for (int w = 0; w < (BLOCK / 64) * 3; w++)
            best = top3_add(best, sTop[w]);
        candOut[0] = best.a;
        candOut[1] = best.b;
        candOut[2] = best.c;
    }
}

// ---- fused kernel: pairs and votes in one workgroup per key point (the
// path for key points whose hit list overflows the split path's capacity,
// and a tuning choice on its own)
template <int BLOCK, bool MULTI>
__global__ void __launch_bounds__(BLOCK) vote_kernel(VoteArgs a) {
    constexpr int NW   = BLOCK / 64;
    constexpr int QCAP = 128;  // >= 63 queued + 64 new hits
    extern __shared__ uint32_t acc[];
    __shared__ int      sCellStart[VOTE_MAX_CELLS];
    __shared__ int      sCellPre[VOTE_MAX_CELLS + 1];
    __shared__ uint32_t qStart[NW][QCAP];
    __shared__ uint32_t qLen[NW][QCAP];
    __shared__ float    qAlpha[NW][QCAP];
    __shared__ int      qPre[NW][QCAP];
    __shared__ int      pQueue[NW][QCAP];  // sorted-slot ids of queued pairs
    __shared__ int      sChunkCell[VOTE_MAX_CHUNKS];  // cell holding candidate 64 t
    __shared__ int      sQn[NW];
    __shared__ int      sScan[NW + 1];
    __shared__ int      sNext;
    __shared__ unsigned long long sTop[NW * 3];
    __shared__ int      sRed[NW];

    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int       tile;
    const int k = block_keypoint(a, a.nTiles, tile);
    if (k < 0)
        return;
    const int r0     = tile * a.tileRefs;
    const int r1     = min(r0 + a.tileRefs, a.refNum);
    const int nrTile = r1 - r0;
    unsigned long long *candOut = a.cand + ((size_t)k * a.nTiles + tile) * 3;

    const int   kpId = a.kp[k];
    const float qx = a.kpX[k], qy = a.kpY[k], qz = a.kpZ[k];
    const float nx1 = a.kpNX[k], ny1 = a.kpNY[k], nz1 = a.kpNZ[k];
    if (ppfc::is_nan(nx1) || ppfc::is_nan(ny1) || ppfc::is_nan(nz1)) {
        if (tid < 3)
            candOut[tid] = 0ull;
        return;
    }
    float rt[12];
    ppfc::frame(qx, qy, qz, nx1, ny1, nz1, rt);

    for (int i = tid; i < nrTile * a.angleNum; i += BLOCK)
        acc[i] = 0u;
    const GridView &g = MULTI ? a.grids[a.kpScene[k]] : a.g;
    int             nCells;
    const int total = vote_cells<BLOCK>(a, g, qx, qy, qz, sCellStart, sCellPre, sChunkCell, sScan,
                                        nCells);
    if (tid == 0)
        sNext = 0;
    __syncthreads();

    // ---- pairs + votes.  Each wave takes chunks of 64 candidates from a
    // shared counter, queues its hits and votes them once 64 are queued, with
    // no workgroup barrier until all candidates are done.  The neighbour
    // count of the reference's skip rule (radiusSearch incl. self <
    // voteThreshold, ppf.cpp:273-276) is taken in the same pass; a key point
    // that fails it has its accumulator discarded below.
    const int nChunks = (total + 63) >> 6;
    unsigned long long nPairs = 0, nHits = 0, nVotes = 0;
    int inSphere = 0;
    int pn = 0, qn = 0;  // wave-uniform lengths of the pair and hit queues
    int *pq = pQueue[wv];
    // features, key lookup and hit queueing of the first cnt (<= 64) queued
    // pairs, one per lane
    auto pair_stage = [&](int cnt) {
        int      hit    = 0;
        uint32_t bStart = 0, bLen = 0;
        float    alphaS = 0.0f;
        if (lane < cnt) {
            const int s  = pq[lane];
            float     px = g.x[s], py = g.y[s], pz = g.z[s];
            alphaS       = ppfc::alpha_of(rt, px, py, pz);
            uint32_t key = ppfc::ppf_key(qx, qy, qz, nx1, ny1, nz1, px, py, pz, g.nx[s], g.ny[s],
                                         g.nz[s], a.angleStep, a.distStep, false);
            uint32_t bucket;
            if (table_find(a.slots, a.tmask, key, bStart, bLen, bucket) && !ppfc::is_nan(alphaS))
                hit = 1;
        }
        const unsigned long long m = __ballot(hit);
        if (hit) {
            const int pos   = qn + __popcll(m & ((1ull << lane) - 1ull));
            qStart[wv][pos] = bStart;
            qLen[wv][pos]   = bLen;
            qAlpha[wv][pos] = alphaS;
        }
        nHits += hit;
        qn += __popcll(m);
        if (qn >= 64) {
            wave_sync();
            wave_vote<QCAP>(a, acc, r0, r1, qStart[wv], qLen[wv], qAlpha[wv], qPre[wv], qn,
                            nVotes);
            qn = 0;
        }
    };
    for (;;) {
        int t = 0;
        if (lane == 0)
            t = atomicAdd(&sNext, 1);
        t               = __shfl(t, 0, 64);
        const bool more = t < nChunks;
        const int  c    = t * 64 + lane;
        int        pc = 0, s = 0;
        if (more && c < total) {
            s               = candidate_slot(t, c, nCells, sCellStart, sCellPre, sChunkCell);
            const float px  = g.x[s], py = g.y[s], pz = g.z[s];
            const int   pid = g.id[s];
            bool        in  = ppfc::dist2(qx, qy, qz, px, py, pz) < a.D2;
            inSphere += in ? 1 : 0;
            if (in && pid != kpId) {
                pc = 1;
                nPairs++;
            }
        }
        // compact the pairs so the feature stage runs on full waves
        const unsigned long long m = __ballot(pc);
        if (pc)
            pq[pn + __popcll(m & ((1ull << lane) - 1ull))] = s;
        pn += __popcll(m);
        if (pn >= 64) {
            wave_sync();
            pair_stage(64);
            const int rest = pn - 64;
            const int v    = lane < rest ? pq[64 + lane] : 0;
            wave_sync();
            if (lane < rest)
                pq[lane] = v;
            pn = rest;
        }
        if (!more)
            break;
    }
    if (pn > 0) {
        wave_sync();
        pair_stage(pn);
    }
    // every wave holds < 64 queued hits: vote them all with the whole
    // workgroup, so no wave's last queue delays the others
    if (lane == 0)
        sQn[wv] = qn;
    __syncthreads();
    {
        const int w = tid >> 6, i = tid & 63;  // entry i of wave w's queue (BLOCK = 64 NW)
        const int L = i < sQn[w] ? (int)qLen[w][i] : 0;
        int       totSeg;
        const int pre = block_excl_scan<BLOCK>((L + SEG - 1) / SEG, sScan, totSeg);
        int      *flat = &pQueue[0][0];  // free now: segment prefix of entry 64 w + i
        flat[tid]      = pre;
        __syncthreads();
        for (int t = tid; t < totSeg; t += BLOCK) {
            int lo = 0, hi = BLOCK - 1;  // last entry whose prefix <= t (entries hold >= 1 segment)
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if (flat[mid] <= t)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const int      ew = lo >> 6, ei = lo & 63;
            const uint32_t st = qStart[ew][ei], L = qLen[ew][ei];
            SegRun         gr;
            gr.e0        = st + (uint32_t)(t - flat[lo]) * SEG;
            gr.vlo       = 0;
            gr.vhi       = (int)min((uint32_t)SEG, st + L - gr.e0);
            gr.tailStart = st + L - (L & 7u);
            gr.offBody   = ppfc::bin_offset_body(qAlpha[ew][ei]);
            gr.offTail   = ppfc::bin_offset_tail(qAlpha[ew][ei]);
            const int r0A = r0 * a.angleNum, r1A = r1 * a.angleNum;
            nVotes += (unsigned long long)(a.nTiles > 1 ? vote_run<true>(a, acc, r0A, r1A, gr)
                                                        : vote_run<false>(a, acc, r0A, r1A, gr));
        }
    }
    __syncthreads();

    for (int o = 32; o > 0; o >>= 1)
        inSphere += __shfl_xor(inSphere, o, 64);
    if ((tid & 63) == 0)
        sRed[tid >> 6] = inSphere;
    __syncthreads();
    int searched = 0;
    for (int w = 0; w < NW; w++)
        searched += sRed[w];
    const bool voted = !((float)searched < a.voteThreshold);

    if (k < a.dumpCount && a.accDump) {
        for (int i = tid; i < nrTile * a.angleNum; i += BLOCK)
            a.accDump[((size_t)k * a.refNum + r0) * a.angleNum + i] = (int32_t)acc[i];
    }
    if (!voted) {
        if (tid < 3)
            candOut[tid] = 0ull;
        return;
    }
    if (a.stats && tile == 0 && a.countPairs) {
        for (int o = 32; o > 0; o >>= 1) {
            nPairs += __shfl_xor(nPairs, o, 64);
            nHits += __shfl_xor(nHits, o, 64);
        }
        if ((tid & 63) == 0) {
            atomicAdd(a.stats + 0, nPairs);
            atomicAdd(a.stats + 1, nHits);
        }
    }
    if (a.stats && tile == 0 && tid == 0)
        atomicAdd(a.stats + 3, 1ull);
    if (a.stats) {
        for (int o = 32; o > 0; o >>= 1)
            nVotes += __shfl_xor(nVotes, o, 64);
        if ((tid & 63) == 0)
            atomicAdd(a.stats + 2, nVotes);
    }
    peaks_top3<BLOCK>(a, acc, nrTile, r0, sTop, candOut);
}

// ---- split path, stage 1: the pairs of every key point, its neighbour
// count and its key hits (bucket start, length, scene alpha), written to
// HBM.  No accumulator, so several small workgroups share a CU and hide
// each other's latency.  A key point with more than hitCap hits gets
// nHit = -1 and is voted by the fused kernel instead.
//
// The features of a pair are computed in three stages, each on full waves
// drawn from a per-wave LDS queue: (A) the distance and the normal angle,
// checked against the model's (q4) and (q3, q4) pre-filters; (B) the angle
// of the first normal to the pair direction, checked against (q1, q3, q4);
// (C) the last angle, the key and the table probe, and the scene alpha for
// hits only.  Most pairs leave after stage A or B, so the expensive tail of
// the feature sequence runs on a fraction of them.  The features are the
// canonical ones (ppfc::ppf_key's operations, in pieces).
constexpr int PAIR_BLOCK = 256;
constexpr int HITS_INFLIGHT      = 2;  // segments whose loads a voting lane issues together
constexpr int VOTE_LANES_PER_SEG = 2;  // lanes sharing one 8-entry segment in packed voting

// quantized feature values packed 10 bits each; 1023 stands for the NaN
// code 0x80000000 (finite values stay below 1000 when the pre-filters are on)
__device__ __forceinline__ uint32_t qpack(uint32_t q) { return q > 1022u ? 1023u : q; }
__device__ __forceinline__ uint32_t qunpack(uint32_t c) { return c == 1023u ? 0x80000000u : c; }
__device__ __forceinline__ bool filter_bit(const uint32_t *w, uint32_t i) {
    return (w[i >> 5] >> (i & 31u)) & 1u;
}

template <bool MULTI>
__global__ void __launch_bounds__(PAIR_BLOCK) pair_kernel(VoteArgs a) {
    constexpr int NW   = PAIR_BLOCK / 64;
    constexpr int QCAP = 128;
    __shared__ int sCellStart[VOTE_MAX_CELLS];
    __shared__ int sCellPre[VOTE_MAX_CELLS + 1];
    __shared__ int sChunkCell[VOTE_MAX_CHUNKS];
    __shared__ int pQueue[NW][QCAP];  // in-sphere pairs: candidate slot
    // stage B queue: slot, |d|, packed (q3, q4)
    __shared__ int      bSlot[NW][QCAP];
    __shared__ float    bDn[NW][QCAP];
    __shared__ uint32_t bCode[NW][QCAP];
    // stage C queue: slot, packed (q1, q3, q4), unit direction
    __shared__ int      cSlot[NW][QCAP];
    __shared__ uint32_t cCode[NW][QCAP];
    __shared__ float    cU[NW][3][QCAP];
    __shared__ int sScan[NW + 1];
    __shared__ int sNext, sHits, sIn;
    __shared__ unsigned long long sPairs[NW];

    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int       tile;
    const int k = block_keypoint(a, 1, tile);
    if (k < 0)
        return;
    const int   kpId = a.kp[k];
    const float qx = a.kpX[k], qy = a.kpY[k], qz = a.kpZ[k];
    const float nx1 = a.kpNX[k], ny1 = a.kpNY[k], nz1 = a.kpNZ[k];
    if (ppfc::is_nan(nx1) || ppfc::is_nan(ny1) || ppfc::is_nan(nz1)) {
        if (tid == 0) {
            a.nHit[k] = 0;
            a.nbr[k]  = 0;
        }
        return;
    }
    float rt[12];
    ppfc::frame(qx, qy, qz, nx1, ny1, nz1, rt);
    const GridView &g = MULTI ? a.grids[a.kpScene[k]] : a.g;
    int             nCells;
    const int total = vote_cells<PAIR_BLOCK>(a, g, qx, qy, qz, sCellStart, sCellPre, sChunkCell,
                                             sScan, nCells);
    if (tid == 0) {
        sNext = 0;
        sHits = 0;
        sIn   = 0;
    }
    __syncthreads();

    const int          nChunks = (total + 63) >> 6;
    const size_t       base    = (size_t)k * a.hitCap;
    const unsigned long long below = (1ull << lane) - 1ull;
    const bool         filt    = a.filter != nullptr;
    const uint32_t     fA = (uint32_t)a.fA, fD = (uint32_t)a.fD;
    const uint32_t    *f4 = a.filter, *f34 = f4 + (fD + 31) / 32,
                   *f134 = f34 + (fA * fD + 31) / 32;
    auto aidx = [&](uint32_t q) { return q < fA - 1 ? q : fA - 1; };
    auto didx = [&](uint32_t q) { return q < fD - 1 ? q : fD - 1; };
    unsigned long long nPairs  = 0;
    int                inSphere = 0, pn = 0, bn = 0, cn = 0;
    int               *pq = pQueue[wv];

    // (C) last angle, key, probe; alpha of the hits; hits to HBM
    auto stageC = [&](int cnt) {
        int      hit    = 0;
        uint32_t bStart = 0, bLen = 0, bucket = 0;
        float    alphaS = 0.0f;
        if (lane < cnt) {
            const int      s    = cSlot[wv][lane];
            const uint32_t code = cCode[wv][lane];
            const float4   N  = g.n4[s];
            const float    f2 = ppfc::angle_a2b(cU[wv][0][lane], cU[wv][1][lane], cU[wv][2][lane],
                                                N.x, N.y, N.z);
            const uint32_t key = ppfc::murmur_key4(
                qunpack(code >> 20), ppfc::quant(f2, a.angleStep, false),
                qunpack((code >> 10) & 1023u), qunpack(code & 1023u));
            if (table_find(a.slots, a.tmask, key, bStart, bLen, bucket)) {
                const float4 P = g.p4[s];
                alphaS         = ppfc::alpha_of(rt, P.x, P.y, P.z);
                hit    = ppfc::is_nan(alphaS) ? 0 : 1;
            }
        }
        const unsigned long long m  = __ballot(hit);
        const int                nh = __popcll(m);
        if (nh == 0)
            return;
        int at = 0;
        if (lane == 0)
            at = atomicAdd(&sHits, nh);
        at = __shfl(at, 0, 64);
        if (hit) {
            const int pos = at + __popcll(m & below);
            if (pos < a.hitCap) {
                a.hitStart[base + pos]  = bStart;
                a.hitLen[base + pos]    = bLen;
                a.hitAlpha[base + pos]  = alphaS;
                a.hitBucket[base + pos] = bucket;
            }
        }
    };
    // 64 entries of the C queue voted, the rest moved to its front
    auto drainC = [&](bool all) {
        while (cn >= 64 || (all && cn > 0)) {
            wave_sync();
            stageC(min(cn, 64));
            const int rest = max(cn - 64, 0);
            int       s0 = 0;
            uint32_t  c0 = 0;
            float     u0 = 0.0f, u1 = 0.0f, u2 = 0.0f;
            if (lane < rest) {
                s0 = cSlot[wv][64 + lane];
                c0 = cCode[wv][64 + lane];
                u0 = cU[wv][0][64 + lane];
                u1 = cU[wv][1][64 + lane];
                u2 = cU[wv][2][64 + lane];
            }
            wave_sync();
            if (lane < rest) {
                cSlot[wv][lane]    = s0;
                cCode[wv][lane]    = c0;
                cU[wv][0][lane]    = u0;
                cU[wv][1][lane]    = u1;
                cU[wv][2][lane]    = u2;
            }
            cn = rest;
        }
    };
    // (B) angle of the first normal to the pair direction
    auto stageB = [&](int cnt) {
        bool     pass = false;
        int      s    = 0;
        uint32_t code = 0;
        float    ux = 0.0f, uy = 0.0f, uz = 0.0f;
        if (lane < cnt) {
            s              = bSlot[wv][lane];
            const float dn = bDn[wv][lane];
            code           = bCode[wv][lane];
            const float4 P = g.p4[s];
            ux             = (P.x - qx) / dn;
            uy             = (P.y - qy) / dn;
            uz             = (P.z - qz) / dn;
            const float    f1 = ppfc::angle_a2b(ux, uy, uz, nx1, ny1, nz1);
            const uint32_t q1 = qpack(ppfc::quant(f1, a.angleStep, false));
            code |= q1 << 20;
            pass = !filt || filter_bit(f134, (aidx(q1) * fA + aidx((code >> 10) & 1023u)) * fD +
                                                 didx(code & 1023u));
        }
        const unsigned long long m = __ballot(pass);
        if (pass) {
            const int pos      = cn + __popcll(m & below);
            cSlot[wv][pos]     = s;
            cCode[wv][pos]     = code;
            cU[wv][0][pos]     = ux;
            cU[wv][1][pos]     = uy;
            cU[wv][2][pos]     = uz;
        }
        cn += __popcll(m);
        drainC(false);
    };
    auto drainB = [&](bool all) {
        while (bn >= 64 || (all && bn > 0)) {
            wave_sync();
            stageB(min(bn, 64));
            const int rest = max(bn - 64, 0);
            int       s0 = 0;
            float     d0 = 0.0f;
            uint32_t  c0 = 0;
            if (lane < rest) {
                s0 = bSlot[wv][64 + lane];
                d0 = bDn[wv][64 + lane];
                c0 = bCode[wv][64 + lane];
            }
            wave_sync();
            if (lane < rest) {
                bSlot[wv][lane] = s0;
                bDn[wv][lane]   = d0;
                bCode[wv][lane] = c0;
            }
            bn = rest;
        }
    };
    // (A) distance and normal angle
    auto stageA = [&](int cnt) {
        bool     pass = false;
        int      s    = 0;
        float    dn   = 0.0f;
        uint32_t code = 0;
        if (lane < cnt) {
            s              = pq[lane];
            const float4 P  = g.p4[s];
            const float  dx = P.x - qx, dy = P.y - qy, dz = P.z - qz;
            dn             = sqrtf(dx * dx + dy * dy + dz * dz);
            const uint32_t q4 = qpack(ppfc::quant(dn, a.distStep, false));
            const float4   N  = g.n4[s];
            const float    f3 = ppfc::angle_a2b(nx1, ny1, nz1, N.x, N.y, N.z);
            const uint32_t q3 = qpack(ppfc::quant(f3, a.angleStep, false));
            code              = (q3 << 10) | q4;
            pass = !filt || (filter_bit(f4, didx(q4)) && filter_bit(f34, aidx(q3) * fD + didx(q4)));
        }
        const unsigned long long m = __ballot(pass);
        if (pass) {
            const int pos   = bn + __popcll(m & below);
            bSlot[wv][pos]  = s;
            bDn[wv][pos]    = dn;
            bCode[wv][pos]  = code;
        }
        bn += __popcll(m);
        drainB(false);
    };
    for (;;) {
        int t = 0;
        if (lane == 0)
            t = atomicAdd(&sNext, 1);
        t               = __shfl(t, 0, 64);
        const bool more = t < nChunks;
        const int  c    = t * 64 + lane;
        int        pc = 0, s = 0;
        if (more && c < total) {
            s               = candidate_slot(t, c, nCells, sCellStart, sCellPre, sChunkCell);
            const float4 P   = g.p4[s];  // {x, y, z, id bits}: one 16-byte load
            const float  px  = P.x, py = P.y, pz = P.z;
            const int    pid = __float_as_int(P.w);
            bool        in  = ppfc::dist2(qx, qy, qz, px, py, pz) < a.D2;
            inSphere += in ? 1 : 0;
            if (in && pid != kpId) {
                pc = 1;
                nPairs++;
            }
        }
        const unsigned long long m = __ballot(pc);
        if (pc)
            pq[pn + __popcll(m & below)] = s;
        pn += __popcll(m);
        if (pn >= 64) {
            wave_sync();
            stageA(64);
            const int rest = pn - 64;
            const int v    = lane < rest ? pq[64 + lane] : 0;
            wave_sync();
            if (lane < rest)
                pq[lane] = v;
            pn = rest;
        }
        if (!more)
            break;
    }
    if (pn > 0) {
        wave_sync();
        stageA(pn);
    }
    drainB(true);
    drainC(true);
    for (int o = 32; o > 0; o >>= 1) {
        inSphere += __shfl_xor(inSphere, o, 64);
        nPairs += __shfl_xor(nPairs, o, 64);
    }
    if (lane == 0) {
        atomicAdd(&sIn, inSphere);
        sPairs[wv] = nPairs;
    }
    __syncthreads();
    if (tid == 0) {
        const int nh = sHits;
        a.nbr[k]     = sIn;
        a.nHit[k]    = nh > a.hitCap ? -1 : nh;
        if (nh > a.hitCap)
            a.overflow[atomicAdd(a.nOverflow, 1)] = k;
        // pairs and hits of voted key points, as the fused kernel counts them
        if (a.stats && !((float)sIn < a.voteThreshold)) {
            unsigned long long np = 0;
            for (int w = 0; w < NW; w++)
                np += sPairs[w];
            atomicAdd(a.stats + 0, np);
            atomicAdd(a.stats + 1, (unsigned long long)nh);
        }
    }
}

// ---- split path, stage 2: vote the hits of one key point (and tile) into
// the LDS accumulator, all lanes striped over the buckets' segments, then
// the peaks.
// The kernel's LDS is one dynamic block: the accumulator at offset 0 (so
// the voting loop's LDS addresses need no base added), then its own arrays
// (HitsLds).  The current batch of hits: with one tile (PACKED) each bucket
// is voted whole and a hit needs only {start, length, scene alpha}, and a
// batch holds BLOCK / 2 hits (2.2 KB at 256 threads, which leaves room for
// three workgroups per CU beside the 50 KB accumulator of an 832-point
// model; LDS is allocated in 512-byte granules); with reference tiles
// {lo, hi, tail start, body offset, tail offset} of the tile's slice, BLOCK
// hits a batch.
__host__ __device__ inline int hits_acc_words(int tileRefs, int angleNum, bool packed) {
    const int cells = (tileRefs + 1) * angleNum;  // + the spare row
    const int w     = packed ? (cells + 1) / 2 : cells;
    return (w + 3) & ~3;  // 16-byte alignment: cleared as uint4, 8-byte words follow
}
template <int BLOCK, bool PACKED> struct HitsLds {
    static constexpr int HB = PACKED ? BLOCK / 2 : BLOCK;
    static constexpr int HW = PACKED ? 3 : 5;
    // words after the accumulator: sTop (u64), sOvf, sScan, sPre, sHit
    static constexpr int words = 2 * (BLOCK / 64) * 3 + 1 + (BLOCK / 64 + 1) + HB + HW * HB;
};

template <int BLOCK, bool PACKED>
__global__ void __launch_bounds__(BLOCK) vote_hits_kernel(VoteArgs a) {
    extern __shared__ uint32_t acc[];
    using LL                 = HitsLds<BLOCK, PACKED>;
    constexpr int HB         = LL::HB;
    constexpr int HW         = LL::HW;
    uint32_t     *extra      = acc + hits_acc_words(a.tileRefs, a.angleNum, PACKED);
    unsigned long long *sTop = reinterpret_cast<unsigned long long *>(extra);
    uint32_t           &sOvf = extra[2 * (BLOCK / 64) * 3];
    int *sScan               = reinterpret_cast<int *>(extra + 2 * (BLOCK / 64) * 3 + 1);
    int *sPre                = sScan + (BLOCK / 64 + 1);
    uint32_t *sHit           = reinterpret_cast<uint32_t *>(sPre + HB);

    const int tid = threadIdx.x;
    int       tile;
    const int k = block_keypoint(a, a.nTiles, tile);
    if (k < 0)
        return;
    const int r0     = tile * a.tileRefs;
    const int r1     = min(r0 + a.tileRefs, a.refNum);
    const int nrTile = r1 - r0;
    unsigned long long *candOut = a.cand + ((size_t)k * a.nTiles + tile) * 3;
    const int           n       = a.nHit[k];
    // skip rule (ppf.cpp:273-276); an overflowed key point is voted by the
    // fused kernel, which rewrites its candidates
    const bool voted = !((float)a.nbr[k] < a.voteThreshold);
    if (n < 0 || (!voted && !(k < a.dumpCount && a.accDump))) {
        if (tid < 3)
            candOut[tid] = 0ull;
        return;
    }
    const int realCells = nrTile * a.angleNum;
    // packed: the spare row too, whose returned words take part in the
    // overflow test; the whole accumulator block, 16 bytes a store
    if (PACKED) {
        uint4    *a4 = reinterpret_cast<uint4 *>(acc);
        const int n4 = hits_acc_words(a.tileRefs, a.angleNum, true) >> 2;
        for (int i = tid; i < n4; i += BLOCK)
            a4[i] = make_uint4(0u, 0u, 0u, 0u);
    } else {
        for (int i = tid; i < realCells; i += BLOCK)
            acc[i] = 0u;
    }
    if (tid == 0)
        sOvf = 0u;
    uint32_t           ovf    = 0u;
    unsigned long long nVotes = 0;
    const size_t       base   = (size_t)k * a.hitCap;
    // the voting loop, with returning adds (RET) or without: a key point
    // whose hits cannot lift a 16-bit counter past 0xFFFF (n * maxMult <=
    // 0xFFFF) needs no overflow test
    auto voteAll = [&](auto retTag) {
        constexpr bool RET = decltype(retTag)::value;
        for (int h0 = 0; h0 < n; h0 += HB) {
            // this tile's entries of each hit's bucket: the whole bucket, or with
            // ref-sorted buckets only the slice holding the tile's references
            const int i   = h0 + tid;
            int       cnt = 0;
            if (tid < HB && i < n) {
                const uint32_t st = a.hitStart[base + i], L = a.hitLen[base + i];
                uint32_t       lo = st, hi = st + L;
                if (a.splits) {
                    const uint32_t *sp = a.splits + (size_t)a.hitBucket[base + i] * (a.nTiles + 1);
                    lo                 = sp[tile];
                    hi                 = sp[tile + 1];
                }
                // aligned segments [lo & ~7, roundup8(hi)) with entries outside
                // [lo, hi) masked
                const uint32_t al = lo & ~7u, ah = (hi + 7u) & ~7u;
                cnt            = hi > lo ? (int)((ah - al) / SEG) : 0;
                const float aS = a.hitAlpha[base + i];
                if (PACKED) {
                    sHit[tid]             = st;
                    sHit[HB + tid]     = L;
                    sHit[2 * HB + tid] = __float_as_uint(aS);
                } else {
                    sHit[tid]             = lo;
                    sHit[HB + tid]     = hi;
                    sHit[2 * HB + tid] = st + L - (L & 7u);
                    sHit[3 * HB + tid] = __float_as_uint(ppfc::bin_offset_body(aS));
                    sHit[4 * HB + tid] = __float_as_uint(ppfc::bin_offset_tail(aS));
                }
            }
            int       totSeg;
            const int pre = block_excl_scan<BLOCK>(cnt, sScan, totSeg);
            if (tid < HB)
                sPre[tid] = pre;
            __syncthreads();
            // each thread votes a contiguous run of segments (SEG entries of one
            // bucket each; different lanes, different buckets, so the LDS atomics
            // of a wave scatter over the accumulator): one search for the run's
            // first hit, then a walk (hits with no segment are skipped); the
            // loads of HITS_INFLIGHT segments are issued before any of them is
            // voted (the accumulator leaves room for two workgroups of waves per
            // CU at most, so each lane keeps several loads in flight instead)
            const int nb  = min(HB, n - h0);
            // packed: LPS lanes share a segment, each loading and voting EPL
            // of its entries, so one load instruction of a wave touches
            // 64 / LPS segments (64-byte lines) instead of 64: the L1's tag
            // lookups, not the ALUs, bound this loop
            constexpr int LPS = PACKED ? VOTE_LANES_PER_SEG : 1;
            constexpr int EPL = SEG / LPS;
            const int     grp = tid / LPS, part = tid % LPS;
            const int spt = (totSeg + BLOCK / LPS - 1) / (BLOCK / LPS);
            const int t0  = grp * spt, t1 = min(t0 + spt, totSeg);
            const int r0A = r0 * a.angleNum, r1A = r1 * a.angleNum;
            int       nv  = 0;
            if (t0 < t1) {
                int lo = 0, hi = nb - 1;  // last hit whose prefix <= t0
                while (lo < hi) {
                    int mid = (lo + hi + 1) >> 1;
                    if (sPre[mid] <= t0)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                // the current hit's descriptor stays in registers; LDS is read
                // again only when the run crosses into the next hit
                int      h       = lo;
                int      hPre    = sPre[h];
                int      nextPre = h + 1 < nb ? sPre[h + 1] : 0x7fffffff;
                uint32_t hLo, hHi, hTail;
                float    hOffB, hOffT;
                auto     loadHit = [&](int q) {
                    if (PACKED) {
                        const uint32_t st = sHit[q], L = sHit[HB + q];
                        const float    aS = __uint_as_float(sHit[2 * HB + q]);
                        hLo                = st;
                        hHi                = st + L;
                        hTail              = st + L - (L & 7u);
                        hOffB              = ppfc::bin_offset_body(aS);
                        hOffT              = ppfc::bin_offset_tail(aS);
                    } else {
                        hLo   = sHit[q];
                        hHi   = sHit[HB + q];
                        hTail = sHit[2 * HB + q];
                        hOffB = __uint_as_float(sHit[3 * HB + q]);
                        hOffT = __uint_as_float(sHit[4 * HB + q]);
                    }
                };
                loadHit(h);
                for (int t = t0; t < t1; t += HITS_INFLIGHT) {
                    SegRun g[HITS_INFLIGHT];
                    uint2  ent[HITS_INFLIGHT][EPL];
#pragma unroll
                    for (int j = 0; j < HITS_INFLIGHT; j++) {
                        const int tq = t + j;
                        g[j].vlo     = 0;
                        g[j].vhi     = 0;
                        if (tq < t1) {
                            if (tq >= nextPre) {
                                // hits without segments share their prefix with
                                // the next hit and are stepped over
                                do {
                                    h++;
                                    hPre    = nextPre;
                                    nextPre = h + 1 < nb ? sPre[h + 1] : 0x7fffffff;
                                } while (tq >= nextPre);
                                loadHit(h);
                            }
                            // entry offsets stay below 2^31: 32-bit differences
                            g[j].e0        = (hLo & ~7u) + (uint32_t)(tq - hPre) * SEG;
                            g[j].vlo       = max((int)(hLo - g[j].e0), 0);
                            g[j].vhi       = min((int)(hHi - g[j].e0), SEG);
                            g[j].tailStart = hTail;
                            g[j].offBody   = hOffB;
                            g[j].offTail   = hOffT;
                            load_entries<EPL>(a, g[j].e0 + (uint32_t)(part * EPL), ent[j]);
                        }
                    }
#pragma unroll
                    for (int j = 0; j < HITS_INFLIGHT; j++)
                        if (t + j < t1) {
                            if constexpr (PACKED) {
                                const int c = vote_run_packed<RET, EPL>(a, g[j], ent[j], ovf);
                                nv += part == 0 ? c : 0;
                            } else
                                nv += a.nTiles > 1
                                          ? vote_run_loaded<true>(a, acc, r0A, r1A, g[j], ent[j])
                                          : vote_run_loaded<false>(a, acc, r0A, r1A, g[j], ent[j]);
                        }
                }
            }
            nVotes += (unsigned long long)nv;
            __syncthreads();
        }
    };
    const bool ret = PACKED && (unsigned long long)n * a.maxMult > 0xFFFFull;
    if (ret)
        voteAll(std::true_type{});
    else
        voteAll(std::false_type{});
    if (PACKED) {
        // a 16-bit counter overflowed: the fused kernel votes this key point
        // again (and writes its candidates, dump and counts)
        if (ovf & 0x80008000u)
            atomicOr(&sOvf, 1u);
        __syncthreads();
        if (sOvf) {
            if (tid == 0)
                a.overflow[atomicAdd(a.nOverflow, 1)] = k;
            return;
        }
    }
    if (k < a.dumpCount && a.accDump) {
        for (int i = tid; i < nrTile * a.angleNum; i += BLOCK)
            a.accDump[((size_t)k * a.refNum + r0) * a.angleNum + i] =
                (int32_t)acc_get<PACKED>(acc, i);
    }
    if (!voted) {
        if (tid < 3)
            candOut[tid] = 0ull;
        return;
    }
    if (a.stats) {
        for (int o = 32; o > 0; o >>= 1)
            nVotes += __shfl_xor(nVotes, o, 64);
        if ((tid & 63) == 0)
            atomicAdd(a.stats + 2, nVotes);
        if (tid == 0 && tile == 0)
            atomicAdd(a.stats + 3, 1ull);
    }
    peaks_top3<BLOCK, PACKED>(a, acc, nrTile, r0, sTop, candOut);
}

// Merge tiles, apply the 0.95 filter (ppf.cpp:393-397) and build the poses
// T = inv(Ts) * Rx(alpha) * Tm (ppf.cpp:399-405).  One thread per keypoint.
__global__ void poses_kernel(const unsigned long long *cand, int nk, int nTiles, const float *KX,
                             const float *KY, const float *KZ, const float *KNX, const float *KNY,
                             const float *KNZ, const float *frames, const float *cosSin,
                             int angleNum, PoseRec *slots, int *counts) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk)
        return;
    Top3 best{0ull, 0ull, 0ull};
    for (int t = 0; t < nTiles * 3; t++)
        best = top3_add(best, cand[(size_t)k * nTiles * 3 + t]);
    const unsigned long long b0 = best.a;
    unsigned long long bs[3] = {best.a, best.b, best.c};
    int n = 0;
    if (b0) {
        float bestVote = (float)(uint32_t)(b0 >> 32);
        float thr      = (float)((double)bestVote * 0.95);
        float Ts[12];
        ppfc::frame(KX[k], KY[k], KZ[k], KNX[k], KNY[k], KNZ[k], Ts);
        for (int r = 0; r < 3; r++) {
            unsigned long long c = bs[r];
            if (!c)
                break;
            float vote = (float)(uint32_t)(c >> 32);
            if (vote < thr)
                continue;
            int ref = (int)(0xFFFFFu - (uint32_t)((c >> 11) & 0x1FFFFFu));
            int bin = (int)(c & 0x7FFu);
            PoseRec &P = slots[(size_t)k * 3 + n];
            float Tm[12];
            for (int x = 0; x < 12; x++)
                Tm[x] = frames[(size_t)ref * 12 + x];
            ppfc::compose_pose(Ts, Tm, cosSin[2 * bin], cosSin[2 * bin + 1], P.pose);
            float R[9] = {P.pose[0], P.pose[1], P.pose[2], P.pose[4], P.pose[5],
                          P.pose[6], P.pose[8], P.pose[9], P.pose[10]};
            ppfc::quat_from_rot(R, P.q);
            P.angle    = ppfc::angle_of_quat(P.q);
            P.votes    = vote;
            P.keypoint = k;
            P.rank     = n;
            P.refId    = ref;
            P.angleId  = bin;
            n++;
        }
    }
    counts[k] = n;
}

__global__ void compact_poses_kernel(const PoseRec *slots, const int *counts, const int *offs,
                                     int nk, PoseRec *out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk)
        return;
    for (int r = 0; r < counts[k]; r++)
        out[offs[k] + r] = slots[(size_t)k * 3 + r];
}

template <int BLOCK>
static void launch_vote(const VoteArgs &a, int blocks, size_t lds, hipStream_t s) {
    if (a.grids) {
        PPF_HIP_CHECK(hipFuncSetAttribute((const void *)vote_kernel<BLOCK, true>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        vote_kernel<BLOCK, true><<<blocks, BLOCK, lds, s>>>(a);
    } else {
        PPF_HIP_CHECK(hipFuncSetAttribute((const void *)vote_kernel<BLOCK, false>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        vote_kernel<BLOCK, false><<<blocks, BLOCK, lds, s>>>(a);
    }
}

// split points of every bucket for tiles of tileRefs references: the first
// entry of tile t is the first entry whose ref >= t * tileRefs (refs ascend)
__global__ void bucket_split_kernel(const uint2 *buckets, int K, const uint2 *ventries,
                                    int tileRefs, int angleNum, int nTiles, uint32_t *splits) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= K)
        return;
    const uint2 bk = buckets[b];
    for (int t = 0; t <= nTiles; t++) {
        uint32_t lo = bk.x, hi = bk.x + bk.y;  // first e in [lo, hi) with ref >= t * tileRefs
        const uint32_t r = (uint32_t)t * (uint32_t)tileRefs * (uint32_t)angleNum;
        const uint2   *entries = ventries;
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (entries[mid].x < r)
                lo = mid + 1;
            else
                hi = mid;
        }
        splits[(size_t)b * (nTiles + 1) + t] = lo;
    }
}

// 16-bit packed accumulators for the split path when one tile holds every
// reference (half the LDS, so twice the workgroups per CU)
constexpr bool VOTE_PACKED = true;

template <int BLOCK>
static void launch_vote_hits(const VoteArgs &a, int blocks, bool packed, hipStream_t s) {
    const size_t lds =
        4 * ((size_t)hits_acc_words(a.tileRefs, a.angleNum, packed) +
             (packed ? HitsLds<BLOCK, true>::words : HitsLds<BLOCK, false>::words));
    // the packed voting addresses its accumulator from LDS address 0: the
    // kernel must have no static LDS ahead of its dynamic block
    static const bool noStatic = [] {
        hipFuncAttributes at{};
        PPF_HIP_CHECK(hipFuncGetAttributes(&at, (const void *)vote_hits_kernel<BLOCK, true>));
        return at.sharedSizeBytes == 0;
    }();
    if (packed && !noStatic)
        throw HipError("vote_hits_kernel has static LDS ahead of its accumulator");
    if (packed) {
        PPF_HIP_CHECK(hipFuncSetAttribute((const void *)vote_hits_kernel<BLOCK, true>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        vote_hits_kernel<BLOCK, true><<<blocks, BLOCK, lds, s>>>(a);
    } else {
        PPF_HIP_CHECK(hipFuncSetAttribute((const void *)vote_hits_kernel<BLOCK, false>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        vote_hits_kernel<BLOCK, false><<<blocks, BLOCK, lds, s>>>(a);
    }
}

// key point data (position, normal) gathered into flat arrays
__global__ void kp_gather_kernel(const int *kp, int nk, const float *PX, const float *PY,
                                 const float *PZ, const float *NX, const float *NY, const float *NZ,
                                 float *kx, float *ky, float *kz, float *knx, float *kny,
                                 float *knz) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk)
        return;
    int i  = kp[k];
    kx[k]  = PX[i];
    ky[k]  = PY[i];
    kz[k]  = PZ[i];
    knx[k] = NX[i];
    kny[k] = NY[i];
    knz[k] = NZ[i];
}

__device__ __forceinline__ uint32_t morton_spread10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Visiting order of the key points: sort key = scene index (high bits), then
// the Morton code of the position on a 1024^3 lattice over the scene's vote
// grid.  Workgroups in flight then hold spatially close key points, whose
// vote spheres cover the same scene cells, so the candidates' positions and
// normals are reused from L2 instead of refetched.  The order only permutes
// which workgroup takes which key point; every output stays indexed by key
// point.
__global__ void kp_order_kernel(const float *KX, const float *KY, const float *KZ,
                                const int *kpScene, const GridView *grids, GridView g, int nk,
                                unsigned long long *keys, int *idx) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk)
        return;
    const int       sc  = kpScene ? kpScene[k] : 0;
    const GridView &v   = grids ? grids[sc] : g;
    const float     ext = v.cell * (float)max(v.dx, max(v.dy, v.dz));
    const float     inv = 1024.0f / ext;
    auto q = [&](float x, float o) {
        float f = floorf((x - o) * inv);
        f       = fminf(fmaxf(f, 0.0f), 1023.0f);
        return (uint32_t)f;
    };
    const uint32_t m = morton_spread10(q(KX[k], v.ox)) | (morton_spread10(q(KY[k], v.oy)) << 1) |
                       (morton_spread10(q(KZ[k], v.oz)) << 2);
    keys[k] = ((unsigned long long)sc << 30) | m;
    idx[k]  = k;
}

void gather_keypoints(const int *kp, int nk, const float *PX, const float *PY, const float *PZ,
                      const float *NX, const float *NY, const float *NZ, float *kx, float *ky,
                      float *kz, float *knx, float *kny, float *knz, hipStream_t s) {
    if (nk > 0)
        kp_gather_kernel<<<div_up(nk, 256), 256, 0, s>>>(kp, nk, PX, PY, PZ, NX, NY, NZ, kx, ky,
                                                         kz, knx, kny, knz);
}

namespace {
__global__ void kp_gather_multi_kernel(const KpSceneIn *sc, const int *ko, int S, int K, float *kx,
                                       float *ky, float *kz, float *knx, float *kny, float *knz,
                                       int *kpScene, int *kpId) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= K)
        return;
    const int        w = owner_of(ko, S, t);
    const KpSceneIn &d = sc[w];
    const int        i = d.kp[t - ko[w]];
    kx[t]      = d.PX[i];
    ky[t]      = d.PY[i];
    kz[t]      = d.PZ[i];
    knx[t]     = d.NX[i];
    kny[t]     = d.NY[i];
    knz[t]     = d.NZ[i];
    kpScene[t] = w;
    kpId[t]    = i;
}
}  // namespace

void gather_keypoints_multi(const std::vector<KpSceneIn> &sc, const std::vector<int> &ko,
                            DevBuf &desc, float *kx, float *ky, float *kz, float *knx, float *kny,
                            float *knz, int *kpScene, int *kpId, hipStream_t s) {
    const int S = (int)sc.size();
    if (S == 0 || (int)ko.size() != S + 1)
        throw std::logic_error("gather_keypoints_multi: offsets do not match the scenes");
    const int K = ko[S];
    if (K <= 0)
        return;
    const size_t dBytes = sizeof(KpSceneIn) * S, oBytes = sizeof(int) * (S + 1);
    unsigned char *blob = desc.as<unsigned char>(dBytes + oBytes);
    std::vector<unsigned char> hostBlob(dBytes + oBytes);
    std::memcpy(hostBlob.data(), sc.data(), dBytes);
    std::memcpy(hostBlob.data() + dBytes, ko.data(), oBytes);
    PPF_HIP_CHECK(hipMemcpyAsync(blob, hostBlob.data(), hostBlob.size(), hipMemcpyHostToDevice, s));
    kp_gather_multi_kernel<<<div_up(K, 256), 256, 0, s>>>(
        reinterpret_cast<const KpSceneIn *>(blob), reinterpret_cast<const int *>(blob + dBytes), S,
        K, kx, ky, kz, knx, kny, knz, kpScene, kpId);
    PPF_HIP_CHECK(hipGetLastError());
}

int vote_and_poses(Workspace &ws, const DevModel &m, const GridView &g, const float *PX,
                   const float *PY, const float *PZ, const float *NX, const float *NY,
                   const float *NZ, const int *keypoints, int nk, float voteThreshold,
                   const VoteConfig &cfg, VoteStats *stats, int32_t *accDump, int dumpCount,
                   hipStream_t s) {
    if (nk <= 0 || m.nSampled <= 0)
        return 0;
    VoteKeypoints in;
    in.g       = g;
    in.gridsDev = nullptr;
    in.kpScene = nullptr;
    in.minCell = g.cell;
    in.kx  = ws.kpX.as<float>(nk);
    in.ky  = ws.kpY.as<float>(nk);
    in.kz  = ws.kpZ.as<float>(nk);
    in.knx = ws.kpNX.as<float>(nk);
    in.kny = ws.kpNY.as<float>(nk);
    in.knz = ws.kpNZ.as<float>(nk);
    in.kp  = keypoints;
    in.nk  = nk;
    kp_gather_kernel<<<div_up(nk, 256), 256, 0, s>>>(keypoints, nk, PX, PY, PZ, NX, NY, NZ,
                                                     ws.kpX.get<float>(), ws.kpY.get<float>(),
                                                     ws.kpZ.get<float>(), ws.kpNX.get<float>(),
                                                     ws.kpNY.get<float>(), ws.kpNZ.get<float>());
    return vote_keypoints(ws, m, in, voteThreshold, cfg, stats, accDump, dumpCount, s);
}

int vote_keypoints(Workspace &ws, const DevModel &m, const VoteKeypoints &in, float voteThreshold,
                   const VoteConfig &cfg, VoteStats *stats, int32_t *accDump, int dumpCount,
                   hipStream_t s) {
    const int nk = in.nk;
    if (nk <= 0 || m.nSampled <= 0)
        return 0;
    // the pair kernel reads the scene through the vote grids' point copies
    if (in.g.n > 0 && (!in.g.p4 || !in.g.n4))
        throw std::logic_error("vote grid built without its point copies");
    int tileRefs = std::min(m.nSampled, vote_max_tile_refs(m.angleNum, cfg.block));
    int nTiles   = std::max(div_up(m.nSampled, tileRefs), std::max(cfg.tiles, 1));
    tileRefs     = div_up(m.nSampled, nTiles);
    nTiles       = div_up(m.nSampled, tileRefs);
    float cell   = in.minCell;
    float D      = m.diameter;
    float searchR = D * 1.0001f + 1e-3f * cell;
    // cells per axis over [q - R, q + R]: floor(2R / cell) + 2, plus one for
    // rounding of the two end coordinates
    if ((int)std::floor(2.0 * searchR / cell) + 3 > VOTE_AXIS_CELLS)
        throw std::runtime_error("vote grid cell is smaller than a quarter of the model diameter");
    VoteArgs a;
    a.g       = in.g;
    a.grids   = in.gridsDev;
    a.kpScene = in.kpScene;
    a.kpX  = in.kx;
    a.kpY  = in.ky;
    a.kpZ  = in.kz;
    a.kpNX = in.knx;
    a.kpNY = in.kny;
    a.kpNZ = in.knz;
    a.kp   = in.kp;
    a.nk   = nk;
    a.D    = D;
    a.slots    = m.slots.get<uint4>();
    a.tmask    = m.tmask;
    a.ventries = m.ventries.get<uint2>();
    a.refNum   = m.nSampled;
    a.angleNum = m.angleNum;
    a.tileRefs = tileRefs;
    a.nTiles   = nTiles;
    a.angleStep = m.angleStep;
    a.distStep  = m.distStep;
    a.sMaxId    = (float)(m.angleNum - 1) / ppfc::kTwoPi;
    a.maxBin    = (float)(m.angleNum - 1);
    a.maxMult   = m.maxMult;
    a.D2        = D * D;            // ppf.cpp:191
    a.voteThreshold = voteThreshold;
    a.cand  = ws.cand.as<unsigned long long>((size_t)nk * nTiles * 3);
    a.stats = nullptr;
    if (stats) {
        a.stats = ws.voteStats.as<unsigned long long>(4);
        PPF_HIP_CHECK(hipMemsetAsync(a.stats, 0, 32, s));
    }
    a.accDump    = accDump;
    a.dumpCount  = accDump ? dumpCount : 0;
    a.xcdRemap   = cfg.xcdRemap;
    a.kpSel      = nullptr;
    a.countPairs = 1;
    a.filter     = m.filterA ? m.keyFilter.get<uint32_t>() : nullptr;
    a.fA         = m.filterA;
    a.fD         = m.filterD;
    a.hitCap     = 0;
    a.hitStart = a.hitLen = a.hitBucket = nullptr;
    a.hitAlpha = nullptr;
    a.splits   = nullptr;
    a.nHit = a.nbr = a.overflow = a.nOverflow = nullptr;
    size_t lds      = (size_t)(tileRefs + 1) * m.angleNum * 4;  // + the spare row
    auto   padded   = [&](int n) { return cfg.xcdRemap ? 8 * ((n + 7) / 8) : n; };
    int    blocks   = padded(nk) * nTiles;
    for (auto &e : ws.voteEv)
        if (!e)
            PPF_HIP_CHECK(hipEventCreate(&e));
    auto fused = [&](const VoteArgs &x, int nb) {
        switch (cfg.block) {
            case 256: launch_vote<256>(x, nb, lds, s); break;
            case 512: launch_vote<512>(x, nb, lds, s); break;
            default: launch_vote<1024>(x, nb, lds, s); break;
        }
    };
    const int *order = nullptr;
    if (cfg.order && nk > 1) {
        unsigned long long *k0 = ws.kpKey.as<unsigned long long>(nk);
        unsigned long long *k1 = ws.kpKey2.as<unsigned long long>(nk);
        int                *i0 = ws.kpOrd.as<int>(nk), *i1 = ws.kpOrd2.as<int>(nk);
        kp_order_kernel<<<div_up(nk, 256), 256, 0, s>>>(in.kx, in.ky, in.kz, in.kpScene,
                                                        in.gridsDev, in.g, nk, k0, i0);
        PPF_HIP_CHECK(hipGetLastError());
        int endBit = 30;
        while ((1 << (endBit - 30)) < std::max(in.nScenes, 1))
            endBit++;
        size_t tb = 0;
        PPF_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tb, k0, k1, i0, i1, nk, 0, endBit, s));
        PPF_HIP_CHECK(rocprim::radix_sort_pairs(ws.cubTmp.reserve(tb), tb, k0, k1, i0, i1, nk, 0, endBit, s));
        order = i1;
    }
    a.kpSel = order;
    int nOverflow = 0;
    PPF_HIP_CHECK(hipEventRecord(ws.voteEv[0], s));
    if (cfg.split) {
        a.hitCap    = cfg.hitCap;
        a.hitStart  = ws.hitStart.as<uint32_t>((size_t)nk * a.hitCap);
        a.hitLen    = ws.hitLen.as<uint32_t>((size_t)nk * a.hitCap);
        a.hitAlpha  = ws.hitAlpha.as<float>((size_t)nk * a.hitCap);
        a.hitBucket = ws.hitBucket.as<uint32_t>((size_t)nk * a.hitCap);
        if (nTiles > 1 && m.refsSorted && m.nKeys > 0) {
            if (ws.splitGen != m.generation || ws.splitRefs != tileRefs ||
                ws.splitTiles != nTiles) {
                uint32_t *sp = ws.splits.as<uint32_t>((size_t)m.nKeys * (nTiles + 1));
                bucket_split_kernel<<<div_up(m.nKeys, 256), 256, 0, s>>>(
                    m.buckets.get<uint2>(), m.nKeys, m.ventries.get<uint2>(), tileRefs,
                    m.angleNum, nTiles, sp);
                PPF_HIP_CHECK(hipGetLastError());
                ws.splitGen   = m.generation;
                ws.splitRefs  = tileRefs;
                ws.splitTiles = nTiles;
            }
            a.splits = ws.splits.get<uint32_t>();
        }
        a.nHit      = ws.hitCount.as<int>((size_t)3 * nk + 1);
        a.nbr       = a.nHit + nk;
        a.overflow  = a.nHit + 2 * (size_t)nk;
        a.nOverflow = a.nHit + 3 * (size_t)nk;
        PPF_HIP_CHECK(hipMemsetAsync(a.nOverflow, 0, sizeof(int), s));
        if (a.grids)
            pair_kernel<true><<<padded(nk), PAIR_BLOCK, 0, s>>>(a);
        else
            pair_kernel<false><<<padded(nk), PAIR_BLOCK, 0, s>>>(a);
        PPF_HIP_CHECK(hipGetLastError());
        PPF_HIP_CHECK(hipEventRecord(ws.voteEv[2], s));
        ws.pairQueued.store(1, std::memory_order_release);
        const bool packed = VOTE_PACKED && nTiles == 1;
        switch (cfg.block) {
            case 256: launch_vote_hits<256>(a, blocks, packed, s); break;
            case 512: launch_vote_hits<512>(a, blocks, packed, s); break;
            default: launch_vote_hits<1024>(a, blocks, packed, s); break;
        }
        PPF_HIP_CHECK(hipGetLastError());
        PPF_HIP_CHECK(hipMemcpyAsync(&nOverflow, a.nOverflow, sizeof(int), hipMemcpyDeviceToHost, s));
        PPF_HIP_CHECK(hipStreamSynchronize(s));
        if (nOverflow > 0) {
            // key points with more hits than the capacity: the fused kernel
            VoteArgs f   = a;
            f.kpSel      = a.overflow;
            f.nk         = nOverflow;
            f.countPairs = 0;
            fused(f, padded(nOverflow) * nTiles);
        }
    } else {
        fused(a, blocks);
    }
    PPF_HIP_CHECK(hipGetLastError());
    PPF_HIP_CHECK(hipEventRecord(ws.voteEv[1], s));

    PoseRec *pslots = reinterpret_cast<PoseRec *>(ws.candTile.as<PoseRec>((size_t)nk * 3));
    int     *counts = ws.voteCnt.as<int>((size_t)nk * 2 + 2);
    int     *offs   = counts + nk + 1;
    poses_kernel<<<div_up(nk, 128), 128, 0, s>>>(a.cand, nk, nTiles, in.kx, in.ky, in.kz, in.knx,
                                                  in.kny, in.knz, m.frames.get<float>(),
                                                  m.cosSin.get<float>(), m.angleNum, pslots,
                                                  counts);
    size_t tb = 0;
    PPF_HIP_CHECK(rocprim::exclusive_scan(nullptr, tb, counts, offs, 0, nk + 1, rocprim::plus<int>(), s));
    // counts[nk] must be 0 so offs[nk] is the total
    PPF_HIP_CHECK(hipMemsetAsync(counts + nk, 0, sizeof(int), s));
    PPF_HIP_CHECK(rocprim::exclusive_scan(ws.cubTmp.reserve(tb), tb, counts, offs, 0, nk + 1, rocprim::plus<int>(), s));
    PoseRec *out = ws.poses.as<PoseRec>((size_t)nk * 3);
    compact_poses_kernel<<<div_up(nk, 128), 128, 0, s>>>(pslots, counts, offs, nk, out);
    int np = 0;
    PPF_HIP_CHECK(hipMemcpyAsync(&np, offs + nk, sizeof(int), hipMemcpyDeviceToHost, s));
    if (stats) {
        unsigned long long st[4];
        PPF_HIP_CHECK(hipMemcpyAsync(st, a.stats, 32, hipMemcpyDeviceToHost, s));
        PPF_HIP_CHECK(hipStreamSynchronize(s));
        stats->pairs = st[0];
        stats->hits  = st[1];
        stats->votes = st[2];
        stats->voted = (int)st[3];
    }
    PPF_HIP_CHECK(hipStreamSynchronize(s));
    if (stats) {
        stats->candidates = (unsigned long long)np;
        float ms          = 0.0f;
        PPF_HIP_CHECK(hipEventElapsedTime(&ms, ws.voteEv[0], ws.voteEv[1]));
        stats->kernelMs = ms;
        stats->pairMs   = 0.0f;
        stats->hitsMs   = 0.0f;
        if (cfg.split) {
            PPF_HIP_CHECK(hipEventElapsedTime(&stats->pairMs, ws.voteEv[0], ws.voteEv[2]));
            stats->hitsMs = ms - stats->pairMs;
        }
        stats->overflow = nOverflow;
    }
    return np;
}

}  // namespace ppfd
