// dev_common.h -- device plumbing shared by the PPF kernels: error checks,
// growable device buffers, block-level scans and the hashed uniform grid.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>
#include <cstdint>
#include <cfloat>

#include "canon.h"

namespace ppfd {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define PPF_HIP_CHECK(expr)                                                                    \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::ppfd::HipError(std::string("HIP error ") + hipGetErrorString(e_) + " at " + \
                                   __FILE__ + ":" + std::to_string(__LINE__) + " (" #expr ")"); \
    } while (0)

// A device allocation that only ever grows; reused across calls so the
// steady-state match path performs no hipMalloc.
struct DevBuf {
    void  *ptr   = nullptr;
    size_t bytes = 0;
    DevBuf()     = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (ptr)
            (void)hipFree(ptr);
    }
    void *reserve(size_t b) {
        if (b <= bytes)
            return ptr;
        if (ptr)
            PPF_HIP_CHECK(hipFree(ptr));
        ptr   = nullptr;
        bytes = 0;
        size_t nb = b < 256 ? 256 : b + b / 8;
        PPF_HIP_CHECK(hipMalloc(&ptr, nb));
        bytes = nb;
        return ptr;
    }
    template <class T> T *as(size_t n) { return reinterpret_cast<T *>(reserve(n * sizeof(T))); }
    template <class T> T *get() const { return reinterpret_cast<T *>(ptr); }
};

inline int div_up(long long a, long long b) { return (int)((a + b - 1) / b); }

// --------------------------------------------------------------------------
// hashed uniform grid
// --------------------------------------------------------------------------
// Points of a cloud (or a subset) are sorted by cell: grids of at most
// GRID_DENSE_MAX cells by a counting sort (the order of the points inside a
// cell is that of the per-cell atomic counters, which no search depends
// on: every search orders its candidates by (dist2, id)), larger ones by a
// stable radix sort with their occupied cells in an open-addressing table
// {key+1, start, count, 0} (key+1 == 0 marks an empty slot).  Cells are
// floor((p - org) / cell) per axis, clamped into [0, dim).
struct GridView {
    float           ox, oy, oz, cell;
    int             dx, dy, dz;
    uint32_t        hmask;
    const uint4    *htab;
    const float    *x, *y, *z;
    const float    *nx, *ny, *nz;  // optional (nullptr)
    const int      *id;            // original point id of each sorted slot
    // occupancy of blocks of 2^occShift cells per axis (occShift 0: of every
    // cell), one bit per block, x fastest: lets searches skip empty space
    // without table probes
    const uint32_t *occ;
    int             odx, ody, occShift;
    int             n;  // (no padding bytes: views are compared bytewise as graph keys)
    // grids of at most GRID_DENSE_MAX cells: the first sorted slot of every
    // cell (linear key), cstart[key + 1] - cstart[key] points; neighbouring
    // cells of a row are then neighbouring words instead of scattered table
    // slots.  nullptr: only the hash table.
    const int      *cstart;
    // slot-ordered copies, one 16-byte load per point (grids built with
    // them, else nullptr): {x, y, z, id bits} and, with normals,
    // {nx, ny, nz, 0}
    const float4   *p4, *n4;
};
static_assert(sizeof(GridView) == 144, "GridView holds no padding bytes");

constexpr size_t GRID_DENSE_MAX = size_t(1) << 24;

// Several scenes in one launch: scene w owns the flat elements
// [off[w], off[w + 1]).  The scene owning element t is the last w with
// off[w] <= t (empty scenes share their offset with the next one and are
// never chosen).
__device__ __forceinline__ int owner_of(const int *off, int S, int t) {
    int lo = 0, hi = S - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// occupancy bitmaps up to this many bits use one bit per cell; larger grids
// one bit per 2x2x2 or 4x4x4 block
constexpr size_t OCC_MAX_BITS = size_t(1) << 26;

__device__ __forceinline__ bool coarse_empty(const GridView &g, int cx, int cy, int cz) {
    const int      sh = g.occShift;
    const uint32_t b  = ((uint32_t)(cz >> sh) * (uint32_t)g.ody + (uint32_t)(cy >> sh)) *
                           (uint32_t)g.odx +
                       (uint32_t)(cx >> sh);
    return ((g.occ[b >> 5] >> (b & 31u)) & 1u) == 0u;
}

__device__ __forceinline__ int cell_coord(float v, float o, float cell, int dim) {
    int c = (int)floorf((v - o) / cell);
    c     = c < 0 ? 0 : c;
    return c >= dim ? dim - 1 : c;
}
__device__ __forceinline__ int cell_coord_unclamped(float v, float o, float cell) {
    float f = floorf((v - o) / cell);
    f       = fminf(fmaxf(f, -1.0e6f), 1.0e6f);
    return (int)f;
}
__device__ __forceinline__ uint32_t cell_hash(uint32_t key) {
    key *= 0x9E3779B1u;
    return key ^ (key >> 16);
}
__device__ __forceinline__ bool cell_lookup(const GridView &g, int cx, int cy, int cz, int &start,
                                            int &count) {
    if (g.cstart) {
        const uint32_t c = (uint32_t)((cz * g.dy + cy) * g.dx + cx);
        start            = g.cstart[c];
        count            = g.cstart[c + 1] - start;
        return count > 0;
    }
    uint32_t key = (uint32_t)((cz * g.dy + cy) * g.dx + cx) + 1u;
    uint32_t h   = cell_hash(key) & g.hmask;
    for (;;) {
        uint4 s = g.htab[h];
        if (s.x == 0u)
            return false;
        if (s.x == key) {
            start = (int)s.y;
            count = (int)s.z;
            return true;
        }
        h = (h + 1u) & g.hmask;
    }
}

// squared distance from q to the axis-aligned box of cell (cx,cy,cz)
__device__ __forceinline__ float cell_min_d2(const GridView &g, int cx, int cy, int cz, float qx,
                                             float qy, float qz) {
    // box grown by 1e-3 cell so float binning at cell faces never prunes a point
    float m  = 1e-3f * g.cell;
    float lx = g.ox + (float)cx * g.cell - m, ly = g.oy + (float)cy * g.cell - m,
          lz = g.oz + (float)cz * g.cell - m, e = g.cell + 2.0f * m;
    float ex = fmaxf(fmaxf(lx - qx, qx - (lx + e)), 0.0f);
    float ey = fmaxf(fmaxf(ly - qy, qy - (ly + e)), 0.0f);
    float ez = fmaxf(fmaxf(lz - qz, qz - (lz + e)), 0.0f);
    return ex * ex + ey * ey + ez * ez;
}

// k nearest neighbours by (dist2, id) with ring expansion.  Results are
// sorted ascending; returns the count found (< k only if the grid holds
// fewer points).  KMAX is a compile-time bound on k; every array index is a
// compile-time constant after unrolling, so the list stays in registers.
template <int KMAX>
struct KnnList {
    float d[KMAX];
    int   id[KMAX];
    int   cnt;
    float wd;  // the k-th entry once the list is full (else +inf)
    int   wi;
    __device__ __forceinline__ void init() {
        cnt = 0;
        wd  = FLT_MAX;
        wi  = 0x7fffffff;
    }
    __device__ __forceinline__ void push(float dd, int ii, int k) {
        if (cnt == k && !(dd < wd || (dd == wd && ii < wi)))
            return;
        // insertion by carrying: the carried entry swaps with every larger one
        // on its way down; slot cnt (when free) takes the last carried entry
        float cd = dd;
        int   ci = ii;
#pragma unroll
        for (int i = 0; i < KMAX; i++) {
            if (i < k && i <= cnt) {
                const bool take = (i == cnt) || cd < d[i] || (cd == d[i] && ci < id[i]);
                if (take) {
                    const float td = d[i];
                    const int   ti = id[i];
                    d[i]           = cd;
                    id[i]          = ci;
                    cd             = td;
                    ci             = ti;
                }
            }
        }
        cnt = cnt < k ? cnt + 1 : k;
        if (cnt == k) {
#pragma unroll
            for (int i = 0; i < KMAX; i++)
                if (i == k - 1) {
                    wd = d[i];
                    wi = id[i];
                }
        }
    }
};

// Visit the cells of row (cy, cz) of Chebyshev ring r around (c0x,c0y,c0z)
// that lie in [xlo, xhi], calling f(start,count) for occupied cells whose box
// is within bound2 of q: a row on the shell's z/y faces over its whole x
// range, a row off them at its two ends.  With a per-cell occupancy bitmap
// (occShift 0) a row's cells are read 32 at a time from the bitmap words and
// only the occupied ones are visited, so empty space costs one load per word
// rather than work per cell.
template <class F>
__device__ __forceinline__ void visit_row(const GridView &g, int c0x, int c0y, int c0z, int r,
                                          int cy, int cz, int xlo, int xhi, float qx, float qy,
                                          float qz, float bound2, F &f) {
    const bool face = (cz == c0z - r) || (cz == c0z + r) || (cy == c0y - r) || (cy == c0y + r);
    const uint32_t rowBase = ((uint32_t)cz * (uint32_t)g.ody + (uint32_t)cy) * (uint32_t)g.odx;
    for (int seg = 0; seg < (face ? 1 : 2); seg++) {
        // the segment [a0, a1] of this row: the whole row on a face, else
        // one end of it
        int a0 = xlo, a1 = xhi;
        if (!face) {
            a0 = a1 = seg == 0 ? c0x - r : c0x + r;
            if (a0 < xlo || a0 > xhi)
                continue;
        }
        for (int x = a0; x <= a1;) {
            uint32_t  bits;
            const int x0 = x;
            if (g.occShift == 0) {
                const uint32_t b   = rowBase + (uint32_t)x;
                const int      off = (int)(b & 31u);
                const int      n   = min(32 - off, a1 - x + 1);
                bits               = g.occ[b >> 5] >> off;
                bits &= n == 32 ? ~0u : ((1u << n) - 1u);
                x += n;
            } else {
                bits = coarse_empty(g, x, cy, cz) ? 0u : 1u;
                x += 1;
            }
            while (bits) {
                const int cx = x0 + __ffs(bits) - 1;
                bits &= bits - 1u;
                if (cell_min_d2(g, cx, cy, cz, qx, qy, qz) > bound2)
                    continue;
                int s0, c0;
                if (cell_lookup(g, cx, cy, cz, s0, c0))
                    f(s0, c0);
            }
        }
    }
}

// Visit the cells of Chebyshev ring r around (c0x,c0y,c0z), calling
// f(start,count) for occupied cells whose box is within bound2 of q, row by
// row (visit_row).  A single call site of f keeps the register use down.
template <class F>
__device__ __forceinline__ void visit_ring(const GridView &g, int c0x, int c0y, int c0z, int r,
                                           float qx, float qy, float qz, float bound2, F &&f) {
    const int zlo = max(c0z - r, 0), zhi = min(c0z + r, g.dz - 1);
    const int ylo = max(c0y - r, 0), yhi = min(c0y + r, g.dy - 1);
    const int xlo = max(c0x - r, 0), xhi = min(c0x + r, g.dx - 1);
    if (zlo > zhi || ylo > yhi || xlo > xhi)
        return;
    for (int cz = zlo; cz <= zhi; cz++)
        for (int cy = ylo; cy <= yhi; cy++)
            visit_row(g, c0x, c0y, c0z, r, cy, cz, xlo, xhi, qx, qy, qz, bound2, f);
}

// The rows row0, row0 + step, row0 + 2 step, ... of ring r (rows in z-major
// order of the ring's in-grid (cy, cz) pairs): lanes of a group that each
// take every step-th row visit the ring's cells between them exactly once.
template <class F>
__device__ __forceinline__ void visit_ring_rows(const GridView &g, int c0x, int c0y, int c0z,
                                                int r, float qx, float qy, float qz, float bound2,
                                                int row0, int step, F &&f) {
    const int zlo = max(c0z - r, 0), zhi = min(c0z + r, g.dz - 1);
    const int ylo = max(c0y - r, 0), yhi = min(c0y + r, g.dy - 1);
    const int xlo = max(c0x - r, 0), xhi = min(c0x + r, g.dx - 1);
    if (zlo > zhi || ylo > yhi || xlo > xhi)
        return;
    const int ny = yhi - ylo + 1;
    int       jz = row0 / ny, jy = row0 - jz * ny;
    for (int cz = zlo + jz; cz <= zhi;) {
        visit_row(g, c0x, c0y, c0z, r, ylo + jy, cz, xlo, xhi, qx, qy, qz, bound2, f);
        jy += step;
        while (jy >= ny) {
            jy -= ny;
            cz++;
        }
    }
}

// distance from q to the boundary of the (2r+1)^3 block of cells around c0;
// every point not yet visited is at least this far away.  Returns a huge
// value once the block covers the whole grid.
__device__ __forceinline__ float ring_safe(const GridView &g, int c0x, int c0y, int c0z, int r,
                                           float qx, float qy, float qz) {
    if (c0x - r <= 0 && c0x + r >= g.dx - 1 && c0y - r <= 0 && c0y + r >= g.dy - 1 &&
        c0z - r <= 0 && c0z + r >= g.dz - 1)
        return FLT_MAX;
    float s = FLT_MAX;
    if (c0x - r > 0)        s = fminf(s, qx - (g.ox + (float)(c0x - r) * g.cell));
    if (c0x + r < g.dx - 1) s = fminf(s, (g.ox + (float)(c0x + r + 1) * g.cell) - qx);
    if (c0y - r > 0)        s = fminf(s, qy - (g.oy + (float)(c0y - r) * g.cell));
    if (c0y + r < g.dy - 1) s = fminf(s, (g.oy + (float)(c0y + r + 1) * g.cell) - qy);
    if (c0z - r > 0)        s = fminf(s, qz - (g.oz + (float)(c0z - r) * g.cell));
    if (c0z + r < g.dz - 1) s = fminf(s, (g.oz + (float)(c0z + r + 1) * g.cell) - qz);
    s = s - 1e-3f * g.cell;
    return s > 0.0f ? s : 0.0f;
}

// smallest Chebyshev ring around c0 that reaches a cell of the grid
// (rings before it are empty when the query lies outside the grid)
__device__ __forceinline__ int first_ring(const GridView &g, int c0x, int c0y, int c0z) {
    int rx = c0x < 0 ? -c0x : (c0x >= g.dx ? c0x - (g.dx - 1) : 0);
    int ry = c0y < 0 ? -c0y : (c0y >= g.dy ? c0y - (g.dy - 1) : 0);
    int rz = c0z < 0 ? -c0z : (c0z >= g.dz ? c0z - (g.dz - 1) : 0);
    return max(max(rx, ry), rz);
}

// The ring searches walk O(r^2) cells a ring: a query far from every other
// point (an outlier of a large grid, or a query outside it) would walk
// O(R^3) cells.  Once the cube of rings up to r holds more cells than the
// grid has points, a kNN search scans the grid's points instead.
__device__ __forceinline__ bool knn_rings_exceed_points(const GridView &g, int r) {
    const long long w = 2LL * r + 1LL;
    return r > 1 && w * w * w > (long long)g.n;
}
// the grid's first slot: its points are the slots [first, first + n) (a
// view of a several-scene grid starts where its scene's cells start)
__device__ __forceinline__ int knn_first_slot(const GridView &g) {
    return g.cstart ? g.cstart[0] : 0;
}

template <int KMAX>
__device__ void grid_knn(const GridView &g, float qx, float qy, float qz, int k,
                         KnnList<KMAX> &L) {
    L.init();
    if (g.n == 0 || !(isfinite(qx) && isfinite(qy) && isfinite(qz)))
        return;
    int c0x = cell_coord_unclamped(qx, g.ox, g.cell);
    int c0y = cell_coord_unclamped(qy, g.oy, g.cell);
    int c0z = cell_coord_unclamped(qz, g.oz, g.cell);
    int maxR = max(max(g.dx, g.dy), g.dz) + max(max(abs(c0x), abs(c0y)), abs(c0z)) + 1;
    for (int r = first_ring(g, c0x, c0y, c0z); r <= maxR; r++) {
        if (knn_rings_exceed_points(g, r)) {
            L.init();
            const int t0 = knn_first_slot(g);
            for (int t = t0; t < t0 + g.n; t++)
                L.push(ppfc::dist2(qx, qy, qz, g.x[t], g.y[t], g.z[t]), g.id[t], k);
            break;
        }
        float bound2 = L.wd;
        visit_ring(g, c0x, c0y, c0z, r, qx, qy, qz, bound2, [&](int s, int c) {
            for (int t = s; t < s + c; t++) {
                float dd = ppfc::dist2(qx, qy, qz, g.x[t], g.y[t], g.z[t]);
                L.push(dd, g.id[t], k);
            }
        });
        float safe = ring_safe(g, c0x, c0y, c0z, r, qx, qy, qz);
        if (safe == FLT_MAX)
            break;
        if (L.cnt == k && L.wd < safe * safe)
            break;
    }
}

// nearest neighbour by (dist2, id) among points with dist2 <= maxD2;
// returns the sorted slot (-1 if none).  With maxD2 = FLT_MAX it is the
// unbounded 1-NN (findClosestPoint, util.cpp:603-626); a finite bound keeps
// far queries from walking empty rings.
__device__ inline int grid_nn1(const GridView &g, float qx, float qy, float qz, float &bestD,
                               int &bestId, float maxD2 = FLT_MAX) {
    bestD      = FLT_MAX;
    bestId     = 0x7fffffff;
    int bestS  = -1;
    if (g.n == 0 || !(isfinite(qx) && isfinite(qy) && isfinite(qz)))
        return -1;
    int c0x  = cell_coord_unclamped(qx, g.ox, g.cell);
    int c0y  = cell_coord_unclamped(qy, g.oy, g.cell);
    int c0z  = cell_coord_unclamped(qz, g.oz, g.cell);
    int maxR = max(max(g.dx, g.dy), g.dz) + max(max(abs(c0x), abs(c0y)), abs(c0z)) + 1;
    for (int r = first_ring(g, c0x, c0y, c0z); r <= maxR; r++) {
        visit_ring(g, c0x, c0y, c0z, r, qx, qy, qz, fminf(bestD, maxD2), [&](int s, int c) {
            for (int t = s; t < s + c; t++) {
                const float4 P  = g.p4[t];
                const float  dd = ppfc::dist2(qx, qy, qz, P.x, P.y, P.z);
                const int    ii = __float_as_int(P.w);
                if (dd > maxD2)
                    continue;
                if (dd < bestD || (dd == bestD && ii < bestId)) {
                    bestD  = dd;
                    bestId = ii;
                    bestS  = t;
                }
            }
        });
        float safe = ring_safe(g, c0x, c0y, c0z, r, qx, qy, qz);
        if (safe == FLT_MAX)
            break;
        if (bestS >= 0 && bestD < safe * safe)
            break;
        if (maxD2 < safe * safe)  // nothing unvisited can be within the bound
            break;
    }
    return bestS;
}

// grid_nn1 by a group of G consecutive lanes (G a power of two <= 64) that
// all pass the same query: each lane visits every G-th row of a ring, the
// group merges its candidates by (dist2, id) after every ring, and every
// lane returns the same result as grid_nn1 would.  Control flow is uniform
// within a group.
template <int G>
__device__ int grid_nn1_group(const GridView &g, float qx, float qy, float qz, float &bestD,
                              int &bestId, float maxD2) {
    const int gl = (int)(threadIdx.x & (G - 1));
    bestD        = FLT_MAX;
    bestId       = 0x7fffffff;
    int bestS    = -1;
    if (g.n == 0 || !(isfinite(qx) && isfinite(qy) && isfinite(qz)))
        return -1;
    const int c0x  = cell_coord_unclamped(qx, g.ox, g.cell);
    const int c0y  = cell_coord_unclamped(qy, g.oy, g.cell);
    const int c0z  = cell_coord_unclamped(qz, g.oz, g.cell);
    const int maxR = max(max(g.dx, g.dy), g.dz) + max(max(abs(c0x), abs(c0y)), abs(c0z)) + 1;
    int       r0   = first_ring(g, c0x, c0y, c0z);
    if (G == 8 && r0 == 0 && g.cstart) {
        // rings 0 and 1 as the 3x3x3 block around the query cell: nine rows
        // of up to three cells, each row one contiguous run of sorted slots
        // (a row's cells are consecutive keys), trimmed to the cells the ring
        // walk would visit (cell_min_d2 within the bound).  The group scans
        // every run together, lane l taking slots l, l + G, ... (neighbouring
        // lanes read neighbouring points), first the query's own row, then
        // the other eight rows (one per lane to find their runs) under the
        // best distance found so far.
        float bound2 = maxD2;
        for (int phase = 0; phase < 2; phase++) {
            // this lane's row: the centre row in phase 0, row gl of the
            // other eight in phase 1
            int dyz = phase == 0 ? 4 : (gl < 4 ? gl : gl + 1);
            const int cz = c0z - 1 + dyz / 3, cy = c0y - 1 + dyz % 3;
            int       t0 = 0, t1 = 0;
            if (cz >= 0 && cz < g.dz && cy >= 0 && cy < g.dy) {
                int lo = 2, hi = 0;
                for (int d = 0; d < 3; d++) {
                    const int cx = c0x - 1 + d;
                    if (cx < 0 || cx >= g.dx || cell_min_d2(g, cx, cy, cz, qx, qy, qz) > bound2)
                        continue;
                    lo = min(lo, d);
                    hi = max(hi, d);
                }
                if (lo <= hi) {
                    const int *cs = g.cstart + (cz * g.dy + cy) * g.dx + c0x - 1;
                    t0            = cs[lo];
                    t1            = cs[hi + 1];
                }
            }
            const int base = (int)(threadIdx.x & 63u) & ~(G - 1);
            for (int q = 0; q < (phase == 0 ? 1 : G); q++) {
                const int s0 = __shfl(t0, base + q, 64), s1 = __shfl(t1, base + q, 64);
                for (int t = s0 + gl; t < s1; t += G) {
                    const float4 P  = g.p4[t];
                    const float  dd = ppfc::dist2(qx, qy, qz, P.x, P.y, P.z);
                    const int    ii = __float_as_int(P.w);
                    if (dd > maxD2)
                        continue;
                    if (dd < bestD || (dd == bestD && ii < bestId)) {
                        bestD  = dd;
                        bestId = ii;
                        bestS  = t;
                    }
                }
            }
#pragma unroll
            for (int o = G / 2; o > 0; o >>= 1) {
                const float od = __shfl_xor(bestD, o, 64);
                const int   oi = __shfl_xor(bestId, o, 64), os = __shfl_xor(bestS, o, 64);
                if (od < bestD || (od == bestD && oi < bestId)) {
                    bestD  = od;
                    bestId = oi;
                    bestS  = os;
                }
            }
            bound2 = fminf(bestD, maxD2);
        }
        const float safe = ring_safe(g, c0x, c0y, c0z, 1, qx, qy, qz);
        if (safe == FLT_MAX || (bestS >= 0 && bestD < safe * safe) || maxD2 < safe * safe)
            return bestS;
        r0 = 2;
    }
    for (int r = r0; r <= maxR; r++) {
        // each lane walks every G-th row of the ring (the ring's cells
        // between them once), so empty space costs bitmap words, not cells
        visit_ring_rows(g, c0x, c0y, c0z, r, qx, qy, qz, fminf(bestD, maxD2), gl, G,
                        [&](int s0, int c) {
                            for (int t = s0; t < s0 + c; t++) {
                                const float4 P  = g.p4[t];
                                const float  dd = ppfc::dist2(qx, qy, qz, P.x, P.y, P.z);
                                const int    ii = __float_as_int(P.w);
                                if (dd > maxD2)
                                    continue;
                                if (dd < bestD || (dd == bestD && ii < bestId)) {
                                    bestD  = dd;
                                    bestId = ii;
                                    bestS  = t;
                                }
                            }
                        });
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) {
            const float od = __shfl_xor(bestD, o, 64);
            const int   oi = __shfl_xor(bestId, o, 64), os = __shfl_xor(bestS, o, 64);
            if (od < bestD || (od == bestD && oi < bestId)) {
                bestD  = od;
                bestId = oi;
                bestS  = os;
            }
        }
        const float safe = ring_safe(g, c0x, c0y, c0z, r, qx, qy, qz);
        if (safe == FLT_MAX)
            break;
        if (bestS >= 0 && bestD < safe * safe)
            break;
        if (maxD2 < safe * safe)
            break;
    }
    return bestS;
}

// grid_any_within by a group of G consecutive lanes passing the same query:
// the rows of the query's box of cells in turn, the group scanning each row's
// points together (lane l taking points l, l + G, ...; with cstart a row's
// cells are one contiguous run of sorted slots); every lane returns the
// group's answer.
template <int G>
__device__ bool grid_any_within_group(const GridView &g, float qx, float qy, float qz, float r2) {
    static_assert(G >= 1 && G < 64 && (G & (G - 1)) == 0, "group of 1..32 lanes");
    if (g.n == 0 || !(isfinite(qx) && isfinite(qy) && isfinite(qz)))
        return false;
    const int      gl    = (int)(threadIdx.x & (G - 1));
    const uint32_t gbase = threadIdx.x & 63u & ~(uint32_t)(G - 1);
    float          r     = sqrtf(r2) * 1.0001f + 1e-3f * g.cell;
    int xlo = cell_coord(qx - r, g.ox, g.cell, g.dx), xhi = cell_coord(qx + r, g.ox, g.cell, g.dx);
    int ylo = cell_coord(qy - r, g.oy, g.cell, g.dy), yhi = cell_coord(qy + r, g.oy, g.cell, g.dy);
    int zlo = cell_coord(qz - r, g.oz, g.cell, g.dz), zhi = cell_coord(qz + r, g.oz, g.cell, g.dz);
    const int ny = yhi - ylo + 1, nrows = ny * (zhi - zlo + 1);
    const int ncx = xhi - xlo + 1;
    for (int j = 0; j < nrows; j++) {
        const int cz = zlo + j / ny, cy = ylo + j % ny;
        bool      f  = false;
        for (int cxi = 0; cxi < (g.cstart ? 1 : ncx); cxi++) {
            int s0, s1;
            if (g.cstart) {
                const int *cs = g.cstart + (cz * g.dy + cy) * g.dx;
                s0            = cs[xlo];
                s1            = cs[xhi + 1];
            } else {
                int c;
                if (coarse_empty(g, xlo + cxi, cy, cz) || !cell_lookup(g, xlo + cxi, cy, cz, s0, c))
                    continue;
                s1 = s0 + c;
            }
            for (int t = s0 + gl; t < s1; t += G) {
                const float4 P = g.p4[t];
                if (ppfc::dist2(qx, qy, qz, P.x, P.y, P.z) < r2) {
                    f = true;
                    break;
                }
            }
        }
        if (((__ballot(f) >> gbase) & ((1ull << G) - 1ull)) != 0ull)
            return true;
    }
    return false;
}

// true if any indexed point has dist2 < r2
__device__ inline bool grid_any_within(const GridView &g, float qx, float qy, float qz, float r2) {
    if (g.n == 0 || !(isfinite(qx) && isfinite(qy) && isfinite(qz)))
        return false;
    float r   = sqrtf(r2) * 1.0001f + 1e-3f * g.cell;
    int   xlo = cell_coord(qx - r, g.ox, g.cell, g.dx), xhi = cell_coord(qx + r, g.ox, g.cell, g.dx);
    int   ylo = cell_coord(qy - r, g.oy, g.cell, g.dy), yhi = cell_coord(qy + r, g.oy, g.cell, g.dy);
    int   zlo = cell_coord(qz - r, g.oz, g.cell, g.dz), zhi = cell_coord(qz + r, g.oz, g.cell, g.dz);
    for (int cz = zlo; cz <= zhi; cz++)
        for (int cy = ylo; cy <= yhi; cy++)
            for (int cx = xlo; cx <= xhi; cx++) {
                int s, c;
                if (coarse_empty(g, cx, cy, cz) || !cell_lookup(g, cx, cy, cz, s, c))
                    continue;
                for (int t = s; t < s + c; t++) {
                    const float4 P = g.p4[t];
                    if (ppfc::dist2(qx, qy, qz, P.x, P.y, P.z) < r2)
                        return true;
                }
            }
    return false;
}

// --------------------------------------------------------------------------
// host side of the grid
// --------------------------------------------------------------------------
struct Grid {
    DevBuf   keys, keys2, vals, vals2, htab, x, y, z, nx, ny, nz, tmp, bbox, occ, cstart, p4, n4,
        cnt;
    GridView view{};
    int      capN = 0;
};

// Build a grid over points (X,Y,Z SoA of a cloud) restricted to `subset`
// (device int list of length m; nullptr = all m points).  cellHint is the
// requested cell edge; it grows if the grid would exceed 1023 cells per
// axis.  withNormals gathers NX/NY/NZ too.  Synchronises `s` once (bbox).
// hostBox6 (min xyz, max xyz), when given, is a box known to contain the
// points; it saves the device bounding box and its read-back.  aos also
// writes the view's p4 (and, with normals, n4) copies, which the pair kernel
// and the refiner's searches (grid_nn1, grid_nn1_group, grid_any_within)
// read.
void grid_build(Grid &g, const float *X, const float *Y, const float *Z, const float *NX,
                const float *NY, const float *NZ, const int *subset, int m, float cellHint,
                hipStream_t s, const float *hostBox6 = nullptr, bool aos = false);

// block-wide exclusive scan of one int per thread (blockDim a multiple of 64, <= 1024)
template <int BLOCK>
__device__ inline int block_excl_scan(int v, int *lds /* >= BLOCK/64+1 */, int &total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int       x    = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o)
            x += y;
    }
    if (lane == 63)
        lds[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int w = 0; w < BLOCK / 64; w++) {
            int t  = lds[w];
            lds[w] = acc;
            acc += t;
        }
        lds[BLOCK / 64] = acc;
    }
    __syncthreads();
    int res = lds[wid] + x - v;
    total   = lds[BLOCK / 64];
    __syncthreads();
    return res;
}

}  // namespace ppfd
