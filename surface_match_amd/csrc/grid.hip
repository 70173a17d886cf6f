// grid.hip -- hashed uniform grid build, voxel-grid downsampling and
// kNN-PCA normal estimation for gfx950.
//
// Voxel sampling replaces the reference's kd-tree greedy Poisson sampling
// (util.cpp:107-139); normal estimation follows util.cpp:234-334 with
// snapshot semantics (all reads see the positions/normals at stage start,
// all write-backs happen afterwards).
#include "ppf_dev.h"

#include <cstring>
#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cmath>

namespace ppfd {

// ---------------------------------------------------------------- bbox

__device__ __forceinline__ int f2ord(float f) {
    int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float ord2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

__global__ void bbox_init_kernel(int *bb) {
    if (threadIdx.x < 3)
        bb[threadIdx.x] = 0x7fffffff;
    else if (threadIdx.x < 6)
        bb[threadIdx.x] = (int)0x80000000;
}

__global__ void __launch_bounds__(256) bbox_kernel(const float *X, const float *Y, const float *Z,
                                                   const int *subset, int m, int *bb) {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < m; t += gridDim.x * blockDim.x) {
        int   i = subset ? subset[t] : t;
        float v[3] = {X[i], Y[i], Z[i]};
        for (int a = 0; a < 3; a++) {
            lo[a] = fminf(lo[a], v[a]);
            hi[a] = fmaxf(hi[a], v[a]);
        }
    }
    for (int a = 0; a < 3; a++) {
        for (int o = 32; o > 0; o >>= 1) {
            lo[a] = fminf(lo[a], __shfl_xor(lo[a], o, 64));
            hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], o, 64));
        }
    }
    // one atomic per block and bound (a few hundred blocks contend per word)
    __shared__ float sLo[4][3], sHi[4][3];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int a = 0; a < 3; a++) {
            sLo[w][a] = lo[a];
            sHi[w][a] = hi[a];
        }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int a = threadIdx.x;
        float     l = sLo[0][a], h = sHi[0][a];
        for (int q = 1; q < 4; q++) {
            l = fminf(l, sLo[q][a]);
            h = fmaxf(h, sHi[q][a]);
        }
        atomicMin(bb + a, f2ord(l));
        atomicMax(bb + 3 + a, f2ord(h));
    }
}

__global__ void bbox_fin_kernel(int *bb) {
    if (threadIdx.x < 6) {
        float f = ord2f(bb[threadIdx.x]);
        bb[threadIdx.x] = __float_as_int(f);
    }
}

void device_bbox(const float *X, const float *Y, const float *Z, const int *subset, int m,
                 float *bb6dev, hipStream_t s) {
    int *bb = reinterpret_cast<int *>(bb6dev);
    bbox_init_kernel<<<1, 64, 0, s>>>(bb);
    if (m > 0) {
        int blocks = std::min(div_up(m, 256), 256);
        bbox_kernel<<<blocks, 256, 0, s>>>(X, Y, Z, subset, m, bb);
    }
    bbox_fin_kernel<<<1, 64, 0, s>>>(bb);
}

// ---------------------------------------------------------------- grid

__global__ void grid_keys_kernel(const float *X, const float *Y, const float *Z, const int *subset,
                                 int m, float ox, float oy, float oz, float cell, int dx, int dy,
                                 int dz, uint32_t *keys, int *vals) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m)
        return;
    int i  = subset ? subset[t] : t;
    int cx = cell_coord(X[i], ox, cell, dx);
    int cy = cell_coord(Y[i], oy, cell, dy);
    int cz = cell_coord(Z[i], oz, cell, dz);
    keys[t] = (uint32_t)((cz * dy + cy) * dx + cx);
    vals[t] = i;
}

// counting sort of a dense grid, first pass: each point's cell key, its
// rank among its cell's points (a per-cell atomic counter) and, for the
// first point of a cell, the cell's coarse occupancy bit
__global__ void grid_rank_kernel(const float *X, const float *Y, const float *Z, const int *subset,
                                 int m, float ox, float oy, float oz, float cell, int dx, int dy,
                                 int dz, int odx, int ody, int sh, uint32_t *keys, int *vals,
                                 int *rank, int *cnt, uint32_t *occ) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m)
        return;
    int            i   = subset ? subset[t] : t;
    int            cx  = cell_coord(X[i], ox, cell, dx);
    int            cy  = cell_coord(Y[i], oy, cell, dy);
    int            cz  = cell_coord(Z[i], oz, cell, dz);
    const uint32_t key = (uint32_t)((cz * dy + cy) * dx + cx);
    keys[t]            = key;
    vals[t]            = i;
    const int r        = atomicAdd(&cnt[key], 1);
    rank[t]            = r;
    if (r == 0) {
        uint32_t b = ((uint32_t)(cz >> sh) * (uint32_t)ody + (uint32_t)(cy >> sh)) * (uint32_t)odx +
                     (uint32_t)(cx >> sh);
        atomicOr(&occ[b >> 5], 1u << (b & 31u));
    }
}

// second pass: every point to its slot (cell start + rank) with the copies
// the grid keeps (ids, positions, normals, the 16-byte {x, y, z, id} and
// {nx, ny, nz, 0} copies), all optional but the ids
__global__ void grid_scatter_kernel(const uint32_t *keys, const int *vals, const int *rank, int m,
                                    const int *cstart, const float *X, const float *Y,
                                    const float *Z, const float *NX, const float *NY,
                                    const float *NZ, int *id, float *x, float *y, float *z,
                                    float *nx, float *ny, float *nz, float4 *p4, float4 *n4) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m)
        return;
    const int i    = vals[t];
    const int slot = cstart[keys[t]] + rank[t];
    id[slot]       = i;
    const float px = X[i], py = Y[i], pz = Z[i];
    x[slot]        = px;
    y[slot]        = py;
    z[slot]        = pz;
    if (p4)
        p4[slot] = make_float4(px, py, pz, __int_as_float(i));
    if (nx) {
        const float qx = NX[i], qy = NY[i], qz = NZ[i];
        nx[slot]       = qx;
        ny[slot]       = qy;
        nz[slot]       = qz;
        if (n4)
            n4[slot] = make_float4(qx, qy, qz, 0.0f);
    }
}

// coarse occupancy bit of every occupied cell (first slot of each cell run)
__global__ void grid_occ_kernel(const uint32_t *keys, int m, int dx, int dy, int odx, int ody,
                                int sh, uint32_t *occ) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m || (s > 0 && keys[s] == keys[s - 1]))
        return;
    uint32_t k  = keys[s];
    int      cx = (int)(k % (uint32_t)dx), cy = (int)((k / (uint32_t)dx) % (uint32_t)dy),
        cz      = (int)(k / ((uint32_t)dx * (uint32_t)dy));
    uint32_t b  = ((uint32_t)(cz >> sh) * (uint32_t)ody + (uint32_t)(cy >> sh)) * (uint32_t)odx +
                 (uint32_t)(cx >> sh);
    atomicOr(&occ[b >> 5], 1u << (b & 31u));
}

// first sorted slot of every cell c in [0, P]: the lower bound of c in the
// sorted keys (P: m)
__global__ void grid_cstart_kernel(const uint32_t *keys, int m, uint32_t P, int *cstart) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > P)
        return;
    int lo = 0, hi = m;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < c)
            lo = mid + 1;
        else
            hi = mid;
    }
    cstart[c] = lo;
}

__global__ void grid_insert_kernel(const uint32_t *keys, int m, uint4 *htab, uint32_t hmask) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m)
        return;
    if (s > 0 && keys[s] == keys[s - 1])
        return;
    uint32_t key = keys[s] + 1u;
    uint32_t h   = cell_hash(key) & hmask;
    for (;;) {
        uint32_t prev = atomicCAS(&htab[h].x, 0u, key);
        if (prev == 0u) {
            htab[h].y = (uint32_t)s;
            return;
        }
        h = (h + 1u) & hmask;
    }
}

__global__ void grid_count_kernel(const uint32_t *keys, int m, uint4 *htab, uint32_t hmask) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m)
        return;
    if (s < m - 1 && keys[s] == keys[s + 1])
        return;
    uint32_t key = keys[s] + 1u;
    uint32_t h   = cell_hash(key) & hmask;
    for (;;) {
        if (htab[h].x == key) {
            htab[h].z = (uint32_t)(s + 1) - htab[h].y;
            return;
        }
        h = (h + 1u) & hmask;
    }
}

__global__ void gather3_kernel(const int *idx, int m, const float *A, const float *B,
                               const float *Cc, float *a, float *b, float *c) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m)
        return;
    int i = idx[s];
    a[s]  = A[i];
    b[s]  = B[i];
    c[s]  = Cc[i];
}

__global__ void grid_clear_kernel(uint32_t *a, size_t na, uint32_t *b, size_t nb) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < na + nb; i += stride) {
        if (i < na)
            a[i] = 0u;
        else
            b[i - na] = 0u;
    }
}

// slot-ordered {x, y, z, id bits} and, with normals, {nx, ny, nz, 0}
__global__ void gather_aos_kernel(const int *idx, int m, const float *X, const float *Y,
                                  const float *Z, const float *NX, const float *NY,
                                  const float *NZ, float4 *p4, float4 *n4) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m)
        return;
    int i = idx[s];
    p4[s] = make_float4(X[i], Y[i], Z[i], __int_as_float(i));
    if (n4)
        n4[s] = make_float4(NX[i], NY[i], NZ[i], 0.0f);
}

void grid_build(Grid &g, const float *X, const float *Y, const float *Z, const float *NX,
                const float *NY, const float *NZ, const int *subset, int m, float cellHint,
                hipStream_t s, const float *hostBox6, bool aos) {
    float hb[6];
    if (hostBox6) {
        // a box known to hold the points, widened by 1e-4 of its extent (a
        // centroid can land a few ulps outside its neighbours' box)
        for (int a = 0; a < 3; a++) {
            const float pad = 1e-4f * (hostBox6[3 + a] - hostBox6[a]) + 1e-6f;
            hb[a]           = hostBox6[a] - pad;
            hb[3 + a]       = hostBox6[3 + a] + pad;
        }
    } else {
        float *bb = g.bbox.as<float>(8);
        device_bbox(X, Y, Z, subset, m, bb, s);
        PPF_HIP_CHECK(hipMemcpyAsync(hb, bb, sizeof hb, hipMemcpyDeviceToHost, s));
        PPF_HIP_CHECK(hipStreamSynchronize(s));
    }
    if (m == 0)
        for (int a = 0; a < 6; a++)
            hb[a] = 0.0f;
    float ext  = std::max({hb[3] - hb[0], hb[4] - hb[1], hb[5] - hb[2], 0.0f});
    float cell = cellHint > 0.0f ? cellHint : 1.0f;
    if (ext / cell > 1000.0f)
        cell = ext / 1000.0f;
    // cap the total number of cells at 2^30 so cell keys fit 30 bits
    int dims[3];
    for (int a = 0; a < 3; a++)
        dims[a] = (int)std::floor((hb[3 + a] - hb[a]) / cell) + 1;
    uint32_t hsize = 64;
    while (hsize < 2u * (uint32_t)std::max(m, 1))
        hsize <<= 1;

    uint32_t *keys  = g.keys.as<uint32_t>(std::max(m, 1));
    uint32_t *keys2 = g.keys2.as<uint32_t>(std::max(m, 1));
    int      *vals  = g.vals.as<int>(std::max(m, 1));
    int      *vals2 = g.vals2.as<int>(std::max(m, 1));
    uint4    *htab  = g.htab.as<uint4>(hsize);
    float    *x     = g.x.as<float>(std::max(m, 1));
    float    *y     = g.y.as<float>(std::max(m, 1));
    float    *z     = g.z.as<float>(std::max(m, 1));
    const size_t cells = (size_t)dims[0] * dims[1] * dims[2];
    const bool   dense = cells <= GRID_DENSE_MAX;  // cell lookups by a dense start array
    int sh = 0;
    while (sh < 2 && ((size_t)(((dims[0] - 1) >> sh) + 1) * (size_t)(((dims[1] - 1) >> sh) + 1) *
                      (size_t)(((dims[2] - 1) >> sh) + 1)) > OCC_MAX_BITS)
        sh++;
    const int odx = ((dims[0] - 1) >> sh) + 1, ody = ((dims[1] - 1) >> sh) + 1,
              odz = ((dims[2] - 1) >> sh) + 1;
    const size_t occWords = ((size_t)odx * ody * odz + 31) / 32;
    uint32_t    *occ      = g.occ.as<uint32_t>(occWords);
    int         *cstart   = nullptr;
    if (dense && m > 0) {
        // counting sort: per-cell counts (cleared with the occupancy bits),
        // their exclusive scan as the cell starts, then one scatter that
        // also writes the point copies
        int *cnt = g.cnt.as<int>(cells + 1);
        cstart   = g.cstart.as<int>(cells + 1);
        grid_clear_kernel<<<(int)std::min<size_t>((cells + 1 + occWords + 255) / 256, 4096), 256, 0,
                            s>>>(reinterpret_cast<uint32_t *>(cnt), cells + 1, occ, occWords);
        {
            int *rank = reinterpret_cast<int *>(keys2);
            grid_rank_kernel<<<div_up(m, 256), 256, 0, s>>>(X, Y, Z, subset, m, hb[0], hb[1], hb[2],
                                                             cell, dims[0], dims[1], dims[2], odx,
                                                             ody, sh, keys, vals, rank, cnt, occ);
            size_t tb = 0;
            PPF_HIP_CHECK(rocprim::exclusive_scan(nullptr, tb, cnt, cstart, 0, cells + 1,
                                                  rocprim::plus<int>(), s));
            void *tmp = g.tmp.reserve(tb);
            PPF_HIP_CHECK(rocprim::exclusive_scan(tmp, tb, cnt, cstart, 0, cells + 1,
                                                  rocprim::plus<int>(), s));
            float  *nx = NX ? g.nx.as<float>(m) : nullptr, *ny = NX ? g.ny.as<float>(m) : nullptr,
                  *nz = NX ? g.nz.as<float>(m) : nullptr;
            float4 *p4 = aos ? g.p4.as<float4>(m) : nullptr,
                   *n4 = (aos && NX) ? g.n4.as<float4>(m) : nullptr;
            grid_scatter_kernel<<<div_up(m, 256), 256, 0, s>>>(keys, vals, rank, m, cstart, X, Y, Z,
                                                                NX, NY, NZ, vals2, x, y, z, nx, ny,
                                                                nz, p4, n4);
        }
    } else {
        // hashed cells (or an empty grid): the cell table and the occupancy
        // bits, cleared in one launch
        grid_clear_kernel<<<(int)std::min<size_t>(((size_t)4 * hsize + occWords + 255) / 256, 2048),
                            256, 0, s>>>(reinterpret_cast<uint32_t *>(htab), (size_t)4 * hsize, occ,
                                         occWords);
        if (m > 0) {
            grid_keys_kernel<<<div_up(m, 256), 256, 0, s>>>(X, Y, Z, subset, m, hb[0], hb[1],
                                                             hb[2], cell, dims[0], dims[1],
                                                             dims[2], keys, vals);
            int endBit = 1;
            while ((1ull << endBit) < (unsigned long long)dims[0] * dims[1] * dims[2])
                endBit++;
            size_t tb = 0;
            PPF_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tb, keys, keys2, vals, vals2, m, 0,
                                                    endBit, s));
            void *tmp = g.tmp.reserve(tb);
            PPF_HIP_CHECK(rocprim::radix_sort_pairs(tmp, tb, keys, keys2, vals, vals2, m, 0, endBit,
                                                    s));
            grid_insert_kernel<<<div_up(m, 256), 256, 0, s>>>(keys2, m, htab, hsize - 1);
            grid_count_kernel<<<div_up(m, 256), 256, 0, s>>>(keys2, m, htab, hsize - 1);
            grid_occ_kernel<<<div_up(m, 256), 256, 0, s>>>(keys2, m, dims[0], dims[1], odx, ody,
                                                            sh, occ);
            gather3_kernel<<<div_up(m, 256), 256, 0, s>>>(vals2, m, X, Y, Z, x, y, z);
            if (NX)
                gather3_kernel<<<div_up(m, 256), 256, 0, s>>>(vals2, m, NX, NY, NZ,
                                                               g.nx.as<float>(m), g.ny.as<float>(m),
                                                               g.nz.as<float>(m));
            if (aos)
                gather_aos_kernel<<<div_up(m, 256), 256, 0, s>>>(
                    vals2, m, X, Y, Z, NX, NY, NZ, g.p4.as<float4>(m),
                    NX ? g.n4.as<float4>(m) : nullptr);
        }
    }
    GridView &v = g.view;
    v.ox = hb[0];
    v.oy = hb[1];
    v.oz = hb[2];
    v.cell  = cell;
    v.dx    = dims[0];
    v.dy    = dims[1];
    v.dz    = dims[2];
    v.hmask = hsize - 1;
    v.htab  = htab;
    v.x = x;
    v.y = y;
    v.z = z;
    v.id = vals2;
    v.n  = m;
    v.occ = occ;
    v.odx = odx;
    v.ody = ody;
    v.occShift = sh;
    v.cstart   = cstart;
    v.nx = v.ny = v.nz = nullptr;
    v.p4 = v.n4 = nullptr;
    if (NX && m > 0) {
        v.nx = g.nx.get<float>();
        v.ny = g.ny.get<float>();
        v.nz = g.nz.get<float>();
    }
    if (aos && m > 0) {
        v.p4 = g.p4.get<float4>();
        v.n4 = NX ? g.n4.get<float4>() : nullptr;
    }
    PPF_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------- voxel sampling

__device__ __forceinline__ unsigned long long voxel_key(float px, float py, float pz,
                                                        const float *bmin, float step,
                                                        float &d) {
    float              p[3] = {px, py, pz}, c[3];
    unsigned long long key  = 0;
    for (int a = 0; a < 3; a++) {
        int vi = (int)floorf((p[a] - bmin[a]) / step);
        vi     = vi < -(1 << 20) ? -(1 << 20) : vi;
        vi     = vi > (1 << 20) - 1 ? (1 << 20) - 1 : vi;
        c[a]   = bmin[a] + ((float)vi + 0.5f) * step;
        key    = (key << 21) | (unsigned long long)(vi + (1 << 20));
    }
    d = ppfc::dist2(px, py, pz, c[0], c[1], c[2]);
    return key;
}

__device__ __forceinline__ uint32_t vox_hash(unsigned long long k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    return (uint32_t)k;
}

struct VoxArgs {
    const float *X, *Y, *Z;
    const int   *subset;
    int          m;
    float        step;
    float        bmin[3];
    unsigned long long *tkey;   // key + 1, 0 = empty
    unsigned long long *tbest;  // (dist bits << 32) | id
    uint32_t     hmask;
};

__global__ void voxel_insert_kernel(VoxArgs a) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.m)
        return;
    int                i = a.subset ? a.subset[t] : t;
    float              d;
    unsigned long long key = voxel_key(a.X[i], a.Y[i], a.Z[i], a.bmin, a.step, d) + 1ull;
    unsigned long long pk  = ((unsigned long long)__float_as_uint(d) << 32) | (uint32_t)i;
    uint32_t           h   = vox_hash(key) & a.hmask;
    for (;;) {
        unsigned long long prev = atomicCAS(&a.tkey[h], 0ull, key);
        if (prev == 0ull || prev == key) {
            atomicMin(&a.tbest[h], pk);
            return;
        }
        h = (h + 1u) & a.hmask;
    }
}

__global__ void voxel_flag_kernel(VoxArgs a, char *flags) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.m)
        return;
    int                i = a.subset ? a.subset[t] : t;
    float              d;
    unsigned long long key = voxel_key(a.X[i], a.Y[i], a.Z[i], a.bmin, a.step, d) + 1ull;
    unsigned long long pk  = ((unsigned long long)__float_as_uint(d) << 32) | (uint32_t)i;
    uint32_t           h   = vox_hash(key) & a.hmask;
    for (;;) {
        if (a.tkey[h] == key) {
            flags[t] = (a.tbest[h] == pk) ? 1 : 0;
            return;
        }
        h = (h + 1u) & a.hmask;
    }
}

__global__ void iota_kernel(int *o, int m) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m)
        o[t] = t;
}

// an empty voxel table: no key, no best point
__global__ void vs_clear_kernel(unsigned long long *tkey, unsigned long long *tbest, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        tkey[i]  = 0ull;
        tbest[i] = ~0ull;
    }
}

int voxel_sample(Workspace &ws, const float *X, const float *Y, const float *Z,
                 const int *subset, int m, float step, const float bmin[3], int *out,
                 int *countDev, hipStream_t s) {
    uint32_t hsize = 64;
    while (hsize < 2u * (uint32_t)std::max(m, 1))
        hsize <<= 1;
    VoxArgs a;
    a.X = X;
    a.Y = Y;
    a.Z = Z;
    a.subset = subset;
    a.m      = m;
    a.step   = step;
    for (int k = 0; k < 3; k++)
        a.bmin[k] = bmin[k];
    a.tkey  = ws.voxKey.as<unsigned long long>(hsize);
    a.tbest = ws.voxBest.as<unsigned long long>(hsize);
    a.hmask = hsize - 1;
    char *flags = ws.voxFlag.as<char>(std::max(m, 1));
    const int *src = subset;
    if (!src) {
        int *io = ws.voxIota.as<int>(std::max(m, 1));
        if (m > 0)
            iota_kernel<<<div_up(m, 256), 256, 0, s>>>(io, m);
        src = io;
    }
    vs_clear_kernel<<<std::min<uint32_t>(div_up(hsize, 256), 2048), 256, 0, s>>>(a.tkey, a.tbest,
                                                                                  hsize);
    if (m > 0) {
        voxel_insert_kernel<<<div_up(m, 256), 256, 0, s>>>(a);
        voxel_flag_kernel<<<div_up(m, 256), 256, 0, s>>>(a, flags);
    }
    size_t tb = 0;
    PPF_HIP_CHECK(rocprim::select(nullptr, tb, src, flags, out, countDev, m, s));
    void *tmp = ws.cubTmp.reserve(tb);
    PPF_HIP_CHECK(rocprim::select(tmp, tb, src, flags, out, countDev, m, s));
    int cnt = 0;
    PPF_HIP_CHECK(hipMemcpyAsync(&cnt, countDev, sizeof(int), hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
    return cnt;
}

// ---- several scenes at once.  Scene w owns the flat elements
// [off[w], off[w + 1]) of a launch; its voxel table is its own slice of the
// shared table (the size voxel_sample would give it), and its grid its own
// cells, counters and bitmap words, so each scene's result is the
// single-scene one.
namespace {
struct VoxSceneDev {
    const float *X, *Y, *Z;
    const int   *subset;
    int         *out;
    float        bmin[3];
    uint32_t     tabOff, hmask;
};

__global__ void vsm_insert_kernel(const VoxSceneDev *sc, const int *off, int S, int M, float step,
                                  unsigned long long *tkey, unsigned long long *tbest) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M)
        return;
    const int          w = owner_of(off, S, t);
    const VoxSceneDev &d = sc[w];
    const int          j = t - off[w];
    const int          i = d.subset ? d.subset[j] : j;
    float              dd;
    unsigned long long key = voxel_key(d.X[i], d.Y[i], d.Z[i], d.bmin, step, dd) + 1ull;
    unsigned long long pk  = ((unsigned long long)__float_as_uint(dd) << 32) | (uint32_t)i;
    uint32_t           h   = vox_hash(key) & d.hmask;
    for (;;) {
        unsigned long long prev = atomicCAS(&tkey[d.tabOff + h], 0ull, key);
        if (prev == 0ull || prev == key) {
            atomicMin(&tbest[d.tabOff + h], pk);
            return;
        }
        h = (h + 1u) & d.hmask;
    }
}

__global__ void vsm_flag_kernel(const VoxSceneDev *sc, const int *off, int S, int M, float step,
                                const unsigned long long *tkey, const unsigned long long *tbest,
                                char *flags) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M)
        return;
    const int          w = owner_of(off, S, t);
    const VoxSceneDev &d = sc[w];
    const int          j = t - off[w];
    const int          i = d.subset ? d.subset[j] : j;
    float              dd;
    unsigned long long key = voxel_key(d.X[i], d.Y[i], d.Z[i], d.bmin, step, dd) + 1ull;
    unsigned long long pk  = ((unsigned long long)__float_as_uint(dd) << 32) | (uint32_t)i;
    uint32_t           h   = vox_hash(key) & d.hmask;
    for (;;) {
        if (tkey[d.tabOff + h] == key) {
            flags[t] = (tbest[d.tabOff + h] == pk) ? 1 : 0;
            return;
        }
        h = (h + 1u) & d.hmask;
    }
}

// base[w] = the first selected position whose flat element is scene w's
// (selected flat elements ascend); base[S] = the selected count
__global__ void vsm_bases_kernel(const int *sel, const int *nSel, const int *off, int S,
                                 int *base) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w > S)
        return;
    const int n = *nSel;
    if (w == S) {
        base[S] = n;
        return;
    }
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (sel[mid] < off[w])
            lo = mid + 1;
        else
            hi = mid;
    }
    base[w] = lo;
}

__global__ void vsm_scatter_kernel(const VoxSceneDev *sc, const int *off, int S, const int *sel,
                                   const int *nSel, const int *base) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= *nSel)
        return;
    const int          t = sel[p];
    const int          w = owner_of(off, S, t);
    const VoxSceneDev &d = sc[w];
    const int          j = t - off[w];
    d.out[p - base[w]]   = d.subset ? d.subset[j] : j;
}
}  // namespace

void voxel_sample_multi(Workspace &ws, const std::vector<VoxSceneIn> &sc, float step,
                        std::vector<int> &counts, hipStream_t s) {
    const int S = (int)sc.size();
    counts.assign(S, 0);
    if (S == 0)
        return;
    std::vector<VoxSceneDev> d(S);
    std::vector<int>         off(S + 1, 0);
    uint32_t                 tab = 0;
    for (int w = 0; w < S; w++) {
        uint32_t hsize = 64;  // voxel_sample's table size for this scene
        while (hsize < 2u * (uint32_t)std::max(sc[w].m, 1))
            hsize <<= 1;
        d[w].X      = sc[w].X;
        d[w].Y      = sc[w].Y;
        d[w].Z      = sc[w].Z;
        d[w].subset = sc[w].subset;
        d[w].out    = sc[w].out;
        for (int a = 0; a < 3; a++)
            d[w].bmin[a] = sc[w].bmin[a];
        d[w].tabOff = tab;
        d[w].hmask  = hsize - 1;
        if ((uint64_t)tab + hsize > 0xFFFFFFFFull)
            throw std::runtime_error("voxel tables of the batch too large");
        tab += hsize;
        off[w + 1] = off[w] + sc[w].m;
    }
    const int M = off[S];
    // one upload: descriptors, offsets, then room for the S + 1 bases
    const size_t dBytes = sizeof(VoxSceneDev) * S, oBytes = sizeof(int) * (S + 1);
    unsigned char *blob = ws.voxMulti.as<unsigned char>(dBytes + 2 * oBytes + 256);
    VoxSceneDev   *dD   = reinterpret_cast<VoxSceneDev *>(blob);
    int           *offD = reinterpret_cast<int *>(blob + dBytes);
    int           *base = offD + (S + 1);
    std::vector<unsigned char> hostBlob(dBytes + oBytes);
    std::memcpy(hostBlob.data(), d.data(), dBytes);
    std::memcpy(hostBlob.data() + dBytes, off.data(), oBytes);
    PPF_HIP_CHECK(hipMemcpyAsync(blob, hostBlob.data(), hostBlob.size(), hipMemcpyHostToDevice, s));
    unsigned long long *tkey  = ws.voxKey.as<unsigned long long>(tab);
    unsigned long long *tbest = ws.voxBest.as<unsigned long long>(tab);
    char *flags = ws.voxFlag.as<char>(std::max(M, 1));
    int  *sel   = ws.voxIota.as<int>(std::max(M, 1));
    int  *nSel  = ws.smallCnt.as<int>(8);
    vs_clear_kernel<<<std::min<uint32_t>(div_up(tab, 256), 4096), 256, 0, s>>>(tkey, tbest, tab);
    if (M > 0) {
        vsm_insert_kernel<<<div_up(M, 256), 256, 0, s>>>(dD, offD, S, M, step, tkey, tbest);
        vsm_flag_kernel<<<div_up(M, 256), 256, 0, s>>>(dD, offD, S, M, step, tkey, tbest, flags);
    }
    size_t tb = 0;
    rocprim::counting_iterator<int> it(0);
    PPF_HIP_CHECK(rocprim::select(nullptr, tb, it, flags, sel, nSel, M, s));
    PPF_HIP_CHECK(rocprim::select(ws.cubTmp.reserve(tb), tb, it, flags, sel, nSel, M, s));
    vsm_bases_kernel<<<div_up(S + 1, 64), 64, 0, s>>>(sel, nSel, offD, S, base);
    if (M > 0)
        vsm_scatter_kernel<<<div_up(M, 256), 256, 0, s>>>(dD, offD, S, sel, nSel, base);
    std::vector<int> b(S + 1);
    PPF_HIP_CHECK(hipMemcpyAsync(b.data(), base, oBytes, hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
    for (int w = 0; w < S; w++)
        counts[w] = b[w + 1] - b[w];
}

namespace {
struct GridSceneDev {
    const float *X, *Y, *Z, *NX, *NY, *NZ;
    const int   *subset;
    float        ox, oy, oz, cell;
    int          dx, dy, dz, odx, ody, sh;
    uint32_t     cellOff, occOff;
    int          idOff;
};

__global__ void gm_rank_kernel(const GridSceneDev *sc, const int *off, int S, int M,
                               uint32_t *keys, int *vals, int *rank, int *cnt, uint32_t *occ) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M)
        return;
    const int           w  = owner_of(off, S, t);
    const GridSceneDev &d  = sc[w];
    const int           j  = t - off[w];
    const int           i  = d.subset ? d.subset[j] : j;
    const int           cx = cell_coord(d.X[i], d.ox, d.cell, d.dx);
    const int           cy = cell_coord(d.Y[i], d.oy, d.cell, d.dy);
    const int           cz = cell_coord(d.Z[i], d.oz, d.cell, d.dz);
    const uint32_t      key = d.cellOff + (uint32_t)((cz * d.dy + cy) * d.dx + cx);
    keys[t]                 = key;
    vals[t]                 = i;
    const int r             = atomicAdd(&cnt[key], 1);
    rank[t]                 = r;
    if (r == 0) {
        const uint32_t b = ((uint32_t)(cz >> d.sh) * (uint32_t)d.ody + (uint32_t)(cy >> d.sh)) *
                               (uint32_t)d.odx + (uint32_t)(cx >> d.sh);
        atomicOr(&occ[d.occOff + (b >> 5)], 1u << (b & 31u));
    }
}

__global__ void gm_scatter_kernel(const GridSceneDev *sc, const int *off, int S, int M,
                                  const uint32_t *keys, const int *vals, const int *rank,
                                  const int *cstart, int *id, float *x, float *y, float *z,
                                  float *nx, float *ny, float *nz, float4 *p4, float4 *n4) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M)
        return;
    const GridSceneDev &d    = sc[owner_of(off, S, t)];
    const int           i    = vals[t];
    const int           slot = cstart[keys[t]] + rank[t];
    const float px = d.X[i], py = d.Y[i], pz = d.Z[i];
    if (x) {
        id[slot] = i + d.idOff;
        x[slot]  = px;
        y[slot]  = py;
        z[slot]  = pz;
    }
    if (p4)
        p4[slot] = make_float4(px, py, pz, __int_as_float(i + d.idOff));
    if (nx) {
        const float qx = d.NX[i], qy = d.NY[i], qz = d.NZ[i];
        nx[slot]       = qx;
        ny[slot]       = qy;
        nz[slot]       = qz;
        if (n4)
            n4[slot] = make_float4(qx, qy, qz, 0.0f);
    }
}
}  // namespace

bool grid_build_multi(MultiGrid &mg, const std::vector<GridSceneIn> &sc, float cellHint,
                      bool aos, hipStream_t s, bool soa) {
    if (!aos && !soa)
        throw std::logic_error("grid_build_multi: a grid needs its point copies");
    const int S = (int)sc.size();
    mg.views.assign(S, GridView{});
    mg.total = 0;
    if (S == 0)
        return true;
    std::vector<GridSceneDev> d(S);
    std::vector<int>          off(S + 1, 0);
    size_t                    cells = 0, occWords = 0;
    const bool                normals = sc[0].NX != nullptr;
    for (int w = 0; w < S; w++) {
        const GridSceneIn &g = sc[w];
        if ((g.NX != nullptr) != normals)
            throw std::logic_error("grid_build_multi: normals for some scenes only");
        // grid_build's choices for this scene (box widened by 1e-4 of its
        // extent, cell grown past 1000 cells an axis, occupancy blocks)
        float hb[6];
        for (int a = 0; a < 3; a++) {
            const float pad = 1e-4f * (g.box[3 + a] - g.box[a]) + 1e-6f;
            hb[a]           = g.box[a] - pad;
            hb[3 + a]       = g.box[3 + a] + pad;
        }
        if (g.m == 0)
            for (int a = 0; a < 6; a++)
                hb[a] = 0.0f;
        const float ext  = std::max({hb[3] - hb[0], hb[4] - hb[1], hb[5] - hb[2], 0.0f});
        float       cell = cellHint > 0.0f ? cellHint : 1.0f;
        if (ext / cell > 1000.0f)
            cell = ext / 1000.0f;
        int dims[3];
        for (int a = 0; a < 3; a++)
            dims[a] = (int)std::floor((hb[3 + a] - hb[a]) / cell) + 1;
        const size_t c = (size_t)dims[0] * dims[1] * dims[2];
        if (c > GRID_DENSE_MAX)
            return false;
        int sh = 0;
        while (sh < 2 && ((size_t)(((dims[0] - 1) >> sh) + 1) * (size_t)(((dims[1] - 1) >> sh) + 1) *
                          (size_t)(((dims[2] - 1) >> sh) + 1)) > OCC_MAX_BITS)
            sh++;
        const int odx = ((dims[0] - 1) >> sh) + 1, ody = ((dims[1] - 1) >> sh) + 1,
                  odz = ((dims[2] - 1) >> sh) + 1;
        d[w] = GridSceneDev{g.X, g.Y, g.Z, g.NX, g.NY, g.NZ, g.subset, hb[0], hb[1], hb[2], cell,
                            dims[0], dims[1], dims[2], odx, ody, sh, (uint32_t)cells,
                            (uint32_t)occWords, g.idOff};
        GridView &v = mg.views[w];
        v.ox = hb[0];
        v.oy = hb[1];
        v.oz = hb[2];
        v.cell = cell;
        v.dx = dims[0];
        v.dy = dims[1];
        v.dz = dims[2];
        v.odx = odx;
        v.ody = ody;
        v.occShift = sh;
        v.n = g.m;
        cells += c + 1;
        occWords += ((size_t)odx * ody * odz + 31) / 32;
        off[w + 1] = off[w] + g.m;
        if (cells > 0x7FFFFFFFull || occWords > 0xFFFFFFFFull)
            return false;
    }
    // all scenes' counters together stay within a budget of 2^28 cells (2 GB
    // of counters and starts; the single-scene builds of the same grids
    // would touch as many, one scene at a time)
    if (cells > (size_t(1) << 28))
        return false;
    const int M = off[S];
    mg.total    = M;
    const size_t dBytes = sizeof(GridSceneDev) * S, oBytes = sizeof(int) * (S + 1);
    unsigned char *blob = mg.tmp.as<unsigned char>(dBytes + oBytes + 256);
    std::vector<unsigned char> hostBlob(dBytes + oBytes);
    std::memcpy(hostBlob.data(), d.data(), dBytes);
    std::memcpy(hostBlob.data() + dBytes, off.data(), oBytes);
    PPF_HIP_CHECK(hipMemcpyAsync(blob, hostBlob.data(), hostBlob.size(), hipMemcpyHostToDevice, s));
    const GridSceneDev *dD   = reinterpret_cast<const GridSceneDev *>(blob);
    const int          *offD = reinterpret_cast<const int *>(blob + dBytes);
    int      *cnt    = mg.cnt.as<int>(cells);
    int      *cstart = mg.cstart.as<int>(cells);
    uint32_t *occ    = mg.occ.as<uint32_t>(std::max<size_t>(occWords, 1));
    uint32_t *keys   = mg.keys.as<uint32_t>(std::max(M, 1));
    int      *vals   = mg.vals.as<int>(std::max(M, 1));
    int      *rank   = mg.rank.as<int>(std::max(M, 1));
    int      *id     = soa ? mg.id.as<int>(std::max(M, 1)) : nullptr;
    float    *x = soa ? mg.x.as<float>(std::max(M, 1)) : nullptr,
          *y = soa ? mg.y.as<float>(std::max(M, 1)) : nullptr,
          *z = soa ? mg.z.as<float>(std::max(M, 1)) : nullptr;
    float  *nx = normals ? mg.nx.as<float>(std::max(M, 1)) : nullptr,
          *ny = normals ? mg.ny.as<float>(std::max(M, 1)) : nullptr,
          *nz = normals ? mg.nz.as<float>(std::max(M, 1)) : nullptr;
    float4 *p4 = aos ? mg.p4.as<float4>(std::max(M, 1)) : nullptr,
           *n4 = (aos && normals) ? mg.n4.as<float4>(std::max(M, 1)) : nullptr;
    const uint4 *htab = mg.htab.as<uint4>(64);  // never probed: the views have cstart
    grid_clear_kernel<<<(int)std::min<size_t>((cells + occWords + 255) / 256, 4096), 256, 0, s>>>(
        reinterpret_cast<uint32_t *>(cnt), cells, occ, occWords);
    if (M > 0)
        gm_rank_kernel<<<div_up(M, 256), 256, 0, s>>>(dD, offD, S, M, keys, vals, rank, cnt, occ);
    size_t tb = 0;
    PPF_HIP_CHECK(rocprim::exclusive_scan(nullptr, tb, cnt, cstart, 0, cells, rocprim::plus<int>(), s));
    void *tmp = mg.scanTmp.reserve(tb);
    PPF_HIP_CHECK(rocprim::exclusive_scan(tmp, tb, cnt, cstart, 0, cells, rocprim::plus<int>(), s));
    if (M > 0)
        gm_scatter_kernel<<<div_up(M, 256), 256, 0, s>>>(dD, offD, S, M, keys, vals, rank, cstart,
                                                          id, x, y, z, nx, ny, nz, p4, n4);
    PPF_HIP_CHECK(hipGetLastError());
    // every view indexes the shared slot arrays: scene w's points are the
    // slots [off[w], off[w + 1]), which its cell starts name directly
    for (int w = 0; w < S; w++) {
        GridView &v = mg.views[w];
        v.hmask     = 63;
        v.htab      = htab;
        v.x         = x;
        v.y         = y;
        v.z         = z;
        v.id        = id;
        v.nx        = nx;
        v.ny        = ny;
        v.nz        = nz;
        v.occ       = occ + d[w].occOff;
        v.cstart    = cstart + d[w].cellOff;
        v.p4        = p4;
        v.n4        = n4;
    }
    return true;
}

// ---- voxel sampling into caller-provided scratch: no allocation, no
// memset, no host synchronisation (the same kernels as voxel_sample; the
// order-keeping compaction is the library's own: per-block counts, one
// workgroup's scan, a scatter)
namespace {
constexpr int VS_BLOCK = 256;

__global__ void __launch_bounds__(VS_BLOCK) vs_count_kernel(const char *flags, int m, int *blockCnt) {
    __shared__ int sScan[VS_BLOCK / 64 + 1];
    const int      t = blockIdx.x * VS_BLOCK + threadIdx.x;
    int            total;
    block_excl_scan<VS_BLOCK>(t < m ? (int)flags[t] : 0, sScan, total);
    if (threadIdx.x == 0)
        blockCnt[blockIdx.x] = total;
}

// one workgroup: exclusive scan of the block counts; the total to *count
__global__ void __launch_bounds__(1024) vs_scan_kernel(int *blockCnt, int nb, int *count) {
    __shared__ int sScan[1024 / 64 + 1];
    int            carry = 0;
    for (int base = 0; base < nb; base += 1024) {
        const int b = base + threadIdx.x;
        const int v = b < nb ? blockCnt[b] : 0;
        int       total;
        const int pre = block_excl_scan<1024>(v, sScan, total);
        if (b < nb)
            blockCnt[b] = carry + pre;
        carry += total;
    }
    if (threadIdx.x == 0)
        *count = carry;
}

__global__ void __launch_bounds__(VS_BLOCK)
    vs_scatter_kernel(const char *flags, const int *subset, int m, const int *blockOff, int *out) {
    __shared__ int sScan[VS_BLOCK / 64 + 1];
    const int      t = blockIdx.x * VS_BLOCK + threadIdx.x;
    const int      f = t < m ? (int)flags[t] : 0;
    int            total;
    const int      pre = block_excl_scan<VS_BLOCK>(f, sScan, total);
    if (f)
        out[blockOff[blockIdx.x] + pre] = subset ? subset[t] : t;
}
}  // namespace

size_t voxel_scratch_bytes(int m) {
    uint32_t hsize = 64;
    while (hsize < 2u * (uint32_t)std::max(m, 1))
        hsize <<= 1;
    const size_t nb = (size_t)div_up(std::max(m, 1), VS_BLOCK);
    return 16 * (size_t)hsize + (((size_t)std::max(m, 1) + 255) & ~(size_t)255) + 4 * nb;
}

void voxel_sample_into(const float *X, const float *Y, const float *Z, const int *subset, int m,
                       float step, const float bmin[3], int *out, int *countDev,
                       unsigned char *scratch, hipStream_t s) {
    uint32_t hsize = 64;
    while (hsize < 2u * (uint32_t)std::max(m, 1))
        hsize <<= 1;
    const int nb = div_up(std::max(m, 1), VS_BLOCK);
    VoxArgs   a;
    a.X      = X;
    a.Y      = Y;
    a.Z      = Z;
    a.subset = subset;
    a.m      = m;
    a.step   = step;
    for (int k = 0; k < 3; k++)
        a.bmin[k] = bmin[k];
    // scratch: keys, best codes, per-point flags, per-block counts
    a.tkey          = reinterpret_cast<unsigned long long *>(scratch);
    a.tbest         = a.tkey + hsize;
    char *flags     = reinterpret_cast<char *>(a.tbest + hsize);
    int  *blockCnt  = reinterpret_cast<int *>(flags + (((size_t)std::max(m, 1) + 255) & ~(size_t)255));
    a.hmask         = hsize - 1;
    vs_clear_kernel<<<std::min<uint32_t>(div_up(hsize, 256), 2048), 256, 0, s>>>(a.tkey, a.tbest,
                                                                                  hsize);
    if (m > 0) {
        voxel_insert_kernel<<<div_up(m, 256), 256, 0, s>>>(a);
        voxel_flag_kernel<<<div_up(m, 256), 256, 0, s>>>(a, flags);
        vs_count_kernel<<<nb, VS_BLOCK, 0, s>>>(flags, m, blockCnt);
    }
    vs_scan_kernel<<<1, 1024, 0, s>>>(blockCnt, m > 0 ? nb : 0, countDev);
    if (m > 0)
        vs_scatter_kernel<<<nb, VS_BLOCK, 0, s>>>(flags, subset, m, blockCnt, out);
    PPF_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------- normals

constexpr int KNN_MAX = 32;

// PCA of a point's k nearest neighbours (computeNormal, util.cpp:234-270):
// FP32 centroid + shifted covariance in kNN order over the snapshot
// positions (g.nx.. = positions indexed by id), FP64 Jacobi eigen-solve,
// eigenvector of the smallest |eigenvalue|.
// The neighbours' coordinates come from px/py/pz (entry t of the list) when
// given, else from the snapshot by id; they are the same values.
template <int KMAX>
__device__ void pca_of_list(const GridView &g, const KnnList<KMAX> &L, float cen[3],
                            float nrm[3], const float *px = nullptr, const float *py = nullptr,
                            const float *pz = nullptr) {
    const int cnt = L.cnt;
    float     sx = 0.0f, sy = 0.0f, sz = 0.0f;
#pragma unroll
    for (int t = 0; t < KMAX; t++) {
        if (t < cnt) {
            int id = L.id[t];
            sx     = sx + (px ? px[t] : g.nx[id]);
            sy     = sy + (px ? py[t] : g.ny[id]);
            sz     = sz + (px ? pz[t] : g.nz[id]);
        }
    }
    float fk = (float)cnt;
    cen[0] = sx / fk;
    cen[1] = sy / fk;
    cen[2] = sz / fk;
    float c00 = 0, c01 = 0, c02 = 0, c11 = 0, c12 = 0, c22 = 0;
#pragma unroll
    for (int t = 0; t < KMAX; t++) {
        if (t < cnt) {
            int   id = L.id[t];
            float x  = (px ? px[t] : g.nx[id]) - cen[0];
            float y  = (px ? py[t] : g.ny[id]) - cen[1];
            float z  = (px ? pz[t] : g.nz[id]) - cen[2];
            c00 = c00 + x * x;
            c01 = c01 + x * y;
            c02 = c02 + x * z;
            c11 = c11 + y * y;
            c12 = c12 + y * z;
            c22 = c22 + z * z;
        }
    }
    double A[9] = {c00, c01, c02, c01, c11, c12, c02, c12, c22}, V[9], W[3];
    ppfc::jacobi_eig<3>(A, V, W);
    int mi = 0;
    for (int t = 1; t < 3; t++)
        if (fabs(W[t]) < fabs(W[mi]))
            mi = t;
    double vx = V[0 * 3 + mi], vy = V[1 * 3 + mi], vz = V[2 * 3 + mi];
    double nn = sqrt(vx * vx + vy * vy + vz * vz);
    nrm[0]    = (float)(vx / nn);
    nrm[1]    = (float)(vy / nn);
    nrm[2]    = (float)(vz / nn);
}

__device__ __forceinline__ bool finite3(float a, float b, float c) {
    return isfinite(a) && isfinite(b) && isfinite(c);
}

struct NormArgs {
    GridView     g;        // grid over valid points; g.nx/ny/nz = snapshot positions (full arrays)
    const float *N0x, *N0y, *N0z;  // snapshot normals (full arrays)
    const int   *query;
    int          nq, k, smooth;
    float       *rawX, *rawY, *rawZ, *cenX, *cenY, *cenZ;  // per point
    int         *have;     // per point: 1 = fresh PCA of a query, 2 = of a neighbour
    int         *nbq;      // nq * k: kNN ids of each query
    int         *cntq;     // nq: kNN count of each query (0: keeps its normal)
    int         *touch;    // per point: neighbour needing a PCA
    int         *mark;     // per point: 1 = query to estimate, 2 = neighbour to estimate
    int         *qidx;     // per point: index in the query list
    // several scenes (estimate_normals_multi): point ids are global, scene w
    // owning [poff[w], poff[w + 1]); its grid is views[w], which shares g's
    // slot arrays and id space (nullptr: one scene, g)
    const GridView *views;
    const int      *poff;
    int             S;
};

// store one estimate (want 1: a query, with its kNN list for smoothing;
// want 2: a neighbour needed by the smoothing)
template <int KMAX>
__device__ void store_estimate(const NormArgs &a, int id, int want, const KnnList<KMAX> &L,
                               const float *px = nullptr, const float *py = nullptr,
                               const float *pz = nullptr) {
    const int cnt = L.cnt;
    if (want == 1)
        a.cntq[a.qidx[id]] = cnt < 3 ? 0 : cnt;
    if (cnt < 3)
        return;
    float cen[3], nrm[3];
    pca_of_list<KMAX>(a.g, L, cen, nrm, px, py, pz);
    a.rawX[id] = nrm[0];
    a.rawY[id] = nrm[1];
    a.rawZ[id] = nrm[2];
    a.cenX[id] = cen[0];
    a.cenY[id] = cen[1];
    a.cenZ[id] = cen[2];
    a.have[id] = want;
    if (want != 1)
        return;
    const size_t t = (size_t)a.qidx[id];
#pragma unroll
    for (int x = 0; x < KMAX; x++) {
        if (x < cnt) {
            const int j     = L.id[x];
            a.nbq[t * a.k + x] = j;
            if (a.smooth && !finite3(a.N0x[j], a.N0y[j], a.N0z[j]))
                a.touch[j] = 1;
        }
    }
}

// queries with a finite normal keep it; the others are marked for
// estimation and counted
__global__ void normals_mark_kernel(NormArgs a, int *nMarked) {
    int  t    = blockIdx.x * blockDim.x + threadIdx.x;
    bool mark = false;
    if (t < a.nq) {
        int id     = a.query[t];
        a.qidx[id] = t;
        if (finite3(a.N0x[id], a.N0y[id], a.N0z[id]))
            a.cntq[t] = 0;
        else {
            a.mark[id] = 1;
            mark       = true;
        }
    }
    const unsigned long long m = __ballot(mark);
    if ((threadIdx.x & 63) == 0 && m)
        atomicAdd(nMarked, __popcll(m));
}

// the grid's points (in cell order) whose mark equals want: estimating them
// in this order lets neighbouring threads search the same cells
__global__ void normals_order_flags_kernel(const int *gid, int n, const int *mark, int want,
                                           char *flags) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n)
        flags[t] = mark[gid[t]] == want ? 1 : 0;
}

// neighbours that the smoothing needs and that have no estimate yet
__global__ void normals_touch_mark_kernel(NormArgs a, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && a.touch[i] && !a.have[i])
        a.mark[i] = 2;
}

// Estimation of one point by the general grid kNN (ring expansion)
template <int KMAX, bool MULTI>
__global__ void __launch_bounds__(128) normals_point_kernel(NormArgs a, const int *list, int nl,
                                                            int want) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nl)
        return;
    int id = list[t];
    if (a.mark[id] != want)
        return;
    KnnList<KMAX> L;
    if constexpr (MULTI) {
        const GridView g = a.views[owner_of(a.poff, a.S, id)];
        grid_knn<KMAX>(g, a.g.nx[id], a.g.ny[id], a.g.nz[id], a.k, L);
    } else
        grid_knn<KMAX>(a.g, a.g.nx[id], a.g.ny[id], a.g.nz[id], a.k, L);
    store_estimate<KMAX>(a, id, want, L);
}

// ---- cooperative kNN: KNN_GROUP lanes per point
//
// The k-best list (k <= KNN_GROUP) is spread over the group, lane i holding
// entry i.  Each batch of KNN_GROUP consecutive slots of a visited cell is
// tested by the group at once (contiguous loads); the candidates that beat
// the current k-th entry are inserted one by one: a ballot finds the suffix
// of entries the candidate precedes, entries shift one lane up.  The list
// ends as the k smallest by (dist2, id), the same set and order as
// grid_knn's, since both visit every cell that can hold a better point.
constexpr int  KNN_GROUP     = 16;
constexpr bool NORMALS_COOP  = true;

__device__ __forceinline__ void coop_knn(const GridView &g, float qx, float qy, float qz, int k,
                                         int l, int gsh, float &myD, int &myId, int &mySlot,
                                         int &cnt) {
    constexpr uint32_t GMASK = (1u << KNN_GROUP) - 1u;
    myD      = FLT_MAX;
    myId     = 0x7fffffff;
    mySlot   = -1;
    cnt      = 0;
    float wd = FLT_MAX;
    int   wi = 0x7fffffff;
    if (g.n == 0 || !(isfinite(qx) && isfinite(qy) && isfinite(qz)))
        return;
    // the points of one cell, KNN_GROUP consecutive slots at a time
    auto scan_cell = [&](int s, int c) {
        for (int base = s; base < s + c; base += KNN_GROUP) {
            const int t    = base + l;
            float     dd   = FLT_MAX;
            int       ii   = 0x7fffffff;
            bool      pass = false;
            if (t < s + c) {
                dd   = ppfc::dist2(qx, qy, qz, g.x[t], g.y[t], g.z[t]);
                ii   = g.id[t];
                pass = cnt < k || dd < wd || (dd == wd && ii < wi);
            }
            uint32_t m = (uint32_t)(__ballot(pass) >> gsh) & GMASK;
            while (m) {
                const int src = __ffs(m) - 1;
                m &= m - 1u;
                const float cd = __shfl(dd, src, KNN_GROUP);
                const int   ci = __shfl(ii, src, KNN_GROUP);
                const int   cs = __shfl(t, src, KNN_GROUP);
                if (cnt == k && !(cd < wd || (cd == wd && ci < wi)))
                    continue;
                const bool before = l < cnt && (cd < myD || (cd == myD && ci < myId));
                const uint32_t bm = (uint32_t)(__ballot(before) >> gsh) & GMASK;
                const int      pos = bm ? __ffs(bm) - 1 : cnt;
                const float    upD = __shfl_up(myD, 1, KNN_GROUP);
                const int      upI = __shfl_up(myId, 1, KNN_GROUP);
                const int      upS = __shfl_up(mySlot, 1, KNN_GROUP);
                if (l > pos && l < k) {
                    myD    = upD;
                    myId   = upI;
                    mySlot = upS;
                }
                if (l == pos) {
                    myD    = cd;
                    myId   = ci;
                    mySlot = cs;
                }
                cnt = cnt < k ? cnt + 1 : k;
                if (cnt == k) {
                    wd = __shfl(myD, k - 1, KNN_GROUP);
                    wi = __shfl(myId, k - 1, KNN_GROUP);
                }
            }
        }
    };
    const int c0x  = cell_coord_unclamped(qx, g.ox, g.cell);
    const int c0y  = cell_coord_unclamped(qy, g.oy, g.cell);
    const int c0z  = cell_coord_unclamped(qz, g.oz, g.cell);
    const int maxR = max(max(g.dx, g.dy), g.dz) + max(max(abs(c0x), abs(c0y)), abs(c0z)) + 1;
    for (int r = first_ring(g, c0x, c0y, c0z); r <= maxR; r++) {
        if (knn_rings_exceed_points(g, r)) {
            // an isolated query (a point far from the rest, or outside the
            // grid): the remaining rings would cost more cells than the grid
            // has points, so the group scans every point once instead (the
            // same k best by (dist2, id))
            myD    = FLT_MAX;
            myId   = 0x7fffffff;
            mySlot = -1;
            cnt    = 0;
            wd     = FLT_MAX;
            wi     = 0x7fffffff;
            scan_cell(knn_first_slot(g), g.n);
            break;
        }
        const int zlo = max(c0z - r, 0), zhi = min(c0z + r, g.dz - 1);
        const int ylo = max(c0y - r, 0), yhi = min(c0y + r, g.dy - 1);
        const int xlo = max(c0x - r, 0), xhi = min(c0x + r, g.dx - 1);
        if (g.occShift != 0 || xhi - xlo >= 32) {
            // rows wider than a bitmap word, or a coarse bitmap: every lane
            // walks the whole ring
            visit_ring(g, c0x, c0y, c0z, r, qx, qy, qz, wd, scan_cell);
        } else if (zlo <= zhi && ylo <= yhi && xlo <= xhi) {
            // the ring's rows split over the group: each lane reads the
            // occupancy bits of its row (the whole row on a z/y face of the
            // shell, else its two end cells) and keeps the cells within the
            // current k-th distance; then the group scans the kept cells of
            // every lane together, in lane order
            const int ny = yhi - ylo + 1, nrows = (zhi - zlo + 1) * ny;
            for (int rb = 0; rb < nrows; rb += KNN_GROUP) {
                const int j  = rb + l;
                uint32_t  bits = 0;
                int       cy = 0, cz = 0;
                if (j < nrows) {
                    cz              = zlo + j / ny;
                    cy              = ylo + j % ny;
                    const bool face = (cz == c0z - r) || (cz == c0z + r) || (cy == c0y - r) ||
                                      (cy == c0y + r);
                    const uint32_t b0 = ((uint32_t)cz * (uint32_t)g.ody + (uint32_t)cy) *
                                            (uint32_t)g.odx + (uint32_t)xlo;
                    const int      n  = xhi - xlo + 1;
                    // bits xlo..xhi of the row (at most two bitmap words)
                    uint32_t       w0 = g.occ[b0 >> 5] >> (b0 & 31u);
                    if ((int)(32u - (b0 & 31u)) < n)
                        w0 |= g.occ[(b0 >> 5) + 1] << (32u - (b0 & 31u));
                    bits = w0 & (n == 32 ? ~0u : ((1u << n) - 1u));
                    if (!face) {
                        uint32_t ends = 0;
                        if (c0x - r >= xlo)
                            ends |= 1u << (c0x - r - xlo);
                        if (c0x + r <= xhi)
                            ends |= 1u << (c0x + r - xlo);
                        bits &= ends;
                    }
                    for (uint32_t b = bits; b;) {
                        const int cx = xlo + __ffs(b) - 1;
                        b &= b - 1u;
                        if (cell_min_d2(g, cx, cy, cz, qx, qy, qz) > wd)
                            bits &= ~(1u << (cx - xlo));
                    }
                }
                uint32_t lanes = (uint32_t)(__ballot(bits != 0u) >> gsh) & GMASK;
                while (lanes) {
                    const int src = __ffs(lanes) - 1;
                    lanes &= lanes - 1u;
                    uint32_t  sb  = (uint32_t)__shfl((int)bits, src, KNN_GROUP);
                    const int scy = __shfl(cy, src, KNN_GROUP), scz = __shfl(cz, src, KNN_GROUP);
                    while (sb) {
                        const int cx = xlo + __ffs(sb) - 1;
                        sb &= sb - 1u;
                        // the bound may have shrunk since the row was read
                        if (cell_min_d2(g, cx, scy, scz, qx, qy, qz) > wd)
                            continue;
                        int s0, c;
                        if (cell_lookup(g, cx, scy, scz, s0, c))
                            scan_cell(s0, c);
                    }
                }
            }
        }
        const float safe = ring_safe(g, c0x, c0y, c0z, r, qx, qy, qz);
        if (safe == FLT_MAX)
            break;
        if (cnt == k && wd < safe * safe)
            break;
    }
}

// the kNN lists of the points list[0..nl) with mark == want: each
// neighbour as {x, y, z, id bits} from the grid's slot-ordered copies (k
// per point, ascending by (dist2, id); the PCA then reads its list as one
// contiguous run instead of gathering by id) and counts (-1: not estimated)
template <bool MULTI>
__global__ void __launch_bounds__(256) normals_knn_coop_kernel(NormArgs a, const int *list,
                                                               int nl, int want, float4 *nbr,
                                                               int *cnts) {
    const int t    = (int)((blockIdx.x * blockDim.x + threadIdx.x) / KNN_GROUP);
    const int lane = (int)(threadIdx.x & 63u);
    const int l    = lane & (KNN_GROUP - 1);
    const int gsh  = lane & ~(KNN_GROUP - 1);
    if (t >= nl)
        return;
    const int id = list[t];
    if (a.mark[id] != want) {
        if (l == 0)
            cnts[t] = -1;
        return;
    }
    float d;
    int   i, slot, cnt;
    if constexpr (MULTI) {
        const GridView g = a.views[owner_of(a.poff, a.S, id)];
        coop_knn(g, a.g.nx[id], a.g.ny[id], a.g.nz[id], a.k, l, gsh, d, i, slot, cnt);
    } else
        coop_knn(a.g, a.g.nx[id], a.g.ny[id], a.g.nz[id], a.k, l, gsh, d, i, slot, cnt);
    if (l < cnt)
        nbr[(size_t)t * a.k + l] =
            make_float4(a.g.x[slot], a.g.y[slot], a.g.z[slot], __int_as_float(i));
    if (l == 0)
        cnts[t] = cnt;
}

// PCA of each point from its kNN list (one thread per point)
template <int KMAX>
__global__ void __launch_bounds__(128) normals_pca_kernel(NormArgs a, const int *list, int nl,
                                                          int want, const float4 *nbr,
                                                          const int *cnts) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nl)
        return;
    const int cnt = cnts[t];
    if (cnt < 0)
        return;
    KnnList<KMAX> L;
    L.init();
    L.cnt = cnt;
    float px[KMAX], py[KMAX], pz[KMAX];
#pragma unroll
    for (int x = 0; x < KMAX; x++)
        if (x < cnt) {
            const float4 P = nbr[(size_t)t * a.k + x];
            px[x]          = P.x;
            py[x]          = P.y;
            pz[x]          = P.z;
            L.id[x]        = __float_as_int(P.w);
        }
    store_estimate<KMAX>(a, list[t], want, L, px, py, pz);
}

// final normal of each query (smoothing, util.cpp:292-306)
__global__ void normals_final_kernel(NormArgs a, float *outN /* nq*3 */) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nq)
        return;
    int id  = a.query[t];
    int cnt = a.cntq[t];
    if (!cnt) {
        outN[3 * t + 0] = a.N0x[id];
        outN[3 * t + 1] = a.N0y[id];
        outN[3 * t + 2] = a.N0z[id];
        return;
    }
    float n0x = a.rawX[id], n0y = a.rawY[id], n0z = a.rawZ[id];
    if (!a.smooth) {
        outN[3 * t + 0] = n0x;
        outN[3 * t + 1] = n0y;
        outN[3 * t + 2] = n0z;
        return;
    }
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
    for (int x = 0; x < cnt; x++) {
        int   j = a.nbq[(size_t)t * a.k + x];
        float jx, jy, jz;
        if (a.have[j]) {
            jx = a.rawX[j];
            jy = a.rawY[j];
            jz = a.rawZ[j];
        } else {
            jx = a.N0x[j];
            jy = a.N0y[j];
            jz = a.N0z[j];
        }
        float d = jx * n0x + jy * n0y + jz * n0z;
        if (d > 0.0f) {
            s0 = s0 + jx;
            s1 = s1 + jy;
            s2 = s2 + jz;
        } else {
            s0 = s0 - jx;
            s1 = s1 - jy;
            s2 = s2 - jz;
        }
    }
    float l = sqrtf(s0 * s0 + s1 * s1 + s2 * s2);
    outN[3 * t + 0] = s0 / l;
    outN[3 * t + 1] = s1 / l;
    outN[3 * t + 2] = s2 / l;
}

// write-back of every fresh PCA (point := centroid, normal := raw)
__global__ void normals_writeback_kernel(NormArgs a, int n, float *PX, float *PY, float *PZ,
                                         float *NX, float *NY, float *NZ) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !a.have[i])
        return;
    PX[i] = a.cenX[i];
    PY[i] = a.cenY[i];
    PZ[i] = a.cenZ[i];
    NX[i] = a.rawX[i];
    NY[i] = a.rawY[i];
    NZ[i] = a.rawZ[i];
}

// a query's final normal oriented towards the viewpoint, or +z without
// one (util.cpp:309-333)
__device__ __forceinline__ void orient_store(float nx, float ny, float nz, float px, float py,
                                             float pz, int invert, int vpFinite, float vx,
                                             float vy, float vz, float *NX, float *NY, float *NZ,
                                             int id) {
    float d;
    if (vpFinite) {
        float dx = vx - px, dy = vy - py, dz = vz - pz;
        d        = nx * dx + ny * dy + nz * dz;
    } else {
        d = nx * 0.0f + ny * 0.0f + nz * 1.0f;
    }
    bool flip = invert ? (d > 0.0f) : (d < 0.0f);
    if (flip) {
        nx = -nx;
        ny = -ny;
        nz = -nz;
    }
    NX[id] = nx;
    NY[id] = ny;
    NZ[id] = nz;
}

// queries take their final normal, then orientation
__global__ void normals_orient_kernel(NormArgs a, const float *outN, float *PX, float *PY,
                                      float *PZ, float *NX, float *NY, float *NZ, int invert,
                                      int vpFinite, float vx, float vy, float vz) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nq)
        return;
    int id = a.query[t];
    orient_store(outN[3 * t + 0], outN[3 * t + 1], outN[3 * t + 2], PX[id], PY[id], PZ[id],
                 invert, vpFinite, vx, vy, vz, NX, NY, NZ, id);
}

// the call's snapshot of positions and normals
__global__ void snapshot6_kernel(int n, const float *a0, const float *a1, const float *a2,
                                 const float *a3, const float *a4, const float *a5, float *b0,
                                 float *b1, float *b2, float *b3, float *b4, float *b5) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        b0[i] = a0[i];
        b1[i] = a1[i];
        b2[i] = a2[i];
        b3[i] = a3[i];
        b4[i] = a4[i];
        b5[i] = a5[i];
    }
}

// per-point flags and the four counters, zeroed
__global__ void normals_clear_kernel(int n, int *have, int *touch, int *mark, int *cnt) {
    const int m = n > 4 ? n : 4;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        if (i < n) {
            have[i]  = 0;
            touch[i] = 0;
            mark[i]  = 0;
        }
        if (i < 4)
            cnt[i] = 0;
    }
}

namespace {
// The estimation proper over prepared arguments (snapshot, grid, query
// list): clears the per-point state of n points, marks the queries,
// estimates them in the grid's cell order and, for the smoothing, the
// neighbours it needs; leaves each query's final normal in outN.  Two
// read-backs (the counts of the points to estimate).
template <bool MULTI>
void normals_core(Workspace &ws, NormArgs &a, int n, float *outN, hipStream_t s) {
    const int nq   = a.nq;
    const int k    = a.k;
    a.rawX = ws.rawX.as<float>(n);
    a.rawY = ws.rawY.as<float>(n);
    a.rawZ = ws.rawZ.as<float>(n);
    a.cenX = ws.cenX.as<float>(n);
    a.cenY = ws.cenY.as<float>(n);
    a.cenZ = ws.cenZ.as<float>(n);
    a.have  = ws.have.as<int>(n);
    a.touch = ws.touch.as<int>(n);
    a.nbq   = ws.nbq.as<int>((size_t)std::max(nq, 1) * k);
    a.cntq  = ws.cntq.as<int>(std::max(nq, 1));
    a.mark  = ws.mark.as<int>(n);
    a.qidx  = ws.qidx.as<int>(n);
    int *cntD = ws.counter.as<int>(4);
    normals_clear_kernel<<<std::min(div_up(std::max(n, 4), 256), 4096), 256, 0, s>>>(
        n, a.have, a.touch, a.mark, cntD);
    if (nq > 0)
        normals_mark_kernel<<<div_up(nq, 256), 256, 0, s>>>(a, cntD + 2);
    // ids with mark == want in the grid's cell order (their count to cntD[0])
    const int ng    = a.g.n;
    int      *olist = ws.touchList.as<int>(std::max(ng, 1));
    auto ordered = [&](int want) {
        char *flags = ws.voxFlag.as<char>(std::max(ng, 1));
        if (ng > 0)
            normals_order_flags_kernel<<<div_up(ng, 256), 256, 0, s>>>(a.g.id, ng, a.mark, want,
                                                                       flags);
        size_t tb = 0;
        PPF_HIP_CHECK(rocprim::select(nullptr, tb, a.g.id, flags, olist, cntD, ng, s));
        PPF_HIP_CHECK(rocprim::select(ws.cubTmp.reserve(tb), tb, a.g.id, flags, olist, cntD, ng, s));
    };

    const bool small = k <= 16;  // kNN list bound: 16 or 32 entries
    auto estimate = [&](const int *list, int nl, int want) {
        if (nl <= 0)
            return;
        if (NORMALS_COOP && k <= KNN_GROUP) {
            float4 *nbr  = ws.knnIds.as<float4>((size_t)nl * k);
            int    *cnts = ws.knnCnt.as<int>(nl);
            normals_knn_coop_kernel<MULTI><<<div_up(nl, 256 / KNN_GROUP), 256, 0, s>>>(
                a, list, nl, want, nbr, cnts);
            normals_pca_kernel<16><<<div_up(nl, 128), 128, 0, s>>>(a, list, nl, want, nbr, cnts);
        } else if (small)
            normals_point_kernel<16, MULTI><<<div_up(nl, 128), 128, 0, s>>>(a, list, nl, want);
        else
            normals_point_kernel<KNN_MAX, MULTI><<<div_up(nl, 128), 128, 0, s>>>(a, list, nl, want);
    };
    // the queries to estimate, in cell order when every one of them is a
    // point of the grid (a query outside the valid set keeps list order)
    if (nq > 0) {
        ordered(1);
        int c[3] = {0, 0, 0};
        PPF_HIP_CHECK(hipMemcpyAsync(c, cntD, sizeof(int) * 3, hipMemcpyDeviceToHost, s));
        PPF_HIP_CHECK(hipStreamSynchronize(s));
        if (c[0] == c[2])
            estimate(olist, c[0], 1);
        else
            estimate(a.query, nq, 1);
    }
    if (a.smooth) {
        // the neighbours the smoothing needs that have no estimate yet (all
        // of them are grid points: they come from kNN lists)
        normals_touch_mark_kernel<<<div_up(n, 256), 256, 0, s>>>(a, n);
        ordered(2);
        int nl = 0;
        PPF_HIP_CHECK(hipMemcpyAsync(&nl, cntD, sizeof(int), hipMemcpyDeviceToHost, s));
        PPF_HIP_CHECK(hipStreamSynchronize(s));
        estimate(olist, nl, 2);
    }
    if (nq > 0)
        normals_final_kernel<<<div_up(nq, 256), 256, 0, s>>>(a, outN);
}
}  // namespace

void estimate_normals(Workspace &ws, int n, float *PX, float *PY, float *PZ, float *NX, float *NY,
                      float *NZ, const int *valid, int nValid, const int *query, int nq, int k,
                      int smooth, int invert, const float vp[3], float cellHint, hipStream_t s,
                      const float *hostBox6) {
    if (k < 3)
        k = 3;
    if (k > KNN_MAX)
        throw std::range_error("knnNormal above 32 is not supported");
    // snapshot copies
    float *P0x = ws.p0x.as<float>(n), *P0y = ws.p0y.as<float>(n), *P0z = ws.p0z.as<float>(n);
    float *N0x = ws.n0x.as<float>(n), *N0y = ws.n0y.as<float>(n), *N0z = ws.n0z.as<float>(n);
    if (n > 0)
        snapshot6_kernel<<<std::min(div_up(n, 256), 4096), 256, 0, s>>>(
            n, PX, PY, PZ, NX, NY, NZ, P0x, P0y, P0z, N0x, N0y, N0z);
    grid_build(ws.gNormal, P0x, P0y, P0z, nullptr, nullptr, nullptr, valid, nValid, cellHint, s,
               hostBox6);

    NormArgs a{};
    a.g    = ws.gNormal.view;
    a.g.nx = P0x;  // kNN coordinates come from the snapshot, indexed by id
    a.g.ny = P0y;
    a.g.nz = P0z;
    a.N0x  = N0x;
    a.N0y  = N0y;
    a.N0z  = N0z;
    a.query  = query;
    a.nq     = nq;
    a.k      = k;
    a.smooth = smooth;
    float *outN = ws.outN.as<float>((size_t)std::max(nq, 1) * 3);
    normals_core<false>(ws, a, n, outN, s);
    normals_writeback_kernel<<<div_up(n, 256), 256, 0, s>>>(a, n, PX, PY, PZ, NX, NY, NZ);
    int vpFinite = vp && std::isfinite(vp[0]) && std::isfinite(vp[1]) && std::isfinite(vp[2]);
    if (nq > 0)
        normals_orient_kernel<<<div_up(nq, 256), 256, 0, s>>>(
            a, outN, PX, PY, PZ, NX, NY, NZ, invert, vpFinite, vpFinite ? vp[0] : 0.0f,
            vpFinite ? vp[1] : 0.0f, vpFinite ? vp[2] : 0.0f);
    PPF_HIP_CHECK(hipGetLastError());
}

namespace {
struct NormSceneDev {
    float     *PX, *PY, *PZ, *NX, *NY, *NZ;
    const int *query;
    float      vx, vy, vz;
    int        vpFinite;
};

// every scene's positions and normals side by side (global id poff[w] + i
// is point i of scene w)
__global__ void nm_snapshot_kernel(const NormSceneDev *sc, const int *poff, int S, int n,
                                   float *P0x, float *P0y, float *P0z, float *N0x, float *N0y,
                                   float *N0z) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const int           w = owner_of(poff, S, t);
        const NormSceneDev &d = sc[w];
        const int           i = t - poff[w];
        P0x[t] = d.PX[i];
        P0y[t] = d.PY[i];
        P0z[t] = d.PZ[i];
        N0x[t] = d.NX[i];
        N0y[t] = d.NY[i];
        N0z[t] = d.NZ[i];
    }
}

// every scene's queries in one list of global ids
__global__ void nm_query_kernel(const NormSceneDev *sc, const int *poff, const int *qoff, int S,
                                int nq, int *gq) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq)
        return;
    const int w = owner_of(qoff, S, t);
    gq[t]       = sc[w].query[t - qoff[w]] + poff[w];
}

// write-back of every fresh PCA into its scene
__global__ void nm_writeback_kernel(NormArgs a, const NormSceneDev *sc, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n || !a.have[t])
        return;
    const int           w = owner_of(a.poff, a.S, t);
    const NormSceneDev &d = sc[w];
    const int           i = t - a.poff[w];
    d.PX[i] = a.cenX[t];
    d.PY[i] = a.cenY[t];
    d.PZ[i] = a.cenZ[t];
    d.NX[i] = a.rawX[t];
    d.NY[i] = a.rawY[t];
    d.NZ[i] = a.rawZ[t];
}

__global__ void nm_orient_kernel(NormArgs a, const NormSceneDev *sc, const float *outN,
                                 int invert) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nq)
        return;
    const int           id = a.query[t];
    const int           w  = owner_of(a.poff, a.S, id);
    const NormSceneDev &d  = sc[w];
    const int           i  = id - a.poff[w];
    orient_store(outN[3 * t + 0], outN[3 * t + 1], outN[3 * t + 2], d.PX[i], d.PY[i], d.PZ[i],
                 invert, d.vpFinite, d.vx, d.vy, d.vz, d.NX, d.NY, d.NZ, i);
}
}  // namespace

void estimate_normals_multi(Workspace &ws, MultiGrid &mg, const std::vector<NormSceneIn> &sc, int k,
                            int smooth, int invert, float cellHint, hipStream_t s) {
    if (k < 3)
        k = 3;
    if (k > KNN_MAX)
        throw std::range_error("knnNormal above 32 is not supported");
    const int S = (int)sc.size();
    if (S == 0)
        return;
    std::vector<int> poff(S + 1, 0), qoff(S + 1, 0);
    for (int w = 0; w < S; w++) {
        if ((long long)poff[w] + sc[w].n > 0x7FFFFFFFLL ||
            (long long)qoff[w] + sc[w].nq > 0x7FFFFFFFLL)
            throw std::runtime_error("estimate_normals_multi: batch too large");
        poff[w + 1] = poff[w] + sc[w].n;
        qoff[w + 1] = qoff[w] + sc[w].nq;
    }
    // the kNN grid of every scene's valid points, over the positions the
    // call starts from (the snapshot's), ids global
    std::vector<GridSceneIn> gin(S);
    for (int w = 0; w < S; w++) {
        gin[w] = GridSceneIn{sc[w].PX, sc[w].PY, sc[w].PZ, nullptr, nullptr, nullptr,
                             sc[w].valid, sc[w].nValid, {}, poff[w]};
        std::memcpy(gin[w].box, sc[w].box, sizeof gin[w].box);
    }
    if (!grid_build_multi(mg, gin, cellHint, false, s)) {
        for (const NormSceneIn &c : sc)
            estimate_normals(ws, c.n, c.PX, c.PY, c.PZ, c.NX, c.NY, c.NZ, c.valid, c.nValid,
                             c.query, c.nq, k, smooth, invert, c.vp, cellHint, s, c.box);
        return;
    }
    const int n = poff[S], nq = qoff[S];
    std::vector<NormSceneDev> d(S);
    for (int w = 0; w < S; w++) {
        const NormSceneIn &c = sc[w];
        const bool fin = std::isfinite(c.vp[0]) && std::isfinite(c.vp[1]) && std::isfinite(c.vp[2]);
        d[w] = NormSceneDev{c.PX, c.PY, c.PZ, c.NX, c.NY, c.NZ, c.query,
                            fin ? c.vp[0] : 0.0f, fin ? c.vp[1] : 0.0f, fin ? c.vp[2] : 0.0f,
                            fin ? 1 : 0};
    }
    // one upload: descriptors, point offsets, query offsets, the views
    const size_t dBytes = sizeof(NormSceneDev) * S, oBytes = sizeof(int) * (S + 1),
                 vOff = (dBytes + 2 * oBytes + 255) / 256 * 256, vBytes = sizeof(GridView) * S;
    unsigned char *blob = ws.voxMulti.as<unsigned char>(vOff + vBytes);
    std::vector<unsigned char> hostBlob(vOff + vBytes, 0);
    std::memcpy(hostBlob.data(), d.data(), dBytes);
    std::memcpy(hostBlob.data() + dBytes, poff.data(), oBytes);
    std::memcpy(hostBlob.data() + dBytes + oBytes, qoff.data(), oBytes);
    std::memcpy(hostBlob.data() + vOff, mg.views.data(), vBytes);
    PPF_HIP_CHECK(hipMemcpyAsync(blob, hostBlob.data(), hostBlob.size(), hipMemcpyHostToDevice, s));
    const NormSceneDev *dD    = reinterpret_cast<const NormSceneDev *>(blob);
    const int          *poffD = reinterpret_cast<const int *>(blob + dBytes);
    const int          *qoffD = poffD + (S + 1);
    const GridView     *views = reinterpret_cast<const GridView *>(blob + vOff);

    float *P0x = ws.p0x.as<float>(n), *P0y = ws.p0y.as<float>(n), *P0z = ws.p0z.as<float>(n);
    float *N0x = ws.n0x.as<float>(n), *N0y = ws.n0y.as<float>(n), *N0z = ws.n0z.as<float>(n);
    if (n > 0)
        nm_snapshot_kernel<<<std::min(div_up(n, 256), 4096), 256, 0, s>>>(
            dD, poffD, S, n, P0x, P0y, P0z, N0x, N0y, N0z);
    int *gq = ws.voxIota.as<int>(std::max(nq, 1));
    if (nq > 0)
        nm_query_kernel<<<div_up(nq, 256), 256, 0, s>>>(dD, poffD, qoffD, S, nq, gq);

    NormArgs a{};
    a.g    = mg.views[0];  // the shared slot arrays and ids (its cells are scene 0's)
    a.g.n  = mg.total;
    a.g.nx = P0x;          // kNN coordinates come from the snapshot, indexed by global id
    a.g.ny = P0y;
    a.g.nz = P0z;
    a.N0x  = N0x;
    a.N0y  = N0y;
    a.N0z  = N0z;
    a.query  = gq;
    a.nq     = nq;
    a.k      = k;
    a.smooth = smooth;
    a.views  = views;
    a.poff   = poffD;
    a.S      = S;
    float *outN = ws.outN.as<float>((size_t)std::max(nq, 1) * 3);
    normals_core<true>(ws, a, n, outN, s);
    if (n > 0)
        nm_writeback_kernel<<<div_up(n, 256), 256, 0, s>>>(a, dD, n);
    if (nq > 0)
        nm_orient_kernel<<<div_up(nq, 256), 256, 0, s>>>(a, dD, outN, invert);
    PPF_HIP_CHECK(hipGetLastError());
}

}  // namespace ppfd
