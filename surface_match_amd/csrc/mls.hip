// Moving-least-squares normals: estimateNormalMLS (util.cpp:336-457).
//
// For each query point: its neighbours within `radius` (dist2 < r^2, the
// query itself included), their centroid and covariance (PCA normal = the
// eigenvector of the smallest |eigenvalue|), the query projected onto the
// PCA plane, a Darboux frame (v = unitOrthogonal(n), u = n x v), and a
// weighted least-squares fit, weights exp(-|p - mean|^2 / r^2), of the
// neighbours' heights above the plane by a bivariate polynomial of the
// given order in the frame's (u, v) coordinates.  The query moves to the
// fitted surface (mean + c0 n) and its normal becomes
// normalize(n - c_u u - c_v v), oriented towards the view point (or +z).
//
// Canonical semantics, as for the kNN normals: the neighbour positions are
// a snapshot taken before any point moves (the reference's kd-tree reads
// the cloud it is updating from several threads); every sum runs over the
// neighbours in ascending id order in FP32 (the reference's order is
// Eigen's, which is implementation-defined); the 3x3 eigenproblem and the
// Cholesky solve of the normal equations (lower triangle, as Eigen's LLT
// reads it) run in FP64; exp is expf_c.  A query with fewer than three
// neighbours keeps its point and normal; one whose fit has a non-positive
// pivot or a non-finite solution keeps its point and takes the PCA normal,
// as the reference's `continue` paths leave them.
//
// Pipeline: grid over the valid points (cell = radius) -> neighbour counts
// -> exclusive scan -> neighbour ids -> per-query segmented radix sort of
// the ids -> one thread per query for the fit.
#include "ppf_dev.h"

#include <cstring>
#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace ppfd {
namespace {

__device__ __forceinline__ void cell_range(const GridView &g, float qx, float qy, float qz,
                                           float r2, int lo[3], int hi[3]) {
    const float r = sqrtf(r2) * 1.0001f + 1e-3f * g.cell;
    lo[0] = cell_coord(qx - r, g.ox, g.cell, g.dx), hi[0] = cell_coord(qx + r, g.ox, g.cell, g.dx);
    lo[1] = cell_coord(qy - r, g.oy, g.cell, g.dy), hi[1] = cell_coord(qy + r, g.oy, g.cell, g.dy);
    lo[2] = cell_coord(qz - r, g.oz, g.cell, g.dz), hi[2] = cell_coord(qz + r, g.oz, g.cell, g.dz);
}

// FILL = false: counts[t] = neighbours of query t; true: their ids into
// ids[offs[t] ..] (in no particular order; they are sorted afterwards).
// MLS_GROUP lanes share a query: they test consecutive slots of each run of
// candidate slots, and a wave ballot gives every hit its output position,
// so loads and stores of a group are contiguous.
constexpr int MLS_GROUP = 16;

template <bool FILL>
__global__ void __launch_bounds__(256) mls_neighbours_kernel(GridView g, const float *X,
                                                             const float *Y, const float *Z,
                                                             const int *query, int nq, float r2,
                                                             int *counts, const long long *offs,
                                                             int *ids) {
    const int t    = (int)((blockIdx.x * blockDim.x + threadIdx.x) / MLS_GROUP);
    const int lane = (int)(threadIdx.x & 63u);
    const int l    = lane & (MLS_GROUP - 1);
    const int gsh  = lane & ~(MLS_GROUP - 1);
    if (t >= nq)
        return;
    const int   q  = query[t];
    const float qx = X[q], qy = Y[q], qz = Z[q];
    long long   o  = FILL ? offs[t] : 0;
    int         c  = 0;
    if (g.n > 0 && isfinite(qx) && isfinite(qy) && isfinite(qz)) {
        int lo[3], hi[3];
        cell_range(g, qx, qy, qz, r2, lo, hi);
        auto scan = [&](int s0, int e0) {
            for (int base = s0; base < e0; base += MLS_GROUP) {
                const int  k   = base + l;
                const bool hit = k < e0 && ppfc::dist2(qx, qy, qz, g.x[k], g.y[k], g.z[k]) < r2;
                const uint64_t m  = __ballot(hit);
                const uint32_t gm = (uint32_t)(m >> gsh) & ((1u << MLS_GROUP) - 1u);
                if (FILL && hit)
                    ids[o + __popc(gm & ((1u << l) - 1u))] = g.id[k];
                o += __popc(gm);
                c += __popc(gm);
            }
        };
        for (int cz = lo[2]; cz <= hi[2]; cz++)
            for (int cy = lo[1]; cy <= hi[1]; cy++) {
                if (g.cstart) {  // a row of cells is one run of sorted slots
                    const int row = (cz * g.dy + cy) * g.dx;
                    scan(g.cstart[row + lo[0]], g.cstart[row + hi[0] + 1]);
                } else {
                    for (int cx = lo[0]; cx <= hi[0]; cx++) {
                        int s0, c0;
                        if (cell_lookup(g, cx, cy, cz, s0, c0))
                            scan(s0, s0 + c0);
                    }
                }
            }
    }
    if (!FILL && l == 0)
        counts[t] = c;
}

// Morton code of each query's grid cell (queries at non-finite positions
// last), so that the queries a wave processes are close in space and share
// neighbours in cache
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void mls_query_key_kernel(GridView g, const float *X, const float *Y, const float *Z,
                                     const int *query, int nq, uint32_t *key, int *val) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq)
        return;
    const int   q  = query[t];
    const float qx = X[q], qy = Y[q], qz = Z[q];
    uint32_t    k  = 0xFFFFFFFFu;
    if (isfinite(qx) && isfinite(qy) && isfinite(qz))
        k = spread10((uint32_t)cell_coord(qx, g.ox, g.cell, g.dx)) |
            (spread10((uint32_t)cell_coord(qy, g.oy, g.cell, g.dy)) << 1) |
            (spread10((uint32_t)cell_coord(qz, g.oz, g.cell, g.dz)) << 2);
    key[t] = k;
    val[t] = q;
}

__global__ void mls_pack_kernel(const float *X, const float *Y, const float *Z, int n,
                                float4 *P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        P[i] = make_float4(X[i], Y[i], Z[i], 0.0f);
}

struct MlsArgs {
    const float     *X, *Y, *Z;        // snapshot positions (all points)
    const float4    *P4;               // the same, one 16-byte load per point
    const int       *query;
    const long long *offs;             // nq + 1
    const int       *nb;               // neighbour ids, ascending per query
    int              nq, invert, hasVp;
    float            r2, vp[3];
    float           *OX, *OY, *OZ;     // output positions (start as copies)
    float           *NX, *NY, *NZ;     // normals, in place (never read as neighbours)
};

// body(p) for the neighbours b..e-1 in order, with the loads of four
// neighbours in flight at a time
template <class F>
__device__ __forceinline__ void for_each_nb(const int *__restrict__ nb,
                                            const float4 *__restrict__ P4, long long b,
                                            long long e, F &&body) {
    long long j = b;
    for (; j + 4 <= e; j += 4) {
        const int    i0 = nb[j], i1 = nb[j + 1], i2 = nb[j + 2], i3 = nb[j + 3];
        const float4 w0 = P4[i0], w1 = P4[i1], w2 = P4[i2], w3 = P4[i3];
        body(w0);
        body(w1);
        body(w2);
        body(w3);
    }
    for (; j < e; j++)
        body(P4[nb[j]]);
}

template <int ORDER>
__global__ void __launch_bounds__(128) mls_kernel(MlsArgs a) {
    constexpr int NC = (ORDER + 1) * (ORDER + 2) / 2;
    constexpr int NA = NC * (NC + 1) / 2;
    const int     t  = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nq)
        return;
    const int       q   = a.query[t];
    const long long b   = a.offs[t], e = a.offs[t + 1];
    const int       cnt = (int)(e - b);
    if (cnt < 3)
        return;
    // centroid and covariance (util.cpp:364-367)
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for_each_nb(a.nb, a.P4, b, e, [&](const float4 &w) {
        sx = sx + w.x;
        sy = sy + w.y;
        sz = sz + w.z;
    });
    const float fk     = (float)cnt;
    const float cen[3] = {sx / fk, sy / fk, sz / fk};
    float       c00 = 0.0f, c01 = 0.0f, c02 = 0.0f, c11 = 0.0f, c12 = 0.0f, c22 = 0.0f;
    for_each_nb(a.nb, a.P4, b, e, [&](const float4 &w) {
        const float x = w.x - cen[0], y = w.y - cen[1], z = w.z - cen[2];
        c00           = c00 + x * x;
        c01           = c01 + x * y;
        c02           = c02 + x * z;
        c11           = c11 + y * y;
        c12           = c12 + y * z;
        c22           = c22 + z * z;
    });
    double A3[9] = {c00, c01, c02, c01, c11, c12, c02, c12, c22}, V[9], W[3];
    ppfc::jacobi_eig<3>(A3, V, W);
    int mi = 0;
    for (int k = 1; k < 3; k++)
        if (fabs(W[k]) < fabs(W[mi]))
            mi = k;
    const double vl  = sqrt(V[mi] * V[mi] + V[3 + mi] * V[3 + mi] + V[6 + mi] * V[6 + mi]);
    float        n[3] = {(float)(V[mi] / vl), (float)(V[3 + mi] / vl), (float)(V[6 + mi] / vl)};
    a.NX[q] = n[0];
    a.NY[q] = n[1];
    a.NZ[q] = n[2];
    if (!(isfinite(n[0]) && isfinite(n[1]) && isfinite(n[2])))
        return;
    // projection onto the PCA plane and the Darboux frame (:375-385)
    const float p[3] = {a.X[q], a.Y[q], a.Z[q]};
    const float d    = -ppfc::dot3(n, cen);
    const float dist = ppfc::dot3(p, n) + d;
    const float mean[3] = {p[0] - dist * n[0], p[1] - dist * n[1], p[2] - dist * n[2]};
    float       vA[3], uA[3];
    ppfc::unit_orthogonal(n, vA);
    ppfc::cross3(n, vA, uA);
    if (ORDER <= 1 && NC > cnt)
        return;
    // weighted normal equations, lower triangle (:390-432)
    float M[NA], R[NC];
#pragma unroll
    for (int i = 0; i < NA; i++)
        M[i] = 0.0f;
#pragma unroll
    for (int i = 0; i < NC; i++)
        R[i] = 0.0f;
    for_each_nb(a.nb, a.P4, b, e, [&](const float4 &pw4) {
        const float de[3] = {pw4.x - mean[0], pw4.y - mean[1], pw4.z - mean[2]};
        const float sq    = (de[0] * de[0] + de[1] * de[1]) + de[2] * de[2];
        const float w     = ppfc::expf_c(-sq / a.r2);
        const float uc = ppfc::dot3(de, uA), vc = ppfc::dot3(de, vA), f = ppfc::dot3(de, n);
        float       T[NC];
        int         idx  = 0;
        float       uPow = 1.0f;
#pragma unroll
        for (int ui = 0; ui <= ORDER; ui++) {
            float vPow = 1.0f;
#pragma unroll
            for (int vi = 0; vi <= ORDER - ui; vi++) {
                T[idx++] = uPow * vPow;
                vPow     = vPow * vc;
            }
            uPow = uPow * uc;
        }
#pragma unroll
        for (int i = 0; i < NC; i++) {
            const float pw = T[i] * w;
            R[i]           = R[i] + pw * f;
#pragma unroll
            for (int k = 0; k <= i; k++)
                M[i * (i + 1) / 2 + k] = M[i * (i + 1) / 2 + k] + pw * T[k];
        }
    });
    double Md[NA], c[NC];
#pragma unroll
    for (int i = 0; i < NA; i++)
        Md[i] = M[i];
#pragma unroll
    for (int i = 0; i < NC; i++)
        c[i] = R[i];
    if (!ppfc::cholesky_solve<NC>(Md, c))
        return;
    float cf[NC];
#pragma unroll
    for (int i = 0; i < NC; i++) {
        cf[i] = (float)c[i];
        if (!isfinite(cf[i]))
            return;
    }
    // the point onto the fitted surface, the normal from its slopes (:440-446)
    const float np[3] = {mean[0] + cf[0] * n[0], mean[1] + cf[0] * n[1], mean[2] + cf[0] * n[2]};
    float       nn[3];
    for (int k = 0; k < 3; k++)
        nn[k] = (n[k] - cf[ORDER + 1] * uA[k]) - cf[1] * vA[k];
    const float s2 = (nn[0] * nn[0] + nn[1] * nn[1]) + nn[2] * nn[2];
    if (s2 > 0.0f) {
        const float l = sqrtf(s2);
        nn[0] = nn[0] / l, nn[1] = nn[1] / l, nn[2] = nn[2] / l;
    }
    float val;
    if (a.hasVp) {
        const float to[3] = {a.vp[0] - np[0], a.vp[1] - np[1], a.vp[2] - np[2]};
        val               = ppfc::dot3(nn, to);
    } else {
        const float z[3] = {0.0f, 0.0f, 1.0f};
        val              = ppfc::dot3(nn, z);
    }
    if (a.invert ? val > 0.0f : val < 0.0f)
        nn[0] = -nn[0], nn[1] = -nn[1], nn[2] = -nn[2];
    a.OX[q] = np[0];
    a.OY[q] = np[1];
    a.OZ[q] = np[2];
    a.NX[q] = nn[0];
    a.NY[q] = nn[1];
    a.NZ[q] = nn[2];
}

}  // namespace

void estimate_normals_mls(Workspace &ws, int n, const float *PX, const float *PY, const float *PZ,
                          float *OX, float *OY, float *OZ, float *NX, float *NY, float *NZ,
                          const int *valid, int nValid, const int *query, int nq, float radius,
                          int order, int invert, const float vp[3], MlsStats *st, hipStream_t s) {
    if (order < 1 || order > MLS_MAX_ORDER)
        throw std::range_error("Invalid Input: MLS order out of range (1..3)");
    if (!(radius > 0.0f) || !std::isfinite(radius))
        throw std::range_error("Invalid Input: MLS radius must be positive");
    if (st)
        *st = MlsStats{};
    if (nq <= 0)
        return;
    const float r2 = radius * radius;
    Grid       &g  = ws.gMls;
    grid_build(g, PX, PY, PZ, nullptr, nullptr, nullptr, valid, nValid, radius, s);
    // process the queries in Morton order of their cells (each query's result
    // depends only on itself, so the order changes no value)
    {
        uint32_t *k1 = ws.mlsKey.as<uint32_t>(nq), *k2 = ws.mlsKey2.as<uint32_t>(nq);
        int      *v1 = ws.mlsQ.as<int>(nq), *v2 = ws.mlsQ2.as<int>(nq);
        mls_query_key_kernel<<<div_up(nq, 256), 256, 0, s>>>(g.view, PX, PY, PZ, query, nq, k1,
                                                             v1);
        PPF_HIP_CHECK(hipGetLastError());
        size_t tb = 0;
        PPF_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tb, k1, k2, v1, v2, nq, 0, 32, s));
        PPF_HIP_CHECK(rocprim::radix_sort_pairs(ws.cubTmp.reserve(tb), tb, k1, k2, v1, v2, nq, 0, 32, s));
        query = v2;
    }
    int       *cnt  = ws.mlsCnt.as<int>((size_t)nq + 1);
    long long *offs = ws.mlsOff.as<long long>((size_t)nq + 1);
    const int  blocks  = div_up(nq, 128);
    const int  nblocks = div_up(nq, 256 / MLS_GROUP);
    mls_neighbours_kernel<false><<<nblocks, 256, 0, s>>>(g.view, PX, PY, PZ, query, nq, r2, cnt,
                                                          nullptr, nullptr);
    PPF_HIP_CHECK(hipGetLastError());
    // exclusive scan of the counts in 64-bit, total in offs[nq]
    PPF_HIP_CHECK(hipMemsetAsync(cnt + nq, 0, sizeof(int), s));
    size_t tb = 0;
    PPF_HIP_CHECK(rocprim::exclusive_scan(nullptr, tb, cnt, offs, (long long)0, nq + 1, rocprim::plus<long long>(), s));
    PPF_HIP_CHECK(rocprim::exclusive_scan(ws.cubTmp.reserve(tb), tb, cnt, offs, (long long)0, nq + 1, rocprim::plus<long long>(), s));
    long long total = 0;
    PPF_HIP_CHECK(hipMemcpyAsync(&total, offs + nq, sizeof total, hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
    if (total >= (1ll << 31))
        throw std::runtime_error("MLS neighbourhoods exceed 2^31 entries; use a smaller radius");
    int *ids  = ws.mlsIds.as<int>((size_t)std::max<long long>(total, 1));
    int *ids2 = ws.mlsIds2.as<int>((size_t)std::max<long long>(total, 1));
    if (total > 0) {
        mls_neighbours_kernel<true><<<nblocks, 256, 0, s>>>(g.view, PX, PY, PZ, query, nq, r2,
                                                             nullptr, offs, ids);
        PPF_HIP_CHECK(hipGetLastError());
        int bits = 1;
        while (bits < 31 && (1ll << bits) < (long long)n)
            bits++;
        // ids are non-negative: sort them as unsigned words
        uint32_t *k1 = reinterpret_cast<uint32_t *>(ids), *k2 = reinterpret_cast<uint32_t *>(ids2);
        tb = 0;
        PPF_HIP_CHECK(rocprim::segmented_radix_sort_keys(nullptr, tb, k1, k2, (int)total, nq, offs, offs + 1, 0, bits, s));
        PPF_HIP_CHECK(rocprim::segmented_radix_sort_keys(ws.cubTmp.reserve(tb), tb, k1, k2, (int)total, nq, offs, offs + 1, 0, bits, s));
    }
    float4 *P4 = ws.mlsP4.as<float4>((size_t)std::max(n, 1));
    mls_pack_kernel<<<div_up(std::max(n, 1), 256), 256, 0, s>>>(PX, PY, PZ, n, P4);
    PPF_HIP_CHECK(hipGetLastError());
    MlsArgs a{};
    a.X = PX, a.Y = PY, a.Z = PZ, a.P4 = P4;
    a.query = query, a.offs = offs, a.nb = ids2, a.nq = nq, a.invert = invert;
    a.hasVp = vp && std::isfinite(vp[0]) && std::isfinite(vp[1]) && std::isfinite(vp[2]);
    a.r2    = r2;
    for (int k = 0; k < 3; k++)
        a.vp[k] = a.hasVp ? vp[k] : 0.0f;
    a.OX = OX, a.OY = OY, a.OZ = OZ, a.NX = NX, a.NY = NY, a.NZ = NZ;
    if (order == 1)
        mls_kernel<1><<<blocks, 128, 0, s>>>(a);
    else if (order == 2)
        mls_kernel<2><<<blocks, 128, 0, s>>>(a);
    else
        mls_kernel<3><<<blocks, 128, 0, s>>>(a);
    PPF_HIP_CHECK(hipGetLastError());
    if (st) {
        st->neighbours = total;
        st->queries    = nq;
    }
}

}  // namespace ppfd
