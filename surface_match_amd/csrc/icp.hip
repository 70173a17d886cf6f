// icp.hip -- batched ICP refinement and inlier scoring on gfx950.
//
// Reference: icp.cpp:123-278 (ICP::regist with findCorresponds, the
// symmetric objective of symmetric_rigid_matching :39-87, the "mse" of
// :182-194 and the convergence logic of :233-272) and util.cpp:628-641
// (inliner).  One workgroup owns one hypothesis and runs every iteration of
// it without leaving the kernel, so all hypotheses of a match are refined
// by a single launch.  Per iteration:
//   1. every source point is moved by the current transform and its nearest
//      scene point found on the hashed grid; pairs beyond rejectDist are
//      dropped; each scene point keeps its closest source point (ties: the
//      lowest source index), through a per-hypothesis hash table with 64-bit
//      atomicMin on (dist2 bits, source index);
//   2. the normal equations are summed in FP32 with compensated (two-sum)
//      accumulation and reduced in a fixed order (deterministic);
//   3. one lane solves the 6x6 system in FP64 (Jacobi pseudo-inverse, the
//      threshold of JacobiSVD::solve) and composes the update in FP64;
//   4. the mean nearest-neighbour distance after the update feeds the
//      reference's convergence rules, quirks included.
#include "ppf_dev.h"

#include <algorithm>
#include <cstring>

namespace ppfd {

// compensated FP32 accumulator: value = hi + lo
struct CF {
    float hi, lo;
};
__device__ __forceinline__ void cf_add(CF &a, float x) {
    float s  = a.hi + x;
    float bb = s - a.hi;
    float e  = (a.hi - (s - bb)) + (x - bb);
    a.hi     = s;
    a.lo     = a.lo + e;
}
__device__ __forceinline__ CF cf_merge(CF a, CF b) {
    float s  = a.hi + b.hi;
    float bb = s - a.hi;
    float e  = (a.hi - (s - bb)) + (b.hi - bb);
    CF r;
    r.hi = s;
    r.lo = (a.lo + b.lo) + e;
    return r;
}

// fixed-order block reduction of M compensated sums; result valid in red[]
// for every thread after the call
template <int BLOCK, int M>
__device__ void block_reduce_cf(CF (&v)[M], CF *lds /* (BLOCK/64) * M */, double *red) {
#pragma unroll
    for (int k = 0; k < M; k++) {
        for (int o = 1; o < 64; o <<= 1) {
            CF u;
            u.hi = __shfl_xor(v[k].hi, o, 64);
            u.lo = __shfl_xor(v[k].lo, o, 64);
            // order the two operands by lane so every lane computes the same value
            v[k] = ((threadIdx.x & o) == 0) ? cf_merge(v[k], u) : cf_merge(u, v[k]);
        }
    }
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < M; k++)
            lds[wid * M + k] = v[k];
    __syncthreads();
    if (threadIdx.x < M) {
        CF acc = lds[threadIdx.x];
        for (int w = 1; w < BLOCK / 64; w++)
            acc = cf_merge(acc, lds[w * M + threadIdx.x]);
        red[threadIdx.x] = (double)acc.hi + (double)acc.lo;
    }
    __syncthreads();
}

__device__ __forceinline__ uint32_t icp_hash(uint32_t k) {
    k *= 0x85ebca6bu;
    return k ^ (k >> 13);
}

struct IcpArgs {
    int          nSrc, H;
    const float *sx, *sy, *sz, *snx, *sny, *snz;
    GridView     dst, full;
    const float *init;  // H x 12
    IcpCriteria  c;
    float        scoringDist;
    int          mode;
    uint32_t     T;      // hash slots per hypothesis (power of two)
    uint32_t    *hkey;   // H x T
    unsigned long long *hbest;  // H x T
    int         *nnSlot; // H x nSrc: dst sorted slot of the pair (-1 = none)
    int         *hslot;  // H x nSrc: hash slot used
    int         *wslot;  // H x nSrc: dst slot of a winning pair, -1 otherwise (set by add_means)
    float       *nnD;    // H x nSrc
    IcpResultDev *out;
    // several scenes: the target grids of hypothesis h are dstG[hypScene[h]]
    // and fullG[hypScene[h]] (nullptr: dst and full for every hypothesis)
    const GridView *dstG, *fullG;
    const int      *hypScene;
};

// the arguments with hypothesis h's target grids in dst / full
__device__ __forceinline__ IcpArgs args_for(const IcpArgs &a, int h) {
    IcpArgs b = a;
    if (a.dstG) {
        const int sc = a.hypScene[h];
        b.dst        = a.dstG[sc];
        b.full       = a.fullG[sc];
    }
    return b;
}

// u = A^-1 rhs by an FP64 Cholesky factorisation A = L L^T, taken only when
// no eigenvalue of A can fall below the pseudo-inverse threshold of
// sym_solve6 (then pinv(A) = A^-1 and both give the same u up to FP64
// rounding): lambda_min >= 1 / ||A^-1||_F >= 1 / ||L^-1||_F^2 and
// lambda_max <= ||A||_inf.  Returns false (u untouched) otherwise.  A
// few hundred dependent flops instead of the Jacobi sweeps' ~10^4.
__device__ __forceinline__ bool chol_solve6(const double A[36], const double rhs[6], double u[6]) {
    // every loop is unrolled, so L (its lower triangle, packed by rows) and
    // the columns of L^-1 stay in registers: the last workgroup's single
    // lane runs this, and arrays in scratch memory would cost it a memory
    // round trip per access
    double L[21];
#define PPF_L(i, j) L[(i) * ((i) + 1) / 2 + (j)]
    double amax = 0.0;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        double rs = 0.0;
#pragma unroll
        for (int j = 0; j < 6; j++)
            rs += fabs(A[i * 6 + j]);
        amax = fmax(amax, rs);
    }
#pragma unroll
    for (int j = 0; j < 6; j++) {
        double d = A[j * 6 + j];
#pragma unroll
        for (int k = 0; k < j; k++)
            d -= PPF_L(j, k) * PPF_L(j, k);
        if (!(d > 0.0))
            return false;
        const double ljj = sqrt(d);
        PPF_L(j, j)      = ljj;
#pragma unroll
        for (int i = j + 1; i < 6; i++) {
            double s = A[i * 6 + j];
#pragma unroll
            for (int k = 0; k < j; k++)
                s -= PPF_L(i, k) * PPF_L(j, k);
            PPF_L(i, j) = s / ljj;
        }
    }
    // squared Frobenius norm of L^-1, column j = L^-1 e_j by forward
    // substitution
    double fro = 0.0;
#pragma unroll
    for (int j = 0; j < 6; j++) {
        double x[6];
#pragma unroll
        for (int i = j; i < 6; i++) {
            double s = i == j ? 1.0 : 0.0;
#pragma unroll
            for (int k = j; k < i; k++)
                s -= PPF_L(i, k) * x[k];
            x[i] = i == j ? 1.0 / PPF_L(i, i) : s / PPF_L(i, i);
            fro += x[i] * x[i];
        }
    }
    // a margin of 1e-3 over the threshold absorbs the rounding of the bounds
    if (!(1.0 / fro > 6.0 * 1.1920928955078125e-07 * amax * 1.001))
        return false;
    double y[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        double s = rhs[i];
#pragma unroll
        for (int k = 0; k < i; k++)
            s -= PPF_L(i, k) * y[k];
        y[i] = s / PPF_L(i, i);
    }
#pragma unroll
    for (int i = 5; i >= 0; i--) {
        double s = y[i];
#pragma unroll
        for (int k = i + 1; k < 6; k++)
            s -= PPF_L(k, i) * u[k];
        u[i] = s / PPF_L(i, i);
    }
#undef PPF_L
    return true;
}

// u = pinv(A) rhs for symmetric PSD A (FP64 Jacobi), dropping eigenvalues
// below 6 * eps_f32 * max |eigenvalue| as JacobiSVD::solve does.  Not
// inlined: its arrays live in scratch memory, which the Cholesky path of
// sym_solve6 then never touches.
__device__ __noinline__ void pinv_solve6(const double *A, const double *rhs, double *u) {
    double a[36], V[36], W[6];
    for (int i = 0; i < 36; i++)
        a[i] = A[i];
    ppfc::jacobi_eig<6>(a, V, W);
    double mx = 0.0;
    for (int i = 0; i < 6; i++)
        mx = fmax(mx, fabs(W[i]));
    double thr = mx * 6.0 * 1.1920928955078125e-07;
    for (int i = 0; i < 6; i++)
        u[i] = 0.0;
    for (int k = 0; k < 6; k++) {
        if (!(fabs(W[k]) > thr))
            continue;
        double proj = 0.0;
        for (int i = 0; i < 6; i++)
            proj += V[i * 6 + k] * rhs[i];
        proj /= W[k];
        for (int i = 0; i < 6; i++)
            u[i] += V[i * 6 + k] * proj;
    }
}

// u = pinv(A) rhs: the Cholesky path when the threshold provably drops no
// eigenvalue, else the Jacobi pseudo-inverse
__device__ __forceinline__ void sym_solve6(const double A[36], const double rhs[6], double u[6]) {
    if (!chol_solve6(A, rhs, u))
        pinv_solve6(A, rhs, u);
}

__device__ void compose3x4(const double *A, const double *B, double *C) {
    double t[12];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 4; c++) {
            double v = A[r * 4 + 0] * B[0 * 4 + c] + A[r * 4 + 1] * B[1 * 4 + c] +
                       A[r * 4 + 2] * B[2 * 4 + c];
            if (c == 3)
                v += A[r * 4 + 3];
            t[r * 4 + c] = v;
        }
    for (int i = 0; i < 12; i++)
        C[i] = t[i];
}

__device__ void rodrigues(const double w[3], double R[9]) {
    double th = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (!(th > 0.0)) {
        for (int i = 0; i < 9; i++)
            R[i] = (i % 4 == 0) ? 1.0 : 0.0;
        return;
    }
    double k[3] = {w[0] / th, w[1] / th, w[2] / th}, c = cos(th), s = sin(th);
    double K[9] = {0, -k[2], k[1], k[2], 0, -k[0], -k[1], k[0], 0};
    for (int r = 0; r < 3; r++)
        for (int q = 0; q < 3; q++) {
            double kk = 0.0;
            for (int m = 0; m < 3; m++)
                kk += K[r * 3 + m] * K[m * 3 + q];
            R[r * 3 + q] = ((r == q) ? 1.0 : 0.0) + s * K[r * 3 + q] + (1.0 - c) * kk;
        }
}

// ---------------------------------------------------------------- shared per-point steps
//
// The batched kernel (one workgroup per hypothesis, state in LDS) and the
// plain multi-kernel refiner (one hypothesis at a time, state in HBM) call
// the same functions below, so they perform the same arithmetic.

struct IcpState {
    double pose[12], cur[12], U[12];
    float  curF[12], UF[12], means[6];
    float  mse, rate;
    int    iters, type, conv, stop, valid;
};

__device__ void state_init(IcpState &st, const float *init) {
    bool ident = true;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 4; c++) {
            float v = init[r * 4 + c], e = (r == c) ? 1.0f : 0.0f;
            if (fabsf(v - e) > 1e-5f)
                ident = false;
        }
    for (int i = 0; i < 12; i++) {
        st.pose[i] = init[i];
        st.cur[i]  = ident ? ((i % 5 == 0) ? 1.0 : 0.0) : st.pose[i];
        st.curF[i] = (float)st.cur[i];
    }
    st.mse   = FLT_MAX;
    st.rate  = 1.0f;
    st.iters = 0;
    st.type  = 0;
    st.conv  = 0;
    st.stop  = 0;
    st.valid = 0;
}

__device__ __forceinline__ void xform(const float *T, float x, float y, float z, float &ox,
                                      float &oy, float &oz) {
    ox = T[0] * x + T[1] * y + T[2] * z + T[3];
    oy = T[4] * x + T[5] * y + T[6] * z + T[7];
    oz = T[8] * x + T[9] * y + T[10] * z + T[11];
}

struct HypoScratch {
    uint32_t           *hkey;
    unsigned long long *hbest;
    int                *nnSlot, *hslot, *wslot;
    float              *nnD;
    uint32_t            T;
};

// `table` selects one of the two dedupe tables of a hypothesis (the spread
// refiner alternates them between iterations so one can be cleared while
// the other is in use)
__device__ __forceinline__ HypoScratch scratch_of(const IcpArgs &a, int h, int table = 0) {
    HypoScratch x;
    size_t      base = ((size_t)table * a.H + h) * a.T;
    x.hkey   = a.hkey + base;
    x.hbest  = a.hbest + base;
    x.nnSlot = a.nnSlot + (size_t)h * a.nSrc;
    x.hslot  = a.hslot + (size_t)h * a.nSrc;
    x.wslot  = a.wslot + (size_t)h * a.nSrc;
    x.nnD    = a.nnD + (size_t)h * a.nSrc;
    x.T      = a.T;
    return x;
}

// findCorresponds for one source point (icp.cpp:123-167): nearest scene
// point, rejection beyond rejectDist, per-scene-point dedupe by 64-bit
// atomicMin on (dist2 bits, source index)
__device__ __forceinline__ void record_corr(const HypoScratch &x, int i, int s, float d2, int did,
                                            float rej2) {
    int hs = -1;
    if (s >= 0 && !(d2 > rej2)) {
        uint32_t key = (uint32_t)did + 1u;
        uint32_t hh  = icp_hash(key) & (x.T - 1);
        for (;;) {
            uint32_t prev = atomicCAS(&x.hkey[hh], 0u, key);
            if (prev == 0u || prev == key)
                break;
            hh = (hh + 1u) & (x.T - 1);
        }
        atomicMin(&x.hbest[hh], ((unsigned long long)__float_as_uint(d2) << 32) | (uint32_t)i);
        hs = (int)hh;
    } else {
        s = -1;
    }
    x.nnSlot[i] = s;
    x.hslot[i]  = hs;
    x.nnD[i]    = d2;
}

__device__ void corr_point(const IcpArgs &a, const HypoScratch &x, int i, const float *curF,
                           float rej2) {
    float px, py, pz;
    xform(curF, a.sx[i], a.sy[i], a.sz[i], px, py, pz);
    float d2;
    int   did;
    // the pair is rejected beyond rejectDist, so the search stops there
    int s = grid_nn1(a.dst, px, py, pz, d2, did, rej2);
    record_corr(x, i, s, d2, did, rej2);
}

// the same by a group of G lanes sharing the search; lane 0 records
template <int G>
__device__ void corr_point_group(const IcpArgs &a, const HypoScratch &x, int i, const float *curF,
                                 float rej2) {
    float px, py, pz;
    xform(curF, a.sx[i], a.sy[i], a.sz[i], px, py, pz);
    float d2;
    int   did;
    int   s = grid_nn1_group<G>(a.dst, px, py, pz, d2, did, rej2);
    if ((threadIdx.x & (G - 1)) == 0)
        record_corr(x, i, s, d2, did, rej2);
}

__device__ __forceinline__ bool is_winner(const HypoScratch &x, int i) {
    if (x.nnSlot[i] < 0)
        return false;
    unsigned long long pk = ((unsigned long long)__float_as_uint(x.nnD[i]) << 32) | (uint32_t)i;
    return x.hbest[x.hslot[i]] == pk;
}

// a winning pair's scene point (and -1 for every other source point) is
// kept in wslot for the iteration's later steps
__device__ __forceinline__ void add_means(const IcpArgs &a, const HypoScratch &x, int i,
                                          const float *curF, CF (&m7)[7]) {
    const bool win = is_winner(x, i);
    const int  s   = win ? x.nnSlot[i] : -1;
    x.wslot[i]     = s;
    if (!win)
        return;
    float px, py, pz;
    xform(curF, a.sx[i], a.sy[i], a.sz[i], px, py, pz);
    const float4 Q = a.dst.p4[s];
    cf_add(m7[0], px);
    cf_add(m7[1], py);
    cf_add(m7[2], pz);
    cf_add(m7[3], Q.x);
    cf_add(m7[4], Q.y);
    cf_add(m7[5], Q.z);
    cf_add(m7[6], 1.0f);
}

// one row of the normal equations: symmetric objective (icp.cpp:56-65) or
// point-to-plane about the source centroid
__device__ __forceinline__ void eq_row(const IcpArgs &a, int i, int s, const float *curF,
                                       const float *means, int mode, float J[6], float &r) {
    float px, py, pz;
    xform(curF, a.sx[i], a.sy[i], a.sz[i], px, py, pz);
    const float4 Q = a.dst.p4[s], NQ = a.dst.n4[s];
    float qx = Q.x, qy = Q.y, qz = Q.z;
    float nqx = NQ.x, nqy = NQ.y, nqz = NQ.z;
    float ux = px - means[0], uy = py - means[1], uz = pz - means[2];
    if (mode == 0) {
        float mx = a.snx[i], my = a.sny[i], mz = a.snz[i];
        float npx = curF[0] * mx + curF[1] * my + curF[2] * mz;
        float npy = curF[4] * mx + curF[5] * my + curF[6] * mz;
        float npz = curF[8] * mx + curF[9] * my + curF[10] * mz;
        float Nx = npx + nqx, Ny = npy + nqy, Nz = npz + nqz;
        float vx = qx - means[3], vy = qy - means[4], vz = qz - means[5];
        float sx = ux + vx, sy = uy + vy, sz = uz + vz;
        r    = (ux - vx) * Nx + (uy - vy) * Ny + (uz - vz) * Nz;
        J[0] = sy * Nz - sz * Ny;
        J[1] = sz * Nx - sx * Nz;
        J[2] = sx * Ny - sy * Nx;
        J[3] = Nx;
        J[4] = Ny;
        J[5] = Nz;
    } else {
        r    = (px - qx) * nqx + (py - qy) * nqy + (pz - qz) * nqz;
        J[0] = uy * nqz - uz * nqy;
        J[1] = uz * nqx - ux * nqz;
        J[2] = ux * nqy - uy * nqx;
        J[3] = nqx;
        J[4] = nqy;
        J[5] = nqz;
    }
}

__device__ __forceinline__ void add_eq(const IcpArgs &a, const HypoScratch &x, int i,
                                       const float *curF, const float *means, int mode,
                                       CF (&e27)[27]) {
    const int s = x.wslot[i];
    if (s < 0)
        return;
    float J[6], r;
    eq_row(a, i, s, curF, means, mode, J, r);
    int k = 0;
#pragma unroll
    for (int p = 0; p < 6; p++)
#pragma unroll
        for (int q = p; q < 6; q++)
            cf_add(e27[k++], J[p] * J[q]);
#pragma unroll
    for (int p = 0; p < 6; p++)
        cf_add(e27[21 + p], r * J[p]);
}

// distance of the updated point to its nearest scene point (icp.cpp:182-194).
// The point's previous partner lies at dist2(updated point, partner), so the
// true nearest neighbour is within that distance: the search is bounded by
// it and returns the same (dist2, id) as an unbounded one.
__device__ __forceinline__ void add_mse(const IcpArgs &a, const HypoScratch &x, int i,
                                        const float *curF, const float *UF, CF (&m1)[1]) {
    const int s = x.wslot[i];
    if (s < 0)
        return;
    float px, py, pz, qx, qy, qz;
    xform(curF, a.sx[i], a.sy[i], a.sz[i], px, py, pz);
    xform(UF, px, py, pz, qx, qy, qz);
    const float4 Q     = a.dst.p4[s];
    float        bound = ppfc::dist2(qx, qy, qz, Q.x, Q.y, Q.z);
    float     d2;
    int       did;
    if (!(bound < FLT_MAX))
        bound = FLT_MAX;
    grid_nn1(a.dst, qx, qy, qz, d2, did, bound);
    cf_add(m1[0], sqrtf(d2));
}

// the same by a group of G lanes sharing the search; lane 0 adds
template <int G>
__device__ __forceinline__ void add_mse_group(const IcpArgs &a, const HypoScratch &x, int i,
                                              const float *curF, const float *UF, CF (&m1)[1]) {
    const int s = x.wslot[i];
    if (s < 0)
        return;
    float px, py, pz, qx, qy, qz;
    xform(curF, a.sx[i], a.sy[i], a.sz[i], px, py, pz);
    xform(UF, px, py, pz, qx, qy, qz);
    const float4 Q     = a.dst.p4[s];
    float        bound = ppfc::dist2(qx, qy, qz, Q.x, Q.y, Q.z);
    float     d2;
    int       did;
    if (!(bound < FLT_MAX))
        bound = FLT_MAX;
    grid_nn1_group<G>(a.dst, qx, qy, qz, d2, did, bound);
    if ((threadIdx.x & (G - 1)) == 0)
        cf_add(m1[0], sqrtf(d2));
}

__device__ __forceinline__ void clear_point(const HypoScratch &x, int i) {
    int hs = x.hslot[i];
    if (hs >= 0) {
        x.hkey[hs]  = 0u;
        x.hbest[hs] = ~0ull;
    }
}

// FP64 solve + update (icp.cpp:67-86 for the symmetric objective; Rodrigues
// about the source centroid for point-to-plane)
__device__ __forceinline__ void solve_update(const double *red, const float *means, int mode,
                                             IcpState &st) {
    double A[36], rhs[6], u[6];
#pragma unroll
    for (int p = 0; p < 6; p++)
#pragma unroll
        for (int q = p; q < 6; q++) {
            const int k  = p * 6 - p * (p - 1) / 2 + (q - p);  // row-major upper triangle
            A[p * 6 + q] = red[k];
            A[q * 6 + p] = red[k];
        }
#pragma unroll
    for (int p = 0; p < 6; p++)
        rhs[p] = -red[21 + p];
    sym_solve6(A, rhs, u);
    double U[12];
    if (mode == 0) {
        double an    = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        double theta = atan(an);
        double ax[3] = {0, 0, 0};
        if (an > 0)
            for (int i = 0; i < 3; i++)
                ax[i] = u[i] / an;
        double W[9] = {0, -ax[2], ax[1], ax[2], 0, -ax[0], -ax[1], ax[0], 0}, W2[9], Ri[9], R[9];
        for (int r = 0; r < 3; r++)
            for (int q = 0; q < 3; q++) {
                double s2 = 0;
                for (int m = 0; m < 3; m++)
                    s2 += W[r * 3 + m] * W[m * 3 + q];
                W2[r * 3 + q] = s2;
            }
        double st_ = sin(theta), ct = cos(theta);
        for (int i = 0; i < 9; i++)
            Ri[i] = ((i % 4 == 0) ? 1.0 : 0.0) + st_ * W[i] + (1.0 - ct) * W2[i];
        for (int r = 0; r < 3; r++)
            for (int q = 0; q < 3; q++) {
                double s2 = 0;
                for (int m = 0; m < 3; m++)
                    s2 += Ri[r * 3 + m] * Ri[m * 3 + q];
                R[r * 3 + q] = s2;
            }
        double pm[3] = {means[0], means[1], means[2]}, qm[3] = {means[3], means[4], means[5]};
        double t2[3] = {ct * u[3], ct * u[4], ct * u[5]};
        for (int r = 0; r < 3; r++) {
            double v = 0;
            for (int m = 0; m < 3; m++)
                v += R[r * 3 + m] * (-pm[m]) + Ri[r * 3 + m] * t2[m];
            U[r * 4 + 0] = R[r * 3 + 0];
            U[r * 4 + 1] = R[r * 3 + 1];
            U[r * 4 + 2] = R[r * 3 + 2];
            U[r * 4 + 3] = v + qm[r];
        }
    } else {
        double R[9];
        rodrigues(u, R);
        double c[3] = {means[0], means[1], means[2]};
        for (int r = 0; r < 3; r++) {
            double Rc = R[r * 3 + 0] * c[0] + R[r * 3 + 1] * c[1] + R[r * 3 + 2] * c[2];
            U[r * 4 + 0] = R[r * 3 + 0];
            U[r * 4 + 1] = R[r * 3 + 1];
            U[r * 4 + 2] = R[r * 3 + 2];
            U[r * 4 + 3] = c[r] - Rc + u[3 + r];
        }
    }
    for (int i = 0; i < 12; i++) {
        st.U[i]  = U[i];
        st.UF[i] = (float)U[i];
    }
}

// ICP::regist's convergence rules for one iteration (icp.cpp:236-271)
__device__ void converge(IcpState &st, float mse, const IcpCriteria &c) {
    int vp = st.valid;
    if (vp > 6 && mse < c.mseMax)
        st.conv = 1;
    bool stop = false;
    if (mse < c.mseMin) {
        st.type = 1;
        stop    = true;
    }
    float rate = (st.mse - mse) / st.mse;
    if (rate < c.tolerance) {
        st.type = 2;
        stop    = true;
    }
    if (st.iters++ >= c.iterations) {
        st.type = 0;
        stop    = true;
    }
    if (rate > 0.0f) {
        st.mse  = mse;
        st.rate = rate;
        compose3x4(st.U, st.pose, st.pose);
    }
    if (!stop) {
        compose3x4(st.U, st.cur, st.cur);
        for (int i = 0; i < 12; i++)
            st.curF[i] = (float)st.cur[i];
    }
    st.stop = stop ? 1 : 0;
}

__device__ __forceinline__ void write_result(IcpResultDev &o, const IcpState &st, int inl) {
    for (int i = 0; i < 12; i++)
        o.pose[i] = (float)st.pose[i];
    o.type         = st.type;
    o.mse          = st.mse;
    o.convergeRate = st.rate;
    o.iterations   = st.iters;
    o.converged    = st.conv;
    o.inliers      = inl;
}

__device__ __forceinline__ bool inlier_point(const IcpArgs &a, int i, const float *PF, float r2) {
    float px, py, pz;
    xform(PF, a.sx[i], a.sy[i], a.sz[i], px, py, pz);
    return grid_any_within(a.full, px, py, pz, r2);
}

// ---------------------------------------------------------------- batched

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) icp_kernel(IcpArgs a0) {
    const int h = blockIdx.x;
    if (h >= a0.H)
        return;
    const IcpArgs a = args_for(a0, h);
    __shared__ IcpState st;
    __shared__ double   sRed[40];
    __shared__ CF       sLds[(BLOCK / 64) * 27];
    __shared__ int      sCnt[BLOCK / 64];
    __shared__ float    sPF[12];
    const int         tid = threadIdx.x;
    const int         n   = a.nSrc;
    const HypoScratch x   = scratch_of(a, h);
    if (tid == 0)
        state_init(st, a.init + (size_t)h * 12);
    for (uint32_t t = tid; t < a.T; t += BLOCK) {
        x.hkey[t]  = 0u;
        x.hbest[t] = ~0ull;
    }
    __syncthreads();
    const float rej2 = a.c.rejectDist * a.c.rejectDist;
    while (!st.stop && st.iters < a.c.iterations) {
        for (int i = tid; i < n; i += BLOCK)
            corr_point(a, x, i, st.curF, rej2);
        __syncthreads();
        CF m7[7];
        for (int k = 0; k < 7; k++)
            m7[k] = CF{0.0f, 0.0f};
        for (int i = tid; i < n; i += BLOCK)
            add_means(a, x, i, st.curF, m7);
        block_reduce_cf<BLOCK, 7>(m7, sLds, sRed);
        const int valid = (int)sRed[6];
        if (valid < 6) {
            if (tid == 0) {
                st.type = 3;  // NO_CORRESPONDS
                st.stop = 1;
            }
            __syncthreads();
            break;
        }
        if (tid < 6)
            st.means[tid] = (float)(sRed[tid] / valid);
        __syncthreads();
        CF e27[27];
        for (int k = 0; k < 27; k++)
            e27[k] = CF{0.0f, 0.0f};
        for (int i = tid; i < n; i += BLOCK)
            add_eq(a, x, i, st.curF, st.means, a.mode, e27);
        block_reduce_cf<BLOCK, 27>(e27, sLds, sRed);
        if (tid == 0) {
            solve_update(sRed, st.means, a.mode, st);
            st.valid = valid;
        }
        __syncthreads();
        CF m1[1] = {CF{0.0f, 0.0f}};
        for (int i = tid; i < n; i += BLOCK)
            add_mse(a, x, i, st.curF, st.UF, m1);
        block_reduce_cf<BLOCK, 1>(m1, sLds, sRed);
        for (int i = tid; i < n; i += BLOCK)
            clear_point(x, i);
        if (tid == 0)
            converge(st, (float)(sRed[0] / (double)st.valid), a.c);
        __syncthreads();
    }
    if (tid < 12)
        sPF[tid] = (float)st.pose[tid];
    __syncthreads();
    float r2  = a.scoringDist * a.scoringDist;
    int   cnt = 0;
    for (int i = tid; i < n; i += BLOCK)
        cnt += inlier_point(a, i, sPF, r2) ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1)
        cnt += __shfl_xor(cnt, o, 64);
    if ((tid & 63) == 0)
        sCnt[tid >> 6] = cnt;
    __syncthreads();
    if (tid == 0) {
        int tot = 0;
        for (int w = 0; w < BLOCK / 64; w++)
            tot += sCnt[w];
        write_result(a.out[h], st, tot);
    }
}

// ---------------------------------------------------------------- spread
//
// The spread batched refiner: every hypothesis is cut into chunks of BLOCK
// source points and one launch covers all (hypothesis, chunk) pairs, so a
// few hypotheses over many points fill the chip as well as many over few.
// An iteration is four launches (correspondences, means, normal equations,
// mse); the serial step of each (means, FP64 solve, convergence) is done by
// the last workgroup of the hypothesis to finish, which reduces the
// per-chunk compensated partials in chunk order (deterministic).  States,
// counters and stop flags stay in HBM; a stopped hypothesis costs one load
// per workgroup.  The launch sequence is fixed by the iteration cap, so it
// is captured once into a hipGraph and replayed.

struct SpreadArgs {
    IcpArgs   b;
    IcpState *st;      // H
    uint32_t *arrive;  // H arrival counters, zero between launches
    CF       *part;    // H x nChunk x 27 per-chunk partial sums
    int      *inl;     // H inlier counts
    int       nChunk;  // workgroups per hypothesis, one point per lane
    int       nChunkG; // workgroups per hypothesis, one point per ICP_G lanes
};

// lanes sharing one nearest-neighbour search in the spread refiner
constexpr int ICP_G = 8;

__device__ __forceinline__ bool hypo_active(const IcpState &st, const IcpCriteria &c) {
    return !st.stop && st.iters < c.iterations;
}

// Publish this workgroup's M partial sums; returns true (uniformly) in the
// last workgroup of hypothesis h to arrive, with the chunk-ordered totals in
// red[0..M).
template <int BLOCK, int M>
__device__ bool arrive_reduce(const SpreadArgs &w, int h, int chunk, int nParts, CF (&v)[M],
                              CF *lds, double *red, int *sLast) {
#pragma unroll
    for (int k = 0; k < M; k++)
        for (int o = 1; o < 64; o <<= 1) {
            CF u;
            u.hi = __shfl_xor(v[k].hi, o, 64);
            u.lo = __shfl_xor(v[k].lo, o, 64);
            v[k] = ((threadIdx.x & o) == 0) ? cf_merge(v[k], u) : cf_merge(u, v[k]);
        }
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < M; k++)
            lds[wid * M + k] = v[k];
    __syncthreads();
    CF *mine = w.part + ((size_t)h * nParts + chunk) * 27;
    if (threadIdx.x < M) {
        CF acc = lds[threadIdx.x];
        for (int q = 1; q < BLOCK / 64; q++)
            acc = cf_merge(acc, lds[q * M + threadIdx.x]);
        mine[threadIdx.x] = acc;
        __threadfence();
    }
    __syncthreads();
    if (threadIdx.x == 0)
        *sLast = (atomicAdd(&w.arrive[h], 1u) == (uint32_t)(nParts - 1)) ? 1 : 0;
    __syncthreads();
    if (!*sLast)
        return false;
    __threadfence();
    if (threadIdx.x < M) {
        const CF *all = w.part + (size_t)h * nParts * 27;
        CF        acc = all[threadIdx.x];
        for (int q = 1; q < nParts; q++)
            acc = cf_merge(acc, all[(size_t)q * 27 + threadIdx.x]);
        red[threadIdx.x] = (double)acc.hi + (double)acc.lo;
    }
    if (threadIdx.x == 0)
        w.arrive[h] = 0u;
    __syncthreads();
    return true;
}

__global__ void spread_init_kernel(SpreadArgs w) {
    const size_t gs = (size_t)gridDim.x * blockDim.x;
    const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t h = t0; h < (size_t)w.b.H; h += gs) {
        state_init(w.st[h], w.b.init + h * 12);
        w.arrive[h] = 0u;
        w.inl[h]    = 0;
    }
    const size_t nt = (size_t)2 * w.b.H * w.b.T;
    for (size_t t = t0; t < nt; t += gs) {
        w.b.hkey[t]  = 0u;
        w.b.hbest[t] = ~0ull;
    }
}

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) spread_corr_kernel(SpreadArgs w) {
    const int h = blockIdx.x / w.nChunkG, chunk = blockIdx.x - h * w.nChunkG;
    const IcpArgs b = args_for(w.b, h);
    __shared__ float sCur[12];
    __shared__ int   sTab;
    const IcpState  &st = w.st[h];
    if (!hypo_active(st, b.c))
        return;
    if (threadIdx.x < 12)
        sCur[threadIdx.x] = st.curF[threadIdx.x];
    if (threadIdx.x == 0)
        sTab = st.iters & 1;
    __syncthreads();
    const int i = chunk * (BLOCK / ICP_G) + (int)threadIdx.x / ICP_G;
    if (i < b.nSrc)
        corr_point_group<ICP_G>(b, scratch_of(b, h, sTab), i, sCur,
                                b.c.rejectDist * b.c.rejectDist);
}

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) spread_means_kernel(SpreadArgs w) {
    const int h = blockIdx.x / w.nChunk, chunk = blockIdx.x - h * w.nChunk;
    const IcpArgs b = args_for(w.b, h);
    __shared__ float  sCur[12];
    __shared__ int    sTab, sLast;
    __shared__ CF     sLds[(BLOCK / 64) * 7];
    __shared__ double sRed[7];
    IcpState &st = w.st[h];
    if (!hypo_active(st, b.c))
        return;
    if (threadIdx.x < 12)
        sCur[threadIdx.x] = st.curF[threadIdx.x];
    if (threadIdx.x == 0)
        sTab = st.iters & 1;
    __syncthreads();
    // the other table was last used by the previous iteration: clear this
    // chunk's share of it for the next one
    {
        const HypoScratch o   = scratch_of(b, h, sTab ^ 1);
        const uint32_t    per = (b.T + w.nChunk - 1) / w.nChunk;
        const uint32_t    lo = per * chunk, hi = min(b.T, lo + per);
        for (uint32_t t = lo + threadIdx.x; t < hi; t += BLOCK) {
            o.hkey[t]  = 0u;
            o.hbest[t] = ~0ull;
        }
    }
    CF m7[7];
    for (int k = 0; k < 7; k++)
        m7[k] = CF{0.0f, 0.0f};
    const int i = chunk * BLOCK + threadIdx.x;
    if (i < b.nSrc)
        add_means(b, scratch_of(b, h, sTab), i, sCur, m7);
    if (!arrive_reduce<BLOCK, 7>(w, h, chunk, w.nChunk, m7, sLds, sRed, &sLast))
        return;
    if (threadIdx.x == 0) {
        const int valid = (int)sRed[6];
        st.valid        = valid;
        if (valid < 6) {
            st.type = 3;  // NO_CORRESPONDS
            st.stop = 1;
        } else {
            for (int k = 0; k < 6; k++)
                st.means[k] = (float)(sRed[k] / valid);
        }
    }
}

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) spread_eq_kernel(SpreadArgs w) {
    const int h = blockIdx.x / w.nChunk, chunk = blockIdx.x - h * w.nChunk;
    const IcpArgs b = args_for(w.b, h);
    __shared__ float  sCur[12], sMeans[6];
    __shared__ int    sTab, sLast;
    __shared__ CF     sLds[(BLOCK / 64) * 27];
    __shared__ double sRed[27];
    IcpState &st = w.st[h];
    if (!hypo_active(st, b.c))
        return;
    if (threadIdx.x < 12)
        sCur[threadIdx.x] = st.curF[threadIdx.x];
    if (threadIdx.x < 6)
        sMeans[threadIdx.x] = st.means[threadIdx.x];
    if (threadIdx.x == 0)
        sTab = st.iters & 1;
    __syncthreads();
    CF e27[27];
    for (int k = 0; k < 27; k++)
        e27[k] = CF{0.0f, 0.0f};
    const int i = chunk * BLOCK + threadIdx.x;
    if (i < b.nSrc)
        add_eq(b, scratch_of(b, h, sTab), i, sCur, sMeans, b.mode, e27);
    if (!arrive_reduce<BLOCK, 27>(w, h, chunk, w.nChunk, e27, sLds, sRed, &sLast))
        return;
    if (threadIdx.x == 0)
        solve_update(sRed, sMeans, b.mode, st);
}

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) spread_mse_kernel(SpreadArgs w) {
    const int h = blockIdx.x / w.nChunkG, chunk = blockIdx.x - h * w.nChunkG;
    const IcpArgs b = args_for(w.b, h);
    __shared__ float  sCur[12], sU[12];
    __shared__ int    sTab, sLast;
    __shared__ CF     sLds[BLOCK / 64];
    __shared__ double sRed[1];
    IcpState &st = w.st[h];
    if (!hypo_active(st, b.c))
        return;
    if (threadIdx.x < 12) {
        sCur[threadIdx.x] = st.curF[threadIdx.x];
        sU[threadIdx.x]   = st.UF[threadIdx.x];
    }
    if (threadIdx.x == 0)
        sTab = st.iters & 1;
    __syncthreads();
    CF        m1[1] = {CF{0.0f, 0.0f}};
    const int i     = chunk * (BLOCK / ICP_G) + (int)threadIdx.x / ICP_G;
    if (i < b.nSrc)
        add_mse_group<ICP_G>(b, scratch_of(b, h, sTab), i, sCur, sU, m1);
    if (!arrive_reduce<BLOCK, 1>(w, h, chunk, w.nChunkG, m1, sLds, sRed, &sLast))
        return;
    if (threadIdx.x == 0)
        converge(st, (float)(sRed[0] / (double)st.valid), b.c);
}

// inlier count of the final poses: ICP_G lanes per source point share the
// search (lane 0 of each group counts)
template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) spread_inlier_kernel(SpreadArgs w) {
    const int h = blockIdx.x / w.nChunkG, chunk = blockIdx.x - h * w.nChunkG;
    const IcpArgs b = args_for(w.b, h);
    __shared__ float sPF[12];
    __shared__ int   sCnt[BLOCK / 64];
    const IcpState  &st = w.st[h];
    if (threadIdx.x < 12)
        sPF[threadIdx.x] = (float)st.pose[threadIdx.x];
    __syncthreads();
    const int i = chunk * (BLOCK / ICP_G) + (int)threadIdx.x / ICP_G;
    int       c = 0;
    if (i < b.nSrc) {
        float px, py, pz;
        xform(sPF, b.sx[i], b.sy[i], b.sz[i], px, py, pz);
        const bool in = grid_any_within_group<ICP_G>(b.full, px, py, pz,
                                                     b.scoringDist * b.scoringDist);
        c = (in && (threadIdx.x & (ICP_G - 1)) == 0) ? 1 : 0;
    }
    for (int o = 32; o > 0; o >>= 1)
        c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0)
        sCnt[threadIdx.x >> 6] = c;
    __syncthreads();
    __shared__ int sLast;
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int q = 0; q < BLOCK / 64; q++)
            tot += sCnt[q];
        atomicAdd(&w.inl[h], tot);
        __threadfence();
        sLast = (atomicAdd(&w.arrive[h], 1u) == (uint32_t)(w.nChunkG - 1)) ? 1 : 0;
        if (sLast) {
            __threadfence();
            write_result(b.out[h], st, atomicAdd(&w.inl[h], 0));
            w.inl[h]    = 0;
            w.arrive[h] = 0u;
        }
    }
}

// ---------------------------------------------------------------- plain
//
// The conventional GPU refiner: one hypothesis at a time, each step a grid
// of 256-thread blocks over the source points with per-block partial sums,
// and a host check of the stop flag after every iteration.

constexpr int PLAIN_BLOCK = 256;

__global__ void __launch_bounds__(PLAIN_BLOCK) plain_init_kernel(IcpArgs a, IcpState *st) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        state_init(*st, a.init);
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < a.T; t += gridDim.x * blockDim.x) {
        a.hkey[t]  = 0u;
        a.hbest[t] = ~0ull;
    }
}

__global__ void __launch_bounds__(PLAIN_BLOCK) plain_corr_kernel(IcpArgs a, const IcpState *st) {
    int i = blockIdx.x * PLAIN_BLOCK + threadIdx.x;
    if (i < a.nSrc)
        corr_point(a, scratch_of(a, 0), i, st->curF, a.c.rejectDist * a.c.rejectDist);
}

template <int PASS>  // 0 means, 1 normal equations, 2 mse
__global__ void __launch_bounds__(PLAIN_BLOCK) plain_sum_kernel(IcpArgs a, const IcpState *st,
                                                                double *partial) {
    __shared__ CF     sLds[(PLAIN_BLOCK / 64) * 27];
    __shared__ double sRed[40];
    int               i = blockIdx.x * PLAIN_BLOCK + threadIdx.x;
    HypoScratch       x = scratch_of(a, 0);
    if (PASS == 0) {
        CF v[7];
        for (int k = 0; k < 7; k++)
            v[k] = CF{0.0f, 0.0f};
        if (i < a.nSrc)
            add_means(a, x, i, st->curF, v);
        block_reduce_cf<PLAIN_BLOCK, 7>(v, sLds, sRed);
        if (threadIdx.x < 7)
            partial[blockIdx.x * 27 + threadIdx.x] = sRed[threadIdx.x];
    } else if (PASS == 1) {
        CF v[27];
        for (int k = 0; k < 27; k++)
            v[k] = CF{0.0f, 0.0f};
        if (i < a.nSrc)
            add_eq(a, x, i, st->curF, st->means, a.mode, v);
        block_reduce_cf<PLAIN_BLOCK, 27>(v, sLds, sRed);
        if (threadIdx.x < 27)
            partial[blockIdx.x * 27 + threadIdx.x] = sRed[threadIdx.x];
    } else {
        CF v[1] = {CF{0.0f, 0.0f}};
        if (i < a.nSrc)
            add_mse(a, x, i, st->curF, st->UF, v);
        block_reduce_cf<PLAIN_BLOCK, 1>(v, sLds, sRed);
        if (threadIdx.x == 0)
            partial[blockIdx.x * 27] = sRed[0];
    }
}

// reduce per-block partials in block order, then the serial step of PASS
template <int PASS>
__global__ void plain_serial_kernel(IcpArgs a, IcpState *st, const double *partial, int nb) {
    if (threadIdx.x != 0)
        return;
    double red[27];
    int    M = PASS == 0 ? 7 : (PASS == 1 ? 27 : 1);
    for (int k = 0; k < M; k++) {
        double s = 0.0;
        for (int b = 0; b < nb; b++)
            s += partial[b * 27 + k];
        red[k] = s;
    }
    if (PASS == 0) {
        int valid = (int)red[6];
        st->valid = valid;
        if (valid < 6) {
            st->type = 3;
            st->stop = 1;
            return;
        }
        for (int k = 0; k < 6; k++)
            st->means[k] = (float)(red[k] / valid);
    } else if (PASS == 1) {
        solve_update(red, st->means, a.mode, *st);
    } else {
        converge(*st, (float)(red[0] / (double)st->valid), a.c);
    }
}

__global__ void __launch_bounds__(PLAIN_BLOCK) plain_clear_kernel(IcpArgs a) {
    int i = blockIdx.x * PLAIN_BLOCK + threadIdx.x;
    if (i < a.nSrc)
        clear_point(scratch_of(a, 0), i);
}

__global__ void __launch_bounds__(PLAIN_BLOCK) plain_inlier_kernel(IcpArgs a, const IcpState *st,
                                                                   int *count) {
    int   i = blockIdx.x * PLAIN_BLOCK + threadIdx.x;
    float PF[12];
    for (int k = 0; k < 12; k++)
        PF[k] = (float)st->pose[k];
    int c = (i < a.nSrc && inlier_point(a, i, PF, a.scoringDist * a.scoringDist)) ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1)
        c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0 && c)
        atomicAdd(count, c);
}

__global__ void plain_result_kernel(const IcpState *st, const int *count, IcpResultDev *out) {
    if (threadIdx.x == 0)
        write_result(*out, *st, *count);
}

// ---------------------------------------------------------------- host

static_assert(sizeof(IcpArgs) <= sizeof(GraphCache::args), "graph key too small");

static IcpArgs make_args(Workspace &ws, int nSrc, const float *sx, const float *sy,
                         const float *sz, const float *snx, const float *sny, const float *snz,
                         const GridView &dst, const GridView &full, int H, const IcpCriteria &c,
                         float scoringDist, int mode) {
    if (c.iterations < 1 || c.mseMin < 0 || c.mseMax < c.mseMin || c.tolerance > 1 ||
        c.tolerance < 0 || c.rejectDist < 0)
        throw std::runtime_error("Invalid ConvergenceCriteria at ICP::regist");
    if (!dst.nx)
        throw std::runtime_error("PointCloud empty or no normal at ICP::regist");
    IcpArgs a;
    std::memset(&a, 0, sizeof a);  // padding too: the bytes key the graph cache
    a.nSrc = nSrc;
    a.H    = H;
    a.sx = sx;
    a.sy = sy;
    a.sz = sz;
    a.snx = snx;
    a.sny = sny;
    a.snz = snz;
    a.dst  = dst;
    a.full = full;
    a.c    = c;
    a.scoringDist = scoringDist;
    a.mode = mode;
    uint32_t T = 64;
    while (T < 2u * (uint32_t)std::max(nSrc, 1))
        T <<= 1;
    a.T      = T;
    // two dedupe tables per hypothesis (the spread refiner alternates them)
    a.hkey   = ws.icpHash.as<uint32_t>((size_t)2 * H * T);
    a.hbest  = ws.icpPartial.as<unsigned long long>((size_t)2 * H * T);
    int *ib  = ws.icpState.as<int>((size_t)H * nSrc * 4);
    a.nnSlot = ib;
    a.hslot  = ib + (size_t)H * nSrc;
    a.nnD    = reinterpret_cast<float *>(ib + 2 * (size_t)H * nSrc);
    a.wslot  = ib + 3 * (size_t)H * nSrc;
    a.init   = nullptr;
    a.out    = nullptr;
    return a;
}

static void launch_icp(const IcpArgs &a, int block, hipStream_t s) {
    switch (block) {
        case 64: icp_kernel<64><<<a.H, 64, 0, s>>>(a); break;
        case 128: icp_kernel<128><<<a.H, 128, 0, s>>>(a); break;
        case 512: icp_kernel<512><<<a.H, 512, 0, s>>>(a); break;
        case 1024: icp_kernel<1024><<<a.H, 1024, 0, s>>>(a); break;
        default: icp_kernel<256><<<a.H, 256, 0, s>>>(a); break;
    }
}

template <int BLOCK>
static void launch_spread_t(const SpreadArgs &w, hipStream_t s) {
    const long long nb = (long long)w.b.H * w.nChunkG;
    if (nb > 0x7fffffffLL)
        throw std::runtime_error("ICP batch too large for one launch");
    const int g  = (int)((long long)w.b.H * w.nChunk);
    const int gG = (int)nb;
    spread_init_kernel<<<std::min(4096, div_up((long long)2 * w.b.H * w.b.T, 256)), 256, 0, s>>>(w);
    for (int it = 0; it < w.b.c.iterations; it++) {
        spread_corr_kernel<BLOCK><<<gG, BLOCK, 0, s>>>(w);
        spread_means_kernel<BLOCK><<<g, BLOCK, 0, s>>>(w);
        spread_eq_kernel<BLOCK><<<g, BLOCK, 0, s>>>(w);
        spread_mse_kernel<BLOCK><<<gG, BLOCK, 0, s>>>(w);
    }
    spread_inlier_kernel<BLOCK><<<gG, BLOCK, 0, s>>>(w);
}

static void launch_spread(const SpreadArgs &w, int block, hipStream_t s) {
    switch (block) {
        case 64: launch_spread_t<64>(w, s); break;
        case 128: launch_spread_t<128>(w, s); break;
        case 512: launch_spread_t<512>(w, s); break;
        case 1024: launch_spread_t<1024>(w, s); break;
        default: launch_spread_t<256>(w, s); break;
    }
}

static int norm_block(int b) {
    return (b == 64 || b == 128 || b == 512 || b == 1024) ? b : 256;
}

// the refiner's searches read the grids' {x, y, z, id} copies
static void require_copies(const GridView &g) {
    if (g.n > 0 && (!g.p4 || (g.nx && !g.n4)))
        throw std::logic_error("refinement grid built without its point copies");
}

void icp_batch(Workspace &ws, int nSrc, const float *sx, const float *sy, const float *sz,
               const float *snx, const float *sny, const float *snz, const GridView &dst,
               const GridView &full, const float *init12, int H, const IcpCriteria &c,
               float scoringDist, const IcpConfig &cfg, IcpResultDev *out, hipStream_t s,
               const GridView *dstG, const GridView *fullG, const int *hypScene) {
    if (H <= 0)
        return;
    require_copies(dst);
    require_copies(full);
    const int block = norm_block(cfg.block);
    IcpArgs a = make_args(ws, nSrc, sx, sy, sz, snx, sny, snz, dst, full, H, c, scoringDist,
                          cfg.mode);
    a.dstG     = dstG;
    a.fullG    = fullG;
    a.hypScene = hypScene;
    float *init = ws.icpPose.as<float>((size_t)H * 12);
    PPF_HIP_CHECK(hipMemcpyAsync(init, init12, sizeof(float) * 12 * H, hipMemcpyHostToDevice, s));
    a.init = init;
    IcpResultDev *od = ws.icpOut.as<IcpResultDev>(H);
    a.out = od;
    SpreadArgs w;
    std::memset(&w, 0, sizeof w);
    if (cfg.spread) {
        w.b      = a;
        w.nChunk  = div_up(std::max(nSrc, 1), block);
        w.nChunkG = div_up(std::max(nSrc, 1), block / ICP_G);
        w.st     = ws.icpSt.as<IcpState>(H);
        w.arrive = ws.icpArrive.as<uint32_t>((size_t)2 * H);
        w.inl    = reinterpret_cast<int *>(w.arrive + H);
        w.part   = ws.icpPart.as<CF>((size_t)H * w.nChunkG * 27);
    }
    auto launch = [&] {
        if (cfg.spread)
            launch_spread(w, block, s);
        else
            launch_icp(a, block, s);
    };
    if (cfg.graph) {
        // the launch sequence captured once per shape and argument set, replayed after
        static_assert(sizeof(SpreadArgs) <= sizeof(GraphCache::args), "graph key too small");
        unsigned char key[sizeof(GraphCache::args)] = {};
        if (cfg.spread)
            std::memcpy(key, &w, sizeof w);
        else
            std::memcpy(key, &a, sizeof a);
        GraphCache *gc = nullptr;
        for (auto &e : ws.icpGraph)
            if (e.H == H && e.nSrc == nSrc && e.block == block && e.spread == cfg.spread &&
                std::memcmp(e.args, key, sizeof key) == 0)
                gc = &e;
        if (!gc) {
            // a key seen for the first time is only remembered and launched
            // directly: capturing and instantiating costs more than one
            // direct launch sequence, and pays only for keys that come back
            // (the same scene matched again)
            gc = &ws.icpGraph[ws.icpGraphNext];
            ws.icpGraphNext = (ws.icpGraphNext + 1) % (int)(sizeof ws.icpGraph / sizeof ws.icpGraph[0]);
            if (gc->exec)
                PPF_HIP_CHECK(hipGraphExecDestroy(gc->exec));
            gc->exec   = nullptr;
            gc->H      = H;
            gc->nSrc   = nSrc;
            gc->block  = block;
            gc->spread = cfg.spread;
            std::memcpy(gc->args, key, sizeof key);
            launch();
        } else if (!gc->exec) {
            hipGraph_t g;
            PPF_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            launch();
            PPF_HIP_CHECK(hipStreamEndCapture(s, &g));
            PPF_HIP_CHECK(hipGraphInstantiate(&gc->exec, g, nullptr, nullptr, 0));
            PPF_HIP_CHECK(hipGraphDestroy(g));
        }
        if (gc->exec)
            PPF_HIP_CHECK(hipGraphLaunch(gc->exec, s));
    } else {
        launch();
    }
    PPF_HIP_CHECK(hipGetLastError());
    if (out) {
        PPF_HIP_CHECK(hipMemcpyAsync(out, od, sizeof(IcpResultDev) * H, hipMemcpyDeviceToHost, s));
        PPF_HIP_CHECK(hipStreamSynchronize(s));
    }
}

void icp_plain(Workspace &ws, int nSrc, const float *sx, const float *sy, const float *sz,
               const float *snx, const float *sny, const float *snz, const GridView &dst,
               const GridView &full, const float *init12, int H, const IcpCriteria &c,
               float scoringDist, int mode, IcpResultDev *out, hipStream_t s) {
    if (H <= 0)
        return;
    require_copies(dst);
    require_copies(full);
    IcpArgs a = make_args(ws, nSrc, sx, sy, sz, snx, sny, snz, dst, full, 1, c, scoringDist, mode);
    int nb = div_up(std::max(nSrc, 1), PLAIN_BLOCK);
    float *init = ws.icpPose.as<float>((size_t)H * 12);
    PPF_HIP_CHECK(hipMemcpyAsync(init, init12, sizeof(float) * 12 * H, hipMemcpyHostToDevice, s));
    IcpState     *st      = reinterpret_cast<IcpState *>(ws.icpCur.as<IcpState>(1));
    double       *partial = ws.icpInl.as<double>((size_t)nb * 27 + 32);
    int          *count   = reinterpret_cast<int *>(partial + (size_t)nb * 27);
    IcpResultDev *od      = ws.icpOut.as<IcpResultDev>(H);
    for (int h = 0; h < H; h++) {
        a.init = init + 12 * (size_t)h;
        plain_init_kernel<<<std::max(1, div_up(a.T, PLAIN_BLOCK)), PLAIN_BLOCK, 0, s>>>(a, st);
        for (int it = 0; it < c.iterations; it++) {
            plain_corr_kernel<<<nb, PLAIN_BLOCK, 0, s>>>(a, st);
            plain_sum_kernel<0><<<nb, PLAIN_BLOCK, 0, s>>>(a, st, partial);
            plain_serial_kernel<0><<<1, 64, 0, s>>>(a, st, partial, nb);
            int stop = 0;
            PPF_HIP_CHECK(hipMemcpyAsync(&stop, &st->stop, sizeof(int), hipMemcpyDeviceToHost, s));
            PPF_HIP_CHECK(hipStreamSynchronize(s));
            if (stop)
                break;
            plain_sum_kernel<1><<<nb, PLAIN_BLOCK, 0, s>>>(a, st, partial);
            plain_serial_kernel<1><<<1, 64, 0, s>>>(a, st, partial, nb);
            plain_sum_kernel<2><<<nb, PLAIN_BLOCK, 0, s>>>(a, st, partial);
            plain_clear_kernel<<<nb, PLAIN_BLOCK, 0, s>>>(a);
            plain_serial_kernel<2><<<1, 64, 0, s>>>(a, st, partial, nb);
            PPF_HIP_CHECK(hipMemcpyAsync(&stop, &st->stop, sizeof(int), hipMemcpyDeviceToHost, s));
            PPF_HIP_CHECK(hipStreamSynchronize(s));
            if (stop)
                break;
        }
        PPF_HIP_CHECK(hipMemsetAsync(count, 0, sizeof(int), s));
        plain_inlier_kernel<<<nb, PLAIN_BLOCK, 0, s>>>(a, st, count);
        plain_result_kernel<<<1, 64, 0, s>>>(st, count, od + h);
    }
    PPF_HIP_CHECK(hipGetLastError());
    if (out) {
        PPF_HIP_CHECK(hipMemcpyAsync(out, od, sizeof(IcpResultDev) * H, hipMemcpyDeviceToHost, s));
        PPF_HIP_CHECK(hipStreamSynchronize(s));
    }
}

// ---------------------------------------------------------------- scoring

// The refiner's scoring step (inlier_point) for given poses: blockIdx.y is
// the pose, x the chunk of source points.
__global__ void __launch_bounds__(256) inliers_kernel(GridView full, const float *sx,
                                                      const float *sy, const float *sz, int nSrc,
                                                      const float *poses12, float r2, int *out) {
    const int   h = blockIdx.y;
    __shared__ float sPF[12];
    __shared__ int   sCnt[4];
    if (threadIdx.x < 12)
        sPF[threadIdx.x] = poses12[12 * h + threadIdx.x];
    __syncthreads();
    const int i = blockIdx.x * 256 + threadIdx.x;
    int       c = 0;
    if (i < nSrc) {
        float px, py, pz;
        xform(sPF, sx[i], sy[i], sz[i], px, py, pz);
        c = grid_any_within(full, px, py, pz, r2) ? 1 : 0;
    }
    for (int o = 32; o > 0; o >>= 1)
        c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0)
        sCnt[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicAdd(&out[h], sCnt[0] + sCnt[1] + sCnt[2] + sCnt[3]);
}

void icp_inliers(Workspace &ws, const GridView &full, int nSrc, const float *sx, const float *sy,
                 const float *sz, const float *poses12Host, int H, float dist, int *inliersHost,
                 hipStream_t s) {
    if (H <= 0)
        return;
    if (H > 65535)
        throw std::runtime_error("at most 65535 poses per scoring call");
    require_copies(full);
    float *P   = ws.icpPose.as<float>((size_t)H * 12);
    int   *cnt = ws.icpInl.as<int>(H);
    PPF_HIP_CHECK(hipMemcpyAsync(P, poses12Host, sizeof(float) * 12 * H, hipMemcpyHostToDevice, s));
    PPF_HIP_CHECK(hipMemsetAsync(cnt, 0, sizeof(int) * H, s));
    if (nSrc > 0) {
        dim3 grid(div_up(nSrc, 256), H);
        inliers_kernel<<<grid, 256, 0, s>>>(full, sx, sy, sz, nSrc, P, dist * dist, cnt);
        PPF_HIP_CHECK(hipGetLastError());
    }
    PPF_HIP_CHECK(hipMemcpyAsync(inliersHost, cnt, sizeof(int) * H, hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------- adjoint
//
// Point-to-plane cost E = 1/2 sum_i r_i^2, r_i = n_q.(x_i - q_i), over the
// refiner's correspondences at the pose T = U(xi) T0, where
// U(xi) = [R(w), v] (Rodrigues, xi = (w, v)).  The kernel returns J^T r with
// J_i = [(x_i - c) x n_q ; n_q] -- the Jacobian rows of the reference's
// point-to-plane solver (icp.cpp:318-320) -- the gradient with respect to a
// perturbation [R(om), c - R(om) c + u] applied on the left of T about the
// centre c (the origin or the winners' centroid), and J^T J, the
// Gauss-Newton matrix the refiner solves with (mode 1).

struct AdjArgs {
    IcpArgs       b;
    const double *xi;      // H x 6
    const float  *T0;      // H x 12
    int           center;  // 0 origin, 1 centroid of the winners
    double       *out;     // H x 36: cost, count, c[3], grad[6], JtJ[21], pad[4]
    int          *corr;    // H x nSrc dst id of each winner (-1 = none), optional
};

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) adjoint_kernel(AdjArgs a) {
    const int h = blockIdx.x;
    if (h >= a.b.H)
        return;
    __shared__ float  sT[12], sC[3];
    __shared__ double sRed[40];
    __shared__ CF     sLds[(BLOCK / 64) * 28];
    const int         tid = threadIdx.x;
    const int         n   = a.b.nSrc;
    const HypoScratch x   = scratch_of(a.b, h);
    if (tid == 0) {
        // T = U(xi) T0 (U about the origin; the centre c only enters the
        // perturbation the gradient is taken in)
        double w[3] = {a.xi[h * 6 + 0], a.xi[h * 6 + 1], a.xi[h * 6 + 2]}, R[9];
        rodrigues(w, R);
        double U[12], T0[12], T[12];
        for (int i = 0; i < 12; i++)
            T0[i] = a.T0[(size_t)h * 12 + i];
        for (int r = 0; r < 3; r++) {
            for (int c = 0; c < 3; c++)
                U[r * 4 + c] = R[r * 3 + c];
            U[r * 4 + 3] = a.xi[h * 6 + 3 + r];
        }
        compose3x4(U, T0, T);
        for (int i = 0; i < 12; i++)
            sT[i] = (float)T[i];
    }
    for (uint32_t t = tid; t < a.b.T; t += BLOCK) {
        x.hkey[t]  = 0u;
        x.hbest[t] = ~0ull;
    }
    __syncthreads();
    const float rej2 = a.b.c.rejectDist * a.b.c.rejectDist;
    for (int i = tid; i < n; i += BLOCK)
        corr_point(a.b, x, i, sT, rej2);
    __syncthreads();
    CF m7[7];
    for (int k = 0; k < 7; k++)
        m7[k] = CF{0.0f, 0.0f};
    for (int i = tid; i < n; i += BLOCK)
        add_means(a.b, x, i, sT, m7);
    block_reduce_cf<BLOCK, 7>(m7, sLds, sRed);
    const double cnt = sRed[6];
    if (tid < 3)
        sC[tid] = (a.center && cnt > 0) ? (float)(sRed[tid] / cnt) : 0.0f;
    __syncthreads();
    float means[6] = {sC[0], sC[1], sC[2], 0.0f, 0.0f, 0.0f};
    CF    e28[28];
    for (int k = 0; k < 28; k++)
        e28[k] = CF{0.0f, 0.0f};
    for (int i = tid; i < n; i += BLOCK) {
        bool win = is_winner(x, i);
        if (a.corr)
            a.corr[(size_t)h * n + i] = win ? a.b.dst.id[x.nnSlot[i]] : -1;
        if (!win)
            continue;
        float J[6], r;
        eq_row(a.b, i, x.nnSlot[i], sT, means, 1, J, r);
        cf_add(e28[0], 0.5f * r * r);
        for (int p = 0; p < 6; p++)
            cf_add(e28[1 + p], r * J[p]);
        int k = 7;
        for (int p = 0; p < 6; p++)
            for (int q = p; q < 6; q++)
                cf_add(e28[k++], J[p] * J[q]);
    }
    block_reduce_cf<BLOCK, 28>(e28, sLds, sRed);
    if (tid == 0) {
        double *o = a.out + (size_t)h * 36;
        o[0] = sRed[0];
        o[1] = cnt;
        o[2] = sC[0];
        o[3] = sC[1];
        o[4] = sC[2];
        for (int k = 0; k < 27; k++)
            o[5 + k] = sRed[1 + k];
        for (int k = 32; k < 36; k++)
            o[k] = 0.0;
    }
    for (int i = tid; i < n; i += BLOCK)
        clear_point(x, i);
}

void icp_adjoint(Workspace &ws, int nSrc, const float *sx, const float *sy, const float *sz,
                 const GridView &dst, const float *T0host, const double *xiHost, int H,
                 float rejectDist, int center, double *outHost, int *corrHost, hipStream_t s) {
    if (H <= 0)
        return;
    require_copies(dst);
    IcpCriteria c{1, rejectDist, 0.0f, 0.0f, 0.0f};
    AdjArgs     a;
    a.b = make_args(ws, nSrc, sx, sy, sz, sx, sy, sz, dst, dst, H, c, 0.0f, 1);
    float  *T0 = ws.icpPose.as<float>((size_t)H * 12);
    double *xi = ws.icpCur.as<double>((size_t)H * 6);
    double *od = ws.icpInl.as<double>((size_t)H * 36);
    int    *cd = corrHost ? ws.icpOut.as<int>((size_t)H * nSrc) : nullptr;
    PPF_HIP_CHECK(hipMemcpyAsync(T0, T0host, sizeof(float) * 12 * H, hipMemcpyHostToDevice, s));
    PPF_HIP_CHECK(hipMemcpyAsync(xi, xiHost, sizeof(double) * 6 * H, hipMemcpyHostToDevice, s));
    a.xi     = xi;
    a.T0     = T0;
    a.center = center;
    a.out    = od;
    a.corr   = cd;
    adjoint_kernel<256><<<H, 256, 0, s>>>(a);
    PPF_HIP_CHECK(hipGetLastError());
    PPF_HIP_CHECK(hipMemcpyAsync(outHost, od, sizeof(double) * 36 * H, hipMemcpyDeviceToHost, s));
    if (corrHost)
        PPF_HIP_CHECK(hipMemcpyAsync(corrHost, cd, sizeof(int) * (size_t)H * nSrc,
                                     hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace ppfd
