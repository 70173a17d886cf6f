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
    float       *nnD;    // H x nSrc
    IcpResultDev *out;
};

// u = pinv(A) rhs for symmetric PSD A (FP64 Jacobi), dropping eigenvalues
// below 6 * eps_f32 * max |eigenvalue| as JacobiSVD::solve does.
__device__ void sym_solve6(const double A[36], const double rhs[6], double u[6]) {
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

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) icp_kernel(IcpArgs a) {
    const int h = blockIdx.x;
    if (h >= a.H)
        return;
    __shared__ double sPose[12], sCur[12], sU[12];
    __shared__ float  sCurF[12], sUF[12];
    __shared__ double sRed[40];
    __shared__ CF     sLds[(BLOCK / 64) * 34];
    __shared__ float  sMse, sRate;
    __shared__ int    sIters, sType, sConv, sStop, sValid;
    __shared__ int    sCnt[BLOCK / 64];

    const int tid = threadIdx.x;
    const int n   = a.nSrc;
    uint32_t *hkey = a.hkey + (size_t)h * a.T;
    unsigned long long *hbest = a.hbest + (size_t)h * a.T;
    int   *nnSlot = a.nnSlot + (size_t)h * n;
    int   *hslot  = a.hslot + (size_t)h * n;
    float *nnD    = a.nnD + (size_t)h * n;

    if (tid == 0) {
        bool ident = true;
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 4; c++) {
                float v = a.init[(size_t)h * 12 + r * 4 + c], e = (r == c) ? 1.0f : 0.0f;
                if (fabsf(v - e) > 1e-5f)
                    ident = false;
            }
        for (int i = 0; i < 12; i++) {
            sPose[i] = a.init[(size_t)h * 12 + i];
            sCur[i]  = ident ? ((i % 5 == 0) ? 1.0 : 0.0) : sPose[i];
        }
        sMse   = FLT_MAX;
        sRate  = 1.0f;
        sIters = 0;
        sType  = 0;
        sConv  = 0;
        sStop  = 0;
    }
    for (uint32_t t = tid; t < a.T; t += BLOCK) {
        hkey[t]  = 0u;
        hbest[t] = ~0ull;
    }
    __syncthreads();
    const float rej2 = a.c.rejectDist * a.c.rejectDist;

    while (!sStop && sIters < a.c.iterations) {
        if (tid < 12)
            sCurF[tid] = (float)sCur[tid];
        __syncthreads();
        // ---- 1. correspondences
        for (int i = tid; i < n; i += BLOCK) {
            float x = a.sx[i], y = a.sy[i], z = a.sz[i];
            float px = sCurF[0] * x + sCurF[1] * y + sCurF[2] * z + sCurF[3];
            float py = sCurF[4] * x + sCurF[5] * y + sCurF[6] * z + sCurF[7];
            float pz = sCurF[8] * x + sCurF[9] * y + sCurF[10] * z + sCurF[11];
            float d2;
            int   did;
            int   s = grid_nn1(a.dst, px, py, pz, d2, did);
            int   hs = -1;
            if (s >= 0 && !(d2 > rej2)) {
                uint32_t key = (uint32_t)did + 1u;
                uint32_t hh  = icp_hash(key) & (a.T - 1);
                for (;;) {
                    uint32_t prev = atomicCAS(&hkey[hh], 0u, key);
                    if (prev == 0u || prev == key)
                        break;
                    hh = (hh + 1u) & (a.T - 1);
                }
                atomicMin(&hbest[hh], ((unsigned long long)__float_as_uint(d2) << 32) | (uint32_t)i);
                hs = (int)hh;
            } else {
                s = -1;
            }
            nnSlot[i] = s;
            hslot[i]  = hs;
            nnD[i]    = d2;
        }
        __syncthreads();
        // ---- 2. winners, means
        CF m7[7];
        for (int k = 0; k < 7; k++)
            m7[k] = CF{0.0f, 0.0f};
        for (int i = tid; i < n; i += BLOCK) {
            int s = nnSlot[i];
            if (s < 0)
                continue;
            unsigned long long pk = ((unsigned long long)__float_as_uint(nnD[i]) << 32) | (uint32_t)i;
            if (hbest[hslot[i]] != pk)
                continue;
            float x = a.sx[i], y = a.sy[i], z = a.sz[i];
            cf_add(m7[0], sCurF[0] * x + sCurF[1] * y + sCurF[2] * z + sCurF[3]);
            cf_add(m7[1], sCurF[4] * x + sCurF[5] * y + sCurF[6] * z + sCurF[7]);
            cf_add(m7[2], sCurF[8] * x + sCurF[9] * y + sCurF[10] * z + sCurF[11]);
            cf_add(m7[3], a.dst.x[s]);
            cf_add(m7[4], a.dst.y[s]);
            cf_add(m7[5], a.dst.z[s]);
            cf_add(m7[6], 1.0f);
        }
        block_reduce_cf<BLOCK, 7>(m7, sLds, sRed);
        const int   valid = (int)sRed[6];
        const float pmx = (float)(sRed[0] / valid), pmy = (float)(sRed[1] / valid),
                    pmz = (float)(sRed[2] / valid);
        const float qmx = (float)(sRed[3] / valid), qmy = (float)(sRed[4] / valid),
                    qmz = (float)(sRed[5] / valid);
        if (valid < 6) {
            if (tid == 0) {
                sType = 3;  // NO_CORRESPONDS
                sStop = 1;
            }
            __syncthreads();
            break;
        }
        // ---- 3. normal equations (27 compensated sums: 21 of A, 6 of b)
        CF e27[27];
        for (int k = 0; k < 27; k++)
            e27[k] = CF{0.0f, 0.0f};
        for (int i = tid; i < n; i += BLOCK) {
            int s = nnSlot[i];
            if (s < 0)
                continue;
            unsigned long long pk = ((unsigned long long)__float_as_uint(nnD[i]) << 32) | (uint32_t)i;
            if (hbest[hslot[i]] != pk)
                continue;
            float x = a.sx[i], y = a.sy[i], z = a.sz[i];
            float px = sCurF[0] * x + sCurF[1] * y + sCurF[2] * z + sCurF[3];
            float py = sCurF[4] * x + sCurF[5] * y + sCurF[6] * z + sCurF[7];
            float pz = sCurF[8] * x + sCurF[9] * y + sCurF[10] * z + sCurF[11];
            float qx = a.dst.x[s], qy = a.dst.y[s], qz = a.dst.z[s];
            float nqx = a.dst.nx[s], nqy = a.dst.ny[s], nqz = a.dst.nz[s];
            float J[6], r;
            if (a.mode == 0) {
                float mx = a.snx[i], my = a.sny[i], mz = a.snz[i];
                float npx = sCurF[0] * mx + sCurF[1] * my + sCurF[2] * mz;
                float npy = sCurF[4] * mx + sCurF[5] * my + sCurF[6] * mz;
                float npz = sCurF[8] * mx + sCurF[9] * my + sCurF[10] * mz;
                float Nx = npx + nqx, Ny = npy + nqy, Nz = npz + nqz;
                float ux = px - pmx, uy = py - pmy, uz = pz - pmz;
                float vx = qx - qmx, vy = qy - qmy, vz = qz - qmz;
                float sx = ux + vx, sy = uy + vy, sz = uz + vz;
                r    = (ux - vx) * Nx + (uy - vy) * Ny + (uz - vz) * Nz;
                J[0] = sy * Nz - sz * Ny;
                J[1] = sz * Nx - sx * Nz;
                J[2] = sx * Ny - sy * Nx;
                J[3] = Nx;
                J[4] = Ny;
                J[5] = Nz;
            } else {
                // point-to-plane about the source centroid: r = n_q.(p - q),
                // J = [(p - c) x n_q ; n_q]
                float ux = px - pmx, uy = py - pmy, uz = pz - pmz;
                r    = (px - qx) * nqx + (py - qy) * nqy + (pz - qz) * nqz;
                J[0] = uy * nqz - uz * nqy;
                J[1] = uz * nqx - ux * nqz;
                J[2] = ux * nqy - uy * nqx;
                J[3] = nqx;
                J[4] = nqy;
                J[5] = nqz;
            }
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
        block_reduce_cf<BLOCK, 27>(e27, sLds, sRed);
        // ---- 4. FP64 solve + update (one lane)
        if (tid == 0) {
            double A[36], rhs[6], u[6];
            int    k = 0;
            for (int p = 0; p < 6; p++)
                for (int q = p; q < 6; q++) {
                    A[p * 6 + q] = sRed[k];
                    A[q * 6 + p] = sRed[k];
                    k++;
                }
            for (int p = 0; p < 6; p++)
                rhs[p] = -sRed[21 + p];
            sym_solve6(A, rhs, u);
            double U[12];
            if (a.mode == 0) {
                // symmetric_rigid_matching's update (icp.cpp:70-86)
                double an    = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
                double theta = atan(an);
                double ax[3] = {0, 0, 0};
                if (an > 0)
                    for (int i = 0; i < 3; i++)
                        ax[i] = u[i] / an;
                double W[9] = {0, -ax[2], ax[1], ax[2], 0, -ax[0], -ax[1], ax[0], 0}, W2[9], Ri[9],
                       R[9];
                for (int r = 0; r < 3; r++)
                    for (int q = 0; q < 3; q++) {
                        double s2 = 0;
                        for (int m = 0; m < 3; m++)
                            s2 += W[r * 3 + m] * W[m * 3 + q];
                        W2[r * 3 + q] = s2;
                    }
                double st = sin(theta), ct = cos(theta);
                for (int i = 0; i < 9; i++)
                    Ri[i] = ((i % 4 == 0) ? 1.0 : 0.0) + st * W[i] + (1.0 - ct) * W2[i];
                for (int r = 0; r < 3; r++)
                    for (int q = 0; q < 3; q++) {
                        double s2 = 0;
                        for (int m = 0; m < 3; m++)
                            s2 += Ri[r * 3 + m] * Ri[m * 3 + q];
                        R[r * 3 + q] = s2;
                    }
                double pm[3] = {sRed[0] * 0 + pmx, pmy, pmz}, qm[3] = {qmx, qmy, qmz};
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
                double c[3] = {pmx, pmy, pmz};
                for (int r = 0; r < 3; r++) {
                    double Rc = R[r * 3 + 0] * c[0] + R[r * 3 + 1] * c[1] + R[r * 3 + 2] * c[2];
                    U[r * 4 + 0] = R[r * 3 + 0];
                    U[r * 4 + 1] = R[r * 3 + 1];
                    U[r * 4 + 2] = R[r * 3 + 2];
                    U[r * 4 + 3] = c[r] - Rc + u[3 + r];
                }
            }
            for (int i = 0; i < 12; i++) {
                sU[i]  = U[i];
                sUF[i] = (float)U[i];
            }
            sValid = valid;
        }
        __syncthreads();
        // ---- 5. mean NN distance after the update (icp.cpp:182-194)
        CF m1[1] = {CF{0.0f, 0.0f}};
        for (int i = tid; i < n; i += BLOCK) {
            int s = nnSlot[i];
            if (s < 0)
                continue;
            unsigned long long pk = ((unsigned long long)__float_as_uint(nnD[i]) << 32) | (uint32_t)i;
            if (hbest[hslot[i]] != pk)
                continue;
            float x = a.sx[i], y = a.sy[i], z = a.sz[i];
            float px = sCurF[0] * x + sCurF[1] * y + sCurF[2] * z + sCurF[3];
            float py = sCurF[4] * x + sCurF[5] * y + sCurF[6] * z + sCurF[7];
            float pz = sCurF[8] * x + sCurF[9] * y + sCurF[10] * z + sCurF[11];
            float qx = sUF[0] * px + sUF[1] * py + sUF[2] * pz + sUF[3];
            float qy = sUF[4] * px + sUF[5] * py + sUF[6] * pz + sUF[7];
            float qz = sUF[8] * px + sUF[9] * py + sUF[10] * pz + sUF[11];
            float d2;
            int   did;
            grid_nn1(a.dst, qx, qy, qz, d2, did);
            cf_add(m1[0], sqrtf(d2));
        }
        block_reduce_cf<BLOCK, 1>(m1, sLds, sRed);
        // clear this iteration's hash slots
        for (int i = tid; i < n; i += BLOCK) {
            int hs = hslot[i];
            if (hs >= 0) {
                hkey[hs]  = 0u;
                hbest[hs] = ~0ull;
            }
        }
        // ---- 6. convergence (icp.cpp:233-272)
        if (tid == 0) {
            float mse = (float)(sRed[0] / (double)sValid);
            int   vp  = sValid;
            if (vp > 6 && mse < a.c.mseMax)
                sConv = 1;
            bool stop = false;
            if (mse < a.c.mseMin) {
                sType = 1;
                stop  = true;
            }
            float rate = (sMse - mse) / sMse;
            if (rate < a.c.tolerance) {
                sType = 2;
                stop  = true;
            }
            if (sIters++ >= a.c.iterations) {
                sType = 0;
                stop  = true;
            }
            if (rate > 0.0f) {
                sMse  = mse;
                sRate = rate;
                compose3x4(sU, sPose, sPose);
            }
            if (!stop)
                compose3x4(sU, sCur, sCur);
            sStop = stop ? 1 : 0;
        }
        __syncthreads();
    }
    // ---- scoring: inliner (util.cpp:628-641) of the refined pose
    __shared__ float sPF[12];
    if (tid < 12)
        sPF[tid] = (float)sPose[tid];
    __syncthreads();
    float r2  = a.scoringDist * a.scoringDist;
    int   cnt = 0;
    for (int i = tid; i < n; i += BLOCK) {
        float x = a.sx[i], y = a.sy[i], z = a.sz[i];
        float px = sPF[0] * x + sPF[1] * y + sPF[2] * z + sPF[3];
        float py = sPF[4] * x + sPF[5] * y + sPF[6] * z + sPF[7];
        float pz = sPF[8] * x + sPF[9] * y + sPF[10] * z + sPF[11];
        if (grid_any_within(a.full, px, py, pz, r2))
            cnt++;
    }
    for (int o = 32; o > 0; o >>= 1)
        cnt += __shfl_xor(cnt, o, 64);
    if ((tid & 63) == 0)
        sCnt[tid >> 6] = cnt;
    __syncthreads();
    if (tid == 0) {
        int tot = 0;
        for (int w = 0; w < BLOCK / 64; w++)
            tot += sCnt[w];
        IcpResultDev &o = a.out[h];
        for (int i = 0; i < 12; i++)
            o.pose[i] = sPF[i];
        o.type         = sType;
        o.mse          = sMse;
        o.convergeRate = sRate;
        o.iterations   = sIters;
        o.converged    = sConv;
        o.inliers      = tot;
    }
}

void icp_batch(Workspace &ws, int nSrc, const float *sx, const float *sy, const float *sz,
               const float *snx, const float *sny, const float *snz, const GridView &dst,
               const GridView &full, const float *init12, int H, const IcpCriteria &c,
               float scoringDist, const IcpConfig &cfg, IcpResultDev *out, hipStream_t s) {
    if (H <= 0)
        return;
    if (c.iterations < 1 || c.mseMin < 0 || c.mseMax < c.mseMin || c.tolerance > 1 ||
        c.tolerance < 0 || c.rejectDist < 0)
        throw std::runtime_error("Invalid ConvergenceCriteria at ICP::regist");
    if (!dst.nx)
        throw std::runtime_error("PointCloud empty or no normal at ICP::regist");
    IcpArgs a;
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
    a.mode = cfg.mode;
    uint32_t T = 64;
    while (T < 2u * (uint32_t)std::max(nSrc, 1))
        T <<= 1;
    a.T      = T;
    a.hkey   = ws.icpHash.as<uint32_t>((size_t)H * T * 3);
    a.hbest  = ws.icpPartial.as<unsigned long long>((size_t)H * T);
    int *ib  = ws.icpState.as<int>((size_t)H * nSrc * 3);
    a.nnSlot = ib;
    a.hslot  = ib + (size_t)H * nSrc;
    a.nnD    = reinterpret_cast<float *>(ib + 2 * (size_t)H * nSrc);
    float *init = ws.icpPose.as<float>((size_t)H * 12);
    PPF_HIP_CHECK(hipMemcpyAsync(init, init12, sizeof(float) * 12 * H, hipMemcpyHostToDevice, s));
    a.init = init;
    IcpResultDev *od = ws.icpOut.as<IcpResultDev>(H);
    a.out = od;
    switch (cfg.block) {
        case 64: icp_kernel<64><<<H, 64, 0, s>>>(a); break;
        case 128: icp_kernel<128><<<H, 128, 0, s>>>(a); break;
        case 512: icp_kernel<512><<<H, 512, 0, s>>>(a); break;
        case 1024: icp_kernel<1024><<<H, 1024, 0, s>>>(a); break;
        default: icp_kernel<256><<<H, 256, 0, s>>>(a); break;
    }
    PPF_HIP_CHECK(hipGetLastError());
    if (out) {
        PPF_HIP_CHECK(hipMemcpyAsync(out, od, sizeof(IcpResultDev) * H, hipMemcpyDeviceToHost, s));
        PPF_HIP_CHECK(hipStreamSynchronize(s));
    }
}

}  // namespace ppfd
