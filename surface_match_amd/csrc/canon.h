// canon.h -- the PPF path's canonical FP32 arithmetic, shared by host and
// gfx950 device code.
//
// Every function here is a fixed sequence of IEEE-754 single-precision
// operations (+ - * / sqrt floor ceil, correctly rounded on both the host
// and gfx950) with no fused multiply-add: the library is compiled with
// -ffp-contract=off and this header also pins contraction off.  Hash keys,
// alphas and angle bins computed here are therefore identical bit for bit
// on every device and on the host.
//
// Reference: SurfaceMan/surface_match util.cpp:459-484 (transformRT,
// computeAlpha), :643-777 (murmurhash3, hashPPF), :697-827 (angleA2B,
// computePPF), ppf.cpp:316-346 (angle bins), :393-405 (poses).
#pragma once

#include <stdint.h>
#include <math.h>
#ifndef __HIP_DEVICE_COMPILE__
#include <cmath>
#endif

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PPF_HD __host__ __device__ inline
#else
#define PPF_HD inline
#endif

#pragma clang fp contract(off)

namespace ppfc {

constexpr float kPi    = 3.14159265358979323846f;
constexpr float kTwoPi = 2.0f * 3.14159265358979323846f;  // ppf.cpp:24 const float M_2PI
constexpr float kPiO2  = 1.5707963267948966192f;
constexpr float kPiO4  = 0.7853981633974483096f;

PPF_HD bool is_nan(float v) { return v != v; }
PPF_HD bool sign_bit(float v) {
    union { float f; uint32_t u; } c{v};
    return (c.u >> 31) != 0u;
}

// arctangent by the method of the Cephes library's atanf (S. L. Moshier):
// two range reductions to |x| <= tan(pi/8) and an odd degree-9 polynomial.
PPF_HD float atanf_c(float xx) {
    float x   = xx;
    bool  neg = false;
    if (x < 0.0f) {
        neg = true;
        x   = -x;
    }
    // the reductions -1/x (x > tan 3pi/8) and (x-1)/(x+1) (x > tan pi/8) as
    // one division; -1/x == -(1/x) and x/1 == x exactly under round to nearest
    const bool big = x > 2.414213562373095f, mid = x > 0.4142135623730950f;
    float      y   = big ? kPiO2 : (mid ? kPiO4 : 0.0f);
    float      num = big ? -1.0f : (mid ? x - 1.0f : x);
    float      den = big ? x : (mid ? x + 1.0f : 1.0f);
    x              = num / den;
    float z = x * x;
    float p = 8.05374449538e-2f;
    p       = p * z - 1.38776856032e-1f;
    p       = p * z + 1.99777106478e-1f;
    p       = p * z - 3.33329491539e-1f;
    p       = p * z;
    p       = p * x;
    p       = p + x;
    y       = y + p;
    return neg ? -y : y;
}

// atan2 with the C99 Annex F treatment of signed zeros.
PPF_HD float atan2f_c(float y, float x) {
    if (y == 0.0f) {
        if (x > 0.0f || (x == 0.0f && !sign_bit(x)))
            return y;
        return sign_bit(y) ? -kPi : kPi;
    }
    if (x == 0.0f)
        return y > 0.0f ? kPiO2 : -kPiO2;
    float z = atanf_c(y / x);
    if (x < 0.0f) {
        float w = y > 0.0f ? kPi : -kPi;
        return w + z;
    }
    return z;
}

// MurmurHash3_x86_32 of four little-endian 32-bit words with seed 42
// (util.cpp:643-688 with len = 16, :732-763).
PPF_HD uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
PPF_HD uint32_t murmur_key4(uint32_t q0, uint32_t q1, uint32_t q2, uint32_t q3) {
    const uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
    uint32_t       h  = 42u;
    uint32_t       ks[4] = {q0, q1, q2, q3};
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int i = 0; i < 4; i++) {
        uint32_t k = ks[i] * c1;
        k          = rotl32(k, 15) * c2;
        h ^= k;
        h = rotl32(h, 13) * 5u + 0xe6546b64u;
    }
    h ^= 16u;
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

// quantisation: SIMD body ceil + cast to uint32 (NaN lane -> 0x80000000 as
// cvttps2dq gives), util.cpp:771-774; scalar tail int() truncation, :691-692.
PPF_HD uint32_t quant(float v, float step, bool tail) {
    float r = v / step;
    if (is_nan(r))
        return 0x80000000u;
    if (tail)
        return (uint32_t)(int32_t)r;
    return (uint32_t)(int32_t)ceilf(r);
}

PPF_HD float angle_a2b(float ax, float ay, float az, float bx, float by, float bz) {
    float i     = ay * bz - az * by;
    float j     = az * bx - ax * bz;
    float k     = ax * by - ay * bx;
    float cross = sqrtf(i * i + j * j + k * k);
    float dot   = ax * bx + ay * by + az * bz;
    return atan2f_c(cross, dot);
}

// computePPF for one partner (body util.cpp:802-818, tail :712-730).
PPF_HD uint32_t ppf_key(float p1x, float p1y, float p1z, float n1x, float n1y, float n1z,
                        float p2x, float p2y, float p2z, float n2x, float n2y, float n2z,
                        float angleStep, float distStep, bool tail) {
    float dx = p2x - p1x;
    float dy = p2y - p1y;
    float dz = p2z - p1z;
    float dn = sqrtf(dx * dx + dy * dy + dz * dz);
    float f1, f2, f3;
    if (tail && !(dn > 0.0f)) {
        f1 = f2 = f3 = 0.0f;
    } else {
        float ux = dx / dn, uy = dy / dn, uz = dz / dn;
        f1 = angle_a2b(ux, uy, uz, n1x, n1y, n1z);
        f2 = angle_a2b(ux, uy, uz, n2x, n2y, n2z);
        f3 = angle_a2b(n1x, n1y, n1z, n2x, n2y, n2z);
    }
    return murmur_key4(quant(f1, angleStep, tail), quant(f2, angleStep, tail),
                       quant(f3, angleStep, tail), quant(dn, distStep, tail));
}

// Local frame (transformRT, util.cpp:459-475): minimal rotation taking n to
// +x, t = -R p; 3x4 row-major.  Closed-form Rodrigues, identical to the
// oracle's restatement.
PPF_HD void frame(float px, float py, float pz, float c, float ny, float nz, float rt[12]) {
    float R[9];
    if (ny == 0.0f && nz == 0.0f) {
        float sx = c >= 0.0f ? 1.0f : -1.0f;
        R[0] = sx;   R[1] = 0.0f; R[2] = 0.0f;
        R[3] = 0.0f; R[4] = 1.0f; R[5] = 0.0f;
        R[6] = 0.0f; R[7] = 0.0f; R[8] = sx;
    } else {
        float s2  = ny * ny + nz * nz;
        float k   = (1.0f - c) / s2;
        float kny = k * ny, knz = k * nz;
        float off = -(kny * nz);
        R[0] = c;   R[1] = ny;           R[2] = nz;
        R[3] = -ny; R[4] = c + knz * nz; R[5] = off;
        R[6] = -nz; R[7] = off;          R[8] = c + kny * ny;
    }
    for (int r = 0; r < 3; r++) {
        rt[r * 4 + 0] = R[r * 3 + 0];
        rt[r * 4 + 1] = R[r * 3 + 1];
        rt[r * 4 + 2] = R[r * 3 + 2];
        rt[r * 4 + 3] = -(R[r * 3 + 0] * px + R[r * 3 + 1] * py + R[r * 3 + 2] * pz);
    }
}

// computeAlpha (util.cpp:477-484, :853-866): sign flip only at +-fl(pi)
// with z != 0 (the value of sin(alpha)*z > 0 for every finite alpha).
PPF_HD float alpha_of(const float rt[12], float x, float y, float z) {
    float yy = rt[4] * x + rt[5] * y + rt[6] * z + rt[7];
    float zz = rt[8] * x + rt[9] * y + rt[10] * z + rt[11];
    float a  = atan2f_c(-zz, yy);
    if ((a == kPi || a == -kPi) && zz != 0.0f)
        a = -a;
    return a;
}

// the scene alpha's part of an angle bin, once per scene pair: the body
// subtracts alphaScene - fl(pi) in float, the tail alphaScene - M_PI in double
// rounded to float (ppf.cpp:316-340)
PPF_HD float bin_offset_body(float alphaScene) { return alphaScene - kPi; }
PPF_HD float bin_offset_tail(float alphaScene) {
    return (float)((double)alphaScene - 3.14159265358979323846);
}
PPF_HD int bin_of(float a, float sMaxId, int angleNum) {
    int bin = (int)floorf(sMaxId * a);
    bin     = bin < 0 ? 0 : bin;
    bin     = bin > angleNum - 1 ? angleNum - 1 : bin;
    return bin;
}
// the same bins as angle_bin from a precomputed offset
PPF_HD int bin_body(float alphaModel, float offBody, float sMaxId, int angleNum) {
    float a = alphaModel - offBody;
    if (a > kTwoPi)
        a = a - kTwoPi;
    if (a < 0.0f)
        a = a + kTwoPi;
    return bin_of(a, sMaxId, angleNum);
}
PPF_HD int bin_tail(float alphaModel, float offTail, float sMaxId, int angleNum) {
    float a = alphaModel - offTail;
    if (a < 0.0f)
        a = a + kTwoPi;
    if (a > kTwoPi)
        a = a - kTwoPi;
    return bin_of(a, sMaxId, angleNum);
}

// bin_body and bin_tail in one branch-free form, given the rule's offset:
// with |alphaModel| <= pi and the offset in [-2pi, 0], a = alphaModel - off
// lies in [-pi, 3pi], so at most one of the two wraps applies (a > 2pi
// leaves a - 2pi > 0; a < 0 gives a + 2pi <= 2pi) and the rules' different
// wrap orders agree.  After the wrap a >= 0, where truncation is floor;
// non-finite inputs clamp as bin_of clamps them.
PPF_HD int bin_wrapped(float alphaModel, float off, float sMaxId, int angleNum) {
    float a = alphaModel - off;
    a       = a > kTwoPi ? a - kTwoPi : (a < 0.0f ? a + kTwoPi : a);
    int bin = (int)(sMaxId * a);
    bin     = bin < 0 ? 0 : bin;
    return bin > angleNum - 1 ? angleNum - 1 : bin;
}

// bin_wrapped for finite inputs in fewer instructions (the voting loop's
// form).  The wrap adds one selected constant: a - 2pi is a + (-2pi) in
// IEEE arithmetic, and a + 0 == a except for a = -0, whose bin is 0 either
// way.  The clamp acts on the product in float: for finite x,
// trunc(clamp(x, 0, maxBin)) == clamp(trunc(x), 0, maxBin) because trunc is
// monotone and maps [0, maxBin] onto integers within the same bounds.
// maxBin = angleNum - 1 as a float.
PPF_HD int bin_wrapped_fast(float alphaModel, float off, float sMaxId, float maxBin) {
    float       a = alphaModel - off;
    const float w = a > kTwoPi ? -kTwoPi : (a < 0.0f ? kTwoPi : 0.0f);
    a             = a + w;
#if defined(__HIP_DEVICE_COMPILE__)
    const float x = __builtin_amdgcn_fmed3f(sMaxId * a, 0.0f, maxBin);
#else
    const float x = std::fmin(std::fmax(sMaxId * a, 0.0f), maxBin);
#endif
    return (int)x;
}

// angle bin of a bucket entry (ppf.cpp:316-340); tail = the entry is among
// the bucket's last size%8.
PPF_HD int angle_bin(float alphaModel, float alphaScene, float sMaxId, int angleNum, bool tail) {
    return tail ? bin_tail(alphaModel, bin_offset_tail(alphaScene), sMaxId, angleNum)
                : bin_body(alphaModel, bin_offset_body(alphaScene), sMaxId, angleNum);
}

// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC 2011): ten rounds of
// two 32x32->64 multiplies, the key bumped by the Weyl constants between
// rounds.  c (4 words) is replaced by the output block.
PPF_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1, n3 = (uint32_t)p0;
        c[0] = n0;
        c[1] = n1;
        c[2] = n2;
        c[3] = n3;
    }
}
// a float in [0, 1) from the top 24 bits of a random word
PPF_HD float uniform01(uint32_t u) { return (float)(u >> 8) * (1.0f / 16777216.0f); }

// a triangle's area |u x v| / 2 and unit normal (u x v) / |u x v|, u = p1 - p0,
// v = p2 - p0, as sampleMesh forms them (util.cpp:29-42): Eigen's cross,
// |k| = sqrt((k0^2 + k1^2) + k2^2), and normalized() leaving a zero k as is
PPF_HD float tri_area_normal(const float *p0, const float *p1, const float *p2, float nrm[3]) {
    const float u0 = p1[0] - p0[0], u1 = p1[1] - p0[1], u2 = p1[2] - p0[2];
    const float v0 = p2[0] - p0[0], v1 = p2[1] - p0[1], v2 = p2[2] - p0[2];
    const float k0 = u1 * v2 - u2 * v1, k1 = u2 * v0 - u0 * v2, k2 = u0 * v1 - u1 * v0;
    const float n2 = (k0 * k0 + k1 * k1) + k2 * k2;
    const float l  = sqrtf(n2);
    if (n2 > 0.0f) {
        nrm[0] = k0 / l;
        nrm[1] = k1 / l;
        nrm[2] = k2 / l;
    } else {
        nrm[0] = k0;
        nrm[1] = k1;
        nrm[2] = k2;
    }
    return l / 2.f;
}

// one point of triangle (p0, p1, p2) from two uniforms, as sampleMesh
// forms it (util.cpp:60-71): a = 1 - sqrt(r1) and b = sqrt(r1) (1 - r2) in
// double, c = sqrt(r1) r2 in float, each weight rounded to float before it
// scales its vertex, summed left to right in float
PPF_HD void mesh_point(const float *p0, const float *p1, const float *p2, float r1, float r2,
                       float out[3]) {
    const float  s = sqrtf(r1);
    const double a = 1. - (double)s;
    const double b = (double)s * (1. - (double)r2);
    const float  c = s * r2;
    const float  fa = (float)a, fb = (float)b;
    for (int k = 0; k < 3; k++)
        out[k] = (fa * p0[k] + fb * p1[k]) + c * p2[k];
}

// nanoflann L2 metric for DIM = 3 (nanoflann.hpp:545-581): ((0+dx^2)+dy^2)+dz^2.
PPF_HD float dist2(float qx, float qy, float qz, float px, float py, float pz) {
    float dx = qx - px, dy = qy - py, dz = qz - pz;
    float r  = 0.0f;
    r        = r + dx * dx;
    r        = r + dy * dy;
    r        = r + dz * dz;
    return r;
}

// Shoemake's quaternion of a rotation matrix (Eigen's branch structure),
// q = (w, x, y, z), R row-major 3x3.
PPF_HD void quat_from_rot(const float R[9], float q[4]) {
    float t = R[0] + R[4] + R[8];
    if (t > 0.0f) {
        t    = sqrtf(t + 1.0f);
        q[0] = 0.5f * t;
        t    = 0.5f / t;
        q[1] = (R[7] - R[5]) * t;
        q[2] = (R[2] - R[6]) * t;
        q[3] = (R[3] - R[1]) * t;
    } else {
        int i = 0;
        if (R[4] > R[0])
            i = 1;
        if (R[8] > R[i * 3 + i])
            i = 2;
        int j = (i + 1) % 3, k = (j + 1) % 3;
        t = sqrtf(R[i * 3 + i] - R[j * 3 + j] - R[k * 3 + k] + 1.0f);
        float v[3];
        v[i] = 0.5f * t;
        t    = 0.5f / t;
        q[0] = (R[k * 3 + j] - R[j * 3 + k]) * t;
        v[j] = (R[j * 3 + i] + R[i * 3 + j]) * t;
        v[k] = (R[k * 3 + i] + R[i * 3 + k]) * t;
        q[1] = v[0];
        q[2] = v[1];
        q[3] = v[2];
    }
}

PPF_HD void rot_from_quat(const float q[4], float R[9]) {
    float w = q[0], x = q[1], y = q[2], z = q[3];
    float tx = 2.0f * x, ty = 2.0f * y, tz = 2.0f * z;
    float twx = tx * w, twy = ty * w, twz = tz * w;
    float txx = tx * x, txy = ty * x, txz = tz * x;
    float tyy = ty * y, tyz = tz * y, tzz = tz * z;
    R[0] = 1.0f - (tyy + tzz);
    R[1] = txy - twz;
    R[2] = txz + twy;
    R[3] = txy + twz;
    R[4] = 1.0f - (txx + tzz);
    R[5] = tyz - twx;
    R[6] = txz - twy;
    R[7] = tyz + twx;
    R[8] = 1.0f - (txx + tyy);
}

PPF_HD float angle_of_quat(const float q[4]) {
    float n = sqrtf(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    return 2.0f * atan2f_c(n, fabsf(q[0]));
}

// T = inv(Ts) * (Rx(alpha) * Tm), rigid inverse; cs = (cos, sin) of the
// bin's pose angle.  3x4 row-major.
PPF_HD void compose_pose(const float Ts[12], const float Tm[12], float c, float s, float out[12]) {
    float M[12];
    for (int col = 0; col < 4; col++) {
        M[0 + col] = Tm[0 + col];
        M[4 + col] = c * Tm[4 + col] - s * Tm[8 + col];
        M[8 + col] = s * Tm[4 + col] + c * Tm[8 + col];
    }
    float d0 = M[3] - Ts[3], d1 = M[7] - Ts[7], d2 = M[11] - Ts[11];
    for (int r = 0; r < 3; r++) {
        for (int col = 0; col < 3; col++)
            out[r * 4 + col] = Ts[0 * 4 + r] * M[0 * 4 + col] + Ts[1 * 4 + r] * M[1 * 4 + col] +
                               Ts[2 * 4 + r] * M[2 * 4 + col];
        out[r * 4 + 3] = Ts[0 * 4 + r] * d0 + Ts[1 * 4 + r] * d1 + Ts[2 * 4 + r] * d2;
    }
}

// e^x by the method of the Cephes library's expf.c (S. L. Moshier): x =
// g + n ln2 with n = floor(x log2(e) + 1/2) and ln2 split into
// 0.693359375 - 2.12194440e-4 (Cody-Waite), a degree-5 polynomial for
// e^g on [-ln2/2, ln2/2], then the exact scaling by 2^n.  Fixed operation
// order, so host and device agree bit for bit.
PPF_HD float expf_c(float xx) {
    if (xx != xx)
        return xx;
    if (xx > 88.72283905206835f)
        return INFINITY;
    if (xx < -103.278929903431851103f)
        return 0.0f;
    float       x = xx;
    const float z = floorf(1.44269504088896341f * x + 0.5f);
    x             = x - z * 0.693359375f;
    x             = x - z * -2.12194440e-4f;
    const int   n = (int)z;
    const float q = x * x;
    float       p = 1.9875691500e-4f * x + 1.3981999507e-3f;
    p             = p * x + 8.3334519073e-3f;
    p             = p * x + 4.1665795894e-2f;
    p             = p * x + 1.6666665459e-1f;
    p             = p * x + 5.0000001201e-1f;
    float r       = (p * q + x) + 1.0f;
    // r * 2^n in two exact power-of-two steps (n in [-150, 128])
    const int n1 = n / 2, n2 = n - n / 2;
    union { uint32_t u; float f; } s1{(uint32_t)(n1 + 127) << 23}, s2{(uint32_t)(n2 + 127) << 23};
    return (r * s1.f) * s2.f;
}

// Eigen's unitOrthogonal() of a 3-vector: (-y, x, 0) / |(x, y)| unless both
// |x| and |y| are at most 1e-5 |z|, else (0, -z, y) / |(y, z)|; the
// division is a multiplication by the reciprocal of the norm
PPF_HD void unit_orthogonal(const float n[3], float o[3]) {
    const float az = fabsf(n[2]) * 1e-5f;
    if (!(fabsf(n[0]) <= az) || !(fabsf(n[1]) <= az)) {
        const float inv = 1.0f / sqrtf(n[0] * n[0] + n[1] * n[1]);
        o[0]            = -n[1] * inv;
        o[1]            = n[0] * inv;
        o[2]            = 0.0f;
    } else {
        const float inv = 1.0f / sqrtf(n[1] * n[1] + n[2] * n[2]);
        o[0]            = 0.0f;
        o[1]            = -n[2] * inv;
        o[2]            = n[1] * inv;
    }
}

// a x b as Eigen's cross() forms it
PPF_HD void cross3(const float a[3], const float b[3], float o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

PPF_HD float dot3(const float a[3], const float b[3]) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

// Solve A x = b for a symmetric positive definite N x N matrix given by its
// lower triangle (row-major packed: A[i*(i+1)/2 + j], j <= i) by Cholesky in
// FP64, sums in ascending index order.  False when a pivot is not positive.
template <int N>
PPF_HD bool cholesky_solve(const double *A, double *x) {
    double L[N * (N + 1) / 2];
    for (int j = 0; j < N; j++) {
        double d = A[j * (j + 1) / 2 + j];
        for (int k = 0; k < j; k++)
            d = d - L[j * (j + 1) / 2 + k] * L[j * (j + 1) / 2 + k];
        if (!(d > 0.0))
            return false;
        const double ljj       = sqrt(d);
        L[j * (j + 1) / 2 + j] = ljj;
        for (int i = j + 1; i < N; i++) {
            double s = A[i * (i + 1) / 2 + j];
            for (int k = 0; k < j; k++)
                s = s - L[i * (i + 1) / 2 + k] * L[j * (j + 1) / 2 + k];
            L[i * (i + 1) / 2 + j] = s / ljj;
        }
    }
    for (int i = 0; i < N; i++) {  // L y = b
        double s = x[i];
        for (int k = 0; k < i; k++)
            s = s - L[i * (i + 1) / 2 + k] * x[k];
        x[i] = s / L[i * (i + 1) / 2 + i];
    }
    for (int i = N - 1; i >= 0; i--) {  // L^T x = y
        double s = x[i];
        for (int k = i + 1; k < N; k++)
            s = s - L[k * (k + 1) / 2 + i] * x[k];
        x[i] = s / L[i * (i + 1) / 2 + i];
    }
    return true;
}

// Cyclic Jacobi eigen-decomposition of a symmetric n x n FP64 matrix (Golub
// & Van Loan, Matrix Computations, sec. 8.5).  a is destroyed; v receives the
// eigenvectors as columns, w the eigenvalues.
template <int N>
PPF_HD void jacobi_eig(double *a, double *v, double *w) {
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            v[i * N + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < N; i++) {
            diag = diag + a[i * N + i] * a[i * N + i];
            for (int j = i + 1; j < N; j++)
                off = off + a[i * N + j] * a[i * N + j];
        }
        if (!(off > 1e-30 * diag) || off == 0.0)
            break;
        for (int p = 0; p < N - 1; p++)
            for (int q = p + 1; q < N; q++) {
                double apq = a[p * N + q];
                if (apq == 0.0)
                    continue;
                double app = a[p * N + p], aqq = a[q * N + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t     = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
                if (theta < 0.0)
                    t = -t;
                double c = 1.0 / sqrt(t * t + 1.0);
                double s = t * c;
                for (int k = 0; k < N; k++) {
                    double akp = a[k * N + p], akq = a[k * N + q];
                    a[k * N + p] = c * akp - s * akq;
                    a[k * N + q] = s * akp + c * akq;
                }
                for (int k = 0; k < N; k++) {
                    double apk = a[p * N + k], aqk = a[q * N + k];
                    a[p * N + k] = c * apk - s * aqk;
                    a[q * N + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < N; k++) {
                    double vkp = v[k * N + p], vkq = v[k * N + q];
                    v[k * N + p] = c * vkp - s * vkq;
                    v[k * N + q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < N; i++)
        w[i] = a[i * N + i];
}

}  // namespace ppfc
