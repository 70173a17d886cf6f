/*
 * ppf_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for the PPF path.
 *
 * A C restatement of SurfaceMan/surface_match's trainModel/matchScene
 * (ppf.cpp:36-147, 149-510) and the helpers it calls (util.cpp, icp.cpp).
 * It is the checker the HIP library is compared against; the product never
 * links, loads or calls it.  Build: oracle/Makefile (gcc -O2
 * -ffp-contract=off, no FMA, so every FP32 operation is one IEEE rounding).
 *
 * Where the reference's result depends on kd-tree traversal order or on
 * OpenMP scheduling, this file defines an order-free "canonical" behaviour
 * (noted per function).  The reference itself cannot be built (its Eigen and
 * xsimd submodules are empty and gtl is fetched over the network), so the
 * pins listed below are what fixes each function:
 *
 *   orc_atanf/atan2f   error vs numpy float64 arctan2 <= 2 ulp, IEEE zero/sign
 *                      cases, odd symmetry (tests/test_oracle_math.py)
 *   murmur3/hash_key4  KAT vs sklearn murmurhash3_32(seed=42) (golden file)
 *   orc_ppf_key        features vs float64 numpy angles; invariants
 *                      (coplanar pair -> pi/2,pi/2,0; rigid invariance;
 *                      quantisation closed form ceil vs trunc)
 *   orc_frame/alpha    closed forms R n = e_x, R p + t = 0, alpha = -atan2(z,y)
 *                      of the transformed point; flip only at +-fl(pi)
 *   orc_angle_bin      brute-force float64 evaluation of the wrap rule
 *   orc_build_table    brute-force dictionary build in Python (test)
 *   orc_vote/peaks     brute-force Python vote on tiny models (test)
 *   grid radius/knn    brute force over all points (test)
 *   voxel sample       numpy brute force (test)
 *   normals            numpy eigh on the same covariance; plane/sphere
 *   normal smoothing   a hand case whose sum is known in closed form, with
 *                      the sign rule, self term and perpendicular neighbour
 *                      (tests/test_oracle_icp_rules.py)
 *   remove_nan/bbox    hand cases: NaN and inf coordinates, -0, ties of the
 *                      extreme point (tests/test_oracle_stages.py)
 *   clustering/NMS     hand cases + Python brute force (test)
 *   jacobi eig         numpy eigh
 *   ICP                exact-correspondence recovery; ground-truth pose
 *                      recovery on seeded scenes; ICP::regist's rules by
 *                      hand cases (5/6/7 correspondences, the mse and rate
 *                      stops, the iteration limit, a near-identity initial
 *                      pose; tests/test_oracle_icp_rules.py)
 *   orc_philox4x32_10  Random123's published known-answer vectors
 *   orc_sample_mesh    closed forms of the point weights, per-triangle
 *                      counts worked by hand, points on their triangles,
 *                      right-hand normals, uniform-over-area statistics
 *                      (tests/test_mesh.py)
 *   orc_expf           below one ulp of numpy's float64 exp; IEEE cases
 *   orc_unit_orthogonal  Eigen's two branches by closed form; orthonormality
 *   MLS normals        plane and paraboloid-apex fixed points, the order-1
 *                      lift in closed form, sphere denoising, an independent
 *                      float64 formulation (eigh + lstsq) (tests/test_mls.py)
 * Bitwise parity against the reference's xsimd atan2/sin and Eigen
 * AngleAxis/EigenSolver/JacobiSVD/SelfAdjointEigenSolver/LLT and glibc's
 * expf is unpinned (those libraries are absent), and sampleMesh's stream
 * is seeded Philox where the reference's std::mt19937 is seeded from
 * std::random_device (no fixed reference output exists).
 */
#include "ppf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI_F 3.14159265358979323846f
#define ORC_2PI_F (2.0f * 3.14159265358979323846f) /* const float M_2PI, ppf.cpp:24 */
#define ORC_PIO2_F 1.5707963267948966192f
#define ORC_PIO4_F 0.7853981633974483096f

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* ====================================================================== */
/* L0: scalar math                                                        */
/* ====================================================================== */

/* arctangent, method of the Cephes library atanf.c (S. L. Moshier, Cephes
 * Math Library 2.2): reduce |x| to [0, tan(pi/8)] with the identities
 * atan(x) = pi/2 + atan(-1/x) and atan(x) = pi/4 + atan((x-1)/(x+1)), then a
 * degree-9 odd polynomial.  Operation order is fixed (no contraction). */
float orc_atanf(float xx) {
    float x = xx, y, z;
    int   neg = 0;
    if (x < 0.0f) {
        neg = 1;
        x   = -x;
    }
    if (x > 2.414213562373095f) {
        y = ORC_PIO2_F;
        x = -(1.0f / x);
    } else if (x > 0.4142135623730950f) {
        y = ORC_PIO4_F;
        x = (x - 1.0f) / (x + 1.0f);
    } else {
        y = 0.0f;
    }
    z = x * x;
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

/* atan2 with C99 Annex F treatment of signed zeros (the reference calls
 * libm atan2f in util.cpp:479,698 and xsimd::atan2 in :709,862). */
float orc_atan2f(float y, float x) {
    if (y == 0.0f) {
        if (x > 0.0f || (x == 0.0f && !signbit(x)))
            return y; /* +-0 */
        return signbit(y) ? -ORC_PI_F : ORC_PI_F;
    }
    if (x == 0.0f)
        return y > 0.0f ? ORC_PIO2_F : -ORC_PIO2_F;
    float z = orc_atanf(y / x);
    if (x < 0.0f) {
        float w = y > 0.0f ? ORC_PI_F : -ORC_PI_F;
        return w + z;
    }
    return z;
}

/* MurmurHash3_x86_32 (Austin Appleby), util.cpp:643-688. */
uint32_t orc_murmur3_x86_32(const uint8_t *data, uint32_t len, uint32_t seed) {
    const uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
    uint32_t       h  = seed;
    uint32_t       nb = len / 4;
    for (uint32_t i = 0; i < nb; i++) {
        uint32_t k;
        memcpy(&k, data + 4 * i, 4);
        k *= c1;
        k = (k << 15) | (k >> 17);
        k *= c2;
        h ^= k;
        h = (h << 13) | (h >> 19);
        h = h * 5u + 0xe6546b64u;
    }
    const uint8_t *tail = data + 4 * nb;
    uint32_t       k1   = 0;
    switch (len & 3u) {
        case 3: k1 ^= (uint32_t)tail[2] << 16; /* fallthrough */
        case 2: k1 ^= (uint32_t)tail[1] << 8;  /* fallthrough */
        case 1:
            k1 ^= tail[0];
            k1 *= c1;
            k1 = (k1 << 15) | (k1 >> 17);
            k1 *= c2;
            h ^= k1;
    }
    h ^= len;
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

/* hashPPF: the 4 quantised features as a 16-byte little-endian key with
 * seed 42 (util.cpp:690-695, :732-777). */
uint32_t orc_hash_key4(const uint32_t q[4]) {
    uint8_t b[16];
    memcpy(b, q, 16);
    return orc_murmur3_x86_32(b, 16, 42u);
}

/* ====================================================================== */
/* L1: pair features, frame, alpha, bins                                 */
/* ====================================================================== */

/* quantisation: SIMD body ceil + batch_cast<uint32> (util.cpp:771-774; a NaN
 * lane converts to 0x80000000 as cvttps2dq does); scalar tail int() trunc
 * (util.cpp:691-692). */
static uint32_t quant(float v, float step, int tail) {
    float r = v / step;
    if (isnan(r))
        return 0x80000000u;
    if (tail)
        return (uint32_t)(int32_t)r;
    return (uint32_t)(int32_t)ceilf(r);
}

/* angleA2B: atan2(|a x b|, a . b) with Eigen/xsimd operation order
 * (util.cpp:697-710). */
static float angle_a2b(float ax, float ay, float az, float bx, float by, float bz) {
    float i     = ay * bz - az * by;
    float j     = az * bx - ax * bz;
    float k     = ax * by - ay * bx;
    float cross = sqrtf(i * i + j * j + k * k);
    float dot   = ax * bx + ay * by + az * bz;
    return orc_atan2f(cross, dot);
}

/* computePPF for one partner: body = util.cpp:802-818, tail = :712-730. */
uint32_t orc_ppf_key(const float p1[3], const float n1[3], const float p2[3], const float n2[3],
                     float angleStep, float distStep, int tail, float f[4], uint32_t q[4]) {
    float dx = p2[0] - p1[0];
    float dy = p2[1] - p1[1];
    float dz = p2[2] - p1[2];
    float dn = sqrtf(dx * dx + dy * dy + dz * dz);
    float f1, f2, f3;
    if (tail && !(dn > 0.0f)) {
        f1 = f2 = f3 = 0.0f;
    } else {
        float ux = dx / dn, uy = dy / dn, uz = dz / dn;
        f1 = angle_a2b(ux, uy, uz, n1[0], n1[1], n1[2]);
        f2 = angle_a2b(ux, uy, uz, n2[0], n2[1], n2[2]);
        f3 = angle_a2b(n1[0], n1[1], n1[2], n2[0], n2[1], n2[2]);
    }
    uint32_t qq[4] = {quant(f1, angleStep, tail), quant(f2, angleStep, tail),
                      quant(f3, angleStep, tail), quant(dn, distStep, tail)};
    if (f) {
        f[0] = f1;
        f[1] = f2;
        f[2] = f3;
        f[3] = dn;
    }
    if (q)
        memcpy(q, qq, sizeof qq);
    return orc_hash_key4(qq);
}

/* transformRT (util.cpp:459-475): the rotation about axis (0, n_z, -n_y)
 * by acos(n_x), i.e. the minimal rotation taking n to +x, with t = -R p.
 * Evaluated in closed form (Rodrigues with cos = n_x, sin^2 = n_y^2+n_z^2):
 *   R = [[c, ny, nz], [-ny, c + k nz^2, -k ny nz], [-nz, -k ny nz, c + k ny^2]],
 *   k = (1 - c) / (ny^2 + nz^2).
 * For n_y = n_z = 0 the reference rotates about +y by acos(n_x): identity
 * for n_x > 0, diag(-1, 1, -1) otherwise. */
void orc_frame(const float p[3], const float n[3], float rt[12]) {
    float c = n[0], ny = n[1], nz = n[2];
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
        R[0] = c;   R[1] = ny;            R[2] = nz;
        R[3] = -ny; R[4] = c + knz * nz;  R[5] = off;
        R[6] = -nz; R[7] = off;           R[8] = c + kny * ny;
    }
    for (int r = 0; r < 3; r++) {
        rt[r * 4 + 0] = R[r * 3 + 0];
        rt[r * 4 + 1] = R[r * 3 + 1];
        rt[r * 4 + 2] = R[r * 3 + 2];
        rt[r * 4 + 3] = -(R[r * 3 + 0] * p[0] + R[r * 3 + 1] * p[1] + R[r * 3 + 2] * p[2]);
    }
}

/* computeAlpha (util.cpp:477-484, :853-866): y,z of rt*p2, alpha =
 * atan2(-z, y), sign flipped when sin(alpha)*z > 0.  For a finite alpha
 * that test is true exactly when alpha = +-fl(pi) and z != 0 (sin(fl(pi))
 * < 0 because fl(pi) > pi; every other float in (-pi,pi) has sin of its own
 * sign, which atan2 makes opposite to z's). */
float orc_alpha(const float rt[12], const float p2[3]) {
    float y = rt[4] * p2[0] + rt[5] * p2[1] + rt[6] * p2[2] + rt[7];
    float z = rt[8] * p2[0] + rt[9] * p2[1] + rt[10] * p2[2] + rt[11];
    float a = orc_atan2f(-z, y);
    if ((a == ORC_PI_F || a == -ORC_PI_F) && z != 0.0f)
        a = -a;
    return a;
}

/* angle bin (ppf.cpp:316-340).  Body (bucket position < size - size%8):
 * float alpha_s - fl(pi), wrap >2pi then <0.  Tail: alpha_s - M_PI in
 * double rounded to float, wrap <0 then >2pi.  bin = floor(fl(maxIdx/2pi) a). */
int orc_angle_bin(float alphaModel, float alphaScene, int angleNum, int tail) {
    const float twoPi  = ORC_2PI_F;
    const float maxIdx = (float)(angleNum - 1);
    const float sMaxId = maxIdx / twoPi;
    float       a;
    if (!tail) {
        float vs = alphaScene - ORC_PI_F;
        a        = alphaModel - vs;
        if (a > twoPi)
            a = a - twoPi;
        if (a < 0.0f)
            a = a + twoPi;
    } else {
        float s = (float)((double)alphaScene - 3.14159265358979323846);
        a       = alphaModel - s;
        if (a < 0.0f)
            a = a + twoPi;
        if (a > twoPi)
            a = a - twoPi;
    }
    float fb  = floorf(sMaxId * a);
    int   bin = (int)fb;
    if (bin < 0)
        bin = 0;
    if (bin > angleNum - 1)
        bin = angleNum - 1;
    return bin;
}

/* ====================================================================== */
/* L3: model hash table                                                   */
/* ====================================================================== */

typedef struct {
    uint32_t key;
    uint32_t seq;
} keyseq;

static int cmp_keyseq(const void *a, const void *b) {
    const keyseq *x = (const keyseq *)a, *y = (const keyseq *)b;
    if (x->key != y->key)
        return x->key < y->key ? -1 : 1;
    return x->seq < y->seq ? -1 : (x->seq > y->seq);
}

static int is_finite3(const float *v) {
    return isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]);
}

/* trainModel hash-table stage, ppf.cpp:105-146: for every i with finite
 * n_i, every j != i with finite alpha_ij, push (i, alpha_ij) to bucket
 * key_ij.  Single-thread insertion order (i asc, j asc) is kept inside each
 * bucket; buckets are ordered by ascending key (the reference's gtl order is
 * implementation-defined). */
int orc_build_table(int n, const float *pts, const float *nrm, float angleStep, float distStep,
                    int refTail, orc_table *out) {
    memset(out, 0, sizeof *out);
    size_t  cap  = (size_t)n * (size_t)(n > 0 ? n - 1 : 0);
    keyseq *ks   = (keyseq *)malloc(sizeof(keyseq) * (cap ? cap : 1));
    uint32_t *rf = (uint32_t *)malloc(sizeof(uint32_t) * (cap ? cap : 1));
    float   *al  = (float *)malloc(sizeof(float) * (cap ? cap : 1));
    if (!ks || !rf || !al)
        return -1;
    int    tailStart = n - n % 8;
    size_t m         = 0;
    for (int i = 0; i < n; i++) {
        const float *p1 = pts + 3 * i, *n1 = nrm + 3 * i;
        if (!is_finite3(n1))
            continue;
        float rt[12];
        orc_frame(p1, n1, rt);
        for (int j = 0; j < n; j++) {
            if (j == i)
                continue;
            float a = orc_alpha(rt, pts + 3 * j);
            if (isnan(a))
                continue;
            int      tail = refTail && j >= tailStart;
            uint32_t key  = orc_ppf_key(p1, n1, pts + 3 * j, nrm + 3 * j, angleStep, distStep,
                                        tail, NULL, NULL);
            ks[m].key = key;
            ks[m].seq = (uint32_t)m;
            rf[m]     = (uint32_t)i;
            al[m]     = a;
            m++;
        }
    }
    qsort(ks, m, sizeof(keyseq), cmp_keyseq);
    int nk = 0;
    for (size_t e = 0; e < m; e++)
        if (e == 0 || ks[e].key != ks[e - 1].key)
            nk++;
    out->nKeys    = nk;
    out->nEntries = (int)m;
    out->keys     = (uint32_t *)malloc(sizeof(uint32_t) * (nk ? nk : 1));
    out->offs     = (int *)malloc(sizeof(int) * (nk + 1));
    out->refs     = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
    out->alphas   = (float *)malloc(sizeof(float) * (m ? m : 1));
    int kk        = -1;
    for (size_t e = 0; e < m; e++) {
        if (e == 0 || ks[e].key != ks[e - 1].key) {
            kk++;
            out->keys[kk] = ks[e].key;
            out->offs[kk] = (int)e;
        }
        out->refs[e]   = rf[ks[e].seq];
        out->alphas[e] = al[ks[e].seq];
    }
    out->offs[nk] = (int)m;
    free(ks);
    free(rf);
    free(al);
    return 0;
}

void orc_free_table(orc_table *t) {
    free(t->keys);
    free(t->offs);
    free(t->refs);
    free(t->alphas);
    memset(t, 0, sizeof *t);
}

int orc_table_find(const orc_table *t, uint32_t key) {
    int lo = 0, hi = t->nKeys - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        if (t->keys[mid] == key)
            return mid;
        if (t->keys[mid] < key)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    return -1;
}

/* ====================================================================== */
/* L4: vote, peaks, poses                                                 */
/* ====================================================================== */

/* One keypoint of the vote loop, ppf.cpp:301-346.  Every scene pair uses the
 * SIMD-body key formula (the reference's scalar-tail partners depend on
 * nanoflann traversal order, which a grid search does not reproduce); the
 * bucket-entry tail rule (last size%8 entries, ppf.cpp:317,332-340) depends
 * only on the table and is reproduced. */
void orc_vote_keypoint(const orc_table *t, int refNum, int angleNum, const float kp[3],
                       const float kn[3], int nNb, const float *nbPts, const float *nbNrm,
                       float angleStep, float distStep, int32_t *acc) {
    int stride = angleNum + 1;
    memset(acc, 0, sizeof(int32_t) * (size_t)refNum * stride);
    float rt[12];
    orc_frame(kp, kn, rt);
    for (int j = 0; j < nNb; j++) {
        const float *p2     = nbPts + 3 * j;
        float        alphaS = orc_alpha(rt, p2);
        uint32_t     key = orc_ppf_key(kp, kn, p2, nbNrm + 3 * j, angleStep, distStep, 0, NULL, NULL);
        int          b   = orc_table_find(t, key);
        if (b < 0 || isnan(alphaS))
            continue;
        int s = t->offs[b], e = t->offs[b + 1];
        int size = e - s, vec = size - size % 8;
        for (int x = s; x < e; x++) {
            int bin = orc_angle_bin(t->alphas[x], alphaS, angleNum, (x - s) >= vec);
            int32_t *a = acc + (size_t)t->refs[x] * stride;
            a[0]++;
            a[bin + 1]++;
        }
    }
}

/* Peak extraction (ppf.cpp:363-397): per ref, skip if count < thre; first
 * argmax bin j; skip if < thre; vote = bin j + both circular neighbours;
 * keep the 3 best by (vote desc, ref asc) (std::multiset with a
 * vote-greater comparator keeps insertion order among equals and erases
 * the last element); then keep those >= 0.95 * best. */
int orc_peaks(const int32_t *acc, int refNum, int angleNum, float voteThreshold,
              orc_candidate out[3]) {
    float         thre = voteThreshold / 2.0f;
    int           maxI = angleNum - 1;
    orc_candidate best[4];
    int           nb = 0;
    for (int i = 0; i < refNum; i++) {
        const int32_t *el = acc + (size_t)i * (angleNum + 1);
        if ((float)el[0] < thre)
            continue;
        const int32_t *b = el + 1;
        int            j = 0;
        for (int x = 1; x < angleNum; x++)
            if (b[x] > b[j])
                j = x;
        int v = b[j];
        if ((float)v < thre)
            continue;
        v += (j == 0) ? b[maxI] : b[j - 1];
        v += (j == maxI) ? b[0] : b[j + 1];
        orc_candidate c = {i, j, (float)v};
        /* insert after all entries with vote >= c.vote */
        int pos = nb;
        while (pos > 0 && best[pos - 1].vote < c.vote)
            pos--;
        for (int x = nb; x > pos; x--)
            best[x] = best[x - 1];
        best[pos] = c;
        if (nb < 3)
            nb++;
    }
    if (nb == 0)
        return 0;
    float t2 = (float)(best[0].vote * 0.95);
    int   no = 0;
    for (int x = 0; x < nb; x++)
        if (!(best[x].vote < t2))
            out[no++] = best[x];
    return no;
}

/* cos/sin of the pose angle of bin j (ppf.cpp:402, privateUtil.h:18-23):
 * alpha = (float)((double)(fl(2pi) * j / maxIdx) - pi); cos and sin are
 * evaluated in double and rounded to float. */
static void bin_cos_sin(int j, int angleNum, float *c, float *s) {
    float  a0 = ORC_2PI_F * (float)j / (float)(angleNum - 1);
    float  a  = (float)((double)a0 - 3.14159265358979323846);
    *c        = (float)cos((double)a);
    *s        = (float)sin((double)a);
}

/* T = inv(Ts) * (Rx(alpha) * Tm), rigid inverse [R^T, -R^T t]. */
void orc_pose(const float Ts[12], const float Tm[12], int angleId, int angleNum, float out[12]) {
    float c, s;
    bin_cos_sin(angleId, angleNum, &c, &s);
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

/* ====================================================================== */
/* L2: uniform grid                                                       */
/* ====================================================================== */

float orc_dist2(const float q[3], const float p[3]) {
    float dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
    float r  = 0.0f;
    r        = r + dx * dx;
    r        = r + dy * dy;
    r        = r + dz * dz;
    return r;
}

typedef struct {
    long long key;
    int       id;
} cellid;

static int cmp_cellid(const void *a, const void *b) {
    const cellid *x = (const cellid *)a, *y = (const cellid *)b;
    if (x->key != y->key)
        return x->key < y->key ? -1 : 1;
    return x->id < y->id ? -1 : (x->id > y->id);
}

static void grid_cell(const orc_grid *g, const float *p, int c[3]) {
    for (int a = 0; a < 3; a++) {
        int v = (int)floorf((p[a] - g->org[a]) / g->cell);
        if (v < 0)
            v = 0;
        if (v >= g->dim[a])
            v = g->dim[a] - 1;
        c[a] = v;
    }
}

int orc_grid_build(orc_grid *g, int n, const float *pts, const int *subset, int nSubset,
                   float cell) {
    memset(g, 0, sizeof *g);
    g->n    = n;
    g->pts  = pts;
    int m   = subset ? nSubset : n;
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int t = 0; t < m; t++) {
        const float *p = pts + 3 * (subset ? subset[t] : t);
        for (int a = 0; a < 3; a++) {
            if (p[a] < lo[a]) lo[a] = p[a];
            if (p[a] > hi[a]) hi[a] = p[a];
        }
    }
    if (m == 0)
        lo[0] = lo[1] = lo[2] = hi[0] = hi[1] = hi[2] = 0.0f;
    /* cap the cell count at ~8M */
    double ext = 1.0;
    for (int a = 0; a < 3; a++)
        ext *= (hi[a] - lo[a]) / cell + 1.0;
    while (ext > 8.0e6) {
        cell *= 1.5f;
        ext = 1.0;
        for (int a = 0; a < 3; a++)
            ext *= (hi[a] - lo[a]) / cell + 1.0;
    }
    g->cell = cell;
    for (int a = 0; a < 3; a++) {
        g->org[a] = lo[a];
        g->dim[a] = (int)((hi[a] - lo[a]) / cell) + 1;
    }
    long long nc = (long long)g->dim[0] * g->dim[1] * g->dim[2];
    cellid   *ci = (cellid *)malloc(sizeof(cellid) * (m ? m : 1));
    for (int t = 0; t < m; t++) {
        int id = subset ? subset[t] : t, c[3];
        grid_cell(g, pts + 3 * id, c);
        ci[t].key = ((long long)c[2] * g->dim[1] + c[1]) * g->dim[0] + c[0];
        ci[t].id  = id;
    }
    qsort(ci, m, sizeof(cellid), cmp_cellid);
    g->start = (int *)calloc((size_t)nc + 1, sizeof(int));
    g->idx   = (int *)malloc(sizeof(int) * (m ? m : 1));
    for (int t = 0; t < m; t++) {
        g->idx[t] = ci[t].id;
        g->start[ci[t].key + 1]++;
    }
    for (long long c = 0; c < nc; c++)
        g->start[c + 1] += g->start[c];
    free(ci);
    return 0;
}

void orc_grid_free(orc_grid *g) {
    free(g->start);
    free(g->idx);
    memset(g, 0, sizeof *g);
}

static int cmp_int(const void *a, const void *b) {
    int x = *(const int *)a, y = *(const int *)b;
    return x < y ? -1 : (x > y);
}

/* cell range covering [q - r, q + r] */
static void grid_range(const orc_grid *g, const float q[3], float r, int lo[3], int hi[3]) {
    for (int a = 0; a < 3; a++) {
        int l = (int)floorf((q[a] - r - g->org[a]) / g->cell);
        int h = (int)floorf((q[a] + r - g->org[a]) / g->cell);
        lo[a] = l < 0 ? 0 : l;
        hi[a] = h >= g->dim[a] ? g->dim[a] - 1 : h;
    }
}

/* Exact radius set (RadiusResultSet::addPoint keeps dist < r2,
 * nanoflann.hpp:396-400).  Points clamped into border cells are covered
 * because clamping only happens for points outside the build bbox, which
 * cannot occur for the indexed points themselves. */
int orc_grid_radius(const orc_grid *g, const float q[3], float r2, int *out, int cap) {
    int   lo[3], hi[3], cnt = 0;
    float r = sqrtf(r2) * 1.0001f + 1e-6f;
    grid_range(g, q, r, lo, hi);
    for (int z = lo[2]; z <= hi[2]; z++)
        for (int y = lo[1]; y <= hi[1]; y++)
            for (int x = lo[0]; x <= hi[0]; x++) {
                long long c = ((long long)z * g->dim[1] + y) * g->dim[0] + x;
                for (int t = g->start[c]; t < g->start[c + 1]; t++) {
                    int id = g->idx[t];
                    if (orc_dist2(q, g->pts + 3 * id) < r2) {
                        if (out && cnt < cap)
                            out[cnt] = id;
                        cnt++;
                    }
                }
            }
    if (out)
        qsort(out, cnt < cap ? cnt : cap, sizeof(int), cmp_int);
    return cnt;
}

static void knn_insert(int k, int *ids, float *ds, int *cnt, int id, float d) {
    int n = *cnt;
    if (n == k && !(d < ds[k - 1] || (d == ds[k - 1] && id < ids[k - 1])))
        return;
    int pos = n < k ? n : k - 1;
    while (pos > 0 && (ds[pos - 1] > d || (ds[pos - 1] == d && ids[pos - 1] > id))) {
        ds[pos]  = ds[pos - 1];
        ids[pos] = ids[pos - 1];
        pos--;
    }
    ds[pos]  = d;
    ids[pos] = id;
    if (n < k)
        *cnt = n + 1;
}

/* k nearest by (dist2, id): the reference's KNNResultSet keeps the first
 * traversed among equal distances; the canonical tie-break is lowest id. */
int orc_grid_knn(const orc_grid *g, const float q[3], int k, int *out, float *d2) {
    int cnt = 0;
    int c0[3];
    grid_cell(g, q, c0);
    int ring = 0, maxRing = g->dim[0] + g->dim[1] + g->dim[2];
    for (;;) {
        for (int z = c0[2] - ring; z <= c0[2] + ring; z++)
            for (int y = c0[1] - ring; y <= c0[1] + ring; y++)
                for (int x = c0[0] - ring; x <= c0[0] + ring; x++) {
                    if (x < 0 || y < 0 || z < 0 || x >= g->dim[0] || y >= g->dim[1] ||
                        z >= g->dim[2])
                        continue;
                    int onShell = (abs(x - c0[0]) == ring) || (abs(y - c0[1]) == ring) ||
                                  (abs(z - c0[2]) == ring);
                    if (!onShell)
                        continue;
                    long long c = ((long long)z * g->dim[1] + y) * g->dim[0] + x;
                    for (int t = g->start[c]; t < g->start[c + 1]; t++) {
                        int id = g->idx[t];
                        knn_insert(k, out, d2, &cnt, id, orc_dist2(q, g->pts + 3 * id));
                    }
                }
        /* every unvisited point is farther than ring*cell from q's cell */
        if (cnt == k) {
            float safe = (float)ring * g->cell;
            if (safe * safe > d2[k - 1] * 1.0001f + 1e-6f)
                break;
        }
        if (ring > maxRing)
            break;
        ring++;
    }
    return cnt;
}

/* ====================================================================== */
/* sampling                                                               */
/* ====================================================================== */

typedef struct {
    long long key;
    float     d;
    int       id;
} vox;

static int cmp_vox(const void *a, const void *b) {
    const vox *x = (const vox *)a, *y = (const vox *)b;
    if (x->key != y->key)
        return x->key < y->key ? -1 : 1;
    if (x->d != y->d)
        return x->d < y->d ? -1 : 1;
    return x->id < y->id ? -1 : (x->id > y->id);
}

/* Voxel-grid downsampling (replaces util.cpp:107-139's greedy Poisson
 * sampling): voxel v = floor((p - boxMin) / step) per axis; keep the point
 * nearest the voxel centre boxMin + (v + 0.5) step, ties to the lowest id;
 * output in ascending id.  Voxel coordinates are offset by 2^20 into a
 * 63-bit key. */
int orc_voxel_sample(int n, const float *pts, const int *subset, int nSubset, float step,
                     const float boxMin[3], int *out) {
    int  m = subset ? nSubset : n;
    vox *v = (vox *)malloc(sizeof(vox) * (m ? m : 1));
    for (int t = 0; t < m; t++) {
        int          id = subset ? subset[t] : t;
        const float *p  = pts + 3 * id;
        long long    key = 0;
        float        c[3];
        for (int a = 0; a < 3; a++) {
            int vi = (int)floorf((p[a] - boxMin[a]) / step);
            if (vi < -(1 << 20)) vi = -(1 << 20);
            if (vi > (1 << 20) - 1) vi = (1 << 20) - 1;
            c[a] = boxMin[a] + ((float)vi + 0.5f) * step;
            key  = (key << 21) | (long long)(vi + (1 << 20));
        }
        v[t].key = key;
        v[t].d   = orc_dist2(p, c);
        v[t].id  = id;
    }
    qsort(v, m, sizeof(vox), cmp_vox);
    int cnt = 0;
    for (int t = 0; t < m; t++)
        if (t == 0 || v[t].key != v[t - 1].key)
            out[cnt++] = v[t].id;
    free(v);
    qsort(out, cnt, sizeof(int), cmp_int);
    return cnt;
}

/* The reference's greedy Poisson-disk sampling (util.cpp:107-139) for a
 * given visiting order: keep a point unless an earlier kept point lies
 * within dist2 < step^2.  Output in visiting order. */
int orc_greedy_sample(int n, const float *pts, const int *order, int nOrder, float step, int *out) {
    char    *sup = (char *)calloc(n, 1);
    orc_grid g;
    orc_grid_build(&g, n, pts, order, nOrder, step);
    int  cnt = 0;
    int *nb  = (int *)malloc(sizeof(int) * (n ? n : 1));
    float r2 = step * step;
    for (int t = 0; t < nOrder; t++) {
        int id = order[t];
        if (sup[id])
            continue;
        out[cnt++] = id;
        int k      = orc_grid_radius(&g, pts + 3 * id, r2, nb, n);
        for (int x = 0; x < k; x++)
            sup[nb[x]] = 1;
    }
    orc_grid_free(&g);
    free(nb);
    free(sup);
    return cnt;
}

/* ====================================================================== */
/* symmetric eigen (cyclic Jacobi, Golub & Van Loan "Matrix Computations"  */
/* sec. 8.5), FP64                                                          */
/* ====================================================================== */

int orc_jacobi_eig(int n, double *a, double *v, double *w) {
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            v[i * n + j] = (i == j) ? 1.0 : 0.0;
    int sweep;
    for (sweep = 0; sweep < 60; sweep++) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < n; i++) {
            diag = diag + a[i * n + i] * a[i * n + i];
            for (int j = i + 1; j < n; j++)
                off = off + a[i * n + j] * a[i * n + j];
        }
        if (!(off > 1e-30 * diag) || off == 0.0)
            break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = a[p * n + q];
                if (apq == 0.0)
                    continue;
                double app = a[p * n + p], aqq = a[q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t     = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
                if (theta < 0.0)
                    t = -t;
                double c = 1.0 / sqrt(t * t + 1.0);
                double s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = c * akp - s * akq;
                    a[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = c * apk - s * aqk;
                    a[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = v[k * n + p], vkq = v[k * n + q];
                    v[k * n + p] = c * vkp - s * vkq;
                    v[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; i++)
        w[i] = a[i * n + i];
    return sweep;
}

/* ====================================================================== */
/* normals (util.cpp:234-334), snapshot semantics                         */
/* ====================================================================== */

/* PCA normal of one point from its k nearest neighbours (computeNormal,
 * util.cpp:234-270): centroid = sum / k (float, in (dist,id) order);
 * covariance of the centred points accumulated in FP32; eigen-solve in
 * FP64; eigenvector of the smallest |eigenvalue| (first on ties). */
static int pca_normal(const orc_grid *g, const float *pos, int id, int k, float cen[3],
                      float nrm[3], int *nbOut) {
    int   ids[64];
    float d2[64];
    int   cnt = orc_grid_knn(g, pos + 3 * id, k, ids, d2);
    if (cnt < 3)
        return 0;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (int t = 0; t < cnt; t++) {
        sx = sx + pos[3 * ids[t] + 0];
        sy = sy + pos[3 * ids[t] + 1];
        sz = sz + pos[3 * ids[t] + 2];
    }
    float fk = (float)cnt;
    cen[0] = sx / fk;
    cen[1] = sy / fk;
    cen[2] = sz / fk;
    float c00 = 0, c01 = 0, c02 = 0, c11 = 0, c12 = 0, c22 = 0;
    for (int t = 0; t < cnt; t++) {
        float x = pos[3 * ids[t] + 0] - cen[0];
        float y = pos[3 * ids[t] + 1] - cen[1];
        float z = pos[3 * ids[t] + 2] - cen[2];
        c00 = c00 + x * x;
        c01 = c01 + x * y;
        c02 = c02 + x * z;
        c11 = c11 + y * y;
        c12 = c12 + y * z;
        c22 = c22 + z * z;
    }
    double A[9] = {c00, c01, c02, c01, c11, c12, c02, c12, c22}, V[9], W[3];
    orc_jacobi_eig(3, A, V, W);
    int    mi = 0;
    for (int t = 1; t < 3; t++)
        if (fabs(W[t]) < fabs(W[mi]))
            mi = t;
    double vx = V[0 * 3 + mi], vy = V[1 * 3 + mi], vz = V[2 * 3 + mi];
    double nn = sqrt(vx * vx + vy * vy + vz * vz);
    nrm[0]    = (float)(vx / nn);
    nrm[1]    = (float)(vy / nn);
    nrm[2]    = (float)(vz / nn);
    if (nbOut)
        memcpy(nbOut, ids, sizeof(int) * cnt);
    return cnt;
}

/* ====================================================================== */
/* MLS normals                                                            */
/* ====================================================================== */

/* e^x after Cephes expf.c (S. L. Moshier): n = floor(x log2 e + 1/2),
 * g = (x - n*0.693359375) - n*(-2.12194440e-4), e^g by a degree-5
 * polynomial, times 2^n applied as two exact power-of-two factors. */
float orc_expf(float xx) {
    if (isnan(xx))
        return xx;
    if (xx > 88.72283905206835f)
        return INFINITY;
    if (xx < -103.278929903431851103f)
        return 0.0f;
    float t = 1.44269504088896341f * xx;
    float z = floorf(t + 0.5f);
    float x = xx - z * 0.693359375f;
    x       = x - z * -2.12194440e-4f;
    int   n = (int)z;
    float xx2 = x * x;
    float p   = 1.9875691500e-4f;
    p = p * x + 1.3981999507e-3f;
    p = p * x + 8.3334519073e-3f;
    p = p * x + 4.1665795894e-2f;
    p = p * x + 1.6666665459e-1f;
    p = p * x + 5.0000001201e-1f;
    float r  = p * xx2 + x;
    r        = r + 1.0f;
    int   na = n / 2, nb = n - na;
    union { uint32_t u; float f; } fa, fb;
    fa.u = (uint32_t)(na + 127) << 23;
    fb.u = (uint32_t)(nb + 127) << 23;
    return (r * fa.f) * fb.f;
}

/* Eigen's unitOrthogonal for 3-vectors: unless |x| and |y| are both
 * <= 1e-5 |z| (NumTraits<float>::dummy_precision), (-y, x, 0) times the
 * reciprocal of |(x, y)|; else (0, -z, y) times that of |(y, z)|. */
void orc_unit_orthogonal(const float n[3], float o[3]) {
    float lim = fabsf(n[2]) * 1e-5f;
    if (!(fabsf(n[0]) <= lim) || !(fabsf(n[1]) <= lim)) {
        float inv = 1.0f / sqrtf(n[0] * n[0] + n[1] * n[1]);
        o[0] = -n[1] * inv;
        o[1] = n[0] * inv;
        o[2] = 0.0f;
    } else {
        float inv = 1.0f / sqrtf(n[1] * n[1] + n[2] * n[2]);
        o[0] = 0.0f;
        o[1] = -n[2] * inv;
        o[2] = n[1] * inv;
    }
}

static float dot3f(const float a[3], const float b[3]) {
    float s = a[0] * b[0] + a[1] * b[1];
    return s + a[2] * b[2];
}

/* Cholesky solve of the packed lower triangle A (A[i(i+1)/2 + j], j <= i)
 * in double; 0 when a pivot is not positive. */
static int chol_solve(int nc, const double *A, double *x) {
    double L[64];
#define LL(i, j) L[(i) * ((i) + 1) / 2 + (j)]
    for (int j = 0; j < nc; j++) {
        double d = A[j * (j + 1) / 2 + j];
        for (int k = 0; k < j; k++)
            d -= LL(j, k) * LL(j, k);
        if (!(d > 0.0))
            return 0;
        LL(j, j) = sqrt(d);
        for (int i = j + 1; i < nc; i++) {
            double s = A[i * (i + 1) / 2 + j];
            for (int k = 0; k < j; k++)
                s -= LL(i, k) * LL(j, k);
            LL(i, j) = s / LL(j, j);
        }
    }
    for (int i = 0; i < nc; i++) {
        double s = x[i];
        for (int k = 0; k < i; k++)
            s -= LL(i, k) * x[k];
        x[i] = s / LL(i, i);
    }
    for (int i = nc - 1; i >= 0; i--) {
        double s = x[i];
        for (int k = i + 1; k < nc; k++)
            s -= LL(k, i) * x[k];
        x[i] = s / LL(i, i);
    }
#undef LL
    return 1;
}

/* estimateNormalMLS, util.cpp:336-457, per query (single thread, snapshot
 * positions; sums in ascending neighbour id order):
 *   neighbours: valid ids with dist2 < r^2 (RadiusResultSet); < 3: skip
 *   centroid = (sum p) / k, cov = sum (p - c)(p - c)^T in float (:364-367)
 *   n = eigenvector of the smallest |eigenvalue| (FP64 Jacobi; :369-374),
 *       written at once; non-finite: stop (:375-376)
 *   d = -(n.c); dist = p.n + d; mean = p - dist n (:377-381)
 *   v = unitOrthogonal(n), u = n x v (:383-385)
 *   for each neighbour (:400-426): e = q - mean, w = exp(-|e|^2 / r^2),
 *       (uc, vc, f) = (e.u, e.v, e.n), terms T = u^i v^j in the order
 *       i = 0..order, j = 0..order-i
 *   M[i][k] = sum (T_i w) T_k (k <= i), b[i] = sum (T_i w) f (:429-431)
 *   c = M^-1 b by FP64 Cholesky (:432); pivot <= 0 or c not finite: stop
 *   p' = mean + c0 n; n' = normalize((n - c[order+1] u) - c[1] v)
 *   flip n' when n'.(vp - p') < 0 (> 0 with invert), n'.z without a
 *   finite view point (:443-455). */
long orc_estimate_normals_mls(int n, float *pts, float *nrm, const int *valid, int nValid,
                              const int *query, int nQuery, float radius, int order, int invert,
                              const float vp[3]) {
    if (order < 1 || order > 3 || !(radius > 0.0f) || isinf(radius))
        return -1;
    const int nc = (order + 1) * (order + 2) / 2;
    float     r2 = radius * radius;
    float    *P0 = (float *)malloc(sizeof(float) * 3 * (size_t)(n ? n : 1));
    memcpy(P0, pts, sizeof(float) * 3 * (size_t)n);
    orc_grid g;
    orc_grid_build(&g, n, P0, valid, nValid, radius);
    int *nb = (int *)malloc(sizeof(int) * (size_t)(nValid ? nValid : 1));
    int  hasVp = vp && isfinite(vp[0]) && isfinite(vp[1]) && isfinite(vp[2]);
    long total = 0;
    for (int t = 0; t < nQuery; t++) {
        int          q  = query[t];
        const float *p  = P0 + 3 * (size_t)q;
        float       *on = nrm + 3 * (size_t)q;
        int cnt = 0;
        if (isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]))
            cnt = orc_grid_radius(&g, p, r2, nb, nValid);
        total += cnt;
        if (cnt < 3)
            continue;
        float s[3] = {0.0f, 0.0f, 0.0f};
        for (int j = 0; j < cnt; j++)
            for (int a = 0; a < 3; a++)
                s[a] = s[a] + P0[3 * (size_t)nb[j] + a];
        float fk     = (float)cnt;
        float cen[3] = {s[0] / fk, s[1] / fk, s[2] / fk};
        float C[6]   = {0, 0, 0, 0, 0, 0}; /* 00 01 02 11 12 22 */
        for (int j = 0; j < cnt; j++) {
            const float *w = P0 + 3 * (size_t)nb[j];
            float x = w[0] - cen[0], y = w[1] - cen[1], z = w[2] - cen[2];
            C[0] = C[0] + x * x;
            C[1] = C[1] + x * y;
            C[2] = C[2] + x * z;
            C[3] = C[3] + y * y;
            C[4] = C[4] + y * z;
            C[5] = C[5] + z * z;
        }
        double A[9] = {C[0], C[1], C[2], C[1], C[3], C[4], C[2], C[4], C[5]}, V[9], W[3];
        orc_jacobi_eig(3, A, V, W);
        int mi = 0;
        for (int k = 1; k < 3; k++)
            if (fabs(W[k]) < fabs(W[mi]))
                mi = k;
        double vl = sqrt(V[mi] * V[mi] + V[3 + mi] * V[3 + mi] + V[6 + mi] * V[6 + mi]);
        float  nn[3] = {(float)(V[mi] / vl), (float)(V[3 + mi] / vl), (float)(V[6 + mi] / vl)};
        memcpy(on, nn, sizeof nn);
        if (!(isfinite(nn[0]) && isfinite(nn[1]) && isfinite(nn[2])))
            continue;
        float d    = -dot3f(nn, cen);
        float dist = dot3f(p, nn) + d;
        float mean[3];
        for (int a = 0; a < 3; a++)
            mean[a] = p[a] - dist * nn[a];
        float va[3], ua[3];
        orc_unit_orthogonal(nn, va);
        ua[0] = nn[1] * va[2] - nn[2] * va[1];
        ua[1] = nn[2] * va[0] - nn[0] * va[2];
        ua[2] = nn[0] * va[1] - nn[1] * va[0];
        if (order <= 1 && nc > cnt)
            continue;
        float M[55], B[10];
        memset(M, 0, sizeof M);
        memset(B, 0, sizeof B);
        for (int j = 0; j < cnt; j++) {
            const float *w = P0 + 3 * (size_t)nb[j];
            float e[3]     = {w[0] - mean[0], w[1] - mean[1], w[2] - mean[2]};
            float sq       = e[0] * e[0] + e[1] * e[1];
            sq             = sq + e[2] * e[2];
            float wt       = orc_expf(-sq / r2);
            float uc = dot3f(e, ua), vc = dot3f(e, va), f = dot3f(e, nn);
            float T[10];
            int   m    = 0;
            float upow = 1.0f;
            for (int ui = 0; ui <= order; ui++) {
                float vpow = 1.0f;
                for (int vi = 0; vi <= order - ui; vi++) {
                    T[m++] = upow * vpow;
                    vpow   = vpow * vc;
                }
                upow = upow * uc;
            }
            for (int i = 0; i < nc; i++) {
                float pw = T[i] * wt;
                B[i]     = B[i] + pw * f;
                for (int k = 0; k <= i; k++)
                    M[i * (i + 1) / 2 + k] = M[i * (i + 1) / 2 + k] + pw * T[k];
            }
        }
        double Md[55], c[10];
        for (int i = 0; i < nc * (nc + 1) / 2; i++)
            Md[i] = M[i];
        for (int i = 0; i < nc; i++)
            c[i] = B[i];
        if (!chol_solve(nc, Md, c))
            continue;
        float cf[10];
        int   ok = 1;
        for (int i = 0; i < nc; i++) {
            cf[i] = (float)c[i];
            ok    = ok && isfinite(cf[i]);
        }
        if (!ok)
            continue;
        float np[3], n2[3];
        for (int a = 0; a < 3; a++) {
            np[a] = mean[a] + cf[0] * nn[a];
            n2[a] = (nn[a] - cf[order + 1] * ua[a]) - cf[1] * va[a];
        }
        float l2 = n2[0] * n2[0] + n2[1] * n2[1];
        l2       = l2 + n2[2] * n2[2];
        if (l2 > 0.0f) {
            float l = sqrtf(l2);
            for (int a = 0; a < 3; a++)
                n2[a] = n2[a] / l;
        }
        float val;
        if (hasVp) {
            float to[3] = {vp[0] - np[0], vp[1] - np[1], vp[2] - np[2]};
            val         = dot3f(n2, to);
        } else {
            float zz[3] = {0.0f, 0.0f, 1.0f};
            val         = dot3f(n2, zz);
        }
        if (invert ? val > 0.0f : val < 0.0f)
            for (int a = 0; a < 3; a++)
                n2[a] = -n2[a];
        memcpy(pts + 3 * (size_t)q, np, sizeof np);
        memcpy(on, n2, sizeof n2);
    }
    orc_grid_free(&g);
    free(nb);
    free(P0);
    return total;
}

static void normalize3(float *n) {
    float l = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    n[0]    = n[0] / l;
    n[1]    = n[1] / l;
    n[2]    = n[2] / l;
}

/* estimateNormal (util.cpp:272-334) with snapshot semantics: every kNN and
 * PCA reads the positions and normals as they were when the call started;
 * all write-backs (point := centroid, normal) happen afterwards.  The
 * reference mutates points while other threads read them (util.cpp:269,
 * 292-295).  `valid` lists the points the search runs over. */
int orc_estimate_normals(int n, float *pts, float *nrm, const int *valid, int nValid,
                         const int *query, int nQuery, int k, int smooth, int invert,
                         const float vp[3]) {
    if (k < 3)
        k = 3;
    if (k > 64)
        k = 64;
    float *pos0 = (float *)malloc(sizeof(float) * 3 * (size_t)n);
    float *nrm0 = (float *)malloc(sizeof(float) * 3 * (size_t)n);
    memcpy(pos0, pts, sizeof(float) * 3 * (size_t)n);
    memcpy(nrm0, nrm, sizeof(float) * 3 * (size_t)n);
    orc_grid g;
    float    ext = 0.0f;
    {
        /* cell ~ expected kNN radius: cube root of volume per point * 2 */
        float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (int t = 0; t < nValid; t++)
            for (int a = 0; a < 3; a++) {
                float v = pos0[3 * valid[t] + a];
                if (v < lo[a]) lo[a] = v;
                if (v > hi[a]) hi[a] = v;
            }
        double vol = 1.0;
        for (int a = 0; a < 3; a++)
            vol *= (hi[a] - lo[a]) + 1e-3;
        ext = (float)cbrt(vol / (nValid > 0 ? nValid : 1)) * 2.0f;
        if (!(ext > 0.0f))
            ext = 1.0f;
    }
    orc_grid_build(&g, n, pos0, valid, nValid, ext);

    /* raw PCA results, computed lazily for every point that needs one */
    float *rawN = (float *)malloc(sizeof(float) * 3 * (size_t)n);
    float *cen  = (float *)malloc(sizeof(float) * 3 * (size_t)n);
    char  *have = (char *)calloc(n, 1);
    int   *nbq  = (int *)malloc(sizeof(int) * 64 * (size_t)(nQuery ? nQuery : 1));
    int   *cntq = (int *)calloc(nQuery ? nQuery : 1, sizeof(int));
#pragma omp parallel for schedule(dynamic, 64)
    for (int t = 0; t < nQuery; t++) {
        int id = query[t];
        if (is_finite3(nrm0 + 3 * id))
            continue; /* computeNormal returns early, util.cpp:238-239 */
        cntq[t] = pca_normal(&g, pos0, id, k, cen + 3 * id, rawN + 3 * id, nbq + 64 * t);
    }
    for (int t = 0; t < nQuery; t++)
        if (cntq[t])
            have[query[t]] = 1;
    if (smooth) {
        /* neighbours still lacking a normal get their own PCA (util.cpp:292-295) */
        char *need = (char *)calloc(n, 1);
        for (int t = 0; t < nQuery; t++)
            for (int x = 0; x < cntq[t]; x++) {
                int j = nbq[64 * t + x];
                if (!have[j] && !is_finite3(nrm0 + 3 * j))
                    need[j] = 1;
            }
#pragma omp parallel for schedule(dynamic, 256)
        for (int j = 0; j < n; j++)
            if (need[j] && pca_normal(&g, pos0, j, k, cen + 3 * j, rawN + 3 * j, NULL))
                need[j] = 2;
        for (int j = 0; j < n; j++)
            if (need[j] == 2)
                have[j] = 1;
        free(need);
    }
    /* final normals of the queries */
    float *outN = (float *)malloc(sizeof(float) * 3 * (size_t)(nQuery ? nQuery : 1));
    for (int t = 0; t < nQuery; t++) {
        int id = query[t];
        if (!cntq[t]) {
            memcpy(outN + 3 * t, nrm0 + 3 * id, sizeof(float) * 3);
            continue;
        }
        float *n0 = rawN + 3 * id;
        if (!smooth) {
            memcpy(outN + 3 * t, n0, sizeof(float) * 3);
            continue;
        }
        float s[3] = {0.0f, 0.0f, 0.0f};
        for (int x = 0; x < cntq[t]; x++) {
            int          j  = nbq[64 * t + x];
            const float *nj = have[j] ? rawN + 3 * j : nrm0 + 3 * j;
            float        d  = nj[0] * n0[0] + nj[1] * n0[1] + nj[2] * n0[2];
            if (d > 0.0f) {
                s[0] = s[0] + nj[0];
                s[1] = s[1] + nj[1];
                s[2] = s[2] + nj[2];
            } else {
                s[0] = s[0] - nj[0];
                s[1] = s[1] - nj[1];
                s[2] = s[2] - nj[2];
            }
        }
        normalize3(s);
        memcpy(outN + 3 * t, s, sizeof s);
    }
    /* write-back: every point with a fresh PCA moves to its centroid and
     * takes its raw normal; queries then take their final normal */
    for (int i = 0; i < n; i++)
        if (have[i]) {
            memcpy(pts + 3 * i, cen + 3 * i, sizeof(float) * 3);
            memcpy(nrm + 3 * i, rawN + 3 * i, sizeof(float) * 3);
        }
    for (int t = 0; t < nQuery; t++)
        memcpy(nrm + 3 * query[t], outN + 3 * t, sizeof(float) * 3);
    /* orientation (util.cpp:309-333), against the moved position */
    int vpFinite = vp && isfinite(vp[0]) && isfinite(vp[1]) && isfinite(vp[2]);
    for (int t = 0; t < nQuery; t++) {
        float *nn = nrm + 3 * query[t];
        float *pp = pts + 3 * query[t];
        float  d;
        if (vpFinite) {
            float vx = vp[0] - pp[0], vy = vp[1] - pp[1], vz = vp[2] - pp[2];
            d        = nn[0] * vx + nn[1] * vy + nn[2] * vz;
        } else {
            d = nn[0] * 0.0f + nn[1] * 0.0f + nn[2] * 1.0f;
        }
        int flip = invert ? (d > 0.0f) : (d < 0.0f);
        if (flip) {
            nn[0] = -nn[0];
            nn[1] = -nn[1];
            nn[2] = -nn[2];
        }
    }
    orc_grid_free(&g);
    free(pos0);
    free(nrm0);
    free(rawN);
    free(cen);
    free(have);
    free(nbq);
    free(cntq);
    free(outN);
    return 0;
}

/* ====================================================================== */
/* rotations, quaternions                                                 */
/* ====================================================================== */

/* quaternion of a rotation matrix (Shoemake's method, the branch structure
 * Eigen uses for Quaternion(Matrix3)), q = (w, x, y, z). */
void orc_quat_from_rot(const float R[9], float q[4]) {
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
        t        = sqrtf(R[i * 3 + i] - R[j * 3 + j] - R[k * 3 + k] + 1.0f);
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

/* rotation matrix of a unit quaternion (the standard expansion used by
 * Quaternion::toRotationMatrix). */
void orc_rot_from_quat(const float q[4], float R[9]) {
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

/* AngleAxis(R).angle(): 2 atan2(|q.vec|, |q.w|) of R's quaternion. */
float orc_rot_angle(const float R[9]) {
    float q[4];
    orc_quat_from_rot(R, q);
    float n = sqrtf(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    return 2.0f * orc_atan2f(n, fabsf(q[0]));
}

static void pose_fill(orc_pose_t *p) {
    float R[9] = {p->pose[0], p->pose[1], p->pose[2], p->pose[4], p->pose[5],
                  p->pose[6], p->pose[8], p->pose[9], p->pose[10]};
    orc_quat_from_rot(R, p->q);
    float n  = sqrtf(p->q[1] * p->q[1] + p->q[2] * p->q[2] + p->q[3] * p->q[3]);
    p->angle = 2.0f * orc_atan2f(n, fabsf(p->q[0]));
}

/* ====================================================================== */
/* pose clustering (util.cpp:486-601)                                      */
/* ====================================================================== */

typedef struct {
    float votes;
    int   seq;
} vs_t;
static int cmp_vs(const void *a, const void *b) {
    const vs_t *x = (const vs_t *)a, *y = (const vs_t *)b;
    if (x->votes != y->votes)
        return x->votes > y->votes ? -1 : 1;
    return x->seq < y->seq ? -1 : (x->seq > y->seq);
}

/* stable order by votes descending (sortPoses, util.cpp:486-491, uses the
 * unstable std::sort; ties keep input order here). */
static int *order_by_votes(const orc_pose_t *in, int n) {
    vs_t *v = (vs_t *)malloc(sizeof(vs_t) * (n ? n : 1));
    for (int i = 0; i < n; i++) {
        v[i].votes = in[i].votes;
        v[i].seq   = i;
    }
    qsort(v, n, sizeof(vs_t), cmp_vs);
    int *o = (int *)malloc(sizeof(int) * (n ? n : 1));
    for (int i = 0; i < n; i++)
        o[i] = v[i].seq;
    free(v);
    return o;
}

static int compare_pose(const orc_pose_t *a, const orc_pose_t *b, float dThr, float aThr) {
    float dx = a->pose[3] - b->pose[3], dy = a->pose[7] - b->pose[7], dz = a->pose[11] - b->pose[11];
    float d   = sqrtf(dx * dx + dy * dy + dz * dz);
    float phi = fabsf(a->angle - b->angle);
    return d < dThr && phi < aThr;
}

/* clusterPose (util.cpp:499-520) + avgClusters (:579-601) with the Markley
 * quaternion mean (:553-577, largest eigenvector of sum(q q^T)/M, here an
 * FP64 Jacobi solve with the sign fixed to w >= 0).  clusterOf[i] gets the
 * cluster of input pose i; centersOut[c] the averaged pose of cluster c
 * (clusters in creation order). */
int orc_cluster_poses(const orc_pose_t *in, int n, float dThr, float aThr, int *clusterOf,
                      orc_pose_t *out) {
    int *ord    = order_by_votes(in, n);
    int *center = (int *)malloc(sizeof(int) * (n ? n : 1));
    int  nc     = 0;
    for (int t = 0; t < n; t++) {
        const orc_pose_t *p  = in + ord[t];
        int               cl = -1;
        for (int c = 0; c < nc; c++)
            if (compare_pose(p, in + center[c], dThr, aThr)) {
                cl = c;
                break;
            }
        if (cl < 0) {
            cl           = nc;
            center[nc++] = ord[t];
        }
        clusterOf[ord[t]] = cl;
    }
    for (int c = 0; c < nc; c++) {
        float  tx = 0, ty = 0, tz = 0, votes = 0;
        float  A[16] = {0};
        int    m     = 0;
        for (int t = 0; t < n; t++) {
            int i = ord[t];
            if (clusterOf[i] != c)
                continue;
            tx    = tx + in[i].pose[3];
            ty    = ty + in[i].pose[7];
            tz    = tz + in[i].pose[11];
            votes = votes + in[i].votes;
            for (int r = 0; r < 4; r++)
                for (int s = 0; s < 4; s++)
                    A[r * 4 + s] = A[r * 4 + s] + in[i].q[r] * in[i].q[s];
            m++;
        }
        double Ad[16], V[16], W[4];
        for (int r = 0; r < 16; r++)
            Ad[r] = (double)(A[r] / (float)m);
        orc_jacobi_eig(4, Ad, V, W);
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
        orc_rot_from_quat(q, R);
        orc_pose_t *o = out + c;
        float       fm = (float)m;
        for (int r = 0; r < 3; r++) {
            o->pose[r * 4 + 0] = R[r * 3 + 0];
            o->pose[r * 4 + 1] = R[r * 3 + 1];
            o->pose[r * 4 + 2] = R[r * 3 + 2];
        }
        o->pose[3]  = tx / fm;
        o->pose[7]  = ty / fm;
        o->pose[11] = tz / fm;
        o->votes    = votes;
        memcpy(o->q, q, sizeof q);
        o->angle = orc_rot_angle(R);
    }
    free(ord);
    free(center);
    return nc;
}

/* sortPoses + clusterPose2 (util.cpp:522-551): greedy overlap NMS on
 * R c + t; merged votes go to the earlier pose. */
int orc_nms_poses(const orc_pose_t *in, int n, const float c[3], float thr, orc_pose_t *out) {
    int   *ord = order_by_votes(in, n);
    float *tr  = (float *)malloc(sizeof(float) * 3 * (n ? n : 1));
    char  *used = (char *)calloc(n ? n : 1, 1);
    for (int t = 0; t < n; t++) {
        const float *P = in[ord[t]].pose;
        for (int r = 0; r < 3; r++)
            tr[3 * t + r] = P[r * 4 + 0] * c[0] + P[r * 4 + 1] * c[1] + P[r * 4 + 2] * c[2] + P[r * 4 + 3];
    }
    float sq  = thr * thr;
    int   cnt = 0;
    for (int i = 0; i < n; i++) {
        if (used[i])
            continue;
        orc_pose_t pi = in[ord[i]];
        for (int j = i + 1; j < n; j++) {
            if (used[j])
                continue;
            float dx = tr[3 * i] - tr[3 * j], dy = tr[3 * i + 1] - tr[3 * j + 1],
                  dz = tr[3 * i + 2] - tr[3 * j + 2];
            if (dx * dx + dy * dy + dz * dz < sq) {
                pi.votes = pi.votes + in[ord[j]].votes;
                used[j]  = 1;
            }
        }
        out[cnt++] = pi;
    }
    free(ord);
    free(tr);
    free(used);
    return cnt;
}

/* ====================================================================== */
/* ICP (icp.cpp), FP64 refinement                                          */
/* ====================================================================== */

static void xform_pt(const double T[12], const float p[3], float o[3]) {
    for (int r = 0; r < 3; r++)
        o[r] = (float)(T[r * 4] * p[0] + T[r * 4 + 1] * p[1] + T[r * 4 + 2] * p[2] + T[r * 4 + 3]);
}
static void xform_dir(const double T[12], const float p[3], float o[3]) {
    for (int r = 0; r < 3; r++)
        o[r] = (float)(T[r * 4] * p[0] + T[r * 4 + 1] * p[1] + T[r * 4 + 2] * p[2]);
}
static void compose(const double A[12], const double B[12], double C[12]) {
    double t[12];
    for (int r = 0; r < 3; r++) {
        for (int c = 0; c < 4; c++) {
            double v = A[r * 4 + 0] * B[0 * 4 + c] + A[r * 4 + 1] * B[1 * 4 + c] +
                       A[r * 4 + 2] * B[2 * 4 + c];
            if (c == 3)
                v += A[r * 4 + 3];
            t[r * 4 + c] = v;
        }
    }
    memcpy(C, t, sizeof t);
}

/* 1-NN (findClosestPoint, util.cpp:603-626): smallest (dist2, id). */
static int nn1(const orc_grid *g, const float q[3], float *d2) {
    int   id;
    int   c = orc_grid_knn(g, q, 1, &id, d2);
    return c ? id : -1;
}

/* u = pinv(A) rhs for symmetric PSD A via FP64 Jacobi; singular values
 * below 6*eps_f32*max are dropped (JacobiSVD::solve's threshold). */
static void sym_solve6(const double A[36], const double rhs[6], double u[6]) {
    double a[36], V[36], W[6];
    memcpy(a, A, sizeof a);
    orc_jacobi_eig(6, a, V, W);
    double mx = 0;
    for (int i = 0; i < 6; i++)
        if (fabs(W[i]) > mx)
            mx = fabs(W[i]);
    double thr = mx * 6.0 * 1.1920928955078125e-07;
    for (int i = 0; i < 6; i++)
        u[i] = 0;
    for (int k = 0; k < 6; k++) {
        if (!(fabs(W[k]) > thr))
            continue;
        double proj = 0;
        for (int i = 0; i < 6; i++)
            proj += V[i * 6 + k] * rhs[i];
        proj /= W[k];
        for (int i = 0; i < 6; i++)
            u[i] += V[i * 6 + k] * proj;
    }
}

/* one ICP step: correspondences (icp.cpp:123-167), solve (symmetric
 * :39-87 or point-to-plane), mean NN distance after the step (:182-194). */
typedef struct {
    double pose[12];
    int    valid;
    float  mse;
} iter_t;

static iter_t icp_iteration(int nSrc, const float *sp, const float *sn, const orc_grid *dst,
                            const float *dstN, float rejectDist, int mode, int *nnId,
                            float *nnD, int *best) {
    iter_t r;
    memset(&r, 0, sizeof r);
    for (int i = 0; i < 12; i++)
        r.pose[i] = (i % 5 == 0) ? 1.0 : 0.0;
    r.mse = FLT_MAX;
    float rej2 = rejectDist * rejectDist;
    /* dedupe: per dst point keep the closest src (ties: lowest src index);
     * pairs ordered by first occurrence of the dst id in src order. */
    int npair = 0;
    int *order = (int *)malloc(sizeof(int) * (nSrc ? nSrc : 1));
#pragma omp parallel for schedule(static, 64)
    for (int i = 0; i < nSrc; i++) {
        float d;
        nnId[i] = nn1(dst, sp + 3 * i, &d);
        nnD[i]  = d;
    }
    for (int i = 0; i < nSrc; i++) {
        float d = nnD[i];
        if (nnId[i] < 0 || d > rej2)
            continue;
        int q = nnId[i];
        if (best[q] < 0) {
            best[q]        = i;
            order[npair++] = q;
        } else if (d < nnD[best[q]]) {
            best[q] = i;
        }
    }
    r.valid = npair;
    if (npair < 6) {
        for (int t = 0; t < npair; t++)
            best[order[t]] = -1;
        free(order);
        return r;
    }
    double U[12];
    if (mode == 0) {
        /* symmetric objective (Rusinkiewicz 2019), icp.cpp:39-87 */
        double pm[3] = {0, 0, 0}, qm[3] = {0, 0, 0};
        for (int t = 0; t < npair; t++) {
            int q = order[t], i = best[q];
            for (int a = 0; a < 3; a++) {
                pm[a] += sp[3 * i + a];
                qm[a] += dst->pts[3 * q + a];
            }
        }
        for (int a = 0; a < 3; a++) {
            pm[a] /= npair;
            qm[a] /= npair;
        }
        double A[36] = {0}, b[6] = {0};
        for (int t = 0; t < npair; t++) {
            int    q = order[t], i = best[q];
            double p[3], qq[3], nn[3], s[3], x[6];
            for (int a = 0; a < 3; a++) {
                p[a]  = sp[3 * i + a] - pm[a];
                qq[a] = dst->pts[3 * q + a] - qm[a];
                nn[a] = (double)sn[3 * i + a] + (double)dstN[3 * q + a];
                s[a]  = p[a] + qq[a];
            }
            double bi = (p[0] - qq[0]) * nn[0] + (p[1] - qq[1]) * nn[1] + (p[2] - qq[2]) * nn[2];
            x[0] = s[1] * nn[2] - s[2] * nn[1];
            x[1] = s[2] * nn[0] - s[0] * nn[2];
            x[2] = s[0] * nn[1] - s[1] * nn[0];
            x[3] = nn[0];
            x[4] = nn[1];
            x[5] = nn[2];
            for (int r2 = 0; r2 < 6; r2++) {
                for (int c2 = 0; c2 < 6; c2++)
                    A[r2 * 6 + c2] += x[r2] * x[c2];
                b[r2] += bi * x[r2];
            }
        }
        double rhs[6], u[6];
        for (int k = 0; k < 6; k++)
            rhs[k] = -b[k];
        sym_solve6(A, rhs, u);
        double an    = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        double theta = atan(an);
        double ax[3] = {0, 0, 0};
        if (an > 0)
            for (int a = 0; a < 3; a++)
                ax[a] = u[a] / an;
        double W[9] = {0, -ax[2], ax[1], ax[2], 0, -ax[0], -ax[1], ax[0], 0}, W2[9], Ri[9], R[9];
        for (int r2 = 0; r2 < 3; r2++)
            for (int c2 = 0; c2 < 3; c2++) {
                double s2 = 0;
                for (int k = 0; k < 3; k++)
                    s2 += W[r2 * 3 + k] * W[k * 3 + c2];
                W2[r2 * 3 + c2] = s2;
            }
        double st = sin(theta), ct = cos(theta);
        for (int k = 0; k < 9; k++)
            Ri[k] = ((k % 4 == 0) ? 1.0 : 0.0) + st * W[k] + (1.0 - ct) * W2[k];
        for (int r2 = 0; r2 < 3; r2++)
            for (int c2 = 0; c2 < 3; c2++) {
                double s2 = 0;
                for (int k = 0; k < 3; k++)
                    s2 += Ri[r2 * 3 + k] * Ri[k * 3 + c2];
                R[r2 * 3 + c2] = s2;
            }
        double t2[3] = {ct * u[3], ct * u[4], ct * u[5]};
        for (int r2 = 0; r2 < 3; r2++) {
            double v = 0;
            for (int k = 0; k < 3; k++)
                v += R[r2 * 3 + k] * (-pm[k]) + Ri[r2 * 3 + k] * t2[k];
            U[r2 * 4 + 0] = R[r2 * 3 + 0];
            U[r2 * 4 + 1] = R[r2 * 3 + 1];
            U[r2 * 4 + 2] = R[r2 * 3 + 2];
            U[r2 * 4 + 3] = v + qm[r2];
        }
    } else {
        /* point-to-plane about the source centroid c (the de-meaning of the
         * reference's point_to_plane, icp.cpp:304-308, with unit weights):
         * r = n_q.(p - q), J = [(p - c) x n_q ; n_q], x = -(J^T J)^+ J^T r,
         * update p -> R(w)(p - c) + c + v */
        double A[36] = {0}, b[6] = {0}, cm[3] = {0, 0, 0};
        for (int t = 0; t < npair; t++)
            for (int a = 0; a < 3; a++)
                cm[a] += sp[3 * best[order[t]] + a];
        for (int a = 0; a < 3; a++)
            cm[a] /= npair;
        for (int t = 0; t < npair; t++) {
            int    q = order[t], i = best[q];
            double p[3], qq[3], nn[3], x[6], d[3];
            for (int a = 0; a < 3; a++) {
                p[a]  = sp[3 * i + a];
                qq[a] = dst->pts[3 * q + a];
                nn[a] = dstN[3 * q + a];
                d[a]  = p[a] - cm[a];
            }
            double ri = (p[0] - qq[0]) * nn[0] + (p[1] - qq[1]) * nn[1] + (p[2] - qq[2]) * nn[2];
            x[0] = d[1] * nn[2] - d[2] * nn[1];
            x[1] = d[2] * nn[0] - d[0] * nn[2];
            x[2] = d[0] * nn[1] - d[1] * nn[0];
            x[3] = nn[0];
            x[4] = nn[1];
            x[5] = nn[2];
            for (int r2 = 0; r2 < 6; r2++) {
                for (int c2 = 0; c2 < 6; c2++)
                    A[r2 * 6 + c2] += x[r2] * x[c2];
                b[r2] += ri * x[r2];
            }
        }
        double rhs[6], u[6];
        for (int k = 0; k < 6; k++)
            rhs[k] = -b[k];
        sym_solve6(A, rhs, u);
        double th = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        double R[9];
        if (th > 0) {
            double k[3] = {u[0] / th, u[1] / th, u[2] / th}, c = cos(th), s = sin(th);
            double K[9] = {0, -k[2], k[1], k[2], 0, -k[0], -k[1], k[0], 0};
            for (int r2 = 0; r2 < 3; r2++)
                for (int c2 = 0; c2 < 3; c2++) {
                    double kk = 0;
                    for (int m = 0; m < 3; m++)
                        kk += K[r2 * 3 + m] * K[m * 3 + c2];
                    R[r2 * 3 + c2] = ((r2 == c2) ? 1.0 : 0.0) + s * K[r2 * 3 + c2] + (1 - c) * kk;
                }
        } else {
            for (int k = 0; k < 9; k++)
                R[k] = (k % 4 == 0) ? 1.0 : 0.0;
        }
        for (int r2 = 0; r2 < 3; r2++) {
            double Rc = R[r2 * 3 + 0] * cm[0] + R[r2 * 3 + 1] * cm[1] + R[r2 * 3 + 2] * cm[2];
            U[r2 * 4 + 0] = R[r2 * 3 + 0];
            U[r2 * 4 + 1] = R[r2 * 3 + 1];
            U[r2 * 4 + 2] = R[r2 * 3 + 2];
            U[r2 * 4 + 3] = cm[r2] - Rc + u[3 + r2];
        }
    }
    memcpy(r.pose, U, sizeof U);
    /* mean NN distance of the corresponded src points after the step */
    double *dd = (double *)malloc(sizeof(double) * npair);
#pragma omp parallel for schedule(static, 64)
    for (int t = 0; t < npair; t++) {
        int   i = best[order[t]];
        float pp[3], d;
        xform_pt(U, sp + 3 * i, pp);
        nn1(dst, pp, &d);
        dd[t] = sqrt((double)d);
    }
    double sum = 0;
    for (int t = 0; t < npair; t++)
        sum += dd[t];
    free(dd);
    r.mse = (float)(sum / npair);
    for (int t = 0; t < npair; t++)
        best[order[t]] = -1;
    free(order);
    return r;
}

static int near_identity(const float T[12]) {
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 4; c++) {
            float v = T[r * 4 + c], e = (r == c) ? 1.0f : 0.0f;
            if (fabsf(v - e) > 1e-5f)
                return 0;
        }
    return 1;
}

/* ICP::regist (icp.cpp:214-278) including its quirks: a near-identity
 * initial pose leaves src untransformed (:231); src keeps moving by each
 * step even when the step is rejected (rate <= 0, :261-271).  mode 0 =
 * symmetric objective (the reference's live solver), 1 = point-to-plane. */
int orc_icp_regist(int nSrc, const float *srcP, const float *srcN, const orc_grid *dst,
                   const float *dstN, const float init[12], const orc_icp_criteria *c, int mode,
                   orc_icp_result *res) {
    double pose[12], cur[12];
    for (int i = 0; i < 12; i++)
        pose[i] = init[i];
    if (near_identity(init)) {
        for (int i = 0; i < 12; i++)
            cur[i] = (i % 5 == 0) ? 1.0 : 0.0;
    } else {
        memcpy(cur, pose, sizeof pose);
    }
    float *sp   = (float *)malloc(sizeof(float) * 3 * (size_t)(nSrc ? nSrc : 1));
    float *sn   = (float *)malloc(sizeof(float) * 3 * (size_t)(nSrc ? nSrc : 1));
    int   *nnId = (int *)malloc(sizeof(int) * (nSrc ? nSrc : 1));
    float *nnD  = (float *)malloc(sizeof(float) * (nSrc ? nSrc : 1));
    int   *best = (int *)malloc(sizeof(int) * (dst->n ? dst->n : 1));
    for (int i = 0; i < dst->n; i++)
        best[i] = -1;
    res->type         = 0;
    res->mse          = FLT_MAX;
    res->convergeRate = 1.0f;
    res->iterations   = 0;
    res->converged    = 0;
    while (res->iterations < c->iterations) {
        for (int i = 0; i < nSrc; i++) {
            xform_pt(cur, srcP + 3 * i, sp + 3 * i);
            xform_dir(cur, srcN + 3 * i, sn + 3 * i);
        }
        iter_t it = icp_iteration(nSrc, sp, sn, dst, dstN, c->rejectDist, mode, nnId, nnD, best);
        if (it.valid > 6 && it.mse < c->mseMax)
            res->converged = 1;
        int stop = 0;
        if (it.valid < 6) {
            res->type = 3;
            stop      = 1;
        } else {
            if (it.mse < c->mseMin) {
                res->type = 1;
                stop      = 1;
            }
            float rate = (res->mse - it.mse) / res->mse;
            if (rate < c->tolerance) {
                res->type = 2;
                stop      = 1;
            }
            if (res->iterations++ >= c->iterations) {
                res->type = 0;
                stop      = 1;
            }
            if (rate > 0) {
                res->mse          = it.mse;
                res->convergeRate = rate;
                compose(it.pose, pose, pose);
            }
        }
        if (stop)
            break;
        compose(it.pose, cur, cur);
    }
    for (int i = 0; i < 12; i++)
        res->pose[i] = (float)pose[i];
    free(sp);
    free(sn);
    free(nnId);
    free(nnD);
    free(best);
    return 0;
}

/* inliner (util.cpp:628-641): count of transformed src points whose nearest
 * neighbour is within dist (dist2 < dist^2). */
int orc_inliers(int nSrc, const float *srcP, const float pose[12], const orc_grid *g, float dist) {
    double T[12];
    for (int i = 0; i < 12; i++)
        T[i] = pose[i];
    float r2  = dist * dist;
    int   cnt = 0;
#pragma omp parallel for reduction(+ : cnt) schedule(static, 64)
    for (int i = 0; i < nSrc; i++) {
        float p[3];
        /* transformPointCloud in float: R p + t */
        for (int r = 0; r < 3; r++)
            p[r] = pose[r * 4] * srcP[3 * i] + pose[r * 4 + 1] * srcP[3 * i + 1] +
                   pose[r * 4 + 2] * srcP[3 * i + 2] + pose[r * 4 + 3];
        (void)T;
        if (orc_grid_radius(g, p, r2, NULL, 0) > 0)
            cnt++;
    }
    return cnt;
}

/* Point-to-plane system in FP64 for fixed correspondences (the system of
 * icp.cpp:289-324): x_i = U(xi) T p_i with U(xi) = [R(w), c - R(w) c + v],
 * r_i = n_q.(x_i - q_i), J_i = [(x_i - c) x n_q ; n_q] (icp.cpp:318-320),
 * weights 1 or TukeyLossWeight(|r|, 2) (icp.cpp:284-287).
 * out[28] = {cost = 1/2 sum w r^2, grad[6] = sum w r J, JtJ[21] upper}. */
void orc_p2pl_system(int nSrc, const float *srcP, const double T[12], const int *corr,
                     const float *dstP, const float *dstN, const double c[3], const double xi[6],
                     int tukey, double out[28]) {
    double R[9], th = sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]);
    if (th > 0) {
        double k[3] = {xi[0] / th, xi[1] / th, xi[2] / th}, cs = cos(th), sn = sin(th);
        double K[9] = {0, -k[2], k[1], k[2], 0, -k[0], -k[1], k[0], 0};
        for (int r = 0; r < 3; r++)
            for (int q = 0; q < 3; q++) {
                double kk = 0;
                for (int m = 0; m < 3; m++)
                    kk += K[r * 3 + m] * K[m * 3 + q];
                R[r * 3 + q] = ((r == q) ? 1.0 : 0.0) + sn * K[r * 3 + q] + (1 - cs) * kk;
            }
    } else {
        for (int i = 0; i < 9; i++)
            R[i] = (i % 4 == 0) ? 1.0 : 0.0;
    }
    for (int i = 0; i < 28; i++)
        out[i] = 0;
    for (int i = 0; i < nSrc; i++) {
        int q = corr[i];
        if (q < 0)
            continue;
        double y[3], x[3], d[3], n[3], J[6];
        for (int r = 0; r < 3; r++)
            y[r] = T[r * 4] * srcP[3 * i] + T[r * 4 + 1] * srcP[3 * i + 1] +
                   T[r * 4 + 2] * srcP[3 * i + 2] + T[r * 4 + 3];
        for (int r = 0; r < 3; r++)
            x[r] = R[r * 3] * (y[0] - c[0]) + R[r * 3 + 1] * (y[1] - c[1]) +
                   R[r * 3 + 2] * (y[2] - c[2]) + c[r] + xi[3 + r];
        for (int r = 0; r < 3; r++) {
            n[r] = dstN[3 * q + r];
            d[r] = x[r] - c[r];
        }
        double res = (x[0] - dstP[3 * q]) * n[0] + (x[1] - dstP[3 * q + 1]) * n[1] +
                     (x[2] - dstP[3 * q + 2]) * n[2];
        double w = 1.0;
        if (tukey) {
            double e = fabs(res) / 2.0;
            e        = e < 1.0 ? e : 1.0;
            w        = (1.0 - e * e) * (1.0 - e * e);
        }
        J[0] = d[1] * n[2] - d[2] * n[1];
        J[1] = d[2] * n[0] - d[0] * n[2];
        J[2] = d[0] * n[1] - d[1] * n[0];
        J[3] = n[0];
        J[4] = n[1];
        J[5] = n[2];
        out[0] += 0.5 * w * res * res;
        for (int k = 0; k < 6; k++)
            out[1 + k] += w * res * J[k];
        int m = 7;
        for (int a = 0; a < 6; a++)
            for (int b = a; b < 6; b++)
                out[m++] += w * J[a] * J[b];
    }
}

/* ====================================================================== */
/* end to end                                                             */
/* ====================================================================== */

static void bbox_of(int n, const float *pts, const int *idx, int m, float lo[3], float hi[3]) {
    /* computeBoundingBox (util.cpp:178-207) */
    const float *p0 = pts + 3 * (m ? idx[0] : 0);
    for (int a = 0; a < 3; a++)
        lo[a] = hi[a] = p0[a];
    for (int t = 0; t < m; t++) {
        const float *p = pts + 3 * idx[t];
        for (int a = 0; a < 3; a++) {
            if (p[a] > hi[a])
                hi[a] = p[a];
            else if (p[a] < lo[a])
                lo[a] = p[a];
        }
    }
    (void)n;
}

static float diameter(const float lo[3], const float hi[3]) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return sqrtf(dx * dx + dy * dy + dz * dz);
}

static int *valid_indices(int n, const float *pts, const float *nrm, int *m) {
    int *v = (int *)malloc(sizeof(int) * (n ? n : 1));
    int  c = 0;
    for (int i = 0; i < n; i++) {
        if (!is_finite3(pts + 3 * i))
            continue;
        if (nrm && !is_finite3(nrm + 3 * i))
            continue;
        v[c++] = i;
    }
    *m = c;
    return v;
}

/* removeNan (util.cpp:86-105): ascending indices of the points with a finite
 * position and, when nrm != NULL, a finite normal; returns their count */
int orc_remove_nan(int n, const float *pts, const float *nrm, int *out) {
    int  m;
    int *v = valid_indices(n, pts, nrm, &m);
    memcpy(out, v, sizeof(int) * (size_t)m);
    free(v);
    return m;
}

/* computeBoundingBox (util.cpp:178-207) over the given indices:
 * box6 = {min xyz, max xyz} */
void orc_bounding_box(int n, const float *pts, const int *idx, int m, float box6[6]) {
    bbox_of(n, pts, idx, m, box6, box6 + 3);
}

static void gather(const float *src, const int *idx, int m, float *dst, int normalize) {
    for (int t = 0; t < m; t++) {
        memcpy(dst + 3 * t, src + 3 * idx[t], sizeof(float) * 3);
        if (normalize)
            normalize3(dst + 3 * t);
    }
}

/* ====================================================================== */
/* mesh sampling                                                          */
/* ====================================================================== */

/* Philox4x32-10: ten rounds of two 32x32->64 multiplies (multipliers
 * 0xD2511F53, 0xCD9E8D57) mixing the counter, with the key bumped by the
 * Weyl constants 0x9E3779B9, 0xBB67AE85 between rounds (J. Salmon et al.,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11, and its Random123
 * known-answer tests). */
static void orc_mulhilo32(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi        = (uint32_t)(p >> 32);
    *lo        = (uint32_t)p;
}

void orc_philox4x32_10(uint32_t ctr[4], const uint32_t key[2]) {
    uint32_t k[2] = {key[0], key[1]};
    for (int round = 0; round < 10; round++) {
        uint32_t hi0, lo0, hi1, lo1;
        orc_mulhilo32(0xD2511F53u, ctr[0], &hi0, &lo0);
        orc_mulhilo32(0xCD9E8D57u, ctr[2], &hi1, &lo1);
        uint32_t x[4] = {hi1 ^ ctr[1] ^ k[0], lo1, hi0 ^ ctr[3] ^ k[1], lo0};
        memcpy(ctr, x, sizeof x);
        k[0] += 0x9E3779B9u;
        k[1] += 0xBB67AE85u;
    }
}

/* sampleMesh, util.cpp:20-84.  Per triangle (util.cpp:29-42): u = p1-p0,
 * v = p2-p0, k = u x v, area = |k| / 2 with |k| = sqrt((k0^2+k1^2)+k2^2),
 * normal = k / |k| (k itself when |k|^2 == 0, Eigen's normalized()), the
 * total area summed in triangle order.  CDF (:45-48): c0 = a0/A,
 * ci = ai/A + c(i-1), the last set to 1.  nPoints = int(A / r^2) (:50).
 * Triangle i fills outputs up to round(ci * nPoints) (:61-62; clamped to
 * nPoints where the reference would write past its buffer).  Point j
 * (:63-74): r1, r2 = top 24 bits of Philox words 0 and 1 times 2^-24;
 * a = 1 - sqrt(r1), b = sqrt(r1)(1 - r2) in double, c = sqrt(r1) r2 in
 * float; a and b rounded to float, point = (a p0 + b p1) + c p2. */
int orc_sample_mesh(int nPts, const float *pts, int nFaces, const int *faces, float radius,
                    uint64_t seed, float *outP, float *outN, int cap) {
    if (nFaces <= 0)
        return 0;
    float *cdf = (float *)malloc(sizeof(float) * (size_t)nFaces);
    float *tn  = (float *)malloc(sizeof(float) * 3 * (size_t)nFaces);
    float  A   = 0.0f;
    for (int i = 0; i < nFaces; i++) {
        const int *f = faces + 3 * (size_t)i;
        if (f[0] < 0 || f[1] < 0 || f[2] < 0 || f[0] >= nPts || f[1] >= nPts || f[2] >= nPts) {
            free(cdf);
            free(tn);
            return -1;
        }
        const float *p0 = pts + 3 * (size_t)f[0], *p1 = pts + 3 * (size_t)f[1],
                    *p2 = pts + 3 * (size_t)f[2];
        float u0 = p1[0] - p0[0], u1 = p1[1] - p0[1], u2 = p1[2] - p0[2];
        float v0 = p2[0] - p0[0], v1 = p2[1] - p0[1], v2 = p2[2] - p0[2];
        float k0 = u1 * v2 - u2 * v1;
        float k1 = u2 * v0 - u0 * v2;
        float k2 = u0 * v1 - u1 * v0;
        float sq = k0 * k0 + k1 * k1;
        sq       = sq + k2 * k2;
        float l  = sqrtf(sq);
        cdf[i]   = l / 2.0f;
        if (sq > 0.0f) {
            tn[3 * i] = k0 / l, tn[3 * i + 1] = k1 / l, tn[3 * i + 2] = k2 / l;
        } else {
            tn[3 * i] = k0, tn[3 * i + 1] = k1, tn[3 * i + 2] = k2;
        }
        A += cdf[i];
    }
    float q = A / (radius * radius);
    if (!(q >= 0.0f && q < 2147483648.0f)) {
        free(cdf);
        free(tn);
        return -2;
    }
    int nPoints = (int)q;
    if (nPoints == 0) {
        free(cdf);
        free(tn);
        return 0;
    }
    cdf[0] = cdf[0] / A;
    for (int i = 1; i < nFaces; i++)
        cdf[i] = cdf[i] / A + cdf[i - 1];
    cdf[nFaces - 1] = 1.0f;
    /* first pass: the count */
    long total = 0;
    for (int i = 0; i < nFaces; i++) {
        long e = (long)roundf(cdf[i] * (float)nPoints);
        if (e > nPoints)
            e = nPoints;
        if (e > total)
            total = e;
    }
    if (!outP || !outN || cap < total) {
        free(cdf);
        free(tn);
        return (int)total;
    }
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    long           j      = 0;
    for (int i = 0; i < nFaces; i++) {
        long e = (long)roundf(cdf[i] * (float)nPoints);
        if (e > nPoints)
            e = nPoints;
        const int   *f  = faces + 3 * (size_t)i;
        const float *p0 = pts + 3 * (size_t)f[0], *p1 = pts + 3 * (size_t)f[1],
                    *p2 = pts + 3 * (size_t)f[2];
        for (; j < e; j++) {
            uint32_t ctr[4] = {(uint32_t)j, 0u, 0u, 0u};
            orc_philox4x32_10(ctr, key);
            float  r1 = (float)(ctr[0] >> 8) * 0x1p-24f;
            float  r2 = (float)(ctr[1] >> 8) * 0x1p-24f;
            float  s  = sqrtf(r1);
            double a  = 1.0 - (double)s;
            double b  = (double)s * (1.0 - (double)r2);
            float  c  = s * r2;
            float  fa = (float)a, fb = (float)b;
            for (int d = 0; d < 3; d++) {
                float t              = fa * p0[d] + fb * p1[d];
                outP[3 * (size_t)j + d] = t + c * p2[d];
                outN[3 * (size_t)j + d] = tn[3 * i + d];
            }
        }
    }
    free(cdf);
    free(tn);
    return (int)total;
}

/* Detector::trainModel (ppf.cpp:36-147) with voxel sampling. */
int orc_train(int n, const float *pts, const float *nrm, const float *box6, const float vp[3],
              float rel, const orc_train_param *p, int refTail, orc_model *out) {
    memset(out, 0, sizeof *out);
    if (rel > 1 || rel < 0 || n <= 0)
        return -1;
    float *P = (float *)malloc(sizeof(float) * 3 * (size_t)n);
    float *N = (float *)malloc(sizeof(float) * 3 * (size_t)n);
    memcpy(P, pts, sizeof(float) * 3 * (size_t)n);
    int hasNormal = nrm != NULL;
    int m;
    int *valid = valid_indices(n, P, hasNormal ? nrm : NULL, &m);
    float lo[3], hi[3];
    if (box6 && diameter(box6, box6 + 3) != 0.0f) {
        memcpy(lo, box6, sizeof lo);
        memcpy(hi, box6 + 3, sizeof hi);
    } else {
        bbox_of(n, P, valid, m, lo, hi);
    }
    float D            = diameter(lo, hi);
    float sampleStep   = D * rel;
    float reSampleStep = D * p->poseRefRelSamplingDistance;
    float distanceStep = D * p->featDistanceStepRel;
    float angleStep    = ORC_2PI_F / (float)p->featAngleResolution;
    int  *s1           = (int *)malloc(sizeof(int) * (m ? m : 1));
    int  *s2           = (int *)malloc(sizeof(int) * (m ? m : 1));
    int   n1           = orc_voxel_sample(n, P, valid, m, sampleStep, lo, s1);
    int   n2           = orc_voxel_sample(n, P, valid, m, reSampleStep, lo, s2);
    if (hasNormal) {
        memcpy(N, nrm, sizeof(float) * 3 * (size_t)n);
        for (int i = 0; i < n; i++)
            normalize3(N + 3 * i);
    } else {
        for (size_t i = 0; i < 3 * (size_t)n; i++)
            N[i] = NAN;
        orc_estimate_normals(n, P, N, valid, m, s1, n1, p->knnNormal, p->smoothNormal, 0, vp);
        orc_estimate_normals(n, P, N, valid, m, s2, n2, p->knnNormal, p->smoothNormal, 0, vp);
    }
    out->samplingDistanceRel = rel;
    out->param               = *p;
    out->nSampled            = n1;
    out->nReSampled          = n2;
    out->sampledP            = (float *)malloc(sizeof(float) * 3 * (size_t)(n1 ? n1 : 1));
    out->sampledN            = (float *)malloc(sizeof(float) * 3 * (size_t)(n1 ? n1 : 1));
    out->reSampledP          = (float *)malloc(sizeof(float) * 3 * (size_t)(n2 ? n2 : 1));
    out->reSampledN          = (float *)malloc(sizeof(float) * 3 * (size_t)(n2 ? n2 : 1));
    gather(P, s1, n1, out->sampledP, 0);
    gather(N, s1, n1, out->sampledN, 1);
    gather(P, s2, n2, out->reSampledP, 0);
    gather(N, s2, n2, out->reSampledN, 1);
    memcpy(out->boxMin, lo, sizeof lo);
    memcpy(out->boxMax, hi, sizeof hi);
    if (vp)
        memcpy(out->viewPoint, vp, sizeof(float) * 3);
    else
        out->viewPoint[0] = out->viewPoint[1] = out->viewPoint[2] = NAN;
    orc_build_table(n1, out->sampledP, out->sampledN, angleStep, distanceStep, refTail,
                    &out->table);
    free(P);
    free(N);
    free(valid);
    free(s1);
    free(s2);
    return 0;
}

void orc_free_model(orc_model *m) {
    free(m->sampledP);
    free(m->sampledN);
    free(m->reSampledP);
    free(m->reSampledN);
    orc_free_table(&m->table);
    memset(m, 0, sizeof *m);
}

typedef struct {
    float score;
    float pose[12];
    int   seq;
} result_t;
static int cmp_result(const void *a, const void *b) {
    const result_t *x = (const result_t *)a, *y = (const result_t *)b;
    if (x->score != y->score)
        return x->score > y->score ? -1 : 1;
    return x->seq < y->seq ? -1 : (x->seq > y->seq);
}

/* Detector::matchScene (ppf.cpp:149-510). */
int orc_match(const orc_model *m, int n, const float *pts, const float *nrm, const float *box6,
              const float vp[3], float rel, float kpFrac, float minScore,
              const orc_match_param *p, int cap, float *poses12, float *scores,
              orc_match_stats *st, int nThreads) {
    orc_match_stats stl;
    if (!st)
        st = &stl;
    memset(st, 0, sizeof *st);
    if (rel > 1 || rel < 0 || kpFrac > 1 || kpFrac < 0 || minScore > 1 || minScore < 0 || n <= 0)
        return -1;
#ifdef _OPENMP
    int prevThreads = omp_get_max_threads();
    omp_set_num_threads(nThreads > 0 ? nThreads : 1);
#endif
    double t0 = now_ms();
    float *P  = (float *)malloc(sizeof(float) * 3 * (size_t)n);
    float *N  = (float *)malloc(sizeof(float) * 3 * (size_t)n);
    memcpy(P, pts, sizeof(float) * 3 * (size_t)n);
    int hasNormal = nrm != NULL;
    int nv;
    int *valid = valid_indices(n, P, hasNormal ? nrm : NULL, &nv);
    float lo[3], hi[3];
    if (box6 && diameter(box6, box6 + 3) != 0.0f) {
        memcpy(lo, box6, sizeof lo);
        memcpy(hi, box6 + 3, sizeof hi);
    } else {
        bbox_of(n, P, valid, nv, lo, hi);
    }
    float D            = diameter(m->boxMin, m->boxMax);
    float angleStep    = ORC_2PI_F / (float)m->param.featAngleResolution;
    float distanceStep = D * m->param.featDistanceStepRel;
    int   angleNum     = m->param.featAngleResolution;
    int   refNum       = m->nSampled;
    float sqD          = D * D;
    float reSampleStep = D * m->param.poseRefRelSamplingDistance;
    float sampleStep   = D * rel;

    int *samp = (int *)malloc(sizeof(int) * (nv ? nv : 1));
    int  ns   = orc_voxel_sample(n, P, valid, nv, sampleStep, lo, samp);
    if (hasNormal) {
        memcpy(N, nrm, sizeof(float) * 3 * (size_t)n);
        for (int i = 0; i < n; i++)
            normalize3(N + 3 * i);
    } else {
        for (size_t i = 0; i < 3 * (size_t)n; i++)
            N[i] = NAN;
        orc_estimate_normals(n, P, N, valid, nv, samp, ns, p->knnNormal, p->smoothNormal,
                             p->invertNormal, vp);
    }
    double t1 = now_ms();
    float keySampleStep = sqrtf(1.0f / kpFrac) * sampleStep;
    int  *kp            = (int *)malloc(sizeof(int) * (ns ? ns : 1));
    int   nk            = orc_voxel_sample(n, P, samp, ns, keySampleStep, lo, kp);
    st->nSampled        = ns;
    st->nKeypoints      = nk;

    float voteThreshold = (float)refNum * minScore;
    float maxOverlapDist = 0;
    if (p->maxOverlapDistRel > 0) maxOverlapDist = D * p->maxOverlapDistRel;
    if (p->maxOverlapDistAbs > 0) maxOverlapDist = p->maxOverlapDistAbs;
    float poseRefDist = 0;
    if (p->poseRefDistThresholdRel > 0) poseRefDist = D * p->poseRefDistThresholdRel;
    if (p->poseRefDistThresholdAbs > 0) poseRefDist = p->poseRefDistThresholdAbs;
    float scoringDist = 0;
    if (p->poseRefScoringDistRel > 0) scoringDist = D * p->poseRefScoringDistRel;
    if (p->poseRefScoringDistAbs > 0) scoringDist = p->poseRefScoringDistAbs;

    /* vote grid over the sampled scene (moved positions) */
    orc_grid gs;
    orc_grid_build(&gs, n, P, samp, ns, D * 0.5f);
    float *modelRt = (float *)malloc(sizeof(float) * 12 * (size_t)(refNum ? refNum : 1));
    for (int r = 0; r < refNum; r++)
        orc_frame(m->sampledP + 3 * r, m->sampledN + 3 * r, modelRt + 12 * r);
    orc_pose_t *poseList = (orc_pose_t *)malloc(sizeof(orc_pose_t) * 3 * (size_t)(nk ? nk : 1));
    int        *posePer  = (int *)calloc(nk ? nk : 1, sizeof(int));
    long        pairs = 0, hits = 0, votes = 0;
    int         voted = 0;
    (void)nThreads;
#pragma omp parallel num_threads(nThreads > 0 ? nThreads : 1) reduction(+ : pairs, hits, votes, voted)
    {
        int32_t *acc  = (int32_t *)malloc(sizeof(int32_t) * (size_t)refNum * (angleNum + 1));
        int     *nb   = (int *)malloc(sizeof(int) * (ns ? ns : 1));
        float   *nbP  = (float *)malloc(sizeof(float) * 3 * (size_t)(ns ? ns : 1));
        float   *nbN  = (float *)malloc(sizeof(float) * 3 * (size_t)(ns ? ns : 1));
#pragma omp for schedule(dynamic, 4)
        for (int c = 0; c < nk; c++) {
            int          id = kp[c];
            const float *p1 = P + 3 * id, *n1 = N + 3 * id;
            if (!is_finite3(n1))
                continue;
            int cnt = orc_grid_radius(&gs, p1, sqD, nb, ns);
            if ((float)cnt < voteThreshold)
                continue;
            int r = 0;
            for (int t = 0; t < cnt; t++) {
                if (nb[t] == id)
                    continue;
                memcpy(nbP + 3 * r, P + 3 * nb[t], sizeof(float) * 3);
                memcpy(nbN + 3 * r, N + 3 * nb[t], sizeof(float) * 3);
                r++;
            }
            voted++;
            pairs += r;
            orc_vote_keypoint(&m->table, refNum, angleNum, p1, n1, r, nbP, nbN, angleStep,
                              distanceStep, acc);
            for (int ref = 0; ref < refNum; ref++)
                votes += acc[(size_t)ref * (angleNum + 1)];
            orc_candidate cand[3];
            int           nc = orc_peaks(acc, refNum, angleNum, voteThreshold, cand);
            float         Ts[12];
            orc_frame(p1, n1, Ts);
            for (int x = 0; x < nc; x++) {
                orc_pose_t *pp = poseList + 3 * (size_t)c + x;
                orc_pose(Ts, modelRt + 12 * cand[x].refId, cand[x].angleId, angleNum, pp->pose);
                pp->votes = cand[x].vote;
                pose_fill(pp);
            }
            posePer[c] = nc;
        }
        free(acc);
        free(nb);
        free(nbP);
        free(nbN);
    }
    (void)hits;
    st->nPairs          = pairs;
    st->nVotes          = votes;
    st->nVotedKeypoints = voted;
    int np              = 0;
    for (int c = 0; c < nk; c++)
        for (int x = 0; x < posePer[c]; x++)
            poseList[np++] = poseList[3 * (size_t)c + x];
    st->nPoses = np;
    double t2  = now_ms();

    int        *clOf     = (int *)malloc(sizeof(int) * (np ? np : 1));
    orc_pose_t *clusters = (orc_pose_t *)malloc(sizeof(orc_pose_t) * (np ? np : 1));
    int         ncl      = orc_cluster_poses(poseList, np, 0.1f * D, angleStep, clOf, clusters);
    float       center[3];
    for (int a = 0; a < 3; a++)
        center[a] = (m->boxMin[a] + m->boxMax[a]) / 2.0f;
    orc_pose_t *nms = (orc_pose_t *)malloc(sizeof(orc_pose_t) * (ncl ? ncl : 1));
    int         nn  = orc_nms_poses(clusters, ncl, center, maxOverlapDist, nms);
    st->nClusters   = ncl;
    st->nNms        = nn;
    double t3       = now_ms();

    /* dense prep (ppf.cpp:427-434) */
    int *dense = NULL, nd = 0;
    if (p->densePoseRefinement) {
        dense = (int *)malloc(sizeof(int) * (nv ? nv : 1));
        nd    = orc_voxel_sample(n, P, valid, nv, reSampleStep, lo, dense);
        if (!hasNormal)
            orc_estimate_normals(n, P, N, valid, nv, dense, nd, p->knnNormal, 0, p->invertNormal,
                                 vp);
    }
    /* grids over the final positions */
    orc_grid gSparse, gDense, gFull;
    orc_grid_build(&gSparse, n, P, samp, ns, sampleStep);
    orc_grid_build(&gFull, n, P, valid, nv, scoringDist > 0 ? scoringDist * 4.0f : sampleStep);
    if (p->densePoseRefinement)
        orc_grid_build(&gDense, n, P, dense, nd, reSampleStep * 2.0f);
    orc_icp_criteria sparseC = {5, poseRefDist, sampleStep * 0.5f, sampleStep, 0.05f};
    orc_icp_criteria denseC  = {p->poseRefNumSteps, poseRefDist, reSampleStep * 0.5f, sampleStep,
                               0.05f};
    result_t *res = (result_t *)malloc(sizeof(result_t) * (nn ? nn : 1));
    int       nr  = 0;
    for (int c = 0; c < nn; c++) {
        float pose[12];
        memcpy(pose, nms[c].pose, sizeof pose);
        float score = nms[c].votes;
        if (score < voteThreshold)
            continue;
        if (p->sparsePoseRefinement) {
            orc_icp_result ir;
            orc_icp_regist(refNum, m->sampledP, m->sampledN, &gSparse, N, pose, &sparseC, 0, &ir);
            if (!ir.converged)
                continue;
            memcpy(pose, ir.pose, sizeof pose);
            int inl = orc_inliers(refNum, m->sampledP, pose, &gFull, scoringDist);
            score   = (float)inl / (float)refNum;
            if (score > 1.0f)
                score = 1.0f;
        }
        if (p->sparsePoseRefinement && score < minScore)
            continue;
        if (p->sparsePoseRefinement && p->densePoseRefinement) {
            orc_icp_result ir;
            orc_icp_regist(m->nReSampled, m->reSampledP, m->reSampledN, &gDense, N, pose, &denseC,
                           0, &ir);
            if (!ir.converged)
                continue;
            memcpy(pose, ir.pose, sizeof pose);
            int inl = orc_inliers(m->nReSampled, m->reSampledP, pose, &gFull, scoringDist);
            score   = (float)inl / (float)m->nReSampled;
            if (score > 1.0f)
                score = 1.0f;
        }
        if ((p->sparsePoseRefinement || p->densePoseRefinement) && score < minScore)
            continue;
        res[nr].score = score;
        memcpy(res[nr].pose, pose, sizeof pose);
        res[nr].seq = nr;
        nr++;
        if (nr >= p->numMatches)
            break;
    }
    qsort(res, nr, sizeof(result_t), cmp_result);
    int nout = nr < cap ? nr : cap;
    for (int i = 0; i < nout; i++) {
        if (poses12)
            memcpy(poses12 + 12 * i, res[i].pose, sizeof(float) * 12);
        if (scores)
            scores[i] = res[i].score;
    }
    double t4 = now_ms();
    st->timeMs[0] = (float)(t1 - t0);
    st->timeMs[1] = (float)(t2 - t1);
    st->timeMs[2] = (float)(t3 - t2);
    st->timeMs[3] = (float)(t4 - t3);
    orc_grid_free(&gs);
    orc_grid_free(&gSparse);
    orc_grid_free(&gFull);
    if (p->densePoseRefinement)
        orc_grid_free(&gDense);
    free(P);
    free(N);
    free(valid);
    free(samp);
    free(kp);
    free(modelRt);
    free(poseList);
    free(posePer);
    free(clOf);
    free(clusters);
    free(nms);
    free(dense);
    free(res);
#ifdef _OPENMP
    omp_set_num_threads(prevThreads);
#endif
    return nr;
}
