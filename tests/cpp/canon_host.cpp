// Host build of the device math header (surface_match_amd/csrc/canon.h) so
// its FP32 sequences can be compared with the oracle on the CPU.
#include "../../surface_match_amd/csrc/canon.h"

#include <cmath>

// The reference's scalar-tail feature sequence (util.cpp:697-730) with
// glibc's atan2f in place of the canonical one: the keys and alphas the
// reference's libm path gives, to measure how often they differ from the
// canonical ones.
static float angle_a2b_libm(float ax, float ay, float az, float bx, float by, float bz) {
    float i     = ay * bz - az * by;
    float j     = az * bx - ax * bz;
    float k     = ax * by - ay * bx;
    float cross = sqrtf(i * i + j * j + k * k);
    float dot   = ax * bx + ay * by + az * bz;
    return atan2f(cross, dot);
}

extern "C" {
uint32_t ch_ppf_key_libm(const float *p1, const float *n1, const float *p2, const float *n2,
                         float as, float ds, int tail) {
    float dx = p2[0] - p1[0], dy = p2[1] - p1[1], dz = p2[2] - p1[2];
    float dn = sqrtf(dx * dx + dy * dy + dz * dz);
    float f1 = 0.0f, f2 = 0.0f, f3 = 0.0f;
    if (!(tail && !(dn > 0.0f))) {
        float ux = dx / dn, uy = dy / dn, uz = dz / dn;
        f1 = angle_a2b_libm(ux, uy, uz, n1[0], n1[1], n1[2]);
        f2 = angle_a2b_libm(ux, uy, uz, n2[0], n2[1], n2[2]);
        f3 = angle_a2b_libm(n1[0], n1[1], n1[2], n2[0], n2[1], n2[2]);
    }
    return ppfc::murmur_key4(ppfc::quant(f1, as, tail), ppfc::quant(f2, as, tail),
                             ppfc::quant(f3, as, tail), ppfc::quant(dn, ds, tail));
}
float ch_alpha_libm(const float *rt, const float *p) {
    float yy = rt[4] * p[0] + rt[5] * p[1] + rt[6] * p[2] + rt[7];
    float zz = rt[8] * p[0] + rt[9] * p[1] + rt[10] * p[2] + rt[11];
    float a  = atan2f(-zz, yy);
    if ((a == ppfc::kPi || a == -ppfc::kPi) && zz != 0.0f)
        a = -a;
    return a;
}
// counts over all pairs (i, j), i != j, of the n points: keys that differ,
// alphas (of j in i's frame) that differ, and bins of those alphas against
// alphaModel = 0 that differ, for the canonical and the libm sequences
void ch_libm_divergence(const float *p, const float *n, int np, float as, float ds, int tail,
                        long long *out4) {
    long long pairs = 0, keys = 0, alphas = 0, bins = 0;
    const float sMaxId = 29.0f / ppfc::kTwoPi;
    for (int i = 0; i < np; i++) {
        float rt[12];
        ppfc::frame(p[3 * i], p[3 * i + 1], p[3 * i + 2], n[3 * i], n[3 * i + 1], n[3 * i + 2], rt);
        for (int j = 0; j < np; j++) {
            if (j == i)
                continue;
            pairs++;
            const float *a = p + 3 * i, *b = p + 3 * j, *na = n + 3 * i, *nb = n + 3 * j;
            uint32_t k1 = ppfc::ppf_key(a[0], a[1], a[2], na[0], na[1], na[2], b[0], b[1], b[2],
                                        nb[0], nb[1], nb[2], as, ds, tail != 0);
            uint32_t k2 = ch_ppf_key_libm(a, na, b, nb, as, ds, tail);
            keys += k1 != k2;
            float x1 = ppfc::alpha_of(rt, b[0], b[1], b[2]), x2 = ch_alpha_libm(rt, b);
            alphas += x1 != x2;
            bins += ppfc::angle_bin(0.0f, x1, sMaxId, 30, false) !=
                    ppfc::angle_bin(0.0f, x2, sMaxId, 30, false);
        }
    }
    out4[0] = pairs;
    out4[1] = keys;
    out4[2] = alphas;
    out4[3] = bins;
}
float    ch_atan2f(float y, float x) { return ppfc::atan2f_c(y, x); }
uint32_t ch_ppf_key(const float *p1, const float *n1, const float *p2, const float *n2,
                    float as, float ds, int tail) {
    return ppfc::ppf_key(p1[0], p1[1], p1[2], n1[0], n1[1], n1[2], p2[0], p2[1], p2[2], n2[0],
                         n2[1], n2[2], as, ds, tail != 0);
}
void  ch_frame(const float *p, const float *n, float *rt) { ppfc::frame(p[0], p[1], p[2], n[0], n[1], n[2], rt); }
float ch_alpha(const float *rt, const float *p) { return ppfc::alpha_of(rt, p[0], p[1], p[2]); }
int   ch_bin(float am, float as, int n, int tail) {
    return ppfc::angle_bin(am, as, (float)(n - 1) / ppfc::kTwoPi, n, tail != 0);
}
int ch_bin_wrapped(float am, float as, int n, int tail) {
    const float off = tail ? ppfc::bin_offset_tail(as) : ppfc::bin_offset_body(as);
    return ppfc::bin_wrapped(am, off, (float)(n - 1) / ppfc::kTwoPi, n);
}
int ch_bin_wrapped_fast(float am, float as, int n, int tail) {
    const float off = tail ? ppfc::bin_offset_tail(as) : ppfc::bin_offset_body(as);
    return ppfc::bin_wrapped_fast(am, off, (float)(n - 1) / ppfc::kTwoPi, (float)(n - 1));
}
void ch_pose(const float *Ts, const float *Tm, float c, float s, float *out) {
    ppfc::compose_pose(Ts, Tm, c, s, out);
}
void ch_quat(const float *R, float *q) { ppfc::quat_from_rot(R, q); }
void ch_philox(uint32_t *c, uint32_t k0, uint32_t k1) { ppfc::philox4x32_10(c, k0, k1); }
float ch_expf(float x) { return ppfc::expf_c(x); }
void  ch_unit_orthogonal(const float *n, float *o) { ppfc::unit_orthogonal(n, o); }
float ch_uniform01(uint32_t u) { return ppfc::uniform01(u); }
void ch_mesh_point(const float *p0, const float *p1, const float *p2, float r1, float r2,
                   float *out) {
    ppfc::mesh_point(p0, p1, p2, r1, r2, out);
}
}
