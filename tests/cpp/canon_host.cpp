// Host build of the device math header (surface_match_amd/csrc/canon.h) so
// its FP32 sequences can be compared with the oracle on the CPU.
#include "../../surface_match_amd/csrc/canon.h"

extern "C" {
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
