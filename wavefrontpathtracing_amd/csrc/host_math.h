// Host-side float helpers with the reference's exact FP semantics: glm's
// min/max (extern/glm/detail/func_common.inl:17-35), mat4*vec4 association
// (extern/glm/detail/type_mat4x4.inl:537-573) and 4x4 inverse
// (extern/glm/detail/func_matrix.inl:388-446).  Built with -ffp-contract=off.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace wf {

inline float gmin(float x, float y) { return (y < x) ? y : x; }
inline float gmax(float x, float y) { return (x < y) ? y : x; }

// as_util::get_aabb_volume (as_util.h:7-11): half the surface area
inline float half_area(float ex, float ey, float ez) { return ex * ey + ey * ez + ez * ex; }

// column-major m[c*4 + r]
inline void mat4_mul_vec4(const float* m, const float* v, float* out)
{
    float r4[4];
    for (int r = 0; r < 4; ++r) {
        const float lo = m[r] * v[0] + m[4 + r] * v[1];
        const float hi = m[8 + r] * v[2] + m[12 + r] * v[3];
        r4[r] = lo + hi;
    }
    std::memcpy(out, r4, sizeof(r4));
}

inline void mat4_mul(const float* a, const float* b, float* out)
{
    float t[16];
    for (int c = 0; c < 4; ++c) mat4_mul_vec4(a, b + 4 * c, t + 4 * c);
    std::memcpy(out, t, sizeof(t));
}

inline void mat4_inverse(const float* m, float* out)
{
    auto e = [m](int c, int r) { return m[c * 4 + r]; };
    // 2x2 minors of columns 2..3 over rows (r0,r1)
    const float k00 = e(2,2) * e(3,3) - e(3,2) * e(2,3);
    const float k02 = e(1,2) * e(3,3) - e(3,2) * e(1,3);
    const float k03 = e(1,2) * e(2,3) - e(2,2) * e(1,3);
    const float k04 = e(2,1) * e(3,3) - e(3,1) * e(2,3);
    const float k06 = e(1,1) * e(3,3) - e(3,1) * e(1,3);
    const float k07 = e(1,1) * e(2,3) - e(2,1) * e(1,3);
    const float k08 = e(2,1) * e(3,2) - e(3,1) * e(2,2);
    const float k10 = e(1,1) * e(3,2) - e(3,1) * e(1,2);
    const float k11 = e(1,1) * e(2,2) - e(2,1) * e(1,2);
    const float k12 = e(2,0) * e(3,3) - e(3,0) * e(2,3);
    const float k14 = e(1,0) * e(3,3) - e(3,0) * e(1,3);
    const float k15 = e(1,0) * e(2,3) - e(2,0) * e(1,3);
    const float k16 = e(2,0) * e(3,2) - e(3,0) * e(2,2);
    const float k18 = e(1,0) * e(3,2) - e(3,0) * e(1,2);
    const float k19 = e(1,0) * e(2,2) - e(2,0) * e(1,2);
    const float k20 = e(2,0) * e(3,1) - e(3,0) * e(2,1);
    const float k22 = e(1,0) * e(3,1) - e(3,0) * e(1,1);
    const float k23 = e(1,0) * e(2,1) - e(2,0) * e(1,1);
    const float F0[4] = {k00, k00, k02, k03}, F1[4] = {k04, k04, k06, k07}, F2[4] = {k08, k08, k10, k11};
    const float F3[4] = {k12, k12, k14, k15}, F4[4] = {k16, k16, k18, k19}, F5[4] = {k20, k20, k22, k23};
    const float V0[4] = {e(1,0), e(0,0), e(0,0), e(0,0)};
    const float V1[4] = {e(1,1), e(0,1), e(0,1), e(0,1)};
    const float V2[4] = {e(1,2), e(0,2), e(0,2), e(0,2)};
    const float V3[4] = {e(1,3), e(0,3), e(0,3), e(0,3)};
    float inv[16];
    for (int k = 0; k < 4; ++k) {
        const float sA = (k & 1) ? -1.0f : 1.0f;
        const float sB = -sA;
        inv[k]      = ((V1[k] * F0[k] - V2[k] * F1[k]) + V3[k] * F2[k]) * sA;
        inv[4 + k]  = ((V0[k] * F0[k] - V2[k] * F3[k]) + V3[k] * F4[k]) * sB;
        inv[8 + k]  = ((V0[k] * F1[k] - V1[k] * F3[k]) + V3[k] * F5[k]) * sA;
        inv[12 + k] = ((V0[k] * F2[k] - V1[k] * F4[k]) + V2[k] * F5[k]) * sB;
    }
    const float d0 = e(0,0) * inv[0], d1 = e(0,1) * inv[4], d2 = e(0,2) * inv[8], d3 = e(0,3) * inv[12];
    const float det = (d0 + d1) + (d2 + d3);
    const float s = 1.0f / det;
    for (int i = 0; i < 16; ++i) out[i] = inv[i] * s;
}

// glm dot / cross / normalize (extern/glm/detail/func_geometric.inl)
inline float dot3(const float* a, const float* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
inline void cross3(const float* x, const float* y, float* o)
{
    const float c0 = x[1] * y[2] - y[1] * x[2];
    const float c1 = x[2] * y[0] - y[2] * x[0];
    const float c2 = x[0] * y[1] - y[0] * x[1];
    o[0] = c0; o[1] = c1; o[2] = c2;
}
inline void normalize3(const float* v, float* o)
{
    const float s = 1.0f / std::sqrt(dot3(v, v));
    o[0] = v[0] * s; o[1] = v[1] * s; o[2] = v[2] * s;
}

} // namespace wf
