// camera_host.h -- projection, view and inverse matrices (host side).
//
// Restates Camera::UpdateProjectionMatrix / UpdateViewMatrix
// (Engine/src/Camera.cpp:134-159): glm::perspective (right-handed, clip z in
// [-1,1]), glm::lookAt (right-handed, up = +y) and glm::inverse (cofactor
// expansion with 2x2 sub-determinants).  Column-major, m[col*4+row].
// Compiled without contraction so the matrices round as GLM's code does.
#pragma once

#include <math.h>
#include <string.h>
#include "path_math.h"

namespace atx {

inline void mat_identity(float m[16])
{
    memset(m, 0, 16 * sizeof(float));
    m[0] = m[5] = m[10] = m[15] = 1.0f;
}

inline void perspective(float fov_deg, float aspect, float znear, float zfar, float out[16])
{
    const float fovy = fov_deg * 0x1.1df46ap-6f;  // glm::radians: deg * float(pi/180)
    const float tan_half = tanf(fovy / 2.0f);
    memset(out, 0, 16 * sizeof(float));
    out[0 * 4 + 0] = 1.0f / (aspect * tan_half);
    out[1 * 4 + 1] = 1.0f / tan_half;
    out[2 * 4 + 2] = -(zfar + znear) / (zfar - znear);
    out[2 * 4 + 3] = -1.0f;
    out[3 * 4 + 2] = -(2.0f * zfar * znear) / (zfar - znear);
}

inline void look_at(f3 eye, f3 center, f3 up, float out[16])
{
    const f3 f = normalize(sub(center, eye));
    const f3 s = normalize(cross(f, up));
    const f3 u = cross(s, f);
    mat_identity(out);
    out[0 * 4 + 0] = s.x; out[1 * 4 + 0] = s.y; out[2 * 4 + 0] = s.z;
    out[0 * 4 + 1] = u.x; out[1 * 4 + 1] = u.y; out[2 * 4 + 1] = u.z;
    out[0 * 4 + 2] = -f.x; out[1 * 4 + 2] = -f.y; out[2 * 4 + 2] = -f.z;
    out[3 * 4 + 0] = -dot(s, eye);
    out[3 * 4 + 1] = -dot(u, eye);
    out[3 * 4 + 2] = dot(f, eye);
}

inline void inverse4(const float* m, float out[16])
{
    auto M = [m](int c, int r) { return m[c * 4 + r]; };
    const float c00 = M(2, 2) * M(3, 3) - M(3, 2) * M(2, 3);
    const float c02 = M(1, 2) * M(3, 3) - M(3, 2) * M(1, 3);
    const float c03 = M(1, 2) * M(2, 3) - M(2, 2) * M(1, 3);
    const float c04 = M(2, 1) * M(3, 3) - M(3, 1) * M(2, 3);
    const float c06 = M(1, 1) * M(3, 3) - M(3, 1) * M(1, 3);
    const float c07 = M(1, 1) * M(2, 3) - M(2, 1) * M(1, 3);
    const float c08 = M(2, 1) * M(3, 2) - M(3, 1) * M(2, 2);
    const float c10 = M(1, 1) * M(3, 2) - M(3, 1) * M(1, 2);
    const float c11 = M(1, 1) * M(2, 2) - M(2, 1) * M(1, 2);
    const float c12 = M(2, 0) * M(3, 3) - M(3, 0) * M(2, 3);
    const float c14 = M(1, 0) * M(3, 3) - M(3, 0) * M(1, 3);
    const float c15 = M(1, 0) * M(2, 3) - M(2, 0) * M(1, 3);
    const float c16 = M(2, 0) * M(3, 2) - M(3, 0) * M(2, 2);
    const float c18 = M(1, 0) * M(3, 2) - M(3, 0) * M(1, 2);
    const float c19 = M(1, 0) * M(2, 2) - M(2, 0) * M(1, 2);
    const float c20 = M(2, 0) * M(3, 1) - M(3, 0) * M(2, 1);
    const float c22 = M(1, 0) * M(3, 1) - M(3, 0) * M(1, 1);
    const float c23 = M(1, 0) * M(2, 1) - M(2, 0) * M(1, 1);
    const float fac[6][4] = {{c00, c00, c02, c03}, {c04, c04, c06, c07}, {c08, c08, c10, c11},
                             {c12, c12, c14, c15}, {c16, c16, c18, c19}, {c20, c20, c22, c23}};
    const float vec[4][4] = {{M(1, 0), M(0, 0), M(0, 0), M(0, 0)},
                             {M(1, 1), M(0, 1), M(0, 1), M(0, 1)},
                             {M(1, 2), M(0, 2), M(0, 2), M(0, 2)},
                             {M(1, 3), M(0, 3), M(0, 3), M(0, 3)}};
    float inv[4][4];
    for (int i = 0; i < 4; ++i) {
        const float sa = (i & 1) ? -1.0f : 1.0f;  // SignA = (+,-,+,-)
        const float sb = -sa;                      // SignB = (-,+,-,+)
        inv[0][i] = ((vec[1][i] * fac[0][i] - vec[2][i] * fac[1][i]) + vec[3][i] * fac[2][i]) * sa;
        inv[1][i] = ((vec[0][i] * fac[0][i] - vec[2][i] * fac[3][i]) + vec[3][i] * fac[4][i]) * sb;
        inv[2][i] = ((vec[0][i] * fac[1][i] - vec[1][i] * fac[3][i]) + vec[3][i] * fac[5][i]) * sa;
        inv[3][i] = ((vec[0][i] * fac[2][i] - vec[1][i] * fac[4][i]) + vec[2][i] * fac[5][i]) * sb;
    }
    const float d0 = M(0, 0) * inv[0][0], d1 = M(0, 1) * inv[1][0];
    const float d2 = M(0, 2) * inv[2][0], d3 = M(0, 3) * inv[3][0];
    const float det = (d0 + d1) + (d2 + d3);
    const float one_over_det = 1.0f / det;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r)
            out[c * 4 + r] = inv[c][r] * one_over_det;
}

// Inverse projection and inverse view of a camera (Camera.cpp:134-159).
inline void camera_inverse_matrices(const float pos[3], const float dir[3], float fov_deg, float znear,
                                    float zfar, uint32_t width, uint32_t height, float inv_proj[16],
                                    float inv_view[16])
{
    float proj[16], view[16];
    const float aspect = (float)width / (float)height;
    perspective(fov_deg, aspect, znear, zfar, proj);
    inverse4(proj, inv_proj);
    const f3 eye = mk(pos[0], pos[1], pos[2]);
    const f3 center = add(eye, mk(dir[0], dir[1], dir[2]));
    look_at(eye, center, mk(0.0f, 1.0f, 0.0f), view);
    inverse4(view, inv_view);
}

}  // namespace atx
