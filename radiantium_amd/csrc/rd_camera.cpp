// rd_camera.cpp — host camera setup; matrices are row-major M11..M44 with
// the row-vector convention of System.Numerics (v' = v * M).
#include "rd_camera.h"

#include <cmath>

namespace rd {

// Matrix4x4 multiply: each row of the product is a[i][0]*b.row0 + a[i][1]*b.row1
// + a[i][2]*b.row2 + a[i][3]*b.row3, summed as (r0 + r1) + (r2 + r3).
void mul4(const float* a, const float* b, float* out) {
    float r[16];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            r[4 * i + j] = (a[4 * i + 0] * b[0 + j] + a[4 * i + 1] * b[4 + j]) +
                           (a[4 * i + 2] * b[8 + j] + a[4 * i + 3] * b[12 + j]);
    for (int i = 0; i < 16; i++) out[i] = r[i];
}

bool invert4(const float* M, float* out) {
    const float a = M[0], b = M[1], c = M[2], d = M[3];
    const float e = M[4], f = M[5], g = M[6], h = M[7];
    const float i = M[8], j = M[9], k = M[10], l = M[11];
    const float m = M[12], n = M[13], o = M[14], p = M[15];
    // 2x2 minors of the lower two rows
    const float kp_lo = k * p - l * o, jp_ln = j * p - l * n, jo_kn = j * o - k * n;
    const float ip_lm = i * p - l * m, io_km = i * o - k * m, in_jm = i * n - j * m;
    const float c11 = (f * kp_lo - g * jp_ln + h * jo_kn);
    const float c12 = -(e * kp_lo - g * ip_lm + h * io_km);
    const float c13 = (e * jp_ln - f * ip_lm + h * in_jm);
    const float c14 = -(e * jo_kn - f * io_km + g * in_jm);
    const float det = a * c11 + b * c12 + c * c13 + d * c14;
    if (std::fabs(det) < kFloatEps) return false;
    const float s = 1.0f / det;
    float r[16];
    r[0] = c11 * s; r[4] = c12 * s; r[8] = c13 * s; r[12] = c14 * s;
    r[1] = -(b * kp_lo - c * jp_ln + d * jo_kn) * s;
    r[5] = (a * kp_lo - c * ip_lm + d * io_km) * s;
    r[9] = -(a * jp_ln - b * ip_lm + d * in_jm) * s;
    r[13] = (a * jo_kn - b * io_km + c * in_jm) * s;
    // minors of rows 2 and 4
    const float gp_ho = g * p - h * o, fp_hn = f * p - h * n, fo_gn = f * o - g * n;
    const float ep_hm = e * p - h * m, eo_gm = e * o - g * m, en_fm = e * n - f * m;
    r[2] = (b * gp_ho - c * fp_hn + d * fo_gn) * s;
    r[6] = -(a * gp_ho - c * ep_hm + d * eo_gm) * s;
    r[10] = (a * fp_hn - b * ep_hm + d * en_fm) * s;
    r[14] = -(a * fo_gn - b * eo_gm + c * en_fm) * s;
    // minors of rows 2 and 3
    const float gl_hk = g * l - h * k, fl_hj = f * l - h * j, fk_gj = f * k - g * j;
    const float el_hi = e * l - h * i, ek_gi = e * k - g * i, ej_fi = e * j - f * i;
    r[3] = -(b * gl_hk - c * fl_hj + d * fk_gj) * s;
    r[7] = (a * gl_hk - c * el_hi + d * ek_gi) * s;
    r[11] = -(a * fl_hj - b * el_hi + d * ej_fi) * s;
    r[15] = (a * fk_gj - b * ek_gi + c * ej_fi) * s;
    for (int q = 0; q < 16; q++) out[q] = r[q];
    return true;
}

// Vector3.Transform(v, M) for a row-major 4x4: x*M11 + y*M21 + z*M31 + M41, ...
static void xform_point4(const float* M, const float* v, float* o) {
    for (int j = 0; j < 3; j++) o[j] = v[0] * M[j] + v[1] * M[4 + j] + v[2] * M[8 + j] + M[12 + j];
}

bool setup_camera(const rd_camera_desc& c, DevCamera* out, LtCamera* lt) {
    float c2w[16];
    if (c.use_matrix) {
        for (int q = 0; q < 16; q++) c2w[q] = c.to_world[q];
    } else {
        f3 origin = mk(c.origin[0], c.origin[1], c.origin[2]);
        f3 target = mk(c.target[0], c.target[1], c.target[2]);
        f3 up = mk(c.up[0], c.up[1], c.up[2]);
        f3 dir = normalize(target - origin);
        f3 left = normalize(cross(normalize(up), dir));
        f3 real_up = normalize(cross(dir, left));
        const float m[16] = {left.x, left.y, left.z, 0, real_up.x, real_up.y, real_up.z, 0,
                             dir.x, dir.y, dir.z, 0, origin.x, origin.y, origin.z, 1};
        for (int q = 0; q < 16; q++) c2w[q] = m[q];
    }
    const float W = (float)c.width, H = (float)c.height;
    out->inv_res_x = 1.0f / W;
    out->inv_res_y = 1.0f / H;
    out->width = c.width;
    out->height = c.height;
    out->near_z = c.near_z;
    out->far_z = c.far_z;
    const float aspect = c.width / (float)c.height;
    const float recip = 1.0f / (c.far_z - c.near_z);
    const float cot = 1.0f / std::tan((c.fov / 2.0f) * (kPi / 180.0f));
    const float persp[16] = {cot, 0, 0, 0, 0, cot, 0, 0, 0, 0, c.far_z * recip, 1, 0, 0, -c.near_z * c.far_z * recip, 0};
    const float scale[16] = {-0.5f, 0, 0, 0, 0, -0.5f * aspect, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    const float trans[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, -1, -1 / aspect, 0, 1};
    float pt[16], p[16];
    mul4(persp, trans, pt);
    mul4(pt, scale, p);
    if (!invert4(p, out->clip_to_camera)) return false;
    float w2c[16];
    if (!invert4(c2w, w2c)) return false;
    for (int q = 0; q < 16; q++) out->camera_to_world[q] = c2w[q];
    if (lt) {
        for (int q = 0; q < 16; q++) {
            lt->camera_to_clip[q] = p[q];
            lt->world_to_camera[q] = w2c[q];
            lt->camera_to_world[q] = c2w[q];
        }
        const float zero[3] = {0, 0, 0}, one[3] = {1, 1, 0};
        float pmin[3], pmax[3];
        xform_point4(out->clip_to_camera, zero, pmin);
        xform_point4(out->clip_to_camera, one, pmax);
        const float zmin = pmin[2], zmax = pmax[2];
        for (int j = 0; j < 3; j++) { pmin[j] = pmin[j] / zmin; pmax[j] = pmax[j] / zmax; }
        lt->min_x = pmax[0]; lt->min_y = pmax[1];
        lt->max_x = pmin[0]; lt->max_y = pmin[1];
        const float area = std::fabs((lt->max_x - lt->min_x) * (lt->max_y - lt->min_y));
        lt->normalization = 1 / area;
        lt->near_z = c.near_z;
        lt->far_z = c.far_z;
        lt->width = c.width;
        lt->height = c.height;
    }
    return true;
}

}  // namespace rd
