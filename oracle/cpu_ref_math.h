// cpu_ref_math.h — TEST INFRASTRUCTURE (the oracle).  Scalar float math that
// restates the reference's System.Numerics / MathExt semantics.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may use anything
// under oracle/.  Compiled with -O2 -ffp-contract=off, no fast-math.
#pragma once
#include <cmath>
#include <cstdint>
#include <cfloat>
#include <algorithm>

namespace cref {

static const float PI_F = 3.14159265f;         // System.MathF.PI
static const float EPS_F = 1.40129846e-45f;    // float.Epsilon (smallest denormal)

// System.MathF.Max / Min (IEEE-754:2019 maximum/minimum: NaN propagates,
// +0 > -0).  Used wherever the reference calls MathF.Max/Min/Math.Max/Min.
inline float nmax(float a, float b) {
    if (a != b) { if (!std::isnan(a)) return b < a ? a : b; return a; }
    return std::signbit(b) ? a : b;
}
inline float nmin(float a, float b) {
    if (a != b) { if (!std::isnan(a)) return a < b ? a : b; return a; }
    return std::signbit(a) ? a : b;
}
// Math.Clamp(float) (value passes through when NaN).
inline float nclamp(float v, float lo, float hi) {
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}
inline int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
inline float sqr(float a) { return a * a; }                               // MathExt.Sqr
inline float mulsign(float a, float b) { return a * std::copysign(1.0f, b); } // MathExt.MulSign
inline float safe_sqrt(float a) { return std::sqrt(nmax(a, 0.0f)); }        // MathExt.SafeSqrt
inline float radian(float v) { return v * (PI_F / 180.0f); }                // MathExt.Radian

struct V3 {
    float x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    explicit V3(float a) : x(a), y(a), z(a) {}
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return V3(-a.x, -a.y, -a.z); }
inline V3 operator*(V3 a, V3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 operator*(V3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(float s, V3 a) { return V3(s * a.x, s * a.y, s * a.z); }
inline V3 operator/(V3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float length_sq(V3 a) { return dot(a, a); }
inline float length(V3 a) { return std::sqrt(length_sq(a)); }
inline V3 normalize(V3 a) { return a / length(a); }           // Vector3.Normalize
inline float abs_dot(V3 a, V3 b) { return std::fabs(dot(a, b)); }
inline V3 vmin(V3 a, V3 b) { return V3(a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z); }
inline V3 vmax(V3 a, V3 b) { return V3(a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z); }
// Vector3.Reflect: v - (2*dot(v,n))*n
inline V3 reflect(V3 v, V3 n) { float d = dot(v, n); return v - (2.0f * d) * n; }

struct V2 { float x, y; V2() : x(0), y(0) {} V2(float a, float b) : x(a), y(b) {} };
inline V2 operator+(V2 a, V2 b) { return V2(a.x + b.x, a.y + b.y); }
inline V2 operator*(float s, V2 a) { return V2(s * a.x, s * a.y); }

// Color3F (Radiantium.Core/Color3F.cs)
struct C3 {
    float r, g, b;
    C3() : r(0), g(0), b(0) {}
    C3(float a, float b_, float c) : r(a), g(b_), b(c) {}
    explicit C3(float a) : r(a), g(a), b(a) {}
    bool black() const { return r == 0.0f && g == 0.0f && b == 0.0f; }  // == Color3F.Black
    bool valid() const {                                                  // Color3F.IsValid
        if (r < 0 || std::isinf(r) || std::isnan(r)) return false;
        if (g < 0 || std::isinf(g) || std::isnan(g)) return false;
        if (b < 0 || std::isinf(b) || std::isnan(b)) return false;
        return true;
    }
    float max_elem() const { return nmax(nmax(r, g), b); }               // MathExt.MaxElement
    float lum() const { return r * 0.212671f + g * 0.715160f + b * 0.072169f; }
};
inline C3 operator+(C3 a, C3 b) { return C3(a.r + b.r, a.g + b.g, a.b + b.b); }
inline C3 operator-(C3 a, C3 b) { return C3(a.r - b.r, a.g - b.g, a.b - b.b); }
inline C3 operator*(C3 a, C3 b) { return C3(a.r * b.r, a.g * b.g, a.b * b.b); }
inline C3 operator/(C3 a, C3 b) { return C3(a.r / b.r, a.g / b.g, a.b / b.b); }
inline C3 operator*(C3 a, float s) { return C3(a.r * s, a.g * s, a.b * s); }
inline C3 operator*(float s, C3 a) { return C3(s * a.r, s * a.g, s * a.b); }
inline C3 operator/(C3 a, float s) { return C3(a.r / s, a.g / s, a.b / s); }
inline C3 operator-(float s, C3 a) { return C3(s - a.r, s - a.g, s - a.b); }
inline C3 operator+(C3 a, float s) { return C3(a.r + s, a.g + s, a.b + s); }
inline C3 operator-(C3 a, float s) { return C3(a.r - s, a.g - s, a.b - s); }
inline C3 cabs(C3 a) { return C3(std::fabs(a.r), std::fabs(a.g), std::fabs(a.b)); }
inline C3 csqrt(C3 a) { return C3(std::sqrt(a.r), std::sqrt(a.g), std::sqrt(a.b)); }
inline C3 cexp(C3 a) { return C3(std::exp(a.r), std::exp(a.g), std::exp(a.b)); }

// Matrix4x4, row-vector convention (v' = v * M), fields M11..M44 row-major.
struct M4 {
    float m[4][4];
};
inline M4 m4_identity() { M4 r{}; for (int i = 0; i < 4; i++) r.m[i][i] = 1.0f; return r; }
// operator*: row i of the product = a[i][0]*b.row0 + a[i][1]*b.row1 + a[i][2]*b.row2 + a[i][3]*b.row3,
// summed pairwise ((r0+r1)+(r2+r3)) as the SSE path of Matrix4x4.Multiply.
inline M4 m4_mul(const M4& a, const M4& b) {
    M4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            r.m[i][j] = (a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j]) +
                        (a.m[i][2] * b.m[2][j] + a.m[i][3] * b.m[3][j]);
    return r;
}
// Matrix4x4.Invert via the adjugate of 2x2 minors.  Returns false when
// |det| < float.Epsilon.
inline bool m4_invert(const M4& mat, M4* out) {
    float a = mat.m[0][0], b = mat.m[0][1], c = mat.m[0][2], d = mat.m[0][3];
    float e = mat.m[1][0], f = mat.m[1][1], g = mat.m[1][2], h = mat.m[1][3];
    float i = mat.m[2][0], j = mat.m[2][1], k = mat.m[2][2], l = mat.m[2][3];
    float m = mat.m[3][0], n = mat.m[3][1], o = mat.m[3][2], p = mat.m[3][3];
    float kp_lo = k * p - l * o, jp_ln = j * p - l * n, jo_kn = j * o - k * n;
    float ip_lm = i * p - l * m, io_km = i * o - k * m, in_jm = i * n - j * m;
    float a11 = (f * kp_lo - g * jp_ln + h * jo_kn);
    float a12 = -(e * kp_lo - g * ip_lm + h * io_km);
    float a13 = (e * jp_ln - f * ip_lm + h * in_jm);
    float a14 = -(e * jo_kn - f * io_km + g * in_jm);
    float det = a * a11 + b * a12 + c * a13 + d * a14;
    if (std::fabs(det) < EPS_F) return false;
    float inv = 1.0f / det;
    M4 r;
    r.m[0][0] = a11 * inv; r.m[1][0] = a12 * inv; r.m[2][0] = a13 * inv; r.m[3][0] = a14 * inv;
    r.m[0][1] = -(b * kp_lo - c * jp_ln + d * jo_kn) * inv;
    r.m[1][1] = (a * kp_lo - c * ip_lm + d * io_km) * inv;
    r.m[2][1] = -(a * jp_ln - b * ip_lm + d * in_jm) * inv;
    r.m[3][1] = (a * jo_kn - b * io_km + c * in_jm) * inv;
    float gp_ho = g * p - h * o, fp_hn = f * p - h * n, fo_gn = f * o - g * n;
    float ep_hm = e * p - h * m, eo_gm = e * o - g * m, en_fm = e * n - f * m;
    r.m[0][2] = (b * gp_ho - c * fp_hn + d * fo_gn) * inv;
    r.m[1][2] = -(a * gp_ho - c * ep_hm + d * eo_gm) * inv;
    r.m[2][2] = (a * fp_hn - b * ep_hm + d * en_fm) * inv;
    r.m[3][2] = -(a * fo_gn - b * eo_gm + c * en_fm) * inv;
    float gl_hk = g * l - h * k, fl_hj = f * l - h * j, fk_gj = f * k - g * j;
    float el_hi = e * l - h * i, ek_gi = e * k - g * i, ej_fi = e * j - f * i;
    r.m[0][3] = -(b * gl_hk - c * fl_hj + d * fk_gj) * inv;
    r.m[1][3] = (a * gl_hk - c * el_hi + d * ek_gi) * inv;
    r.m[2][3] = -(a * fl_hj - b * el_hi + d * ej_fi) * inv;
    r.m[3][3] = (a * fk_gj - b * ek_gi + c * ej_fi) * inv;
    *out = r;
    return true;
}
// Vector3.Transform(position, M): x*M11 + y*M21 + z*M31 + M41 (no w divide)
inline V3 xform_point(V3 v, const M4& M) {
    return V3(v.x * M.m[0][0] + v.y * M.m[1][0] + v.z * M.m[2][0] + M.m[3][0],
              v.x * M.m[0][1] + v.y * M.m[1][1] + v.z * M.m[2][1] + M.m[3][1],
              v.x * M.m[0][2] + v.y * M.m[1][2] + v.z * M.m[2][2] + M.m[3][2]);
}
// Vector3.TransformNormal(normal, M): x*M11 + y*M21 + z*M31
inline V3 xform_vector(V3 v, const M4& M) {
    return V3(v.x * M.m[0][0] + v.y * M.m[1][0] + v.z * M.m[2][0],
              v.x * M.m[0][1] + v.y * M.m[1][1] + v.z * M.m[2][1],
              v.x * M.m[0][2] + v.y * M.m[1][2] + v.z * M.m[2][2]);
}

// ---- counter-based RNG (replaces the unseeded per-thread System.Random of
// BlockBasedRenderer.cs:42; the stream is keyed by (seed, pixel, sample) so a
// GPU implementation of the same key schedule draws identical numbers).
inline uint64_t mix64(uint64_t z) {
    z ^= z >> 30; z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27; z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z;
}
inline uint64_t path_key(uint32_t seed, uint32_t pixel, uint32_t sample) {
    uint64_t k0 = ((uint64_t)seed << 32) | (uint64_t)sample;
    return mix64(k0 ^ mix64((uint64_t)pixel + 0x9E3779B97F4A7C15ull));
}
// Light-path key: a particle's (or a bidirectional sample's light subpath's)
// 64-bit index over the pixel and sample words, the sample word's top bit
// set so that no camera path (sample < 2^31) draws from the same stream.
inline uint64_t particle_key(uint32_t seed, uint64_t index) {
    return path_key(seed, (uint32_t)index, 0x80000000u | (uint32_t)(index >> 32));
}
struct Rng {
    uint64_t key;
    uint32_t n;
    uint64_t draw() { n++; return mix64(key + (uint64_t)n * 0x9E3779B97F4A7C15ull); }
    float next_float() { return (float)(draw() >> 40) * 0x1p-24f; }          // (float)NextDouble()
    int next_int(int m) { return (int)(((draw() >> 32) * (uint64_t)m) >> 32); } // Next(m)
};

}  // namespace cref
