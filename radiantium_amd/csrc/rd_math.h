// rd_math.h — float math for host and gfx950 device code of the product.
//
// Rounding contract: all code is compiled with -ffp-contract=off and
// correctly rounded f32 division / sqrt, so every expression rounds like the
// reference's .NET code (System.Numerics, which never contracts to FMA).
// Operation order inside each helper mirrors the reference's (e.g. Vector3.Dot
// sums x, y, then z; Vector3.Normalize divides each component by Length()).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#define RD_HD __host__ __device__ __forceinline__

namespace rd {

constexpr float kPi = 3.14159265f;          // System.MathF.PI
constexpr float kFloatEps = 1.40129846e-45f;  // float.Epsilon
constexpr float kFloatMax = 3.40282347e+38f;

struct f3 {
    float x, y, z;
};
RD_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
RD_HD f3 mk(float a) { return f3{a, a, a}; }
RD_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
RD_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
RD_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
RD_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
RD_HD f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
RD_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
RD_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
RD_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
RD_HD f3 operator+(f3 a, float s) { return f3{a.x + s, a.y + s, a.z + s}; }
RD_HD f3 operator-(f3 a, float s) { return f3{a.x - s, a.y - s, a.z - s}; }
RD_HD f3 operator-(float s, f3 a) { return f3{s - a.x, s - a.y, s - a.z}; }
RD_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RD_HD f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
RD_HD float length_sq(f3 a) { return dot(a, a); }
RD_HD float length(f3 a) { return sqrtf(length_sq(a)); }
RD_HD f3 normalize(f3 a) { return a / length(a); }
RD_HD float abs_dot(f3 a, f3 b) { return fabsf(dot(a, b)); }
RD_HD f3 reflect(f3 v, f3 n) { float d = dot(v, n); return v - (2.0f * d) * n; }
RD_HD bool is_black(f3 c) { return c.x == 0.0f && c.y == 0.0f && c.z == 0.0f; }
RD_HD f3 fabs3(f3 a) { return f3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
RD_HD f3 sqrt3(f3 a) { return f3{sqrtf(a.x), sqrtf(a.y), sqrtf(a.z)}; }
RD_HD float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// Color3F.IsValid: no negative, infinite or NaN channel.
RD_HD bool valid_color(f3 c) {
    if (c.x < 0 || isinf(c.x) || isnan(c.x)) return false;
    if (c.y < 0 || isinf(c.y) || isnan(c.y)) return false;
    if (c.z < 0 || isinf(c.z) || isnan(c.z)) return false;
    return true;
}

// MathF.Max / MathF.Min (IEEE maximum / minimum: NaN propagates, +0 > -0).
RD_HD float nmax(float a, float b) {
    if (a != b) { if (!isnan(a)) return b < a ? a : b; return a; }
    return signbit(b) ? a : b;
}
RD_HD float nmin(float a, float b) {
    if (a != b) { if (!isnan(a)) return a < b ? a : b; return a; }
    return signbit(a) ? a : b;
}
RD_HD float nclamp(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
RD_HD int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
RD_HD float sqr(float a) { return a * a; }
RD_HD float mulsign(float a, float b) { return a * copysignf(1.0f, b); }
RD_HD float safe_sqrt(float a) { return sqrtf(nmax(a, 0.0f)); }
RD_HD float max_elem(f3 c) { return nmax(nmax(c.x, c.y), c.z); }

// ---- counter-based RNG keyed by (seed, pixel, sample); draw n uses
// mix64(key + n * golden).  Replaces the per-thread System.Random stream of
// BlockBasedRenderer.cs:42 with a reproducible one.
RD_HD uint64_t mix64(uint64_t z) {
    z ^= z >> 30; z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27; z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z;
}
RD_HD uint64_t path_key(uint32_t seed, uint32_t pixel, uint32_t sample) {
    uint64_t k0 = ((uint64_t)seed << 32) | (uint64_t)sample;
    return mix64(k0 ^ mix64((uint64_t)pixel + 0x9E3779B97F4A7C15ull));
}
// Light-path key: a particle's (or a bidirectional sample's light subpath's)
// 64-bit index over the pixel and sample words, the sample word's top bit
// set so that no camera path (sample < 2^31) draws from the same stream.
RD_HD uint64_t particle_key(uint32_t seed, uint64_t index) {
    return path_key(seed, (uint32_t)index, 0x80000000u | (uint32_t)(index >> 32));
}
struct Rng {
    uint64_t key;
    uint32_t n;
    RD_HD uint64_t draw() { n++; return mix64(key + (uint64_t)n * 0x9E3779B97F4A7C15ull); }
    RD_HD float next_float() { return (float)(draw() >> 40) * 5.9604644775390625e-08f; }  // 2^-24
    RD_HD int next_int(int m) { return (int)(((draw() >> 32) * (uint64_t)m) >> 32); }
};

}  // namespace rd
