// Host/device vector math for the renderer core.
//
// Operation order matters: the closest-hit path must reproduce the
// reference's FP32 results bit for bit (lib/src/Math/VectorOperationsDetail.h,
// VectorFunctionsDetail.h), so every helper spells out its evaluation order,
// vector/scalar is a true per-component division, and the library is built
// with -ffp-contract=off (no fused multiply-add contraction).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define CUPBR_HD __host__ __device__ __forceinline__

namespace cupbr {

constexpr float kEpsilon = 1e-5f;  // EPSILON, lib/include/Math/Functions.h:4
constexpr double kPi = 3.14159265358979323846;

struct v3 {
    float x, y, z;
};
CUPBR_HD v3 mk(float x, float y, float z) { return v3{x, y, z}; }
CUPBR_HD v3 splat(float s) { return v3{s, s, s}; }
CUPBR_HD v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
CUPBR_HD v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
CUPBR_HD v3 operator*(float s, v3 v) { return v3{s * v.x, s * v.y, s * v.z}; }
CUPBR_HD v3 operator*(v3 v, float s) { return v3{s * v.x, s * v.y, s * v.z}; }
CUPBR_HD v3 operator*(v3 a, v3 b) { return v3{a.x * b.x, a.y * b.y, a.z * b.z}; }
CUPBR_HD v3 operator/(v3 v, float s) { return v3{v.x / s, v.y / s, v.z / s}; }
CUPBR_HD v3 operator/(v3 a, v3 b) { return v3{a.x / b.x, a.y / b.y, a.z / b.z}; }
CUPBR_HD void operator+=(v3& a, v3 b) { a.x += b.x; a.y += b.y; a.z += b.z; }
CUPBR_HD float comp(v3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
CUPBR_HD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
CUPBR_HD float norm(v3 a) { return sqrtf(dot(a, a)); }
CUPBR_HD v3 normalize(v3 a) { return a / norm(a); }
CUPBR_HD v3 cross(v3 v, v3 w) { return v3{v.y * w.z - v.z * w.y, v.z * w.x - v.x * w.z, v.x * w.y - v.y * w.x}; }
CUPBR_HD v3 vexp(v3 v) { return v3{expf(v.x), expf(v.y), expf(v.z)}; }
CUPBR_HD float bits_to_float(uint32_t u) { return __builtin_bit_cast(float, u); }
CUPBR_HD uint32_t float_to_bits(float f) { return __builtin_bit_cast(uint32_t, f); }
CUPBR_HD v3 ld3(const float* p) { return v3{p[0], p[1], p[2]}; }
CUPBR_HD void st3(float* p, v3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

// lib/src/Math/FunctionsDetail.h:10-117
CUPBR_HD bool safe_float_equal(float a, float b, float eps = kEpsilon) { return fabsf(a - b) < eps; }
CUPBR_HD float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
CUPBR_HD float delta(float x) { return safe_float_equal(x, 0.0f) ? 1.0f : 0.0f; }
CUPBR_HD v3 reflect(v3 i, v3 n) { return normalize((2.0f * dot(i, n)) * n - i); }
CUPBR_HD v3 refract(float eta, v3 i, v3 n) {
    float ndoti = dot(i, n);
    float k = 1.0f - eta * eta * (1.0f - ndoti * ndoti);
    if (k < 0) return splat(0.0f);
    return (-eta) * i - ((-eta) * ndoti + sqrtf(k)) * n;
}
CUPBR_HD float fresnel_schlick(float f0, float vdoth) {
    float p = fmaxf(0.0f, 1.0f - vdoth);
    return f0 + (1.0f - f0) * p * p * p * p * p;
}
CUPBR_HD v3 fresnel_schlick(v3 f0, float vdoth) {
    float p = fmaxf(0.0f, 1.0f - vdoth);
    return f0 + ((splat(1.0f) - f0) * p) * p * p * p * p;
}
CUPBR_HD uint32_t tea4(uint32_t val0, uint32_t val1) {
    uint32_t v0 = val0, v1 = val1, s0 = 0;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}
CUPBR_HD float rnd(uint32_t& s) {
    s = 1664525u * s + 1013904223u;
    return (float)(s & 0x00FFFFFFu) / (float)0x01000000;
}
// PCG random numbers, the path tracer's alternative to the reference's TEA + LCG (no counterpart in
// the reference; O'Neill 2014, "PCG: a family of simple fast space-efficient statistically good
// algorithms for random number generation"): a 32-bit LCG state (multiplier 747796405, increment
// 2891336453) whose output is the RXS-M-XS permutation of the advanced state; the top 24 bits make
// the float. The state fits the same 32 bits as the reference's LCG state.
struct Pcg32 {
    uint32_t s;
};
CUPBR_HD uint32_t pcg_rxs_m_xs(uint32_t state) {
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
CUPBR_HD uint32_t pcg_hash(uint32_t x) { return pcg_rxs_m_xs(x * 747796405u + 2891336453u); }
CUPBR_HD float rnd(Pcg32& g) {
    g.s = g.s * 747796405u + 2891336453u;
    return (float)(pcg_rxs_m_xs(g.s) >> 8) / (float)0x01000000;
}
// the 32 bits a path keeps between bounces, and back
CUPBR_HD uint32_t state_bits(uint32_t s) { return s; }
CUPBR_HD uint32_t state_bits(Pcg32 g) { return g.s; }
template <class Seed> CUPBR_HD Seed seed_from_bits(uint32_t s);
template <> CUPBR_HD uint32_t seed_from_bits<uint32_t>(uint32_t s) { return s; }
template <> CUPBR_HD Pcg32 seed_from_bits<Pcg32>(uint32_t s) { return Pcg32{s}; }
CUPBR_HD v3 to_local_frame(v3 N, v3 d) {
    const float sz = (N.z >= 0.0f) ? 1.0f : -1.0f;
    const float a = 1.0f / (sz + N.z);
    const float ya = N.y * a;
    const float b = N.x * ya;
    const float c = N.x * sz;
    v3 lx = v3{c * N.x * a - 1.0f, sz * b, c};
    v3 ly = v3{b, N.y * ya - sz, N.y};
    return d.x * lx + d.y * ly + d.z * N;
}

}  // namespace cupbr
