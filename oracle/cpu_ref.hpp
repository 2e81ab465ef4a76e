// ORACLE — test infrastructure only.
//
// A CPU restatement of cupbr's renderer core, used as the checker for the
// HIP path and as the timed CPU baseline. Nothing in the product library
// (cupbr_amd/) includes, links or calls this file; only tests/, the smoke
// check and bench.py's cpu_baseline leg load it, through liboracle.so.
//
// Every function is templated on the scalar type R:
//   R = float  -> the reference's FP32 semantics, operation by operation
//                 (compiled with -ffp-contract=off, no fast-math), so
//                 integer/ID results can be compared bit-exactly;
//   R = double -> the FP64 oracle of the same estimator.
// The RNG is integer arithmetic and identical in both modes.
//
// Each function cites the reference file:line it restates. Paths are
// relative to the reference repository root.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace orc {

static const float kEps = 1e-5f;  // EPSILON, lib/include/Math/Functions.h:4
static const double kPiD = 3.14159265358979323846;  // M_PI

// ---------------------------------------------------------------- vectors
// lib/src/Math/VectorOperationsDetail.h:90-170 (component-wise ops; v/s is
// a true division per component), VectorFunctionsDetail.h:38-73.
template <typename R>
struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a) : x(a), y(a), z(a) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    R& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
template <typename R> inline V3<R> operator+(const V3<R>& a, const V3<R>& b) { return V3<R>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename R> inline V3<R> operator-(const V3<R>& a, const V3<R>& b) { return V3<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename R> inline V3<R> operator*(R s, const V3<R>& v) { return V3<R>(s * v.x, s * v.y, s * v.z); }
template <typename R> inline V3<R> operator*(const V3<R>& v, R s) { return s * v; }
template <typename R> inline V3<R> operator*(const V3<R>& a, const V3<R>& b) { return V3<R>(a.x * b.x, a.y * b.y, a.z * b.z); }
template <typename R> inline V3<R> operator/(const V3<R>& v, R s) { return V3<R>(v.x / s, v.y / s, v.z / s); }
template <typename R> inline V3<R> operator/(const V3<R>& a, const V3<R>& b) { return V3<R>(a.x / b.x, a.y / b.y, a.z / b.z); }
template <typename R> inline void operator+=(V3<R>& a, const V3<R>& b) { a.x += b.x; a.y += b.y; a.z += b.z; }
template <typename R> inline R dot(const V3<R>& a, const V3<R>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> inline R norm(const V3<R>& a) { return std::sqrt(dot(a, a)); }
template <typename R> inline V3<R> normalize(const V3<R>& a) { return a / norm(a); }
template <typename R> inline V3<R> cross(const V3<R>& v, const V3<R>& w) {
    return V3<R>(v.y * w.z - v.z * w.y, v.z * w.x - v.x * w.z, v.x * w.y - v.y * w.x);
}
template <typename R> inline V3<R> vexp(const V3<R>& v) { return V3<R>(std::exp(v.x), std::exp(v.y), std::exp(v.z)); }
template <typename R, typename S> inline V3<R> cast3(const V3<S>& v) { return V3<R>((R)v.x, (R)v.y, (R)v.z); }

// ---------------------------------------------------------------- scalar functions
// lib/src/Math/FunctionsDetail.h:10-117
template <typename R> inline bool safeFloatEqual(R a, R b, R eps = (R)kEps) { return std::abs(a - b) < eps; }
template <typename R> inline R clampf(R x, R lo, R hi) { return std::fmin(std::fmax(x, lo), hi); }
template <typename R> inline R delta(R x) { return safeFloatEqual(x, (R)0) ? (R)1 : (R)0; }
template <typename R> inline V3<R> reflect(const V3<R>& i, const V3<R>& n) { return normalize(((R)2 * dot(i, n)) * n - i); }
template <typename R> inline V3<R> refract(R eta, const V3<R>& i, const V3<R>& n) {
    R NdotI = dot(i, n);
    R k = (R)1 - eta * eta * ((R)1 - NdotI * NdotI);
    if (k < 0) return V3<R>(0);
    return (-eta) * i - ((-eta) * NdotI + std::sqrt(k)) * n;
}
template <typename R> inline R fresnel_schlick(R F0, R VdotH) {
    R p = std::fmax((R)0, (R)1 - VdotH);
    return F0 + ((R)1 - F0) * p * p * p * p * p;
}
template <typename R> inline V3<R> fresnel_schlick(const V3<R>& F0, R VdotH) {
    R p = std::fmax((R)0, (R)1 - VdotH);
    return F0 + ((V3<R>((R)1) - F0) * p) * p * p * p * p;
}
// FunctionsDetail.h:70-87
inline uint32_t tea4(uint32_t val0, uint32_t val1) {
    uint32_t v0 = val0, v1 = val1, s0 = 0;
    for (uint32_t n = 0; n < 4; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}
// The PCG generator of the product's CUPBR_RNG_PCG mode (no counterpart in the reference; O'Neill
// 2014): the 32-bit LCG s = 747796405 s + 2891336453 with the RXS-M-XS output permutation. While
// g_rng_pcg is set (orc_set_rng), rnd() draws from it and path-tracer paths start at
// pcg_hash(pcg_hash(frame) ^ pixel); the tests set it only around path-tracer renders.
extern int g_rng_pcg;
inline uint32_t pcg_rxs_m_xs(uint32_t state) {
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
inline uint32_t pcg_hash(uint32_t x) { return pcg_rxs_m_xs(x * 747796405u + 2891336453u); }
// FunctionsDetail.h:89-98 (or the PCG step)
inline float rnd(uint32_t& prev) {
    if (g_rng_pcg) {
        prev = prev * 747796405u + 2891336453u;
        return (float)(pcg_rxs_m_xs(prev) >> 8) / (float)0x01000000;
    }
    prev = 1664525u * prev + 1013904223u;
    return (float)(prev & 0x00FFFFFFu) / (float)0x01000000;
}
// FunctionsDetail.h:100-117
template <typename R> inline V3<R> toLocalFrame(const V3<R>& N, const V3<R>& d) {
    const R x = N.x, y = N.y, z = N.z;
    const R sz = (z >= (R)0) ? (R)1 : (R)-1;
    const R a = (R)1 / (sz + z);
    const R ya = y * a;
    const R b = x * ya;
    const R c = x * sz;
    V3<R> lx(c * x * a - (R)1, sz * b, c);
    V3<R> ly(b, y * ya - sz, y);
    return d.x * lx + d.y * ly + d.z * N;
}

// ---------------------------------------------------------------- scene data
enum GeomType { SPHERE = 0, PLANE = 1, QUAD = 2, TRIANGLE = 3, MESH = 4 };     // Geometry.h:15-23
enum MatType { LAMBERT = 0, PHONG = 1, MIRROR = 2, GLASS = 3, GGX = 4, VOLUME = 5 };  // Material.h:14-22
enum LightType { POINT = 0, AREA = 1 };                                        // Light.h:11-15

template <typename R>
struct Volume {  // lib/include/DataStructure/Volume.h:15-21
    V3<R> sigma_s = V3<R>(1), sigma_a = V3<R>(1);
    R g = 0;
    int interface_ = 0;  // 0 NONE, 1 GLASS
};
template <typename R>
struct Material {  // lib/include/Geometry/Material.h:35-43
    V3<R> albedo_e = V3<R>(0), albedo_d = V3<R>(1), albedo_s = V3<R>(0);
    R shininess = (R)(128.0f * 0.4f), roughness = 1, eta = (R)1.5f;
    Volume<R> volume;
    int type = LAMBERT;
};
template <typename R>
struct Light {  // lib/include/DataStructure/Light.h:28-34
    int type = POINT;
    V3<R> position, intensity, radiance, halfExtend1, halfExtend2;
};
template <typename R>
struct Object {
    int type = TRIANGLE;
    V3<R> a, b, c, d;       // sphere: a=center,r=rad; plane: a=pos,b=normal; quad: a=pos,b=normal,c=e1,d=e2; tri: a,b,c vertices
    R radius = 0;
    V3<R> tri_normal;       // triangle: normalize(cross(v2-v1, v3-v1))  (TriangleDetail.h:16)
    uint32_t first = 0, count = 0;  // mesh: triangle range in Scene::mesh_tris
    bool finite_box = false;        // Geometry.cu:6-9 (infinite default); SphereDetail.h:16; MeshDetail.h:18
    V3<R> bmin, bmax;
    Material<R> mat;
};
template <typename R>
struct Tri {  // a mesh triangle (ObjLoader.cpp:72-76); its _id is never set (hit records use the mesh id)
    V3<R> v1, v2, v3, n;
};
template <typename R>
struct Hit {  // LocalGeometry, Geometry.h:28-36; tri (index inside a mesh) is an addition
    R depth = std::numeric_limits<R>::infinity();
    V3<R> P, N;
    int obj = -1;
    int tri = -1;
};

// ---------------------------------------------------------------- intersections
// AABB::hit, lib/src/Geometry/AABBDetail.h:18-50 (biased 1/(d+1e-5), no sign fix)
template <typename R>
inline bool aabb_hit(const V3<R>& mn, const V3<R>& mx, const V3<R>& o, const V3<R>& d) {
    bool hit = true;
    R t_min = 0, t_max = std::numeric_limits<R>::infinity();
    for (int a = 0; a < 3; ++a) {
        R invD = (R)1 / (d[a] + (R)kEps);
        R t0 = (mn[a] - o[a]) * invD;
        R t1 = (mx[a] - o[a]) * invD;
        if (invD < (R)0) std::swap(t0, t1);
        t_min = t0 > t_min ? t0 : t_min;
        t_max = t1 < t_max ? t1 : t_max;
        if (t_max <= t_min) hit = false;
    }
    return hit;
}
// Triangle::computeRayIntersection, TriangleDetail.h:21-57 (Moller-Trumbore)
template <typename R>
inline bool tri_intersect(const V3<R>& v1, const V3<R>& v2, const V3<R>& v3, const V3<R>& o, const V3<R>& d, R& t_out) {
    V3<R> v0v1 = v2 - v1;
    V3<R> v0v2 = v3 - v1;
    V3<R> pvec = cross(d, v0v2);
    R det = dot(v0v1, pvec);
    if (safeFloatEqual(det, (R)0)) return false;
    R invDet = (R)1 / det;
    V3<R> tvec = o - v1;
    R u = dot(tvec, pvec) * invDet;
    if (u < (R)0 || u > (R)1) return false;
    V3<R> qvec = cross(tvec, v0v1);
    R v = dot(d, qvec) * invDet;
    if (v < (R)0 || u + v > (R)1) return false;
    R t = dot(v0v2, qvec) * invDet;
    if (t < (R)0) return false;
    t_out = t;
    return true;
}
// Sphere::computeRayIntersection, SphereDetail.h:33-74
template <typename R>
inline bool sphere_intersect(const V3<R>& c, R rad, const V3<R>& o, const V3<R>& d, R& t_out) {
    V3<R> OS = o - c;
    R b = (R)2 * dot(d, OS);
    R cc = dot(OS, OS) - rad * rad;
    R disc = b * b - (R)4 * cc;
    if (disc > 0) {
        R distSqrt = std::sqrt(disc);
        R q = b < 0 ? (-b - distSqrt) / (R)2 : (-b + distSqrt) / (R)2;
        R t0 = q, t1 = cc / q;
        R mn = std::fmin(t0, t1), mx = std::fmax(t0, t1);
        R t = mn < 0 ? mx : mn;
        if (t >= 0) { t_out = t; return true; }
    }
    return false;
}
// Plane::computeRayIntersection, PlaneDetail.h:19-46
template <typename R>
inline bool plane_intersect(const V3<R>& p, const V3<R>& n, const V3<R>& o, const V3<R>& d, R& t_out) {
    R PdotN = dot(p, n), OdotN = dot(o, n), DdotN = dot(d, n);
    if (safeFloatEqual(DdotN, (R)0)) return false;
    R t = (PdotN - OdotN) / DdotN;
    if (t < 0) return false;
    t_out = t;
    return true;
}
// Quad::computeRayIntersection, QuadDetail.h:21-57
template <typename R>
inline bool quad_intersect(const V3<R>& p, const V3<R>& n, const V3<R>& e1, const V3<R>& e2, const V3<R>& o, const V3<R>& d, R& t_out) {
    R PdotN = dot(p, n), OdotN = dot(o, n), DdotN = dot(d, n);
    if (safeFloatEqual(DdotN, (R)0)) return false;
    R t = (PdotN - OdotN) / DdotN;
    if (t < 0) return false;
    V3<R> I = o + t * d;
    R x = dot(I - p, e1), y = dot(I - p, e2);
    R s1 = dot(e1, e1), s2 = dot(e2, e2);
    if (x < -s1 || x > s1 || y < -s2 || y > s2) return false;
    t_out = t;
    return true;
}

// ---------------------------------------------------------------- materials
// lib/src/Geometry/MaterialDetail.h
template <typename R> inline R D_GGX(R NdotH, R roughness) {  // :11-18
    R a2 = roughness * roughness;
    R d = (NdotH * a2 - NdotH) * NdotH + (R)1;
    return a2 / ((R)(float)kPiD * d * d);
}
template <typename R> inline R V_SmithJointGGX(R NdotL, R NdotV, R roughness) {  // :20-31
    R a2 = roughness * roughness;
    R denomA = NdotV * std::sqrt(a2 + ((R)1 - a2) * NdotL * NdotL);
    R denomB = NdotL * std::sqrt(a2 + ((R)1 - a2) * NdotV * NdotV);
    R G = (R)2 * NdotL * NdotV / (denomA + denomB + (R)kEps);
    return G / ((R)4 * NdotV * NdotL + (R)kEps);
}
template <typename R> inline V3<R> brdf_lambert(const Material<R>& m) { return m.albedo_d / (R)(float)kPiD; }  // :202-207
template <typename R> inline V3<R> brdf_phong(const Material<R>& m, const V3<R>& i, const V3<R>& o, const V3<R>& n) {  // :209-215
    V3<R> h = normalize(i + o);
    return m.albedo_s * (R)std::pow(std::fmax((R)0, dot(h, n)), m.shininess);
}
template <typename R> inline V3<R> brdf_mirror(const Material<R>& m, const V3<R>& i, const V3<R>& o, const V3<R>& n) {  // :217-223
    V3<R> r = reflect(i, n);
    return m.albedo_s * delta((R)1 - dot(o, r)) / dot(o, n);
}
template <typename R> inline V3<R> btdf_glass(const Material<R>& m, const V3<R>& i, const V3<R>& o, const V3<R>& n) {  // :225-257
    if (dot(i, n) * dot(o, n) > 0) {
        if (dot(i, n) > 0) return brdf_mirror(m, i, o, n);
        return brdf_mirror(m, i, o, (R)-1 * n);
    }
    V3<R> refr, nn = n;
    if (dot(i, n) > 0) {
        refr = refract((R)1 / m.eta, i, n);
    } else {
        refr = refract(m.eta, i, (R)-1 * n);
        nn = (R)-1 * n;
    }
    return m.albedo_s * delta((R)1 - dot(refr, o)) / dot(i, nn);
}
template <typename R> inline V3<R> brdf_ggx(const Material<R>& m, const V3<R>& i, const V3<R>& o, const V3<R>& n) {  // :259-274
    V3<R> H = normalize(i + o);
    R NdotH = dot(n, H), LdotH = dot(o, H), NdotL = dot(n, o), NdotV = dot(n, i);
    R ndf = D_GGX(NdotH, m.roughness);
    R vis = V_SmithJointGGX(NdotL, NdotV, m.roughness);
    return (ndf * vis) * fresnel_schlick(m.albedo_s, LdotH);
}
template <typename R> inline V3<R> brdf(const Material<R>& m, const V3<R>& i, const V3<R>& o, const V3<R>& n) {  // :35-74
    switch (m.type) {
        case LAMBERT: return brdf_lambert(m);
        case PHONG: return brdf_phong(m, i, o, n) + brdf_lambert(m);
        case MIRROR: return brdf_mirror(m, i, o, n);
        case GLASS: return btdf_glass(m, i, o, n);
        case GGX: return brdf_ggx(m, i, o, n);
        case VOLUME: return V3<R>(0);
    }
    return V3<R>(0);
}
template <typename R> struct V4 { V3<R> v; R w; };
template <typename R> inline V4<R> sample_lambert(uint32_t& seed, const V3<R>& n) {  // :113-132
    const R xi1 = (R)rnd(seed), xi2 = (R)rnd(seed);
    const R r = std::sqrt(xi1);
    const R phi = (R)2.0f * (R)3.14159f * xi2;
    const R x = r * std::cos(phi), y = r * std::sin(phi);
    const R z = std::sqrt(std::fmax((R)0, (R)1 - x * x - y * y));
    V3<R> dir = normalize(toLocalFrame(n, V3<R>(x, y, z)));
    R p = std::fmax((R)kEps, dot(dir, n)) / (R)3.14159f;
    return {dir, p};
}
template <typename R> inline V4<R> sample_mirror(const V3<R>& i, const V3<R>& n) { return {reflect(i, n), (R)1}; }  // :134-141
template <typename R> inline V4<R> sample_glass(const Material<R>& m, uint32_t& seed, const V3<R>& i, const V3<R>& n_) {  // :143-169
    const R NdotV = dot(i, n_);
    bool outside = NdotV > (R)0;
    R eta_ = outside ? (R)1 / m.eta : m.eta;
    V3<R> n = outside ? n_ : (R)-1 * n_;
    R F0 = (m.eta - (R)1) / (m.eta + (R)1);
    F0 *= F0;
    R p_reflect = fresnel_schlick(F0, std::abs(dot(i, n)));
    R xi = (R)rnd(seed);
    V3<R> refr = refract(eta_, i, n);
    V3<R> dir;
    if (xi <= p_reflect || safeFloatEqual(norm(refr), (R)0)) dir = reflect(i, n);
    else dir = normalize(refr);
    return {dir, (R)1};
}
template <typename R> inline V4<R> sample_ggx(const Material<R>& m, uint32_t& seed, const V3<R>& i, const V3<R>& n) {  // :171-200
    R u = (R)rnd(seed), v = (R)rnd(seed);
    R cosTheta = std::sqrt(((R)1 - u) / ((R)1 + (m.roughness * m.roughness - (R)1) * u));
    R sinTheta = std::sqrt(std::fmax((R)0, (R)1 - cosTheta * cosTheta));
    R phi = (R)2.0f * (R)(float)kPiD * v;
    R x = sinTheta * std::cos(phi), y = sinTheta * std::sin(phi), z = cosTheta;
    V3<R> H = toLocalFrame(n, V3<R>(x, y, z));
    V3<R> L = reflect(i, H);
    R LdotH = dot(i, H), NdotH = dot(n, H);
    if (dot(n, L) <= (R)0) return {V3<R>(0), (R)1};
    R p = D_GGX(NdotH, m.roughness) * NdotH / std::abs((R)4 * LdotH);
    return {L, p};
}
template <typename R> inline V4<R> sampleDirection(const Material<R>& m, uint32_t& seed, const V3<R>& i, const V3<R>& n) {  // :76-111
    switch (m.type) {
        case LAMBERT:
        case PHONG: return sample_lambert(seed, n);
        case MIRROR: return sample_mirror(i, n);
        case GLASS: return sample_glass(m, seed, i, n);
        case GGX: return sample_ggx(m, seed, i, n);
        case VOLUME: return m.volume.interface_ == 1 ? sample_glass(m, seed, i, n) : V4<R>{(R)-1 * i, (R)1};
    }
    return {V3<R>(std::numeric_limits<R>::infinity()), std::numeric_limits<R>::infinity()};
}
// :276-286 (0.25 / float(M_PI) is evaluated in double, then narrowed)
template <typename R> inline R hg_phase(R g, R cos_theta) {
    if (std::abs(g) < (R)1e-3f) return (R)(0.25 / (double)(float)kPiD);
    R g2 = g * g;
    R area = (R)4.0f * (R)3.14159f;
    return ((R)1 - g2) / area * (R)std::pow((R)1 + g2 - (R)2 * g * cos_theta, (R)-1.5f);
}
// :288-313
template <typename R> inline V4<R> hg_sample(R g_, const V3<R>& fwd, uint32_t& seed) {
    R g = g_ < (R)0 ? (R)std::fmin((R)(-1e-3), g_) : (R)std::fmax((R)1e-3, g_);
    R u1 = (R)rnd(seed), u2 = (R)rnd(seed);
    R g2 = g * g;
    R d = ((R)1 - g2) / ((R)1 - g + (R)2 * g * u1);
    R cos_theta = clampf((R)0.5f / g * ((R)1 + g2 - d * d), (R)-1, (R)1);
    R sin_theta = std::sqrt(std::fmax((R)0, (R)1 - cos_theta * cos_theta));
    R phi = (R)2.0f * (R)3.14159f * u2;
    R x = sin_theta * std::cos(phi), y = sin_theta * std::sin(phi), z = cos_theta;
    V3<R> res = normalize(toLocalFrame(fwd, V3<R>(x, y, z)));
    R pdf = hg_phase(g, dot(fwd, res));
    return {res, pdf};
}

// ---------------------------------------------------------------- lights
// lib/src/DataStructure/LightDetail.h:8-52
template <typename R> inline V3<R> light_sample(const Light<R>& L, uint32_t& seed, const V3<R>& pos, V3<R>& dir, R& dist) {
    if (L.type == POINT) {
        dir = normalize(L.position - pos);
        dist = norm(L.position - pos);
        return L.intensity / (dist * dist);
    }
    R xi1 = (R)(rnd(seed) * 2.0f - 1.0f);
    R xi2 = (R)(rnd(seed) * 2.0f - 1.0f);
    V3<R> s = L.position + xi1 * L.halfExtend1 + xi2 * L.halfExtend2;
    V3<R> n = normalize(cross(L.halfExtend1, L.halfExtend2));
    R area = (R)4.0f * norm(L.halfExtend1) * norm(L.halfExtend2);
    dir = normalize(s - pos);
    dist = norm(s - pos);
    R NdotL = dot(dir, n);
    if (NdotL < 0) NdotL *= (R)-1;
    R solid = area * NdotL / (dist * dist);
    return L.radiance * solid;
}

// ---------------------------------------------------------------- camera
// lib/src/DataStructure/Camera.cpp:6-11 + Camera.h:72-76
template <typename R>
struct Camera {
    V3<R> pos = V3<R>(0), xa = V3<R>(1, 0, 0), ya = V3<R>(0, 1, 0), za = V3<R>(0, 0, 1);
};
struct Ray3 { V3<float> o, d; };
// Tracing::launchRay, lib/src/Core/TracingDetail.h:15-45 (+ Ray ctor normalisation, RayDetail.h:6-13)
template <typename R>
inline void launch_ray(uint32_t px, uint32_t py, uint32_t w, uint32_t h, const Camera<R>& cam, uint32_t& seed, V3<R>& o, V3<R>& d) {
    float jx = rnd(seed), jy = rnd(seed);
    const R rx = (R)2.0f * (((R)(float)px + (R)jx) / (R)w - (R)0.5f);
    const R ry = (R)2.0f * (((R)(float)py + (R)jy) / (R)h - (R)0.5f);
    const V3<R> wp = cam.pos + cam.za + rx * cam.xa + ry * cam.ya;
    o = cam.pos;
    d = normalize(wp - cam.pos);
}
// Tracing::direction2UV, TracingDetail.h:185-196 (double arithmetic from M_PI, narrowed to float)
template <typename R>
inline void direction2uv(const V3<R>& d, uint32_t w, uint32_t h, uint32_t& x, uint32_t& y) {
    float theta, phi;
    if (sizeof(R) == 4) {
        theta = (float)((double)std::acos((float)d.y) / kPiD);
        phi = (float)(((double)std::atan2((float)d.z, (float)d.x) + kPiD) / (2.0 * kPiD));
    } else {
        theta = (float)(std::acos((double)d.y) / kPiD);
        phi = (float)((std::atan2((double)d.z, (double)d.x) + kPiD) / (2.0 * kPiD));
    }
    // the float -> unsigned casts; a NaN direction (undefined in the reference) maps to texel 0
    float fx = (float)(w - 1) * phi, fy = (float)(h - 1) * theta;
    x = fx >= 0.0f ? (uint32_t)fx : 0u;
    y = fy >= 0.0f ? (uint32_t)fy : 0u;
}

}  // namespace orc
