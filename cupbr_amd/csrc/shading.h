// Device-side shading: BSDFs, phase function, light sampling, camera rays and
// the environment lookup, following the reference estimator operation by
// operation (lib/src/Geometry/MaterialDetail.h, lib/src/DataStructure/LightDetail.h,
// lib/src/Core/TracingDetail.h). Transcendentals use the device math library.
#pragma once

#include "scene_types.h"

namespace cupbr {

struct Mat {  // register copy of cupbr_material (only the fields a lobe reads)
    v3 albedo_e, albedo_d, albedo_s;
    float shininess, roughness, eta;
    int32_t type;
    int32_t interface_type;
};
__device__ __forceinline__ Mat load_mat(const cupbr_material* m) {
    Mat r;
    r.albedo_e = ld3(m->albedo_e);
    r.albedo_d = ld3(m->albedo_d);
    r.albedo_s = ld3(m->albedo_s);
    r.shininess = m->shininess;
    r.roughness = m->roughness;
    r.eta = m->eta;
    r.type = m->type;
    r.interface_type = m->volume.interface_type;
    return r;
}

__device__ __forceinline__ float d_ggx(float ndoth, float rough) {  // MaterialDetail.h:11-18
    float a2 = rough * rough;
    float d = (ndoth * a2 - ndoth) * ndoth + 1.0f;
    return a2 / ((float)kPi * d * d);
}
__device__ __forceinline__ float v_smith_joint_ggx(float ndotl, float ndotv, float rough) {  // :20-31
    float a2 = rough * rough;
    float da = ndotv * sqrtf(a2 + (1.0f - a2) * ndotl * ndotl);
    float db = ndotl * sqrtf(a2 + (1.0f - a2) * ndotv * ndotv);
    float g = 2.0f * ndotl * ndotv / (da + db + kEpsilon);
    return g / (4.0f * ndotv * ndotl + kEpsilon);
}
__device__ __forceinline__ v3 brdf_mirror(const Mat& m, v3 i, v3 o, v3 n) {  // :217-223
    v3 r = reflect(i, n);
    return m.albedo_s * delta(1.0f - dot(o, r)) / dot(o, n);
}
// Material::brdf, MaterialDetail.h:35-74
__device__ __forceinline__ v3 brdf(const Mat& m, v3 i, v3 o, v3 n) {
    switch (m.type) {
        case CUPBR_LAMBERT: return m.albedo_d / (float)kPi;
        case CUPBR_PHONG: {
            v3 h = normalize(i + o);
            return m.albedo_s * powf(fmaxf(0.0f, dot(h, n)), m.shininess) + m.albedo_d / (float)kPi;
        }
        case CUPBR_MIRROR: return brdf_mirror(m, i, o, n);
        case CUPBR_GLASS: {  // btdf_glass :225-257
            if (dot(i, n) * dot(o, n) > 0) {
                if (dot(i, n) > 0) return brdf_mirror(m, i, o, n);
                return brdf_mirror(m, i, o, -1.0f * n);
            }
            v3 refr, nn = n;
            if (dot(i, n) > 0) refr = refract(1.0f / m.eta, i, n);
            else { refr = refract(m.eta, i, -1.0f * n); nn = -1.0f * n; }
            return m.albedo_s * delta(1.0f - dot(refr, o)) / dot(i, nn);
        }
        case CUPBR_GGX: {  // brdf_ggx :259-274
            v3 h = normalize(i + o);
            float ndoth = dot(n, h), ldoth = dot(o, h), ndotl = dot(n, o), ndotv = dot(n, i);
            float ndf = d_ggx(ndoth, m.roughness);
            float vis = v_smith_joint_ggx(ndotl, ndotv, m.roughness);
            return (ndf * vis) * fresnel_schlick(m.albedo_s, ldoth);
        }
        default: return splat(0.0f);
    }
}

struct DirPdf {
    v3 d;
    float pdf;
};
template <class Seed>
__device__ __forceinline__ DirPdf sample_glass(const Mat& m, Seed& seed, v3 i, v3 n0) {  // :143-169
    const float ndotv = dot(i, n0);
    bool outside = ndotv > 0.0f;
    float eta = outside ? 1.0f / m.eta : m.eta;
    v3 n = outside ? n0 : -1.0f * n0;
    float f0 = (m.eta - 1.0f) / (m.eta + 1.0f);
    f0 *= f0;
    float p_reflect = fresnel_schlick(f0, fabsf(dot(i, n)));
    float xi = rnd(seed);
    v3 refr = refract(eta, i, n);
    v3 dir;
    if (xi <= p_reflect || safe_float_equal(norm(refr), 0.0f)) dir = reflect(i, n);
    else dir = normalize(refr);
    return {dir, 1.0f};
}
// Material::sampleDirection, MaterialDetail.h:76-111
template <class Seed>
__device__ __forceinline__ DirPdf sample_direction(const Mat& m, Seed& seed, v3 i, v3 n) {
    switch (m.type) {
        case CUPBR_LAMBERT:
        case CUPBR_PHONG: {  // sample_lambert :113-132
            const float xi1 = rnd(seed), xi2 = rnd(seed);
            const float r = sqrtf(xi1);
            const float phi = 2.0f * 3.14159f * xi2;
            const float x = r * cosf(phi), y = r * sinf(phi);
            const float z = sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y));
            v3 dir = normalize(to_local_frame(n, mk(x, y, z)));
            return {dir, fmaxf(kEpsilon, dot(dir, n)) / 3.14159f};
        }
        case CUPBR_MIRROR: return {reflect(i, n), 1.0f};
        case CUPBR_GLASS: return sample_glass(m, seed, i, n);
        case CUPBR_GGX: {  // sample_ggx :171-200
            float u = rnd(seed), v = rnd(seed);
            float cos_t = sqrtf((1.0f - u) / (1.0f + (m.roughness * m.roughness - 1.0f) * u));
            float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
            float phi = 2.0f * (float)kPi * v;
            v3 h = to_local_frame(n, mk(sin_t * cosf(phi), sin_t * sinf(phi), cos_t));
            v3 l = reflect(i, h);
            float ldoth = dot(i, h), ndoth = dot(n, h);
            if (dot(n, l) <= 0.0f) return {splat(0.0f), 1.0f};
            return {l, d_ggx(ndoth, m.roughness) * ndoth / fabsf(4.0f * ldoth)};
        }
        case CUPBR_VOLUME_MATERIAL:
            if (m.interface_type == CUPBR_INTERFACE_GLASS) return sample_glass(m, seed, i, n);
            return {-1.0f * i, 1.0f};
        default: return {splat(__builtin_inff()), __builtin_inff()};
    }
}

// Material::henyeyGreensteinPhaseFunction, :276-286 (0.25 / float(pi) evaluated in double)
__device__ __forceinline__ float hg_phase(float g, float cos_t) {
    if (fabsf(g) < 1e-3f) return (float)(0.25 / (double)(float)kPi);
    float g2 = g * g;
    float area = 4.0f * 3.14159f;
    return (1.0f - g2) / area * powf(1.0f + g2 - 2.0f * g * cos_t, -1.5f);
}
// sampleHenyeyGreensteinPhaseFunction, :288-313
template <class Seed>
__device__ __forceinline__ DirPdf hg_sample(float g_, v3 fwd, Seed& seed) {
    float g = g_ < 0.0f ? fminf(-1e-3f, g_) : fmaxf(1e-3f, g_);
    float u1 = rnd(seed), u2 = rnd(seed);
    float g2 = g * g;
    float d = (1.0f - g2) / (1.0f - g + 2.0f * g * u1);
    float cos_t = clampf(0.5f / g * (1.0f + g2 - d * d), -1.0f, 1.0f);
    float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
    float phi = 2.0f * 3.14159f * u2;
    v3 res = normalize(to_local_frame(fwd, mk(sin_t * cosf(phi), sin_t * sinf(phi), cos_t)));
    return {res, hg_phase(g, dot(fwd, res))};
}

// Light::sample, lib/src/DataStructure/LightDetail.h:8-52
template <class Seed>
__device__ __forceinline__ v3 light_sample(const cupbr_light& L, Seed& seed, v3 pos, v3& dir, float& dist) {
    if (L.type == CUPBR_POINT) {
        v3 lp = ld3(L.position);
        dir = normalize(lp - pos);
        dist = norm(lp - pos);
        return ld3(L.intensity) / (dist * dist);
    }
    float xi1 = rnd(seed) * 2.0f - 1.0f;
    float xi2 = rnd(seed) * 2.0f - 1.0f;
    v3 e1 = ld3(L.half_extend1), e2 = ld3(L.half_extend2);
    v3 s = ld3(L.position) + xi1 * e1 + xi2 * e2;
    v3 n = normalize(cross(e1, e2));
    float area = 4.0f * norm(e1) * norm(e2);
    dir = normalize(s - pos);
    dist = norm(s - pos);
    float ndotl = dot(dir, n);
    if (ndotl < 0) ndotl *= -1.0f;
    float solid = area * ndotl / (dist * dist);
    return ld3(L.radiance) * solid;
}

// Tracing::direction2UV (TracingDetail.h:185-196; the pi terms are double) + Image lookup.
// Non-finite directions map to texel 0 (the reference's float-to-unsigned cast is undefined there).
__device__ __forceinline__ v3 env_lookup(const DevScene& S, v3 d) {
    float theta = (float)((double)acosf(d.y) / kPi);
    float phi = (float)(((double)atan2f(d.z, d.x) + kPi) / (2.0 * kPi));
    float fx = (float)(S.env_w - 1) * phi, fy = (float)(S.env_h - 1) * theta;
    uint32_t x = fx >= 0.0f ? (uint32_t)fx : 0u, y = fy >= 0.0f ? (uint32_t)fy : 0u;
    x = min(x, S.env_w - 1);
    y = min(y, S.env_h - 1);
    return ld3(S.env + 3 * ((size_t)y * S.env_w + x));
}

// Light choice and sample of every NEE site (PathTracer.cu:43-62 and its copies in
// VolumeRenderer.cu / GradientDomain.cu). Returns false when there is nothing to sample:
// no lights and no environment, or an environment direction from a failed BSDF sample
// (a zero vector, whose shadow ray would carry NaN).
template <class Seed>
__device__ __forceinline__ bool nee_sample(const DevScene& S, const Mat& m, Seed& seed, v3 P, v3 inc, v3 N, v3& ldir,
                                           v3& le, float& dist) {
    const uint32_t pick = (uint32_t)(rnd(seed) * (float)(S.n_lights + (S.use_env ? 1u : 0u)));
    if (pick != S.n_lights && S.n_lights) {
        const cupbr_light L = S.lights[min(pick, S.n_lights - 1)];
        le = light_sample(L, seed, P, ldir, dist);
        return true;
    }
    if (!S.use_env) return false;
    const DirPdf sp = sample_direction(m, seed, inc, N);
    ldir = sp.d;
    dist = __builtin_inff();
    le = env_lookup(S, ldir) / sp.pdf;
    return norm(ldir) > 0.0f;
}

// Tracing::launchRay, TracingDetail.h:15-45 (+ Ray's normalisation)
template <class Seed>
__device__ __forceinline__ void launch_ray(uint32_t px, uint32_t py, uint32_t w, uint32_t h, const cupbr_camera& cam,
                                           Seed& seed, v3& o, v3& d) {
    float jx = rnd(seed), jy = rnd(seed);
    const float rx = 2.0f * (((float)px + jx) / (float)w - 0.5f);
    const float ry = 2.0f * (((float)py + jy) / (float)h - 0.5f);
    const v3 pos = ld3(cam.position);
    const v3 wp = pos + ld3(cam.z_axis) + rx * ld3(cam.x_axis) + ry * ld3(cam.y_axis);
    o = pos;
    d = normalize(wp - pos);
}

// Normal and object of a closest hit (LocalGeometry, lib/include/Geometry/Geometry.h:28-36):
// planes and quads from their record, triangles from the stored face normal, spheres from the
// hit point (SphereDetail.h:66). One 16-byte record per triangle hit.
__device__ __forceinline__ void hit_geometry(const DevScene& S, uint32_t prim, v3 P, v3& N, uint32_t& obj) {
    if (prim & kUnbBit) {
        const UnbRec& u = S.unb[prim & ~kUnbBit];
        N = mk(u.n.x, u.n.y, u.n.z);
        obj = __float_as_uint(u.e1.w);
        return;
    }
    const float4 n = S.normals[prim];
    const uint32_t bits = __float_as_uint(n.w);
    obj = bits & ~kNormalSphereBit;
    if (bits & kNormalSphereBit) {
        const float4 a = S.prims[prim].a;
        N = normalize(P - mk(a.x, a.y, a.z));
    } else {
        N = mk(n.x, n.y, n.z);
    }
}

// Pixel of work item j of a w x h image, for kernels whose 64-wide waves take 64 consecutive
// items: 8x8 pixel tiles (row-major tile order) over the part of the image whole tiles cover, then
// the right-hand strip and the bottom rows in row order. A bijection of [0, w*h); a wave then
// holds one 8x8 tile, whose rays reach neighbouring geometry (the path tracer's camera-ray queue:
// 9% faster than 64x1 pixel rows at 1080p / 1M triangles, since the order carries over to the
// later bounces' queues).
__device__ __forceinline__ void tile_pixel(uint32_t j, uint32_t w, uint32_t h, uint32_t& px, uint32_t& py) {
    const uint32_t w8 = w & ~7u, h8 = h & ~7u, n_full = w8 * h8;
    if (j < n_full) {
        const uint32_t t = j >> 6, k = j & 63u, tiles_x = w8 >> 3;
        px = (t % tiles_x) * 8u + (k & 7u);
        py = (t / tiles_x) * 8u + (k >> 3);
        return;
    }
    const uint32_t r = j - n_full, wr = w - w8, n_right = wr * h8;
    if (r < n_right) {
        py = r / wr;
        px = w8 + r % wr;
    } else {
        const uint32_t r2 = r - n_right;
        py = h8 + r2 / w;
        px = r2 % w;
    }
}

// Inverse of tile_pixel: the work item of pixel (px, py).
__device__ __forceinline__ uint32_t tile_slot(uint32_t px, uint32_t py, uint32_t w, uint32_t h) {
    const uint32_t w8 = w & ~7u, h8 = h & ~7u, n_full = w8 * h8, wr = w - w8;
    if (px < w8 && py < h8) return (((py >> 3) * (w8 >> 3) + (px >> 3)) << 6) + ((py & 7u) << 3) + (px & 7u);
    if (py < h8) return n_full + py * wr + (px - w8);
    return n_full + wr * h8 + (py - h8) * w + px;
}

// Queue slot of sample s (frame of a batch of nb frames) of the pixel at work item j (tile_pixel
// order), and its inverse. With f frames per wave (f a power of two dividing nb, 64 / f dividing
// w * h) a 64-wide wave holds f consecutive frames of 64 / f consecutive work items, and the waves
// run over the frame groups of a pixel group before the next pixel group: at 64 frames per batch
// a wave traces the 64 samples of one pixel. With f = 1 the queue is frame-major.
struct SampleOrder {
    uint32_t npix, nb, f;
    __device__ __forceinline__ uint32_t slot(uint32_t s, uint32_t j) const {
        if (f == 1u) return s * npix + j;
        const uint32_t pp = 64u / f, g = nb / f;
        return (((j / pp) * g + s / f) << 6) + (j % pp) * f + s % f;
    }
    __device__ __forceinline__ void sample(uint32_t p, uint32_t& s, uint32_t& j) const {
        if (f == 1u) {
            s = p / npix;
            j = p - s * npix;
            return;
        }
        const uint32_t pp = 64u / f, g = nb / f, wv = p >> 6, k = p & 63u;
        const uint32_t pg = wv / g;
        s = (wv - pg * g) * f + k % f;
        j = pg * pp + k / f;
    }
};

// coefficients of the current medium of the volume renderer: the scene's (obj < 0) or an
// object's interior (Volume.h:15-21; VolumeRenderer.cu:88-98)
struct VolCoef {
    v3 sa, ss;
    float g;
};
__device__ __forceinline__ VolCoef vol_coef(const DevScene& S, int32_t obj) {
    VolCoef c;
    if (obj < 0) {
        c.sa = ld3(S.volume.sigma_a);
        c.ss = ld3(S.volume.sigma_s);
        c.g = S.volume.g;
    } else {
        const cupbr_volume& v = S.mats[obj].volume;
        c.sa = ld3(v.sigma_a);
        c.ss = ld3(v.sigma_s);
        c.g = v.g;
    }
    return c;
}

}  // namespace cupbr
