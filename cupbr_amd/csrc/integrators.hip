// Volume and ray-tracing integrators (one thread per pixel) and the gradient-domain reconstruction.
//
//   volume   volume_kernel + handleMediumInteraction, lib/src/Renderer/VolumeRenderer.cu:11-329
//   gdpt     init_kernel, optimization_kernel, lib/src/Renderer/GradientDomain.cu:248-402 (the
//            paths of gdpt_kernel run in wavefront form, pathtracer.hip)
//   raytracer raytracing_kernel, lib/src/Renderer/RayTracer.cu:11-65
// They share the BVH4 traversal and the shading code of the path tracer.
#include "renderer.h"
#include "shading.h"
#include "traverse.h"

namespace cupbr {

constexpr int kIntBlock = 256;
constexpr int kIntLds = 32;
// The volume kernel holds a whole path (and its shading) per thread; left to itself the compiler
// uses ~160 VGPRs (2-3 waves per SIMD). Capping it at 128 VGPRs (4 waves per SIMD) spills a few
// values and runs 1.2x faster; a 96-VGPR cap (5 waves) spills far more and is slower.
#define CUPBR_INT_OCCUPANCY __attribute__((amdgpu_waves_per_eu(4, 8)))

// kCount: BVH4 node visits and primitive tests of closest-hit (te) and visibility (ts) queries
template <bool kCount = false>
struct TracerT {
    const DevScene& S;
    LaneStack<kIntLds>& st;
    uint32_t ext = 0, shadow = 0;  // ray counts of this lane
    TravCount te, ts;

    __device__ __forceinline__ bool closest(v3 o, v3 d, TraceResult& r) {
        ++ext;
        traverse<false, kIntLds, kCount>(S, o, d, 0.0f, st, r, &te);
        return r.prim != kNoHit;
    }
    // Tracing::traceVisibility, TracingDetail.h:118-183
    __device__ __forceinline__ bool visible(v3 o, v3 d, float dist) {
        ++shadow;
        TraceResult r;
        return !traverse<true, kIntLds, kCount>(S, o, d, dist, st, r, &ts);
    }
    // Tracing::traceRay(scene, ray, index), TracingDetail.h:68-116: one object, no box test;
    // a mesh keeps its first strictly-closer triangle (MeshDetail.h:28-35)
    __device__ bool object(uint32_t obj, v3 o, v3 d, TraceResult& r) {
        ++ext;
        r.t = __builtin_inff();
        r.prim = kNoHit;
        r.rank = 0xFFFFFFFFu;
        const ObjInfo& oi = S.objs[obj];
        for (uint32_t k = 0; k < oi.rank_count; ++k) {
            const uint32_t prim = S.rank_prim[oi.rank_base + k];
            if (prim == kNoHit) continue;
            float t;
            bool h;
            if (prim & kUnbBit) {
                h = plane_quad_test(S.unb[prim & ~kUnbBit], o, d, t);
            } else {
                const PrimRec pr = S.prims[prim];
                if ((__float_as_uint(pr.c.w) & 3u) == kPrimTriangle)
                    h = mt_test(mk(pr.a.x, pr.a.y, pr.a.z), mk(pr.a.w, pr.b.x, pr.b.y), mk(pr.b.z, pr.b.w, pr.c.x), o, d, t);
                else
                    h = sphere_test(mk(pr.a.x, pr.a.y, pr.a.z), pr.a.w, o, d, t);
            }
            if (h && t < r.t) {
                r.t = t;
                r.prim = prim;
                r.rank = oi.rank_base + k;
            }
        }
        return r.prim != kNoHit;
    }
};
using Tracer = TracerT<false>;

template <bool kCount>
__device__ __forceinline__ void add_visit_counts(Counters* c, const TracerT<kCount>& T) {
    if (!kCount) return;
    unsigned long long v[4] = {T.te.nodes, T.te.prims, T.ts.nodes, T.ts.prims};
    for (int k = 0; k < 4; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off);
    if (__lane_id() == 0) {
        atomicAdd(&c->ext_nodes, v[0]);
        atomicAdd(&c->ext_prims, v[1]);
        atomicAdd(&c->shadow_nodes, v[2]);
        atomicAdd(&c->shadow_prims, v[3]);
    }
}

__device__ __forceinline__ void add_ray_counts(Counters* c, uint32_t ext, uint32_t shadow) {
    unsigned long long e = ext, s = shadow;
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_xor(e, off);
        s += __shfl_xor(s, off);
    }
    if (__lane_id() == 0) {
        atomicAdd(&c->extension_rays, e);
        atomicAdd(&c->shadow_rays, s);
    }
}

// ------------------------------------------------------------------ volume
// coefficients of the current medium: the scene's (obj < 0) or an object's interior
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

template <class TR>
__device__ v3 volume_path(TR& T, const cupbr_camera& cam, uint32_t tid, uint32_t w, uint32_t h, uint32_t frame,
                          uint32_t max_depth, bool rr) {
    const DevScene& S = T.S;
    const float inf = __builtin_inff();
    uint32_t seed = tea4(tid, frame);
    v3 o, d;
    launch_ray(tid % w, tid / w, w, h, cam, seed, o, d);
    v3 radiance = splat(0.0f), weight = splat(1.0f), ray_start = splat(0.0f), out_dir = splat(0.0f);
    bool inside = false;
    uint32_t obj_index = 0;
    const float nl = (float)(S.n_lights + (S.use_env ? 1u : 0u));
    int32_t vol_obj = -1;  // current medium: -1 the scene's, else the interior of that object
    uint32_t depth = 0;
    do {
        bool next_valid = false;
        TraceResult hr;
        const bool hit = inside ? T.object(obj_index, o, d, hr) : T.closest(o, d, hr);
        const v3 inc = -1.0f * d;
        if (!hit) {
            if (S.use_env) radiance += weight * env_lookup(S, d);
            break;
        }
        const v3 P = o + hr.t * d;
        v3 N;
        uint32_t obj;
        hit_geometry(S, hr.prim, P, N, obj);
        const cupbr_material* mp = S.mats + obj;
        const Mat m = load_mat(mp);
        if (inside) vol_obj = (int32_t)obj;  // the payload's volume pointer follows the current hit (:277)
        // handleMediumInteraction :88-212
        bool medium_event = false;
        {
            const VolCoef vc = vol_coef(S, vol_obj);
            const float g = vc.g;
            const v3 sa = vc.sa, ss = vc.ss, st = sa + ss;
            uint32_t ch = (uint32_t)(rnd(seed) * 3.0f);
            ch = min(ch, 2u);
            const float stc = comp(st, (int)ch);
            if (!safe_float_equal(stc, 0.0f)) {
                const float t = -logf(1.0f - rnd(seed)) / stc;
                if (t < hr.t) {
                    medium_event = true;
                    const v3 ep = o + t * d;
                    const float sprob = comp(ss, (int)ch) / stc;
                    if (rnd(seed) < sprob) {
                        float pdf = 0.0f;
                        for (int i = 0; i < 3; ++i) pdf += comp(st, i) * expf(-1.0f * comp(st, i) * t);
                        pdf /= 3.0f;
                        weight = weight * ss / sprob * vexp(-1.0f * st * t) / pdf;
                    } else {
                        weight = splat(0.0f);
                        break;
                    }
                    v3 ldir, le;
                    float dist;
                    if (nee_sample(S, m, seed, ep, inc, N, ldir, le, dist)) {
                        v3 so, sd = normalize(ldir), att = splat(1.0f);
                        if (inside) {
                            TraceResult ge;
                            float gd = 0.0f;
                            v3 gp = ep;
                            if (T.object(obj_index, ep, sd, ge)) {
                                gd = ge.t;
                                gp = ep + ge.t * sd;
                            }
                            att = vexp(-1.0f * st * gd) * hg_phase(g, -dot(inc, ldir));
                            so = gp + 0.001f * ldir;
                            dist -= gd;
                        } else {
                            so = ep + 0.001f * ldir;
                        }
                        const v3 sst = ld3(S.volume.sigma_a) + ld3(S.volume.sigma_s);
                        if (T.visible(so, sd, dist))
                            radiance += nl * vexp(-1.0f * sst * fminf(dist, 100000000.0f)) * att * le * weight;
                    }
                    const DirPdf hs = hg_sample(g, -1.0f * inc, seed);
                    out_dir = hs.d;
                    ray_start = ep;
                    next_valid = true;
                } else {
                    float pdf = 0.0f;
                    for (int i = 0; i < 3; ++i) pdf += expf(-comp(st, i) * hr.t);
                    pdf /= 3.0f;
                    weight = weight * vexp(-1.0f * st * hr.t) / pdf;
                    ray_start = P;
                }
            }
        }
        if (!medium_event) {
            if (m.type == CUPBR_VOLUME_MATERIAL) {  // medium interface :269-278
                const DirPdf s = sample_direction(m, seed, inc, N);
                out_dir = s.d;
                ray_start = P;
                const bool refl = dot(inc, N) * dot(out_dir, N) > 0;
                inside = refl ? inside : !inside;
                obj_index = obj;
                next_valid = true;
                vol_obj = inside ? (int32_t)obj : -1;
            } else {
                if (m.type != CUPBR_GLASS && dot(N, inc) <= 0.0f) {  // directIlluminationVolumetric :24-70
                    weight = splat(0.0f);
                } else {
                    v3 ldir, le;
                    float dist;
                    if (nee_sample(S, m, seed, P, inc, N, ldir, le, dist)) {
                        if (T.visible(P + 0.001f * ldir, normalize(ldir), dist)) {
                            const VolCoef vc = vol_coef(S, vol_obj);
                            radiance += (nl * fmaxf(0.0f, dot(N, ldir))) *
                                        vexp(-1.0f * (vc.ss + vc.sa) * fminf(dist, 100000000.0f)) *
                                        brdf(m, inc, ldir, N) * le * weight;
                        }
                    }
                }
                const DirPdf s = sample_direction(m, seed, inc, N);  // indirectIlluminationVolumetric :72-86
                if (!(norm(s.d) == 0.0f)) {
                    weight = weight * fabsf(dot(s.d, N)) * brdf(m, inc, s.d, N) / s.pdf;
                    out_dir = s.d;
                    ray_start = P;
                    next_valid = true;
                }
            }
            if (rr) {
                const float alpha = clampf(fmaxf(weight.x, fmaxf(weight.y, weight.z)), 0.0f, 1.0f);
                if (rnd(seed) > alpha || safe_float_equal(alpha, 0.0f)) break;
                weight = weight / alpha;
            }
        }
        o = ray_start + 0.001f * out_dir;  // :298
        d = normalize(out_dir);
        if (!next_valid) break;
        ++depth;
    } while (depth < max_depth);
    return radiance;
}

__device__ __forceinline__ void running_mean(float* px, v3 L, uint32_t f) {  // PathTracer.cu:162-168
    if (f > 0) {
        const float a = 1.0f / ((float)f + 1.0f);
        st3(px, (1.0f - a) * ld3(px) + a * L);
    } else {
        st3(px, L);
    }
}

template <bool kCount>
__global__ __launch_bounds__(kIntBlock) CUPBR_INT_OCCUPANCY void k_volume(DevScene S, cupbr_camera cam, uint32_t w, uint32_t h,
                                                      uint32_t frame0, uint32_t nspp, uint32_t max_depth, int rr,
                                                      float* __restrict__ img, Counters* c, uint32_t* spill) {
    __shared__ uint32_t lds[kIntLds * kIntBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    LaneStack<kIntLds> st{lds + threadIdx.x, kIntBlock, spill ? spill + gtid : nullptr, gridDim.x * blockDim.x};
    TracerT<kCount> T{S, st};
    const uint32_t npix = w * h;
    if (gtid < npix) {
        for (uint32_t s = 0; s < nspp; ++s) {
            const v3 L = volume_path(T, cam, gtid, w, h, frame0 + s, max_depth, rr != 0);
            running_mean(img + 3 * (size_t)gtid, L, frame0 + s);
        }
    }
    add_ray_counts(c, T.ext, T.shadow);
    add_visit_counts(c, T);
}

// ------------------------------------------------------------------ gradient-domain
// The paths of gdpt_kernel (GradientDomain.cu:110-246) run in wavefront form (pathtracer.hip,
// gdpt_paths_wavefront); the reconstruction kernels follow.

// init_kernel, GradientDomain.cu:248-273
__global__ void k_gdpt_init(float* recon, const float* base, const float* gxf, const float* gxb, const float* gyf,
                            const float* gyb, float* gx, float* gy, uint32_t w, uint32_t h) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    st3(recon + 3 * (size_t)i, ld3(base + 3 * (size_t)i));
    const uint32_t x = i % w, y = i / w;
    if (x == 0 || y == 0 || x == w - 1 || y == h - 1) return;
    st3(gx + 3 * (size_t)i, ld3(gxf + 3 * (size_t)i) + ld3(gxb + 3 * (size_t)i));
    st3(gy + 3 * (size_t)i, ld3(gyf + 3 * (size_t)i) + ld3(gyb + 3 * (size_t)i));
}

__device__ __forceinline__ float median4(float v0, float v1, float v2, float v3_) {  // compute_median :275-301
    const float v[4] = {v0, v1, v2, v3_};
    float m = 0.0f;
    uint32_t amin = 0, amax = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        m += v[i];
        if (v[i] < v[amin]) amin = i;
        if (v[i] > v[amax]) amax = i;
    }
    return (m - v[amin] - v[amax]) / 2.0f;
}

// optimization_kernel, GradientDomain.cu:303-350 (interior pixels; the border of out stays zero)
__global__ void k_gdpt_opt(const float* __restrict__ recon, const float* __restrict__ gx, const float* __restrict__ gy,
                           float* __restrict__ out, uint32_t w, uint32_t h) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const uint32_t x = i % w, y = i / w;
    if (x == 0 || y == 0 || x == w - 1 || y == h - 1) return;
    const v3 v1 = ld3(recon + 3 * (size_t)(i - 1)) + ld3(gx + 3 * (size_t)(i - 1));
    const v3 v2 = ld3(recon + 3 * (size_t)(i - w)) + ld3(gy + 3 * (size_t)(i - w));
    const v3 v3_ = ld3(recon + 3 * (size_t)(i + 1)) - ld3(gx + 3 * (size_t)i);
    const v3 v4 = ld3(recon + 3 * (size_t)(i + w)) - ld3(gy + 3 * (size_t)i);
    st3(out + 3 * (size_t)i, mk(median4(v1.x, v2.x, v3_.x, v4.x), median4(v1.y, v2.y, v3_.y, v4.y),
                                median4(v1.z, v2.z, v3_.z, v4.z)));
}

// Screened-Poisson (L2) reconstruction, one Jacobi iteration over the interior pixels (no
// counterpart in the reference): (a2 + 4) I(p) = a2 B(p) + (I(L) + gx(L)) + (I(D) + gy(D)) +
// (I(R) - gx(p)) + (I(U) - gy(p)), i.e. the minimiser of a2 |I - B|^2 + |dx I - gx|^2 + |dy I - gy|^2
// with the gradients of init_kernel read as forward differences, as the trimmed mean reads them.
// The border of out stays zero.
__global__ void k_gdpt_l2(const float* __restrict__ recon, const float* __restrict__ base,
                          const float* __restrict__ gx, const float* __restrict__ gy, float* __restrict__ out,
                          uint32_t w, uint32_t h, float a2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const uint32_t x = i % w, y = i / w;
    if (x == 0 || y == 0 || x == w - 1 || y == h - 1) return;
    const v3 v1 = ld3(recon + 3 * (size_t)(i - 1)) + ld3(gx + 3 * (size_t)(i - 1));
    const v3 v2 = ld3(recon + 3 * (size_t)(i - w)) + ld3(gy + 3 * (size_t)(i - w));
    const v3 v3_ = ld3(recon + 3 * (size_t)(i + 1)) - ld3(gx + 3 * (size_t)i);
    const v3 v4 = ld3(recon + 3 * (size_t)(i + w)) - ld3(gy + 3 * (size_t)i);
    const v3 b = ld3(base + 3 * (size_t)i);
    const float den = a2 + 4.0f;
    st3(out + 3 * (size_t)i, (a2 * b + ((v1 + v2) + (v3_ + v4))) / den);
}

// zero border of an image (the trimmed-mean and L2 iterates keep a zero border)
__global__ void k_zero_border(float* img, uint32_t w, uint32_t h) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const uint32_t x = i % w, y = i / w;
    if (x == 0 || y == 0 || x == w - 1 || y == h - 1) st3(img + 3 * (size_t)i, splat(0.0f));
}

// ------------------------------------------------------------------ ray tracer
// raytracing_kernel, RayTracer.cu:11-54: primary hit, every light as a point light, shadow rays
__global__ __launch_bounds__(kIntBlock) void k_raytrace(DevScene S, cupbr_camera cam, uint32_t w, uint32_t h,
                                                        float* __restrict__ img, Counters* c, uint32_t* spill) {
    __shared__ uint32_t lds[kIntLds * kIntBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    LaneStack<kIntLds> st{lds + threadIdx.x, kIntBlock, spill ? spill + gtid : nullptr, gridDim.x * blockDim.x};
    Tracer T{S, st};
    if (gtid < w * h) {
        const uint32_t px = gtid % w, py = gtid / w;
        const float rx = 2.0f * (((float)px + 0.0f) / (float)w - 0.5f);
        const float ry = 2.0f * (((float)py + 0.0f) / (float)h - 0.5f);
        const v3 pos = ld3(cam.position);
        const v3 wp = pos + ld3(cam.z_axis) + rx * ld3(cam.x_axis) + ry * ld3(cam.y_axis);
        const v3 d = normalize(wp - pos);
        TraceResult hr;
        v3 radiance = splat(0.0f);
        if (T.closest(pos, d, hr)) {
            const v3 P = pos + hr.t * d;
            v3 N;
            uint32_t obj;
            hit_geometry(S, hr.prim, P, N, obj);
            const Mat m = load_mat(S.mats + obj);
            const v3 inc = normalize(pos - P);
            for (uint32_t l = 0; l < S.n_lights; ++l) {
                const v3 lp = ld3(S.lights[l].position);
                const v3 ldir = normalize(lp - P);
                const v3 b = brdf(m, inc, ldir, N);
                const float dist = norm(P - lp);
                const v3 lrad = ld3(S.lights[l].intensity) / (dist * dist);
                const float cos_t = fmaxf(0.0f, dot(N, ldir));
                if (T.visible(P - kEpsilon * d, normalize(ldir), dist)) radiance += b * lrad * cos_t;
            }
        }
        st3(img + 3 * (size_t)gtid, radiance);
    }
    add_ray_counts(c, T.ext, T.shadow);
}

// ------------------------------------------------------------------ host launchers
static int prepare(cupbr_renderer* r, uint32_t threads, uint32_t** spill, Counters** c) {
    int rc;
    if ((rc = r->wf.counters.ensure(sizeof(Counters)))) return rc;
    *c = r->wf.counters.as<Counters>();
    CUPBR_HIP_CHECK(hipMemsetAsync(*c, 0, sizeof(Counters), r->stream));
    *spill = nullptr;
    const uint32_t entries = r->scene->stack_entries;
    if (entries > (uint32_t)kIntLds) {
        if ((rc = r->wf.spill.ensure((size_t)(entries - kIntLds) * threads * 4))) return rc;
        *spill = r->wf.spill.as<uint32_t>();
    }
    return CUPBR_OK;
}

int render_volume(cupbr_renderer* r, const cupbr_camera& cam, uint32_t frame0, uint32_t spp) {
    const uint32_t npix = r->width * r->height;
    const uint32_t blocks = (npix + kIntBlock - 1) / kIntBlock;
    uint32_t* spill;
    Counters* c;
    int rc = prepare(r, blocks * kIntBlock, &spill, &c);
    if (rc) return rc;
    // one launch per 4 samples keeps each launch short
    for (uint32_t s = 0; s < spp; s += 4) {
        hipLaunchKernelGGL(r->counting ? k_volume<true> : k_volume<false>, dim3(blocks), dim3(kIntBlock), 0, r->stream,
                           r->scene->dev, cam, r->width,
                           r->height, frame0 + s, std::min(4u, spp - s), std::max(1u, r->max_trace_depth),
                           (int)r->russian_roulette, r->images[CUPBR_IMAGE_OUTPUT].as<float>(), c, spill);
    }
    CUPBR_HIP_CHECK(hipGetLastError());
    return CUPBR_OK;
}

int render_gdpt(cupbr_renderer* r, const cupbr_camera& cam, uint32_t frame0, uint32_t spp) {
    const uint32_t w = r->width, h = r->height, npix = w * h;
    const uint32_t blocks = (npix + kIntBlock - 1) / kIntBlock;
    int rc;
    float* im[kNumImages];
    for (int k = 0; k < kNumImages; ++k) im[k] = r->images[k].as<float>();
    // the five paths of every (pixel, frame) in wavefront form (pathtracer.hip)
    if ((rc = gdpt_paths_wavefront(r, cam, frame0, spp))) return rc;
    // reconstruction (:380-401); a frame's reconstruction depends only on the current running
    // means, so it runs once after the call's last frame
    hipLaunchKernelGGL(k_gdpt_init, dim3(blocks), dim3(kIntBlock), 0, r->stream, im[CUPBR_IMAGE_OUTPUT],
                       im[CUPBR_IMAGE_BASE], im[CUPBR_IMAGE_GRADIENT_X_FORWARD], im[CUPBR_IMAGE_GRADIENT_X_BACKWARD],
                       im[CUPBR_IMAGE_GRADIENT_Y_FORWARD], im[CUPBR_IMAGE_GRADIENT_Y_BACKWARD],
                       im[CUPBR_IMAGE_GRADIENT_X], im[CUPBR_IMAGE_GRADIENT_Y], w, h);
    float* a = im[CUPBR_IMAGE_OUTPUT];
    float* b = r->temp.as<float>();
    if (r->gdpt_mode == CUPBR_GDPT_TRIMMED_MEAN) {
        for (int it = 0; it < 50; ++it) {  // ping-pong; 50 is even so the result lands in the output image
            hipLaunchKernelGGL(k_gdpt_opt, dim3(blocks), dim3(kIntBlock), 0, r->stream, a, im[CUPBR_IMAGE_GRADIENT_X],
                               im[CUPBR_IMAGE_GRADIENT_Y], b, w, h);
            std::swap(a, b);
        }
    } else {
        // the iterates hold a zero border (temp's border is never written by the Jacobi kernel)
        hipLaunchKernelGGL(k_zero_border, dim3(blocks), dim3(kIntBlock), 0, r->stream, a, w, h);
        hipLaunchKernelGGL(k_zero_border, dim3(blocks), dim3(kIntBlock), 0, r->stream, b, w, h);
        const float a2 = r->gdpt_alpha * r->gdpt_alpha;
        for (uint32_t it = 0; it < r->gdpt_iterations; ++it) {
            hipLaunchKernelGGL(k_gdpt_l2, dim3(blocks), dim3(kIntBlock), 0, r->stream, a, im[CUPBR_IMAGE_BASE],
                               im[CUPBR_IMAGE_GRADIENT_X], im[CUPBR_IMAGE_GRADIENT_Y], b, w, h, a2);
            std::swap(a, b);
        }
        if (a != im[CUPBR_IMAGE_OUTPUT])
            CUPBR_HIP_CHECK(hipMemcpyAsync(im[CUPBR_IMAGE_OUTPUT], a, (size_t)npix * 12, hipMemcpyDeviceToDevice,
                                           r->stream));
    }
    CUPBR_HIP_CHECK(hipGetLastError());
    return CUPBR_OK;
}

int render_raytracer(cupbr_renderer* r, const cupbr_camera& cam) {
    const uint32_t npix = r->width * r->height;
    const uint32_t blocks = (npix + kIntBlock - 1) / kIntBlock;
    uint32_t* spill;
    Counters* c;
    int rc = prepare(r, blocks * kIntBlock, &spill, &c);
    if (rc) return rc;
    hipLaunchKernelGGL(k_raytrace, dim3(blocks), dim3(kIntBlock), 0, r->stream, r->scene->dev, cam, r->width, r->height,
                       r->images[CUPBR_IMAGE_OUTPUT].as<float>(), c, spill);
    CUPBR_HIP_CHECK(hipGetLastError());
    return CUPBR_OK;
}

}  // namespace cupbr
