// The ray-tracing integrator (one thread per pixel), the gradient-domain reconstruction, and the
// host entry points of the volume and gradient-domain renderers, whose paths run in wavefront
// form on the path tracer's queues (pathtracer.hip).
//
//   volume   volume_kernel + handleMediumInteraction, lib/src/Renderer/VolumeRenderer.cu:11-329
//   gdpt     init_kernel, optimization_kernel, lib/src/Renderer/GradientDomain.cu:248-402
//   raytracer raytracing_kernel, lib/src/Renderer/RayTracer.cu:11-65
// The ray tracer shares the BVH4 traversal and the shading code of the path tracer.
#include "renderer.h"
#include "shading.h"
#include "traverse.h"

namespace cupbr {

constexpr int kIntBlock = 256;
constexpr int kIntLds = 32;

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
};
using Tracer = TracerT<false>;

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
    // volume_kernel's paths in wavefront form (pathtracer.hip)
    return volume_paths_wavefront(r, cam, frame0, spp);
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
