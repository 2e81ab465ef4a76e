// Heterogeneous-medium path tracer with a path-replay adjoint. The reference has no counterpart:
// its volumes are homogeneous (lib/include/DataStructure/Volume.h:15-21,
// lib/src/Renderer/VolumeRenderer.cu:97-124); this integrator keeps its camera, RNG seeding,
// light sampling and Henyey-Greenstein functions and adds a voxel grid.
//
// Model
//   surfaces   every scene primitive is a two-sided Lambertian surface whose RGB albedo is a
//              parameter: surface index = object id, or n_objects + global triangle index for a
//              mesh triangle
//   medium     a box [bmin, bmax] with sigma_t(x) = scale * density(x), density trilinear between
//              the centres of an nx*ny*nz grid (x fastest, clamped at the faces); sigma_s = alpha
//              * sigma_t; phase Henyey-Greenstein(g)
//   lights     the scene's lights, one picked uniformly per vertex (weight n_lights)
//
// Estimator (one camera path per (pixel, frame), seed tea<4>(pixel, frame) + LCG)
//   Free flight by weighted delta tracking against the constant majorant m: tentative collisions
//   at exponential distances of rate m; each is real with probability
//   P = clamp(proposal(x) * scale / m, eps, 1 - eps) and then scatters, else null. Path weights
//   are multiplied by alpha*sigma/(m*P) (real) or (1 - sigma/m)/(1 - P) (null). The proposal is
//   a density grid held fixed (the parameters by default), so for fixed seeds every sampling
//   decision is independent of the parameters and the estimate is a smooth function of the
//   density and albedo values: same-seed finite differences converge to the adjoint gradient.
//   With proposal = density the weights reduce to alpha and 1 (ordinary delta tracking).
//   NEE at every vertex; shadow transmittance by ratio tracking, prod (1 - sigma/m) over
//   tentative points, after an any-hit visibility test against the surfaces. Surface
//   continuation is cosine-sampled (weight = albedo), medium continuation HG-sampled (weight 1).
//
// Adjoint (path replay, constant memory)
//   Each path is run twice with the same seed: first for its radiance L, then again carrying
//   L_rem, the radiance still to be collected. A throughput factor f adds
//   <dI, L_rem> * dlog(f)/dsigma * scale * w_k to the 8 voxels k of its trilinear stencil (and
//   <dI, L_rem>/albedo per channel for a surface's continuation factor); an NEE term C adds
//   its own derivative and is then subtracted from L_rem. Gradients are accumulated with
//   device atomics. A factor that is exactly zero ends the path, so the derivative through
//   that factor is dropped (a zero-measure event for densities below the majorant).
#include <algorithm>
#include <cstring>
#include <cmath>

#include "device.h"
#include "scene.h"
#include "shading.h"
#include "traverse.h"

namespace cupbr {

constexpr int kGridBlock = 256;
constexpr int kGridLds = 32;
constexpr float kSurfaceOffset = 1e-4f;
constexpr int kAlbLds = 3 * 1024;  // surfaces whose albedo gradients a block accumulates in LDS
constexpr size_t kMaxGridSamples = (size_t)1 << 24;  // samples per forward chunk (192 MB buffer)
// Bound on tentative collisions per segment. The host requires majorant * box diagonal <= 1e5,
// so the expected count per segment is below 1e5 and the bound is never reached in practice;
// reaching it ends the path (or zeroes the shadow ray's transmittance), so every loop terminates.
constexpr uint32_t kMaxTrackSteps = 1u << 22;
// The adjoint's density-gradient atomics go to one private copy of the gradient per group of
// blocks, block b -> copy b % kGradCopies. Workgroups are dispatched to the 8 XCDs round-robin,
// so each copy is written by one XCD and its lines stay in that XCD's L2 instead of moving
// between L2s on every atomic; a reduction adds the copies to the caller's gradient in copy
// order. The mapping only affects speed: any assignment of blocks to copies sums the same terms.
#ifndef CUPBR_GRID_GRAD_COPIES
#define CUPBR_GRID_GRAD_COPIES 8
#endif
constexpr uint32_t kGradCopies = CUPBR_GRID_GRAD_COPIES;
constexpr size_t kMaxGradCopyBytes = (size_t)1 << 30;  // larger grids scatter into the output directly

struct GridDev {
    const float* rho;   // density parameters
    const float* prop;  // proposal density (fixed)
    const float* alb;   // surface albedo, 3 per surface
    float bmin[3], ext[3];
    uint32_t n[3];
    float scale, maj, alpha, g, eps;
    uint32_t n_objs;
};

struct Stencil {
    uint32_t idx[8];
    float w[8];
};

// trilinear stencil of x: corner k = (dx, dy, dz) = (k & 1, (k >> 1) & 1, k >> 2)
__device__ __forceinline__ void stencil(const GridDev& G, v3 x, Stencil& s) {
    const float p[3] = {x.x, x.y, x.z};
    uint32_t i0[3], i1[3];
    float f[3];
    for (int a = 0; a < 3; ++a) {
        const float u = (p[a] - G.bmin[a]) / G.ext[a] * (float)G.n[a] - 0.5f;
        const float fl = floorf(u);
        f[a] = u - fl;
        const int lo = (int)fl;
        const int hi_max = (int)G.n[a] - 1;
        i0[a] = (uint32_t)min(max(lo, 0), hi_max);
        i1[a] = (uint32_t)min(max(lo + 1, 0), hi_max);
    }
    for (int k = 0; k < 8; ++k) {
        const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
        const uint32_t ix = dx ? i1[0] : i0[0], iy = dy ? i1[1] : i0[1], iz = dz ? i1[2] : i0[2];
        s.idx[k] = (iz * G.n[1] + iy) * G.n[0] + ix;
        s.w[k] = ((dx ? f[0] : 1.0f - f[0]) * (dy ? f[1] : 1.0f - f[1])) * (dz ? f[2] : 1.0f - f[2]);
    }
}

__device__ __forceinline__ float sample_grid(const float* g, const Stencil& s) {
    float v = 0.0f;
    for (int k = 0; k < 8; ++k) v += s.w[k] * g[s.idx[k]];
    return v;
}

__device__ __forceinline__ uint32_t scatter_grad(float* grad, const Stencil& s, float c) {
    if (c == 0.0f || !(fabsf(c) < __builtin_inff())) return 0;
    uint32_t n = 0;
    for (int k = 0; k < 8; ++k)
        if (s.w[k] != 0.0f) {
            atomicAdd(grad + s.idx[k], c * s.w[k]);
            ++n;
        }
    return n;
}

// parametric range [t0, t1] of o + t d inside the medium box, clipped to [0, tmax]
__device__ __forceinline__ bool clip_box(const GridDev& G, v3 o, v3 d, float tmax, float& t0, float& t1) {
    const float po[3] = {o.x, o.y, o.z}, pd[3] = {d.x, d.y, d.z};
    t0 = 0.0f;
    t1 = tmax;
    for (int a = 0; a < 3; ++a) {
        const float lo = G.bmin[a], hi = G.bmin[a] + G.ext[a];
        if (pd[a] == 0.0f) {
            if (po[a] < lo || po[a] > hi) return false;
            continue;
        }
        float ta = (lo - po[a]) / pd[a], tb = (hi - po[a]) / pd[a];
        if (ta > tb) {
            const float tt = ta;
            ta = tb;
            tb = tt;
        }
        t0 = fmaxf(t0, ta);
        t1 = fminf(t1, tb);
    }
    return t0 < t1;
}

// adjoint state of one path
struct Adj {
    v3 delta;     // dLoss/dI of the pixel, divided by spp
    v3 rem;       // radiance of the path not yet collected
    float* grad_rho;
    float* grad_alb;
};

struct GridTracer {
    const DevScene& S;
    LaneStack<kGridLds>& st;
    // work of this lane: closest-hit rays, shadow rays, grid stencil reads, gradient atomics
    uint32_t n_ext = 0, n_shadow = 0, n_reads = 0, n_atomics = 0;
    __device__ __forceinline__ bool closest(v3 o, v3 d, TraceResult& r) {
        ++n_ext;
        traverse<false>(S, o, d, 0.0f, st, r);
        return r.prim != kNoHit;
    }
    __device__ __forceinline__ bool visible(v3 o, v3 d, float dist) {
        ++n_shadow;
        TraceResult r;
        return !traverse<true>(S, o, d, dist, st, r);
    }
};

// shadow-ray transmittance through the medium by ratio tracking
__device__ __forceinline__ float ratio_track(const GridDev& G, GridTracer& Tr, v3 o, v3 d, float dist, uint32_t& seed) {
    float t0, t1;
    if (!clip_box(G, o, d, dist, t0, t1)) return 1.0f;
    const float len = t1 - t0;
    float T = 1.0f, u = 0.0f;
    for (uint32_t it = 0;; ++it) {
        if (it == kMaxTrackSteps) return 0.0f;
        u = u - logf(1.0f - rnd(seed)) / G.maj;
        if (!(u < len)) break;
        Stencil s;
        stencil(G, o + (t0 + u) * d, s);
        const float sigma = G.scale * sample_grid(G.rho, s);
        ++Tr.n_reads;
        T = T * (1.0f - sigma / G.maj);
    }
    return T;
}

// the same tentative points again, adding coef * dlog(1 - sigma/m)/dsigma * scale to the voxels
__device__ __forceinline__ void ratio_track_grad(const GridDev& G, GridTracer& Tr, v3 o, v3 d, float dist,
                                                 uint32_t seed, float coef, float* grad) {
    float t0, t1;
    if (!clip_box(G, o, d, dist, t0, t1)) return;
    const float len = t1 - t0;
    float u = 0.0f;
    for (uint32_t it = 0; it < kMaxTrackSteps; ++it) {
        u = u - logf(1.0f - rnd(seed)) / G.maj;
        if (!(u < len)) break;
        Stencil s;
        stencil(G, o + (t0 + u) * d, s);
        const float sigma = G.scale * sample_grid(G.rho, s);
        ++Tr.n_reads;
        Tr.n_atomics += scatter_grad(grad, s, coef * (-1.0f / (G.maj - sigma)) * G.scale);
    }
}

__device__ __forceinline__ float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// transmittance-weighted NEE from x towards a light; returns the light term (Le * n_lights * T)
// and sets the direction; T includes surface visibility
template <bool kAdj>
__device__ __forceinline__ bool nee(const GridDev& G, const DevScene& S, GridTracer& T, v3 x, uint32_t& seed, v3& ldir,
                                    v3& le, float& tr, uint32_t& track_seed, float& dist, bool surface, v3 N) {
    if (S.n_lights == 0) return false;
    uint32_t k = (uint32_t)(rnd(seed) * (float)S.n_lights);
    k = min(k, S.n_lights - 1);
    le = light_sample(S.lights[k], seed, x, ldir, dist) * (float)S.n_lights;
    if (surface && !(dot(N, ldir) > 0.0f)) return false;
    const v3 so = surface ? x + kSurfaceOffset * N : x;
    if (!T.visible(so, ldir, dist)) return false;
    track_seed = seed;
    tr = ratio_track(G, T, so, ldir, dist, seed);
    return true;
}

template <bool kAdj>
__device__ v3 grid_path(const DevScene& S, const GridDev& G, GridTracer& T, const cupbr_camera& cam, uint32_t px,
                        uint32_t py, uint32_t w, uint32_t h, uint32_t frame, uint32_t max_depth, Adj* A) {
    const float pi = (float)kPi;
    uint32_t seed = tea4(py * w + px, frame);
    v3 o, d;
    launch_ray(px, py, w, h, cam, seed, o, d);
    v3 L = splat(0.0f), beta = splat(1.0f);
    for (uint32_t depth = 0; depth < max_depth; ++depth) {
        TraceResult hr;
        const bool hit = T.closest(o, d, hr);
        const float ts = hit ? hr.t : __builtin_inff();
        // free flight through the medium up to the surface
        bool scattered = false;
        v3 x;
        float t0, t1;
        if (clip_box(G, o, d, ts, t0, t1)) {
            const float len = t1 - t0;
            float u = 0.0f;
            for (uint32_t it = 0;; ++it) {
                if (it == kMaxTrackSteps) return L;
                u = u - logf(1.0f - rnd(seed)) / G.maj;
                if (!(u < len)) break;
                x = o + (t0 + u) * d;
                Stencil s;
                stencil(G, x, s);
                const float sigma = G.scale * sample_grid(G.rho, s);
                const float P = clampf(G.scale * sample_grid(G.prop, s) / G.maj, G.eps, 1.0f - G.eps);
                T.n_reads += 2;
                float f, dlog;
                const bool real = rnd(seed) < P;
                if (real) {
                    f = G.alpha * sigma / (G.maj * P);
                    dlog = 1.0f / sigma;
                } else {
                    f = (1.0f - sigma / G.maj) / (1.0f - P);
                    dlog = -1.0f / (G.maj - sigma);
                }
                if (f == 0.0f) return L;
                beta = beta * f;
                if (kAdj) T.n_atomics += scatter_grad(A->grad_rho, s, dot3(A->delta, A->rem) * dlog * G.scale);
                if (real) {
                    scattered = true;
                    break;
                }
            }
        }
        if (scattered) {
            v3 ldir, le;
            float tr, dist;
            uint32_t tseed;
            if (nee<kAdj>(G, S, T, x, seed, ldir, le, tr, tseed, dist, false, splat(0.0f))) {
                const float ph = hg_phase(G.g, dot(d, ldir));
                const v3 C = beta * (le * (ph * tr));
                L += C;
                if (kAdj) {
                    const float c = dot3(A->delta, C);
                    if (c != 0.0f) ratio_track_grad(G, T, x, ldir, dist, tseed, c, A->grad_rho);
                    A->rem = A->rem - C;
                }
            }
            if (depth + 1 == max_depth) break;
            const DirPdf hs = hg_sample(G.g, d, seed);
            o = x;
            d = hs.d;
            continue;
        }
        if (!hit) break;
        // Lambertian surface
        const v3 P = o + ts * d;
        v3 N;
        uint32_t surf;
        hit_geometry(S, hr.prim, P, N, surf);
        if (!(hr.prim & kUnbBit)) {
            const ObjInfo& oi = S.objs[surf];
            if (oi.type == CUPBR_MESH) surf = G.n_objs + oi.first_triangle + (hr.rank - oi.rank_base);
        }
        if (dot(N, d) > 0.0f) N = -1.0f * N;
        const v3 a = ld3(G.alb + 3 * (size_t)surf);
        v3 ldir, le;
        float tr, dist;
        uint32_t tseed;
        if (nee<kAdj>(G, S, T, P, seed, ldir, le, tr, tseed, dist, true, N)) {
            const v3 E0 = le * (dot(N, ldir) * tr / pi);  // light term without the albedo
            const v3 C = beta * (a * E0);
            L += C;
            if (kAdj) {
                const v3 ga = A->delta * beta * E0;
                float* gp = A->grad_alb + 3 * (size_t)surf;
                if (ga.x != 0.0f) atomicAdd(gp + 0, ga.x);
                if (ga.y != 0.0f) atomicAdd(gp + 1, ga.y);
                if (ga.z != 0.0f) atomicAdd(gp + 2, ga.z);
                T.n_atomics += 3;
                const float c = dot3(A->delta, C);
                if (c != 0.0f) ratio_track_grad(G, T, P + kSurfaceOffset * N, ldir, dist, tseed, c, A->grad_rho);
                A->rem = A->rem - C;
            }
        }
        if (depth + 1 == max_depth) break;
        const float u1 = rnd(seed), u2 = rnd(seed);
        const float r = sqrtf(u1), phi = 2.0f * pi * u2;
        const v3 nd = normalize(to_local_frame(N, mk(r * cosf(phi), r * sinf(phi), sqrtf(fmaxf(0.0f, 1.0f - u1)))));
        if (a.x == 0.0f && a.y == 0.0f && a.z == 0.0f) break;
        beta = beta * a;
        if (kAdj) {
            const v3 rem = A->rem;
            float* gp = A->grad_alb + 3 * (size_t)surf;
            if (a.x > 0.0f && A->delta.x * rem.x != 0.0f) atomicAdd(gp + 0, A->delta.x * rem.x / a.x);
            if (a.y > 0.0f && A->delta.y * rem.y != 0.0f) atomicAdd(gp + 1, A->delta.y * rem.y / a.y);
            if (a.z > 0.0f && A->delta.z * rem.z != 0.0f) atomicAdd(gp + 2, A->delta.z * rem.z / a.z);
        }
        o = P + kSurfaceOffset * N;
        d = nd;
    }
    return L;
}

// work item i -> (pixel, sample): sample-major, so the 64 lanes of a wave trace 64 neighbouring
// pixels (not 64 samples of one pixel, whose identical first segments would send the same
// gradient atomics to the same voxels from every lane)
template <bool kAdj>
__global__ __launch_bounds__(kGridBlock) void k_grid(DevScene S, GridDev G, cupbr_camera cam, uint32_t w, uint32_t h,
                                                     uint32_t spp, uint32_t frame0, uint32_t s_begin, uint32_t s_count,
                                                     uint32_t max_depth, float* __restrict__ samples,
                                                     const float* __restrict__ d_image, float* grad_rho,
                                                     float* grad_alb, uint32_t n_alb, size_t copy_stride,
                                                     uint32_t* spill, unsigned long long* counts,
                                                     const float* __restrict__ fwd_samples) {
    __shared__ uint32_t lds[kGridLds * kGridBlock];
    // per-block albedo gradients (few surfaces, hit by every path): LDS atomics, one global
    // atomic per entry at the end
    __shared__ float alb_lds[kAlbLds];
    const bool alb_local = kAdj && n_alb <= (uint32_t)kAlbLds;
    if (alb_local) {
        for (uint32_t k = threadIdx.x; k < n_alb; k += blockDim.x) alb_lds[k] = 0.0f;
        __syncthreads();
    }
    const size_t gtid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    // private gradient copies (copy_stride floats apart; 0: scatter into grad_rho itself)
    if (kAdj && copy_stride) grad_rho += (size_t)(blockIdx.x % kGradCopies) * copy_stride;
    LaneStack<kGridLds> st{lds + threadIdx.x, kGridBlock, spill ? spill + gtid : nullptr, (uint32_t)stride};
    GridTracer T{S, st};
    const uint32_t npix = w * h;
    const size_t n = (size_t)npix * s_count;
    for (size_t i = gtid; i < n; i += stride) {
        const uint32_t p = (uint32_t)(i % npix);
        const uint32_t s = s_begin + (uint32_t)(i / npix);
        const uint32_t px = p % w, py = p / w;
        // the adjoint needs the path's radiance: from the forward pass's samples when given
        // (stored at the same index, sample-major), else by running the path once more
        const v3 L = (kAdj && fwd_samples) ? ld3(fwd_samples + 3 * i)
                                           : grid_path<false>(S, G, T, cam, px, py, w, h, frame0 + s, max_depth, nullptr);
        if (!kAdj) {
            st3(samples + 3 * i, L);  // sample-major: samples[(s - s_begin) * npix + p]
        } else {
            Adj A;
            A.delta = ld3(d_image + 3 * (size_t)p) / (float)spp;
            A.rem = L;
            A.grad_rho = grad_rho;
            A.grad_alb = alb_local ? alb_lds : grad_alb;
            if (A.delta.x != 0.0f || A.delta.y != 0.0f || A.delta.z != 0.0f)
                grid_path<true>(S, G, T, cam, px, py, w, h, frame0 + s, max_depth, &A);
        }
    }
    if (alb_local) {
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < n_alb; k += blockDim.x)
            if (alb_lds[k] != 0.0f) atomicAdd(grad_alb + k, alb_lds[k]);
    }
    if (counts) {
        unsigned long long c[4] = {T.n_ext, T.n_shadow, T.n_reads, T.n_atomics};
        for (int k = 0; k < 4; ++k) {
            for (int off = 32; off > 0; off >>= 1) c[k] += __shfl_xor(c[k], off);
            if (__lane_id() == 0) atomicAdd(counts + k, c[k]);
        }
    }
}

__global__ void k_grid_zero(float4* __restrict__ p, size_t n4) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// grad[v] += sum over copies c (in order) of copies[c * stride + v], v < n_vox <= stride
__global__ void k_grid_reduce(float* __restrict__ grad, const float* __restrict__ copies, size_t n_vox,
                              size_t stride) {
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n_vox; v += (size_t)gridDim.x * blockDim.x) {
        float sum = 0.0f;
        for (uint32_t c = 0; c < kGradCopies; ++c) sum += copies[(size_t)c * stride + v];
        if (sum != 0.0f) grad[v] += sum;
    }
}

// image = mean over samples (stored sample-major), summed in sample order across chunks
__global__ void k_grid_accumulate(float* __restrict__ acc, const float* __restrict__ samples, uint32_t n_px,
                                  uint32_t s_count, int first, int last, uint32_t spp) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_px) return;
    v3 v = first ? splat(0.0f) : ld3(acc + 3 * (size_t)p);
    for (uint32_t s = 0; s < s_count; ++s) v += ld3(samples + 3 * ((size_t)s * n_px + p));
    if (last) v = v / (float)spp;
    st3(acc + 3 * (size_t)p, v);
}

namespace {

int check_medium(const cupbr_grid_medium* m) {
    if (!m) return set_error(CUPBR_ERR_INVALID_ARGUMENT, "null medium");
    for (int a = 0; a < 3; ++a)
        if (!(m->bmax[a] > m->bmin[a])) return set_error(CUPBR_ERR_INVALID_ARGUMENT, "medium box must have bmax > bmin");
    if (!m->nx || !m->ny || !m->nz || (uint64_t)m->nx * m->ny * m->nz > (1ull << 28))
        return set_error(CUPBR_ERR_INVALID_ARGUMENT, "invalid grid resolution");
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(m->bmin[a]) || !std::isfinite(m->bmax[a]))
            return set_error(CUPBR_ERR_INVALID_ARGUMENT, "medium box must be finite");
    if (!(m->majorant > 0.0f) || !std::isfinite(m->majorant) || !(m->sigma_t_scale >= 0.0f) ||
        !std::isfinite(m->sigma_t_scale) || !std::isfinite(m->scattering_albedo))
        return set_error(CUPBR_ERR_INVALID_ARGUMENT, "majorant must be finite and > 0, sigma_t_scale finite and >= 0");
    {
        const double dx = (double)m->bmax[0] - m->bmin[0], dy = (double)m->bmax[1] - m->bmin[1],
                     dz = (double)m->bmax[2] - m->bmin[2];
        if ((double)m->majorant * std::sqrt(dx * dx + dy * dy + dz * dz) > 1e5)
            return set_error(CUPBR_ERR_INVALID_ARGUMENT, "majorant * box diagonal must be <= 1e5");
    }
    if (!(m->min_collision_probability > 0.0f && m->min_collision_probability <= 0.5f))
        return set_error(CUPBR_ERR_INVALID_ARGUMENT, "min_collision_probability must be in (0, 0.5]");
    if (!(m->g > -1.0f && m->g < 1.0f)) return set_error(CUPBR_ERR_INVALID_ARGUMENT, "g must be in (-1, 1)");
    return CUPBR_OK;
}

GridDev make_grid(const cupbr_scene* s, const cupbr_grid_medium* m, const float* rho, const float* prop,
                  const float* alb) {
    GridDev G{};
    G.rho = rho;
    G.prop = prop ? prop : rho;
    G.alb = alb;
    for (int a = 0; a < 3; ++a) {
        G.bmin[a] = m->bmin[a];
        G.ext[a] = m->bmax[a] - m->bmin[a];
    }
    G.n[0] = m->nx;
    G.n[1] = m->ny;
    G.n[2] = m->nz;
    G.scale = m->sigma_t_scale;
    G.maj = m->majorant;
    G.alpha = m->scattering_albedo;
    G.g = m->g;
    G.eps = m->min_collision_probability;
    G.n_objs = (uint32_t)s->objects.size();
    return G;
}

int grid_launch_shape(cupbr_scene* s, size_t n, int* blocks, uint32_t** spill) {
    int b = (int)std::min<size_t>((n + kGridBlock - 1) / kGridBlock, (size_t)device_cu_count() * 8);
    b = std::max(b, 1);
    *spill = nullptr;
    if (s->stack_entries > (uint32_t)kGridLds) {
        const size_t need = (size_t)(s->stack_entries - kGridLds) * b * kGridBlock * sizeof(uint32_t);
        if (s->grid_spill.ensure(need)) return CUPBR_ERR_OUT_OF_MEMORY;
        *spill = s->grid_spill.as<uint32_t>();
    }
    *blocks = b;
    return CUPBR_OK;
}

}  // namespace

int grid_set_counting(cupbr_scene* s, bool enable) {
    s->grid_counting = enable;
    if (!enable) return CUPBR_OK;
    if (s->grid_counts.ensure(4 * sizeof(unsigned long long))) return CUPBR_ERR_OUT_OF_MEMORY;
    CUPBR_HIP_CHECK(hipMemset(s->grid_counts.ptr, 0, 4 * sizeof(unsigned long long)));
    return CUPBR_OK;
}

int grid_get_counts(const cupbr_scene* s, uint64_t out[4]) {
    for (int k = 0; k < 4; ++k) out[k] = 0;
    if (!s->grid_counting || !s->grid_counts.ptr) return CUPBR_OK;
    CUPBR_HIP_CHECK(hipDeviceSynchronize());
    CUPBR_HIP_CHECK(hipMemcpy(out, s->grid_counts.ptr, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost));
    return CUPBR_OK;
}

int grid_render_device(const cupbr_scene* cs, const cupbr_camera* cam, const cupbr_grid_medium* m, const float* rho,
                       const float* prop, const float* alb, uint32_t w, uint32_t h, uint32_t frame0, uint32_t spp,
                       uint32_t max_depth, float* image, hipStream_t stream) {
    cupbr_scene* s = const_cast<cupbr_scene*>(cs);
    if (int rc = scene_upload(s)) return rc;
    const GridDev G = make_grid(s, m, rho, prop, alb);
    unsigned long long* counts = s->grid_counting ? s->grid_counts.as<unsigned long long>() : nullptr;
    const uint32_t n_px = w * h;
    const uint32_t chunk = (uint32_t)std::max<size_t>(1, std::min<size_t>(spp, kMaxGridSamples / n_px));
    if (s->grid_samples.ensure((size_t)n_px * chunk * 12)) return CUPBR_ERR_OUT_OF_MEMORY;
    if (int rc = scene_stream_acquire(s, stream)) return rc;
    for (uint32_t s0 = 0; s0 < spp; s0 += chunk) {
        const uint32_t cnt = std::min(chunk, spp - s0);
        int blocks;
        uint32_t* spill;
        if (grid_launch_shape(s, (size_t)n_px * cnt, &blocks, &spill)) return CUPBR_ERR_OUT_OF_MEMORY;
        hipLaunchKernelGGL(k_grid<false>, dim3(blocks), dim3(kGridBlock), 0, stream, s->dev, G, *cam, w, h, spp, frame0,
                           s0, cnt, max_depth, s->grid_samples.as<float>(), nullptr, nullptr, nullptr, 0u, (size_t)0,
                           spill, counts, nullptr);
        CUPBR_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_grid_accumulate, dim3((n_px + 255) / 256), dim3(256), 0, stream, image,
                           s->grid_samples.as<float>(), n_px, cnt, s0 == 0 ? 1 : 0, s0 + cnt == spp ? 1 : 0, spp);
        CUPBR_HIP_CHECK(hipGetLastError());
    }
    if (int rc = scene_stream_release(s, stream)) return rc;
    // all samples of this call stay in grid_samples when they fit one chunk: an adjoint with the
    // same arguments may take the paths' radiance from there
    auto& f = s->grid_forward;
    f.valid = chunk >= spp;
    f.rho = rho;
    f.prop = prop;
    f.alb = alb;
    f.cam = *cam;
    f.medium = *m;
    f.w = w;
    f.h = h;
    f.frame0 = frame0;
    f.spp = spp;
    f.max_depth = max_depth;
    return CUPBR_OK;
}

int grid_adjoint_device(const cupbr_scene* cs, const cupbr_camera* cam, const cupbr_grid_medium* m, const float* rho,
                        const float* prop, const float* alb, uint32_t w, uint32_t h, uint32_t frame0, uint32_t spp,
                        uint32_t max_depth, const float* d_image, float* grad_rho, float* grad_alb,
                        hipStream_t stream, bool from_forward) {
    cupbr_scene* s = const_cast<cupbr_scene*>(cs);
    const float* fwd_samples = nullptr;
    if (from_forward) {
        const auto& f = s->grid_forward;
        const bool same = f.valid && f.rho == rho && f.prop == prop && f.alb == alb && f.w == w && f.h == h &&
                          f.frame0 == frame0 && f.spp == spp && f.max_depth == max_depth &&
                          std::memcmp(&f.cam, cam, sizeof(cupbr_camera)) == 0 &&
                          std::memcmp(&f.medium, m, sizeof(cupbr_grid_medium)) == 0;
        if (!same)
            return set_error(CUPBR_ERR_INVALID_ARGUMENT,
                             "no preceding grid render with these arguments whose samples are kept (the last render "
                             "had other arguments, or more than 2^24 samples)");
        fwd_samples = s->grid_samples.as<float>();
    }
    if (int rc = scene_upload(s)) return rc;
    const GridDev G = make_grid(s, m, rho, prop, alb);
    unsigned long long* counts = s->grid_counting ? s->grid_counts.as<unsigned long long>() : nullptr;
    int blocks;
    uint32_t* spill;
    if (grid_launch_shape(s, (size_t)w * h * spp, &blocks, &spill)) return CUPBR_ERR_OUT_OF_MEMORY;
    if (int rc = scene_stream_acquire(s, stream)) return rc;
    // private gradient copies (cleared by a kernel, not a memset, so that a captured graph replays it)
    const size_t n_vox = (size_t)m->nx * m->ny * m->nz;
    const size_t copy_floats = (n_vox + 3) & ~(size_t)3;  // 16-byte aligned copies
    const bool private_copies = kGradCopies > 1 && copy_floats * kGradCopies * sizeof(float) <= kMaxGradCopyBytes;
    float* target = grad_rho;
    if (private_copies) {
        if (s->grid_grad.ensure(copy_floats * kGradCopies * sizeof(float))) return CUPBR_ERR_OUT_OF_MEMORY;
        target = s->grid_grad.as<float>();
        const size_t n4 = copy_floats * kGradCopies / 4;
        hipLaunchKernelGGL(k_grid_zero, dim3((unsigned)std::min<size_t>((n4 + 255) / 256, 4096)), dim3(256), 0, stream,
                           reinterpret_cast<float4*>(target), n4);
        CUPBR_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_grid<true>, dim3(blocks), dim3(kGridBlock), 0, stream, s->dev, G, *cam, w, h, spp, frame0, 0u,
                       spp, max_depth, nullptr, d_image, target, grad_alb,
                       (uint32_t)(3 * (s->objects.size() + s->triangles.size() / 9)),
                       private_copies ? copy_floats : (size_t)0, spill, counts, fwd_samples);
    CUPBR_HIP_CHECK(hipGetLastError());
    if (private_copies) {
        hipLaunchKernelGGL(k_grid_reduce, dim3((unsigned)std::min<size_t>((n_vox + 255) / 256, 4096)), dim3(256), 0,
                           stream, grad_rho, target, n_vox, copy_floats);
        CUPBR_HIP_CHECK(hipGetLastError());
    }
    return scene_stream_release(s, stream);
}

int grid_validate(const cupbr_scene* s, const cupbr_camera* cam, const cupbr_grid_medium* m, uint32_t w, uint32_t h,
                  uint32_t spp, uint32_t max_depth) {
    if (!s || !cam) return set_error(CUPBR_ERR_INVALID_ARGUMENT, "null scene or camera");
    if (int rc = check_medium(m)) return rc;
    if (!w || !h || (uint64_t)w * h > (1ull << 26)) return set_error(CUPBR_ERR_INVALID_ARGUMENT, "invalid image size");
    if (!spp) return set_error(CUPBR_ERR_INVALID_ARGUMENT, "spp must be >= 1");
    if (!max_depth || max_depth > 64) return set_error(CUPBR_ERR_INVALID_ARGUMENT, "max_depth must be in [1, 64]");
    return CUPBR_OK;
}

}  // namespace cupbr
