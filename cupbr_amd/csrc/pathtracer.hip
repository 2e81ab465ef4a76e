// PT+NEE in wavefront form (pathtracer_kernel_nee, lib/src/Renderer/PathTracer.cu:97-169).
//
// Per batch of paths (one path = one pixel of one reference frame):
//   raygen      camera rays into queue 0                       (PathTracer.cu:112-117)
//   per bounce: extend   closest hit for every queued ray      (:126)
//               shade    emission, NEE sample -> shadow queue,  (:127-158)
//                        Russian roulette, BSDF sample -> next queue
//               shadow   any-hit visibility; adds the NEE term  (:64-73)
//   accumulate  the reference running mean, frame by frame      (:162-168)
// Queue lengths live on the device, so no host round trip happens between
// bounces. extend is a persistent grid whose waves take 64-ray chunks from a
// work counter (traversal cost varies per ray); shade and shadow walk their
// queues with a static grid stride, and shade appends to the next queues with one
// atomic per 256-entry block (ballot + LDS prefix).
//
// Differences from the reference that cannot change the image: a path whose
// throughput became exactly zero (back-facing hit, :36-41) stops instead of
// bouncing on with zero weight, and a shadow ray whose contribution is exactly
// zero is not traced.
#include "renderer.h"
#include "shading.h"
#include "traverse.h"

namespace cupbr {

constexpr int kPtBlock = 256;
constexpr int kPtWaves = kPtBlock / 64;
#ifndef CUPBR_PT_LDS
#define CUPBR_PT_LDS 32
#endif
#ifndef CUPBR_PT_KEYS
#define CUPBR_PT_KEYS 0
#endif
#ifndef CUPBR_PT_FUSED
#define CUPBR_PT_FUSED 0
#endif
// fused: one persistent trace kernel per bounce takes the bounce's extension rays and the
// previous bounce's shadow rays from one work counter (64-ray chunks, extension first)
constexpr bool kPtFused = CUPBR_PT_FUSED != 0;
// traversal stack entries per lane kept in LDS; with kPtKeys the extension kernel also keeps each
// entry's entry distance (a second LDS array of the same size) to skip entries behind the best hit
constexpr int kPtLds = CUPBR_PT_LDS;
constexpr bool kPtKeys = CUPBR_PT_KEYS != 0;

__device__ __forceinline__ uint32_t wave_fetch(uint32_t* counter) {
    uint32_t base = 0;
    if (__lane_id() == 0) base = atomicAdd(counter, 64u);
    return __shfl(base, 0);
}

// Block-wide append: every thread of the block calls it; lanes with `want` get consecutive
// slots of the queue whose length is *count (one atomic per block).
__device__ __forceinline__ uint32_t block_append(bool want, uint32_t* count, uint32_t* s) {
    const unsigned long long m = __ballot(want);
    const uint32_t wave = threadIdx.x >> 6, lane = __lane_id();
    if (lane == 0) s[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < kPtWaves; ++w) {
            const uint32_t c = s[w];
            s[w] = tot;
            tot += c;
        }
        s[kPtWaves] = tot ? atomicAdd(count, tot) : 0u;
    }
    __syncthreads();
    const uint32_t slot = s[kPtWaves] + s[wave] + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    __syncthreads();
    return slot;
}

// adds a wave's traversal counts to two device counters
__device__ __forceinline__ void wave_add_counts(const TravCount& tc, unsigned long long* nodes,
                                                unsigned long long* prims) {
    unsigned long long a = tc.nodes, b = tc.prims;
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    if (__lane_id() == 0) {
        atomicAdd(nodes, a);
        atomicAdd(prims, b);
    }
}

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// radiance-cache training paths: pixel/frame pairs whose hash falls below the threshold
__device__ __forceinline__ bool is_training_path(uint32_t pix, uint32_t frame, uint32_t thresh) {
    return thresh != 0u && hash32(pix * 0x9E3779B1u ^ hash32(frame + 0x632BE5ABu)) < thresh;
}

// rad[p].w carries the path's flags: bit 0 = radiance-cache training path
__global__ __launch_bounds__(kPtBlock) void k_pt_raygen(cupbr_camera cam, uint32_t w, uint32_t h, uint32_t frame0,
                                                        uint32_t npath, float4* __restrict__ q_o,
                                                        float4* __restrict__ q_d, float4* __restrict__ weight,
                                                        float4* __restrict__ rad, Counters* c, uint32_t train_thresh) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) {
        c->queue_count[0].v = npath;
        c->work_extend.v = 0;
        c->shadow_count[0].v = 0;
        c->shadow_count[1].v = 0;
        c->query_count.v = 0;
        c->row_count.v = 0;
    }
    if (p >= npath) return;
    const uint32_t npix = w * h;
    const uint32_t s = p / npix, pix = p - s * npix;
    uint32_t seed = tea4(pix, frame0 + s);  // PathTracer.cu:112
    v3 o, d;
    launch_ray(pix % w, pix / w, w, h, cam, seed, o, d);
    q_o[p] = make_float4(o.x, o.y, o.z, 0.0f);
    q_d[p] = make_float4(d.x, d.y, d.z, __uint_as_float(p));
    weight[p] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(seed));
    const uint32_t flags = is_training_path(pix, frame0 + s, train_thresh) ? 1u : 0u;
    rad[p] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(flags));
}

template <bool kCount>
__global__ __launch_bounds__(kPtBlock) void k_pt_extend(DevScene S, int cur, const float4* __restrict__ q_o,
                                                        const float4* __restrict__ q_d, float2* __restrict__ q_hit,
                                                        Counters* c, uint32_t* spill, uint32_t spill_stride,
                                                        uint32_t spill_entries) {
    __shared__ uint32_t lds[(kPtKeys ? 2 : 1) * kPtLds * kPtBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = c->queue_count[cur].v;
    if (gtid == 0) {
        c->queue_count[cur ^ 1].v = 0;
        c->shadow_count[cur].v = 0;
        c->extension_rays += n;
    }
    LaneStack<kPtLds> st{lds + threadIdx.x, kPtBlock, spill ? spill + gtid : nullptr, spill_stride};
    // keys: second LDS half; their spill region follows the entries' (spill_entries * spill_stride)
    LaneStack<kPtLds> ks{lds + kPtLds * kPtBlock + threadIdx.x, kPtBlock,
                         spill ? spill + (size_t)spill_entries * spill_stride + gtid : nullptr, spill_stride};
    TravCount tc;
    for (;;) {
        const uint32_t base = wave_fetch(&c->work_extend.v);
        if (base >= n) break;
        const uint32_t i = base + __lane_id();
        if (i < n) {
            const float4 o4 = q_o[i], d4 = q_d[i];
            TraceResult r;
            traverse<false, kPtLds, kCount, kPtKeys>(S, mk(o4.x, o4.y, o4.z), mk(d4.x, d4.y, d4.z), 0.0f, st, r,
                                                     &tc, &ks);
            q_hit[i] = make_float2(r.t, __uint_as_float(r.prim));
        }
    }
    if (kCount) wave_add_counts(tc, &c->ext_nodes, &c->ext_prims);
}

#ifndef CUPBR_PT_REFILL
#define CUPBR_PT_REFILL 0
#endif
// refill: the extension kernel keeps every lane busy: a lane whose ray finished takes a new one
// as soon as kRefillMin lanes of its wave are idle (one atomic per refill), instead of idling
// until the wave's longest ray of a 64-ray chunk is done
constexpr bool kPtRefill = CUPBR_PT_REFILL != 0;
constexpr int kRefillMin = 16;

#define CUPBR_CSWAP(a, b)                                             \
    if (key[b] < key[a]) {                                             \
        float tk = key[a]; key[a] = key[b]; key[b] = tk;               \
        uint32_t tv = val[a]; val[a] = val[b]; val[b] = tv;            \
    }
// closest hit, one node or leaf per step, with per-lane ray refill (same result as traverse<false>)
template <bool kCount>
__global__ __launch_bounds__(kPtBlock) void k_pt_extend_refill(DevScene S, int cur_q, const float4* __restrict__ q_o,
                                                               const float4* __restrict__ q_d,
                                                               float2* __restrict__ q_hit, Counters* c,
                                                               uint32_t* spill, uint32_t spill_stride) {
    __shared__ uint32_t lds[kPtLds * kPtBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = c->queue_count[cur_q].v;
    if (gtid == 0) {
        c->queue_count[cur_q ^ 1].v = 0;
        c->shadow_count[cur_q].v = 0;
        c->extension_rays += n;
    }
    LaneStack<kPtLds> st{lds + threadIdx.x, kPtBlock, spill ? spill + gtid : nullptr, spill_stride};
    TravCount tc;
    const float inf = __builtin_inff();
    const uint32_t lane = __lane_id();
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    constexpr uint32_t kNone = 0xFFFFFFFFu;
    uint32_t ray = kNone, cur = 0;
    int sp = 0;
    v3 o = splat(0.0f), d = splat(0.0f);
    float ix = 0.0f, iy = 0.0f, iz = 0.0f;
    bool sx = true, sy = true, sz = true;
    TraceResult best{inf, kNoHit, 0xFFFFFFFFu};
    bool exhausted = false;  // wave-uniform: the queue has no rays left to hand out
    for (;;) {
        const unsigned long long idle = __ballot(ray == kNone);
        const uint32_t n_idle = (uint32_t)__popcll(idle);
        if (!exhausted && (n_idle >= (uint32_t)kRefillMin || n_idle == 64u)) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(&c->work_extend.v, n_idle);
            base = __shfl(base, 0);
            if (base + n_idle >= n) exhausted = true;
            if (ray == kNone) {
                const uint32_t my = base + (uint32_t)__popcll(idle & lt_mask);
                if (my < n) {
                    ray = my;
                    const float4 o4 = q_o[my], d4 = q_d[my];
                    o = mk(o4.x, o4.y, o4.z);
                    d = mk(d4.x, d4.y, d4.z);
                    best = TraceResult{inf, kNoHit, 0xFFFFFFFFu};
                    for (uint32_t k = 0; k < S.n_unb; ++k) {  // planes and quads first, as traverse()
                        const UnbRec u = S.unb[k];
                        float t;
                        if (!plane_quad_test(u, o, d, t)) continue;
                        const uint32_t rank = __float_as_uint(u.n.w);
                        if (t < best.t || (t == best.t && rank < best.rank)) {
                            best.t = t;
                            best.prim = kUnbBit | k;
                            best.rank = rank;
                        }
                    }
                    ix = 1.0f / d.x;
                    iy = 1.0f / d.y;
                    iz = 1.0f / d.z;
                    sx = ix >= 0.0f;
                    sy = iy >= 0.0f;
                    sz = iz >= 0.0f;
                    sp = 0;
                    cur = 0;
                    if (S.n_nodes == 0) {  // nothing but planes and quads
                        q_hit[ray] = make_float2(best.t, __uint_as_float(best.prim));
                        ray = kNone;
                    }
                }
            }
        }
        if (__ballot(ray != kNone) == 0ull) {
            if (exhausted) break;
            continue;
        }
        if (ray == kNone) continue;
        bool pop = true;
        if (cur & kLeafBit) {
            const uint32_t first = leaf_first(cur), n_leaf = leaf_count(cur);
            if (kCount) tc.prims += n_leaf;
            for (uint32_t k = 0; k < n_leaf; ++k) {
                const PrimRec pr = S.prims[first + k];
                const uint32_t flags = __float_as_uint(pr.c.w);
                float t;
                bool h;
                if ((flags & 3u) == kPrimTriangle)
                    h = mt_test(mk(pr.a.x, pr.a.y, pr.a.z), mk(pr.a.w, pr.b.x, pr.b.y), mk(pr.b.z, pr.b.w, pr.c.x), o, d, t);
                else
                    h = sphere_test(mk(pr.a.x, pr.a.y, pr.a.z), pr.a.w, o, d, t);
                if (!h) continue;
                const uint32_t rank = __float_as_uint(pr.c.y);
                if (t < best.t || (t == best.t && rank < best.rank)) {
                    if (flags & kPrimNeedsBoxCheck) {
                        const ObjInfo& ob = S.objs[__float_as_uint(pr.c.z)];
                        if (!ref_aabb_hit(ob.bmin, ob.bmax, o, d)) continue;
                    }
                    best.t = t;
                    best.prim = first + k;
                    best.rank = rank;
                }
            }
        } else {
            if (kCount) tc.nodes += 1;
            const float4* nf = reinterpret_cast<const float4*>(S.nodes + cur);
            const float4 lx = nf[0], ly = nf[1], lz = nf[2], hx = nf[3], hy = nf[4], hz = nf[5];
            const uint4 ch = *reinterpret_cast<const uint4*>(nf + 6);
            const float4 nx = sx ? lx : hx, fx = sx ? hx : lx;
            const float4 ny = sy ? ly : hy, fy = sy ? hy : ly;
            const float4 nz = sz ? lz : hz, fz = sz ? hz : lz;
            const float lim = best.t * 1.0001f + 1e-6f;
            const float tlo = -1e-5f;
            float key[4];
            uint32_t val[4] = {ch.x, ch.y, ch.z, ch.w};
#define CUPBR_CHILD(i, cc)                                                                              \
    {                                                                                                   \
        float tn = fmaxf(fmaxf((nx.cc - o.x) * ix, (ny.cc - o.y) * iy), fmaxf((nz.cc - o.z) * iz, tlo)); \
        float tf = fminf(fminf((fx.cc - o.x) * ix, (fy.cc - o.y) * iy), fminf((fz.cc - o.z) * iz, lim)); \
        bool ok = val[i] != kChildEmpty && tn <= tf * 1.000002f;                                       \
        key[i] = ok ? tn : inf;                                                                         \
    }
            CUPBR_CHILD(0, x)
            CUPBR_CHILD(1, y)
            CUPBR_CHILD(2, z)
            CUPBR_CHILD(3, w)
#undef CUPBR_CHILD
            CUPBR_CSWAP(0, 1)
            CUPBR_CSWAP(2, 3)
            CUPBR_CSWAP(0, 2)
            CUPBR_CSWAP(1, 3)
            CUPBR_CSWAP(1, 2)
            if (key[0] < inf) {
                if (key[3] < inf) st.put(sp++, val[3]);
                if (key[2] < inf) st.put(sp++, val[2]);
                if (key[1] < inf) st.put(sp++, val[1]);
                cur = val[0];
                pop = false;
            }
        }
        if (pop) {
            if (sp == 0) {
                q_hit[ray] = make_float2(best.t, __uint_as_float(best.prim));
                ray = kNone;
            } else {
                cur = st.get(--sp);
            }
        }
    }
    if (kCount) wave_add_counts(tc, &c->ext_nodes, &c->ext_prims);
}
#undef CUPBR_CSWAP

// geometry of a closest-hit record: world normal and object id (LocalGeometry, Geometry.h:28-36)
__device__ __forceinline__ void hit_geometry(const DevScene& S, uint32_t prim, v3 P, v3& N, uint32_t& obj) {
    if (prim & kUnbBit) {
        const UnbRec& u = S.unb[prim & ~kUnbBit];
        N = mk(u.n.x, u.n.y, u.n.z);
        obj = __float_as_uint(u.e1.w);
        return;
    }
    const float4 c = S.prims[prim].c;
    obj = __float_as_uint(c.z);
    if ((__float_as_uint(c.w) & 3u) == kPrimSphere) {
        const float4 a = S.prims[prim].a;
        N = normalize(P - mk(a.x, a.y, a.z));  // SphereDetail.h:66
    } else {
        const float4 n = S.normals[prim];
        N = mk(n.x, n.y, n.z);
    }
}

// radiance-cache records written by shade: a query (inputs 16, throughput 3, path) ends a rendering
// path at its first diffuse/glossy vertex after the camera hit; a training path leaves a pending
// row (inputs 16, radiance so far 3, throughput 3, path, pad) at each such vertex
struct NrcShade {
    float4* queries;  // 5 float4 per query
    float4* pending;  // 6 float4 per row
    uint32_t max_rows;
};

template <bool kNrc>
__global__ __launch_bounds__(kPtBlock) void k_pt_shade(DevScene S, int cur, uint32_t bounce, uint32_t max_depth, int rr,
                                                       const float4* __restrict__ q_o, const float4* __restrict__ q_d,
                                                       const float2* __restrict__ q_hit, float4* __restrict__ nq_o,
                                                       float4* __restrict__ nq_d, float4* __restrict__ s_o,
                                                       float4* __restrict__ s_d, float4* __restrict__ s_c,
                                                       float4* __restrict__ weight, float4* __restrict__ rad,
                                                       Counters* c, NrcShade nrc) {
    __shared__ uint32_t s_append[kPtWaves + 1];
    const uint32_t n = c->queue_count[cur].v;
    if (blockIdx.x == 0 && threadIdx.x == 0) c->work_extend.v = 0;  // extend of the next bounce starts at 0
    const float nl = (float)(S.n_lights + (S.use_env ? 1u : 0u));
    for (uint32_t base = blockIdx.x * kPtBlock; base < n; base += gridDim.x * kPtBlock) {
        const uint32_t i = base + threadIdx.x;
        bool want_next = false, want_shadow = false, want_query = false, want_row = false;
        float4 no = {}, nd = {}, so = {}, sd = {}, sc = {};
        float4 rec[6] = {};  // cache query or pending row
        if (i < n) {
            const float4 o4 = q_o[i], d4 = q_d[i];
            const float2 hh = q_hit[i];
            const uint32_t p = __float_as_uint(d4.w);
            const v3 o = mk(o4.x, o4.y, o4.z), d = mk(d4.x, d4.y, d4.z);
            const float4 W = weight[p];
            const float4 R = rad[p];
            v3 w3 = mk(W.x, W.y, W.z);
            uint32_t seed = __float_as_uint(W.w);
            v3 L = mk(R.x, R.y, R.z);
            const uint32_t prim = __float_as_uint(hh.y);
            if (prim == kNoHit) {  // :127-135
                if (S.use_env) L += w3 * env_lookup(S, d);
            } else {
                const v3 P = o + hh.x * d;  // geom.P = origin + t * direction
                v3 N;
                uint32_t obj;
                hit_geometry(S, prim, P, N, obj);
                const Mat m = load_mat(S.mats + obj);
                const v3 inc = -1.0f * d;
                L += w3 * m.albedo_e;  // emissiveIllumintation :20-26
                bool alive = true;
                bool cached = false;
                if (kNrc && bounce >= 1 && (m.type == CUPBR_LAMBERT || m.type == CUPBR_PHONG || m.type == CUPBR_GGX) &&
                    !(m.type != CUPBR_GLASS && dot(N, inc) <= 0.0f)) {
                    rec[0] = make_float4(P.x, P.y, P.z, inc.x);
                    rec[1] = make_float4(inc.y, inc.z, N.x, N.y);
                    rec[2] = make_float4(N.z, m.roughness, m.albedo_d.x, m.albedo_d.y);
                    rec[3] = make_float4(m.albedo_d.z, m.albedo_s.x, m.albedo_s.y, m.albedo_s.z);
                    if (__float_as_uint(R.w) & 1u) {  // training path: pending row, the path goes on
                        want_row = true;
                        rec[4] = make_float4(L.x, L.y, L.z, w3.x);
                        rec[5] = make_float4(w3.y, w3.z, __uint_as_float(p), 0.0f);
                    } else {  // rendering path: ends here with throughput * cache
                        want_query = true;
                        cached = true;
                        rec[4] = make_float4(w3.x, w3.y, w3.z, __uint_as_float(p));
                    }
                }
                if (cached) {
                    alive = false;
                } else if (m.type != CUPBR_GLASS && dot(N, inc) <= 0.0f) {  // :36-41
                    w3 = splat(0.0f);
                    alive = false;
                } else {
                    // directIllumination :43-73
                    v3 ldir, le;
                    float dist = 0.0f;
                    if (nee_sample(S, m, seed, P, inc, N, ldir, le, dist)) {
                        const v3 contrib = (nl * fmaxf(0.0f, dot(N, ldir))) * brdf(m, inc, ldir, N) * le * w3;
                        if (!(contrib.x == 0.0f && contrib.y == 0.0f && contrib.z == 0.0f)) {
                            const v3 sorg = P + 0.001f * ldir;
                            const v3 sdir = normalize(ldir);
                            want_shadow = true;
                            so = make_float4(sorg.x, sorg.y, sorg.z, dist);
                            sd = make_float4(sdir.x, sdir.y, sdir.z, __uint_as_float(p));
                            sc = make_float4(contrib.x, contrib.y, contrib.z, 0.0f);
                        }
                    }
                    if (rr) {  // :141-152
                        const float alpha = clampf(fmaxf(w3.x, fmaxf(w3.y, w3.z)), 0.0f, 1.0f);
                        if (rnd(seed) > alpha || safe_float_equal(alpha, 0.0f)) {
                            alive = false;
                            w3 = splat(0.0f);
                        } else {
                            w3 = w3 / alpha;
                        }
                    }
                    if (alive) {  // indirectIllumination :76-95
                        const DirPdf sp = sample_direction(m, seed, inc, N);
                        if (norm(sp.d) < kEpsilon) {
                            alive = false;
                        } else {
                            w3 = w3 * fabsf(dot(sp.d, N)) * brdf(m, inc, sp.d, N) / sp.pdf;
                            if (bounce + 1 < max_depth) {  // :156-160
                                const v3 norg = P + 0.01f * sp.d;
                                const v3 ndir = normalize(sp.d);
                                want_next = true;
                                no = make_float4(norg.x, norg.y, norg.z, 0.0f);
                                nd = make_float4(ndir.x, ndir.y, ndir.z, __uint_as_float(p));
                            }
                        }
                    }
                }
                weight[p] = make_float4(w3.x, w3.y, w3.z, __uint_as_float(seed));
            }
            rad[p] = make_float4(L.x, L.y, L.z, R.w);
        }
        const uint32_t ns = block_append(want_next, &c->queue_count[cur ^ 1].v, s_append);
        if (want_next) {
            nq_o[ns] = no;
            nq_d[ns] = nd;
        }
        const uint32_t ss = block_append(want_shadow, &c->shadow_count[cur].v, s_append);
        if (want_shadow) {
            s_o[ss] = so;
            s_d[ss] = sd;
            s_c[ss] = sc;
        }
        if (kNrc) {
            const uint32_t qs = block_append(want_query, &c->query_count.v, s_append);
            if (want_query)
                for (int k = 0; k < 5; ++k) nrc.queries[(size_t)qs * 5 + k] = rec[k];
            const uint32_t rs = block_append(want_row, &c->row_count.v, s_append);
            if (want_row && rs < nrc.max_rows)
                for (int k = 0; k < 6; ++k) nrc.pending[(size_t)rs * 6 + k] = rec[k];
        }
    }
}

// training rows from the pending rows of finished training paths: target = (final radiance -
// radiance before the vertex's reflection) / throughput at the vertex (0 where the throughput is 0)
__global__ __launch_bounds__(kPtBlock) void k_nrc_rows(const float4* __restrict__ pending, const Counters* c,
                                                       uint32_t max_rows, const float4* __restrict__ rad,
                                                       float4* __restrict__ rows) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = min(c->row_count.v, max_rows);
    if (i >= n) return;
    const float4* r = pending + (size_t)i * 6;
    const float4 a = r[4], b = r[5];
    const uint32_t p = __float_as_uint(b.z);
    const float4 R = rad[p];
    const float beta[3] = {a.w, b.x, b.y}, snap[3] = {a.x, a.y, a.z}, fin[3] = {R.x, R.y, R.z};
    float tgt[3];
    for (int k = 0; k < 3; ++k) tgt[k] = beta[k] > 0.0f ? (fin[k] - snap[k]) / beta[k] : 0.0f;
    float4* o = rows + (size_t)i * 5;
    for (int k = 0; k < 4; ++k) o[k] = r[k];
    o[4] = make_float4(tgt[0], tgt[1], tgt[2], 0.0f);
}

// shadow ray i: visible -> radiance of its path += contribution (traceVisibility, :66-73)
__device__ __forceinline__ void shadow_result(bool occluded, const float4 d4, const float4* __restrict__ s_c,
                                              uint32_t i, float4* __restrict__ rad) {
    if (occluded) return;
    const uint32_t p = __float_as_uint(d4.w);
    const float4 cc = s_c[i];
    const float4 R = rad[p];
    v3 L = mk(R.x, R.y, R.z);
    L += mk(cc.x, cc.y, cc.z);
    rad[p] = make_float4(L.x, L.y, L.z, R.w);
}

// one bounce's extension rays (queue cur) and the previous bounce's shadow rays (parity cur ^ 1)
// under one work counter: chunk k < ceil(n_ext / 64) is extension rays [64k, 64k + 64), the rest
// shadow rays; every wave's chunk is of one kind
template <bool kCount>
__global__ __launch_bounds__(kPtBlock) void k_pt_trace(DevScene S, int cur, const float4* __restrict__ q_o,
                                                       const float4* __restrict__ q_d, float2* __restrict__ q_hit,
                                                       const float4* __restrict__ s_o, const float4* __restrict__ s_d,
                                                       const float4* __restrict__ s_c, float4* __restrict__ rad,
                                                       Counters* c, uint32_t* spill, uint32_t spill_stride,
                                                       uint32_t spill_entries) {
    __shared__ uint32_t lds[(kPtKeys ? 2 : 1) * kPtLds * kPtBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n_ext = c->queue_count[cur].v;
    const uint32_t n_sh = c->shadow_count[cur ^ 1].v;
    if (gtid == 0) {
        c->queue_count[cur ^ 1].v = 0;
        c->shadow_count[cur].v = 0;
        c->extension_rays += n_ext;
        c->shadow_rays += n_sh;
    }
    LaneStack<kPtLds> st{lds + threadIdx.x, kPtBlock, spill ? spill + gtid : nullptr, spill_stride};
    LaneStack<kPtLds> ks{lds + kPtLds * kPtBlock + threadIdx.x, kPtBlock,
                         spill ? spill + (size_t)spill_entries * spill_stride + gtid : nullptr, spill_stride};
    TravCount te, ts;
    const uint32_t ext_span = (n_ext + 63u) & ~63u;
    const uint32_t total = ext_span + n_sh;
    for (;;) {
        const uint32_t base = wave_fetch(&c->work_extend.v);
        if (base >= total) break;
        const uint32_t i = base + __lane_id();
        if (base < ext_span) {
            if (i < n_ext) {
                const float4 o4 = q_o[i], d4 = q_d[i];
                TraceResult r;
                traverse<false, kPtLds, kCount, kPtKeys>(S, mk(o4.x, o4.y, o4.z), mk(d4.x, d4.y, d4.z), 0.0f, st,
                                                         r, &te, &ks);
                q_hit[i] = make_float2(r.t, __uint_as_float(r.prim));
            }
        } else {
            const uint32_t j = i - ext_span;
            if (j < n_sh) {
                const float4 o4 = s_o[j], d4 = s_d[j];
                TraceResult r;
                const bool occluded =
                    traverse<true, kPtLds, kCount>(S, mk(o4.x, o4.y, o4.z), mk(d4.x, d4.y, d4.z), o4.w, st, r, &ts);
                shadow_result(occluded, d4, s_c, j, rad);
            }
        }
    }
    if (kCount) {
        wave_add_counts(te, &c->ext_nodes, &c->ext_prims);
        wave_add_counts(ts, &c->shadow_nodes, &c->shadow_prims);
    }
}

template <bool kCount>
__global__ __launch_bounds__(kPtBlock) void k_pt_shadow(DevScene S, int par, const float4* __restrict__ s_o,
                                                        const float4* __restrict__ s_d, const float4* __restrict__ s_c,
                                                        float4* __restrict__ rad, Counters* c, uint32_t* spill,
                                                        uint32_t spill_stride) {
    __shared__ uint32_t lds[kPtLds * kPtBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = c->shadow_count[par].v;
    if (gtid == 0) c->shadow_rays += n;
    LaneStack<kPtLds> st{lds + threadIdx.x, kPtBlock, spill ? spill + gtid : nullptr, spill_stride};
    TravCount tc;
    const uint32_t stride = gridDim.x * kPtBlock;
    for (uint32_t i = gtid; i < n; i += stride) {
        const float4 o4 = s_o[i], d4 = s_d[i];
        TraceResult r;
        const bool occluded =
            traverse<true, kPtLds, kCount>(S, mk(o4.x, o4.y, o4.z), mk(d4.x, d4.y, d4.z), o4.w, st, r, &tc);
        shadow_result(occluded, d4, s_c, i, rad);
    }
    if (kCount) wave_add_counts(tc, &c->shadow_nodes, &c->shadow_prims);
}

// running mean of PathTracer.cu:162-168, applied frame by frame in order
__global__ __launch_bounds__(kPtBlock) void k_accumulate(float* __restrict__ img, const float4* __restrict__ rad,
                                                         uint32_t npix, uint32_t frame0, uint32_t nspp) {
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    v3 acc = ld3(img + 3 * (size_t)pix);
    for (uint32_t s = 0; s < nspp; ++s) {
        const float4 R = rad[(size_t)s * npix + pix];
        const v3 L = mk(R.x, R.y, R.z);
        const uint32_t f = frame0 + s;
        if (f > 0) {
            const float a = 1.0f / ((float)f + 1.0f);
            acc = (1.0f - a) * acc + a * L;
        } else {
            acc = L;
        }
    }
    st3(img + 3 * (size_t)pix, acc);
}

static int ensure_wavefront(cupbr_renderer* r, uint32_t paths, uint32_t grid_threads) {
    Wavefront& wf = r->wf;
    int rc;
    if (paths > wf.capacity) {
        const size_t n = paths;
        if ((rc = wf.weight.ensure(n * 16))) return rc;
        if ((rc = wf.radiance.ensure(n * 16))) return rc;
        for (int k = 0; k < 2; ++k) {
            if ((rc = wf.q_o[k].ensure(n * 16))) return rc;
            if ((rc = wf.q_d[k].ensure(n * 16))) return rc;
        }
        if ((rc = wf.q_hit.ensure(n * 8))) return rc;
        for (int k = 0; k < (kPtFused ? 2 : 1); ++k) {
            if ((rc = wf.s_o[k].ensure(n * 16))) return rc;
            if ((rc = wf.s_d[k].ensure(n * 16))) return rc;
            if ((rc = wf.s_c[k].ensure(n * 16))) return rc;
        }
        wf.capacity = paths;
    }
    if ((rc = wf.counters.ensure(sizeof(Counters)))) return rc;
    const uint32_t entries = r->scene->stack_entries;
    if (entries > (uint32_t)kPtLds) {
        if ((rc = wf.spill.ensure((size_t)(kPtKeys ? 2 : 1) * (entries - kPtLds) * grid_threads * 4))) return rc;
    }
    return CUPBR_OK;
}

int render_pathtracer(cupbr_renderer* r, const cupbr_camera& cam, uint32_t frame0, uint32_t spp) {
    const uint32_t w = r->width, h = r->height, npix = w * h;
    if (npix == 0 || spp == 0) return CUPBR_OK;
    const bool nrc = r->nrc != nullptr;
    // paths in flight: enough to fill the GPU at the deepest bounce, bounded to ~8M; with a
    // radiance cache one frame per batch (the cache trains between frames)
    const uint32_t target = 1u << 22, cap = 1u << 23;
    uint32_t batch = std::max<uint32_t>(1, (target + npix - 1) / npix);
    batch = std::min(batch, std::max<uint32_t>(1, cap / npix));
    batch = std::min(batch, spp);
    if (nrc) batch = 1;
    const int blocks = device_cu_count() * 8;
    const uint32_t grid_threads = (uint32_t)blocks * kPtBlock;
    int rc = ensure_wavefront(r, batch * npix, grid_threads);
    if (rc) return rc;
    Wavefront& wf = r->wf;
    Counters* c = wf.counters.as<Counters>();
    CUPBR_HIP_CHECK(hipMemsetAsync(c, 0, sizeof(Counters), r->stream));
    uint32_t* spill = r->scene->stack_entries > (uint32_t)kPtLds ? wf.spill.as<uint32_t>() : nullptr;
    const DevScene& S = r->scene->dev;
    float* img = r->images[CUPBR_IMAGE_OUTPUT].as<float>();
    const uint32_t bounces = std::max<uint32_t>(1, r->max_trace_depth);  // do { } while (depth < max)
    auto extend = r->counting ? k_pt_extend<true> : k_pt_extend<false>;
    auto extend_rf = r->counting ? k_pt_extend_refill<true> : k_pt_extend_refill<false>;
    auto shadow = r->counting ? k_pt_shadow<true> : k_pt_shadow<false>;
    auto trace = r->counting ? k_pt_trace<true> : k_pt_trace<false>;
    const uint32_t spill_entries =
        r->scene->stack_entries > (uint32_t)kPtLds ? r->scene->stack_entries - kPtLds : 0u;
    auto shade = nrc ? k_pt_shade<true> : k_pt_shade<false>;
    NrcShade ns{nullptr, nullptr, 0};
    uint32_t train_thresh = 0;
    if (nrc) {
        if ((rc = r->nrc_queries.ensure((size_t)batch * npix * 80)) ||
            (rc = r->nrc_pending.ensure((size_t)r->nrc_max_rows * 96)) ||
            (rc = r->nrc_rows.ensure((size_t)r->nrc_max_rows * 80)))
            return rc;
        ns = NrcShade{r->nrc_queries.as<float4>(), r->nrc_pending.as<float4>(), r->nrc_max_rows};
        const double frac = std::min(1.0, (double)r->nrc_train_paths / (double)npix);
        train_thresh = frac >= 1.0 ? 0xFFFFFFFFu : (uint32_t)(frac * 4294967296.0);
    }
    if ((rc = stage_mark(r, -1))) return rc;
    for (uint32_t s0 = 0; s0 < spp; s0 += batch) {
        const uint32_t nb = std::min(batch, spp - s0);
        const uint32_t npath = nb * npix;
        hipLaunchKernelGGL(k_pt_raygen, dim3((npath + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0, r->stream, cam, w,
                           h, frame0 + s0, npath, wf.q_o[0].as<float4>(), wf.q_d[0].as<float4>(),
                           wf.weight.as<float4>(), wf.radiance.as<float4>(), c, train_thresh);
        if ((rc = stage_mark(r, 3))) return rc;
        for (uint32_t b = 0; b < bounces; ++b) {
            const int cur = (int)(b & 1);
            const int sp = kPtFused ? cur : 0;  // shadow queue shade(b) appends to
            if (kPtFused) {
                hipLaunchKernelGGL(trace, dim3(blocks), dim3(kPtBlock), 0, r->stream, S, cur, wf.q_o[cur].as<float4>(),
                                   wf.q_d[cur].as<float4>(), wf.q_hit.as<float2>(), wf.s_o[cur ^ 1].as<float4>(),
                                   wf.s_d[cur ^ 1].as<float4>(), wf.s_c[cur ^ 1].as<float4>(),
                                   wf.radiance.as<float4>(), c, spill, grid_threads, spill_entries);
            } else if (kPtRefill) {
                hipLaunchKernelGGL(extend_rf, dim3(blocks), dim3(kPtBlock), 0, r->stream, S, cur,
                                   wf.q_o[cur].as<float4>(), wf.q_d[cur].as<float4>(), wf.q_hit.as<float2>(), c, spill,
                                   grid_threads);
            } else {
                hipLaunchKernelGGL(extend, dim3(blocks), dim3(kPtBlock), 0, r->stream, S, cur, wf.q_o[cur].as<float4>(),
                                   wf.q_d[cur].as<float4>(), wf.q_hit.as<float2>(), c, spill, grid_threads,
                                   spill_entries);
            }
            if ((rc = stage_mark(r, 0))) return rc;
            hipLaunchKernelGGL(shade, dim3(blocks), dim3(kPtBlock), 0, r->stream, S, cur, b, r->max_trace_depth,
                               (int)r->russian_roulette, wf.q_o[cur].as<float4>(), wf.q_d[cur].as<float4>(),
                               wf.q_hit.as<float2>(), wf.q_o[cur ^ 1].as<float4>(), wf.q_d[cur ^ 1].as<float4>(),
                               wf.s_o[sp].as<float4>(), wf.s_d[sp].as<float4>(), wf.s_c[sp].as<float4>(),
                               wf.weight.as<float4>(), wf.radiance.as<float4>(), c, ns);
            if ((rc = stage_mark(r, 1))) return rc;
            if (!kPtFused) {
                hipLaunchKernelGGL(shadow, dim3(blocks), dim3(kPtBlock), 0, r->stream, S, cur, wf.s_o[0].as<float4>(),
                                   wf.s_d[0].as<float4>(), wf.s_c[0].as<float4>(), wf.radiance.as<float4>(), c, spill,
                                   grid_threads);
                if ((rc = stage_mark(r, 2))) return rc;
            }
        }
        if (kPtFused) {  // the last bounce's shadow rays (its extension queue is empty)
            const int cur = (int)(bounces & 1);
            hipLaunchKernelGGL(trace, dim3(blocks), dim3(kPtBlock), 0, r->stream, S, cur, wf.q_o[cur].as<float4>(),
                               wf.q_d[cur].as<float4>(), wf.q_hit.as<float2>(), wf.s_o[cur ^ 1].as<float4>(),
                               wf.s_d[cur ^ 1].as<float4>(), wf.s_c[cur ^ 1].as<float4>(), wf.radiance.as<float4>(),
                               c, spill, grid_threads, spill_entries);
            if ((rc = stage_mark(r, 0))) return rc;
        }
        if (nrc) {
            // throughput * cache for the paths that ended at a cache vertex, then the training rows
            if ((rc = nrc_infer(r->nrc, r->nrc_queries.as<float>(), npath, &c->query_count.v, 20,
                                wf.radiance.as<float>(), true, r->stream)))
                return rc;
            hipLaunchKernelGGL(k_nrc_rows, dim3((r->nrc_max_rows + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0,
                               r->stream, r->nrc_pending.as<float4>(), c, r->nrc_max_rows, wf.radiance.as<float4>(),
                               r->nrc_rows.as<float4>());
            if ((rc = stage_mark(r, 3))) return rc;
        }
        hipLaunchKernelGGL(k_accumulate, dim3((npix + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0, r->stream, img,
                           wf.radiance.as<float4>(), npix, frame0 + s0, nb);
        if ((rc = stage_mark(r, 3))) return rc;
        if (nrc) {
            uint32_t counts[2] = {0, 0};
            CUPBR_HIP_CHECK(hipMemcpyAsync(&counts[0], &c->query_count.v, 4, hipMemcpyDeviceToHost, r->stream));
            CUPBR_HIP_CHECK(hipMemcpyAsync(&counts[1], &c->row_count.v, 4, hipMemcpyDeviceToHost, r->stream));
            CUPBR_HIP_CHECK(hipStreamSynchronize(r->stream));
            const uint32_t n_rows = std::min(counts[1], r->nrc_max_rows);
            r->nrc_queries_total += counts[0];
            r->nrc_rows_total += n_rows;
            if (n_rows > 0 && r->nrc_steps > 0) {
                if ((rc = nrc_train(r->nrc, r->nrc_rows.as<float>(), n_rows, r->nrc_batch, r->nrc_steps, r->nrc_lr,
                                    nullptr, r->stream)))
                    return rc;
            }
            if ((rc = stage_mark(r, 3))) return rc;
        }
    }
    CUPBR_HIP_CHECK(hipGetLastError());
    return CUPBR_OK;
}

}  // namespace cupbr
