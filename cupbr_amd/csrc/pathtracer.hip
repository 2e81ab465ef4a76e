// PT+NEE in wavefront form (pathtracer_kernel_nee, lib/src/Renderer/PathTracer.cu:97-169).
//
// Per batch of paths (one path = one pixel of one reference frame):
//   raygen      camera rays into extension queue 0                    (PathTracer.cu:112-117)
//   per bounce: trace    closest hit for the bounce's extension rays   (:126), and any-hit
//                        visibility for the previous bounce's shadow rays, adding their
//                        NEE terms (:64-73), in one launch
//               shade    emission, NEE sample -> shadow queue,         (:127-158)
//                        Russian roulette, BSDF sample -> next extension queue
//   final trace the last bounce's shadow rays
//   accumulate  the reference running mean, frame by frame             (:162-168)
// Queue lengths live on the device, so no host round trip happens between bounces.
// trace is a persistent grid whose waves take 64-ray chunks from 8 work counters, one
// per XCD group (traversal cost varies per ray; one launch per bounce keeps the GPU full
// where the shadow queue alone is short); shade walks its queue with a static grid stride and
// appends to the next queues with one atomic per 256-entry block (ballot + LDS prefix).
// Extension and shadow queues alternate by bounce parity.
//
// Differences from the reference that cannot change the image: a path whose
// throughput became exactly zero (back-facing hit, :36-41) stops instead of
// bouncing on with zero weight, and a shadow ray whose contribution is exactly
// zero is not traced.
#include "renderer.h"
#include "shading.h"
#include "traverse.h"

// Lanes of the refill loop that must be at a leaf before a wave runs a leaf step (inner-node steps
// otherwise), at most: a wave runs its leaf step once half of its rays wait at a leaf, capped at
// this many lanes. 32 (always half) measured best with the half rule (1080p, 1M triangles: cap 12
// / 16 / 20 / 24 / 28 / 32 -> 1143 / 1155-1159 / 1163 / 1164-1168 / 1168 / 1168 Msamples/s). With
// a fixed count instead (an earlier build): 8 / 16 / 24 / 32 -> 901 / 932 / 920 / 843; leaf and
// inner steps taken together every iteration: 889.
#ifndef CUPBR_LEAF_BATCH
#define CUPBR_LEAF_BATCH 32
#endif
// idle lanes of a wave that trigger a refill from the ray queue
#ifndef CUPBR_REFILL_IDLE
#define CUPBR_REFILL_IDLE 32
#endif

namespace cupbr {

constexpr int kPtBlock = 256;
constexpr int kPtWaves = kPtBlock / 64;
// traversal stack entries per lane kept in LDS (16 KB per 256-thread block: the trace kernel's 61
// VGPRs, not LDS, then bound its occupancy at 8 waves per SIMD; deeper entries spill to memory)
constexpr int kPtLds = 16;

// one wave-uniform ticket per call
__device__ __forceinline__ uint32_t wave_fetch(uint32_t* counter) {
    uint32_t base = 0;
    if (__lane_id() == 0) base = atomicAdd(counter, 1u);
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
// a path's first generator state: the reference's tea<4>(pixel, frame) (PathTracer.cu:112), or for
// the PCG mode the PCG hash of the pixel keyed by the hash of the frame
template <class Seed> __device__ __forceinline__ Seed path_seed(uint32_t pix, uint32_t frame);
template <> __device__ __forceinline__ uint32_t path_seed<uint32_t>(uint32_t pix, uint32_t frame) {
    return tea4(pix, frame);
}
template <> __device__ __forceinline__ Pcg32 path_seed<Pcg32>(uint32_t pix, uint32_t frame) {
    return Pcg32{pcg_hash(pcg_hash(frame) ^ pix)};
}

// The camera ray of queue slot p of a batch in SampleOrder (PathTracer.cu:112-126): the pixel it
// samples, frame frame0 + its sample index, and the generator state after launchRay's two draws.
template <class Seed>
__device__ __forceinline__ void camera_path(const cupbr_camera& cam, uint32_t w, uint32_t h, uint32_t frame0,
                                            const SampleOrder& order, uint32_t p, Seed& seed, uint32_t& pix,
                                            uint32_t& frame, v3& o, v3& d) {
    uint32_t s, j, px, py;
    order.sample(p, s, j);
    tile_pixel(j, w, h, px, py);
    pix = py * w + px;
    frame = frame0 + s;
    seed = path_seed<Seed>(pix, frame);
    launch_ray(px, py, w, h, cam, seed, o, d);
}

// the per-batch counter reset that precedes bounce 0
__device__ __forceinline__ void reset_batch_counters(Counters* c, uint32_t npath) {
    c->queue_count[0].v = npath;
    for (int g = 0; g < 8; ++g) c->work[g].v = c->swork[g].v = 0;
    c->shadow_count[0].v = 0;
    c->shadow_count[1].v = 0;
    c->query_count.v = 0;
    c->row_count.v = 0;
    c->iqueue_count[0].v = 0;
    c->iqueue_count[1].v = 0;
}

// A camera path's state as bounce 0 leaves it: direction and path id (queue 0), throughput 1 and
// generator state, radiance 0 and flags (bit 0: radiance-cache training path).
template <class Seed>
__device__ __forceinline__ void store_camera_path(uint32_t p, v3 d, Seed seed, uint32_t pix, uint32_t frame,
                                                  uint32_t train_thresh, float4* __restrict__ q_d,
                                                  float4* __restrict__ weight, float4* __restrict__ rad) {
    q_d[p] = make_float4(d.x, d.y, d.z, __uint_as_float(p));
    weight[p] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(state_bits(seed)));
    const uint32_t flags = is_training_path(pix, frame, train_thresh) ? 1u : 0u;
    rad[p] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(flags));
}

// Camera rays of a batch in SampleOrder: the path id of a path is its queue slot here.
template <class Seed>
__global__ __launch_bounds__(kPtBlock) void k_pt_raygen(cupbr_camera cam, uint32_t w, uint32_t h, uint32_t frame0,
                                                        uint32_t npath, SampleOrder order, float4* __restrict__ q_o,
                                                        float4* __restrict__ q_d, float4* __restrict__ weight,
                                                        float4* __restrict__ rad, Counters* c, uint32_t train_thresh) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) reset_batch_counters(c, npath);
    if (p >= npath) return;  // p: queue slot and path id
    Seed seed;
    uint32_t pix, frame;
    v3 o, d;
    camera_path(cam, w, h, frame0, order, p, seed, pix, frame, o, d);
    q_o[p] = make_float4(o.x, o.y, o.z, 0.0f);
    store_camera_path(p, d, seed, pix, frame, train_thresh, q_d, weight, rad);
}

// radiance-cache records written by shade: a query (inputs 16, throughput 3, path) ends a rendering
// path at its first diffuse/glossy vertex after the camera hit; a training path leaves a pending
// row (inputs 16, radiance so far 3, throughput 3, path, pad) at each such vertex
struct NrcShade {
    float4* queries;  // 5 float4 per query
    float4* pending;  // 6 float4 per row
    uint32_t max_rows;
    // the camera hit is also a cache vertex when its material is Lambertian or GGX with roughness
    // >= camera_min_roughness (+inf: never)
    float camera_min_roughness;
};

template <bool kNrc, class Seed>
__device__ __forceinline__ void pt_shade(const DevScene& S, int cur, uint32_t bounce, uint32_t max_depth, int rr,
                                         const float4* __restrict__ q_o, const float4* __restrict__ q_d,
                                         const float2* __restrict__ q_hit, float4* __restrict__ nq_o,
                                         float4* __restrict__ nq_d, float4* __restrict__ s_o, float4* __restrict__ s_d,
                                         float4* __restrict__ s_c, float4* __restrict__ weight,
                                         float4* __restrict__ rad, Counters* c, const NrcShade& nrc,
                                         uint32_t omask) {
    __shared__ uint32_t s_append[kPtWaves + 1];
    const uint32_t n = c->queue_count[cur].v;
    if (blockIdx.x == 0 && threadIdx.x < 8)  // the next trace starts at 0
        c->work[threadIdx.x].v = c->swork[threadIdx.x].v = 0;
    const float nl = (float)(S.n_lights + (S.use_env ? 1u : 0u));
    for (uint32_t base = blockIdx.x * kPtBlock; base < n; base += gridDim.x * kPtBlock) {
        const uint32_t i = base + threadIdx.x;
        bool want_next = false, want_shadow = false, want_query = false, want_row = false;
        float4 no = {}, nd = {}, so = {}, sd = {}, sc = {};
        float4 rec[6] = {};  // cache query or pending row
        if (i < n) {
            const float4 o4 = q_o[i & omask], d4 = q_d[i];
            const float2 hh = q_hit[i];
            const uint32_t p = __float_as_uint(d4.w);
            const v3 o = mk(o4.x, o4.y, o4.z), d = mk(d4.x, d4.y, d4.z);
            const float4 W = weight[p];
            const float4 R = rad[p];
            v3 w3 = mk(W.x, W.y, W.z);
            Seed seed = seed_from_bits<Seed>(__float_as_uint(W.w));
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
                const bool cache_vertex =
                    bounce >= 1 ? (m.type == CUPBR_LAMBERT || m.type == CUPBR_PHONG || m.type == CUPBR_GGX)
                                : ((m.type == CUPBR_LAMBERT && nrc.camera_min_roughness <= 1.0f) ||
                                   (m.type == CUPBR_GGX && m.roughness >= nrc.camera_min_roughness));
                if (kNrc && cache_vertex && !(dot(N, inc) <= 0.0f)) {
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
                weight[p] = make_float4(w3.x, w3.y, w3.z, __uint_as_float(state_bits(seed)));
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

// Shade without the radiance cache. amdgpu_waves_per_eu(5, 8): 96 VGPRs instead of 102 and 5
// waves per SIMD instead of 4, without spills (23.7 -> 23.1 ms of shading per 1080p 64-spp step;
// 6 and 8 waves spill and are slower). The cache variant keeps its 123 VGPRs (it would spill).
template <class Seed>
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(5, 8)))
void k_pt_shade(DevScene S, int cur, uint32_t bounce, uint32_t max_depth, int rr, const float4* __restrict__ q_o,
                const float4* __restrict__ q_d, const float2* __restrict__ q_hit, float4* __restrict__ nq_o,
                float4* __restrict__ nq_d, float4* __restrict__ s_o, float4* __restrict__ s_d,
                float4* __restrict__ s_c, float4* __restrict__ weight, float4* __restrict__ rad, Counters* c,
                NrcShade nrc, uint32_t omask) {
    pt_shade<false, Seed>(S, cur, bounce, max_depth, rr, q_o, q_d, q_hit, nq_o, nq_d, s_o, s_d, s_c, weight, rad, c,
                          nrc, omask);
}

// Shade with the radiance cache: queries and training rows (NrcShade)
template <class Seed>
__global__ __launch_bounds__(kPtBlock)
void k_pt_shade_nrc(DevScene S, int cur, uint32_t bounce, uint32_t max_depth, int rr, const float4* __restrict__ q_o,
                    const float4* __restrict__ q_d, const float2* __restrict__ q_hit, float4* __restrict__ nq_o,
                    float4* __restrict__ nq_d, float4* __restrict__ s_o, float4* __restrict__ s_d,
                    float4* __restrict__ s_c, float4* __restrict__ weight, float4* __restrict__ rad, Counters* c,
                    NrcShade nrc, uint32_t omask) {
    pt_shade<true, Seed>(S, cur, bounce, max_depth, rr, q_o, q_d, q_hit, nq_o, nq_d, s_o, s_d, s_c, weight, rad, c,
                         nrc, omask);
}

// query count after the second bounce's shade: the queries inferred on the side stream
__global__ void k_snapshot_queries(Counters* c) { c->query_early.v = c->query_count.v; }

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

// One queue of the trace kernel, n rays, with per-lane refill: the queue is split into 8
// contiguous groups, group h's rays handed out by counter[h]. A wave starts on group g0 (its
// block's XCD), moves to the next group once its counter has passed the group's end, and gives
// fresh rays to its idle lanes whenever at least CUPBR_REFILL_IDLE (32) of them are idle (or all are), so that a
// lane whose ray ended does not wait for the slowest ray of a 64-ray chunk. finish(i, hit, r)
// consumes the result of ray i.
template <bool kAny, class Load, class Finish>
__device__ __forceinline__ void trace_queue(const DevScene& S, uint32_t n, uint32_t g0, CounterLine* counter,
                                            LaneStack<kPtLds>& st, Load load, Finish finish) {
    const uint32_t lane = __lane_id();
    uint32_t k = 0, h = g0;
    uint32_t r0 = h * n / 8u, rl = (h + 1u) * n / 8u - r0;
    bool active = false, pool = true, hit = false;
    uint32_t my = 0;
    RayState rs;
    for (;;) {
        const unsigned long long idle = __ballot(!active);
        if (pool && (__popcll(idle) >= CUPBR_REFILL_IDLE || idle == ~0ull)) {
            const uint32_t need = (uint32_t)__popcll(idle);
            const uint32_t leader = (uint32_t)__builtin_ctzll(idle);
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&counter[h].v, need);
            base = (uint32_t)__shfl((int)base, (int)leader);
            if (!active) {
                const uint32_t idx = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                if (idx < rl) {  // ray r0 + idx < (h + 1) n / 8 <= n
                    my = r0 + idx;
                    v3 o, d;
                    float dist;
                    load(my, o, d, dist);
                    active = ray_begin<kAny>(S, rs, o, d, dist, hit);
                    if (!active) finish(my, hit, rs.best);
                }
            }
            if (base + need >= rl) {  // group h has handed out all of its rays
                if (++k == 8) pool = false;
                else {
                    h = (g0 + k) & 7u;
                    r0 = h * n / 8u;
                    rl = (h + 1u) * n / 8u - r0;
                }
            }
        }
        const unsigned long long am = __ballot(active);
        if (am == 0ull) {
            if (!pool) break;
            continue;
        }
        // Wave-uniform phases: a primitive step for the lanes at a leaf once half of the wave's rays
        // are there (at most CUPBR_LEAF_BATCH lanes), or all of them (no lane has an inner node
        // left); an inner-node step for the others otherwise. A lane whose ray reached a leaf waits
        // during inner phases, so that the primitive test runs for many lanes at once instead of a
        // few in every iteration (before: leaf code in 97% of iterations for 9 of 64 lanes). Counted
        // against the wave's own rays, the rule also keeps a nearly empty wave (the last rays of a
        // launch) from holding its leaf lanes back until every other ray reaches a leaf (1080p, 1M
        // triangles: a fixed 16 lanes 1147, half capped at 16 1155-1159, half 1168 Msamples/s).
        const bool at_leaf = active && (rs.cur & kLeafBit);
        const uint32_t n_leaf = (uint32_t)__popcll(__ballot(at_leaf));
        const uint32_t n_act = (uint32_t)__popcll(am);
        const bool leaf_phase = n_leaf >= min((uint32_t)CUPBR_LEAF_BATCH, (n_act + 1u) / 2u) || n_leaf == n_act;
        const bool go = leaf_phase ? at_leaf : (active && !at_leaf);
        if (go && !ray_step<kAny>(S, rs, st, hit)) {
            finish(my, hit, rs.best);
            active = false;
        }
    }
}

// One bounce's extension rays (queue cur), then the previous bounce's shadow rays (parity cur ^ 1),
// each queue handed out by 8 group counters (trace_queue; short launches and the visit-counting
// instantiation hand out 64-ray chunks instead, one ray per lane until the chunk is done).
// amdgpu_waves_per_eu(8, 8): the compiler keeps the kernel within the SGPR budget of 8 waves per
// SIMD (spilling a few SGPRs into VGPR lanes); without it 104 SGPRs held the kernel to 6 resident
// waves per SIMD and a quarter of the persistent blocks only started when the queue was empty.
template <bool kCount>
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(8, 8)))
void k_pt_trace(DevScene S, int cur, const float4* __restrict__ q_o, const float4* __restrict__ q_d,
                float2* __restrict__ q_hit, const float4* __restrict__ s_o, const float4* __restrict__ s_d,
                const float4* __restrict__ s_c, float4* __restrict__ rad, Counters* c, uint32_t* spill,
                uint32_t spill_stride) {
    __shared__ uint32_t lds[kPtLds * kPtBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n_ext = c->queue_count[cur].v;
    const uint32_t n_sh = c->shadow_count[cur ^ 1].v;
    if (gtid == 0) {
        c->queue_count[cur ^ 1].v = 0;
        c->iqueue_count[cur ^ 1].v = 0;
        c->shadow_count[cur].v = 0;
        c->extension_rays += n_ext;
        c->shadow_rays += n_sh;
    }
    LaneStack<kPtLds> st{lds + threadIdx.x, kPtBlock, spill ? spill + gtid : nullptr, spill_stride};
    TravCount te, ts;
    // Each queue is split into 8 groups: group h owns its h-th eighth. Workgroups are dispatched
    // to the 8 XCDs round-robin, so the waves of block b start on group b % 8: each XCD traces
    // neighbouring queue entries (neighbouring pixels and hit points) and its own L2 holds the
    // nodes they share. A group that runs dry takes rays from the next groups in turn; every ray
    // is handed out once, whatever XCD a block runs on.
    const uint32_t g0 = blockIdx.x & 7u;
    // Refill pays off once every lane of the persistent grid has several rays to trace; a short
    // launch (deep bounces, one-frame calls) ends sooner with the one tail of 64-ray chunks than
    // with the two tails of the refill (extension rays, then shadow rays). The decision is the
    // same for every wave (n_ext and n_sh are read at the start).
    const bool refill = !kCount && n_ext + n_sh >= 4u * gridDim.x * kPtBlock;
    if (refill && gtid == 0) c->refill_launches += 1;
    if (refill) {
        trace_queue<false>(
            S, n_ext, g0, c->work, st,
            [&](uint32_t i, v3& o, v3& d, float& dist) {
                const float4 o4 = q_o[i], d4 = q_d[i];
                o = mk(o4.x, o4.y, o4.z);
                d = mk(d4.x, d4.y, d4.z);
                dist = 0.0f;
            },
            [&](uint32_t i, bool, const TraceResult& r) { q_hit[i] = make_float2(r.t, __uint_as_float(r.prim)); });
        trace_queue<true>(
            S, n_sh, g0, c->swork, st,
            [&](uint32_t j, v3& o, v3& d, float& dist) {
                const float4 o4 = s_o[j], d4 = s_d[j];
                o = mk(o4.x, o4.y, o4.z);
                d = mk(d4.x, d4.y, d4.z);
                dist = o4.w;
            },
            [&](uint32_t j, bool occluded, const TraceResult&) { shadow_result(occluded, s_d[j], s_c, j, rad); });
    } else {
        // 64-ray chunks, one ray per lane until the chunk is done (and the visit-counting pass):
        // group h's counter hands out its extension chunks, then its shadow chunks
        const uint32_t e_chunks = (n_ext + 63u) >> 6, s_chunks = (n_sh + 63u) >> 6;
        for (uint32_t k = 0; k < 8; ++k) {
            const uint32_t h = (g0 + k) & 7u;
            const uint32_t e0 = h * e_chunks / 8u, e_len = (h + 1u) * e_chunks / 8u - e0;
            const uint32_t s0 = h * s_chunks / 8u, s_len = (h + 1u) * s_chunks / 8u - s0;
            for (;;) {
                const uint32_t q = wave_fetch(&c->work[h].v);
                if (q >= e_len + s_len) break;
                if (q < e_len) {
                    const uint32_t i = ((e0 + q) << 6) + __lane_id();
                    if (i < n_ext) {
                        const float4 o4 = q_o[i], d4 = q_d[i];
                        TraceResult r;
                        traverse<false, kPtLds, kCount>(S, mk(o4.x, o4.y, o4.z), mk(d4.x, d4.y, d4.z), 0.0f, st, r,
                                                        &te);
                        q_hit[i] = make_float2(r.t, __uint_as_float(r.prim));
                    }
                } else {
                    const uint32_t j = ((s0 + q - e_len) << 6) + __lane_id();
                    if (j < n_sh) {
                        const float4 o4 = s_o[j], d4 = s_d[j];
                        TraceResult r;
                        const bool occluded = traverse<true, kPtLds, kCount>(S, mk(o4.x, o4.y, o4.z),
                                                                              mk(d4.x, d4.y, d4.z), o4.w, st, r, &ts);
                        shadow_result(occluded, d4, s_c, j, rad);
                    }
                }
            }
        }
    }
    if (kCount) {
        wave_add_counts(te, &c->ext_nodes, &c->ext_prims);
        wave_add_counts(ts, &c->shadow_nodes, &c->shadow_prims);
    }
}

// Bounce 0 (PathTracer.cu:112-126): the camera rays, made where they are traced and traced as
// packets. Queue slot chunks of 64 (SampleOrder: the samples of one pixel, or of a few neighbouring
// pixels, or an 8x8 pixel tile; GDPT: the five paths of about 13 neighbouring pixels) hold rays that
// leave the camera position in nearby directions, so each wave runs traverse_packet over one chunk.
// Blocks are dispatched to the 8 XCDs round-robin; block b takes chunk range
// (b % 8) * (gridDim.x / 8) + b / 8, so each XCD traces a contiguous eighth of the queue
// (neighbouring tiles share its L2). gridDim.x is a multiple of 8; n = queue 0's length.
// make(i, o, d) makes slot i's ray and writes the path's state. Queue 0 holds one origin, in slot 0:
// every camera ray leaves the camera position, and the first shade reads it for all (omask = 0).
template <class Make>
__device__ __forceinline__ void camera_packets(const DevScene& S, uint32_t n, float4* __restrict__ q_o,
                                               float2* __restrict__ q_hit, Make make) {
    const uint32_t blk = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    const uint32_t wave_first = blk * kPtBlock + (threadIdx.x & ~63u);
    if (wave_first >= n) return;  // whole waves only: traverse_packet needs all 64 lanes
    const uint32_t i = wave_first + __lane_id();
    const bool valid = i < n;
    v3 o = splat(0.0f), d = mk(0.0f, 0.0f, 1.0f);
    if (valid) make(i, o, d);
    if (i == 0u) q_o[0] = make_float4(o.x, o.y, o.z, 0.0f);
    TraceResult r;
    traverse_packet(S, o, d, valid, r);
    if (valid) q_hit[i] = make_float2(r.t, __uint_as_float(r.prim));
}

// bounce 0 of a path-tracer or volume-renderer batch: the rays of k_pt_raygen, traced at once;
// it also resets the batch's counters
template <class Seed>
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(8, 8)))
void k_pt_camera(DevScene S, uint32_t n, cupbr_camera cam, uint32_t w, uint32_t h, uint32_t frame0, SampleOrder order,
                 uint32_t train_thresh, float4* __restrict__ q_o, float4* __restrict__ q_d,
                 float4* __restrict__ weight, float4* __restrict__ rad, float2* __restrict__ q_hit, Counters* c) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        reset_batch_counters(c, n);
        c->queue_count[1].v = 0;
        c->extension_rays += n;
    }
    camera_packets(S, n, q_o, q_hit, [&](uint32_t i, v3& o, v3& d) {
        Seed seed;
        uint32_t pix, frame;
        camera_path(cam, w, h, frame0, order, i, seed, pix, frame, o, d);
        store_camera_path(i, d, seed, pix, frame, train_thresh, q_d, weight, rad);
    });
}

// The running mean of PathTracer.cu:162-168 per pixel, frame by frame in order. A pixel's samples
// are runs of consecutive slots (SampleOrder: up to f frames); a thread loads up to 8 of them at
// once, so that the 128-byte line they share is fetched once.
__global__ __launch_bounds__(kPtBlock) void k_accumulate(float* __restrict__ img, const float4* __restrict__ rad,
                                                         uint32_t w, uint32_t h, SampleOrder order, uint32_t frame0,
                                                         uint32_t nspp) {
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= w * h) return;
    const uint32_t py = pix / w, j = tile_slot(pix - py * w, py, w, h);
    v3 acc = ld3(img + 3 * (size_t)pix);
    for (uint32_t s = 0; s < nspp;) {
        const uint32_t base = order.slot(s, j);
        const uint32_t run = min(order.f - s % order.f, nspp - s);  // slots base .. base + run - 1
        for (uint32_t k = 0; k < run; k += 8) {
            const uint32_t m = min(8u, run - k);
            float4 R[8];
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q)
                if (q < m) R[q] = rad[base + k + q];
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) {
                if (q >= m) break;
                const v3 L = mk(R[q].x, R[q].y, R[q].z);
                const uint32_t f = frame0 + s + k + q;
                if (f > 0) {
                    const float a = 1.0f / ((float)f + 1.0f);
                    acc = (1.0f - a) * acc + a * L;
                } else {
                    acc = L;
                }
            }
        }
        s += run;
    }
    st3(img + 3 * (size_t)pix, acc);
}

// ------------------------------------------------------------------ gradient-domain paths
// gdpt_kernel (GradientDomain.cu:160-246) on the same queues. A group is one interior pixel of one
// frame: its base path and four shifted paths (left, right, up, down), path ids 5g + k. Each
// follows collect_radiance (:110-158) on its own part of the group's seed chain; a shifted path
// stops at the first depth where the base path has no vertex (:126-127). All five advance bounce
// by bounce: the base path's shade of bounce b records the queue slot of its bounce-(b+1) ray (or
// none) in slot[(b+1) & 1][g], and a shifted path's shade of bounce b+1 reads that ray's hit, one
// launch later. A shifted path alive at bounce b passed the check at b-1, so the base path was
// shaded at b-1 and wrote the slot it reads: no stale entries, no reads of what the same launch
// writes. Queue slots of a frame's groups follow the 8x8 tiles of the interior pixels.
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

// The camera ray of GDPT path pid = 5 g + k (g = frame s of interior pixel gi; k: base, left,
// right, up, down), and its generator state after the ray's draws (GradientDomain.cu:180-213).
__device__ __forceinline__ void gdpt_camera_path(const cupbr_camera& cam, uint32_t w, uint32_t h, uint32_t frame0,
                                                 uint32_t n_int, uint32_t pid, uint32_t& chain, v3& o, v3& d) {
    const uint32_t g = pid / 5u, k = pid - 5u * g;
    const uint32_t s = g / n_int, gi = g - s * n_int;
    uint32_t ix, iy;
    tile_pixel(gi, w - 2u, h - 2u, ix, iy);
    const uint32_t px = ix + 1u, py = iy + 1u;
    // the five camera rays of a group come from one chain, base, left, right, up, down, two draws
    // each; each path continues from the chain as it stood after its own ray (:180-213). Path k
    // skips the 2k draws of the rays before its own.
    chain = tea4(py * w + px, frame0 + s);
    for (uint32_t j = 0; j < 2u * k; ++j) chain = 1664525u * chain + 1013904223u;  // rnd's state step
    const uint32_t nx = k == 1 ? px - 1u : (k == 2 ? px + 1u : px);
    const uint32_t ny = k == 3 ? py + 1u : (k == 4 ? py - 1u : py);
    launch_ray(nx, ny, w, h, cam, chain, o, d);
}

// the GDPT batch's counter reset that precedes bounce 0
__device__ __forceinline__ void reset_gdpt_counters(Counters* c, uint32_t npath) {
    c->queue_count[0].v = npath;
    for (int k = 0; k < 8; ++k) c->work[k].v = c->swork[k].v = 0;
    c->shadow_count[0].v = 0;
    c->shadow_count[1].v = 0;
}

__global__ __launch_bounds__(kPtBlock) void k_gdpt_raygen(cupbr_camera cam, uint32_t w, uint32_t h, uint32_t frame0,
                                                          uint32_t n_int, uint32_t n_groups, float4* __restrict__ q_o,
                                                          float4* __restrict__ q_d, float4* __restrict__ weight,
                                                          float4* __restrict__ rad, uint32_t* __restrict__ slot0,
                                                          Counters* c) {
    // one thread per path, so that the queue and state writes of a wave are contiguous
    const uint32_t pid = blockIdx.x * blockDim.x + threadIdx.x;
    if (pid == 0) reset_gdpt_counters(c, 5u * n_groups);
    if (pid >= 5u * n_groups) return;
    uint32_t chain;
    v3 o, d;
    gdpt_camera_path(cam, w, h, frame0, n_int, pid, chain, o, d);
    q_o[pid] = make_float4(o.x, o.y, o.z, 0.0f);
    q_d[pid] = make_float4(d.x, d.y, d.z, __uint_as_float(pid));
    weight[pid] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(chain));
    rad[pid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (pid % 5u == 0u) slot0[pid / 5u] = pid;
}

// bounce 0 of a GDPT batch: the rays of k_gdpt_raygen, traced at once (camera_packets); it also
// resets the batch's counters. n = 5 x the batch's groups.
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(8, 8)))
void k_gdpt_camera(DevScene S, uint32_t n, cupbr_camera cam, uint32_t w, uint32_t h, uint32_t frame0, uint32_t n_int,
                   float4* __restrict__ q_o, float4* __restrict__ q_d, float4* __restrict__ weight,
                   float4* __restrict__ rad, uint32_t* __restrict__ slot0, float2* __restrict__ q_hit, Counters* c) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        reset_gdpt_counters(c, n);
        c->queue_count[1].v = 0;
        c->iqueue_count[1].v = 0;
        c->extension_rays += n;
    }
    camera_packets(S, n, q_o, q_hit, [&](uint32_t i, v3& o, v3& d) {
        uint32_t chain;
        gdpt_camera_path(cam, w, h, frame0, n_int, i, chain, o, d);
        q_d[i] = make_float4(d.x, d.y, d.z, __uint_as_float(i));
        weight[i] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(chain));
        rad[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (i % 5u == 0u) slot0[i / 5u] = i;
    });
}

// one bounce of collect_radiance (:121-156) for every queued path. amdgpu_waves_per_eu(4, 8): 102
// VGPRs and 4 waves per SIMD (Cornell box 512², 128 spp: 504 Msamples/s; 496 at 5 waves with 96
// VGPRs, 478 at 6 waves with 80 VGPRs and a 24-byte spill).
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(4, 8)))
void k_gdpt_shade(DevScene S, int cur, uint32_t bounce, uint32_t max_depth, const float4* __restrict__ q_o,
                  const float4* __restrict__ q_d, const float2* __restrict__ q_hit, float4* __restrict__ nq_o,
                  float4* __restrict__ nq_d, float4* __restrict__ s_o, float4* __restrict__ s_d,
                  float4* __restrict__ s_c, float4* __restrict__ weight, float4* __restrict__ rad, Counters* c,
                  const uint32_t* __restrict__ slot_now, uint32_t* __restrict__ slot_next, uint32_t omask) {
    __shared__ uint32_t s_append[kPtWaves + 1];
    const uint32_t n = c->queue_count[cur].v;
    if (blockIdx.x == 0 && threadIdx.x < 8) c->work[threadIdx.x].v = c->swork[threadIdx.x].v = 0;
    const float nl = (float)(S.n_lights + (S.use_env ? 1u : 0u));
    for (uint32_t base = blockIdx.x * kPtBlock; base < n; base += gridDim.x * kPtBlock) {
        const uint32_t i = base + threadIdx.x;
        bool want_next = false, want_shadow = false;
        float4 no = {}, nd = {}, so = {}, sd = {}, sc = {};
        if (i < n) {
            const float4 o4 = q_o[i & omask], d4 = q_d[i];
            const float2 hh = q_hit[i];
            const uint32_t p = __float_as_uint(d4.w);
            const uint32_t g = p / 5u, k = p - 5u * g;
            bool go = true;
            if (k != 0) {  // :126-127: the base path needs a vertex at this depth
                const uint32_t bs = slot_now[g];
                go = bs != kNoSlot && __float_as_uint(q_hit[bs].y) != kNoHit;
            }
            if (go) {
                const float4 W = weight[p];
                const float4 R = rad[p];
                v3 w3 = mk(W.x, W.y, W.z);
                uint32_t seed = __float_as_uint(W.w);
                v3 L = mk(R.x, R.y, R.z);
                const v3 o = mk(o4.x, o4.y, o4.z), d = mk(d4.x, d4.y, d4.z);
                const uint32_t prim = __float_as_uint(hh.y);
                if (prim == kNoHit) {  // :128-137
                    if (S.use_env) L += w3 * env_lookup(S, d);
                } else {
                    const v3 P = o + hh.x * d;
                    v3 N;
                    uint32_t obj;
                    hit_geometry(S, prim, P, N, obj);
                    const Mat m = load_mat(S.mats + obj);
                    const v3 inc = -1.0f * d;
                    L += w3 * m.albedo_e;  // emissiveIllumintationGDPT :39-44
                    if (m.type != CUPBR_GLASS && dot(N, inc) <= 0.0f) {  // directIlluminationGDPT :46-88
                        w3 = splat(0.0f);
                    } else {
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
                    }
                    // indirectIlluminationGDPT :90-108 (a zero weight keeps the path going: the
                    // base path's later vertices decide how far its shifted paths go)
                    const DirPdf sp = sample_direction(m, seed, inc, N);
                    if (!(norm(sp.d) == 0.0f)) {
                        w3 = w3 * fabsf(dot(sp.d, N)) * brdf(m, inc, sp.d, N) / sp.pdf;
                        if (bounce + 1 < max_depth) {  // :154-156
                            const v3 norg = P + 0.01f * sp.d;
                            const v3 ndir = normalize(sp.d);
                            want_next = true;
                            no = make_float4(norg.x, norg.y, norg.z, 0.0f);
                            nd = make_float4(ndir.x, ndir.y, ndir.z, __uint_as_float(p));
                        }
                    }
                    weight[p] = make_float4(w3.x, w3.y, w3.z, __uint_as_float(seed));
                }
                rad[p] = make_float4(L.x, L.y, L.z, R.w);
            }
            if (k == 0) slot_next[g] = kNoSlot;  // overwritten below when the base path goes on
        }
        const uint32_t ns = block_append(want_next, &c->queue_count[cur ^ 1].v, s_append);
        if (want_next) {
            nq_o[ns] = no;
            nq_d[ns] = nd;
            const uint32_t p = __float_as_uint(nd.w);
            const uint32_t g = p / 5u;
            if (p == 5u * g) slot_next[g] = ns;
        }
        const uint32_t ss = block_append(want_shadow, &c->shadow_count[cur].v, s_append);
        if (want_shadow) {
            s_o[ss] = so;
            s_d[ss] = sd;
            s_c[ss] = sc;
        }
    }
}

// gradients and running means of the five buffers (:214-245), frame by frame in order
__global__ __launch_bounds__(kPtBlock) void k_gdpt_accumulate(float* __restrict__ base, float* __restrict__ gxf,
                                                              float* __restrict__ gxb, float* __restrict__ gyf,
                                                              float* __restrict__ gyb, const float4* __restrict__ rad,
                                                              uint32_t w, uint32_t h, uint32_t n_int, uint32_t frame0,
                                                              uint32_t nspp) {
    const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= n_int) return;
    uint32_t ix, iy;
    tile_pixel(gi, w - 2u, h - 2u, ix, iy);
    const size_t i3 = 3 * ((size_t)(iy + 1u) * w + ix + 1u);
    v3 acc[5] = {ld3(base + i3), ld3(gxf + i3), ld3(gxb + i3), ld3(gyf + i3), ld3(gyb + i3)};
    for (uint32_t s = 0; s < nspp; ++s) {
        const size_t p = 5 * ((size_t)s * n_int + gi);
        v3 L[5];
        for (int k = 0; k < 5; ++k) {
            const float4 R = rad[p + k];
            L[k] = mk(R.x, R.y, R.z);
        }
        const v3 val[5] = {L[0], 0.5f * (L[0] - L[1]), 0.5f * (L[2] - L[0]), 0.5f * (L[0] - L[4]),
                           0.5f * (L[3] - L[0])};
        const uint32_t f = frame0 + s;
        for (int k = 0; k < 5; ++k) {
            if (f > 0) {
                const float a = 1.0f / ((float)f + 1.0f);
                acc[k] = (1.0f - a) * acc[k] + a * val[k];
            } else {
                acc[k] = val[k];
            }
        }
    }
    st3(base + i3, acc[0]);
    st3(gxf + i3, acc[1]);
    st3(gxb + i3, acc[2]);
    st3(gyf + i3, acc[3]);
    st3(gyb + i3, acc[4]);
}

// ------------------------------------------------------------------ volume renderer paths
// volume_kernel (VolumeRenderer.cu:214-310) on the same queues, one iteration of its loop per
// bounce. A path's medium state lives in rad[p].w: bit 31 = inside an object, bits 0..30 = that
// object (the last interface crossed). Rays outside are traced by the trace kernels against the
// scene; rays inside go to the inside queue and are traced by the shade kernel against their
// object alone (Tracing::traceRay(scene, ray, index), TracingDetail.h:68-116), whose primitives it
// loops over. A path whose weight becomes zero (absorbed, back-facing, Russian roulette) stops:
// every later term would be multiplied by that weight.
constexpr uint32_t kInsideBit = 0x80000000u;

__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(4, 8)))
void k_vol_shade(DevScene S, int cur, uint32_t bounce, uint32_t max_depth, int rr, const float4* __restrict__ q_o,
                 const float4* __restrict__ q_d, const float2* __restrict__ q_hit, const float4* __restrict__ iq_o,
                 const float4* __restrict__ iq_d, float4* __restrict__ nq_o, float4* __restrict__ nq_d,
                 float4* __restrict__ niq_o, float4* __restrict__ niq_d, float4* __restrict__ s_o,
                 float4* __restrict__ s_d, float4* __restrict__ s_c, float4* __restrict__ weight,
                 float4* __restrict__ rad, Counters* c, uint32_t omask) {
    __shared__ uint32_t s_append[kPtWaves + 1];
    const uint32_t n_out = c->queue_count[cur].v, n_in = c->iqueue_count[cur].v, n = n_out + n_in;
    if (blockIdx.x == 0 && threadIdx.x < 8) c->work[threadIdx.x].v = c->swork[threadIdx.x].v = 0;
    const float nl = (float)(S.n_lights + (S.use_env ? 1u : 0u));
    unsigned long long inside_traces = 0;
    for (uint32_t base = blockIdx.x * kPtBlock; base < n; base += gridDim.x * kPtBlock) {
        const uint32_t i = base + threadIdx.x;
        bool want_next = false, want_shadow = false, next_inside = false;
        float4 no = {}, nd = {}, so = {}, sd = {}, sc = {};
        if (i < n) {
            const bool from_inside = i >= n_out;
            const float4 o4 = from_inside ? iq_o[i - n_out] : q_o[i & omask];
            const float4 d4 = from_inside ? iq_d[i - n_out] : q_d[i];
            const uint32_t p = __float_as_uint(d4.w);
            const float4 W = weight[p];
            const float4 R = rad[p];
            v3 w3 = mk(W.x, W.y, W.z);
            uint32_t seed = __float_as_uint(W.w);
            v3 L = mk(R.x, R.y, R.z);
            uint32_t vstate = __float_as_uint(R.w);
            bool inside = (vstate & kInsideBit) != 0u;
            uint32_t obj_index = vstate & ~kInsideBit;
            const v3 o = mk(o4.x, o4.y, o4.z), d = mk(d4.x, d4.y, d4.z);
            TraceResult hr;
            if (from_inside) {
                ++inside_traces;
                object_trace(S, obj_index, o, d, hr);
            } else {
                const float2 hh = q_hit[i];
                hr.t = hh.x;
                hr.prim = __float_as_uint(hh.y);
            }
            const v3 inc = -1.0f * d;
            if (hr.prim == kNoHit) {  // :253-260
                if (S.use_env) L += w3 * env_lookup(S, d);
            } else {
                const v3 P = o + hr.t * d;
                v3 N;
                uint32_t obj;
                hit_geometry(S, hr.prim, P, N, obj);
                const Mat m = load_mat(S.mats + obj);
                const int32_t vol_obj = inside ? (int32_t)obj : -1;  // the payload's volume (:277)
                bool next_valid = false, alive = true, medium_event = false;
                v3 out_dir = splat(0.0f), ray_start = splat(0.0f);
                // handleMediumInteraction :88-212
                const VolCoef vc = vol_coef(S, vol_obj);
                const v3 st = vc.sa + vc.ss;
                uint32_t ch = (uint32_t)(rnd(seed) * 3.0f);
                ch = min(ch, 2u);
                const float stc = comp(st, (int)ch);
                if (!safe_float_equal(stc, 0.0f)) {
                    const float t = -logf(1.0f - rnd(seed)) / stc;
                    if (t < hr.t) {
                        medium_event = true;
                        const v3 ep = o + t * d;
                        const float sprob = comp(vc.ss, (int)ch) / stc;
                        if (rnd(seed) < sprob) {
                            float pdf = 0.0f;
                            for (int k = 0; k < 3; ++k) pdf += comp(st, k) * expf(-1.0f * comp(st, k) * t);
                            pdf /= 3.0f;
                            w3 = w3 * vc.ss / sprob * vexp(-1.0f * st * t) / pdf;
                            v3 ldir, le;
                            float dist;
                            if (nee_sample(S, m, seed, ep, inc, N, ldir, le, dist)) {
                                const v3 sdir = normalize(ldir);
                                v3 sorg, att = splat(1.0f);
                                if (inside) {  // to the object's boundary first (:159-171)
                                    ++inside_traces;
                                    TraceResult ge;
                                    float gd = 0.0f;
                                    v3 gp = ep;
                                    if (object_trace(S, obj_index, ep, sdir, ge)) {
                                        gd = ge.t;
                                        gp = ep + ge.t * sdir;
                                    }
                                    att = vexp(-1.0f * st * gd) * hg_phase(vc.g, -dot(inc, ldir));
                                    sorg = gp + 0.001f * ldir;
                                    dist -= gd;
                                } else {
                                    sorg = ep + 0.001f * ldir;
                                }
                                const v3 sst = ld3(S.volume.sigma_a) + ld3(S.volume.sigma_s);
                                const v3 contrib = nl * vexp(-1.0f * sst * fminf(dist, 100000000.0f)) * att * le * w3;
                                if (!(contrib.x == 0.0f && contrib.y == 0.0f && contrib.z == 0.0f)) {
                                    want_shadow = true;
                                    so = make_float4(sorg.x, sorg.y, sorg.z, dist);
                                    sd = make_float4(sdir.x, sdir.y, sdir.z, __uint_as_float(p));
                                    sc = make_float4(contrib.x, contrib.y, contrib.z, 0.0f);
                                }
                            }
                            const DirPdf hs = hg_sample(vc.g, -1.0f * inc, seed);
                            out_dir = hs.d;
                            ray_start = ep;
                            next_valid = true;
                        } else {
                            alive = false;  // absorbed (:120-125)
                        }
                    } else {
                        float pdf = 0.0f;
                        for (int k = 0; k < 3; ++k) pdf += expf(-comp(st, k) * hr.t);
                        pdf /= 3.0f;
                        w3 = w3 * vexp(-1.0f * st * hr.t) / pdf;
                        ray_start = P;
                    }
                }
                if (alive && !medium_event) {
                    if (m.type == CUPBR_VOLUME_MATERIAL) {  // medium interface :269-278
                        const DirPdf sdp = sample_direction(m, seed, inc, N);
                        out_dir = sdp.d;
                        ray_start = P;
                        const bool refl = dot(inc, N) * dot(out_dir, N) > 0;
                        inside = refl ? inside : !inside;
                        obj_index = obj;
                        next_valid = true;
                    } else {
                        if (m.type != CUPBR_GLASS && dot(N, inc) <= 0.0f) {  // :24-70
                            w3 = splat(0.0f);
                        } else {
                            v3 ldir, le;
                            float dist;
                            if (nee_sample(S, m, seed, P, inc, N, ldir, le, dist)) {
                                const v3 contrib = (nl * fmaxf(0.0f, dot(N, ldir))) *
                                                   vexp(-1.0f * (vc.ss + vc.sa) * fminf(dist, 100000000.0f)) *
                                                   brdf(m, inc, ldir, N) * le * w3;
                                if (!(contrib.x == 0.0f && contrib.y == 0.0f && contrib.z == 0.0f)) {
                                    const v3 sorg = P + 0.001f * ldir;
                                    const v3 sdir = normalize(ldir);
                                    want_shadow = true;
                                    so = make_float4(sorg.x, sorg.y, sorg.z, dist);
                                    sd = make_float4(sdir.x, sdir.y, sdir.z, __uint_as_float(p));
                                    sc = make_float4(contrib.x, contrib.y, contrib.z, 0.0f);
                                }
                            }
                        }
                        const DirPdf sdp = sample_direction(m, seed, inc, N);  // :72-86
                        if (!(norm(sdp.d) == 0.0f)) {
                            w3 = w3 * fabsf(dot(sdp.d, N)) * brdf(m, inc, sdp.d, N) / sdp.pdf;
                            out_dir = sdp.d;
                            ray_start = P;
                            next_valid = true;
                        }
                    }
                    if (rr) {  // :280-289
                        const float alpha = clampf(fmaxf(w3.x, fmaxf(w3.y, w3.z)), 0.0f, 1.0f);
                        if (rnd(seed) > alpha || safe_float_equal(alpha, 0.0f)) {
                            next_valid = false;
                        } else {
                            w3 = w3 / alpha;
                        }
                    }
                }
                if (alive && next_valid && bounce + 1 < max_depth && !(w3.x == 0.0f && w3.y == 0.0f && w3.z == 0.0f)) {
                    const v3 norg = ray_start + 0.001f * out_dir;  // :298
                    const v3 ndir = normalize(out_dir);
                    want_next = true;
                    next_inside = inside;
                    no = make_float4(norg.x, norg.y, norg.z, 0.0f);
                    nd = make_float4(ndir.x, ndir.y, ndir.z, __uint_as_float(p));
                }
                weight[p] = make_float4(w3.x, w3.y, w3.z, __uint_as_float(seed));
            }
            vstate = (inside ? kInsideBit : 0u) | obj_index;
            rad[p] = make_float4(L.x, L.y, L.z, __uint_as_float(vstate));
        }
        const bool app_out = want_next && !next_inside, app_in = want_next && next_inside;
        const uint32_t ns = block_append(app_out, &c->queue_count[cur ^ 1].v, s_append);
        if (app_out) {
            nq_o[ns] = no;
            nq_d[ns] = nd;
        }
        const uint32_t ni = block_append(app_in, &c->iqueue_count[cur ^ 1].v, s_append);
        if (app_in) {
            niq_o[ni] = no;
            niq_d[ni] = nd;
        }
        const uint32_t ss = block_append(want_shadow, &c->shadow_count[cur].v, s_append);
        if (want_shadow) {
            s_o[ss] = so;
            s_d[ss] = sd;
            s_c[ss] = sc;
        }
    }
    // object traces of this launch count as closest-hit rays
    for (int off = 32; off > 0; off >>= 1) inside_traces += __shfl_xor(inside_traces, off);
    if (__lane_id() == 0 && inside_traces) atomicAdd(&c->extension_rays, inside_traces);
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
        for (int k = 0; k < 2; ++k) {
            if ((rc = wf.s_o[k].ensure(n * 16))) return rc;
            if ((rc = wf.s_d[k].ensure(n * 16))) return rc;
            if ((rc = wf.s_c[k].ensure(n * 16))) return rc;
        }
        wf.capacity = paths;
    }
    if ((rc = wf.counters.ensure(sizeof(Counters)))) return rc;
    const uint32_t entries = r->scene->stack_entries;
    if (entries > (uint32_t)kPtLds) {
        if ((rc = wf.spill.ensure((size_t)(entries - kPtLds) * grid_threads * 4))) return rc;
    }
    return CUPBR_OK;
}

// SampleOrder of a batch of nb frames of npix pixels: as many frames per wave as divide nb (a power
// of two up to 64) while 64 / frames still divides the pixel count
SampleOrder sample_order(uint32_t npix, uint32_t nb) {
    uint32_t f = 64;
    while (f > 1u && (nb % f != 0u || npix % (64u / f) != 0u)) f >>= 1;
    return SampleOrder{npix, nb, f};
}

int render_pathtracer(cupbr_renderer* r, const cupbr_camera& cam, uint32_t frame0, uint32_t spp) {
    const uint32_t w = r->width, h = r->height, npix = w * h;
    if (npix == 0 || spp == 0) return CUPBR_OK;
    const bool nrc = r->nrc != nullptr;
    // Paths in flight, in whole frames: up to 128M (200 B of path state and queues per path, so
    // up to ~27 GB). A trace launch ends with the last 64-ray chunks of its slowest waves, and
    // with more paths per launch that tail is a smaller share of it (1080p, 64 spp, Msamples/s:
    // 17 frames per batch 637, 32 frames 651, all 64 frames 656). Bounded to an eighth of the
    // device memory. A call of one frame (interactive use) holds one frame of paths. With a
    // radiance cache the same holds: a call of k frames is one batch, and the cache trains after
    // each batch.
    const uint32_t target = 1u << 27;
    uint32_t cap = 1u << 27;
    {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b > 0)
            cap = (uint32_t)std::min<size_t>(cap, std::max<size_t>(total_b / 8 / 200, 1u << 20));
    }
    uint32_t batch = std::max<uint32_t>(1, (target + npix - 1) / npix);
    batch = std::min(batch, std::max<uint32_t>(1, cap / npix));
    batch = std::min(batch, spp);
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
    auto trace = r->counting ? k_pt_trace<true> : k_pt_trace<false>;
    // camera rays as wave packets (the visit-counting pass keeps the per-ray traversal, whose
    // counts define the algorithmic bytes of a trace launch)
    const bool packet = !r->counting && r->scene->stack_entries <= kPacketStack && S.n_nodes > 0;
    const bool pcg = r->rng == CUPBR_RNG_PCG;
    auto shade = nrc ? (pcg ? k_pt_shade_nrc<Pcg32> : k_pt_shade_nrc<uint32_t>)
                     : (pcg ? k_pt_shade<Pcg32> : k_pt_shade<uint32_t>);
    auto raygen = pcg ? k_pt_raygen<Pcg32> : k_pt_raygen<uint32_t>;
    auto camera = pcg ? k_pt_camera<Pcg32> : k_pt_camera<uint32_t>;
    NrcShade ns{nullptr, nullptr, 0, __builtin_inff()};
    uint32_t train_thresh = 0;
    if (nrc) {
        if ((rc = r->nrc_queries.ensure((size_t)batch * npix * 80)) ||
            (rc = r->nrc_pending.ensure((size_t)r->nrc_max_rows * 96)) ||
            (rc = r->nrc_rows.ensure((size_t)r->nrc_max_rows * 80)))
            return rc;
        ns = NrcShade{r->nrc_queries.as<float4>(), r->nrc_pending.as<float4>(), r->nrc_max_rows,
                      r->nrc_camera_min_roughness};
        if (!r->side_stream) {
            if (hipStreamCreateWithFlags(&r->side_stream, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&r->ev_queries, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&r->ev_side_done, hipEventDisableTiming) != hipSuccess)
                return set_error(CUPBR_ERR_HIP, "could not create the radiance cache's stream");
        }
        const double frac = std::min(1.0, (double)r->nrc_train_paths / (double)npix);
        train_thresh = frac >= 1.0 ? 0xFFFFFFFFu : (uint32_t)(frac * 4294967296.0);
    }
    if ((rc = stage_mark(r, -1))) return rc;
    for (uint32_t s0 = 0; s0 < spp; s0 += batch) {
        const uint32_t nb = std::min(batch, spp - s0);
        bool side_pending = false;  // cache inference of this batch running on the side stream
        const uint32_t npath = nb * npix;
        const SampleOrder order = sample_order(npix, nb);
        // with camera packets, bounce 0 makes its own camera rays (k_pt_camera)
        if (!packet) {
            hipLaunchKernelGGL(raygen, dim3((npath + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0, r->stream, cam, w,
                               h, frame0 + s0, npath, order, wf.q_o[0].as<float4>(), wf.q_d[0].as<float4>(),
                               wf.weight.as<float4>(), wf.radiance.as<float4>(), c, train_thresh);
            if ((rc = stage_mark(r, 3))) return rc;
        }
        // per bounce: trace (this bounce's extension rays + the previous bounce's shadow rays), then
        // shade (appends the next extension rays and this bounce's shadow rays, by bounce parity)
        for (uint32_t b = 0; b <= bounces; ++b) {
            const int cur = (int)(b & 1);
            if (b == 0 && packet) {
                const uint32_t nblk = (npath + kPtBlock - 1) / kPtBlock;
                hipLaunchKernelGGL(camera, dim3((nblk + 7u) / 8u * 8u), dim3(kPtBlock), 0, r->stream, S, npath, cam, w,
                                   h, frame0 + s0, order, train_thresh, wf.q_o[0].as<float4>(), wf.q_d[0].as<float4>(),
                                   wf.weight.as<float4>(), wf.radiance.as<float4>(), wf.q_hit.as<float2>(), c);
            } else {
                hipLaunchKernelGGL(trace, dim3(blocks), dim3(kPtBlock), 0, r->stream, S, cur, wf.q_o[cur].as<float4>(),
                                   wf.q_d[cur].as<float4>(), wf.q_hit.as<float2>(), wf.s_o[cur ^ 1].as<float4>(),
                                   wf.s_d[cur ^ 1].as<float4>(), wf.s_c[cur ^ 1].as<float4>(),
                                   wf.radiance.as<float4>(), c, spill, grid_threads);
            }
            if ((rc = stage_mark(r, 0))) return rc;
            if (b == bounces) break;  // the last trace only had the last bounce's shadow rays
            hipLaunchKernelGGL(shade, dim3(blocks), dim3(kPtBlock), 0, r->stream, S, cur, b, r->max_trace_depth,
                               (int)r->russian_roulette, wf.q_o[cur].as<float4>(),
                               wf.q_d[cur].as<float4>(), wf.q_hit.as<float2>(), wf.q_o[cur ^ 1].as<float4>(),
                               wf.q_d[cur ^ 1].as<float4>(), wf.s_o[cur].as<float4>(), wf.s_d[cur].as<float4>(),
                               wf.s_c[cur].as<float4>(), wf.weight.as<float4>(), wf.radiance.as<float4>(), c, ns,
                               b == 0 && packet ? 0u : ~0u);
            if ((rc = stage_mark(r, 1))) return rc;
            if (nrc && b == 1 && bounces > 2) {
                // The paths that queried the cache so far have no shadow ray left in flight (the
                // shadow rays of bounce 0 were traced before this shade, and a query vertex
                // traces none), and no later kernel touches their radiance until the final pass
                // below: their inference runs on the side stream while the remaining bounces
                // (mostly training paths) are traced.
                hipLaunchKernelGGL(k_snapshot_queries, dim3(1), dim3(1), 0, r->stream, c);
                CUPBR_HIP_CHECK(hipEventRecord(r->ev_queries, r->stream));
                CUPBR_HIP_CHECK(hipStreamWaitEvent(r->side_stream, r->ev_queries, 0));
                if ((rc = nrc_infer(r->nrc, r->nrc_queries.as<float>(), npath, &c->query_early.v, 20,
                                    wf.radiance.as<float>(), true, r->side_stream)))
                    return rc;
                CUPBR_HIP_CHECK(hipEventRecord(r->ev_side_done, r->side_stream));
                side_pending = true;
            }
        }
        if (nrc) {
            // throughput * cache for the paths that ended at a cache vertex (those not already
            // inferred on the side stream), then the training rows
            if ((rc = nrc_infer(r->nrc, r->nrc_queries.as<float>(), npath, &c->query_count.v, 20,
                                wf.radiance.as<float>(), true, r->stream,
                                side_pending ? &c->query_early.v : nullptr)))
                return rc;
            if (side_pending) {
                CUPBR_HIP_CHECK(hipStreamWaitEvent(r->stream, r->ev_side_done, 0));
                side_pending = false;
            }
            hipLaunchKernelGGL(k_nrc_rows, dim3((r->nrc_max_rows + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0,
                               r->stream, r->nrc_pending.as<float4>(), c, r->nrc_max_rows, wf.radiance.as<float4>(),
                               r->nrc_rows.as<float4>());
            if ((rc = stage_mark(r, 3))) return rc;
        }
        hipLaunchKernelGGL(k_accumulate, dim3((npix + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0, r->stream, img,
                           wf.radiance.as<float4>(), w, h, order, frame0 + s0, nb);
        if ((rc = stage_mark(r, 3))) return rc;
        if (nrc) {
            // the query and row counts in one copy (three consecutive counter lines)
            static_assert(offsetof(Counters, row_count) == offsetof(Counters, query_count) + 2 * sizeof(CounterLine),
                          "query_count, query_early and row_count are consecutive");
            CounterLine lines[3];
            CUPBR_HIP_CHECK(hipMemcpyAsync(lines, &c->query_count, sizeof(lines), hipMemcpyDeviceToHost, r->stream));
            CUPBR_HIP_CHECK(hipStreamSynchronize(r->stream));
            const uint32_t counts[2] = {lines[0].v, lines[2].v};
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

int volume_paths_wavefront(cupbr_renderer* r, const cupbr_camera& cam, uint32_t frame0, uint32_t spp) {
    const uint32_t w = r->width, h = r->height, npix = w * h;
    if (npix == 0 || spp == 0) return CUPBR_OK;
    const uint32_t target = 1u << 27;
    uint32_t cap = 1u << 27;
    {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b > 0)
            cap = (uint32_t)std::min<size_t>(cap, std::max<size_t>(total_b / 8 / 264, 1u << 20));
    }
    uint32_t batch = std::max<uint32_t>(1, (target + npix - 1) / npix);
    batch = std::min(batch, std::max<uint32_t>(1, cap / npix));
    batch = std::min(batch, spp);
    const int blocks = device_cu_count() * 8;
    const uint32_t grid_threads = (uint32_t)blocks * kPtBlock;
    int rc = ensure_wavefront(r, batch * npix, grid_threads);
    if (rc) return rc;
    Wavefront& wf = r->wf;
    if (wf.iq_capacity < batch * npix) {
        for (int k = 0; k < 2; ++k) {
            if ((rc = wf.iq_o[k].ensure((size_t)batch * npix * 16)) || (rc = wf.iq_d[k].ensure((size_t)batch * npix * 16)))
                return rc;
        }
        wf.iq_capacity = batch * npix;
    }
    Counters* c = wf.counters.as<Counters>();
    CUPBR_HIP_CHECK(hipMemsetAsync(c, 0, sizeof(Counters), r->stream));
    uint32_t* spill = r->scene->stack_entries > (uint32_t)kPtLds ? wf.spill.as<uint32_t>() : nullptr;
    const DevScene& S = r->scene->dev;
    float* img = r->images[CUPBR_IMAGE_OUTPUT].as<float>();
    const uint32_t bounces = std::max<uint32_t>(1, r->max_trace_depth);
    auto trace = r->counting ? k_pt_trace<true> : k_pt_trace<false>;
    const bool packet = !r->counting && r->scene->stack_entries <= kPacketStack && S.n_nodes > 0;
    if ((rc = stage_mark(r, -1))) return rc;
    for (uint32_t s0 = 0; s0 < spp; s0 += batch) {
        const uint32_t nb = std::min(batch, spp - s0);
        const uint32_t npath = nb * npix;
        // the path tracer's ray generation (camera rays in SampleOrder), seed tea<4>(pixel, frame)
        // (VolumeRenderer.cu:229-234); rad[p].w = 0: outside every object
        // (VolumeRenderer.cu:229-234), made inside the camera trace with camera packets
        const SampleOrder order = sample_order(npix, nb);
        if (!packet) {
            hipLaunchKernelGGL(k_pt_raygen<uint32_t>, dim3((npath + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0,
                               r->stream, cam, w, h, frame0 + s0, npath, order, wf.q_o[0].as<float4>(),
                               wf.q_d[0].as<float4>(), wf.weight.as<float4>(), wf.radiance.as<float4>(), c, 0u);
            if ((rc = stage_mark(r, 3))) return rc;
        }
        for (uint32_t b = 0; b <= bounces; ++b) {
            const int cur = (int)(b & 1);
            if (b == 0 && packet) {
                const uint32_t nblk = (npath + kPtBlock - 1) / kPtBlock;
                hipLaunchKernelGGL(k_pt_camera<uint32_t>, dim3((nblk + 7u) / 8u * 8u), dim3(kPtBlock), 0, r->stream, S,
                                   npath, cam, w, h, frame0 + s0, order, 0u, wf.q_o[0].as<float4>(),
                                   wf.q_d[0].as<float4>(), wf.weight.as<float4>(), wf.radiance.as<float4>(),
                                   wf.q_hit.as<float2>(), c);
            } else {
                hipLaunchKernelGGL(trace, dim3(blocks), dim3(kPtBlock), 0, r->stream, S, cur, wf.q_o[cur].as<float4>(),
                                   wf.q_d[cur].as<float4>(), wf.q_hit.as<float2>(), wf.s_o[cur ^ 1].as<float4>(),
                                   wf.s_d[cur ^ 1].as<float4>(), wf.s_c[cur ^ 1].as<float4>(),
                                   wf.radiance.as<float4>(), c, spill, grid_threads);
            }
            if ((rc = stage_mark(r, 0))) return rc;
            if (b == bounces) break;
            hipLaunchKernelGGL(k_vol_shade, dim3(blocks), dim3(kPtBlock), 0, r->stream, S, cur, b, r->max_trace_depth,
                               (int)r->russian_roulette, wf.q_o[cur].as<float4>(),
                               wf.q_d[cur].as<float4>(),
                               wf.q_hit.as<float2>(), wf.iq_o[cur].as<float4>(), wf.iq_d[cur].as<float4>(),
                               wf.q_o[cur ^ 1].as<float4>(), wf.q_d[cur ^ 1].as<float4>(),
                               wf.iq_o[cur ^ 1].as<float4>(), wf.iq_d[cur ^ 1].as<float4>(), wf.s_o[cur].as<float4>(),
                               wf.s_d[cur].as<float4>(), wf.s_c[cur].as<float4>(), wf.weight.as<float4>(),
                               wf.radiance.as<float4>(), c, b == 0 && packet ? 0u : ~0u);
            if ((rc = stage_mark(r, 1))) return rc;
        }
        hipLaunchKernelGGL(k_accumulate, dim3((npix + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0, r->stream, img,
                           wf.radiance.as<float4>(), w, h, order, frame0 + s0, nb);
        if ((rc = stage_mark(r, 3))) return rc;
    }
    CUPBR_HIP_CHECK(hipGetLastError());
    return CUPBR_OK;
}

int gdpt_paths_wavefront(cupbr_renderer* r, const cupbr_camera& cam, uint32_t frame0, uint32_t spp) {
    const uint32_t w = r->width, h = r->height;
    if (w < 3 || h < 3 || spp == 0) return CUPBR_OK;  // no interior pixel: the buffers stay as they are
    const uint32_t n_int = (w - 2u) * (h - 2u);
    // paths in flight, in whole frames (5 per interior pixel), bounded as in the path tracer
    const uint32_t target = 1u << 27;
    uint32_t cap = 1u << 27;
    {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b > 0)
            cap = (uint32_t)std::min<size_t>(cap, std::max<size_t>(total_b / 8 / 200, 1u << 20));
    }
    const uint32_t per_frame = 5u * n_int;
    uint32_t batch = std::max<uint32_t>(1, target / per_frame);
    batch = std::min(batch, std::max<uint32_t>(1, cap / per_frame));
    batch = std::min(batch, spp);
    const int blocks = device_cu_count() * 8;
    const uint32_t grid_threads = (uint32_t)blocks * kPtBlock;
    int rc = ensure_wavefront(r, batch * per_frame, grid_threads);
    if (rc) return rc;
    Wavefront& wf = r->wf;
    if ((rc = wf.gdpt_slot.ensure((size_t)2 * batch * n_int * sizeof(uint32_t)))) return rc;
    uint32_t* slots[2] = {wf.gdpt_slot.as<uint32_t>(), wf.gdpt_slot.as<uint32_t>() + (size_t)batch * n_int};
    Counters* c = wf.counters.as<Counters>();
    CUPBR_HIP_CHECK(hipMemsetAsync(c, 0, sizeof(Counters), r->stream));
    uint32_t* spill = r->scene->stack_entries > (uint32_t)kPtLds ? wf.spill.as<uint32_t>() : nullptr;
    const DevScene& S = r->scene->dev;
    float* im[kNumImages];
    for (int k = 0; k < kNumImages; ++k) im[k] = r->images[k].as<float>();
    const uint32_t bounces = std::max<uint32_t>(1, r->max_trace_depth);
    auto trace = r->counting ? k_pt_trace<true> : k_pt_trace<false>;
    const bool packet = !r->counting && r->scene->stack_entries <= kPacketStack && S.n_nodes > 0;
    if ((rc = stage_mark(r, -1))) return rc;
    for (uint32_t s0 = 0; s0 < spp; s0 += batch) {
        const uint32_t nb = std::min(batch, spp - s0);
        const uint32_t groups = nb * n_int, npath = 5u * groups;
        if (!packet) {  // with camera packets the camera trace makes the rays (k_gdpt_camera)
            hipLaunchKernelGGL(k_gdpt_raygen, dim3((npath + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0, r->stream,
                               cam, w, h, frame0 + s0, n_int, groups, wf.q_o[0].as<float4>(), wf.q_d[0].as<float4>(),
                               wf.weight.as<float4>(), wf.radiance.as<float4>(), slots[0], c);
            if ((rc = stage_mark(r, 3))) return rc;
        }
        for (uint32_t b = 0; b <= bounces; ++b) {
            const int cur = (int)(b & 1);
            if (b == 0 && packet) {
                const uint32_t nblk = (npath + kPtBlock - 1) / kPtBlock;
                hipLaunchKernelGGL(k_gdpt_camera, dim3((nblk + 7u) / 8u * 8u), dim3(kPtBlock), 0, r->stream, S, npath,
                                   cam, w, h, frame0 + s0, n_int, wf.q_o[0].as<float4>(), wf.q_d[0].as<float4>(),
                                   wf.weight.as<float4>(),
                                   wf.radiance.as<float4>(), slots[0], wf.q_hit.as<float2>(), c);
            } else {
                hipLaunchKernelGGL(trace, dim3(blocks), dim3(kPtBlock), 0, r->stream, S, cur, wf.q_o[cur].as<float4>(),
                                   wf.q_d[cur].as<float4>(), wf.q_hit.as<float2>(), wf.s_o[cur ^ 1].as<float4>(),
                                   wf.s_d[cur ^ 1].as<float4>(), wf.s_c[cur ^ 1].as<float4>(),
                                   wf.radiance.as<float4>(), c, spill, grid_threads);
            }
            if ((rc = stage_mark(r, 0))) return rc;
            if (b == bounces) break;
            hipLaunchKernelGGL(k_gdpt_shade, dim3(blocks), dim3(kPtBlock), 0, r->stream, S, cur, b, r->max_trace_depth,
                               wf.q_o[cur].as<float4>(), wf.q_d[cur].as<float4>(),
                               wf.q_hit.as<float2>(), wf.q_o[cur ^ 1].as<float4>(), wf.q_d[cur ^ 1].as<float4>(),
                               wf.s_o[cur].as<float4>(), wf.s_d[cur].as<float4>(), wf.s_c[cur].as<float4>(),
                               wf.weight.as<float4>(), wf.radiance.as<float4>(), c, slots[b & 1],
                               slots[(b + 1) & 1], b == 0 && packet ? 0u : ~0u);
            if ((rc = stage_mark(r, 1))) return rc;
        }
        hipLaunchKernelGGL(k_gdpt_accumulate, dim3((n_int + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0, r->stream,
                           im[CUPBR_IMAGE_BASE], im[CUPBR_IMAGE_GRADIENT_X_FORWARD], im[CUPBR_IMAGE_GRADIENT_X_BACKWARD],
                           im[CUPBR_IMAGE_GRADIENT_Y_FORWARD], im[CUPBR_IMAGE_GRADIENT_Y_BACKWARD],
                           wf.radiance.as<float4>(), w, h, n_int, frame0 + s0, nb);
        if ((rc = stage_mark(r, 3))) return rc;
    }
    CUPBR_HIP_CHECK(hipGetLastError());
    return CUPBR_OK;
}

}  // namespace cupbr
