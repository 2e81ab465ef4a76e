// Closest-hit and any-hit traversal of the BVH4, plus the reference's
// primitive tests.
//
// Result semantics (what must match the reference bit for bit):
//   The reference walks its object BVH depth-first and keeps the first
//   strictly closer hit (lib/src/DataStructure/BoundingVolumeHierarchyDetail.h:11-101),
//   testing each object's own box with a biased slab test first
//   (lib/src/Geometry/AABBDetail.h:18-50). Inner boxes of that tree contain their
//   children's boxes, so the result is: among primitives whose object box passes
//   that biased test, the minimum of (t, reference rank), where rank is the
//   primitive's position in the reference DFS order (mesh triangles in index
//   order inside their mesh). This traversal computes exactly that minimum over
//   its own SAH BVH4: boxes are culled conservatively (slack on both ends of
//   the interval), and the biased object-box test runs lazily, only for a
//   candidate that would become the new best.
//
// Stack: a short per-lane stack in LDS (lane-minor, so a wave's pushes hit 64
// distinct banks), continued in a global spill area for deep trees.
#pragma once

#include <hip/hip_fp16.h>

#include "scene_types.h"

namespace cupbr {

// Triangle::computeRayIntersection, lib/src/Geometry/TriangleDetail.h:21-57.
// Returns true for a hit with t >= 0; NaN results report true with NaN t and are
// discarded by the callers' ordered comparisons, as in the reference.
__device__ __forceinline__ bool mt_test(v3 v1, v3 e1, v3 e2, v3 o, v3 d, float& t) {
    v3 pvec = cross(d, e2);
    float det = dot(e1, pvec);
    if (safe_float_equal(det, 0.0f)) return false;
    float inv_det = 1.0f / det;
    v3 tvec = o - v1;
    float u = dot(tvec, pvec) * inv_det;
    if (u < 0.0f || u > 1.0f) return false;
    v3 qvec = cross(tvec, e1);
    float v = dot(d, qvec) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot(e2, qvec) * inv_det;
    return !(t < 0.0f);
}

// Sphere::computeRayIntersection, lib/src/Geometry/SphereDetail.h:33-74
__device__ __forceinline__ bool sphere_test(v3 c, float r, v3 o, v3 d, float& t) {
    v3 os = o - c;
    float b = 2.0f * dot(d, os);
    float cc = dot(os, os) - r * r;
    float disc = b * b - 4.0f * cc;
    if (disc > 0) {
        float s = sqrtf(disc);
        float q = b < 0 ? (-b - s) / 2.0f : (-b + s) / 2.0f;
        float t0 = q, t1 = cc / q;
        float mn = fminf(t0, t1), mx = fmaxf(t0, t1);
        float tt = mn < 0 ? mx : mn;
        if (tt >= 0) { t = tt; return true; }
    }
    return false;
}

// Plane::computeRayIntersection (PlaneDetail.h:19-46) and Quad (QuadDetail.h:21-57)
__device__ __forceinline__ bool plane_quad_test(const UnbRec& u, v3 o, v3 d, float& t) {
    v3 p = mk(u.p.x, u.p.y, u.p.z), n = mk(u.n.x, u.n.y, u.n.z);
    float pdotn = dot(p, n), odotn = dot(o, n), ddotn = dot(d, n);
    if (safe_float_equal(ddotn, 0.0f)) return false;
    float tt = (pdotn - odotn) / ddotn;
    if (tt < 0) return false;
    if (__float_as_int(u.p.w) == kUnbQuad) {
        v3 e1 = mk(u.e1.x, u.e1.y, u.e1.z), e2 = mk(u.e2.x, u.e2.y, u.e2.z);
        v3 ip = o + tt * d;
        float x = dot(ip - p, e1), y = dot(ip - p, e2);
        float s1 = dot(e1, e1), s2 = dot(e2, e2);
        if (x < -s1 || x > s1 || y < -s2 || y > s2) return false;
    }
    t = tt;
    return true;
}

// AABB::hit, lib/src/Geometry/AABBDetail.h:18-50: 1/(d+1e-5) with no sign correction
__device__ __forceinline__ bool ref_aabb_hit(const float* mn, const float* mx, v3 o, v3 d) {
    bool hit = true;
    float tmin = 0.0f, tmax = __builtin_inff();
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float inv = 1.0f / (comp(d, a) + kEpsilon);
        float t0 = (mn[a] - comp(o, a)) * inv;
        float t1 = (mx[a] - comp(o, a)) * inv;
        if (inv < 0.0f) { float tmp = t0; t0 = t1; t1 = tmp; }
        tmin = t0 > tmin ? t0 : tmin;
        tmax = t1 < tmax ? t1 : tmax;
        if (tmax <= tmin) hit = false;
    }
    return hit;
}

struct TraceResult {
    float t;
    uint32_t prim;  // leaf-order prim index, kUnbBit|unbounded index, or kNoHit
    uint32_t rank;
};

// triangle or sphere test of one primitive record
__device__ __forceinline__ bool prim_test(const PrimRec& pr, v3 o, v3 d, float& t) {
    if ((__float_as_uint(pr.c.w) & 3u) == kPrimTriangle)
        return mt_test(mk(pr.a.x, pr.a.y, pr.a.z), mk(pr.a.w, pr.b.x, pr.b.y), mk(pr.b.z, pr.b.w, pr.c.x), o, d, t);
    return sphere_test(mk(pr.a.x, pr.a.y, pr.a.z), pr.a.w, o, d, t);
}

// the reference's biased test of the primitive's object box (finite boxes only: spheres, meshes)
__device__ __forceinline__ bool object_box_ok(const DevScene& S, const PrimRec& pr, v3 o, v3 d) {
    if (!(__float_as_uint(pr.c.w) & kPrimNeedsBoxCheck)) return true;
    const ObjInfo& ob = S.objs[__float_as_uint(pr.c.z)];
    return ref_aabb_hit(ob.bmin, ob.bmax, o, d);
}

// closest-hit step of one primitive (leaf-order index idx): best <- (t, idx, rank) when the
// primitive is hit at (t, rank) lexicographically below best and passes its object-box test
// (tested last, only for a candidate that would win)
__device__ __forceinline__ void closest_prim(const DevScene& S, const PrimRec& pr, uint32_t idx, v3 o, v3 d,
                                             TraceResult& best) {
    float t;
    if (!prim_test(pr, o, d, t)) return;
    const uint32_t rank = __float_as_uint(pr.c.y);
    if (t < best.t || (t == best.t && rank < best.rank)) {
        if (!object_box_ok(S, pr, o, d)) return;
        best.t = t;
        best.prim = idx;
        best.rank = rank;
    }
}

// Per-lane stack: entries [0, kLds) in LDS at lds[i * lds_stride], the rest in global
// memory at spill[(i - kLds) * spill_stride].
template <int kLds>
struct LaneStack {
    uint32_t* lds;
    uint32_t lds_stride;
    uint32_t* spill;
    uint32_t spill_stride;
    __device__ __forceinline__ void put(int i, uint32_t v) {
        if (i < kLds) lds[i * lds_stride] = v;
        else spill[(size_t)(i - kLds) * spill_stride] = v;
    }
    // The LDS read is unconditional (its address clamped into the LDS part) and the spill read a
    // relaxed atomic load, so that the compiler cannot merge the two into one generic (flat) load
    // of a selected pointer: a pop stays a ds_read.
    __device__ __forceinline__ uint32_t get(int i) const {
        uint32_t v = lds[min(i, kLds - 1) * lds_stride];
        if (i >= kLds)
            v = __hip_atomic_load(spill + (size_t)(i - kLds) * spill_stride, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_WAVEFRONT);
        return v;
    }
};

// visit counters of one lane (only updated by the counting instantiation)
struct TravCount {
    uint32_t nodes = 0, prims = 0;
};

// The primitives of one leaf (leaf word `word`): closest-hit updates of best, or (kAny) true as
// soon as one is hit within [0, dist].
template <bool kAny>
__device__ __forceinline__ bool visit_leaf(const DevScene& S, uint32_t word, v3 o, v3 d, float dist, TraceResult& best) {
    const uint32_t first = leaf_first(word), n_leaf = leaf_count(word);
    for (uint32_t k = 0; k < n_leaf; ++k) {
        const PrimRec pr = S.prims[first + k];
        if (kAny) {
            float t;
            if (!prim_test(pr, o, d, t) || !(t <= dist) || !object_box_ok(S, pr, o, d)) continue;
            return true;
        }
        closest_prim(S, pr, first + k, o, d, best);
    }
    return false;
}

// One inner node `cur`: the children whose boxes overlap [tlo, lim] (conservatively) are pushed
// farthest first, and cur becomes the nearest (closest hit) or a hit one (any hit); false when
// no child overlaps. The node is read in its NodeH form: a lane loads its reference point, the 16-byte block
// of each axis for the sign of its inverse direction (near planes of the 4 children, then far
// planes, as binary16 offsets) and the child words, and takes each plane's distance as one
// mixed-precision FMA, t = h (i 2^k) + (c - o) i. The decoded planes lie outside the FP32 ones by
// a margin that covers the decode's rounding (half_nodes), so the test stays conservative. An
// empty child slot has the box lo = +inf, hi = -inf, which no ray overlaps.
// Children are ordered by integer keys, the entry distance's float bits read as a signed integer
// (non-negative distances in their order, the slightly negative ones of boxes around the origin
// ahead of them), INT_MAX for a culled child; each compare-exchange of the network is a min, a max
// and two selects of the child words, with no branch. The order only decides which overlapping
// child is visited first; the closest hit is the same for any order.
template <bool kAny, int kLds>
__device__ __forceinline__ bool visit_inner(const DevScene& S, uint32_t& cur, v3 o, float ix, float iy, float iz,
                                            float lim, LaneStack<kLds>& st, int& sp) {
    const float tlo = -1e-5f;
    constexpr int kCulled = 0x7FFFFFFF;
    const char* base = reinterpret_cast<const char*>(S.hnodes);
    const uint32_t nb = cur * 128u;  // 32-bit byte offsets: scene_finalize keeps the nodes below 2^25
    const float4 co = *reinterpret_cast<const float4*>(base + nb);
    const uint4 px = *reinterpret_cast<const uint4*>(base + (nb + (ix >= 0.0f ? 16u : 32u)));
    const uint4 py = *reinterpret_cast<const uint4*>(base + (nb + (iy >= 0.0f ? 48u : 64u)));
    const uint4 pz = *reinterpret_cast<const uint4*>(base + (nb + (iz >= 0.0f ? 80u : 96u)));
    const uint4 ch = *reinterpret_cast<const uint4*>(base + (nb + 112u));
    int key[4];
    uint32_t val[4] = {ch.x, ch.y, ch.z, ch.w};
    // per axis: A = i 2^k (exact: k <= 0, and |i| >= 1 for a unit direction), B = (c - o) i
    const uint32_t ex = __float_as_uint(co.w);
    const float ax = __builtin_ldexpf(ix, (int)(int8_t)(ex & 0xFFu));
    const float ay = __builtin_ldexpf(iy, (int)(int8_t)((ex >> 8) & 0xFFu));
    const float az = __builtin_ldexpf(iz, (int)(int8_t)((ex >> 16) & 0xFFu));
    const float bx = (co.x - o.x) * ix, by = (co.y - o.y) * iy, bz = (co.z - o.z) * iz;
    auto lo16 = [](uint32_t w) { return __low2float(*reinterpret_cast<const __half2*>(&w)); };
    auto hi16 = [](uint32_t w) { return __high2float(*reinterpret_cast<const __half2*>(&w)); };
    // child i's near plane: half i of words 0..1 of its axis block; far plane: half i of words 2..3
#define CUPBR_CHILD(i, half, wn, wf)                                                               \
    {                                                                                              \
        const float tn = fmaxf(fmaxf(fmaf(half(px.wn), ax, bx), fmaf(half(py.wn), ay, by)),        \
                               fmaxf(fmaf(half(pz.wn), az, bz), tlo));                              \
        const float tf = fminf(fminf(fmaf(half(px.wf), ax, bx), fmaf(half(py.wf), ay, by)),        \
                               fminf(fmaf(half(pz.wf), az, bz), lim));                              \
        key[i] = tn <= tf * 1.000002f ? __float_as_int(tn) : kCulled;                              \
    }
    CUPBR_CHILD(0, lo16, x, z)
    CUPBR_CHILD(1, hi16, x, z)
    CUPBR_CHILD(2, lo16, y, w)
    CUPBR_CHILD(3, hi16, y, w)
#undef CUPBR_CHILD
    // The child words are only needed once a child overlaps; without this use the compiler moves
    // their load behind that test, and the step then waits for a second memory round trip.
    asm volatile("" ::"v"(val[0]), "v"(val[1]), "v"(val[2]), "v"(val[3]));
#define CUPBR_KSWAP(a, b)                                   \
    {                                                       \
        const bool sw = key[b] < key[a];                    \
        const int klo = min(key[a], key[b]);                \
        key[b] = max(key[a], key[b]);                       \
        key[a] = klo;                                       \
        const uint32_t va = val[a];                         \
        val[a] = sw ? val[b] : va;                          \
        val[b] = sw ? va : val[b];                          \
    }
    if (!kAny) {  // nearest child first; an any-hit query only needs the ones that were hit
        CUPBR_KSWAP(0, 1)
        CUPBR_KSWAP(2, 3)
        CUPBR_KSWAP(0, 2)
        CUPBR_KSWAP(1, 3)
        CUPBR_KSWAP(1, 2)
    } else {
        CUPBR_KSWAP(0, 1)  // the nearest hit child, if any, into slot 0
        CUPBR_KSWAP(0, 2)
        CUPBR_KSWAP(0, 3)
    }
#undef CUPBR_KSWAP
    if (key[0] == kCulled) return false;
    if (key[3] != kCulled) st.put(sp++, val[3]);
    if (key[2] != kCulled) st.put(sp++, val[2]);
    if (key[1] != kCulled) st.put(sp++, val[1]);
    cur = val[0];
    return true;
}

// kAny=false: closest hit (best.t = +inf, best.prim = kNoHit on a miss).
// kAny=true : any valid hit with t <= dist; returns whether one exists.
template <bool kAny, int kLds, bool kCount = false>
__device__ __forceinline__ bool traverse(const DevScene& S, v3 o, v3 d, float dist, LaneStack<kLds>& st,
                                         TraceResult& best, TravCount* cnt = nullptr) {
    best.t = __builtin_inff();
    best.prim = kNoHit;
    best.rank = 0xFFFFFFFFu;

    // planes and quads first: their hits tighten the interval for the BVH
    for (uint32_t i = 0; i < S.n_unb; ++i) {
        const UnbRec u = S.unb[i];
        float t;
        if (!plane_quad_test(u, o, d, t)) continue;
        if (kAny) {
            if (t <= dist) return true;
            continue;
        }
        uint32_t rank = __float_as_uint(u.n.w);
        if (t < best.t || (t == best.t && rank < best.rank)) {
            best.t = t;
            best.prim = kUnbBit | i;
            best.rank = rank;
        }
    }
    if (S.n_nodes == 0) return false;

    const float ix = 1.0f / d.x, iy = 1.0f / d.y, iz = 1.0f / d.z;
    int sp = 0;
    uint32_t cur = 0;
    for (;;) {
        if (cur & kLeafBit) {
            if (kCount) cnt->prims += leaf_count(cur);
            if (visit_leaf<kAny>(S, cur, o, d, dist, best)) return true;
        } else {
            if (kCount) cnt->nodes += 1;
            const float lim = kAny ? dist * 1.0001f + 1e-6f : best.t * 1.0001f + 1e-6f;
            if (visit_inner<kAny>(S, cur, o, ix, iy, iz, lim, st, sp)) continue;
        }
        if (sp == 0) break;
        cur = st.get(--sp);
    }
    return false;
}

// The same traversal as a resumable state, one node or leaf per step, for kernels that hand a
// new ray to a lane as soon as its previous one ends. best.t holds the query distance of an
// any-hit ray.
struct RayState {
    v3 o, d;
    float ix, iy, iz;
    TraceResult best;
    uint32_t cur;
    int sp;
};

// Starts a ray: planes and quads first (as in traverse). Returns whether BVH steps follow; when
// not, the result is final (closest hit in best; any hit in `hit`).
template <bool kAny>
__device__ __forceinline__ bool ray_begin(const DevScene& S, RayState& r, v3 o, v3 d, float dist, bool& hit) {
    r.o = o;
    r.d = d;
    r.best.t = kAny ? dist : __builtin_inff();
    r.best.prim = kNoHit;
    r.best.rank = 0xFFFFFFFFu;
    hit = false;
    for (uint32_t i = 0; i < S.n_unb; ++i) {
        const UnbRec u = S.unb[i];
        float t;
        if (!plane_quad_test(u, o, d, t)) continue;
        if (kAny) {
            if (t <= dist) {
                hit = true;
                return false;
            }
            continue;
        }
        const uint32_t rank = __float_as_uint(u.n.w);
        if (t < r.best.t || (t == r.best.t && rank < r.best.rank)) {
            r.best.t = t;
            r.best.prim = kUnbBit | i;
            r.best.rank = rank;
        }
    }
    r.ix = 1.0f / d.x;
    r.iy = 1.0f / d.y;
    r.iz = 1.0f / d.z;
    r.cur = 0;
    r.sp = 0;
    return S.n_nodes != 0;
}

// One inner node or one primitive of a started ray; false when its traversal has ended. A leaf
// takes one step per primitive, so that lanes at leaves of different sizes do not hold the wave
// for the largest one.
template <bool kAny, int kLds>
__device__ __forceinline__ bool ray_step(const DevScene& S, RayState& r, LaneStack<kLds>& st, bool& hit) {
    if (r.cur & kLeafBit) {
        // one primitive per step: the leaf word keeps the rest of the leaf (first + 1, count - 1)
        const uint32_t first = leaf_first(r.cur);
        const PrimRec pr = S.prims[first];
        if (kAny) {
            float t;
            if (prim_test(pr, r.o, r.d, t) && t <= r.best.t && object_box_ok(S, pr, r.o, r.d)) {
                hit = true;
                return false;
            }
        } else {
            closest_prim(S, pr, first, r.o, r.d, r.best);
        }
        if (leaf_count(r.cur) > 1u) {
            r.cur += 15u;  // first + 1 (bits 4..30), count - 1 (bits 0..3)
            return true;
        }
    } else {
        const float lim = r.best.t * 1.0001f + 1e-6f;  // the query distance of an any-hit ray
        if (visit_inner<kAny>(S, r.cur, r.o, r.ix, r.iy, r.iz, lim, st, r.sp)) return true;
    }
    if (r.sp == 0) return false;
    r.cur = st.get(--r.sp);
    return true;
}

// a load whose address is the same in every lane, issued as a scalar load (records of the
// BVH and primitive arrays, which no kernel writes while it traverses them)
__device__ __forceinline__ float4 uniform_load4(const void* p) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 v = *(const __attribute__((address_space(4))) f4*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ PrimRec uniform_prim(const PrimRec* p) {
    return PrimRec{uniform_load4(&p->a), uniform_load4(&p->b), uniform_load4(&p->c)};
}

// Entries of the wave-uniform stack of traverse_packet: entry k is held by lane k % 64 of
// register k / 64.
constexpr uint32_t kPacketStack = 128;

// Closest hit for the 64 rays of a wave traversed as one packet (rays with a common origin and
// nearby directions, such as the camera rays of an 8x8 pixel tile). Every node and leaf is
// visited by the whole wave: its record is a wave-uniform scalar load, each lane tests its own
// ray, and a child is entered when its box overlaps the current interval of any lane (the same
// conservative test as traverse), nearest first by the entry distance of the first such lane.
// Every lane therefore tests a superset of the primitives its own traversal would test, none
// outside the object boxes' candidates, and the closest hit, a minimum of (t, reference rank)
// over them, is the same. Lanes with valid = false take part in the traversal (the stack is
// one value per wave, spread over the lanes) but have an empty interval and keep no hit.
// Requires the BVH's worst-case stack (3 entries per BVH4 level + 1) <= kPacketStack and the
// whole wave active. kUniformSigns: every valid lane's direction has the signs (ux, uy, uz), so
// the near and far planes of a box are chosen once per wave.
template <bool kUniformSigns>
__device__ __forceinline__ void traverse_packet_impl(const DevScene& S, v3 o, v3 d, bool valid, bool ux, bool uy,
                                                     bool uz, TraceResult& best) {
    const float inf = __builtin_inff();
    best.t = valid ? inf : -1.0f;
    best.prim = kNoHit;
    best.rank = 0xFFFFFFFFu;
    for (uint32_t i = 0; i < S.n_unb; ++i) {  // planes and quads, as in traverse
        const UnbRec u = S.unb[i];
        float t;
        if (!valid || !plane_quad_test(u, o, d, t)) continue;
        const uint32_t rank = __float_as_uint(u.n.w);
        if (t < best.t || (t == best.t && rank < best.rank)) {
            best.t = t;
            best.prim = kUnbBit | i;
            best.rank = rank;
        }
    }
    if (S.n_nodes == 0) return;

    const float ix = 1.0f / d.x, iy = 1.0f / d.y, iz = 1.0f / d.z;
    const bool sx = kUniformSigns ? ux : ix >= 0.0f, sy = kUniformSigns ? uy : iy >= 0.0f,
               sz = kUniformSigns ? uz : iz >= 0.0f;
    const float tlo = -1e-5f;
    // byte offsets of the NodeH plane blocks for the wave's signs (near planes of the 4 children,
    // then far planes); the mixed-sign instantiation loads both blocks of an axis and selects
    const uint32_t obx = __builtin_amdgcn_readfirstlane(sx ? 16 : 32);
    const uint32_t oby = __builtin_amdgcn_readfirstlane(sy ? 48 : 64);
    const uint32_t obz = __builtin_amdgcn_readfirstlane(sz ? 80 : 96);
    const uint32_t lane = __lane_id();
    uint32_t stk0 = 0, stk1 = 0;  // the wave's stack
    uint32_t sp = 0, cur = 0;     // wave-uniform
    for (;;) {
        if (cur & kLeafBit) {
            const uint32_t first = leaf_first(cur), n_leaf = leaf_count(cur);
            for (uint32_t k = 0; k < n_leaf; ++k) closest_prim(S, uniform_prim(S.prims + first + k), first + k, o, d, best);
        } else {
            // the node in its NodeH form, as in visit_inner: plane t = h (i 2^k) + (c - o) i
            const char* hb = reinterpret_cast<const char*>(S.hnodes) + (size_t)cur * 128u;
            const float4 co = uniform_load4(hb);
            const float4 cw = uniform_load4(hb + 112);
            const uint32_t child[4] = {__float_as_uint(cw.x), __float_as_uint(cw.y), __float_as_uint(cw.z),
                                       __float_as_uint(cw.w)};
            float4 BX, BY, BZ;  // near planes of the 4 children in words 0..1, far planes in 2..3
            // (below, nx .. fz are the distances to those planes)
            if (kUniformSigns) {
                BX = uniform_load4(hb + obx), BY = uniform_load4(hb + oby), BZ = uniform_load4(hb + obz);
            } else {
                const float4 x0 = uniform_load4(hb + 16), x1 = uniform_load4(hb + 32);
                const float4 y0 = uniform_load4(hb + 48), y1 = uniform_load4(hb + 64);
                const float4 z0 = uniform_load4(hb + 80), z1 = uniform_load4(hb + 96);
                BX = sx ? x0 : x1, BY = sy ? y0 : y1, BZ = sz ? z0 : z1;
            }
            const uint32_t ex = __float_as_uint(co.w);
            const float ax = __builtin_ldexpf(ix, (int)(int8_t)(ex & 0xFFu));
            const float ay = __builtin_ldexpf(iy, (int)(int8_t)((ex >> 8) & 0xFFu));
            const float az = __builtin_ldexpf(iz, (int)(int8_t)((ex >> 16) & 0xFFu));
            const float bx = (co.x - o.x) * ix, by = (co.y - o.y) * iy, bz = (co.z - o.z) * iz;
            auto lo16 = [](float w) { const uint32_t u = __float_as_uint(w); return __low2float(*reinterpret_cast<const __half2*>(&u)); };
            auto hi16 = [](float w) { const uint32_t u = __float_as_uint(w); return __high2float(*reinterpret_cast<const __half2*>(&u)); };
            const float nx[4] = {fmaf(lo16(BX.x), ax, bx), fmaf(hi16(BX.x), ax, bx), fmaf(lo16(BX.y), ax, bx),
                                 fmaf(hi16(BX.y), ax, bx)};
            const float fx[4] = {fmaf(lo16(BX.z), ax, bx), fmaf(hi16(BX.z), ax, bx), fmaf(lo16(BX.w), ax, bx),
                                 fmaf(hi16(BX.w), ax, bx)};
            const float ny[4] = {fmaf(lo16(BY.x), ay, by), fmaf(hi16(BY.x), ay, by), fmaf(lo16(BY.y), ay, by),
                                 fmaf(hi16(BY.y), ay, by)};
            const float fy[4] = {fmaf(lo16(BY.z), ay, by), fmaf(hi16(BY.z), ay, by), fmaf(lo16(BY.w), ay, by),
                                 fmaf(hi16(BY.w), ay, by)};
            const float nz[4] = {fmaf(lo16(BZ.x), az, bz), fmaf(hi16(BZ.x), az, bz), fmaf(lo16(BZ.y), az, bz),
                                 fmaf(hi16(BZ.y), az, bz)};
            const float fz[4] = {fmaf(lo16(BZ.z), az, bz), fmaf(hi16(BZ.z), az, bz), fmaf(lo16(BZ.w), az, bz),
                                 fmaf(hi16(BZ.w), az, bz)};
            const float lim = best.t * 1.0001f + 1e-6f;  // < tlo for an invalid lane: nothing overlaps
            // ordering key of a child: the entry distance of the first lane that overlaps it, as
            // float bits (negative distances as 0, so the keys order as unsigned integers on the
            // scalar unit); 0xFFFFFFFF when no lane overlaps it
            uint32_t key[4], val[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const float tn = fmaxf(fmaxf(nx[c], ny[c]), fmaxf(nz[c], tlo));
                const float tf = fminf(fminf(fx[c], fy[c]), fminf(fz[c], lim));
                const unsigned long long m = child[c] != kChildEmpty ? __ballot(tn <= tf * 1.000002f) : 0ull;
                uint32_t k = 0xFFFFFFFFu;
                if (m) {
                    k = (uint32_t)__builtin_amdgcn_readlane(__float_as_int(tn), (int)__builtin_ctzll(m));
                    k = (k >> 31) ? 0u : k;
                }
                key[c] = k;
                val[c] = child[c];
            }
#define CUPBR_PSWAP(a, b)                                                  \
    if (key[b] < key[a]) {                                                  \
        const uint32_t tk = key[a]; key[a] = key[b]; key[b] = tk;           \
        const uint32_t tv = val[a]; val[a] = val[b]; val[b] = tv;           \
    }
            CUPBR_PSWAP(0, 1)
            CUPBR_PSWAP(2, 3)
            CUPBR_PSWAP(0, 2)
            CUPBR_PSWAP(1, 3)
            CUPBR_PSWAP(1, 2)
#undef CUPBR_PSWAP
            if (key[0] != 0xFFFFFFFFu) {
#pragma unroll
                for (int c = 3; c >= 1; --c) {
                    if (key[c] == 0xFFFFFFFFu) continue;
                    if (sp < 64) stk0 = lane == sp ? val[c] : stk0;
                    else stk1 = lane == sp - 64 ? val[c] : stk1;
                    ++sp;
                }
                cur = val[0];
                continue;
            }
        }
        if (sp == 0) break;
        --sp;
        cur = (uint32_t)__builtin_amdgcn_readlane((int)(sp < 64 ? stk0 : stk1), (int)(sp & 63u));
    }
}

__device__ __forceinline__ void traverse_packet(const DevScene& S, v3 o, v3 d, bool valid, TraceResult& best) {
    // an invalid lane's interval is empty whatever planes it is given, so only valid lanes vote
    const unsigned long long v = __ballot(valid);
    const unsigned long long px = __ballot(valid && 1.0f / d.x >= 0.0f), py = __ballot(valid && 1.0f / d.y >= 0.0f),
                             pz = __ballot(valid && 1.0f / d.z >= 0.0f);
    if ((px == 0 || px == v) && (py == 0 || py == v) && (pz == 0 || pz == v))
        traverse_packet_impl<true>(S, o, d, valid, px != 0, py != 0, pz != 0, best);
    else
        traverse_packet_impl<false>(S, o, d, valid, false, false, false, best);
}

// Tracing::traceRay(scene, ray, index), TracingDetail.h:68-116: one object's primitives in order,
// no box test; a mesh keeps its first strictly-closer triangle (MeshDetail.h:28-35)
__device__ __forceinline__ bool object_trace(const DevScene& S, uint32_t obj, v3 o, v3 d, TraceResult& r) {
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

// object id and triangle-in-mesh index of a closest-hit result
__device__ __forceinline__ void hit_ids(const DevScene& S, const TraceResult& r, int32_t& obj, int32_t& tri) {
    if (r.prim == kNoHit) { obj = -1; tri = -1; return; }
    if (r.prim & kUnbBit) {
        obj = (int32_t)__float_as_uint(S.unb[r.prim & ~kUnbBit].e1.w);
        tri = -1;
        return;
    }
    obj = (int32_t)__float_as_uint(S.prims[r.prim].c.z);
    const ObjInfo& ob = S.objs[obj];
    tri = ob.type == CUPBR_MESH ? (int32_t)(r.rank - ob.rank_base) : -1;
}

}  // namespace cupbr
