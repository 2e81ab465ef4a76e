// ORACLE — test infrastructure only (see cpu_ref.hpp). Exposes a C ABI for
// the Python tests (ctypes) and for bench.py's cpu_baseline leg.
//
// Scene traversal comes in three forms that must agree exactly:
//   mode 0  the reference's own object BVH: median split on AABB min.x,
//           DFS left-first with the biased slab test on every node
//           (lib/src/DataStructure/BoundingVolumeHierarchy.cpp:46-136,
//            BoundingVolumeHierarchyDetail.h:11-101);
//   mode 1  the same objects visited in that tree's leaf order with the
//           biased slab test applied to each object's own box only;
//   mode 2  a binned-SAH binary BVH over single primitives with the
//           (t, reference-rank) tie-break and lazy object-box checks.
// Mode 0 is the literal restatement; modes 1 and 2 are checked against it.
#include "cpu_ref.hpp"

#include <atomic>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>

namespace orc {

int g_rng_pcg = 0;

// ------------------------------------------------------------------ scene
template <typename R>
struct RefNode {  // BoundingVolumeHierarchy node: box + two children (>=0 node, <0 object ~idx)
    V3<R> bmin, bmax;
    int left, right;
};
struct Prim {  // one ranked primitive: an object, or one triangle of a mesh
    uint32_t rank;
    int obj, tri;  // tri = -1 unless a mesh triangle
};
template <typename R>
struct ANode {
    V3<R> bmin, bmax;
    uint32_t a, b, count;  // count==0: inner node with children a, b; else leaf prims [a, a+count)
};

template <typename R>
struct SceneT {
    std::vector<Object<R>> objs;
    std::vector<Tri<R>> tris;
    std::vector<Light<R>> lights;
    bool use_env = false;
    uint32_t env_w = 0, env_h = 0;
    std::vector<V3<R>> env;
    Volume<R> volume;
    std::vector<RefNode<R>> ref_nodes;  // mode 0
    int ref_root = -1;
    std::vector<int> leaf_order;        // mode 1
    std::vector<uint32_t> rank_base;
    std::vector<Prim> prims;            // mode 2 (bounded prims in leaf order)
    std::vector<Prim> unbounded;        // planes and quads, tested for every ray
    std::vector<ANode<R>> anodes;
};

template <typename R>
static inline void obj_box(const Object<R>& o, V3<R>& mn, V3<R>& mx) {
    if (o.finite_box) { mn = o.bmin; mx = o.bmax; }
    else { mn = V3<R>(-std::numeric_limits<R>::infinity()); mx = V3<R>(std::numeric_limits<R>::infinity()); }
}

// BoundingVolumeHierarchy.cpp:46-136. box_compare on axis 0: the reference draws
// axis = (rand()/RAND_MAX)%3, which is 0 unless rand()==RAND_MAX (:50).
template <typename R>
static int build_ref(SceneT<R>& S, const std::vector<int>& objs) {
    auto key = [&](int o) { V3<R> mn, mx; obj_box(S.objs[o], mn, mx); return mn.x; };
    auto less = [&](int a, int b) { return key(a) < key(b); };
    int left, right;
    size_t n = objs.size();
    if (n == 1) {
        left = right = ~objs[0];
    } else if (n == 2) {  // :66-86 — equal keys store the pair swapped
        if (less(objs[0], objs[1])) { left = ~objs[0]; right = ~objs[1]; }
        else { left = ~objs[1]; right = ~objs[0]; }
    } else {  // :89-121
        std::vector<size_t> idx(n);
        for (size_t i = 0; i < n; ++i) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return less(objs[a], objs[b]); });
        size_t mid = n / 2;
        std::vector<int> lo, hi;
        for (size_t i = 0; i < n; ++i) (i < mid ? lo : hi).push_back(objs[idx[i]]);
        left = build_ref(S, lo);
        right = build_ref(S, hi);
    }
    auto child_box = [&](int c, V3<R>& mn, V3<R>& mx) {
        if (c < 0) obj_box(S.objs[~c], mn, mx);
        else { mn = S.ref_nodes[c].bmin; mx = S.ref_nodes[c].bmax; }
    };
    V3<R> a0, a1, b0, b1;
    child_box(left, a0, a1);
    child_box(right, b0, b1);
    RefNode<R> node;  // surrounding_box, :31-42
    node.bmin = V3<R>(std::fmin(a0.x, b0.x), std::fmin(a0.y, b0.y), std::fmin(a0.z, b0.z));
    node.bmax = V3<R>(std::fmax(a1.x, b1.x), std::fmax(a1.y, b1.y), std::fmax(a1.z, b1.z));
    node.left = left;
    node.right = right;
    S.ref_nodes.push_back(node);
    return (int)S.ref_nodes.size() - 1;
}

template <typename R>
static void ref_leaf_order(const SceneT<R>& S, int node, std::vector<int>& out) {
    const RefNode<R>& n = S.ref_nodes[node];
    auto visit = [&](int c) { if (c < 0) out.push_back(~c); else ref_leaf_order(S, c, out); };
    visit(n.left);
    if (n.right != n.left) visit(n.right);  // BoundingVolumeHierarchyDetail.h:83-85
}

// ------------------------------------------------------------------ primitive tests
template <typename R>
static inline bool prim_t(const SceneT<R>& S, int obj, int tri, const V3<R>& o, const V3<R>& d, R& t) {
    const Object<R>& ob = S.objs[obj];
    switch (ob.type) {
        case SPHERE: return sphere_intersect(ob.a, ob.radius, o, d, t);
        case PLANE: return plane_intersect(ob.a, ob.b, o, d, t);
        case QUAD: return quad_intersect(ob.a, ob.b, ob.c, ob.d, o, d, t);
        case TRIANGLE: return tri_intersect(ob.a, ob.b, ob.c, o, d, t);
        case MESH: { const Tri<R>& tr = S.tris[ob.first + tri]; return tri_intersect(tr.v1, tr.v2, tr.v3, o, d, t); }
    }
    return false;
}
template <typename R>
static inline void fill_hit(const SceneT<R>& S, Hit<R>& h, const V3<R>& o, const V3<R>& d) {
    const Object<R>& ob = S.objs[h.obj];
    h.P = o + h.depth * d;
    switch (ob.type) {
        case SPHERE: h.N = normalize(h.P - ob.a); break;
        case PLANE:
        case QUAD: h.N = ob.b; break;
        case TRIANGLE: h.N = ob.tri_normal; break;
        case MESH: h.N = S.tris[ob.first + h.tri].n; break;
    }
}
// Geometry::computeRayIntersection per type; a mesh loops its triangles with strict '<'
// (MeshDetail.h:21-40) and the caller keeps the first strictly-closer hit.
template <typename R>
static inline void object_hit(const SceneT<R>& S, int obj, const V3<R>& o, const V3<R>& d, Hit<R>& best) {
    const Object<R>& ob = S.objs[obj];
    if (ob.type == MESH) {
        R bt = std::numeric_limits<R>::infinity();
        int bi = -1;
        for (uint32_t i = 0; i < ob.count; ++i) {
            R t;
            if (prim_t(S, obj, (int)i, o, d, t) && t < bt) { bt = t; bi = (int)i; }
        }
        if (bi >= 0 && bt < best.depth) { best.depth = bt; best.obj = obj; best.tri = bi; }
        return;
    }
    R t;
    if (prim_t(S, obj, -1, o, d, t) && t < best.depth) { best.depth = t; best.obj = obj; best.tri = -1; }
}

// mode 0: BoundingVolumeHierarchyDetail.h:11-101
template <typename R>
static Hit<R> trace_tree(const SceneT<R>& S, const V3<R>& o, const V3<R>& d) {
    Hit<R> best;
    if (S.ref_root < 0) return best;
    std::vector<int> stack;
    stack.reserve(64);
    stack.push_back(S.ref_root);
    while (!stack.empty()) {
        int cur = stack.back();
        stack.pop_back();
        for (;;) {
            V3<R> mn, mx;
            if (cur < 0) obj_box(S.objs[~cur], mn, mx);
            else { mn = S.ref_nodes[cur].bmin; mx = S.ref_nodes[cur].bmax; }
            if (!aabb_hit(mn, mx, o, d)) break;
            if (cur < 0) { object_hit(S, ~cur, o, d, best); break; }
            const RefNode<R>& n = S.ref_nodes[cur];
            if (n.right != n.left) stack.push_back(n.right);
            cur = n.left;
        }
    }
    if (best.obj >= 0) fill_hit(S, best, o, d);
    return best;
}
// mode 1: leaf order, object boxes only
template <typename R>
static Hit<R> trace_flat(const SceneT<R>& S, const V3<R>& o, const V3<R>& d) {
    Hit<R> best;
    for (int obj : S.leaf_order) {
        const Object<R>& ob = S.objs[obj];
        if (ob.finite_box && !aabb_hit(ob.bmin, ob.bmax, o, d)) continue;
        object_hit(S, obj, o, d, best);
    }
    if (best.obj >= 0) fill_hit(S, best, o, d);
    return best;
}

// ------------------------------------------------------------------ mode 2: binned-SAH BVH
template <typename R>
static void prim_bounds(const SceneT<R>& S, const Prim& p, V3<R>& mn, V3<R>& mx) {
    const Object<R>& ob = S.objs[p.obj];
    if (ob.type == SPHERE) { mn = ob.a - V3<R>(ob.radius); mx = ob.a + V3<R>(ob.radius); return; }
    V3<R> a, b, c;
    if (ob.type == MESH) { const Tri<R>& t = S.tris[ob.first + p.tri]; a = t.v1; b = t.v2; c = t.v3; }
    else { a = ob.a; b = ob.b; c = ob.c; }
    for (int k = 0; k < 3; ++k) {
        mn[k] = std::fmin(a[k], std::fmin(b[k], c[k]));
        mx[k] = std::fmax(a[k], std::fmax(b[k], c[k]));
    }
}

template <typename R>
struct SahBuilder {
    SceneT<R>& S;
    std::vector<V3<R>> pmin, pmax, cen;
    explicit SahBuilder(SceneT<R>& s) : S(s) {}
    static R area(const V3<R>& mn, const V3<R>& mx) {
        V3<R> e = mx - mn;
        if (e.x < 0) return 0;
        return (R)2 * (e.x * e.y + e.y * e.z + e.z * e.x);
    }
    void swap_prim(uint32_t i, uint32_t j) {
        std::swap(S.prims[i], S.prims[j]);
        std::swap(pmin[i], pmin[j]);
        std::swap(pmax[i], pmax[j]);
        std::swap(cen[i], cen[j]);
    }
    uint32_t build(uint32_t lo, uint32_t hi, int depth = 0) {
        const R inf = std::numeric_limits<R>::infinity();
        uint32_t me = (uint32_t)S.anodes.size();
        S.anodes.push_back(ANode<R>());
        V3<R> bmn(inf), bmx(-inf), cmn(inf), cmx(-inf);
        for (uint32_t i = lo; i < hi; ++i)
            for (int a = 0; a < 3; ++a) {
                bmn[a] = std::fmin(bmn[a], pmin[i][a]); bmx[a] = std::fmax(bmx[a], pmax[i][a]);
                cmn[a] = std::fmin(cmn[a], cen[i][a]); cmx[a] = std::fmax(cmx[a], cen[i][a]);
            }
        S.anodes[me].bmin = bmn;
        S.anodes[me].bmax = bmx;
        uint32_t n = hi - lo;
        if (n <= 4) { S.anodes[me].a = lo; S.anodes[me].count = n; return me; }
        const int NB = 16;
        int best_axis = -1, best_split = -1;
        R best_cost = inf;
        for (int a = 0; a < 3; ++a) {
            R ext = cmx[a] - cmn[a];
            if (!(ext > 0)) continue;
            uint32_t cnt[NB] = {0};
            V3<R> bn[NB], bx[NB];
            for (int b = 0; b < NB; ++b) { bn[b] = V3<R>(inf); bx[b] = V3<R>(-inf); }
            for (uint32_t i = lo; i < hi; ++i) {
                int b = std::min(NB - 1, (int)((cen[i][a] - cmn[a]) / ext * NB));
                cnt[b]++;
                for (int k = 0; k < 3; ++k) { bn[b][k] = std::fmin(bn[b][k], pmin[i][k]); bx[b][k] = std::fmax(bx[b][k], pmax[i][k]); }
            }
            R rarea[NB];
            uint32_t rcnt[NB];
            V3<R> rn(inf), rx(-inf);
            uint32_t rc = 0;
            for (int b = NB - 1; b > 0; --b) {
                rc += cnt[b];
                for (int k = 0; k < 3; ++k) { rn[k] = std::fmin(rn[k], bn[b][k]); rx[k] = std::fmax(rx[k], bx[b][k]); }
                rarea[b] = area(rn, rx);
                rcnt[b] = rc;
            }
            V3<R> ln(inf), lx(-inf);
            uint32_t lc = 0;
            for (int b = 0; b < NB - 1; ++b) {
                lc += cnt[b];
                for (int k = 0; k < 3; ++k) { ln[k] = std::fmin(ln[k], bn[b][k]); lx[k] = std::fmax(lx[k], bx[b][k]); }
                if (lc == 0 || rcnt[b + 1] == 0) continue;
                R cost = area(ln, lx) * (R)lc + rarea[b + 1] * (R)rcnt[b + 1];
                if (cost < best_cost) { best_cost = cost; best_axis = a; best_split = b; }
            }
        }
        uint32_t mid = lo + n / 2;
        if (depth > 160) best_axis = -1;  // keeps the tree shallower than the traversal stack
        if (best_axis >= 0) {
            R ext = cmx[best_axis] - cmn[best_axis];
            uint32_t i = lo, j = hi;
            while (i < j) {
                int b = std::min(NB - 1, (int)((cen[i][best_axis] - cmn[best_axis]) / ext * NB));
                if (b <= best_split) ++i;
                else { --j; swap_prim(i, j); }
            }
            if (i != lo && i != hi) mid = i;
        }
        uint32_t l = build(lo, mid, depth + 1);
        uint32_t r = build(mid, hi, depth + 1);
        S.anodes[me].a = l;
        S.anodes[me].b = r;
        S.anodes[me].count = 0;
        return me;
    }
};

// Slab test for the SAH BVH: the near/far plane is chosen by the sign of 1/d, and fmin/fmax drop
// the NaN of 0*inf, leaving that axis unconstrained (a ray lying in a slab plane). The upper end is
// widened by a relative 1e-5 so rounding in the interval never culls a primitive hit.
template <typename R>
static inline bool slab(const V3<R>& mn, const V3<R>& mx, const V3<R>& o, const V3<R>& inv, R tlo, R thi) {
    R t0 = tlo, t1 = thi;
    for (int a = 0; a < 3; ++a) {
        R nb = inv[a] >= 0 ? mn[a] : mx[a];
        R fb = inv[a] >= 0 ? mx[a] : mn[a];
        t0 = std::fmax(t0, (nb - o[a]) * inv[a]);
        t1 = std::fmin(t1, (fb - o[a]) * inv[a]);
    }
    return t0 <= t1 * ((R)1 + (R)1e-5);
}
template <typename R>
static inline bool lazy_box_ok(const SceneT<R>& S, int obj, const V3<R>& o, const V3<R>& d) {
    const Object<R>& ob = S.objs[obj];
    return !ob.finite_box || aabb_hit(ob.bmin, ob.bmax, o, d);
}

// Closest hit: lexicographic minimum of (t, reference rank) over primitives whose object box passes
// the reference's biased slab test; equal to the first strictly-closer hit in reference DFS order.
// Any hit (anyhit=true): any primitive with t <= dist whose object box passes.
template <typename R>
static Hit<R> trace_accel(const SceneT<R>& S, const V3<R>& o, const V3<R>& d, bool anyhit, R dist, bool* occluded) {
    Hit<R> best;
    uint32_t best_rank = 0xFFFFFFFFu;
    const R inf = std::numeric_limits<R>::infinity();
    bool found = false;
    auto consider = [&](const Prim& p) {
        R t;
        if (!prim_t(S, p.obj, p.tri, o, d, t)) return;
        if (anyhit) {
            if (t <= dist && lazy_box_ok(S, p.obj, o, d)) found = true;
            return;
        }
        if (!(t < inf)) return;
        if (t < best.depth || (t == best.depth && p.rank < best_rank)) {
            if (!lazy_box_ok(S, p.obj, o, d)) return;
            best.depth = t; best.obj = p.obj; best.tri = p.tri; best_rank = p.rank;
        }
    };
    for (const Prim& p : S.unbounded) {
        consider(p);
        if (found) break;
    }
    if (!found && !S.anodes.empty()) {
        V3<R> inv((R)1 / d.x, (R)1 / d.y, (R)1 / d.z);
        uint32_t stack[256];
        int sp = 0;
        stack[sp++] = 0;
        while (sp && !found) {
            const ANode<R>& n = S.anodes[stack[--sp]];
            R thi = anyhit ? dist : best.depth;
            thi = thi * ((R)1 + (R)1e-4) + (R)1e-6;
            if (!slab(n.bmin, n.bmax, o, inv, (R)-1e-6, thi)) continue;
            if (n.count) {
                for (uint32_t i = 0; i < n.count && !found; ++i) consider(S.prims[n.a + i]);
            } else {
                stack[sp++] = n.b;
                stack[sp++] = n.a;
            }
        }
    }
    if (occluded) *occluded = found;
    if (!anyhit && best.obj >= 0) fill_hit(S, best, o, d);
    return best;
}

// ------------------------------------------------------------------ tracer front-end
template <typename R>
struct Tracer {
    const SceneT<R>& S;
    int mode;
    Hit<R> closest(const V3<R>& o, const V3<R>& d) const {
        if (mode == 0) return trace_tree(S, o, d);
        if (mode == 1) return trace_flat(S, o, d);
        return trace_accel(S, o, d, false, (R)0, nullptr);
    }
    // Tracing::traceVisibility, TracingDetail.h:118-183
    bool visible(const V3<R>& o, const V3<R>& d, R dist) const {
        if (mode == 2) {
            bool occ = false;
            trace_accel(S, o, d, true, dist, &occ);
            return !occ;
        }
        Hit<R> h = closest(o, d);
        return h.depth > dist || h.depth == std::numeric_limits<R>::infinity();
    }
    // Tracing::traceRay(scene, ray, index), TracingDetail.h:68-116 — one object, no box test
    Hit<R> object(int obj, const V3<R>& o, const V3<R>& d) const {
        Hit<R> best;
        object_hit(S, obj, o, d, best);
        if (best.obj >= 0) fill_hit(S, best, o, d);
        return best;
    }
};

template <typename R>
static inline V3<R> env_lookup(const SceneT<R>& S, const V3<R>& d) {
    uint32_t x, y;
    direction2uv(d, S.env_w, S.env_h, x, y);
    if (x >= S.env_w) x = S.env_w - 1;
    if (y >= S.env_h) y = S.env_h - 1;
    return S.env[x + y * S.env_w];
}

// Light choice + sample shared by every NEE site (PathTracer.cu:43-62 and copies).
// Returns false when the pick lands on the environment slot of a scene without one.
template <typename R>
static inline bool nee_sample(const SceneT<R>& S, const Material<R>& mat, uint32_t& seed, const V3<R>& P, const V3<R>& inc,
                              const V3<R>& N, V3<R>& dir, V3<R>& Le, R& dist, bool& degenerate) {
    degenerate = false;
    uint32_t useEnv = S.use_env ? 1u : 0u;
    uint32_t nl = (uint32_t)S.lights.size();
    uint32_t pick = (uint32_t)(rnd(seed) * (float)(nl + useEnv));
    if (pick != nl) {
        if (pick > nl) pick = nl - 1;
        Le = light_sample(S.lights[pick], seed, P, dir, dist);
        return true;
    }
    if (!S.use_env) return false;
    V4<R> s = sampleDirection(mat, seed, inc, N);
    dir = s.v;
    dist = std::numeric_limits<R>::infinity();
    Le = env_lookup(S, dir) / s.w;
    // a failed BSDF sample returns the zero vector; its shadow ray would carry NaN
    degenerate = !(norm(dir) > (R)0);
    return true;
}

// ------------------------------------------------------------------ PT + NEE
// pathtracer_kernel_nee, lib/src/Renderer/PathTracer.cu:97-169 (returns the sample radiance;
// the running mean of :162-168 is applied by the caller)
template <typename R>
static V3<R> pt_sample(const SceneT<R>& S, const Tracer<R>& T, const Camera<R>& cam, uint32_t tid, uint32_t w, uint32_t h,
                       uint32_t frame, uint32_t maxDepth, bool rr) {
    uint32_t seed = g_rng_pcg ? pcg_hash(pcg_hash(frame) ^ tid) : tea4(tid, frame);
    V3<R> o, d;
    launch_ray(tid % w, tid / w, w, h, cam, seed, o, d);
    V3<R> radiance(0), weight(1), out_dir;
    uint32_t depth = 0;
    do {
        bool next_valid = false;
        Hit<R> g = T.closest(o, d);
        if (!(g.depth < std::numeric_limits<R>::infinity())) {  // miss, :127-135
            if (S.use_env) radiance += weight * env_lookup(S, d);
            break;
        }
        const Material<R>& mat = S.objs[g.obj].mat;
        V3<R> inc = (R)-1.0f * d;
        radiance += weight * mat.albedo_e;  // emissiveIllumintation :20-26
        // directIllumination :28-74
        if (mat.type != GLASS && dot(g.N, inc) <= (R)0) {
            weight = V3<R>(0);
        } else {
            V3<R> ldir, Le;
            R dist;
            bool degen;
            if (nee_sample(S, mat, seed, g.P, inc, g.N, ldir, Le, dist, degen) && !degen) {
                V3<R> so = g.P + (R)0.001f * ldir;
                V3<R> sd = normalize(ldir);
                if (T.visible(so, sd, dist)) {
                    R nl = (R)(float)(S.lights.size() + (S.use_env ? 1 : 0));
                    radiance += (nl * std::fmax((R)0, dot(g.N, ldir))) * brdf(mat, inc, ldir, g.N) * Le * weight;
                }
            }
        }
        if (rr) {  // :141-152
            R alpha = clampf(std::fmax(weight.x, std::fmax(weight.y, weight.z)), (R)0, (R)1);
            if ((R)rnd(seed) > alpha || safeFloatEqual(alpha, (R)0)) break;
            weight = weight / alpha;
        }
        // indirectIllumination :76-95
        V4<R> s = sampleDirection(mat, seed, inc, g.N);
        if (!(norm(s.v) < (R)kEps)) {
            weight = weight * std::abs(dot(s.v, g.N)) * brdf(mat, inc, s.v, g.N) / s.w;
            out_dir = s.v;
            next_valid = true;
        }
        if (!next_valid) break;
        o = g.P + (R)0.01f * out_dir;  // ray.traceNew :156
        d = normalize(out_dir);
        ++depth;
    } while (depth < maxDepth);
    return radiance;
}

// ------------------------------------------------------------------ volume renderer
// volume_kernel + handleMediumInteraction, lib/src/Renderer/VolumeRenderer.cu:11-310
template <typename R>
static V3<R> vol_sample(const SceneT<R>& S, const Tracer<R>& T, const Camera<R>& cam, uint32_t tid, uint32_t w, uint32_t h,
                        uint32_t frame, uint32_t maxDepth, bool rr) {
    const R inf = std::numeric_limits<R>::infinity();
    uint32_t seed = g_rng_pcg ? pcg_hash(pcg_hash(frame) ^ tid) : tea4(tid, frame);
    V3<R> o, d;
    launch_ray(tid % w, tid / w, w, h, cam, seed, o, d);
    V3<R> radiance(0), weight(1), ray_start, out_dir;
    bool inside = false;
    int obj_index = 0;
    const Volume<R>* vol = &S.volume;
    uint32_t depth = 0;
    const R nl = (R)(float)(S.lights.size() + (S.use_env ? 1 : 0));
    do {
        bool next_valid = false;
        Hit<R> g = inside ? T.object(obj_index, o, d) : T.closest(o, d);
        V3<R> inc = (R)-1.0f * d;
        if (!(g.depth < inf)) {
            if (S.use_env) radiance += weight * env_lookup(S, d);
            break;
        }
        const Material<R>& mat = S.objs[g.obj].mat;
        if (inside) vol = &mat.volume;  // payload.volume points at the current hit's material volume
        // handleMediumInteraction :88-212
        bool medium_event = false;
        {
            R gg = vol->g;
            V3<R> sa = vol->sigma_a, ss = vol->sigma_s, st = sa + ss;
            uint32_t ch = (uint32_t)(rnd(seed) * 3.0f);
            if (ch > 2) ch = 2;
            if (!safeFloatEqual(st[ch], (R)0)) {
                R t = -(R)std::log((R)1 - (R)rnd(seed)) / st[ch];
                if (t < g.depth) {
                    medium_event = true;
                    V3<R> ep = o + t * d;
                    R sprob = ss[ch] / st[ch];
                    if ((R)rnd(seed) < sprob) {
                        R pdf = 0;
                        for (int i = 0; i < 3; ++i) pdf += st[i] * (R)std::exp(-(R)1 * st[i] * t);
                        pdf /= (R)3;
                        weight = weight * ss / sprob * vexp((R)-1 * st * t) / pdf;
                    } else {
                        weight = V3<R>(0);
                        break;  // next_ray_valid=false -> loop ends after traceNew
                    }
                    // direct illumination at the medium event
                    V3<R> ldir, Le;
                    R dist;
                    bool degen;
                    bool have = nee_sample(S, mat, seed, ep, inc, g.N, ldir, Le, dist, degen) && !degen;
                    if (have) {
                        V3<R> so, sd;
                        V3<R> att(1);
                        if (inside) {
                            V3<R> ld = normalize(ldir);
                            Hit<R> ge = T.object(obj_index, ep, ld);
                            if (!(ge.depth < inf)) { ge.depth = 0; ge.P = ep; }
                            att = vexp((R)-1 * st * ge.depth) * hg_phase(gg, -dot(inc, ldir));
                            so = ge.P + (R)0.001f * ldir;
                            sd = normalize(ldir);
                            dist -= ge.depth;
                        } else {
                            so = ep + (R)0.001f * ldir;
                            sd = normalize(ldir);
                        }
                        V3<R> sst = S.volume.sigma_a + S.volume.sigma_s;
                        if (T.visible(so, sd, dist))
                            radiance += nl * vexp((R)-1 * sst * std::fmin(dist, (R)100000000.0f)) * att * Le * weight;
                    }
                    V4<R> hs = hg_sample(gg, (R)-1 * inc, seed);
                    out_dir = hs.v;
                    ray_start = ep;
                    next_valid = true;
                } else {
                    R pdf = 0;
                    for (int i = 0; i < 3; ++i) pdf += (R)std::exp(-st[i] * g.depth);
                    pdf /= (R)3;
                    weight = weight * vexp((R)-1 * st * g.depth) / pdf;
                    ray_start = g.P;
                }
            }
        }
        if (!medium_event) {
            if (mat.type == VOLUME) {  // medium interface :269-278
                V4<R> s = sampleDirection(mat, seed, inc, g.N);
                out_dir = s.v;
                ray_start = g.P;
                bool refl = dot(inc, g.N) * dot(out_dir, g.N) > 0;
                inside = refl ? inside : !inside;
                obj_index = g.obj;
                next_valid = true;
                vol = inside ? &mat.volume : &S.volume;
            } else {
                // directIlluminationVolumetric :24-70
                if (mat.type != GLASS && dot(g.N, inc) <= (R)0) {
                    weight = V3<R>(0);
                } else {
                    V3<R> ldir, Le;
                    R dist;
                    bool degen;
                    if (nee_sample(S, mat, seed, g.P, inc, g.N, ldir, Le, dist, degen) && !degen) {
                        V3<R> so = g.P + (R)0.001f * ldir;
                        V3<R> sd = normalize(ldir);
                        if (T.visible(so, sd, dist))
                            radiance += (nl * std::fmax((R)0, dot(g.N, ldir))) *
                                        vexp((R)-1 * (vol->sigma_s + vol->sigma_a) * std::fmin(dist, (R)100000000.0f)) *
                                        brdf(mat, inc, ldir, g.N) * Le * weight;
                    }
                }
                // indirectIlluminationVolumetric :72-86
                V4<R> s = sampleDirection(mat, seed, inc, g.N);
                if (!(norm(s.v) == (R)0)) {
                    weight = weight * std::abs(dot(s.v, g.N)) * brdf(mat, inc, s.v, g.N) / s.w;
                    out_dir = s.v;
                    ray_start = g.P;
                    next_valid = true;
                }
            }
            if (rr) {
                R alpha = clampf(std::fmax(weight.x, std::fmax(weight.y, weight.z)), (R)0, (R)1);
                if ((R)rnd(seed) > alpha || safeFloatEqual(alpha, (R)0)) break;
                weight = weight / alpha;
            }
        }
        o = ray_start + (R)0.001f * out_dir;
        d = normalize(out_dir);
        if (!next_valid) break;
        ++depth;
    } while (depth < maxDepth);
    return radiance;
}

// ------------------------------------------------------------------ gradient-domain PT
// lib/src/Renderer/GradientDomain.cu:110-246. valid[] of the base path is false beyond the vertices
// the base path wrote (the reference reads uninitialised local memory there).
template <typename R>
static V3<R> gdpt_collect(const SceneT<R>& S, const Tracer<R>& T, V3<R> o, V3<R> d, uint32_t seed, uint32_t maxDepth,
                          bool* valid, const bool* base_valid) {
    V3<R> radiance(0), weight(1), out_dir;
    uint32_t depth = 0;
    const R nl = (R)(float)(S.lights.size() + (S.use_env ? 1 : 0));
    do {
        bool next_valid = false;
        if (base_valid && !base_valid[depth]) return radiance;  // :127
        Hit<R> g = T.closest(o, d);
        if (!(g.depth < std::numeric_limits<R>::infinity())) {
            if (valid) valid[depth] = false;
            if (S.use_env) radiance += weight * env_lookup(S, d);
            break;
        }
        if (valid) valid[depth] = true;
        const Material<R>& mat = S.objs[g.obj].mat;
        V3<R> inc = (R)-1.0f * d;
        radiance += weight * mat.albedo_e;
        if (mat.type != GLASS && dot(g.N, inc) <= (R)0) {
            weight = V3<R>(0);
        } else {
            V3<R> ldir, Le;
            R dist;
            bool degen;
            if (nee_sample(S, mat, seed, g.P, inc, g.N, ldir, Le, dist, degen) && !degen) {
                V3<R> so = g.P + (R)0.001f * ldir;
                if (T.visible(so, normalize(ldir), dist))
                    radiance += (nl * std::fmax((R)0, dot(g.N, ldir))) * brdf(mat, inc, ldir, g.N) * Le * weight;
            }
        }
        V4<R> s = sampleDirection(mat, seed, inc, g.N);
        if (!(norm(s.v) == (R)0)) {
            weight = weight * std::abs(dot(s.v, g.N)) * brdf(mat, inc, s.v, g.N) / s.w;
            out_dir = s.v;
            next_valid = true;
        }
        o = g.P + (R)0.01f * out_dir;
        d = normalize(out_dir);
        if (!next_valid) break;
        ++depth;
    } while (depth < maxDepth);
    return radiance;
}

// :160-246 — one pixel's five paths; returns base radiance and the four half-differences
template <typename R>
static void gdpt_pixel(const SceneT<R>& S, const Tracer<R>& T, const Camera<R>& cam, uint32_t px, uint32_t py, uint32_t w,
                       uint32_t h, uint32_t frame, uint32_t maxDepth, V3<R> out[5]) {
    uint32_t seed = tea4(py * w + px, frame);
    V3<R> ob, db, ol, dl, orr, dr, ou, du, od, dd;
    launch_ray(px, py, w, h, cam, seed, ob, db);
    uint32_t sb = seed;
    launch_ray(px - 1, py, w, h, cam, seed, ol, dl);
    uint32_t sl = seed;
    launch_ray(px + 1, py, w, h, cam, seed, orr, dr);
    uint32_t sr = seed;
    launch_ray(px, py + 1, w, h, cam, seed, ou, du);
    uint32_t su = seed;
    launch_ray(px, py - 1, w, h, cam, seed, od, dd);
    uint32_t sd = seed;
    bool valid[16];
    for (int i = 0; i < 16; ++i) valid[i] = false;
    V3<R> Lb = gdpt_collect(S, T, ob, db, sb, maxDepth, valid, (const bool*)nullptr);
    V3<R> Ll = gdpt_collect(S, T, ol, dl, sl, maxDepth, (bool*)nullptr, valid);
    V3<R> Lr = gdpt_collect(S, T, orr, dr, sr, maxDepth, (bool*)nullptr, valid);
    V3<R> Lu = gdpt_collect(S, T, ou, du, su, maxDepth, (bool*)nullptr, valid);
    V3<R> Ld = gdpt_collect(S, T, od, dd, sd, maxDepth, (bool*)nullptr, valid);
    out[0] = Lb;
    out[1] = (R)0.5f * (Lb - Ll);  // gx forward
    out[2] = (R)0.5f * (Lr - Lb);  // gx backward
    out[3] = (R)0.5f * (Lb - Ld);  // gy forward
    out[4] = (R)0.5f * (Lu - Lb);  // gy backward
}

// compute_median, :275-301 (mean of the middle two of four)
template <typename R>
static inline R median4(const R* v) {
    R m = 0;
    uint32_t amin = 0, amax = 0;
    for (uint32_t i = 0; i < 4; ++i) {
        m += v[i];
        if (v[i] < v[amin]) amin = i;
        if (v[i] > v[amax]) amax = i;
    }
    return (m - v[amin] - v[amax]) / (R)2;
}

// init_kernel + 50 x optimization_kernel + copy, :248-350, :380-401. recon/temp borders are zero.
template <typename R>
static void gdpt_reconstruct(uint32_t w, uint32_t h, const std::vector<V3<R>>& base, const std::vector<V3<R>>& gxf,
                             const std::vector<V3<R>>& gxb, const std::vector<V3<R>>& gyf, const std::vector<V3<R>>& gyb,
                             int iters, std::vector<V3<R>>& recon, std::vector<V3<R>>& gx, std::vector<V3<R>>& gy) {
    size_t n = (size_t)w * h;
    recon = base;
    gx.assign(n, V3<R>(0));
    gy.assign(n, V3<R>(0));
    for (uint32_t y = 1; y + 1 < h; ++y)
        for (uint32_t x = 1; x + 1 < w; ++x) {
            size_t i = (size_t)y * w + x;
            gx[i] = gxf[i] + gxb[i];
            gy[i] = gyf[i] + gyb[i];
        }
    std::vector<V3<R>> temp(n, V3<R>(0));
    for (int it = 0; it < iters; ++it) {
        for (uint32_t y = 1; y + 1 < h; ++y)
            for (uint32_t x = 1; x + 1 < w; ++x) {
                size_t i = (size_t)y * w + x;
                V3<R> v1 = recon[i - 1] + gx[i - 1];
                V3<R> v2 = recon[i - w] + gy[i - w];
                V3<R> v3 = recon[i + 1] - gx[i];
                V3<R> v4 = recon[i + w] - gy[i];
                V3<R> r;
                for (int c = 0; c < 3; ++c) {
                    R vals[4] = {v1[c], v2[c], v3[c], v4[c]};
                    r[c] = median4(vals);
                }
                temp[i] = r;
            }
        recon = temp;
    }
}

// ------------------------------------------------------------------ post chain
// thresholding_kernel, lib/src/PostProcessing/PyramidConstructor.cu:9-30
static inline V3<float> threshold_px(V3<float> c, const float* t) {
    float b = c.x > c.y ? c.x : c.y;
    b = c.z > b ? c.z : b;
    float rq = clampf(b - t[1], 0.0f, t[2]);
    rq = t[3] * rq * rq;
    float s = std::fmax(rq, b - t[0]) / std::fmax(b, 1e-5f);
    return V3<float>(c.x * s, c.y * s, c.z * s);
}

// ------------------------------------------------------------------ thread pool over pixels
static void parallel_for(size_t n, int threads, const std::function<void(size_t, size_t)>& fn) {
    if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
    if (threads <= 1 || n < 256) { fn(0, n); return; }
    std::atomic<size_t> next(0);
    const size_t chunk = 256;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&]() {
            for (;;) {
                size_t s = next.fetch_add(chunk);
                if (s >= n) break;
                fn(s, std::min(n, s + chunk));
            }
        });
    for (auto& th : pool) th.join();
}

// material layout (21 floats): albedo_e[3] albedo_d[3] albedo_s[3] shininess roughness eta
// sigma_s[3] sigma_a[3] g interface type — the fields of Material.h:35-43 / Volume.h:15-21
template <typename R>
static Material<R> read_mat(const float* m) {
    Material<R> mat;
    mat.albedo_e = V3<R>(m[0], m[1], m[2]);
    mat.albedo_d = V3<R>(m[3], m[4], m[5]);
    mat.albedo_s = V3<R>(m[6], m[7], m[8]);
    mat.shininess = m[9];
    mat.roughness = m[10];
    mat.eta = m[11];
    mat.volume.sigma_s = V3<R>(m[12], m[13], m[14]);
    mat.volume.sigma_a = V3<R>(m[15], m[16], m[17]);
    mat.volume.g = m[18];
    mat.volume.interface_ = (int)m[19];
    mat.type = (int)m[20];
    return mat;
}

// ------------------------------------------------------------------ building a scene
// object params: SPHERE c[3] r; PLANE p[3] n[3]; QUAD p[3] n[3] e1[3] e2[3]; TRIANGLE v1 v2 v3;
// MESH box min[3] max[3] (ObjLoader.cpp:24-47). Normals are normalised as the constructors do
// (PlaneDetail.h:14, QuadDetail.h:14, TriangleDetail.h:16).
template <typename R>
static SceneT<R>* make_scene(const int32_t* types, const float* params, const uint32_t* first, const uint32_t* count,
                             const float* mats, uint32_t n_obj, const float* tris, uint32_t n_tri, const float* lights,
                             uint32_t n_lights, const float* env, uint32_t env_w, uint32_t env_h, const float* volume7) {
    auto* S = new SceneT<R>();
    auto v3 = [](const float* p) { return V3<R>((R)p[0], (R)p[1], (R)p[2]); };
    for (uint32_t i = 0; i < n_tri; ++i) {
        const float* t = tris + 9 * (size_t)i;
        Tri<R> tr;
        tr.v1 = v3(t); tr.v2 = v3(t + 3); tr.v3 = v3(t + 6);
        tr.n = normalize(cross(tr.v2 - tr.v1, tr.v3 - tr.v1));
        S->tris.push_back(tr);
    }
    for (uint32_t i = 0; i < n_obj; ++i) {
        const float* p = params + 12 * (size_t)i;
        Object<R> ob;
        ob.type = types[i];
        ob.mat = read_mat<R>(mats + 21 * (size_t)i);
        switch (ob.type) {
            case SPHERE:
                ob.a = v3(p); ob.radius = (R)p[3];
                ob.finite_box = true;  // SphereDetail.h:16
                ob.bmin = ob.a - V3<R>(ob.radius);
                ob.bmax = ob.a + V3<R>(ob.radius);
                break;
            case PLANE: ob.a = v3(p); ob.b = normalize(v3(p + 3)); break;
            case QUAD: ob.a = v3(p); ob.b = normalize(v3(p + 3)); ob.c = v3(p + 6); ob.d = v3(p + 9); break;
            case TRIANGLE:
                ob.a = v3(p); ob.b = v3(p + 3); ob.c = v3(p + 6);
                ob.tri_normal = normalize(cross(ob.b - ob.a, ob.c - ob.a));
                break;
            case MESH:
                ob.first = first[i]; ob.count = count[i];
                ob.finite_box = true;  // MeshDetail.h:18
                ob.bmin = v3(p); ob.bmax = v3(p + 3);
                break;
        }
        S->objs.push_back(ob);
    }
    for (uint32_t i = 0; i < n_lights; ++i) {
        const float* l = lights + 16 * (size_t)i;
        Light<R> L;
        L.type = (int)l[0];
        L.position = v3(l + 1); L.intensity = v3(l + 4); L.radiance = v3(l + 7);
        L.halfExtend1 = v3(l + 10); L.halfExtend2 = v3(l + 13);
        S->lights.push_back(L);
    }
    if (env && env_w && env_h) {
        S->use_env = true; S->env_w = env_w; S->env_h = env_h;
        S->env.resize((size_t)env_w * env_h);
        for (size_t i = 0; i < S->env.size(); ++i) S->env[i] = v3(env + 3 * i);
    }
    if (volume7) {
        S->volume.sigma_s = v3(volume7);
        S->volume.sigma_a = v3(volume7 + 3);
        S->volume.g = volume7[6];
    }
    // mode 0/1: reference tree and its leaf order; ranks
    if (n_obj) {
        std::vector<int> all(n_obj);
        for (uint32_t i = 0; i < n_obj; ++i) all[i] = (int)i;
        S->ref_root = build_ref(*S, all);
        ref_leaf_order(*S, S->ref_root, S->leaf_order);
    }
    S->rank_base.assign(n_obj, 0);
    uint32_t rank = 0;
    for (int obj : S->leaf_order) {
        S->rank_base[obj] = rank;
        const Object<R>& ob = S->objs[obj];
        if (ob.type == MESH) {
            for (uint32_t t = 0; t < ob.count; ++t) S->prims.push_back({rank + t, obj, (int)t});
            rank += ob.count;
        } else if (ob.type == PLANE || ob.type == QUAD) {
            S->unbounded.push_back({rank, obj, -1});
            rank += 1;
        } else {
            S->prims.push_back({rank, obj, -1});
            rank += 1;
        }
    }
    // mode 2
    if (!S->prims.empty()) {
        SahBuilder<R> B(*S);
        size_t np = S->prims.size();
        B.pmin.resize(np); B.pmax.resize(np); B.cen.resize(np);
        for (size_t i = 0; i < np; ++i) {
            prim_bounds(*S, S->prims[i], B.pmin[i], B.pmax[i]);
            B.cen[i] = (B.pmin[i] + B.pmax[i]) * (R)0.5;
        }
        B.build(0, (uint32_t)np);
    }
    return S;
}


// ------------------------------------------------------------------ heterogeneous medium + adjoint
// The reference has no heterogeneous medium (Volume.h:15-21); this restates the renderer core's
// grid integrator (cupbr_amd/csrc/grid.hip) from its definition: weighted delta tracking with a
// fixed proposal grid, ratio-tracking NEE, Lambertian surfaces with per-surface albedo, and the
// path-replay adjoint. Same operation order as the HIP code so that R=float agrees path for path.
template <typename R>
struct GridT {
    const double* rho;
    const double* prop;
    const double* alb;
    R bmin[3], ext[3];
    uint32_t n[3];
    R scale, maj, alpha, g, eps;
    uint32_t n_objs;
};

static const uint32_t kGridMaxSteps = 1u << 22;  // tentative collisions per segment (path ends there)

struct GridAcc {  // per-thread gradient accumulators
    std::vector<double> rho, alb;
};

template <typename R>
struct StencilT {
    uint32_t idx[8];
    R w[8];
};

template <typename R>
static inline void grid_stencil(const GridT<R>& G, const V3<R>& x, StencilT<R>& s) {
    uint32_t i0[3], i1[3];
    R f[3];
    for (int a = 0; a < 3; ++a) {
        const R u = (x[a] - G.bmin[a]) / G.ext[a] * (R)(float)G.n[a] - (R)0.5f;
        const R fl = std::floor(u);
        f[a] = u - fl;
        const int lo = (int)fl, hi_max = (int)G.n[a] - 1;
        i0[a] = (uint32_t)std::min(std::max(lo, 0), hi_max);
        i1[a] = (uint32_t)std::min(std::max(lo + 1, 0), hi_max);
    }
    for (int k = 0; k < 8; ++k) {
        const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
        const uint32_t ix = dx ? i1[0] : i0[0], iy = dy ? i1[1] : i0[1], iz = dz ? i1[2] : i0[2];
        s.idx[k] = (iz * G.n[1] + iy) * G.n[0] + ix;
        s.w[k] = ((dx ? f[0] : (R)1 - f[0]) * (dy ? f[1] : (R)1 - f[1])) * (dz ? f[2] : (R)1 - f[2]);
    }
}

template <typename R>
static inline R grid_value(const double* g, const StencilT<R>& s) {
    R v = 0;
    for (int k = 0; k < 8; ++k) v += s.w[k] * (R)g[s.idx[k]];
    return v;
}

template <typename R>
static inline void grid_scatter(std::vector<double>& grad, const StencilT<R>& s, R c) {
    if (c == (R)0 || !(std::abs(c) < std::numeric_limits<R>::infinity())) return;
    for (int k = 0; k < 8; ++k)
        if (s.w[k] != (R)0) grad[s.idx[k]] += (double)(c * s.w[k]);
}

template <typename R>
static inline bool grid_clip(const GridT<R>& G, const V3<R>& o, const V3<R>& d, R tmax, R& t0, R& t1) {
    t0 = 0;
    t1 = tmax;
    for (int a = 0; a < 3; ++a) {
        const R lo = G.bmin[a], hi = G.bmin[a] + G.ext[a];
        if (d[a] == (R)0) {
            if (o[a] < lo || o[a] > hi) return false;
            continue;
        }
        R ta = (lo - o[a]) / d[a], tb = (hi - o[a]) / d[a];
        if (ta > tb) std::swap(ta, tb);
        t0 = std::fmax(t0, ta);
        t1 = std::fmin(t1, tb);
    }
    return t0 < t1;
}

template <typename R>
static inline R grid_step(const GridT<R>& G, R t, uint32_t& seed) {
    return t - std::log((R)1.0f - (R)rnd(seed)) / G.maj;
}

template <typename R>
static R grid_ratio_track(const GridT<R>& G, const V3<R>& o, const V3<R>& d, R dist, uint32_t& seed,
                          std::vector<double>* grad, R coef) {
    R t0, t1;
    if (!grid_clip(G, o, d, dist, t0, t1)) return (R)1;
    const R len = t1 - t0;
    R T = 1, u = 0;
    for (uint32_t it = 0;; ++it) {
        if (it == kGridMaxSteps) return (R)0;
        u = grid_step(G, u, seed);
        if (!(u < len)) break;
        StencilT<R> s;
        grid_stencil(G, o + (t0 + u) * d, s);
        const R sigma = G.scale * grid_value(G.rho, s);
        if (grad) grid_scatter(*grad, s, coef * ((R)-1.0f / (G.maj - sigma)) * G.scale);
        else T = T * ((R)1.0f - sigma / G.maj);
    }
    return T;
}

template <typename R>
struct GridAdj {
    V3<R> delta, rem;
    GridAcc* acc;
};

template <typename R>
static bool grid_nee(const SceneT<R>& S, const GridT<R>& G, const Tracer<R>& T, const V3<R>& x, uint32_t& seed,
                     V3<R>& ldir, V3<R>& le, R& tr, uint32_t& tseed, R& dist, bool surface, const V3<R>& N) {
    const uint32_t nl = (uint32_t)S.lights.size();
    if (nl == 0) return false;
    uint32_t k = (uint32_t)(rnd(seed) * (float)nl);
    k = std::min(k, nl - 1);
    le = light_sample(S.lights[k], seed, x, ldir, dist) * (R)(float)nl;
    if (surface && !(dot(N, ldir) > (R)0)) return false;
    const V3<R> so = surface ? x + (R)1e-4f * N : x;
    if (!T.visible(so, ldir, dist)) return false;
    tseed = seed;
    tr = grid_ratio_track(G, so, ldir, dist, seed, nullptr, (R)0);
    return true;
}

template <typename R, bool kAdj>
static V3<R> grid_path(const SceneT<R>& S, const GridT<R>& G, const Tracer<R>& T, const Camera<R>& cam, uint32_t px,
                       uint32_t py, uint32_t w, uint32_t h, uint32_t frame, uint32_t max_depth, GridAdj<R>* A) {
    const R pi = (R)kPiD;
    const R inf = std::numeric_limits<R>::infinity();
    uint32_t seed = tea4(py * w + px, frame);
    V3<R> o, d;
    launch_ray(px, py, w, h, cam, seed, o, d);
    V3<R> L(0), beta(1);
    for (uint32_t depth = 0; depth < max_depth; ++depth) {
        Hit<R> hr = T.closest(o, d);
        const bool hit = hr.obj >= 0;
        const R ts = hit ? hr.depth : inf;
        bool scattered = false;
        V3<R> x;
        R t0, t1;
        if (grid_clip(G, o, d, ts, t0, t1)) {
            const R len = t1 - t0;
            R u = 0;
            for (uint32_t it = 0;; ++it) {
                if (it == kGridMaxSteps) return L;
                u = grid_step(G, u, seed);
                if (!(u < len)) break;
                x = o + (t0 + u) * d;
                StencilT<R> s;
                grid_stencil(G, x, s);
                const R sigma = G.scale * grid_value(G.rho, s);
                const R P = clampf(G.scale * grid_value(G.prop, s) / G.maj, G.eps, (R)1.0f - G.eps);
                R f, dlog;
                const bool real = (R)rnd(seed) < P;
                if (real) {
                    f = G.alpha * sigma / (G.maj * P);
                    dlog = (R)1.0f / sigma;
                } else {
                    f = ((R)1.0f - sigma / G.maj) / ((R)1.0f - P);
                    dlog = (R)-1.0f / (G.maj - sigma);
                }
                if (f == (R)0) return L;
                beta = beta * f;
                if (kAdj) grid_scatter(A->acc->rho, s, dot(A->delta, A->rem) * dlog * G.scale);
                if (real) {
                    scattered = true;
                    break;
                }
            }
        }
        if (scattered) {
            V3<R> ldir, le;
            R tr, dist;
            uint32_t tseed;
            if (grid_nee(S, G, T, x, seed, ldir, le, tr, tseed, dist, false, V3<R>(0))) {
                const R ph = hg_phase(G.g, dot(d, ldir));
                const V3<R> C = beta * (le * (ph * tr));
                L += C;
                if (kAdj) {
                    const R c = dot(A->delta, C);
                    if (c != (R)0) grid_ratio_track(G, x, ldir, dist, tseed, &A->acc->rho, c);
                    A->rem = A->rem - C;
                }
            }
            if (depth + 1 == max_depth) break;
            V4<R> hs = hg_sample(G.g, d, seed);
            o = x;
            d = hs.v;
            continue;
        }
        if (!hit) break;
        const Object<R>& ob = S.objs[hr.obj];
        const uint32_t surf = ob.type == MESH ? G.n_objs + ob.first + (uint32_t)hr.tri : (uint32_t)hr.obj;
        V3<R> N = hr.N;
        if (dot(N, d) > (R)0) N = (R)-1.0f * N;
        const V3<R> P = hr.P;
        const V3<R> a((R)G.alb[3 * surf], (R)G.alb[3 * surf + 1], (R)G.alb[3 * surf + 2]);
        V3<R> ldir, le;
        R tr, dist;
        uint32_t tseed;
        if (grid_nee(S, G, T, P, seed, ldir, le, tr, tseed, dist, true, N)) {
            const V3<R> E0 = le * (dot(N, ldir) * tr / pi);
            const V3<R> C = beta * (a * E0);
            L += C;
            if (kAdj) {
                const V3<R> ga = A->delta * beta * E0;
                for (int c = 0; c < 3; ++c) A->acc->alb[3 * surf + c] += (double)ga[c];
                const R cc = dot(A->delta, C);
                if (cc != (R)0) grid_ratio_track(G, P + (R)1e-4f * N, ldir, dist, tseed, &A->acc->rho, cc);
                A->rem = A->rem - C;
            }
        }
        if (depth + 1 == max_depth) break;
        const R u1 = (R)rnd(seed), u2 = (R)rnd(seed);
        const R r = std::sqrt(u1), phi = (R)2.0f * pi * u2;
        const V3<R> nd = normalize(toLocalFrame(N, V3<R>(r * std::cos(phi), r * std::sin(phi), std::sqrt(std::fmax((R)0, (R)1.0f - u1)))));
        if (a.x == (R)0 && a.y == (R)0 && a.z == (R)0) break;
        beta = beta * a;
        if (kAdj) {
            for (int c = 0; c < 3; ++c)
                if (a[c] > (R)0 && A->delta[c] * A->rem[c] != (R)0)
                    A->acc->alb[3 * surf + c] += (double)(A->delta[c] * A->rem[c] / a[c]);
        }
        o = P + (R)1e-4f * N;
        d = nd;
    }
    return L;
}

}  // namespace orc

// ======================================================================= C ABI
using namespace orc;

struct OrcScene {
    SceneT<float>* f = nullptr;
    SceneT<double>* d = nullptr;
};

template <typename R>
static Camera<R> read_cam(const float* c) {
    Camera<R> cam;
    cam.pos = V3<R>(c[0], c[1], c[2]);
    cam.xa = V3<R>(c[3], c[4], c[5]);
    cam.ya = V3<R>(c[6], c[7], c[8]);
    cam.za = V3<R>(c[9], c[10], c[11]);
    return cam;
}

template <typename R>
static void trace_impl(const SceneT<R>& S, int mode, const float* rays, uint64_t n, int32_t* obj, int32_t* tri, float* t,
                       double* P, double* N, int threads) {
    Tracer<R> T{S, mode};
    parallel_for(n, threads, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            const float* r = rays + 6 * i;
            V3<R> o(r[0], r[1], r[2]), d(r[3], r[4], r[5]);
            Hit<R> h = T.closest(o, d);
            obj[i] = h.obj;
            tri[i] = h.tri;
            t[i] = (float)h.depth;
            if (P) { P[3 * i] = (double)h.P.x; P[3 * i + 1] = (double)h.P.y; P[3 * i + 2] = (double)h.P.z; }
            if (N) { N[3 * i] = (double)h.N.x; N[3 * i + 1] = (double)h.N.y; N[3 * i + 2] = (double)h.N.z; }
        }
    });
}

template <typename R>
static void vis_impl(const SceneT<R>& S, int mode, const float* rays, const float* dist, uint64_t n, uint8_t* vis, int threads) {
    Tracer<R> T{S, mode};
    parallel_for(n, threads, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            const float* r = rays + 6 * i;
            vis[i] = T.visible(V3<R>(r[0], r[1], r[2]), V3<R>(r[3], r[4], r[5]), (R)dist[i]) ? 1 : 0;
        }
    });
}

// Runs frames [frame0, frame0+nframes) of method 2 (PT) or 5 (VOLUME) with the reference's running mean
// (PathTracer.cu:162-168) in precision R; also returns per-pixel sums of the samples and their squares
// (FP64) for statistical tests. img is the running-mean image (w*h*3 doubles, in/out).
template <typename R>
static void render_impl(const SceneT<R>& S, int method, const float* camf, uint32_t w, uint32_t h, uint32_t frame0,
                        uint32_t nframes, uint32_t depth, int rr, int mode, int threads, double* img, double* sum,
                        double* sumsq) {
    Tracer<R> T{S, mode};
    Camera<R> cam = read_cam<R>(camf);
    size_t n = (size_t)w * h;
    parallel_for(n, threads, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            V3<R> acc((R)img[3 * i], (R)img[3 * i + 1], (R)img[3 * i + 2]);
            for (uint32_t f = frame0; f < frame0 + nframes; ++f) {
                V3<R> L = method == 5 ? vol_sample(S, T, cam, (uint32_t)i, w, h, f, depth, rr != 0)
                                      : pt_sample(S, T, cam, (uint32_t)i, w, h, f, depth, rr != 0);
                if (f > 0) {
                    const R a = (R)1.0f / ((R)(float)f + (R)1.0f);
                    acc = ((R)1.0f - a) * acc + a * L;
                } else {
                    acc = L;
                }
                if (sum) for (int c = 0; c < 3; ++c) { sum[3 * i + c] += (double)L[c]; sumsq[3 * i + c] += (double)L[c] * (double)L[c]; }
            }
            img[3 * i] = (double)acc.x; img[3 * i + 1] = (double)acc.y; img[3 * i + 2] = (double)acc.z;
        }
    });
}

// GDPT frames with running means of base + 4 gradient buffers (GradientDomain.cu:231-245); bufs holds
// 5 images (base, gxf, gxb, gyf, gyb), each w*h*3 doubles, in/out. Border pixels are untouched.
template <typename R>
static void gdpt_impl(const SceneT<R>& S, const float* camf, uint32_t w, uint32_t h, uint32_t frame0, uint32_t nframes,
                      uint32_t depth, int mode, int threads, double* bufs) {
    Tracer<R> T{S, mode};
    Camera<R> cam = read_cam<R>(camf);
    size_t n = (size_t)w * h;
    parallel_for(n, threads, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            uint32_t px = (uint32_t)(i % w), py = (uint32_t)(i / w);
            if (px == 0 || px == w - 1 || py == 0 || py == h - 1) continue;
            V3<R> acc[5];
            for (int b = 0; b < 5; ++b) acc[b] = V3<R>((R)bufs[b * 3 * n + 3 * i], (R)bufs[b * 3 * n + 3 * i + 1], (R)bufs[b * 3 * n + 3 * i + 2]);
            for (uint32_t f = frame0; f < frame0 + nframes; ++f) {
                V3<R> v[5];
                gdpt_pixel(S, T, cam, px, py, w, h, f, depth, v);
                if (f > 0) {
                    const R a = (R)1.0f / ((R)(float)f + (R)1.0f);
                    for (int b = 0; b < 5; ++b) acc[b] = ((R)1.0f - a) * acc[b] + a * v[b];
                } else {
                    for (int b = 0; b < 5; ++b) acc[b] = v[b];
                }
            }
            for (int b = 0; b < 5; ++b)
                for (int c = 0; c < 3; ++c) bufs[b * 3 * n + 3 * i + c] = (double)acc[b][c];
        }
    });
}

// grid medium: image = mean over frames frame0..frame0+spp-1 (summed in order, then divided)
template <typename R>
static GridT<R> read_grid(const SceneT<R>& S, const float* med, const uint32_t* dims, const double* rho,
                          const double* prop, const double* alb) {
    GridT<R> G;
    G.rho = rho;
    G.prop = prop ? prop : rho;
    G.alb = alb;
    for (int a = 0; a < 3; ++a) {
        G.bmin[a] = (R)med[a];
        G.ext[a] = (R)med[3 + a] - (R)med[a];
        G.n[a] = dims[a];
    }
    G.scale = (R)med[6];
    G.maj = (R)med[7];
    G.alpha = (R)med[8];
    G.g = (R)med[9];
    G.eps = (R)med[10];
    G.n_objs = (uint32_t)S.objs.size();
    return G;
}

template <typename R>
static void grid_impl(const SceneT<R>& S, int mode, const float* camf, const float* med, const uint32_t* dims,
                      const double* rho, const double* prop, const double* alb, uint32_t w, uint32_t h, uint32_t frame0,
                      uint32_t spp, uint32_t depth, int threads, double* img, const double* d_img, double* g_rho,
                      double* g_alb, size_t n_surf) {
    Tracer<R> T{S, mode};
    Camera<R> cam = read_cam<R>(camf);
    const GridT<R> G = read_grid<R>(S, med, dims, rho, prop, alb);
    const size_t n = (size_t)w * h, nvox = (size_t)dims[0] * dims[1] * dims[2];
    std::mutex mu;
    parallel_for(n, threads, [&](size_t lo, size_t hi) {
        GridAcc acc;
        if (d_img) {
            acc.rho.assign(nvox, 0.0);
            acc.alb.assign(3 * n_surf, 0.0);
        }
        for (size_t i = lo; i < hi; ++i) {
            const uint32_t px = (uint32_t)(i % w), py = (uint32_t)(i / w);
            V3<R> sum(0);
            for (uint32_t s = 0; s < spp; ++s) {
                V3<R> L = grid_path<R, false>(S, G, T, cam, px, py, w, h, frame0 + s, depth, nullptr);
                sum += L;
                if (d_img) {
                    GridAdj<R> A;
                    A.delta = V3<R>((R)d_img[3 * i], (R)d_img[3 * i + 1], (R)d_img[3 * i + 2]) / (R)(float)spp;
                    A.rem = L;
                    A.acc = &acc;
                    if (A.delta.x != (R)0 || A.delta.y != (R)0 || A.delta.z != (R)0)
                        grid_path<R, true>(S, G, T, cam, px, py, w, h, frame0 + s, depth, &A);
                }
            }
            if (img) {
                const V3<R> m = sum / (R)(float)spp;
                for (int c = 0; c < 3; ++c) img[3 * i + c] = (double)m[c];
            }
        }
        if (d_img) {
            std::lock_guard<std::mutex> lk(mu);
            for (size_t k = 0; k < nvox; ++k) g_rho[k] += acc.rho[k];
            for (size_t k = 0; k < 3 * n_surf; ++k) g_alb[k] += acc.alb[k];
        }
    });
}

extern "C" {

// precisions: bit 0 builds the FP32 scene, bit 1 the FP64 scene
void* orc_scene_create(const int32_t* types, const float* params, const uint32_t* first, const uint32_t* count,
                       const float* mats, uint32_t n_obj, const float* tris, uint32_t n_tri, const float* lights,
                       uint32_t n_lights, const float* env, uint32_t env_w, uint32_t env_h, const float* volume7,
                       int precisions) {
    auto* s = new OrcScene();
    if (precisions & 1)
        s->f = make_scene<float>(types, params, first, count, mats, n_obj, tris, n_tri, lights, n_lights, env, env_w, env_h, volume7);
    if (precisions & 2)
        s->d = make_scene<double>(types, params, first, count, mats, n_obj, tris, n_tri, lights, n_lights, env, env_w, env_h, volume7);
    return s;
}

void orc_scene_destroy(void* p) {
    auto* s = (OrcScene*)p;
    delete s->f;
    delete s->d;
    delete s;
}

uint32_t orc_tea4(uint32_t a, uint32_t b) { return tea4(a, b); }
void orc_set_rng(int pcg) { g_rng_pcg = pcg ? 1 : 0; }
uint32_t orc_pcg_hash(uint32_t x) { return pcg_hash(x); }

void orc_rnd_seq(uint32_t seed, uint32_t n, float* out, uint32_t* states) {
    for (uint32_t i = 0; i < n; ++i) {
        out[i] = rnd(seed);
        if (states) states[i] = seed;
    }
}

// reference DFS leaf order of the objects and their rank bases
uint32_t orc_leaf_order(void* p, int32_t* order, uint32_t* rank_base) {
    auto* s = (OrcScene*)p;
    const std::vector<int>& lo = s->f ? s->f->leaf_order : s->d->leaf_order;
    const std::vector<uint32_t>& rb = s->f ? s->f->rank_base : s->d->rank_base;
    for (size_t i = 0; i < lo.size(); ++i) order[i] = lo[i];
    for (size_t i = 0; i < rb.size(); ++i) rank_base[i] = rb[i];
    return (uint32_t)lo.size();
}

// rays: n x {o.xyz, d.xyz} (d normalised by the caller as Ray's constructor does)
void orc_trace(void* p, int precision, int mode, const float* rays, uint64_t n, int32_t* obj, int32_t* tri, float* t,
               double* P, double* N, int threads) {
    auto* s = (OrcScene*)p;
    if (precision == 64) trace_impl(*s->d, mode, rays, n, obj, tri, t, P, N, threads);
    else trace_impl(*s->f, mode, rays, n, obj, tri, t, P, N, threads);
}

void orc_visibility(void* p, int precision, int mode, const float* rays, const float* dist, uint64_t n, uint8_t* vis,
                    int threads) {
    auto* s = (OrcScene*)p;
    if (precision == 64) vis_impl(*s->d, mode, rays, dist, n, vis, threads);
    else vis_impl(*s->f, mode, rays, dist, n, vis, threads);
}

void orc_render(void* p, int precision, int method, const float* cam, uint32_t w, uint32_t h, uint32_t frame0,
                uint32_t nframes, uint32_t depth, int rr, int mode, int threads, double* img, double* sum, double* sumsq) {
    auto* s = (OrcScene*)p;
    if (precision == 64) render_impl(*s->d, method, cam, w, h, frame0, nframes, depth, rr, mode, threads, img, sum, sumsq);
    else render_impl(*s->f, method, cam, w, h, frame0, nframes, depth, rr, mode, threads, img, sum, sumsq);
}

void orc_gdpt(void* p, int precision, const float* cam, uint32_t w, uint32_t h, uint32_t frame0, uint32_t nframes,
              uint32_t depth, int mode, int threads, double* bufs) {
    auto* s = (OrcScene*)p;
    if (precision == 64) gdpt_impl(*s->d, cam, w, h, frame0, nframes, depth, mode, threads, bufs);
    else gdpt_impl(*s->f, cam, w, h, frame0, nframes, depth, mode, threads, bufs);
}

// grid medium forward (img) and adjoint (d_img -> g_rho, g_alb; accumulated). med: bmin[3], bmax[3],
// sigma_t_scale, majorant, scattering_albedo, g, min_collision_probability; dims: nx, ny, nz.
void orc_grid(void* p, int precision, int mode, const float* cam, const float* med, const uint32_t* dims,
              const double* rho, const double* prop, const double* alb, uint32_t w, uint32_t h, uint32_t frame0,
              uint32_t spp, uint32_t depth, int threads, double* img, const double* d_img, double* g_rho,
              double* g_alb) {
    auto* s = (OrcScene*)p;
    if (precision == 64) {
        size_t ns = s->d->objs.size() + s->d->tris.size();
        grid_impl(*s->d, mode, cam, med, dims, rho, prop, alb, w, h, frame0, spp, depth, threads, img, d_img, g_rho, g_alb, ns);
    } else {
        size_t ns = s->f->objs.size() + s->f->tris.size();
        grid_impl(*s->f, mode, cam, med, dims, rho, prop, alb, w, h, frame0, spp, depth, threads, img, d_img, g_rho, g_alb, ns);
    }
}

// ---- probes of single functions (FP64), for the closed-form checks of the tests
void orc_brdf(const float* mat21, const double* i3, const double* o3, const double* n3, double* out3) {
    Material<double> m = read_mat<double>(mat21);
    V3<double> r = brdf(m, V3<double>(i3[0], i3[1], i3[2]), V3<double>(o3[0], o3[1], o3[2]), V3<double>(n3[0], n3[1], n3[2]));
    out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}
// n samples of Material::sampleDirection from consecutive RNG states starting at seed; out4: dir.xyz, pdf
void orc_sample_direction(const float* mat21, uint32_t seed, const double* i3, const double* n3, uint32_t n, double* out4) {
    Material<double> m = read_mat<double>(mat21);
    for (uint32_t k = 0; k < n; ++k) {
        V4<double> s = sampleDirection(m, seed, V3<double>(i3[0], i3[1], i3[2]), V3<double>(n3[0], n3[1], n3[2]));
        out4[4 * k] = s.v.x; out4[4 * k + 1] = s.v.y; out4[4 * k + 2] = s.v.z; out4[4 * k + 3] = s.w;
    }
}
double orc_hg_phase(double g, double c) { return hg_phase(g, c); }
void orc_hg_sample(double g, const double* fwd3, uint32_t seed, uint32_t n, double* out4) {
    for (uint32_t k = 0; k < n; ++k) {
        V4<double> s = hg_sample(g, V3<double>(fwd3[0], fwd3[1], fwd3[2]), seed);
        out4[4 * k] = s.v.x; out4[4 * k + 1] = s.v.y; out4[4 * k + 2] = s.v.z; out4[4 * k + 3] = s.w;
    }
}
double orc_d_ggx(double ndoth, double rough) { return D_GGX(ndoth, rough); }
// n samples of Light::sample at pos; out7: dir.xyz, dist, radiance.xyz
void orc_light_sample(const float* light16, uint32_t seed, const double* pos3, uint32_t n, double* out7) {
    Light<double> L;
    L.type = (int)light16[0];
    L.position = V3<double>(light16[1], light16[2], light16[3]);
    L.intensity = V3<double>(light16[4], light16[5], light16[6]);
    L.radiance = V3<double>(light16[7], light16[8], light16[9]);
    L.halfExtend1 = V3<double>(light16[10], light16[11], light16[12]);
    L.halfExtend2 = V3<double>(light16[13], light16[14], light16[15]);
    for (uint32_t k = 0; k < n; ++k) {
        V3<double> dir;
        double dist;
        V3<double> le = light_sample(L, seed, V3<double>(pos3[0], pos3[1], pos3[2]), dir, dist);
        double* o = out7 + 7 * (size_t)k;
        o[0] = dir.x; o[1] = dir.y; o[2] = dir.z; o[3] = dist; o[4] = le.x; o[5] = le.y; o[6] = le.z;
    }
}
void orc_to_local_frame(const double* n3, const double* d3, double* out3) {
    V3<double> r = toLocalFrame(V3<double>(n3[0], n3[1], n3[2]), V3<double>(d3[0], d3[1], d3[2]));
    out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}
// AABB::hit (biased slab test) in FP32, for n rays against one box
void orc_aabb_hit(const float* bmin3, const float* bmax3, const float* rays, uint64_t n, uint8_t* out) {
    V3<float> mn(bmin3[0], bmin3[1], bmin3[2]), mx(bmax3[0], bmax3[1], bmax3[2]);
    for (uint64_t i = 0; i < n; ++i) {
        const float* r = rays + 6 * i;
        out[i] = aabb_hit(mn, mx, V3<float>(r[0], r[1], r[2]), V3<float>(r[3], r[4], r[5])) ? 1 : 0;
    }
}

// in: 5 images (base, gxf, gxb, gyf, gyb) as float; out: recon, gx, gy as float (FP32 semantics)
void orc_gdpt_reconstruct(uint32_t w, uint32_t h, const float* bufs5, int iters, float* recon, float* gx, float* gy) {
    size_t n = (size_t)w * h;
    std::vector<V3<float>> b[5];
    for (int k = 0; k < 5; ++k) {
        b[k].resize(n);
        for (size_t i = 0; i < n; ++i) b[k][i] = V3<float>(bufs5[k * 3 * n + 3 * i], bufs5[k * 3 * n + 3 * i + 1], bufs5[k * 3 * n + 3 * i + 2]);
    }
    std::vector<V3<float>> r, x, y;
    gdpt_reconstruct<float>(w, h, b[0], b[1], b[2], b[3], b[4], iters, r, x, y);
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) { recon[3 * i + c] = r[i][c]; gx[3 * i + c] = x[i][c]; gy[3 * i + c] = y[i][c]; }
}

// Screened-Poisson (L2) reconstruction, the gradient-domain mode with no counterpart in the
// reference (it keeps init_kernel's gradients, GradientDomain.cu:248-273, and border handling):
// Jacobi iterations on (a^2 + 4) I(p) = a^2 B(p) + (I(L) + gx(L)) + (I(D) + gy(D)) + (I(R) - gx(p))
// + (I(U) - gy(p)) over the interior pixels, the border held at zero. FP32, the device kernel's
// operation order: (a2 * B + ((v1 + v2) + (v3 + v4))) / (a2 + 4).
void orc_gdpt_reconstruct_l2(uint32_t w, uint32_t h, const float* bufs5, float alpha, int iters, float* recon) {
    const size_t n = (size_t)w * h;
    std::vector<V3<float>> b[5];
    for (int k = 0; k < 5; ++k) {
        b[k].resize(n);
        for (size_t i = 0; i < n; ++i) b[k][i] = V3<float>(bufs5[k * 3 * n + 3 * i], bufs5[k * 3 * n + 3 * i + 1], bufs5[k * 3 * n + 3 * i + 2]);
    }
    std::vector<V3<float>> gx(n, V3<float>(0)), gy(n, V3<float>(0)), cur = b[0], nxt(n, V3<float>(0));
    for (uint32_t y = 1; y + 1 < h; ++y)
        for (uint32_t x = 1; x + 1 < w; ++x) {
            size_t i = (size_t)y * w + x;
            gx[i] = b[1][i] + b[2][i];
            gy[i] = b[3][i] + b[4][i];
        }
    // the border of the iterate is zero, as in the trimmed-mean mode
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x)
            if (x == 0 || y == 0 || x + 1 == w || y + 1 == h) cur[(size_t)y * w + x] = V3<float>(0);
    const float a2 = alpha * alpha;
    const float den = a2 + 4.0f;
    for (int it = 0; it < iters; ++it) {
        for (uint32_t y = 1; y + 1 < h; ++y)
            for (uint32_t x = 1; x + 1 < w; ++x) {
                size_t i = (size_t)y * w + x;
                V3<float> v1 = cur[i - 1] + gx[i - 1];
                V3<float> v2 = cur[i - w] + gy[i - w];
                V3<float> v3 = cur[i + 1] - gx[i];
                V3<float> v4 = cur[i + w] - gy[i];
                V3<float> r;
                for (int c = 0; c < 3; ++c) r[c] = (a2 * b[0][i][c] + ((v1[c] + v2[c]) + (v3[c] + v4[c]))) / den;
                nxt[i] = r;
            }
        std::swap(cur, nxt);
    }
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) recon[3 * i + c] = cur[i][c];
}

// PostProcessor::bloom, PostProcessor.cpp:125-142 + PyramidConstructor.cu:9-169, 6 levels, FP32.
// Returns the output buffer and, optionally, every pyramid level (down then up, levels 0..5).
void orc_bloom(const float* hdr, uint32_t w, uint32_t h, const float* thr, float* out) {
    const int depth = 6;
    std::vector<std::vector<V3<float>>> down(depth), up(depth);
    uint32_t lw[depth], lh[depth];
    uint32_t cw = w, ch = h;
    for (int i = 0; i < depth; ++i) { lw[i] = cw; lh[i] = ch; down[i].assign((size_t)cw * ch, V3<float>(0)); up[i].assign((size_t)cw * ch, V3<float>(0)); cw /= 2; ch /= 2; }
    for (size_t i = 0; i < (size_t)w * h; ++i) down[0][i] = threshold_px(V3<float>(hdr[3 * i], hdr[3 * i + 1], hdr[3 * i + 2]), thr);
    for (int l = 1; l < depth; ++l)  // downsample_kernel :32-65
        for (uint32_t y = 0; y < lh[l]; ++y)
            for (uint32_t x = 0; x < lw[l]; ++x) {
                V3<float> r(0);
                for (int k = 0; k < 4; ++k) {
                    uint32_t sx = 2 * x + (k >> 1), sy = 2 * y + (k & 1);  // (0,0),(0,1),(1,0),(1,1)
                    if (sx >= lw[l - 1]) sx = lw[l - 1] - 1;
                    if (sy >= lh[l - 1]) sy = lh[l - 1] - 1;
                    r += down[l - 1][(size_t)sy * lw[l - 1] + sx];
                }
                down[l][(size_t)y * lw[l] + x] = r / 4.0f;
            }
    up[depth - 1] = down[depth - 1];
    const float k3[3] = {1.0f, 2.0f, 1.0f};
    for (int l = depth - 2; l >= 0; --l)  // upscale_and_combine_kernel :67-112 (reads down[l+1])
        for (uint32_t y = 0; y < lh[l] && lw[l + 1] > 0 && lh[l + 1] > 0; ++y)  // small images: empty levels
            for (uint32_t x = 0; x < lw[l]; ++x) {
                V3<float> r(0);
                int dx = (int)x / 2, dy = (int)y / 2;
                for (int u = -1; u < 2; ++u)
                    for (int v = -1; v < 2; ++v) {
                        int sx = dx + u, sy = dy + v;
                        if (sx < 0) sx = 0;
                        if (sx >= (int)lw[l + 1]) sx = (int)lw[l + 1] - 1;
                        if (sy < 0) sy = 0;
                        if (sy >= (int)lh[l + 1]) sy = (int)lh[l + 1] - 1;
                        r += down[l + 1][(size_t)sy * lw[l + 1] + sx] * (k3[u + 1] * k3[v + 1]);
                    }
                up[l][(size_t)y * lw[l] + x] = r / 16.0f + down[l][(size_t)y * lw[l] + x];
            }
    for (size_t i = 0; i < (size_t)w * h; ++i) {  // combine_kernel :114-125
        V3<float> o = V3<float>(hdr[3 * i], hdr[3 * i + 1], hdr[3 * i + 2]) + up[0][i];
        out[3 * i] = o.x; out[3 * i + 1] = o.y; out[3 * i + 2] = o.z;
    }
}

// Gaussian-pyramid bloom: an addition with no counterpart in the reference (the C-ABI's
// CUPBR_BLOOM_GAUSSIAN5). L0 = threshold(hdr); L_k(x, y) = blur(L_{k-1})(2x, 2y), blur = the
// separable binomial [1 4 6 4 1] / 16 with clamped edges (each row's five taps summed, then the five
// rows, then / 256); U_5 = L_5, U_k = L_k + up(U_{k+1}) with up = the coarse texel at (x/2, y/2),
// averaged with its clamped neighbour along an odd coordinate; out = hdr + U_0 / 6.
void orc_bloom_gaussian(const float* hdr, uint32_t w, uint32_t h, const float* thr, float* out) {
    const int levels = 5;
    std::vector<std::vector<V3<float>>> L(levels + 1);
    uint32_t lw[levels + 1], lh[levels + 1];
    lw[0] = w;
    lh[0] = h;
    L[0].resize((size_t)w * h);
    for (size_t i = 0; i < (size_t)w * h; ++i) L[0][i] = threshold_px(V3<float>(hdr[3 * i], hdr[3 * i + 1], hdr[3 * i + 2]), thr);
    const float wt[5] = {1.0f, 4.0f, 6.0f, 4.0f, 1.0f};
    for (int k = 1; k <= levels; ++k) {
        lw[k] = (lw[k - 1] + 1) / 2;
        lh[k] = (lh[k - 1] + 1) / 2;
        L[k].resize((size_t)lw[k] * lh[k]);
        for (uint32_t y = 0; y < lh[k]; ++y)
            for (uint32_t x = 0; x < lw[k]; ++x) {
                V3<float> acc(0);
                for (int j = 0; j < 5; ++j) {
                    const int sy = std::min(std::max((int)(2 * y) + j - 2, 0), (int)lh[k - 1] - 1);
                    V3<float> row(0);
                    for (int i = 0; i < 5; ++i) {
                        const int sx = std::min(std::max((int)(2 * x) + i - 2, 0), (int)lw[k - 1] - 1);
                        row += L[k - 1][(size_t)sy * lw[k - 1] + sx] * wt[i];
                    }
                    acc += row * wt[j];
                }
                L[k][(size_t)y * lw[k] + x] = acc / 256.0f;
            }
    }
    auto up = [&](int k, uint32_t x, uint32_t y) {  // level k sampled at level k-1's pixel (x, y)
        const uint32_t x0 = x / 2, y0 = y / 2;
        const uint32_t x1 = (x & 1u) ? std::min(x0 + 1u, lw[k] - 1u) : x0, y1 = (y & 1u) ? std::min(y0 + 1u, lh[k] - 1u) : y0;
        const auto& C = L[k];
        const V3<float> a = C[(size_t)y0 * lw[k] + x0], b = C[(size_t)y0 * lw[k] + x1];
        const V3<float> c = C[(size_t)y1 * lw[k] + x0], d = C[(size_t)y1 * lw[k] + x1];
        return ((a + b) * 0.5f + (c + d) * 0.5f) * 0.5f;
    };
    for (int k = levels - 1; k >= 1; --k)
        for (uint32_t y = 0; y < lh[k]; ++y)
            for (uint32_t x = 0; x < lw[k]; ++x) L[k][(size_t)y * lw[k] + x] += up(k + 1, x, y);
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            const size_t i = (size_t)y * w + x;
            const V3<float> u0 = L[0][i] + up(1, x, y);
            const V3<float> o = V3<float>(hdr[3 * i], hdr[3 * i + 1], hdr[3 * i + 2]) + u0 / 6.0f;
            out[3 * i] = o.x; out[3 * i + 1] = o.y; out[3 * i + 2] = o.z;
        }
}

// reinhard_kernel / gamma_kernel, lib/src/Renderer/ToneMapper.cu:17-69 (truncating uint8 casts)
void orc_tonemap(const float* hdr, uint64_t n, int type, float exposure, uint8_t* rgba) {
    for (uint64_t i = 0; i < n; ++i) {
        float c[3] = {hdr[3 * i], hdr[3 * i + 1], hdr[3 * i + 2]};
        for (int k = 0; k < 3; ++k) {
            float m;
            if (type == 0) {
                double base = 1.0 - (double)std::exp(-c[k] * exposure);
                float ex = k == 2 ? (float)(1.0f / 2.2) : 1.0f / 2.2f;
                m = std::pow((float)base, ex);
            } else if (type == 2) {  // ACES filmic fit (Narkowicz 2015), then gamma 1/2.2: an addition
                float x = c[k] * exposure;
                float f = (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
                m = std::pow(clampf(f, 0.0f, 1.0f), 1.0f / 2.2f);
            } else {
                float x = c[k];
                m = x <= 0.0031308f ? 12.92f * x : 1.055f * std::pow(x, 1.0f / 2.4f) - 0.055f;
            }
            float v = clampf(m, 0.0f, 1.0f) * 255.0f;
            rgba[4 * i + k] = (uint8_t)v;
        }
        rgba[4 * i + 3] = 255;
    }
}

}  // extern "C"
