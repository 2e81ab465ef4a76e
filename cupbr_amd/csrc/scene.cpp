#include "scene.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "bvh_build.h"

namespace cupbr {

cupbr_volume default_volume() {
    cupbr_volume v;
    for (int k = 0; k < 3; ++k) { v.sigma_s[k] = 1.0f; v.sigma_a[k] = 1.0f; }
    v.g = 0.0f;
    v.interface_type = CUPBR_INTERFACE_NONE;
    return v;
}

cupbr_material default_material() {
    cupbr_material m;
    for (int k = 0; k < 3; ++k) { m.albedo_e[k] = 0.0f; m.albedo_d[k] = 1.0f; m.albedo_s[k] = 0.0f; }
    m.shininess = 128.0f * 0.4f;
    m.roughness = 1.0f;
    m.eta = 1.5f;
    m.volume = default_volume();
    m.type = CUPBR_LAMBERT;
    return m;
}

namespace {

// The boxes the reference gives its objects: finite for spheres (SphereDetail.h:16) and
// meshes (MeshDetail.h:18), infinite otherwise (Geometry.cu:6-9).
bool object_box(const cupbr_object& o, float mn[3], float mx[3]) {
    if (o.type == CUPBR_SPHERE) {
        for (int k = 0; k < 3; ++k) { mn[k] = o.params[k] - o.params[3]; mx[k] = o.params[k] + o.params[3]; }
        return true;
    }
    if (o.type == CUPBR_MESH) {
        for (int k = 0; k < 3; ++k) { mn[k] = o.params[k]; mx[k] = o.params[3 + k]; }
        return true;
    }
    for (int k = 0; k < 3; ++k) { mn[k] = -INFINITY; mx[k] = INFINITY; }
    return false;
}

// Leaf order of the reference's object BVH (BoundingVolumeHierarchy.cpp:46-136): sort key is the
// box minimum on axis 0, stable_sort + median split above two objects, a strictly-ordered pair
// otherwise (equal keys leave the pair swapped).
void reference_order(const std::vector<float>& key, std::vector<int32_t> objs, std::vector<int32_t>& out) {
    const size_t n = objs.size();
    if (n == 0) return;
    if (n == 1) { out.push_back(objs[0]); return; }
    if (n == 2) {
        if (key[objs[0]] < key[objs[1]]) { out.push_back(objs[0]); out.push_back(objs[1]); }
        else { out.push_back(objs[1]); out.push_back(objs[0]); }
        return;
    }
    std::stable_sort(objs.begin(), objs.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
    const size_t mid = n / 2;
    reference_order(key, std::vector<int32_t>(objs.begin(), objs.begin() + mid), out);
    reference_order(key, std::vector<int32_t>(objs.begin() + mid, objs.end()), out);
}

template <typename T>
int upload(cupbr_scene* s, const std::vector<T>& host, const T** dev) {
    *dev = nullptr;
    if (host.empty()) return CUPBR_OK;
    void* p = nullptr;
    if (hipMalloc(&p, host.size() * sizeof(T)) != hipSuccess)
        return set_error(CUPBR_ERR_OUT_OF_MEMORY, "hipMalloc failed while uploading the scene");
    s->allocations.push_back(p);
    if (hipMemcpy(p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return set_error(CUPBR_ERR_HIP, "hipMemcpy failed while uploading the scene");
    *dev = (const T*)p;
    return CUPBR_OK;
}

inline v3 tri_vertex(const cupbr_scene* s, const cupbr_object& o, uint32_t tri, int k) {
    if (o.type == CUPBR_TRIANGLE) return ld3(o.params + 3 * k);
    return ld3(&s->triangles[9 * (size_t)(o.first_triangle + tri) + 3 * k]);
}

}  // namespace

void scene_release_device(cupbr_scene* s) {
    for (void* p : s->allocations) (void)hipFree(p);
    s->allocations.clear();
    s->dev = DevScene{};
    s->uploaded = false;
    if (s->stream_event) (void)hipEventDestroy(s->stream_event);
    s->stream_event = nullptr;
    s->stream_event_recorded = false;
}

int scene_stream_acquire(cupbr_scene* s, hipStream_t stream) {
    if (!s->stream_event) {
        CUPBR_HIP_CHECK(hipEventCreateWithFlags(&s->stream_event, hipEventDisableTiming));
        s->stream_event_recorded = false;
    }
    if (s->stream_event_recorded) CUPBR_HIP_CHECK(hipStreamWaitEvent(stream, s->stream_event, 0));
    return CUPBR_OK;
}

int scene_stream_release(cupbr_scene* s, hipStream_t stream) {
    CUPBR_HIP_CHECK(hipEventRecord(s->stream_event, stream));
    s->stream_event_recorded = true;
    return CUPBR_OK;
}

int scene_upload(const cupbr_scene* cs) {
    cupbr_scene* s = const_cast<cupbr_scene*>(cs);
    if (s->uploaded) return CUPBR_OK;
    auto t0 = std::chrono::steady_clock::now();
    if (hipGetDevice(&s->device) != hipSuccess) return set_error(CUPBR_ERR_HIP, "no HIP device available");
    DevScene& d = s->dev;
    int rc;
    if ((rc = upload(s, s->h_hnodes, &d.hnodes)) ||
        (rc = upload(s, s->h_prims, &d.prims)) ||
        (rc = upload(s, s->h_normals, &d.normals)) || (rc = upload(s, s->h_unb, &d.unb)) ||
        (rc = upload(s, s->h_info, &d.objs)) || (rc = upload(s, s->h_mats, &d.mats)) ||
        (rc = upload(s, s->lights, &d.lights)) || (rc = upload(s, s->env, &d.env)) ||
        (rc = upload(s, s->h_rank_prim, &d.rank_prim))) {
        scene_release_device(s);
        return rc;
    }
    d.n_nodes = (uint32_t)s->h_nodes.size();
    d.n_prims = (uint32_t)s->h_prims.size();
    d.n_unb = (uint32_t)s->h_unb.size();
    d.n_objs = (uint32_t)s->objects.size();
    d.n_lights = (uint32_t)s->lights.size();
    d.env_w = s->env.empty() ? 0 : s->env_w;
    d.env_h = s->env.empty() ? 0 : s->env_h;
    d.use_env = s->env.empty() ? 0 : 1;
    d.volume = s->volume;
    s->uploaded = true;
    s->stats.upload_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return CUPBR_OK;
}

int scene_finalize(cupbr_scene* s) {
    const uint32_t n_obj = (uint32_t)s->objects.size();
    const size_t n_tri_total = s->triangles.size() / 9;
    for (uint32_t i = 0; i < n_obj; ++i) {
        cupbr_object& o = s->objects[i];
        if (o.type < CUPBR_SPHERE || o.type > CUPBR_MESH)
            return set_error(CUPBR_ERR_GEOMETRY_TYPE, "Error while loading scene: object " + std::to_string(i) +
                                                          " is not a valid geometry type!", true);
        if (o.type == CUPBR_MESH) {
            if ((size_t)o.first_triangle + o.triangle_count > n_tri_total)
                return set_error(CUPBR_ERR_INVALID_ARGUMENT, "mesh triangle range exceeds the triangle array");
            bool bad = false;
            for (int k = 0; k < 3; ++k) bad |= !(o.params[k] <= o.params[3 + k]);
            if (bad && o.triangle_count) {  // no usable box given: the min/max over the mesh vertices
                for (int k = 0; k < 3; ++k) { o.params[k] = INFINITY; o.params[3 + k] = -INFINITY; }
                for (uint32_t t = 0; t < o.triangle_count; ++t)
                    for (int v = 0; v < 3; ++v)
                        for (int k = 0; k < 3; ++k) {
                            float x = s->triangles[9 * (size_t)(o.first_triangle + t) + 3 * v + k];
                            o.params[k] = std::min(o.params[k], x);
                            o.params[3 + k] = std::max(o.params[3 + k], x);
                        }
            }
        }
    }
    for (const cupbr_light& l : s->lights)
        if (l.type != CUPBR_POINT && l.type != CUPBR_AREA)
            return set_error(CUPBR_ERR_LIGHT_TYPE, "Error while loading scene: invalid light type!", true);

    auto t0 = std::chrono::steady_clock::now();
    // reference order and ranks
    std::vector<float> key(n_obj);
    std::vector<ObjInfo> info(n_obj);
    for (uint32_t i = 0; i < n_obj; ++i) {
        float mn[3], mx[3];
        bool fin = object_box(s->objects[i], mn, mx);
        key[i] = mn[0];
        ObjInfo& oi = info[i];
        std::memset(&oi, 0, sizeof(oi));
        for (int k = 0; k < 3; ++k) { oi.bmin[k] = mn[k]; oi.bmax[k] = mx[k]; }
        oi.type = s->objects[i].type;
        oi.finite_box = fin ? 1u : 0u;
        oi.first_triangle = s->objects[i].type == CUPBR_MESH ? s->objects[i].first_triangle : 0u;
    }
    s->ref_order.clear();
    std::vector<int32_t> all(n_obj);
    for (uint32_t i = 0; i < n_obj; ++i) all[i] = (int32_t)i;
    reference_order(key, all, s->ref_order);
    s->rank_base.assign(n_obj, 0);

    struct Src { uint32_t obj, tri, rank; };
    std::vector<Src> src;
    std::vector<BuildPrim> bprims;
    std::vector<UnbRec> unb;
    uint32_t rank = 0;
    for (int32_t obj : s->ref_order) {
        const cupbr_object& o = s->objects[obj];
        s->rank_base[obj] = rank;
        info[obj].rank_base = rank;
        if (o.type == CUPBR_PLANE || o.type == CUPBR_QUAD) {
            UnbRec u;
            v3 n = normalize(ld3(o.params + 3));  // PlaneDetail.h:14, QuadDetail.h:14
            u.p = make_float4(o.params[0], o.params[1], o.params[2], bits_to_float((uint32_t)(o.type == CUPBR_QUAD ? kUnbQuad : kUnbPlane)));
            u.n = make_float4(n.x, n.y, n.z, bits_to_float(rank));
            u.e1 = make_float4(o.params[6], o.params[7], o.params[8], bits_to_float((uint32_t)obj));
            u.e2 = make_float4(o.params[9], o.params[10], o.params[11], 0.0f);
            if (o.type == CUPBR_PLANE) u.e1 = make_float4(0, 0, 0, bits_to_float((uint32_t)obj)), u.e2 = make_float4(0, 0, 0, 0);
            unb.push_back(u);
            info[obj].rank_count = 1;
            rank += 1;
            continue;
        }
        const uint32_t cnt = o.type == CUPBR_MESH ? o.triangle_count : 1u;
        info[obj].rank_count = cnt;
        for (uint32_t t = 0; t < cnt; ++t) {
            BuildPrim bp;
            if (o.type == CUPBR_SPHERE) {
                for (int k = 0; k < 3; ++k) { bp.mn[k] = info[obj].bmin[k]; bp.mx[k] = info[obj].bmax[k]; }
            } else {
                v3 a = tri_vertex(s, o, t, 0), b = tri_vertex(s, o, t, 1), c = tri_vertex(s, o, t, 2);
                bp.mn[0] = std::min(a.x, std::min(b.x, c.x)); bp.mx[0] = std::max(a.x, std::max(b.x, c.x));
                bp.mn[1] = std::min(a.y, std::min(b.y, c.y)); bp.mx[1] = std::max(a.y, std::max(b.y, c.y));
                bp.mn[2] = std::min(a.z, std::min(b.z, c.z)); bp.mx[2] = std::max(a.z, std::max(b.z, c.z));
            }
            bool finite = true;
            for (int k = 0; k < 3; ++k) finite &= std::isfinite(bp.mn[k]) && std::isfinite(bp.mx[k]);
            if (!finite) continue;  // a primitive with non-finite coordinates cannot be hit
            src.push_back({(uint32_t)obj, t, rank + t});
            bprims.push_back(bp);
        }
        rank += cnt;
    }

    // leaf words hold the first primitive in 27 bits; inner-node loads use 32-bit byte offsets
    if (bprims.size() >= (size_t(1) << 27))
        return set_error(CUPBR_ERR_INVALID_ARGUMENT, "scene too large: at most 2^27 bounded primitives");
    BvhBuildOptions opt;
    Bvh4 bvh = build_bvh4(bprims, opt);
    if (bvh.nodes.size() >= (size_t(1) << 25))
        return set_error(CUPBR_ERR_INVALID_ARGUMENT, "scene too large: its BVH needs more than 2^25 nodes");
    std::vector<uint32_t> rank_prim(rank, 0xFFFFFFFFu);
    for (size_t k = 0; k < bvh.order.size(); ++k) rank_prim[src[bvh.order[k]].rank] = (uint32_t)k;
    for (size_t k = 0; k < unb.size(); ++k) rank_prim[float_to_bits(unb[k].n.w)] = kUnbBit | (uint32_t)k;
    std::vector<PrimRec> prims(bvh.order.size());
    std::vector<float4> normals(bvh.order.size());
    for (size_t k = 0; k < bvh.order.size(); ++k) {
        const Src& sp = src[bvh.order[k]];
        const cupbr_object& o = s->objects[sp.obj];
        PrimRec pr;
        uint32_t flags;
        if (o.type == CUPBR_SPHERE) {
            pr.a = make_float4(o.params[0], o.params[1], o.params[2], o.params[3]);
            pr.b = make_float4(0, 0, 0, 0);
            pr.c.x = 0.0f;
            flags = kPrimSphere | kPrimNeedsBoxCheck;
            normals[k] = make_float4(0, 0, 0, bits_to_float(sp.obj | kNormalSphereBit));
        } else {
            v3 a = tri_vertex(s, o, sp.tri, 0), b = tri_vertex(s, o, sp.tri, 1), c = tri_vertex(s, o, sp.tri, 2);
            v3 e1 = b - a, e2 = c - a;  // TriangleDetail.h:27-28
            v3 n = normalize(cross(e1, e2));  // TriangleDetail.h:16
            pr.a = make_float4(a.x, a.y, a.z, e1.x);
            pr.b = make_float4(e1.y, e1.z, e2.x, e2.y);
            pr.c.x = e2.z;
            flags = kPrimTriangle | (o.type == CUPBR_MESH ? kPrimNeedsBoxCheck : 0u);
            normals[k] = make_float4(n.x, n.y, n.z, bits_to_float(sp.obj));
        }
        pr.c.y = bits_to_float(sp.rank);
        pr.c.z = bits_to_float(sp.obj);
        pr.c.w = bits_to_float(flags);
        prims[k] = pr;
    }
    auto t1 = std::chrono::steady_clock::now();
    std::vector<cupbr_material> mats(n_obj);
    for (uint32_t i = 0; i < n_obj; ++i) mats[i] = s->objects[i].material;
    scene_release_device(s);
    s->h_nodes = std::move(bvh.nodes);
    s->h_hnodes = half_nodes(s->h_nodes);
    s->h_prims = std::move(prims);
    s->h_normals = std::move(normals);
    s->h_unb = std::move(unb);
    s->h_info = std::move(info);
    s->h_mats = std::move(mats);
    s->h_rank_prim = std::move(rank_prim);
    s->stack_entries = 3 * bvh.depth + 1;
    cupbr_bvh_stats& st = s->stats;
    st.nodes = (uint32_t)s->h_nodes.size();
    st.leaves = bvh.leaves;
    st.primitives = (uint32_t)s->h_prims.size();
    st.unbounded = (uint32_t)s->h_unb.size();
    st.depth = bvh.depth;
    st.stack_entries = s->stack_entries;
    st.build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    st.upload_ms = 0.0;
    st.sah_cost = bvh.sah_cost;
    return CUPBR_OK;
}

}  // namespace cupbr
