// extern "C" entry points (include/cupbr.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "nrc.h"
#include "post.h"
#include "renderer.h"
#include "scene.h"

namespace cupbr {

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg, bool print) {
    g_last_error = msg;
    if (print) std::cerr << msg << std::endl;
    return code;
}

int device_cu_count() {
    static std::mutex mu;
    static std::vector<int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)cache.size() <= dev) cache.resize(dev + 1, 0);
    if (cache[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 1;
        cache[dev] = n;
    }
    return cache[dev];
}

int renderer_clear_images(cupbr_renderer* r) {
    // setMethod clears eight of the nine images (PBRenderer.cu:125-134); base is cleared at allocation
    for (int i = 0; i < kNumImages; ++i) {
        if (i == CUPBR_IMAGE_BASE) continue;
        if (r->images[i].ptr) CUPBR_HIP_CHECK(hipMemsetAsync(r->images[i].ptr, 0, r->images[i].bytes, r->stream));
    }
    if (r->temp.ptr) CUPBR_HIP_CHECK(hipMemsetAsync(r->temp.ptr, 0, r->temp.bytes, r->stream));
    return CUPBR_OK;
}

}  // namespace cupbr

using namespace cupbr;

#define CHECK_ARG(cond, msg) \
    if (!(cond)) return set_error(CUPBR_ERR_INVALID_ARGUMENT, msg)

extern "C" {

const char* cupbr_last_error(void) { return g_last_error.c_str(); }

int cupbr_device_name(char* out, size_t n) {
    CHECK_ARG(out && n, "null buffer");
    int dev = 0;
    CUPBR_HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    CUPBR_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    snprintf(out, n, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return CUPBR_OK;
}

// ------------------------------------------------------------------ camera
int cupbr_camera_init(cupbr_camera* cam, uint32_t width, uint32_t height) {
    CHECK_ARG(cam && height, "invalid camera arguments");
    std::memset(cam, 0, sizeof *cam);
    const float aspect = (float)width / (float)height;  // Camera.cpp:8-10
    cam->x_axis[0] = aspect;
    cam->y_axis[1] = 1.0f;
    cam->z_axis[2] = 1.0f;
    return CUPBR_OK;
}

int cupbr_camera_look_at(cupbr_camera* cam, const float eye[3], const float target[3], const float up[3],
                         float vfov_degrees, uint32_t width, uint32_t height) {
    CHECK_ARG(cam && eye && target && up && height && vfov_degrees > 0.0f && vfov_degrees < 180.0f, "invalid camera arguments");
    std::memset(cam, 0, sizeof *cam);
    const float aspect = (float)width / (float)height;
    v3 f = normalize(ld3(target) - ld3(eye));
    // right = up x forward keeps the reference's handedness (x axis = (z.z, 0, -z.x) for up=+y, Camera.cpp:87)
    v3 rgt = normalize(cross(ld3(up), f));
    v3 u = cross(f, rgt);
    const float th = (float)std::tan(0.5 * (double)vfov_degrees * kPi / 180.0);
    st3(cam->position, ld3(eye));
    st3(cam->z_axis, f);
    st3(cam->x_axis, (aspect * th) * rgt);
    st3(cam->y_axis, th * u);
    return CUPBR_OK;
}

int cupbr_camera_controller_init(cupbr_camera_controller* ctl, uint32_t width, uint32_t height) {
    CHECK_ARG(ctl && height, "invalid controller arguments");
    ctl->aspect_ratio = (float)width / (float)height;
    ctl->pitch = 0.0f;
    ctl->yaw = 3.14159f / 2.0f;
    ctl->first_mouse = 1;
    ctl->last_x = ctl->last_y = 0.0f;
    return CUPBR_OK;
}

// Camera::processInput, Camera.cpp:13-90 (operation order kept: the axes must follow the
// reference's for the same inputs)
int cupbr_camera_process_input(cupbr_camera* cam, cupbr_camera_controller* ctl, const cupbr_input_state* in,
                               float delta_time) {
    CHECK_ARG(cam && ctl && in, "null argument");
    bool moved = false;
    float speed = 4.0f;
    if (in->keys & CUPBR_KEY_LEFT_SHIFT) speed *= 4;
    if (in->keys & CUPBR_KEY_LEFT_CONTROL) speed /= 4;
    v3 pos = ld3(cam->position);
    const v3 xa = ld3(cam->x_axis), za = ld3(cam->z_axis);
    if (in->keys & CUPBR_KEY_D) { pos += speed * xa * delta_time; moved = true; }
    if (in->keys & CUPBR_KEY_A) { pos = pos - speed * xa * delta_time; moved = true; }
    if (in->keys & CUPBR_KEY_S) { pos = pos - speed * za * delta_time; moved = true; }
    if (in->keys & CUPBR_KEY_W) { pos += speed * za * delta_time; moved = true; }
    st3(cam->position, pos);
    const float cx = (float)in->cursor_x, cy = (float)in->cursor_y;
    if (ctl->first_mouse) {
        ctl->last_x = cx;
        ctl->last_y = cy;
        ctl->first_mouse = 0;
    }
    const float sensitivity = 0.002f;
    float xoff = ctl->last_x - cx, yoff = ctl->last_y - cy;
    if (!(safe_float_equal(xoff, 0.0f) && safe_float_equal(yoff, 0.0f))) moved = true;
    ctl->last_x = cx;
    ctl->last_y = cy;
    xoff *= sensitivity;
    yoff *= sensitivity;
    ctl->yaw += xoff;
    ctl->pitch += yoff;
    if (ctl->pitch > 3.14159f / 2.0f) ctl->pitch = 3.14159f / 2.0f;
    if (ctl->pitch < -3.14159f / 2.0f) ctl->pitch = -3.14159f / 2.0f;
    const v3 front = mk(std::cos(ctl->yaw) * std::cos(ctl->pitch), std::sin(ctl->pitch),
                        std::sin(ctl->yaw) * std::cos(ctl->pitch));
    const v3 xn = ctl->aspect_ratio * normalize(mk(front.z, 0.0f, -front.x));
    st3(cam->z_axis, front);
    st3(cam->x_axis, xn);
    st3(cam->y_axis, cross(front, xn) / ctl->aspect_ratio);
    cam->moved = moved ? 1 : 0;
    return CUPBR_OK;
}

// ------------------------------------------------------------------ scene
int cupbr_scene_load_xml(const char* path, cupbr_scene** out) {
    CHECK_ARG(path && out, "null argument");
    *out = nullptr;
    auto* s = new (std::nothrow) cupbr_scene();
    if (!s) return set_error(CUPBR_ERR_OUT_OF_MEMORY, "out of host memory");
    int rc = load_scene_xml(path, s);
    if (rc == CUPBR_OK) rc = scene_finalize(s);
    if (rc != CUPBR_OK) {
        scene_release_device(s);
        delete s;
        return rc;
    }
    *out = s;
    return CUPBR_OK;
}

int cupbr_scene_create(const cupbr_object* objects, uint32_t n_objects, const float* triangles, uint32_t n_triangles,
                       const cupbr_light* lights, uint32_t n_lights, const float* env_rgb, uint32_t env_w,
                       uint32_t env_h, const cupbr_volume* volume, cupbr_scene** out) {
    CHECK_ARG(out, "null output handle");
    *out = nullptr;
    CHECK_ARG(n_objects == 0 || objects, "objects is null");
    CHECK_ARG(n_triangles == 0 || triangles, "triangles is null");
    CHECK_ARG(n_lights == 0 || lights, "lights is null");
    CHECK_ARG(!env_rgb || (env_w && env_h), "environment map without a size");
    auto* s = new (std::nothrow) cupbr_scene();
    if (!s) return set_error(CUPBR_ERR_OUT_OF_MEMORY, "out of host memory");
    s->objects.assign(objects, objects + n_objects);
    s->triangles.assign(triangles, triangles + 9 * (size_t)n_triangles);
    s->lights.assign(lights, lights + n_lights);
    if (env_rgb) {
        s->env.assign(env_rgb, env_rgb + 3 * (size_t)env_w * env_h);
        s->env_w = env_w;
        s->env_h = env_h;
    }
    s->volume = volume ? *volume : default_volume();
    int rc = scene_finalize(s);
    if (rc != CUPBR_OK) {
        scene_release_device(s);
        delete s;
        return rc;
    }
    *out = s;
    return CUPBR_OK;
}

int cupbr_scene_upload(cupbr_scene* s) {
    CHECK_ARG(s, "null scene");
    return scene_upload(s);
}

int cupbr_scene_destroy(cupbr_scene* s) {
    if (!s) return CUPBR_OK;
    scene_release_device(s);
    delete s;
    return CUPBR_OK;
}

int cupbr_scene_get_counts(const cupbr_scene* s, uint32_t* n_objects, uint32_t* n_triangles, uint32_t* n_lights,
                           int32_t* use_env) {
    CHECK_ARG(s, "null scene");
    if (n_objects) *n_objects = (uint32_t)s->objects.size();
    if (n_triangles) *n_triangles = (uint32_t)(s->triangles.size() / 9);
    if (n_lights) *n_lights = (uint32_t)s->lights.size();
    if (use_env) *use_env = s->env.empty() ? 0 : 1;
    return CUPBR_OK;
}

int cupbr_scene_get_objects(const cupbr_scene* s, cupbr_object* out, uint32_t n) {
    CHECK_ARG(s && out, "null argument");
    CHECK_ARG(n >= s->objects.size(), "buffer too small");
    std::memcpy(out, s->objects.data(), s->objects.size() * sizeof(cupbr_object));
    return CUPBR_OK;
}

int cupbr_scene_get_triangles(const cupbr_scene* s, float* out, uint32_t n_triangles) {
    CHECK_ARG(s && (out || s->triangles.empty()), "null argument");
    CHECK_ARG((size_t)n_triangles * 9 >= s->triangles.size(), "buffer too small");
    if (!s->triangles.empty()) std::memcpy(out, s->triangles.data(), s->triangles.size() * sizeof(float));
    return CUPBR_OK;
}

int cupbr_scene_get_lights(const cupbr_scene* s, cupbr_light* out, uint32_t n) {
    CHECK_ARG(s && (out || s->lights.empty()), "null argument");
    CHECK_ARG(n >= s->lights.size(), "buffer too small");
    if (!s->lights.empty()) std::memcpy(out, s->lights.data(), s->lights.size() * sizeof(cupbr_light));
    return CUPBR_OK;
}

int cupbr_scene_get_volume(const cupbr_scene* s, cupbr_volume* out) {
    CHECK_ARG(s && out, "null argument");
    *out = s->volume;
    return CUPBR_OK;
}

int cupbr_scene_get_environment(const cupbr_scene* s, float* out, size_t n, uint32_t* w, uint32_t* h) {
    CHECK_ARG(s, "null scene");
    if (w) *w = s->env.empty() ? 0 : s->env_w;
    if (h) *h = s->env.empty() ? 0 : s->env_h;
    if (out) {
        CHECK_ARG(n >= s->env.size(), "buffer too small");
        if (!s->env.empty()) std::memcpy(out, s->env.data(), s->env.size() * sizeof(float));
    }
    return CUPBR_OK;
}

int cupbr_scene_get_bvh_stats(const cupbr_scene* s, cupbr_bvh_stats* out) {
    CHECK_ARG(s && out, "null argument");
    *out = s->stats;
    return CUPBR_OK;
}

int cupbr_scene_get_bvh_nodes(const cupbr_scene* s, void* fp32_nodes, void* half_nodes, uint32_t n) {
    CHECK_ARG(s, "null argument");
    CHECK_ARG(n >= s->h_nodes.size(), "buffer too small");
    if (fp32_nodes && !s->h_nodes.empty())
        std::memcpy(fp32_nodes, s->h_nodes.data(), s->h_nodes.size() * sizeof(cupbr::Node4));
    if (half_nodes && !s->h_hnodes.empty())
        std::memcpy(half_nodes, s->h_hnodes.data(), s->h_hnodes.size() * sizeof(cupbr::NodeH));
    return CUPBR_OK;
}

int cupbr_scene_get_reference_order(const cupbr_scene* s, int32_t* order, uint32_t n) {
    CHECK_ARG(s && (order || s->ref_order.empty()), "null argument");
    CHECK_ARG(n >= s->ref_order.size(), "buffer too small");
    if (!s->ref_order.empty()) std::memcpy(order, s->ref_order.data(), s->ref_order.size() * sizeof(int32_t));
    return CUPBR_OK;
}

// ------------------------------------------------------------------ batch queries
int cupbr_trace_closest_device(const cupbr_scene* s, const float* rays, size_t n, int32_t* obj, int32_t* tri, float* t,
                               void* stream) {
    CHECK_ARG(s, "null scene");
    CHECK_ARG(n == 0 || (rays && obj && tri && t), "null buffer");
    return trace_closest_device(s, rays, n, obj, tri, t, (hipStream_t)stream);
}

int cupbr_trace_visibility_device(const cupbr_scene* s, const float* rays, const float* dist, size_t n, uint8_t* vis,
                                  void* stream) {
    CHECK_ARG(s, "null scene");
    CHECK_ARG(n == 0 || (rays && dist && vis), "null buffer");
    return trace_visibility_device(s, rays, dist, n, vis, (hipStream_t)stream);
}

int cupbr_trace_closest(const cupbr_scene* s, const float* rays, size_t n, int32_t* obj, int32_t* tri, float* t) {
    CHECK_ARG(s, "null scene");
    CHECK_ARG(n == 0 || (rays && obj && tri && t), "null buffer");
    if (n == 0) return CUPBR_OK;
    DevBuf d_rays, d_obj, d_tri, d_t;
    int rc;
    if ((rc = d_rays.ensure(n * 24)) || (rc = d_obj.ensure(n * 4)) || (rc = d_tri.ensure(n * 4)) || (rc = d_t.ensure(n * 4)))
        return rc;
    CUPBR_HIP_CHECK(hipMemcpy(d_rays.ptr, rays, n * 24, hipMemcpyHostToDevice));
    if ((rc = trace_closest_device(s, d_rays.as<float>(), n, d_obj.as<int32_t>(), d_tri.as<int32_t>(), d_t.as<float>(), nullptr)))
        return rc;
    CUPBR_HIP_CHECK(hipMemcpy(obj, d_obj.ptr, n * 4, hipMemcpyDeviceToHost));
    CUPBR_HIP_CHECK(hipMemcpy(tri, d_tri.ptr, n * 4, hipMemcpyDeviceToHost));
    CUPBR_HIP_CHECK(hipMemcpy(t, d_t.ptr, n * 4, hipMemcpyDeviceToHost));
    return CUPBR_OK;
}

int cupbr_trace_visibility(const cupbr_scene* s, const float* rays, const float* dist, size_t n, uint8_t* vis) {
    CHECK_ARG(s, "null scene");
    CHECK_ARG(n == 0 || (rays && dist && vis), "null buffer");
    if (n == 0) return CUPBR_OK;
    DevBuf d_rays, d_dist, d_vis;
    int rc;
    if ((rc = d_rays.ensure(n * 24)) || (rc = d_dist.ensure(n * 4)) || (rc = d_vis.ensure(n))) return rc;
    CUPBR_HIP_CHECK(hipMemcpy(d_rays.ptr, rays, n * 24, hipMemcpyHostToDevice));
    CUPBR_HIP_CHECK(hipMemcpy(d_dist.ptr, dist, n * 4, hipMemcpyHostToDevice));
    if ((rc = trace_visibility_device(s, d_rays.as<float>(), d_dist.as<float>(), n, d_vis.as<uint8_t>(), nullptr))) return rc;
    CUPBR_HIP_CHECK(hipMemcpy(vis, d_vis.ptr, n, hipMemcpyDeviceToHost));
    return CUPBR_OK;
}

// ------------------------------------------------------------------ picking (Interaction, MousePicker.cu)
// pickMouse (MousePicker.cu:8-35,54-73): the unjittered ray through pixel (x, y) (TracingDetail.h:34-45);
// on a hit the object's material and scene index are written, on a miss both are left unchanged
int cupbr_pick(const cupbr_scene* s, const cupbr_camera* cam, uint32_t x, uint32_t y, uint32_t width,
               uint32_t height, cupbr_material* out_material, int32_t* out_scene_index) {
    CHECK_ARG(s && cam && out_material && out_scene_index, "null argument");
    CHECK_ARG(width && height, "invalid size");
    const float rx = 2.0f * (((float)x + 0.0f) / (float)width - 0.5f);
    const float ry = 2.0f * (((float)y + 0.0f) / (float)height - 0.5f);
    float wp[3], d[3], n2 = 0.0f;
    for (int k = 0; k < 3; ++k) {
        wp[k] = cam->position[k] + cam->z_axis[k] + rx * cam->x_axis[k] + ry * cam->y_axis[k];
        d[k] = wp[k] - cam->position[k];
    }
    n2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const float nrm = std::sqrt(n2);
    const float ray[6] = {cam->position[0], cam->position[1], cam->position[2], d[0] / nrm, d[1] / nrm, d[2] / nrm};
    int32_t obj = -1, tri = -1;
    float t = 0.0f;
    if (int rc = cupbr_trace_closest(s, ray, 1, &obj, &tri, &t)) return rc;
    if (obj >= 0) {
        *out_material = s->objects[obj].material;
        *out_scene_index = obj;
    }
    return CUPBR_OK;
}

// updateMaterial (MousePicker.cu:37-51,76-84): replaces the material of object scene_index; a
// renderer accumulating this scene should be reset by the caller (as the reference's UI does)
int cupbr_scene_update_material(cupbr_scene* s, int32_t scene_index, const cupbr_material* m) {
    CHECK_ARG(s && m, "null argument");
    CHECK_ARG(scene_index >= 0 && (size_t)scene_index < s->objects.size(), "scene index out of range");
    s->objects[scene_index].material = *m;
    s->h_mats[scene_index] = *m;
    if (s->uploaded) {
        CUPBR_HIP_CHECK(hipDeviceSynchronize());
        CUPBR_HIP_CHECK(hipMemcpy(const_cast<cupbr_material*>(s->dev.mats) + scene_index, m, sizeof(cupbr_material),
                                  hipMemcpyHostToDevice));
    }
    return CUPBR_OK;
}

// ------------------------------------------------------------------ renderer
static bool valid_method(int32_t m) { return m >= CUPBR_RAYTRACER && m <= CUPBR_VOLUME; }

int cupbr_renderer_create(int32_t method, cupbr_renderer** out) {
    CHECK_ARG(out, "null output handle");
    *out = nullptr;
    CHECK_ARG(valid_method(method), "unknown rendering method");
    auto* r = new (std::nothrow) cupbr_renderer();
    if (!r) return set_error(CUPBR_ERR_OUT_OF_MEMORY, "out of host memory");
    r->method = method;
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&r->ev_start) != hipSuccess || hipEventCreate(&r->ev_stop) != hipSuccess) {
        delete r;
        return set_error(CUPBR_ERR_HIP, "could not create a HIP stream (no device?)");
    }
    *out = r;
    return CUPBR_OK;
}

int cupbr_renderer_destroy(cupbr_renderer* r) {
    if (!r) return CUPBR_OK;
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    for (auto& im : r->images) im.release();
    if (r->ev_start) (void)hipEventDestroy(r->ev_start);
    if (r->ev_stop) (void)hipEventDestroy(r->ev_stop);
    if (r->side_stream) (void)hipStreamSynchronize(r->side_stream);
    if (r->ev_queries) (void)hipEventDestroy(r->ev_queries);
    if (r->ev_side_done) (void)hipEventDestroy(r->ev_side_done);
    hipStream_t st = r->stream, side = r->side_stream;
    delete r;
    if (st) (void)hipStreamDestroy(st);
    if (side) (void)hipStreamDestroy(side);
    return CUPBR_OK;
}

int cupbr_renderer_set_output_size(cupbr_renderer* r, uint32_t width, uint32_t height) {
    CHECK_ARG(r, "null renderer");
    CHECK_ARG(width && height && (uint64_t)width * height <= (1ull << 28), "invalid output size");
    CUPBR_HIP_CHECK(hipStreamSynchronize(r->stream));
    const size_t bytes = (size_t)width * height * 12;
    for (auto& im : r->images) {
        im.release();
        int rc = im.ensure(bytes);
        if (rc) return rc;
        CUPBR_HIP_CHECK(hipMemsetAsync(im.ptr, 0, bytes, r->stream));
    }
    r->temp.release();
    int rc = r->temp.ensure(bytes);
    if (rc) return rc;
    CUPBR_HIP_CHECK(hipMemsetAsync(r->temp.ptr, 0, bytes, r->stream));
    r->width = width;
    r->height = height;
    r->output_size_set = true;
    return CUPBR_OK;
}

int cupbr_renderer_register_scene(cupbr_renderer* r, const cupbr_scene* s) {
    CHECK_ARG(r && s, "null argument");
    r->scene = s;
    return CUPBR_OK;
}

int cupbr_renderer_set_method(cupbr_renderer* r, int32_t method) {
    CHECK_ARG(r, "null renderer");
    CHECK_ARG(valid_method(method), "unknown rendering method");
    r->method = method;
    r->frame_index = 0;
    return renderer_clear_images(r);
}

int cupbr_renderer_get_method(const cupbr_renderer* r, int32_t* m) {
    CHECK_ARG(r && m, "null argument");
    *m = r->method;
    return CUPBR_OK;
}

int cupbr_renderer_reset(cupbr_renderer* r) {
    CHECK_ARG(r, "null renderer");
    r->frame_index = 0;
    return CUPBR_OK;
}

int cupbr_renderer_set_russian_roulette(cupbr_renderer* r, int32_t enable) {
    CHECK_ARG(r, "null renderer");
    r->russian_roulette = enable != 0;
    return CUPBR_OK;
}

int cupbr_renderer_set_max_trace_depth(cupbr_renderer* r, uint32_t depth) {
    CHECK_ARG(r, "null renderer");
    r->max_trace_depth = depth;
    return CUPBR_OK;
}

int cupbr_renderer_get_max_trace_depth(const cupbr_renderer* r, uint32_t* depth) {
    CHECK_ARG(r && depth, "null argument");
    *depth = r->max_trace_depth;
    return CUPBR_OK;
}

int cupbr_renderer_get_frame_index(const cupbr_renderer* r, uint32_t* f) {
    CHECK_ARG(r && f, "null argument");
    *f = r->frame_index;
    return CUPBR_OK;
}

int cupbr_renderer_get_output_size(const cupbr_renderer* r, uint32_t* width, uint32_t* height) {
    CHECK_ARG(r && width && height, "null argument");
    *width = r->output_size_set ? r->width : 0;
    *height = r->output_size_set ? r->height : 0;
    return CUPBR_OK;
}

// PBRenderer::render, PBRenderer.cu:137-230
int cupbr_render(cupbr_renderer* r, const cupbr_camera* cam, uint32_t spp) {
    CHECK_ARG(r && cam, "null argument");
    if (!r->output_size_set)
        return set_error(CUPBR_ERR_NO_OUTPUT_SIZE, "[PBRenderer]  No output size set. Use setOutputSize()", true);
    if (!r->scene) return set_error(CUPBR_ERR_NO_SCENE, "[PBRenderer]  No scene registered. Use registerScene()", true);
    if (r->method == CUPBR_WHITTED)
        return set_error(CUPBR_ERR_UNSUPPORTED, "[PBRenderer]  WHITTED currently disabled.", true);
    if (r->method == CUPBR_METROPOLIS)
        return set_error(CUPBR_ERR_UNSUPPORTED, "[PBRenderer]  METROPOLIS not supported.", true);
    if (r->method == CUPBR_GRADIENTDOMAIN && r->max_trace_depth > 10)
        return set_error(CUPBR_ERR_INVALID_ARGUMENT, "gradient-domain paths store at most 10 vertices (max trace depth <= 10)");
    if (int rc = scene_upload(r->scene)) return rc;
    if (cam->moved) r->frame_index = 0;  // :174-177
    r->stats = cupbr_render_stats{};
    r->nrc_queries_total = r->nrc_rows_total = 0;
    r->timer.used = 0;
    CUPBR_HIP_CHECK(hipEventRecord(r->ev_start, r->stream));
    int rc = CUPBR_OK;
    switch (r->method) {
        case CUPBR_RAYTRACER:
            for (uint32_t s = 0; s < std::max<uint32_t>(spp, 1) && rc == CUPBR_OK; ++s) rc = render_raytracer(r, *cam);
            break;
        case CUPBR_PATHTRACER: rc = render_pathtracer(r, *cam, r->frame_index, spp); break;
        case CUPBR_GRADIENTDOMAIN: rc = render_gdpt(r, *cam, r->frame_index, spp); break;
        case CUPBR_VOLUME: rc = render_volume(r, *cam, r->frame_index, spp); break;
    }
    if (rc != CUPBR_OK) return rc;
    CUPBR_HIP_CHECK(hipEventRecord(r->ev_stop, r->stream));
    if (r->method != CUPBR_RAYTRACER) r->frame_index += spp;
    r->stats.samples = (uint64_t)spp * r->width * r->height;
    return CUPBR_OK;
}

int cupbr_renderer_synchronize(cupbr_renderer* r) {
    CHECK_ARG(r, "null renderer");
    CUPBR_HIP_CHECK(hipStreamSynchronize(r->stream));
    return CUPBR_OK;
}

int cupbr_renderer_get_stats(const cupbr_renderer* cr, cupbr_render_stats* out) {
    CHECK_ARG(cr && out, "null argument");
    cupbr_renderer* r = const_cast<cupbr_renderer*>(cr);
    CUPBR_HIP_CHECK(hipStreamSynchronize(r->stream));
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, r->ev_start, r->ev_stop) == hipSuccess) r->stats.milliseconds = ms;
    if (r->wf.counters.ptr && (r->method == CUPBR_PATHTRACER || r->method == CUPBR_VOLUME || r->method == CUPBR_GRADIENTDOMAIN)) {
        Counters c;
        CUPBR_HIP_CHECK(hipMemcpy(&c, r->wf.counters.ptr, sizeof c, hipMemcpyDeviceToHost));
        r->stats.extension_rays = c.extension_rays;
        r->stats.shadow_rays = c.shadow_rays;
        r->stats.extension_node_visits = c.ext_nodes;
        r->stats.extension_prim_tests = c.ext_prims;
        r->stats.shadow_node_visits = c.shadow_nodes;
        r->stats.shadow_prim_tests = c.shadow_prims;
        r->stats.refill_launches = c.refill_launches;
    }
    r->stats.cache_queries = r->nrc_queries_total;
    r->stats.cache_training_rows = r->nrc_rows_total;
    for (int k = 0; k < 4; ++k) { r->stats.stage_ms[k] = 0.0; r->stats.stage_launches[k] = 0; }
    for (size_t i = 1; i < r->timer.used; ++i) {
        const int st = r->timer.stage[i];
        if (st < 0 || st > 3) continue;
        float e = 0.0f;
        CUPBR_HIP_CHECK(hipEventElapsedTime(&e, r->timer.pool[i - 1], r->timer.pool[i]));
        r->stats.stage_ms[st] += e;
        r->stats.stage_launches[st] += 1;
    }
    *out = r->stats;
    return CUPBR_OK;
}

int cupbr_renderer_set_gdpt_reconstruction(cupbr_renderer* r, int32_t mode, float alpha, uint32_t iterations) {
    CHECK_ARG(r, "null renderer");
    CHECK_ARG(mode == CUPBR_GDPT_TRIMMED_MEAN || mode == CUPBR_GDPT_SCREENED_POISSON, "unknown reconstruction mode");
    if (mode == CUPBR_GDPT_SCREENED_POISSON) {
        CHECK_ARG(std::isfinite(alpha) && alpha >= 0.0f, "alpha must be finite and >= 0");
        CHECK_ARG(iterations <= (1u << 20), "at most 2^20 iterations");
        r->gdpt_alpha = alpha;
        r->gdpt_iterations = iterations;
    }
    r->gdpt_mode = mode;
    return CUPBR_OK;
}

int cupbr_renderer_set_rng(cupbr_renderer* r, int32_t rng) {
    CHECK_ARG(r, "null renderer");
    CHECK_ARG(rng == CUPBR_RNG_TEA_LCG || rng == CUPBR_RNG_PCG, "unknown random-number generator");
    CUPBR_HIP_CHECK(hipStreamSynchronize(r->stream));
    r->rng = rng;
    return CUPBR_OK;
}

int cupbr_renderer_set_profiling(cupbr_renderer* r, int32_t enable) {
    CHECK_ARG(r, "null renderer");
    r->profiling = enable != 0;
    return CUPBR_OK;
}

int cupbr_renderer_set_traversal_counting(cupbr_renderer* r, int32_t enable) {
    CHECK_ARG(r, "null renderer");
    r->counting = enable != 0;
    return CUPBR_OK;
}

int cupbr_renderer_get_image_device(const cupbr_renderer* r, int32_t which, float** dev) {
    CHECK_ARG(r && dev, "null argument");
    CHECK_ARG(which >= 0 && which < kNumImages, "unknown image");
    if (!r->output_size_set)
        return set_error(CUPBR_ERR_NO_OUTPUT_SIZE, "[PBRenderer]  No output size set. Use setOutputSize()", true);
    *dev = r->images[which].as<float>();
    return CUPBR_OK;
}

int cupbr_renderer_get_image(const cupbr_renderer* r, int32_t which, float* host, size_t n_floats) {
    float* dev = nullptr;
    int rc = cupbr_renderer_get_image_device(r, which, &dev);
    if (rc) return rc;
    CHECK_ARG(host, "null buffer");
    const size_t need = (size_t)r->width * r->height * 3;
    if (n_floats < need) return set_error(CUPBR_ERR_BUFFER_SIZE, "host buffer too small for the image");
    CUPBR_HIP_CHECK(hipStreamSynchronize(r->stream));
    CUPBR_HIP_CHECK(hipMemcpy(host, dev, need * sizeof(float), hipMemcpyDeviceToHost));
    return CUPBR_OK;
}

// PFM (Portable Float Map, "PF" colour): header "PF\n<w> <h>\n-1.0\n" (negative scale = little
// endian), then rows from the bottom up, which is the renderer's own row order
int cupbr_renderer_save_pfm(const cupbr_renderer* r, int32_t which, const char* path) {
    CHECK_ARG(r && path, "null argument");
    const size_t n = (size_t)r->width * r->height * 3;
    std::vector<float> px(n);
    if (int rc = cupbr_renderer_get_image(r, which, px.data(), n)) return rc;
    FILE* f = std::fopen(path, "wb");
    if (!f) return set_error(CUPBR_ERR_FILE, std::string("cannot write ") + path);
    std::fprintf(f, "PF\n%u %u\n-1.0\n", r->width, r->height);
    const size_t wrote = std::fwrite(px.data(), sizeof(float), n, f);
    const int closed = std::fclose(f);
    if (wrote != n || closed != 0) return set_error(CUPBR_ERR_FILE, std::string("write failed: ") + path);
    return CUPBR_OK;
}

// restores an accumulation state: the output image (host FP32 RGB, width*height*3) and the frame
// index, so that the next render continues the running mean where a saved run stopped
int cupbr_renderer_resume(cupbr_renderer* r, const float* host_rgb, size_t n_floats, uint32_t frame_index) {
    CHECK_ARG(r && host_rgb, "null argument");
    if (!r->output_size_set)
        return set_error(CUPBR_ERR_NO_OUTPUT_SIZE, "[PBRenderer]  No output size set. Use setOutputSize()", true);
    const size_t need = (size_t)r->width * r->height * 3;
    if (n_floats < need) return set_error(CUPBR_ERR_BUFFER_SIZE, "image buffer smaller than width*height*3");
    CUPBR_HIP_CHECK(hipStreamSynchronize(r->stream));
    CUPBR_HIP_CHECK(hipMemcpy(r->images[CUPBR_IMAGE_OUTPUT].ptr, host_rgb, need * sizeof(float), hipMemcpyHostToDevice));
    r->frame_index = frame_index;
    return CUPBR_OK;
}

// ------------------------------------------------------------------ post processing
int cupbr_threshold_curve(float threshold, float knee, float curve[4]) {
    CHECK_ARG(curve, "null curve");
    curve[0] = threshold;  // Interactor.cpp:436-440
    curve[1] = knee - threshold;
    curve[2] = 2.0f * knee;
    curve[3] = 0.25f / knee;
    return CUPBR_OK;
}

int cupbr_post_create(cupbr_post** out) {
    CHECK_ARG(out, "null output handle");
    *out = nullptr;
    auto* p = new (std::nothrow) cupbr_post();
    if (!p) return set_error(CUPBR_ERR_OUT_OF_MEMORY, "out of host memory");
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
        delete p;
        return set_error(CUPBR_ERR_HIP, "could not create a HIP stream (no device?)");
    }
    *out = p;
    return CUPBR_OK;
}

int cupbr_post_destroy(cupbr_post* p) {
    if (!p) return CUPBR_OK;
    (void)hipStreamSynchronize(p->stream);
    hipStream_t st = p->stream;
    delete p;
    (void)hipStreamDestroy(st);
    return CUPBR_OK;
}

int cupbr_post_register_image(cupbr_post* p, const float* dev, uint32_t w, uint32_t h) {
    CHECK_ARG(p && dev && w && h, "invalid image registration");
    p->hdr = dev;
    p->width = w;
    p->height = h;
    int rc = p->output.ensure((size_t)w * h * 12);
    if (rc) return rc;
    p->registered = true;
    return CUPBR_OK;
}

int cupbr_post_bloom(cupbr_post* p, const float curve[4]) {
    CHECK_ARG(p && curve, "null argument");
    if (!p->registered) return set_error(CUPBR_ERR_NOT_REGISTERED, "[PostProcessor]  No HDR image has been registered.", true);
    return post_bloom(p, curve, nullptr, 0, 1.0f);
}

int cupbr_post_set_bloom_pyramid(cupbr_post* p, int32_t pyramid) {
    CHECK_ARG(p, "null post processor");
    CHECK_ARG(pyramid == CUPBR_BLOOM_REFERENCE || pyramid == CUPBR_BLOOM_GAUSSIAN5, "unknown bloom pyramid");
    CUPBR_HIP_CHECK(hipStreamSynchronize(p->stream));
    p->pyramid = pyramid;
    return CUPBR_OK;
}

int cupbr_post_get_buffer_device(const cupbr_post* p, float** dev) {
    CHECK_ARG(p && dev, "null argument");
    if (!p->registered) return set_error(CUPBR_ERR_NOT_REGISTERED, "[PostProcessor]  No HDR image has been registered.", true);
    *dev = p->output.as<float>();
    return CUPBR_OK;
}

int cupbr_post_get_buffer(const cupbr_post* p, float* host, size_t n) {
    float* dev;
    int rc = cupbr_post_get_buffer_device(p, &dev);
    if (rc) return rc;
    CHECK_ARG(host, "null buffer");
    const size_t need = (size_t)p->width * p->height * 3;
    if (n < need) return set_error(CUPBR_ERR_BUFFER_SIZE, "host buffer too small for the image");
    CUPBR_HIP_CHECK(hipStreamSynchronize(p->stream));
    CUPBR_HIP_CHECK(hipMemcpy(host, dev, need * 4, hipMemcpyDeviceToHost));
    return CUPBR_OK;
}

// ------------------------------------------------------------------ tone mapping
int cupbr_tonemapper_create(int32_t type, cupbr_tonemapper** out) {
    CHECK_ARG(out, "null output handle");
    *out = nullptr;
    CHECK_ARG(type == CUPBR_REINHARD || type == CUPBR_GAMMA || type == CUPBR_ACES, "unknown tone mapping type");
    auto* t = new (std::nothrow) cupbr_tonemapper();
    if (!t) return set_error(CUPBR_ERR_OUT_OF_MEMORY, "out of host memory");
    t->type = type;
    if (hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess) {
        delete t;
        return set_error(CUPBR_ERR_HIP, "could not create a HIP stream (no device?)");
    }
    *out = t;
    return CUPBR_OK;
}

int cupbr_tonemapper_destroy(cupbr_tonemapper* t) {
    if (!t) return CUPBR_OK;
    (void)hipStreamSynchronize(t->stream);
    hipStream_t st = t->stream;
    delete t;
    (void)hipStreamDestroy(st);
    return CUPBR_OK;
}

int cupbr_tonemapper_register_image(cupbr_tonemapper* t, const float* dev, uint32_t w, uint32_t h) {
    CHECK_ARG(t && dev && w && h, "invalid image registration");
    t->hdr = dev;
    t->width = w;
    t->height = h;
    int rc = t->rgba.ensure((size_t)w * h * 4);
    if (rc) return rc;
    t->registered = true;
    return CUPBR_OK;
}

int cupbr_tonemapper_tonemap(cupbr_tonemapper* t) {
    CHECK_ARG(t, "null tone mapper");
    if (!t->registered)  // ToneMapper.cu:154-157
        return set_error(CUPBR_ERR_NOT_REGISTERED, "[ToneMapper]  No HDR image has been registered. Call registerImage() first!", true);
    return tonemap(t);
}

int cupbr_tonemapper_tonemap_bloom(cupbr_tonemapper* t, cupbr_post* p, const float curve[4]) {
    CHECK_ARG(t && p && curve, "null argument");
    if (!p->registered) return set_error(CUPBR_ERR_NOT_REGISTERED, "[PostProcessor]  No HDR image has been registered.", true);
    if (!t->registered)
        return set_error(CUPBR_ERR_NOT_REGISTERED, "[ToneMapper]  No HDR image has been registered. Call registerImage() first!", true);
    if (t->width != p->width || t->height != p->height)
        return set_error(CUPBR_ERR_INVALID_ARGUMENT, "tone mapper and post processor images differ in size");
    CUPBR_HIP_CHECK(hipStreamSynchronize(t->stream));
    return post_bloom(p, curve, t->rgba.as<uint8_t>(), t->type, t->exposure);
}

int cupbr_tonemapper_get_render_buffer_device(const cupbr_tonemapper* t, uint8_t** dev) {
    CHECK_ARG(t && dev, "null argument");
    if (!t->registered)
        return set_error(CUPBR_ERR_NOT_REGISTERED, "[ToneMapper]  No HDR image has been registered. Call registerImage() first!", true);
    *dev = t->rgba.as<uint8_t>();
    return CUPBR_OK;
}

int cupbr_tonemapper_get_render_buffer(const cupbr_tonemapper* t, uint8_t* host, size_t n) {
    uint8_t* dev;
    int rc = cupbr_tonemapper_get_render_buffer_device(t, &dev);
    if (rc) return rc;
    CHECK_ARG(host, "null buffer");
    const size_t need = (size_t)t->width * t->height * 4;
    if (n < need) return set_error(CUPBR_ERR_BUFFER_SIZE, "host buffer too small for the render buffer");
    CUPBR_HIP_CHECK(hipDeviceSynchronize());
    CUPBR_HIP_CHECK(hipMemcpy(host, dev, need, hipMemcpyDeviceToHost));
    return CUPBR_OK;
}

int cupbr_tonemapper_set_type(cupbr_tonemapper* t, int32_t type) {
    CHECK_ARG(t, "null tone mapper");
    CHECK_ARG(type == CUPBR_REINHARD || type == CUPBR_GAMMA || type == CUPBR_ACES, "unknown tone mapping type");
    t->type = type;
    return CUPBR_OK;
}

int cupbr_tonemapper_get_type(const cupbr_tonemapper* t, int32_t* type) {
    CHECK_ARG(t && type, "null argument");
    *type = t->type;
    return CUPBR_OK;
}

int cupbr_tonemapper_set_exposure(cupbr_tonemapper* t, float e) {
    CHECK_ARG(t, "null tone mapper");
    t->exposure = e;
    return CUPBR_OK;
}

// ToneMapper::saveToFile, ToneMapper.cu:183-198: 32-bit BMP with the rows flipped (row 0 = bottom)
int cupbr_tonemapper_save_to_file(const cupbr_tonemapper* t, const char* path) {
    CHECK_ARG(t && path, "null argument");
    if (!t->registered)
        return set_error(CUPBR_ERR_NOT_REGISTERED, "[ToneMapper]  No HDR image has been registered. Call registerImage() first!", true);
    const uint32_t w = t->width, h = t->height;
    std::vector<uint8_t> px((size_t)w * h * 4);
    int rc = cupbr_tonemapper_get_render_buffer(t, px.data(), px.size());
    if (rc) return rc;
    return write_bmp(path, px.data(), w, h);
}

}  // extern "C"

// ------------------------------------------------------------------ heterogeneous medium + adjoint
static size_t grid_voxels(const cupbr_grid_medium* m) { return (size_t)m->nx * m->ny * m->nz; }
static size_t grid_surfaces(const cupbr_scene* s) {
    size_t n = s->objects.size() + s->triangles.size() / 9;
    return n;
}

int cupbr_grid_surface_count(const cupbr_scene* s, uint32_t* n) {
    CHECK_ARG(s && n, "null argument");
    *n = (uint32_t)grid_surfaces(s);
    return CUPBR_OK;
}

int cupbr_grid_set_counting(cupbr_scene* s, int32_t enable) {
    CHECK_ARG(s, "null scene");
    return grid_set_counting(s, enable != 0);
}

int cupbr_grid_get_counts(const cupbr_scene* s, uint64_t counts[4]) {
    CHECK_ARG(s && counts, "null argument");
    return grid_get_counts(s, counts);
}

int cupbr_grid_render_device(const cupbr_scene* s, const cupbr_camera* cam, const cupbr_grid_medium* m,
                             const float* density, const float* proposal, const float* albedo, uint32_t w, uint32_t h,
                             uint32_t frame0, uint32_t spp, uint32_t max_depth, float* image, void* stream) {
    if (int rc = grid_validate(s, cam, m, w, h, spp, max_depth)) return rc;
    CHECK_ARG(density && albedo && image, "null buffer");
    return grid_render_device(s, cam, m, density, proposal, albedo, w, h, frame0, spp, max_depth, image,
                              (hipStream_t)stream);
}

int cupbr_grid_adjoint_device(const cupbr_scene* s, const cupbr_camera* cam, const cupbr_grid_medium* m,
                              const float* density, const float* proposal, const float* albedo, uint32_t w,
                              uint32_t h, uint32_t frame0, uint32_t spp, uint32_t max_depth, const float* d_image,
                              float* grad_density, float* grad_albedo, void* stream) {
    if (int rc = grid_validate(s, cam, m, w, h, spp, max_depth)) return rc;
    CHECK_ARG(density && albedo && d_image && grad_density && grad_albedo, "null buffer");
    return grid_adjoint_device(s, cam, m, density, proposal, albedo, w, h, frame0, spp, max_depth, d_image,
                               grad_density, grad_albedo, (hipStream_t)stream);
}

int cupbr_grid_adjoint_from_forward_device(const cupbr_scene* s, const cupbr_camera* cam, const cupbr_grid_medium* m,
                                           const float* density, const float* proposal, const float* albedo,
                                           uint32_t w, uint32_t h, uint32_t frame0, uint32_t spp, uint32_t max_depth,
                                           const float* d_image, float* grad_density, float* grad_albedo,
                                           void* stream) {
    if (int rc = grid_validate(s, cam, m, w, h, spp, max_depth)) return rc;
    CHECK_ARG(density && albedo && d_image && grad_density && grad_albedo, "null buffer");
    return grid_adjoint_device(s, cam, m, density, proposal, albedo, w, h, frame0, spp, max_depth, d_image,
                               grad_density, grad_albedo, (hipStream_t)stream, true);
}

namespace {
struct GridHostBufs {
    DevBuf rho, prop, alb;
    int upload(const cupbr_scene* s, const cupbr_grid_medium* m, const float* density, const float* proposal,
               const float* albedo) {
        const size_t nv = grid_voxels(m) * 4, na = grid_surfaces(s) * 12;
        int rc;
        if ((rc = rho.ensure(nv)) || (rc = alb.ensure(std::max<size_t>(na, 4)))) return rc;
        CUPBR_HIP_CHECK(hipMemcpy(rho.ptr, density, nv, hipMemcpyHostToDevice));
        if (na) CUPBR_HIP_CHECK(hipMemcpy(alb.ptr, albedo, na, hipMemcpyHostToDevice));
        if (proposal) {
            if ((rc = prop.ensure(nv))) return rc;
            CUPBR_HIP_CHECK(hipMemcpy(prop.ptr, proposal, nv, hipMemcpyHostToDevice));
        }
        return CUPBR_OK;
    }
};
}  // namespace

int cupbr_grid_render(const cupbr_scene* s, const cupbr_camera* cam, const cupbr_grid_medium* m, const float* density,
                      const float* proposal, const float* albedo, uint32_t w, uint32_t h, uint32_t frame0,
                      uint32_t spp, uint32_t max_depth, float* image) {
    if (int rc = grid_validate(s, cam, m, w, h, spp, max_depth)) return rc;
    CHECK_ARG(density && (albedo || grid_surfaces(s) == 0) && image, "null buffer");
    GridHostBufs b;
    DevBuf img;
    int rc;
    if ((rc = b.upload(s, m, density, proposal, albedo)) || (rc = img.ensure((size_t)w * h * 12))) return rc;
    if ((rc = grid_render_device(s, cam, m, b.rho.as<float>(), proposal ? b.prop.as<float>() : nullptr,
                                 b.alb.as<float>(), w, h, frame0, spp, max_depth, img.as<float>(), nullptr)))
        return rc;
    CUPBR_HIP_CHECK(hipMemcpy(image, img.ptr, (size_t)w * h * 12, hipMemcpyDeviceToHost));
    return CUPBR_OK;
}

int cupbr_grid_adjoint(const cupbr_scene* s, const cupbr_camera* cam, const cupbr_grid_medium* m, const float* density,
                       const float* proposal, const float* albedo, uint32_t w, uint32_t h, uint32_t frame0,
                       uint32_t spp, uint32_t max_depth, const float* d_image, float* grad_density,
                       float* grad_albedo) {
    if (int rc = grid_validate(s, cam, m, w, h, spp, max_depth)) return rc;
    CHECK_ARG(density && (albedo || grid_surfaces(s) == 0) && d_image && grad_density &&
                  (grad_albedo || grid_surfaces(s) == 0),
              "null buffer");
    GridHostBufs b;
    DevBuf di, gr, ga;
    const size_t nv = grid_voxels(m) * 4, na = grid_surfaces(s) * 12, ni = (size_t)w * h * 12;
    int rc;
    if ((rc = b.upload(s, m, density, proposal, albedo)) || (rc = di.ensure(ni)) || (rc = gr.ensure(nv)) ||
        (rc = ga.ensure(std::max<size_t>(na, 4))))
        return rc;
    CUPBR_HIP_CHECK(hipMemcpy(di.ptr, d_image, ni, hipMemcpyHostToDevice));
    CUPBR_HIP_CHECK(hipMemset(gr.ptr, 0, nv));
    CUPBR_HIP_CHECK(hipMemset(ga.ptr, 0, std::max<size_t>(na, 4)));
    if ((rc = grid_adjoint_device(s, cam, m, b.rho.as<float>(), proposal ? b.prop.as<float>() : nullptr,
                                  b.alb.as<float>(), w, h, frame0, spp, max_depth, di.as<float>(), gr.as<float>(),
                                  ga.as<float>(), nullptr)))
        return rc;
    CUPBR_HIP_CHECK(hipMemcpy(grad_density, gr.ptr, nv, hipMemcpyDeviceToHost));
    if (na) CUPBR_HIP_CHECK(hipMemcpy(grad_albedo, ga.ptr, na, hipMemcpyDeviceToHost));
    return CUPBR_OK;
}

// ------------------------------------------------------------------ neural radiance cache
int cupbr_nrc_create(const float lo[3], const float hi[3], uint32_t seed, cupbr_nrc** out) {
    CHECK_ARG(lo && hi && out, "null argument");
    for (int a = 0; a < 3; ++a) CHECK_ARG(std::isfinite(lo[a]) && std::isfinite(hi[a]) && hi[a] > lo[a], "invalid box");
    return nrc_create(lo, hi, seed, out);
}

int cupbr_nrc_destroy(cupbr_nrc* c) {
    delete c;
    return CUPBR_OK;
}

int cupbr_nrc_get_weights(const cupbr_nrc* c, float* w) {
    CHECK_ARG(c && w, "null argument");
    CUPBR_HIP_CHECK(hipDeviceSynchronize());
    CUPBR_HIP_CHECK(hipMemcpy(w, c->w.ptr, kNrcParams * 4, hipMemcpyDeviceToHost));
    return CUPBR_OK;
}

int cupbr_nrc_set_weights(cupbr_nrc* c, const float* w) {
    CHECK_ARG(c && w, "null argument");
    CUPBR_HIP_CHECK(hipDeviceSynchronize());
    int rc = nrc_set_weights(c, w);
    if (rc) return rc;
    CUPBR_HIP_CHECK(hipDeviceSynchronize());
    return CUPBR_OK;
}

int cupbr_nrc_infer_device(cupbr_nrc* c, const float* inputs, uint32_t n, float* out, void* stream) {
    CHECK_ARG(c && (n == 0 || (inputs && out)), "null argument");
    return nrc_infer(c, inputs, n, nullptr, kNrcInputs, out, false, (hipStream_t)stream);
}

int cupbr_nrc_train_device(cupbr_nrc* c, const float* rows, uint32_t n_rows, uint32_t batch, uint32_t steps, float lr,
                           float* loss, void* stream) {
    CHECK_ARG(c && (n_rows == 0 || rows), "null argument");
    CHECK_ARG(batch >= 512 && batch % 512 == 0 && batch <= (1u << 22), "batch must be a multiple of 512, <= 2^22");
    CHECK_ARG(std::isfinite(lr) && lr >= 0.0f, "invalid learning rate");
    return nrc_train(c, rows, n_rows, batch, steps, lr, loss, (hipStream_t)stream);
}

int cupbr_nrc_matrix_core_product(const float* a, const float* b, float* d) {
    CHECK_ARG(a && b && d, "null argument");
    return nrc_mfma_product(a, b, d);
}

int cupbr_renderer_set_radiance_cache(cupbr_renderer* r, cupbr_nrc* cache, uint32_t training_paths,
                                      uint32_t train_steps, uint32_t batch, float lr) {
    CHECK_ARG(r, "null renderer");
    CHECK_ARG(batch >= 512 && batch % 512 == 0 && batch <= (1u << 22), "batch must be a multiple of 512, <= 2^22");
    CHECK_ARG(std::isfinite(lr) && lr >= 0.0f, "invalid learning rate");
    CUPBR_HIP_CHECK(hipStreamSynchronize(r->stream));
    r->nrc = cache;
    r->nrc_train_paths = training_paths;
    r->nrc_steps = train_steps;
    r->nrc_batch = batch;
    r->nrc_lr = lr;
    return CUPBR_OK;
}

int cupbr_renderer_set_radiance_cache_camera_hit(cupbr_renderer* r, float min_roughness) {
    CHECK_ARG(r, "null renderer");
    CHECK_ARG(!std::isnan(min_roughness), "min_roughness is NaN");
    CUPBR_HIP_CHECK(hipStreamSynchronize(r->stream));
    r->nrc_camera_min_roughness = min_roughness;
    return CUPBR_OK;
}
