// cupbr.hpp — the reference's C++ class interface, header-only, over the C ABI in cupbr.h.
//
// The reference's application code (src/source.cu:7-108) is written against SceneLoader,
// PBRenderer, PostProcessor, ToneMapper and Camera. The classes below keep those names,
// argument meanings and call order, so such code switches by changing its includes:
//
//   cupbr::Scene scene = cupbr::SceneLoader::loadFromFile(path);   // SceneLoader.h:15-21
//   cupbr::PBRenderer renderer(cupbr::RenderingMethod::PATHTRACER); // PBRenderer.h:34
//   renderer.setOutputSize(w, h);                                  // PBRenderer.h:46
//   renderer.registerScene(&scene);                                // PBRenderer.h:52
//   renderer.render(&camera);                                      // PBRenderer.h:64
//   cupbr::ToneMapper mapper(cupbr::ToneMappingType::REINHARD);    // ToneMapper.h:29-35
//   mapper.registerImage(renderer.getOutputImage());               // ToneMapper.h:41
//   mapper.toneMap(); mapper.saveToFile("out.bmp");                // ToneMapper.h:47,75
//   cupbr::SceneLoader::destroyScene(scene);                       // SceneLoader.h:21
//
// Differences a caller sees:
//   - Methods that return void in the reference return the C status code (0 = success);
//     messages the reference printed to std::cerr are still printed.
//   - Images are views of device memory owned by the renderer / post processor
//     (Image<Vector3float>* in the reference, equally borrowed); Image::toHost() copies.
//   - render(Camera*, spp) accumulates spp reference frames in one call; render(Camera*) is one.
#ifndef CUPBR_HPP
#define CUPBR_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "cupbr.h"

namespace cupbr {

enum class RenderingMethod : int32_t {  // PBRenderer.h:14-22
    RAYTRACER = CUPBR_RAYTRACER,
    WHITTED = CUPBR_WHITTED,
    PATHTRACER = CUPBR_PATHTRACER,
    METROPOLIS = CUPBR_METROPOLIS,
    GRADIENTDOMAIN = CUPBR_GRADIENTDOMAIN,
    VOLUME = CUPBR_VOLUME
};

enum class ToneMappingType : int32_t {  // ToneMapper.h:13-17 (+ ACES, an addition)
    REINHARD = CUPBR_REINHARD,
    GAMMA = CUPBR_GAMMA,
    ACES = CUPBR_ACES
};

struct Vector4float {
    float x, y, z, w;
};

// Image<Vector3float> as a borrowed view: FP32 RGB, row 0 = bottom.
struct Image {
    float* data = nullptr;  // device pointer
    uint32_t width = 0, height = 0;
    size_t size() const { return (size_t)width * height; }
};

// Camera (Camera.h:14-86): Camera(width, height) is the reference's default pose; lookAt is an
// addition for scripted poses. moved() restarts accumulation (PBRenderer.cu:174-177).
class Camera {
public:
    Camera(uint32_t width, uint32_t height) {
        check(cupbr_camera_init(&c_, width, height));
        check(cupbr_camera_controller_init(&ctl_, width, height));
    }
    static Camera lookAt(const float eye[3], const float target[3], const float up[3], float vfov_degrees,
                         uint32_t width, uint32_t height) {
        Camera cam(width, height);
        check(cupbr_camera_look_at(&cam.c_, eye, target, up, vfov_degrees, width, height));
        return cam;
    }
    bool moved() const { return c_.moved != 0; }
    void setMoved(bool m) { c_.moved = m ? 1 : 0; }
    // processInput(GLFWwindow*, const float&), Camera.cpp:13-90: the window's key and cursor state
    // is passed as a snapshot (keys: CUPBR_KEY_* bits held down)
    void processInput(const cupbr_input_state& input, float delta_time) {
        check(cupbr_camera_process_input(&c_, &ctl_, &input, delta_time));
    }
    const cupbr_camera* get() const { return &c_; }
    cupbr_camera* get() { return &c_; }

private:
    static void check(int rc) {
        if (rc != CUPBR_OK) throw std::runtime_error(cupbr_last_error());
    }
    cupbr_camera c_{};
    cupbr_camera_controller ctl_{};
};

// Scene (Scene.h:15-34) as a handle; owned until SceneLoader::destroyScene.
struct Scene {
    cupbr_scene* handle = nullptr;
};

namespace SceneLoader {
// Scene loadFromFile(const std::string&), SceneLoader.cpp:140-454. On failure the message is
// printed as in the reference and the returned handle is empty.
inline Scene loadFromFile(const std::string& path) {
    Scene s;
    if (cupbr_scene_load_xml(path.c_str(), &s.handle) != CUPBR_OK) s.handle = nullptr;
    return s;
}
// void destroyScene(Scene&), SceneLoader.cpp:456-522
inline void destroyScene(Scene& scene) {
    if (scene.handle) cupbr_scene_destroy(scene.handle);
    scene.handle = nullptr;
}
}  // namespace SceneLoader

// PBRenderer, PBRenderer.h:27-116
class PBRenderer {
public:
    explicit PBRenderer(RenderingMethod method = RenderingMethod::RAYTRACER) {
        if (cupbr_renderer_create((int32_t)method, &r_) != CUPBR_OK) throw std::runtime_error(cupbr_last_error());
    }
    ~PBRenderer() { cupbr_renderer_destroy(r_); }
    PBRenderer(const PBRenderer&) = delete;
    PBRenderer& operator=(const PBRenderer&) = delete;

    int setOutputSize(uint32_t width, uint32_t height) { return cupbr_renderer_set_output_size(r_, width, height); }
    int registerScene(Scene* scene) { return cupbr_renderer_register_scene(r_, scene ? scene->handle : nullptr); }
    int setMethod(RenderingMethod method) { return cupbr_renderer_set_method(r_, (int32_t)method); }
    int render(Camera* camera, uint32_t spp = 1) { return cupbr_render(r_, camera ? camera->get() : nullptr, spp); }
    Image getOutputImage() const { return image(CUPBR_IMAGE_OUTPUT); }
    Image getGradientX() const { return image(CUPBR_IMAGE_GRADIENT_X); }
    Image getGradientY() const { return image(CUPBR_IMAGE_GRADIENT_Y); }
    RenderingMethod getMethod() const {
        int32_t m = 0;
        cupbr_renderer_get_method(r_, &m);
        return (RenderingMethod)m;
    }
    int reset() { return cupbr_renderer_reset(r_); }
    int setRussianRoulette(bool use) { return cupbr_renderer_set_russian_roulette(r_, use ? 1 : 0); }
    int setMaxTraceDepth(uint32_t depth) { return cupbr_renderer_set_max_trace_depth(r_, depth); }
    uint32_t getMaxTraceDepth() const {
        uint32_t d = 0;
        cupbr_renderer_get_max_trace_depth(r_, &d);
        return d;
    }
    // additions: device synchronisation, host copies and statistics
    int synchronize() { return cupbr_renderer_synchronize(r_); }
    std::vector<float> toHost(int32_t which = CUPBR_IMAGE_OUTPUT) const {
        Image im = image(which);
        std::vector<float> out(im.size() * 3);
        if (!out.empty()) cupbr_renderer_get_image(r_, which, out.data(), out.size());
        return out;
    }
    cupbr_render_stats stats() const {
        cupbr_render_stats s{};
        cupbr_renderer_get_stats(r_, &s);
        return s;
    }
    cupbr_renderer* handle() const { return r_; }

private:
    Image image(int32_t which) const {
        Image im;
        if (cupbr_renderer_get_image_device(r_, which, &im.data) != CUPBR_OK) im.data = nullptr;
        cupbr_renderer_get_output_size(r_, &im.width, &im.height);
        return im;
    }

    cupbr_renderer* r_ = nullptr;
};

// PostProcessor, PostProcessor.h:18-47
class PostProcessor {
public:
    PostProcessor() {
        if (cupbr_post_create(&p_) != CUPBR_OK) throw std::runtime_error(cupbr_last_error());
    }
    ~PostProcessor() { cupbr_post_destroy(p_); }
    PostProcessor(const PostProcessor&) = delete;
    PostProcessor& operator=(const PostProcessor&) = delete;

    int registerImage(const Image& hdr) {
        w_ = hdr.width;
        h_ = hdr.height;
        return cupbr_post_register_image(p_, hdr.data, hdr.width, hdr.height);
    }
    int bloom(const Vector4float& threshold) {
        const float curve[4] = {threshold.x, threshold.y, threshold.z, threshold.w};
        return cupbr_post_bloom(p_, curve);
    }
    // an addition: CUPBR_BLOOM_GAUSSIAN5 instead of the reference's pyramid
    int setBloomPyramid(int32_t pyramid) { return cupbr_post_set_bloom_pyramid(p_, pyramid); }
    Image getPostProcessBuffer() const {
        Image im;
        cupbr_post_get_buffer_device(p_, &im.data);
        im.width = w_;
        im.height = h_;
        return im;
    }
    cupbr_post* handle() const { return p_; }

private:
    cupbr_post* p_ = nullptr;
    uint32_t w_ = 0, h_ = 0;
};

// Interactor::Impl::compute_threshold (Interactor.cpp:436-440)
inline Vector4float computeThreshold(float threshold, float knee) {
    float c[4] = {0, 0, 0, 0};
    cupbr_threshold_curve(threshold, knee, c);
    return {c[0], c[1], c[2], c[3]};
}

// ToneMapper, ToneMapper.h:29-75
class ToneMapper {
public:
    explicit ToneMapper(ToneMappingType type = ToneMappingType::REINHARD) {
        if (cupbr_tonemapper_create((int32_t)type, &t_) != CUPBR_OK) throw std::runtime_error(cupbr_last_error());
    }
    ~ToneMapper() { cupbr_tonemapper_destroy(t_); }
    ToneMapper(const ToneMapper&) = delete;
    ToneMapper& operator=(const ToneMapper&) = delete;

    int registerImage(const Image& hdr) {
        w_ = hdr.width;
        h_ = hdr.height;
        return cupbr_tonemapper_register_image(t_, hdr.data, hdr.width, hdr.height);
    }
    int toneMap() { return cupbr_tonemapper_tonemap(t_); }
    // addition: bloom + tone map in one pass over the post processor's registered image
    int toneMapBloom(PostProcessor& post, const Vector4float& threshold) {
        const float curve[4] = {threshold.x, threshold.y, threshold.z, threshold.w};
        return cupbr_tonemapper_tonemap_bloom(t_, post.handle(), curve);
    }
    std::vector<uint8_t> getRenderBuffer() const {
        std::vector<uint8_t> out((size_t)w_ * h_ * 4);
        if (!out.empty()) cupbr_tonemapper_get_render_buffer(t_, out.data(), out.size());
        return out;
    }
    ToneMappingType getType() const {
        int32_t t = 0;
        cupbr_tonemapper_get_type(t_, &t);
        return (ToneMappingType)t;
    }
    int setType(ToneMappingType type) { return cupbr_tonemapper_set_type(t_, (int32_t)type); }
    int setExposure(float exposure) { return cupbr_tonemapper_set_exposure(t_, exposure); }
    int saveToFile(const std::string& path) const { return cupbr_tonemapper_save_to_file(t_, path.c_str()); }

private:
    cupbr_tonemapper* t_ = nullptr;
    uint32_t w_ = 0, h_ = 0;
};

}  // namespace cupbr

#endif  // CUPBR_HPP
