/*
 * cupbr.h — C ABI of the MI355X renderer core.
 *
 * The reference exposes this path as C++ classes (SceneLoader, PBRenderer,
 * PostProcessor, ToneMapper) and plain types (Scene, Camera, Material,
 * Light, Volume, Image). This header replaces them with opaque handles and
 * POD structs; include/cupbr.hpp rebuilds the reference's class interface on
 * top of it. Every entry point names the reference interface it replaces
 * (paths relative to the reference repository root).
 *
 * Conventions
 *   - Every function returns CUPBR_OK (0) or a negative CUPBR_ERR_* code;
 *     cupbr_last_error() returns the message of the last failure on the
 *     calling thread. Where the reference printed a message to std::cerr and
 *     returned, the same message is printed and a code is returned.
 *   - Images are FP32 RGB, 12 bytes per pixel, row-major, row 0 = bottom
 *     (Image<Vector3float>, lib/include/DataStructure/Image.h:12-138).
 *   - Render buffers are RGBA8 (RenderBuffer, lib/include/DataStructure/RenderBuffer.h:8).
 *   - "_device" entry points take and return device pointers (HIP memory on
 *     the current device); the others copy through host memory.
 *   - A handle is used by one host thread at a time. Calls are asynchronous
 *     on the handle's HIP stream unless they return host data.
 */
#ifndef CUPBR_H
#define CUPBR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status codes */
#define CUPBR_OK 0
#define CUPBR_ERR_INVALID_ARGUMENT (-1)
#define CUPBR_ERR_NO_OUTPUT_SIZE (-2)   /* PBRenderer.cu:140-144 */
#define CUPBR_ERR_NO_SCENE (-3)         /* PBRenderer.cu:146-150 */
#define CUPBR_ERR_FILE (-4)             /* SceneLoader.cpp:148-152 */
#define CUPBR_ERR_PARSE (-5)            /* missing XML elements (SceneLoader.cpp:155-156,176) */
#define CUPBR_ERR_GEOMETRY_TYPE (-6)    /* SceneLoader.cpp:264-268 */
#define CUPBR_ERR_LIGHT_TYPE (-7)       /* SceneLoader.cpp:288-292 */
#define CUPBR_ERR_MATERIAL_TYPE (-8)    /* SceneLoader.cpp:66-69 */
#define CUPBR_ERR_NOT_REGISTERED (-9)   /* ToneMapper.cu:154-157 */
#define CUPBR_ERR_UNSUPPORTED (-10)     /* WHITTED / METROPOLIS, PBRenderer.cu:161-170,187-191 */
#define CUPBR_ERR_HIP (-11)
#define CUPBR_ERR_OUT_OF_MEMORY (-12)
#define CUPBR_ERR_BUFFER_SIZE (-13)

/* ------------------------------------------------------------------ enums */
/* RenderingMethod, lib/include/Renderer/PBRenderer.h:14-22 */
enum cupbr_rendering_method {
    CUPBR_RAYTRACER = 0,
    CUPBR_WHITTED = 1,
    CUPBR_PATHTRACER = 2,
    CUPBR_METROPOLIS = 3,
    CUPBR_GRADIENTDOMAIN = 4,
    CUPBR_VOLUME = 5
};
/* GeometryType, lib/include/Geometry/Geometry.h:15-23 */
enum cupbr_geometry_type { CUPBR_SPHERE = 0, CUPBR_PLANE = 1, CUPBR_QUAD = 2, CUPBR_TRIANGLE = 3, CUPBR_MESH = 4 };
/* MaterialType, lib/include/Geometry/Material.h:14-22 */
enum cupbr_material_type {
    CUPBR_LAMBERT = 0, CUPBR_PHONG = 1, CUPBR_MIRROR = 2, CUPBR_GLASS = 3, CUPBR_GGX = 4, CUPBR_VOLUME_MATERIAL = 5
};
/* Interface, lib/include/DataStructure/Volume.h:6-10 */
enum cupbr_interface { CUPBR_INTERFACE_NONE = 0, CUPBR_INTERFACE_GLASS = 1 };
/* LightType, lib/include/DataStructure/Light.h:11-15 */
enum cupbr_light_type { CUPBR_POINT = 0, CUPBR_AREA = 1 };
/* ToneMappingType, lib/include/Renderer/ToneMapper.h:13-17. CUPBR_ACES has no counterpart in the
 * reference: Narkowicz's fit of the ACES filmic curve, f(x) = x(2.51x + 0.03) / (x(2.43x + 0.59) +
 * 0.14) of exposure * radiance, clamped to [0, 1], then the 1/2.2 gamma of the REINHARD mode and the
 * same truncating 8-bit cast. */
enum cupbr_tone_mapping_type { CUPBR_REINHARD = 0, CUPBR_GAMMA = 1, CUPBR_ACES = 2 };
/* images a renderer owns (PBRenderer.cu:35-43) */
enum cupbr_image {
    CUPBR_IMAGE_OUTPUT = 0,     /* getOutputImage(), PBRenderer.cu:232-236 */
    CUPBR_IMAGE_GRADIENT_X = 1, /* getGradientX(), :238-242 */
    CUPBR_IMAGE_GRADIENT_Y = 2, /* getGradientY(), :244-248 */
    CUPBR_IMAGE_BASE = 3,       /* GDPT primal running mean */
    CUPBR_IMAGE_GRADIENT_X_FORWARD = 4,
    CUPBR_IMAGE_GRADIENT_X_BACKWARD = 5,
    CUPBR_IMAGE_GRADIENT_Y_FORWARD = 6,
    CUPBR_IMAGE_GRADIENT_Y_BACKWARD = 7
};

/* ------------------------------------------------------------------ POD types */
/* Volume, lib/include/DataStructure/Volume.h:15-21 (defaults sigma_s = sigma_a = 1, g = 0) */
typedef struct cupbr_volume {
    float sigma_s[3];
    float sigma_a[3];
    float g;
    int32_t interface_type;
} cupbr_volume;

/* Material, lib/include/Geometry/Material.h:35-43 (84 bytes, same fields and order) */
typedef struct cupbr_material {
    float albedo_e[3];
    float albedo_d[3];
    float albedo_s[3];
    float shininess;
    float roughness;
    float eta;
    cupbr_volume volume;
    int32_t type;
} cupbr_material;

/* Light, lib/include/DataStructure/Light.h:28-34 */
typedef struct cupbr_light {
    int32_t type;
    float position[3];
    float intensity[3];
    float radiance[3];
    float half_extend1[3];
    float half_extend2[3];
} cupbr_light;

/* Camera, lib/include/DataStructure/Camera.h:72-85: pinhole at position, image plane at
 * position + z_axis spanned by x_axis (right, scaled by the aspect ratio) and y_axis (up). */
typedef struct cupbr_camera {
    float position[3];
    float x_axis[3];
    float y_axis[3];
    float z_axis[3];
    int32_t moved; /* Camera::moved(): non-zero restarts accumulation (PBRenderer.cu:174-177) */
} cupbr_camera;

/* One scene object (Geometry and its subclasses, lib/include/Geometry/).
 * params by type:
 *   SPHERE   center[3], radius                       (Sphere.h)
 *   PLANE    position[3], normal[3]                  (Plane.h; normal is normalised on load)
 *   QUAD     position[3], normal[3], extend1[3], extend2[3]   (Quad.h)
 *   TRIANGLE vertex1[3], vertex2[3], vertex3[3]      (Triangle.h)
 *   MESH     box min[3], box max[3]; triangles [first_triangle, +triangle_count) of the
 *            triangle array (Mesh.h; the box is the min/max over the mesh's vertices as
 *            ObjLoader.cpp:24-47 computes it). */
typedef struct cupbr_object {
    int32_t type;
    float params[12];
    uint32_t first_triangle;
    uint32_t triangle_count;
    cupbr_material material;
} cupbr_object;

typedef struct cupbr_bvh_stats {
    uint32_t nodes;          /* 128-byte BVH4 nodes */
    uint32_t leaves;
    uint32_t primitives;     /* bounded primitives in the BVH */
    uint32_t unbounded;      /* planes and quads tested for every ray */
    uint32_t depth;          /* BVH4 levels */
    uint32_t stack_entries;  /* traversal stack entries needed in the worst case */
    double build_ms;
    double upload_ms;
    double sah_cost;         /* sum over BVH4 nodes of 1.5 x area + over leaves of area x primitives,
                                areas relative to the root box */
} cupbr_bvh_stats;

typedef struct cupbr_render_stats {
    uint64_t samples;          /* camera paths traced by the last cupbr_render call */
    uint64_t extension_rays;   /* closest-hit rays (device queue counters) */
    uint64_t shadow_rays;      /* NEE visibility rays */
    double milliseconds;       /* device time of the last cupbr_render call (HIP events) */
    /* with cupbr_renderer_set_profiling(r, 1): per-stage device time (HIP events around every
     * launch on the renderer's stream) and launch counts; stage 0 trace (extension and shadow
     * rays), 1 shade, 2 unused, 3 ray generation, accumulation and radiance-cache work */
    double stage_ms[4];
    uint32_t stage_launches[4];
    /* with cupbr_renderer_set_traversal_counting(r, 1): BVH4 nodes visited and primitives
     * tested by the extension (closest-hit) and shadow (any-hit) traversals */
    uint64_t extension_node_visits, extension_prim_tests;
    uint64_t shadow_node_visits, shadow_prim_tests;
    /* with a radiance cache attached: cache queries answered and training rows collected */
    uint64_t cache_queries, cache_training_rows;
    /* path tracer trace launches (after bounce 0) whose rays were handed to a wave's idle lanes
     * as they finished; shorter launches trace 64-ray chunks, one ray per lane */
    uint64_t refill_launches;
} cupbr_render_stats;

typedef struct cupbr_scene cupbr_scene;
typedef struct cupbr_renderer cupbr_renderer;
typedef struct cupbr_post cupbr_post;
typedef struct cupbr_tonemapper cupbr_tonemapper;

/* ------------------------------------------------------------------ errors / device */
const char* cupbr_last_error(void);
int cupbr_device_name(char* out, size_t n);
/* Measurement aid, no counterpart in the reference: runs a kernel of independent FP32 FMAs with
 * every lane active at 8 waves per SIMD (twice; the second launch is timed with HIP events) and
 * returns its time and its wave64 FMA instruction count, i.e. the VALU issue ceiling against which
 * the trace kernels' roofline is stated. */
int cupbr_measure_valu_peak(uint32_t iterations, double* milliseconds, uint64_t* fma_instructions);
/* Measurement aid (no counterpart in the reference): times FP32 atomic adds without return, one per
 * lane and iteration, to pseudo-random words of 8 copies of words_per_copy floats (block b into copy
 * b % 8, the grid adjoint's per-XCD layout); `atomics` = lane atomics of the timed launch. The
 * grid adjoint's roofline is stated against this rate. */
int cupbr_measure_atomic_peak(uint32_t iterations, uint32_t words_per_copy, double* milliseconds, uint64_t* atomics);

/* ------------------------------------------------------------------ camera */
/* Camera(width, height), lib/src/DataStructure/Camera.cpp:6-11: origin, looking +z, 90 deg vertical FOV */
int cupbr_camera_init(cupbr_camera* cam, uint32_t width, uint32_t height);
/* extension: pinhole with pose and vertical field of view (degrees), same axis convention */
int cupbr_camera_look_at(cupbr_camera* cam, const float eye[3], const float target[3], const float up[3],
                         float vfov_degrees, uint32_t width, uint32_t height);

/* Camera::processInput(GLFWwindow*, const float&), lib/src/DataStructure/Camera.cpp:13-90, with
 * the window's key and cursor polling replaced by an input snapshot (the caller reads them from
 * its own window system). The controller holds the state the reference keeps in Camera's private
 * members (Camera.h:78-85): aspect ratio, pitch, yaw, first-mouse flag, last cursor position.
 * WASD move along the x and z axes at 4 units/s (x4 with left shift, /4 with left control); the
 * cursor offset times 0.002 turns yaw and pitch (pitch clamped to +-3.14159/2); the axes are then
 * rebuilt from yaw and pitch, and moved is set when a key moved the camera or the cursor moved. */
typedef struct cupbr_camera_controller {
    float aspect_ratio;
    float pitch;
    float yaw;
    int32_t first_mouse;
    float last_x, last_y;
} cupbr_camera_controller;
#define CUPBR_KEY_W 1u
#define CUPBR_KEY_A 2u
#define CUPBR_KEY_S 4u
#define CUPBR_KEY_D 8u
#define CUPBR_KEY_LEFT_SHIFT 16u
#define CUPBR_KEY_LEFT_CONTROL 32u
typedef struct cupbr_input_state {
    uint32_t keys;            /* CUPBR_KEY_* bits of the keys held down (glfwGetKey == GLFW_PRESS) */
    double cursor_x, cursor_y; /* glfwGetCursorPos */
} cupbr_input_state;
/* the state of Camera(width, height) (Camera.cpp:6-11, Camera.h:78-85) */
int cupbr_camera_controller_init(cupbr_camera_controller* ctl, uint32_t width, uint32_t height);
int cupbr_camera_process_input(cupbr_camera* cam, cupbr_camera_controller* ctl, const cupbr_input_state* input,
                               float delta_time);

/* ------------------------------------------------------------------ scene */
/* Scene SceneLoader::loadFromFile(const std::string&), lib/src/Scene/SceneLoader.cpp:140-454.
 * Relative mesh paths resolve against the working directory, then against the XML file's directory. */
int cupbr_scene_load_xml(const char* path, cupbr_scene** out);
/* the Scene POD (lib/include/Scene/Scene.h:15-34) built from arrays; triangles: n_triangles x 9 floats
 * (v1, v2, v3); env_rgb: env_w x env_h x 3 floats (NULL for none); volume NULL for the defaults */
int cupbr_scene_create(const cupbr_object* objects, uint32_t n_objects, const float* triangles, uint32_t n_triangles,
                       const cupbr_light* lights, uint32_t n_lights, const float* env_rgb, uint32_t env_w,
                       uint32_t env_h, const cupbr_volume* volume, cupbr_scene** out);
/* Scene construction above runs on the host (parsing, reference ranks, BVH build); the device copy
 * is made by the first GPU call that needs it, or here explicitly (the reference uploads inside
 * loadFromFile, SceneLoader.cpp:370-420) */
int cupbr_scene_upload(cupbr_scene* scene);
/* void SceneLoader::destroyScene(Scene&), SceneLoader.cpp:456-522 */
int cupbr_scene_destroy(cupbr_scene* scene);
int cupbr_scene_get_counts(const cupbr_scene* scene, uint32_t* n_objects, uint32_t* n_triangles, uint32_t* n_lights,
                           int32_t* use_environment_map);
int cupbr_scene_get_objects(const cupbr_scene* scene, cupbr_object* out, uint32_t n);
int cupbr_scene_get_triangles(const cupbr_scene* scene, float* out, uint32_t n_triangles);
int cupbr_scene_get_lights(const cupbr_scene* scene, cupbr_light* out, uint32_t n);
int cupbr_scene_get_volume(const cupbr_scene* scene, cupbr_volume* out);
/* Scene::environment (Scene.h:26): width/height (0 without a map) and, when out is not NULL,
 * the FP32 RGB texels, row 0 first as stored in the file */
int cupbr_scene_get_environment(const cupbr_scene* scene, float* out, size_t n_floats, uint32_t* width,
                                uint32_t* height);
int cupbr_scene_get_bvh_stats(const cupbr_scene* scene, cupbr_bvh_stats* out);
/* No counterpart in the reference (inspection of this build's BVH, for tests and tools): copies
 * the BVH4 nodes in both device forms, 128 bytes per node: the FP32 node used by wave-uniform
 * traversal (child boxes as lo x, y, z then hi x, y, z, each a float[4] over the children, then
 * the 4 child words) and the form used by per-ray traversal (reference point c as 3 floats, a word of
 * three signed-byte exponents k, per axis two blocks of 8 binary16 plane offsets, lo0..3 hi0..3
 * then hi0..3 lo0..3, a plane being c + offset * 2^k, then the 4 child words). Either pointer may
 * be NULL; n must be at least cupbr_bvh_stats.nodes. */
int cupbr_scene_get_bvh_nodes(const cupbr_scene* scene, void* fp32_nodes, void* half_nodes, uint32_t n);
/* reference BVH leaf order of the objects (BoundingVolumeHierarchy.cpp:46-136); n >= n_objects */
int cupbr_scene_get_reference_order(const cupbr_scene* scene, int32_t* order, uint32_t n);

/* ------------------------------------------------------------------ closest hit / visibility */
/* Tracing::traceRay(Scene&, const Ray&), lib/src/Core/TracingDetail.h:47-66: rays are n x {o.xyz, d.xyz}
 * with d normalised (Ray's constructor, RayDetail.h:6-13). Outputs scene_index (-1 on a miss),
 * the triangle index inside a mesh (-1 otherwise) and the hit depth (+inf on a miss). */
int cupbr_trace_closest(const cupbr_scene* scene, const float* rays, size_t n, int32_t* object_id,
                        int32_t* triangle_id, float* t);
int cupbr_trace_closest_device(const cupbr_scene* scene, const float* rays, size_t n, int32_t* object_id,
                               int32_t* triangle_id, float* t, void* hip_stream);
/* No counterpart in the reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc). Same results as
 * cupbr_trace_closest_device; the traversal-stack spill lives in a fixed per-device buffer of the
 * library (16 MiB, held until the process ends) instead of a scene-owned allocation, and the grid
 * is cut to the blocks whose stacks fit it. Nothing is allocated: the scene must already be
 * resident on the current device (cupbr_scene_upload, or any GPU call on it) — otherwise
 * CUPBR_ERR_INVALID_ARGUMENT — and a BVH deeper than the buffer allows (more than 16,416 stack
 * entries per ray) gives CUPBR_ERR_BUFFER_SIZE. Calls that overlap on one device (other streams
 * or threads) share the buffer and race; order them. */
int cupbr_trace_closest_device_noalloc(const cupbr_scene* scene, const float* rays, size_t n, int32_t* object_id,
                                       int32_t* triangle_id, float* t, void* hip_stream);
/* Tracing::traceVisibility(Scene&, const float&, const Ray&), TracingDetail.h:118-183 */
int cupbr_trace_visibility(const cupbr_scene* scene, const float* rays, const float* distance, size_t n,
                           uint8_t* visible);
int cupbr_trace_visibility_device(const cupbr_scene* scene, const float* rays, const float* distance, size_t n,
                                  uint8_t* visible, void* hip_stream);

/* ------------------------------------------------------------------ picking (Interaction/MousePicker.h:13-33) */
/* Interaction::pickMouse: closest hit of the unjittered ray through pixel (x, y); on a hit writes
 * the object's material and scene index, on a miss leaves both unchanged (MousePicker.cu:23-34) */
int cupbr_pick(const cupbr_scene* scene, const cupbr_camera* camera, uint32_t x, uint32_t y, uint32_t width,
               uint32_t height, cupbr_material* out_material, int32_t* out_scene_index);
/* Interaction::updateMaterial: replaces the material of object scene_index (MousePicker.cu:37-51) */
int cupbr_scene_update_material(cupbr_scene* scene, int32_t scene_index, const cupbr_material* material);

/* ------------------------------------------------------------------ renderer (PBRenderer.h:27-116) */
int cupbr_renderer_create(int32_t method, cupbr_renderer** out);       /* PBRenderer(RenderingMethod) */
int cupbr_renderer_destroy(cupbr_renderer* r);                          /* ~PBRenderer */
int cupbr_renderer_set_output_size(cupbr_renderer* r, uint32_t width, uint32_t height); /* setOutputSize */
int cupbr_renderer_register_scene(cupbr_renderer* r, const cupbr_scene* scene);         /* registerScene */
int cupbr_renderer_set_method(cupbr_renderer* r, int32_t method);       /* setMethod (clears buffers) */
int cupbr_renderer_get_method(const cupbr_renderer* r, int32_t* method);/* getMethod */
int cupbr_renderer_reset(cupbr_renderer* r);                            /* reset */
int cupbr_renderer_set_russian_roulette(cupbr_renderer* r, int32_t enable); /* setRussianRoulette (default off) */
int cupbr_renderer_set_max_trace_depth(cupbr_renderer* r, uint32_t depth);  /* setMaxTraceDepth */
int cupbr_renderer_get_max_trace_depth(const cupbr_renderer* r, uint32_t* depth); /* getMaxTraceDepth */
int cupbr_renderer_get_frame_index(const cupbr_renderer* r, uint32_t* frame_index);
/* the size set by setOutputSize (0 x 0 before), for views of the renderer's images */
int cupbr_renderer_get_output_size(const cupbr_renderer* r, uint32_t* width, uint32_t* height);
/* render(Camera*), PBRenderer.cu:137-230, repeated spp times: each sample is one reference frame with
 * its own seed tea<4>(pixel, frameIndex) and the running mean of PathTracer.cu:162-168. */
int cupbr_render(cupbr_renderer* r, const cupbr_camera* camera, uint32_t spp);
int cupbr_renderer_get_image(const cupbr_renderer* r, int32_t which, float* host_rgb, size_t n_floats);
int cupbr_renderer_get_image_device(const cupbr_renderer* r, int32_t which, float** device_rgb);
int cupbr_renderer_get_stats(const cupbr_renderer* r, cupbr_render_stats* out);
/* extension: FP32 dumps and resume. save_pfm writes image `which` as a colour PFM (rows bottom-up,
 * little endian); resume loads an output image and frame index so that render continues the
 * running mean of a saved run (PBRenderer.cu:172-186 accumulates by frame index) */
int cupbr_renderer_save_pfm(const cupbr_renderer* r, int32_t which, const char* path);
int cupbr_renderer_resume(cupbr_renderer* r, const float* host_rgb, size_t n_floats, uint32_t frame_index);
/* extension (no reference counterpart): instrumentation for throughput and roofline reports */
int cupbr_renderer_set_profiling(cupbr_renderer* r, int32_t enable);
int cupbr_renderer_set_traversal_counting(cupbr_renderer* r, int32_t enable);
int cupbr_renderer_synchronize(cupbr_renderer* r);

/* Reconstruction of GRADIENTDOMAIN frames. CUPBR_GDPT_TRIMMED_MEAN is the reference's: 50 iterations
 * of the trimmed mean of the four neighbour estimates (GradientDomain.cu:303-350,392-401; alpha and
 * iterations are ignored). CUPBR_GDPT_SCREENED_POISSON has no counterpart in the reference: `iterations`
 * Jacobi iterations of the L2 screened-Poisson problem, minimising
 * alpha^2 |I - base|^2 + |dx I - gradient_x|^2 + |dy I - gradient_y|^2 over the interior pixels (the
 * gradients of init_kernel, GradientDomain.cu:248-273, as forward differences; border pixels zero).
 * Defaults: trimmed mean, alpha 0.2, 50 iterations. */
#define CUPBR_GDPT_TRIMMED_MEAN 0
#define CUPBR_GDPT_SCREENED_POISSON 1
int cupbr_renderer_set_gdpt_reconstruction(cupbr_renderer* r, int32_t mode, float alpha, uint32_t iterations);

/* Random numbers of the PATHTRACER method. CUPBR_RNG_TEA_LCG is the reference's: a path's state is
 * tea<4>(pixel, frame) and each draw one step of the LCG 1664525 s + 1013904223, its low 24 bits
 * divided by 2^24 (FunctionsDetail.h:70-98, PathTracer.cu:112). CUPBR_RNG_PCG has no counterpart in
 * the reference: the state starts at pcg_hash(pcg_hash(frame) ^ pixel) and each draw advances it as
 * s = 747796405 s + 2891336453 and returns the top 24 bits of the RXS-M-XS permutation of s
 * (O'Neill 2014), divided by 2^24, where pcg_hash(x) is that permutation of 747796405 x +
 * 2891336453. The other methods always use the reference's generator. */
#define CUPBR_RNG_TEA_LCG 0
#define CUPBR_RNG_PCG 1
int cupbr_renderer_set_rng(cupbr_renderer* r, int32_t rng);

/* ------------------------------------------------------------------ post processing (PostProcessor.h:12-52) */
int cupbr_post_create(cupbr_post** out);
int cupbr_post_destroy(cupbr_post* p);
/* registerImage(Image<Vector3float>*): borrows a device FP32 RGB image (e.g. a renderer's output) */
int cupbr_post_register_image(cupbr_post* p, const float* device_rgb, uint32_t width, uint32_t height);
/* bloom(Vector4float threshold), PostProcessor.cpp:125-142; curve = Interactor.cpp:436-440 */
int cupbr_post_bloom(cupbr_post* p, const float threshold_curve[4]);
/* Bloom pyramid of cupbr_post_bloom and cupbr_tonemapper_tonemap_bloom. CUPBR_BLOOM_REFERENCE (the
 * default) is the reference's (PostProcessor.cpp:125-142, PyramidConstructor.cu:9-169).
 * CUPBR_BLOOM_GAUSSIAN5 has no counterpart in the reference: L0 = the thresholded image, five
 * levels L_k(x, y) = B_{k-1}(2x, 2y) with B = the separable binomial blur [1 4 6 4 1] / 16
 * (edge-clamped) of L_{k-1}, level sizes (n + 1) / 2 per axis; then U_5 = L_5, U_k = L_k +
 * up(U_{k+1}) where up() samples U_{k+1} at (x / 2, y / 2) bilinearly (the average of the two or
 * four neighbours at odd coordinates, edge-clamped); output = hdr + U_0 / 6 (a constant image c
 * gives c + threshold(c)). */
enum cupbr_bloom_pyramid { CUPBR_BLOOM_REFERENCE = 0, CUPBR_BLOOM_GAUSSIAN5 = 1 };
int cupbr_post_set_bloom_pyramid(cupbr_post* p, int32_t pyramid);
int cupbr_post_get_buffer_device(const cupbr_post* p, float** device_rgb);   /* getPostProcessBuffer */
int cupbr_post_get_buffer(const cupbr_post* p, float* host_rgb, size_t n_floats);
/* threshold curve from (threshold, knee), Interactor::Impl::compute_threshold */
int cupbr_threshold_curve(float threshold, float knee, float curve[4]);

/* ------------------------------------------------------------------ tone mapping (ToneMapper.h:22-80) */
int cupbr_tonemapper_create(int32_t type, cupbr_tonemapper** out);
int cupbr_tonemapper_destroy(cupbr_tonemapper* t);
int cupbr_tonemapper_register_image(cupbr_tonemapper* t, const float* device_rgb, uint32_t width, uint32_t height);
int cupbr_tonemapper_tonemap(cupbr_tonemapper* t);                       /* toneMap */
/* extension: bloom of p's registered image and tone map of the bloomed result in one pass
 * (equals cupbr_post_bloom then registering the post buffer and calling toneMap) */
int cupbr_tonemapper_tonemap_bloom(cupbr_tonemapper* t, cupbr_post* p, const float threshold_curve[4]);
int cupbr_tonemapper_get_render_buffer(const cupbr_tonemapper* t, uint8_t* host_rgba, size_t n_bytes);
int cupbr_tonemapper_get_render_buffer_device(const cupbr_tonemapper* t, uint8_t** device_rgba);
int cupbr_tonemapper_set_type(cupbr_tonemapper* t, int32_t type);
int cupbr_tonemapper_get_type(const cupbr_tonemapper* t, int32_t* type);
int cupbr_tonemapper_set_exposure(cupbr_tonemapper* t, float exposure);
/* saveToFile(path): 32-bit BMP, rows flipped so row 0 of the image is the bottom (ToneMapper.cu:183-198) */
int cupbr_tonemapper_save_to_file(const cupbr_tonemapper* t, const char* path);

/* ------------------------------------------------------------------ heterogeneous medium + adjoint
 * No counterpart in the reference (its volumes are homogeneous: Volume.h:15-21,
 * VolumeRenderer.cu:97-124). A path tracer for scenes whose primitives are two-sided Lambertian
 * surfaces with one RGB albedo parameter each, inside and around a box medium with a voxel density
 * grid, plus its gradient by path replay: the adjoint re-runs every path from its seed instead of
 * storing it. Camera, seeding (tea<4>(pixel, frame0 + sample) + LCG), light sampling and HG phase
 * follow the reference's definitions.
 *
 * Surface index: object id for spheres/planes/quads/triangles; n_objects + global triangle index
 * for a mesh triangle (cupbr_grid_surface_count gives n_objects + n_triangles).
 * Images are the mean over spp samples, FP32 RGB, row 0 = bottom. */
typedef struct cupbr_grid_medium {
    float bmin[3], bmax[3];     /* medium box */
    uint32_t nx, ny, nz;        /* voxels; density(x) is trilinear between voxel centres, x fastest */
    float sigma_t_scale;        /* sigma_t(x) = sigma_t_scale * density(x) */
    float majorant;             /* tracking rate; sigma_t above it gives negative null weights */
    float scattering_albedo;    /* sigma_s / sigma_t */
    float g;                    /* Henyey-Greenstein asymmetry, (-1, 1) */
    float min_collision_probability; /* real-collision probability is clamped to [p, 1 - p], p in (0, 0.5] */
} cupbr_grid_medium;

int cupbr_grid_surface_count(const cupbr_scene* scene, uint32_t* n_surfaces);
/* Forward render. density / proposal: nx*ny*nz floats; proposal NULL = density (the fixed grid that
 * drives real/null decisions); surface_albedo: 3 floats per surface; image: width*height*3 floats. */
int cupbr_grid_render_device(const cupbr_scene* scene, const cupbr_camera* camera, const cupbr_grid_medium* medium,
                             const float* density, const float* proposal, const float* surface_albedo,
                             uint32_t width, uint32_t height, uint32_t frame0, uint32_t spp, uint32_t max_depth,
                             float* image, void* hip_stream);
/* Adjoint: ADDS dLoss/d(density) and dLoss/d(surface_albedo) for the image's cotangent d_image
 * (width*height*3 floats) to grad_density and grad_albedo (same sizes as density and albedo). */
int cupbr_grid_adjoint_device(const cupbr_scene* scene, const cupbr_camera* camera, const cupbr_grid_medium* medium,
                              const float* density, const float* proposal, const float* surface_albedo,
                              uint32_t width, uint32_t height, uint32_t frame0, uint32_t spp, uint32_t max_depth,
                              const float* d_image, float* grad_density, float* grad_albedo, void* hip_stream);
/* Adjoint that takes each path's radiance from the immediately preceding
 * cupbr_grid_render_device call on this scene instead of running every path an extra time (the
 * path-replay pass then runs once). The arguments must equal that call's (same buffers, camera,
 * medium, size, frame0, spp, max_depth), its width * height * spp must not exceed 2^24, the
 * parameter values must be unchanged since it, and the work must be ordered after it (same stream
 * or synchronised). Otherwise CUPBR_ERR_INVALID_ARGUMENT and nothing is launched. */
int cupbr_grid_adjoint_from_forward_device(const cupbr_scene* scene, const cupbr_camera* camera,
                                           const cupbr_grid_medium* medium, const float* density,
                                           const float* proposal, const float* surface_albedo, uint32_t width,
                                           uint32_t height, uint32_t frame0, uint32_t spp, uint32_t max_depth,
                                           const float* d_image, float* grad_density, float* grad_albedo,
                                           void* hip_stream);
/* instrumentation: with counting on, the grid kernels add their work to four counters of the
 * scene: closest-hit rays, shadow rays, grid stencil reads (8 voxels each) and gradient atomics.
 * Enabling resets them; cupbr_grid_get_counts synchronizes the device and reads them. */
int cupbr_grid_set_counting(cupbr_scene* scene, int32_t enable);
int cupbr_grid_get_counts(const cupbr_scene* scene, uint64_t counts[4]);
/* host-pointer variants (copy through device memory; the adjoint OVERWRITES the gradient outputs) */
int cupbr_grid_render(const cupbr_scene* scene, const cupbr_camera* camera, const cupbr_grid_medium* medium,
                      const float* density, const float* proposal, const float* surface_albedo, uint32_t width,
                      uint32_t height, uint32_t frame0, uint32_t spp, uint32_t max_depth, float* image);
int cupbr_grid_adjoint(const cupbr_scene* scene, const cupbr_camera* camera, const cupbr_grid_medium* medium,
                       const float* density, const float* proposal, const float* surface_albedo, uint32_t width,
                       uint32_t height, uint32_t frame0, uint32_t spp, uint32_t max_depth, const float* d_image,
                       float* grad_density, float* grad_albedo);

/* ------------------------------------------------------------------ neural radiance cache
 * No counterpart in the reference. A 64 -> 4 x 64 (ReLU) -> 3 MLP without biases (16,576 weights)
 * predicting outgoing radiance from 16 raw inputs per query: position[3] (world), outgoing
 * direction[3], normal[3], roughness, diffuse albedo[3], specular albedo[3]. BF16 matrix-core
 * products with FP32 accumulation; FP32 master weights trained with Adam on the relative L2 loss
 * sum_c (y_c - t_c)^2 / (y_c^2 + 0.01). Positions are normalised to the box given at creation.
 * Weights: W1[64][64] | W2 | W3 | W4 | W5[3][64], FP32, row = output. */
typedef struct cupbr_nrc cupbr_nrc;
#define CUPBR_NRC_PARAMS 16576
#define CUPBR_NRC_ROW_FLOATS 20 /* training row: 16 inputs, 3 targets, 1 unused */
int cupbr_nrc_create(const float box_min[3], const float box_max[3], uint32_t seed, cupbr_nrc** out);
int cupbr_nrc_destroy(cupbr_nrc* cache);
int cupbr_nrc_get_weights(const cupbr_nrc* cache, float* weights);   /* CUPBR_NRC_PARAMS floats */
int cupbr_nrc_set_weights(cupbr_nrc* cache, const float* weights);   /* also resets Adam */
/* predictions out[n][3] for inputs[n][16] (device pointers) */
int cupbr_nrc_infer_device(cupbr_nrc* cache, const float* inputs, uint32_t n, float* out, void* hip_stream);
/* `steps` Adam steps (learning rate lr, betas 0.9 / 0.99), each on `batch` rows (a multiple of 512)
 * drawn with replacement from rows[n_rows][CUPBR_NRC_ROW_FLOATS] (device); when loss is not NULL,
 * loss[s] += the summed relative L2 loss of step s's batch (device, caller zeroes) */
int cupbr_nrc_train_device(cupbr_nrc* cache, const float* rows, uint32_t n_rows, uint32_t batch, uint32_t steps,
                           float lr, float* loss, void* hip_stream);
/* Attaches a cache to a PATHTRACER renderer (NULL detaches: plain PT). Each frame then traces
 * `training_paths` pixels (chosen by a hash of pixel and frame) as full PT paths whose vertices
 * become training rows (target = the path's radiance after the vertex divided by its throughput
 * there), and ends every other path at its first diffuse or glossy vertex after the camera hit
 * with throughput * cache prediction. The frames of one cupbr_render call are traced together
 * (as one batch, up to the renderer's path bound) with the cache as it was at the start of the
 * call; after each batch the cache takes `train_steps` Adam steps of `batch` rows (a multiple of
 * 512) at learning rate lr. Images stay the running mean of frames. */
int cupbr_renderer_set_radiance_cache(cupbr_renderer* r, cupbr_nrc* cache, uint32_t training_paths,
                                      uint32_t train_steps, uint32_t batch, float learning_rate);
/* Makes the camera hit a cache vertex too, for rough surfaces: a path also ends there (and a
 * training path leaves a row there) when the hit material is Lambertian and min_roughness <= 1, or
 * GGX with roughness >= min_roughness. +INFINITY (the default) keeps the camera hit out of the
 * cache. The pixel is then the camera hit's emission plus the cache's prediction of the light it
 * reflects: no next-event estimation there, and the image carries the cache's approximation error
 * (its rel-MSE against a converged reference includes that bias). */
int cupbr_renderer_set_radiance_cache_camera_hit(cupbr_renderer* r, float min_roughness);
/* self-check of the matrix-core fragment layout: d[16][16] = a[16][32] b[32][16] (host, row-major,
 * inputs rounded to BF16) through one instruction */
int cupbr_nrc_matrix_core_product(const float* a, const float* b, float* d);

#ifdef __cplusplus
}
#endif

#endif /* CUPBR_H */
