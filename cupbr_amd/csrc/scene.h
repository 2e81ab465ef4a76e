// Host-side scene handle (cupbr_scene) and its device copy.
#pragma once

#include <string>
#include <vector>

#include "../../include/cupbr.h"
#include "device.h"
#include "scene_types.h"

struct cupbr_scene {
    // the scene as loaded (the reference's Scene POD, lib/include/Scene/Scene.h:15-34)
    std::vector<cupbr_object> objects;
    std::vector<float> triangles;  // 9 floats per mesh triangle
    std::vector<cupbr_light> lights;
    std::vector<float> env;        // env_w * env_h * 3, row 0 first as stored in the file
    uint32_t env_w = 0, env_h = 0;
    cupbr_volume volume{};

    // derived
    std::vector<int32_t> ref_order;    // reference BVH leaf order of the objects
    std::vector<uint32_t> rank_base;   // first reference rank of each object
    cupbr_bvh_stats stats{};
    uint32_t stack_entries = 0;        // worst-case traversal stack depth
    cupbr::DevScene dev{};
    bool uploaded = false;
    // host copies of the device arrays (built by scene_finalize, uploaded by scene_upload)
    std::vector<cupbr::Node4> h_nodes;
    std::vector<cupbr::NodeH> h_hnodes;
    std::vector<cupbr::PrimRec> h_prims;
    std::vector<float4> h_normals;
    std::vector<cupbr::UnbRec> h_unb;
    std::vector<cupbr::ObjInfo> h_info;
    std::vector<cupbr_material> h_mats;
    std::vector<uint32_t> h_rank_prim;
    std::vector<void*> allocations;
    int device = 0;
    cupbr::DevBuf spill;               // traversal stack spill for the batch query kernels
    hipEvent_t stream_event = nullptr; // end of the last kernel that used the scene's own buffers
    bool stream_event_recorded = false;
    cupbr::DevBuf grid_spill;          // traversal stack spill of the grid-medium kernels
    cupbr::DevBuf grid_samples;        // per-sample radiance of the grid-medium forward pass
    cupbr::DevBuf grid_counts;         // 4 x u64 work counters of the grid-medium kernels
    cupbr::DevBuf grid_grad;           // per-XCD density-gradient copies of the adjoint
    bool grid_counting = false;
    // arguments of the last grid forward render whose per-sample radiance is all in grid_samples
    struct GridForward {
        bool valid = false;
        const float *rho = nullptr, *prop = nullptr, *alb = nullptr;
        cupbr_camera cam{};
        cupbr_grid_medium medium{};
        uint32_t w = 0, h = 0, frame0 = 0, spp = 0, max_depth = 0;
    } grid_forward;
};

namespace cupbr {

// default Volume (Volume.h:15-21) and Material (Material.h:35-43)
cupbr_volume default_volume();
cupbr_material default_material();

// builds ranks and the BVH4 on the host (no GPU needed); returns a CUPBR_ status
int scene_finalize(cupbr_scene* s);
// copies the scene to the current device once; every GPU entry point calls it first
int scene_upload(const cupbr_scene* s);
void scene_release_device(cupbr_scene* s);

// The batch-query spill stack and the grid kernels' spill, sample and gradient buffers belong to
// the scene, not to a call: a call on one stream first waits for the scene's last kernel on any
// stream (acquire), and marks the end of its own kernels (release).
int scene_stream_acquire(cupbr_scene* s, hipStream_t stream);
int scene_stream_release(cupbr_scene* s, hipStream_t stream);

// XML / OBJ / RGBE loading (loader.cpp)
int load_scene_xml(const std::string& path, cupbr_scene* s);
int load_rgbe(const std::string& path, std::vector<float>& rgb, uint32_t& w, uint32_t& h);

// batch queries (trace.hip)
int trace_closest_device(const cupbr_scene* s, const float* rays, size_t n, int32_t* obj, int32_t* tri, float* t,
                         hipStream_t stream);
int trace_visibility_device(const cupbr_scene* s, const float* rays, const float* dist, size_t n, uint8_t* vis,
                            hipStream_t stream);
// the closest-hit kernel of trace_closest_device on a given grid of 256-lane blocks, with a spill
// stack of (stack_entries - 32) words per lane at a stride of blocks * 256 (nullptr when
// stack_entries <= 32)
constexpr int kTraceBlock = 256;
constexpr int kTraceLds = 32;
int trace_closest_launch(const cupbr_scene* s, const float* rays, size_t n, int32_t* obj, int32_t* tri, float* t,
                         uint32_t* spill, int blocks, hipStream_t stream);

// heterogeneous medium + adjoint (grid.hip)
int grid_set_counting(cupbr_scene* s, bool enable);
int grid_get_counts(const cupbr_scene* s, uint64_t out[4]);
int grid_validate(const cupbr_scene* s, const cupbr_camera* cam, const cupbr_grid_medium* m, uint32_t w, uint32_t h,
                  uint32_t spp, uint32_t max_depth);
int grid_render_device(const cupbr_scene* s, const cupbr_camera* cam, const cupbr_grid_medium* m, const float* rho,
                       const float* prop, const float* alb, uint32_t w, uint32_t h, uint32_t frame0, uint32_t spp,
                       uint32_t max_depth, float* image, hipStream_t stream);
// from_forward: take each path's radiance from the preceding grid_render_device call with the same
// arguments instead of replaying the path for it (an error when there is no such call)
int grid_adjoint_device(const cupbr_scene* s, const cupbr_camera* cam, const cupbr_grid_medium* m, const float* rho,
                        const float* prop, const float* alb, uint32_t w, uint32_t h, uint32_t frame0, uint32_t spp,
                        uint32_t max_depth, const float* d_image, float* grad_rho, float* grad_alb,
                        hipStream_t stream, bool from_forward = false);

}  // namespace cupbr
