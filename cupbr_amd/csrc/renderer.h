// Host-side renderer handle (cupbr_renderer): PBRenderer's state
// (lib/src/Renderer/PBRenderer.cu:26-52) plus the wavefront buffers.
#pragma once

#include "device.h"
#include "scene.h"

namespace cupbr {

constexpr int kNumImages = 8;  // cupbr_image

// Device counters of one render call. Work counters hand out 64-entry chunks to
// persistent waves; each is reset by the kernel that runs after its last user.
struct Counters {
    uint32_t queue_count[2];   // extension queues (ping-pong by bounce parity)
    uint32_t shadow_count;
    uint32_t work_extend, work_shade, work_shadow;
    uint32_t pad[2];
    unsigned long long extension_rays;
    unsigned long long shadow_rays;
};

struct Wavefront {
    uint32_t capacity = 0;  // paths
    DevBuf weight;          // float4 per path: throughput.xyz, rng state (bits)
    DevBuf radiance;        // float4 per path
    DevBuf q_o[2], q_d[2];  // float4 per entry: origin | direction.xyz, path (bits)
    DevBuf q_hit;           // float2 per entry: t, prim (bits)
    DevBuf s_o, s_d, s_c;   // shadow queue: origin.xyz+distance | direction.xyz+path | contribution
    DevBuf counters;
    DevBuf spill;
    uint32_t spill_entries = 0;
};

}  // namespace cupbr

struct cupbr_renderer {
    int32_t method = CUPBR_RAYTRACER;
    uint32_t width = 0, height = 0;
    bool output_size_set = false;
    const cupbr_scene* scene = nullptr;
    uint32_t frame_index = 0;
    uint32_t max_trace_depth = 5;   // PBRenderer.cu:59
    bool russian_roulette = false;  // uninitialised in the reference (PBRenderer.cu:51); off here
    cupbr::DevBuf images[cupbr::kNumImages];
    cupbr::DevBuf temp;
    cupbr::Wavefront wf;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    cupbr_render_stats stats{};
};

namespace cupbr {

int renderer_clear_images(cupbr_renderer* r);
// one reference frame per sample: frames [frame0, frame0+spp) of method PATHTRACER
int render_pathtracer(cupbr_renderer* r, const cupbr_camera& cam, uint32_t frame0, uint32_t spp);
int render_volume(cupbr_renderer* r, const cupbr_camera& cam, uint32_t frame0, uint32_t spp);
int render_gdpt(cupbr_renderer* r, const cupbr_camera& cam, uint32_t frame0, uint32_t spp);
int render_raytracer(cupbr_renderer* r, const cupbr_camera& cam);

}  // namespace cupbr
