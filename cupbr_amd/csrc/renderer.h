// Host-side renderer handle (cupbr_renderer): PBRenderer's state
// (lib/src/Renderer/PBRenderer.cu:26-52) plus the wavefront buffers.
#pragma once

#include <vector>

#include "device.h"
#include "nrc.h"
#include "scene.h"

namespace cupbr {

constexpr int kNumImages = 8;  // cupbr_image

// Device counters of one render call, one 128-byte line each so that atomics on
// different counters never serialise on the same L2 line. The trace kernel's work
// counters hand out 64-ray chunks to persistent waves; each counter is reset by
// the kernel that runs after its last user.
struct alignas(128) CounterLine {
    uint32_t v;
    uint32_t pad[31];
};
struct Counters {
    CounterLine queue_count[2];  // extension queues (ping-pong by bounce parity)
    CounterLine shadow_count[2]; // shadow queues (bounce parity)
    CounterLine work[8];         // trace kernel: extension rays handed out, per XCD group
    CounterLine swork[8];        // trace kernel: shadow rays handed out, per XCD group
    CounterLine query_count;  // radiance-cache queries of the batch
    CounterLine query_early;  // queries that existed after the second bounce's shade
    CounterLine row_count;    // radiance-cache training rows of the batch
    CounterLine iqueue_count[2];  // volume renderer: rays inside an object (bounce parity)
    unsigned long long extension_rays;
    unsigned long long shadow_rays;
    unsigned long long ext_nodes, ext_prims, shadow_nodes, shadow_prims;  // traversal counting
    unsigned long long refill_launches;  // trace launches that handed rays to idle lanes
};

// HIP events recorded between consecutive launches on the renderer's stream;
// stage time = time between the event before a launch and the event after it
struct StageTimer {
    std::vector<hipEvent_t> pool;
    std::vector<int> stage;  // stage[i]: stage of the launch that ends at event i (-1: first event)
    size_t used = 0;
    int mark(hipStream_t st, int s) {
        if (used == pool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return CUPBR_ERR_HIP;
            pool.push_back(e);
            stage.push_back(-1);
        }
        stage[used] = s;
        if (hipEventRecord(pool[used], st) != hipSuccess) return CUPBR_ERR_HIP;
        ++used;
        return CUPBR_OK;
    }
    ~StageTimer() {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

struct Wavefront {
    uint32_t capacity = 0;  // paths
    DevBuf weight;          // float4 per path: throughput.xyz, rng state (bits)
    DevBuf radiance;        // float4 per path
    DevBuf q_o[2], q_d[2];  // float4 per entry: origin | direction.xyz, path (bits)
    DevBuf q_hit;           // float2 per entry: t, prim (bits)
    DevBuf s_o[2], s_d[2], s_c[2];  // shadow queues by bounce parity: origin.xyz+distance |
                                    // direction.xyz+path | contribution
    DevBuf counters;
    DevBuf spill;
    DevBuf gdpt_slot;  // gradient-domain groups: queue slot of the base path's ray, by bounce parity
    DevBuf iq_o[2], iq_d[2];  // volume renderer: rays inside an object (traced against that object)
    uint32_t iq_capacity = 0;
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
    int32_t gdpt_mode = CUPBR_GDPT_TRIMMED_MEAN;  // reconstruction of GRADIENTDOMAIN frames
    int32_t rng = CUPBR_RNG_TEA_LCG;              // random numbers of PATHTRACER frames
    float gdpt_alpha = 0.2f;                      // screened-Poisson weight of the primal image
    uint32_t gdpt_iterations = 50;                // screened-Poisson Jacobi iterations
    cupbr::DevBuf images[cupbr::kNumImages];
    cupbr::DevBuf temp;
    cupbr::Wavefront wf;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    cupbr_render_stats stats{};
    bool profiling = false;
    bool counting = false;
    cupbr::StageTimer timer;
    // radiance cache (PATHTRACER only; nullptr = plain PT)
    cupbr_nrc* nrc = nullptr;
    uint32_t nrc_train_paths = 1u << 16;  // training paths per frame (full-length PT paths)
    uint32_t nrc_steps = 16;              // Adam steps per frame
    uint32_t nrc_batch = 1u << 14;        // rows per step
    float nrc_lr = 1e-3f;
    float nrc_camera_min_roughness = __builtin_inff();  // camera-hit cache vertices (+inf: none)
    uint32_t nrc_max_rows = 1u << 18;     // training rows kept per frame
    cupbr::DevBuf nrc_queries, nrc_pending, nrc_rows;
    // second stream for cache inference that overlaps the training paths' later bounces
    hipStream_t side_stream = nullptr;
    hipEvent_t ev_queries = nullptr, ev_side_done = nullptr;
    uint64_t nrc_queries_total = 0, nrc_rows_total = 0;  // of the last cupbr_render call
};

namespace cupbr {

int renderer_clear_images(cupbr_renderer* r);
// records a stage boundary when profiling is on (stage: 0 extend, 1 shade, 2 shadow, 3 other)
inline int stage_mark(cupbr_renderer* r, int stage) { return r->profiling ? r->timer.mark(r->stream, stage) : CUPBR_OK; }
// one reference frame per sample: frames [frame0, frame0+spp) of method PATHTRACER
int render_pathtracer(cupbr_renderer* r, const cupbr_camera& cam, uint32_t frame0, uint32_t spp);
int render_volume(cupbr_renderer* r, const cupbr_camera& cam, uint32_t frame0, uint32_t spp);
int render_gdpt(cupbr_renderer* r, const cupbr_camera& cam, uint32_t frame0, uint32_t spp);
// the gradient-domain paths of frames [frame0, frame0+spp) in wavefront form: running means of the
// base and the four gradient buffers (pathtracer.hip)
int gdpt_paths_wavefront(cupbr_renderer* r, const cupbr_camera& cam, uint32_t frame0, uint32_t spp);
// the volume renderer's frames [frame0, frame0+spp) in wavefront form (pathtracer.hip)
int volume_paths_wavefront(cupbr_renderer* r, const cupbr_camera& cam, uint32_t frame0, uint32_t spp);
int render_raytracer(cupbr_renderer* r, const cupbr_camera& cam);

}  // namespace cupbr
