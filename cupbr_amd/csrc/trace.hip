// Batch closest-hit and visibility queries (the parity entry points).
#include "device.h"
#include "scene.h"
#include "traverse.h"

namespace cupbr {

__global__ __launch_bounds__(kTraceBlock) void k_trace_closest(DevScene S, const float* __restrict__ rays, size_t n,
                                                               int32_t* __restrict__ obj, int32_t* __restrict__ tri,
                                                               float* __restrict__ tout, uint32_t* spill) {
    __shared__ uint32_t lds[kTraceLds * kTraceBlock];
    const size_t gtid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    LaneStack<kTraceLds> st{lds + threadIdx.x, kTraceBlock, spill ? spill + gtid : nullptr, (uint32_t)stride};
    for (size_t i = gtid; i < n; i += stride) {
        const float* r = rays + 6 * i;
        v3 o = mk(r[0], r[1], r[2]), d = mk(r[3], r[4], r[5]);
        TraceResult res;
        traverse<false>(S, o, d, 0.0f, st, res);
        int32_t ob, tr;
        hit_ids(S, res, ob, tr);
        obj[i] = ob;
        tri[i] = tr;
        tout[i] = res.t;
    }
}

// Tracing::traceVisibility: visible iff no valid hit with t <= distance
__global__ __launch_bounds__(kTraceBlock) void k_trace_visibility(DevScene S, const float* __restrict__ rays,
                                                                  const float* __restrict__ dist, size_t n,
                                                                  uint8_t* __restrict__ vis, uint32_t* spill) {
    __shared__ uint32_t lds[kTraceLds * kTraceBlock];
    const size_t gtid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    LaneStack<kTraceLds> st{lds + threadIdx.x, kTraceBlock, spill ? spill + gtid : nullptr, (uint32_t)stride};
    for (size_t i = gtid; i < n; i += stride) {
        const float* r = rays + 6 * i;
        v3 o = mk(r[0], r[1], r[2]), d = mk(r[3], r[4], r[5]);
        TraceResult res;
        bool occluded = traverse<true>(S, o, d, dist[i], st, res);
        vis[i] = occluded ? 0 : 1;
    }
}

int trace_closest_launch(const cupbr_scene* s, const float* rays, size_t n, int32_t* obj, int32_t* tri, float* t,
                         uint32_t* spill, int blocks, hipStream_t stream) {
    hipLaunchKernelGGL(k_trace_closest, dim3(blocks), dim3(kTraceBlock), 0, stream, s->dev, rays, n, obj, tri, t, spill);
    CUPBR_HIP_CHECK(hipGetLastError());
    return CUPBR_OK;
}

static int trace_grid(cupbr_scene* s, size_t n, uint32_t** spill) {
    int blocks = (int)std::min<size_t>((n + kTraceBlock - 1) / kTraceBlock, (size_t)device_cu_count() * 16);
    blocks = std::max(blocks, 1);
    *spill = nullptr;
    if (s->stack_entries > (uint32_t)kTraceLds) {
        size_t need = (size_t)(s->stack_entries - kTraceLds) * blocks * kTraceBlock * sizeof(uint32_t);
        int rc = s->spill.ensure(need);
        if (rc) return -1;
        *spill = s->spill.as<uint32_t>();
    }
    return blocks;
}

int trace_closest_device(const cupbr_scene* cs, const float* rays, size_t n, int32_t* obj, int32_t* tri, float* t,
                         hipStream_t stream) {
    if (n == 0) return CUPBR_OK;
    cupbr_scene* s = const_cast<cupbr_scene*>(cs);
    if (int rc = scene_upload(s)) return rc;
    uint32_t* spill;
    if (int rc = scene_stream_acquire(s, stream)) return rc;
    int blocks = trace_grid(s, n, &spill);
    if (blocks < 0) return CUPBR_ERR_OUT_OF_MEMORY;
    if (int rc = trace_closest_launch(s, rays, n, obj, tri, t, spill, blocks, stream)) return rc;
    return scene_stream_release(s, stream);
}

int trace_visibility_device(const cupbr_scene* cs, const float* rays, const float* dist, size_t n, uint8_t* vis,
                            hipStream_t stream) {
    if (n == 0) return CUPBR_OK;
    cupbr_scene* s = const_cast<cupbr_scene*>(cs);
    if (int rc = scene_upload(s)) return rc;
    uint32_t* spill;
    if (int rc = scene_stream_acquire(s, stream)) return rc;
    int blocks = trace_grid(s, n, &spill);
    if (blocks < 0) return CUPBR_ERR_OUT_OF_MEMORY;
    hipLaunchKernelGGL(k_trace_visibility, dim3(blocks), dim3(kTraceBlock), 0, stream, s->dev, rays, dist, n, vis, spill);
    CUPBR_HIP_CHECK(hipGetLastError());
    return scene_stream_release(s, stream);
}

}  // namespace cupbr
