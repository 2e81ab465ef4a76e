// cupbr_trace_closest_device_noalloc: the batch closest-hit query of cupbr_trace_closest_device for
// callers that cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc). The traversal-stack spill lives in a fixed buffer of the
// library, one per device, instead of a scene-owned allocation; the scene must already be resident
// on the current device. Calls that overlap on one device (other streams, other threads) share the
// buffer and race.
#include <mutex>
#include <string>

#include "device.h"
#include "scene.h"

// 16 MiB: a BVH4 of depth 16 (1M triangles: 49 stack entries per lane, 17 beyond the 32 kept in
// LDS) spills at most 17 words per lane, so the buffer holds the stacks of 963 blocks of 256 lanes
// (3.8 per compute unit on 256 CUs). Host code passes the address hipGetSymbolAddress returns for
// the current device to the kernel, never the symbol itself.
static constexpr size_t kScratchWords = (size_t(16) << 20) / sizeof(uint32_t);
static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace cupbr {
namespace {

constexpr int kMaxDevices = 64;

// device address of `scratch` on the current device, looked up once per device with that device
// current (each device holds its own copy at its own address)
int scratch_address(uint32_t** out) {
    static std::mutex mu;
    static uint32_t* addr[kMaxDevices] = {};
    int dev = 0;
    CUPBR_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_error(CUPBR_ERR_HIP, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        CUPBR_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_error(CUPBR_ERR_HIP, "hipGetSymbolAddress gave no address for the scratch buffer");
        addr[dev] = static_cast<uint32_t*>(p);
    }
    *out = addr[dev];
    return CUPBR_OK;
}

}  // namespace
}  // namespace cupbr

using namespace cupbr;

extern "C" int cupbr_trace_closest_device_noalloc(const cupbr_scene* s, const float* rays, size_t n,
                                                  int32_t* object_id, int32_t* triangle_id, float* t,
                                                  void* hip_stream) {
    if (!s) return set_error(CUPBR_ERR_INVALID_ARGUMENT, "null scene");
    if (n != 0 && !(rays && object_id && triangle_id && t)) return set_error(CUPBR_ERR_INVALID_ARGUMENT, "null buffer");
    if (n == 0) return CUPBR_OK;
    int dev = 0;
    CUPBR_HIP_CHECK(hipGetDevice(&dev));
    if (!s->uploaded || s->device != dev)
        return set_error(CUPBR_ERR_INVALID_ARGUMENT,
                         "the scene is not resident on the current device: call cupbr_scene_upload there first");
    uint32_t* base = nullptr;
    if (int rc = scratch_address(&base)) return rc;
    size_t blocks = std::min<size_t>((n + kTraceBlock - 1) / kTraceBlock, (size_t)device_cu_count() * 16);
    blocks = std::max<size_t>(blocks, 1);
    uint32_t* spill = nullptr;
    if (s->stack_entries > (uint32_t)kTraceLds) {
        // (entries - 32) words per lane at a stride of the grid's lane count: the grid is cut to
        // the number of blocks whose stacks fit the buffer
        const size_t per_block = (size_t)(s->stack_entries - kTraceLds) * kTraceBlock;
        const size_t fit = kScratchWords / per_block;
        if (fit == 0)
            return set_error(CUPBR_ERR_BUFFER_SIZE,
                             "the BVH needs " + std::to_string(s->stack_entries) +
                                 " traversal stack entries per ray; the fixed scratch buffer (" +
                                 std::to_string(kScratchWords * sizeof(uint32_t)) + " bytes) holds at most " +
                                 std::to_string(kScratchWords / kTraceBlock + kTraceLds));
        blocks = std::min(blocks, fit);
        spill = base;
    }
    return trace_closest_launch(s, rays, n, object_id, triangle_id, t, spill, (int)blocks, (hipStream_t)hip_stream);
}
