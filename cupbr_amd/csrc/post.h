// Post chain handles: PostProcessor (lib/src/PostProcessing/PostProcessor.cpp) and
// ToneMapper (lib/src/Renderer/ToneMapper.cu).
#pragma once

#include "device.h"

struct cupbr_post {
    const float* hdr = nullptr;  // borrowed, like PostProcessor::registerImage
    uint32_t width = 0, height = 0;
    bool registered = false;
    cupbr::DevBuf output;        // getPostProcessBuffer()
    hipStream_t stream = nullptr;
    int32_t pyramid = CUPBR_BLOOM_REFERENCE;
    cupbr::DevBuf levels;        // Gaussian pyramid: the thresholded image and five levels
};

struct cupbr_tonemapper {
    int32_t type = CUPBR_REINHARD;
    float exposure = 1.0f;       // ToneMapper.cu:94
    const float* hdr = nullptr;  // borrowed
    uint32_t width = 0, height = 0;
    bool registered = false;
    cupbr::DevBuf rgba;          // RenderBuffer, owned
    hipStream_t stream = nullptr;
};

namespace cupbr {

// bloom of p->hdr into p->output; when rgba is given the tone map of the bloomed image is
// written in the same pass (type/exposure as ToneMapper)
int post_bloom(cupbr_post* p, const float curve[4], uint8_t* rgba, int32_t type, float exposure);
// level k (1..5) of the Gaussian pyramid of a w x h image: ((dim_{k-1} + 1) / 2) per axis
constexpr int kGaussLevels = 5;
int tonemap(cupbr_tonemapper* t);
int write_bmp(const char* path, const uint8_t* rgba, uint32_t w, uint32_t h);

}  // namespace cupbr
