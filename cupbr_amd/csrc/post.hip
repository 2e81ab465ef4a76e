// Bloom and tone mapping.
//
// The reference builds a 6-level pyramid in 13 launches (PostProcessor.cpp:125-142,
// PyramidConstructor.cu:9-169), but its upsampling reads the *down* pyramid
// (`up[i] = tent(down[i+1])/16 + down[i]`, :79-80,111), so the output depends on
// levels 0 and 1 only:
//     out(p) = hdr(p) + ( tent3x3(down1)(p/2) / 16 + down0(p) )
//     down0  = threshold(hdr),  down1(q) = box2x2(down0)(2q)   (edge-clamped)
// This file computes exactly that, in the reference's FP32 operation order, in
// one pass per 64x8 tile: the tile's 34x6 block of down1 (with its one-texel
// halo) is built in LDS from hdr reads, then every output pixel reads its own
// hdr value again (an L1/L2 hit). A tile's reads reach 2 pixels into each
// neighbouring tile, so tiles are assigned to XCDs in contiguous row-major
// bands: the neighbours run on the same XCD at about the same time and their
// shared lines come from that XCD's L2. HBM traffic is then about one FP32 RGB
// read and one write per pixel (+ 4 bytes when the RGBA8 tone map is fused in).
#include <cstdio>
#include <vector>

#include "post.h"
#include "vecmath.h"

namespace cupbr {

constexpr int kTileW = 64, kTileH = 8;
constexpr int kD1W = kTileW / 2 + 2, kD1H = kTileH / 2 + 2;

// thresholding_kernel, PyramidConstructor.cu:9-30
__device__ __forceinline__ v3 threshold_px(v3 c, float4 t) {
    float b = c.x > c.y ? c.x : c.y;
    b = c.z > b ? c.z : b;
    float rq = clampf(b - t.y, 0.0f, t.z);
    rq = t.w * rq * rq;
    const float s = fmaxf(rq, b - t.x) / fmaxf(b, 1e-5f);
    return v3{c.x * s, c.y * s, c.z * s};
}

// reinhard_kernel / gamma_kernel, ToneMapper.cu:17-69 (truncating casts; 1.0 - expf() and the
// blue channel's 1.0f/2.2 are double expressions in the reference)
__device__ __forceinline__ uchar4 tonemap_px(v3 c, int32_t type, float exposure) {
    float m[3];
    const float cc[3] = {c.x, c.y, c.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (type == CUPBR_REINHARD) {
            const float base = (float)(1.0 - (double)expf(-cc[k] * exposure));
            m[k] = powf(base, k == 2 ? (float)(1.0f / 2.2) : 1.0f / 2.2f);
        } else if (type == CUPBR_ACES) {  // Narkowicz's ACES filmic fit, then gamma 1/2.2
            const float x = cc[k] * exposure;
            const float f = (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
            m[k] = powf(clampf(f, 0.0f, 1.0f), 1.0f / 2.2f);
        } else {
            const float x = cc[k];
            m[k] = x <= 0.0031308f ? 12.92f * x : 1.055f * powf(x, 1.0f / 2.4f) - 0.055f;
        }
    }
    return make_uchar4((uint8_t)(clampf(m[0], 0.0f, 1.0f) * 255.0f), (uint8_t)(clampf(m[1], 0.0f, 1.0f) * 255.0f),
                       (uint8_t)(clampf(m[2], 0.0f, 1.0f) * 255.0f), 255);
}

__device__ __forceinline__ v3 ld_px(const float* img, uint32_t w, int x, int y) {
    return ld3(img + 3 * ((size_t)y * w + x));
}

__global__ __launch_bounds__(256) void k_bloom(const float* __restrict__ hdr, float* __restrict__ out,
                                               uchar4* __restrict__ rgba, uint32_t w, uint32_t h, float4 curve,
                                               int32_t type, float exposure) {
    __shared__ float d1[3][kD1H][kD1W];
    // workgroups are dispatched to the 8 XCDs round-robin by block id: block b runs on XCD b % 8,
    // which takes tiles [xcd * q + min(xcd, r), ... + its block count) in row-major order
    // (a bijection onto the n tiles: XCD x has q + (x < r) blocks)
    const uint32_t n = gridDim.x, q = n / 8u, rem = n % 8u, xcd = blockIdx.x % 8u;
    const uint32_t tile = xcd * q + (xcd < rem ? xcd : rem) + blockIdx.x / 8u;
    const uint32_t tiles_x = (w + kTileW - 1) / kTileW;
    const int x0 = (int)(tile % tiles_x) * kTileW, y0 = (int)(tile / tiles_x) * kTileH;
    const int w1 = (int)(w / 2), h1 = (int)(h / 2);
    const int qx0 = x0 / 2 - 1, qy0 = y0 / 2 - 1;
    const int tid = threadIdx.x;
    if (tid < kD1W * kD1H) {
        const int j = tid / kD1W, i = tid - j * kD1W;
        int qx = qx0 + i, qy = qy0 + j;
        qx = qx < 0 ? 0 : (qx >= w1 ? w1 - 1 : qx);  // upscale_and_combine_kernel :101-104
        qy = qy < 0 ? 0 : (qy >= h1 ? h1 - 1 : qy);
        v3 r = splat(0.0f);
#pragma unroll
        for (int k = 0; k < 4; ++k) {  // downsample_kernel :51-64, reads (0,0),(0,1),(1,0),(1,1)
            int sx = 2 * qx + (k >> 1), sy = 2 * qy + (k & 1);
            if (sx >= (int)w) sx = (int)w - 1;
            if (sy >= (int)h) sy = (int)h - 1;
            r += threshold_px(ld_px(hdr, w, sx, sy), curve);
        }
        r = r / 4.0f;
        d1[0][j][i] = r.x;
        d1[1][j][i] = r.y;
        d1[2][j][i] = r.z;
    }
    __syncthreads();
    const int tx = tid & (kTileW - 1);
#pragma unroll
    for (int row = 0; row < kTileH; row += 256 / kTileW) {
        const int x = x0 + tx, y = y0 + row + tid / kTileW;
        if (x >= (int)w || y >= (int)h) continue;
        const v3 c = ld_px(hdr, w, x, y);
        const v3 d0 = threshold_px(c, curve);
        const int cx = x / 2 - qx0, cy = y / 2 - qy0;
        v3 res = splat(0.0f);
        const float kw[3] = {1.0f, 2.0f, 1.0f};
#pragma unroll
        for (int u = -1; u < 2; ++u)
#pragma unroll
            for (int v = -1; v < 2; ++v) {
                const float k = kw[u + 1] * kw[v + 1];
                res += v3{d1[0][cy + v][cx + u], d1[1][cy + v][cx + u], d1[2][cy + v][cx + u]} * k;
            }
        const v3 up0 = res / 16.0f + d0;  // :111
        const v3 o = c + up0;             // combine_kernel :124
        st3(out + 3 * ((size_t)y * w + x), o);
        if (rgba) rgba[(size_t)y * w + x] = tonemap_px(o, type, exposure);
    }
}

__global__ __launch_bounds__(256) void k_tonemap(const float* __restrict__ hdr, uchar4* __restrict__ rgba, size_t n,
                                                 int32_t type, float exposure) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    rgba[i] = tonemap_px(ld3(hdr + 3 * i), type, exposure);
}

// ------------------------------------------------------------------ Gaussian pyramid bloom
// (CUPBR_BLOOM_GAUSSIAN5, an addition: see cupbr_post_set_bloom_pyramid for the definition). Each
// level is one launch over its pixels; every level fits L2 from the second on, so these are
// latency- rather than bandwidth-bound at 1080p.
__global__ __launch_bounds__(256) void k_gauss_threshold(const float* __restrict__ hdr, float* __restrict__ l0, size_t n,
                                                         float4 curve) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) st3(l0 + 3 * i, threshold_px(ld3(hdr + 3 * i), curve));
}

// dst(x, y) = blur(src)(2x, 2y): rows of five taps [1 4 6 4 1] summed first, then the five rows,
// then / 256 (edge-clamped)
__global__ __launch_bounds__(256) void k_gauss_down(const float* __restrict__ src, uint32_t sw, uint32_t sh,
                                                    float* __restrict__ dst, uint32_t dw, uint32_t dh) {
    const uint32_t x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= dw || y >= dh) return;
    const float wt[5] = {1.0f, 4.0f, 6.0f, 4.0f, 1.0f};
    v3 acc = splat(0.0f);
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        const int sy = min(max((int)(2 * y) + j - 2, 0), (int)sh - 1);
        v3 row = splat(0.0f);
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const int sx = min(max((int)(2 * x) + i - 2, 0), (int)sw - 1);
            row += ld_px(src, sw, sx, sy) * wt[i];
        }
        acc += row * wt[j];
    }
    st3(dst + 3 * ((size_t)y * dw + x), acc / 256.0f);
}

// coarse level (cw x ch) sampled at fine pixel (x, y): the texel (x/2, y/2), the mean of two
// texels along an odd coordinate (edge-clamped)
__device__ __forceinline__ v3 gauss_up(const float* coarse, uint32_t cw, uint32_t ch, uint32_t x, uint32_t y) {
    const uint32_t x0 = x >> 1, y0 = y >> 1;
    const uint32_t x1 = (x & 1u) ? min(x0 + 1u, cw - 1u) : x0, y1 = (y & 1u) ? min(y0 + 1u, ch - 1u) : y0;
    const v3 a = ld_px(coarse, cw, (int)x0, (int)y0), b = ld_px(coarse, cw, (int)x1, (int)y0);
    const v3 c = ld_px(coarse, cw, (int)x0, (int)y1), d = ld_px(coarse, cw, (int)x1, (int)y1);
    return ((a + b) * 0.5f + (c + d) * 0.5f) * 0.5f;
}

// fine += up(coarse)
__global__ __launch_bounds__(256) void k_gauss_up_add(const float* __restrict__ coarse, uint32_t cw, uint32_t ch,
                                                      float* __restrict__ fine, uint32_t fw, uint32_t fh) {
    const uint32_t x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= fw || y >= fh) return;
    float* f = fine + 3 * ((size_t)y * fw + x);
    st3(f, ld3(f) + gauss_up(coarse, cw, ch, x, y));
}

// out = hdr + (L0 + up(U1)) / 6, with the tone map fused when rgba is given
__global__ __launch_bounds__(256) void k_gauss_combine(const float* __restrict__ hdr, const float* __restrict__ l0,
                                                       const float* __restrict__ u1, uint32_t w, uint32_t h,
                                                       uint32_t w1, uint32_t h1, float* __restrict__ out,
                                                       uchar4* __restrict__ rgba, int32_t type, float exposure) {
    const uint32_t x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= w || y >= h) return;
    const size_t i = (size_t)y * w + x;
    const v3 u0 = ld3(l0 + 3 * i) + gauss_up(u1, w1, h1, x, y);
    const v3 o = ld3(hdr + 3 * i) + u0 / 6.0f;
    st3(out + 3 * i, o);
    if (rgba) rgba[i] = tonemap_px(o, type, exposure);
}

static int gauss_bloom(cupbr_post* p, float4 curve, uint8_t* rgba, int32_t type, float exposure) {
    uint32_t lw[kGaussLevels + 1], lh[kGaussLevels + 1];
    size_t off[kGaussLevels + 2];
    lw[0] = p->width;
    lh[0] = p->height;
    off[0] = 0;
    for (int k = 0; k <= kGaussLevels; ++k) {
        if (k > 0) {
            lw[k] = (lw[k - 1] + 1) / 2;
            lh[k] = (lh[k - 1] + 1) / 2;
        }
        off[k + 1] = off[k] + (size_t)lw[k] * lh[k] * 3;
    }
    if (int rc = p->levels.ensure(off[kGaussLevels + 1] * sizeof(float))) return rc;
    float* L = p->levels.as<float>();
    auto grid = [](uint32_t w, uint32_t h) { return dim3((w + 63) / 64, (h + 3) / 4); };
    const size_t n = (size_t)p->width * p->height;
    hipLaunchKernelGGL(k_gauss_threshold, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream, p->hdr, L, n, curve);
    for (int k = 1; k <= kGaussLevels; ++k)
        hipLaunchKernelGGL(k_gauss_down, grid(lw[k], lh[k]), dim3(256), 0, p->stream, L + off[k - 1], lw[k - 1],
                           lh[k - 1], L + off[k], lw[k], lh[k]);
    for (int k = kGaussLevels - 1; k >= 1; --k)  // U_k = L_k + up(U_{k+1}), in place
        hipLaunchKernelGGL(k_gauss_up_add, grid(lw[k], lh[k]), dim3(256), 0, p->stream, L + off[k + 1], lw[k + 1],
                           lh[k + 1], L + off[k], lw[k], lh[k]);
    hipLaunchKernelGGL(k_gauss_combine, grid(lw[0], lh[0]), dim3(256), 0, p->stream, p->hdr, L, L + off[1], lw[0],
                       lh[0], lw[1], lh[1], p->output.as<float>(), reinterpret_cast<uchar4*>(rgba), type, exposure);
    CUPBR_HIP_CHECK(hipGetLastError());
    return CUPBR_OK;
}

int post_bloom(cupbr_post* p, const float curve[4], uint8_t* rgba, int32_t type, float exposure) {
    if (p->width < 2 || p->height < 2)
        return set_error(CUPBR_ERR_INVALID_ARGUMENT, "bloom needs an image of at least 2x2 pixels");
    if (p->pyramid == CUPBR_BLOOM_GAUSSIAN5) {
        CUPBR_HIP_CHECK(hipDeviceSynchronize());  // the registered image is borrowed, as below
        if (int rc = gauss_bloom(p, make_float4(curve[0], curve[1], curve[2], curve[3]), rgba, type, exposure))
            return rc;
        CUPBR_HIP_CHECK(hipStreamSynchronize(p->stream));
        return CUPBR_OK;
    }
    const uint32_t tiles = ((p->width + kTileW - 1) / kTileW) * ((p->height + kTileH - 1) / kTileH);
    dim3 grid(tiles);
    float4 c = make_float4(curve[0], curve[1], curve[2], curve[3]);
    // the registered image is borrowed: finish pending writes to it first, and return only when
    // the outputs are complete (the reference synchronises after every launch, PyramidConstructor.cu:134)
    CUPBR_HIP_CHECK(hipDeviceSynchronize());
    hipLaunchKernelGGL(k_bloom, grid, dim3(256), 0, p->stream, p->hdr, p->output.as<float>(),
                       reinterpret_cast<uchar4*>(rgba), p->width, p->height, c, type, exposure);
    CUPBR_HIP_CHECK(hipGetLastError());
    CUPBR_HIP_CHECK(hipStreamSynchronize(p->stream));
    return CUPBR_OK;
}

int tonemap(cupbr_tonemapper* t) {
    const size_t n = (size_t)t->width * t->height;
    CUPBR_HIP_CHECK(hipDeviceSynchronize());
    hipLaunchKernelGGL(k_tonemap, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, t->stream, t->hdr,
                       t->rgba.as<uchar4>(), n, t->type, t->exposure);
    CUPBR_HIP_CHECK(hipGetLastError());
    CUPBR_HIP_CHECK(hipStreamSynchronize(t->stream));  // ToneMapper.cu:172
    return CUPBR_OK;
}

// 32-bit BI_RGB bitmap, bottom-up: buffer row 0 is written first and shown at the bottom,
// as stbi_write_bmp does after stbi_flip_vertically_on_write(true) (ToneMapper.cu:194-195)
int write_bmp(const char* path, const uint8_t* rgba, uint32_t w, uint32_t h) {
    FILE* f = fopen(path, "wb");
    if (!f) return set_error(CUPBR_ERR_FILE, std::string("cannot write ") + path, true);
    const uint32_t data = w * h * 4, off = 14 + 40, size = off + data;
    uint8_t hdr[54] = {'B', 'M'};
    auto put32 = [&](int at, uint32_t v) { for (int k = 0; k < 4; ++k) hdr[at + k] = (uint8_t)(v >> (8 * k)); };
    put32(2, size);
    put32(10, off);
    put32(14, 40);
    put32(18, w);
    put32(22, h);
    hdr[26] = 1;
    hdr[28] = 32;
    put32(34, data);
    put32(38, 2835);
    put32(42, 2835);
    bool ok = fwrite(hdr, 1, 54, f) == 54;
    std::vector<uint8_t> row((size_t)w * 4);
    for (uint32_t y = 0; y < h && ok; ++y) {
        for (uint32_t x = 0; x < w; ++x) {
            const uint8_t* s = rgba + 4 * ((size_t)y * w + x);
            row[4 * x] = s[2];
            row[4 * x + 1] = s[1];
            row[4 * x + 2] = s[0];
            row[4 * x + 3] = s[3];
        }
        ok = fwrite(row.data(), 1, row.size(), f) == row.size();
    }
    fclose(f);
    return ok ? CUPBR_OK : set_error(CUPBR_ERR_FILE, std::string("write failed: ") + path, true);
}

}  // namespace cupbr
