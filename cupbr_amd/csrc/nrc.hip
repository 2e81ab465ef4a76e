// Neural radiance cache: a 64 -> 64 -> 64 -> 64 -> 64 -> 3 ReLU MLP without biases (16,576
// weights), BF16 operands on the CDNA4 matrix cores (v_mfma_f32_16x16x32_bf16) with FP32
// accumulation, trained online with Adam on a relative L2 loss. No counterpart in the reference.
//
// Inputs (16 floats per query): position (3, world), outgoing direction wo (3, unit), shading
// normal (3, unit), roughness (1), diffuse albedo (3), specular albedo (3). Encoding (64 features):
//   0..35   sin/cos(2^k pi p_a), p normalised to the cache's box, a = 0..2, k = 0..5
//   36..43  sin/cos(2^k pi x), x = acos(wo.z)/pi and (atan2(wo.y, wo.x)/pi + 1)/2, k = 0..1
//   44..51  the same for the normal
//   52      1 - exp(-roughness);  53..55 diffuse albedo;  56..58 specular albedo;  59 one; 60..63 zero
//
// Matrix-core layout (one wave, 16 queries per tile; queries are the N dimension):
//   D[16 x 16] += A[16 x 32] * B[32 x 16]; lane l holds A[l%16][8(l/16) + j], B[8(l/16) + j][l%16]
//   (j = 0..7) and D[4(l/16) + v][l%16] (v = 0..3).
// A layer's 64 outputs are 4 D tiles; lane l then holds outputs t*16 + 4(l/16) + v of query l%16,
// t = 0..3. Those 16 values become the next layer's B operand unchanged: K slot kk of chunk kc
// stands for feature perm(kc, kk) = (2kc + (kk%8)/4)*16 + (kk/8)*4 + kk%4, and the weights are
// stored pre-permuted to match, so activations never leave registers between layers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "device.h"
#include "nrc.h"

namespace cupbr {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kFwdLayer = 4096;                           // bf16 per packed 64x64 layer
constexpr int kFwdSize = 4 * kFwdLayer + 2 * 64 * 8;      // + W5 (one output tile)
constexpr int kBwdSize = 3 * kFwdLayer + 4 * 64 * 8;      // W2^T..W4^T + W5^T
constexpr int kNrcBlock = 256;

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// feature i of the encoding (see the file comment); d = derived inputs:
// d[0..2] normalised position, d[3..4] wo angles, d[5..6] normal angles, d[7] roughness, d[8..13] albedos
__device__ __forceinline__ float feature(int i, const float* d) {
    const float pi = 3.14159265358979f;
    if (i < 52) {
        int a, k, sc;
        float x;
        if (i < 36) {
            a = i / 12;
            k = (i % 12) >> 1;
            sc = i & 1;
            x = d[a];
        } else {
            const int r = (i - 36) & 7;
            a = (i < 44 ? 3 : 5) + (r >> 2);
            k = (r & 3) >> 1;
            sc = r & 1;
            x = d[a];
        }
        const float arg = (float)(1 << k) * pi * x;
        return sc ? cosf(arg) : sinf(arg);
    }
    if (i == 52) return 1.0f - expf(-d[7]);
    if (i < 59) return d[8 + (i - 53)];
    return i == 59 ? 1.0f : 0.0f;
}

__device__ __forceinline__ void derive(const NrcBox& box, const float* raw, float* d) {
    const float pi = 3.14159265358979f;
    for (int a = 0; a < 3; ++a) d[a] = (raw[a] - box.lo[a]) * box.inv_ext[a];
    d[3] = acosf(fminf(fmaxf(raw[5], -1.0f), 1.0f)) / pi;
    d[4] = (atan2f(raw[4], raw[3]) / pi + 1.0f) * 0.5f;
    d[5] = acosf(fminf(fmaxf(raw[8], -1.0f), 1.0f)) / pi;
    d[6] = (atan2f(raw[7], raw[6]) / pi + 1.0f) * 0.5f;
    d[7] = raw[9];
    for (int k = 0; k < 6; ++k) d[8 + k] = raw[10 + k];
}

// sin(pi y) and cos(pi y) from the hardware's sine and cosine, whose argument is in revolutions
// (v_sin_f32(r) = sin(2 pi r)): r = y / 2 is exact, and its fraction keeps the argument in [0, 1),
// well inside the instructions' input range. Their error (~1e-6 absolute) is far below the BF16
// rounding the feature then takes.
__device__ __forceinline__ void sincospi_hw(float y, float* s, float* c) {
    float r = 0.5f * y;
    r -= floorf(r);
    *s = __builtin_amdgcn_sinf(r);
    *c = __builtin_amdgcn_cosf(r);
}

// B operand of the first layer for lane group g: slot (kc, j) holds feature perm(kc, 8g + j)
// Slots j and j + 1 (j even) hold features i and i + 1 with i even: below 52 these are the sine
// and cosine of the same argument 2^k pi x, computed together from the exact 2^k x.
__device__ __forceinline__ void encode(const NrcBox& box, const float* raw, int g, bf16x8 b[2]) {
    float d[14];
    derive(box, raw, d);
    for (int kc = 0; kc < 2; ++kc)
        for (int j = 0; j < 8; j += 2) {
            const int i = (2 * kc + (j >> 2)) * 16 + g * 4 + (j & 3);
            float f0, f1;
            if (i < 52) {
                int a, k;
                if (i < 36) {
                    a = i / 12;
                    k = (i % 12) >> 1;
                } else {
                    const int r = (i - 36) & 7;
                    a = (i < 44 ? 3 : 5) + (r >> 2);
                    k = (r & 3) >> 1;
                }
                sincospi_hw((float)(1 << k) * d[a], &f0, &f1);
            } else {
                f0 = feature(i, d);
                f1 = feature(i + 1, d);
            }
            b[kc][j] = to_bf16(f0);
            b[kc][j + 1] = to_bf16(f1);
        }
}

// one hidden layer: out = relu(W x) as the next B operand; w = packed fragments [ot][kc][lane][8]
__device__ __forceinline__ void hidden_layer(const bf16x8* w, int lane, const bf16x8 in[2], bf16x8 out[2],
                                             f32x4* pre = nullptr) {
    f32x4 acc[4];
    for (int ot = 0; ot < 4; ++ot) {
        acc[ot] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int kc = 0; kc < 2; ++kc) acc[ot] = mfma(w[(ot * 2 + kc) * 64 + lane], in[kc], acc[ot]);
    }
    for (int kc = 0; kc < 2; ++kc)
        for (int j = 0; j < 8; ++j) out[kc][j] = to_bf16(fmaxf(acc[2 * kc + (j >> 2)][j & 3], 0.0f));
    if (pre)
        for (int ot = 0; ot < 4; ++ot) pre[ot] = acc[ot];
}

// hidden_layer for two tiles of 16 queries: each weight fragment read from LDS feeds both
__device__ __forceinline__ void hidden_layer2(const bf16x8* w, int lane, const bf16x8 in0[2], const bf16x8 in1[2],
                                              bf16x8 out0[2], bf16x8 out1[2]) {
    f32x4 acc0[4], acc1[4];
    for (int ot = 0; ot < 4; ++ot) {
        acc0[ot] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        acc1[ot] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int kc = 0; kc < 2; ++kc) {
            const bf16x8 wf = w[(ot * 2 + kc) * 64 + lane];
            acc0[ot] = mfma(wf, in0[kc], acc0[ot]);
            acc1[ot] = mfma(wf, in1[kc], acc1[ot]);
        }
    }
    for (int kc = 0; kc < 2; ++kc)
        for (int j = 0; j < 8; ++j) {
            out0[kc][j] = to_bf16(fmaxf(acc0[2 * kc + (j >> 2)][j & 3], 0.0f));
            out1[kc][j] = to_bf16(fmaxf(acc1[2 * kc + (j >> 2)][j & 3], 0.0f));
        }
}

// output layer: lane group 0 gets the 3 outputs (v = 0..2) of query lane%16
__device__ __forceinline__ f32x4 output_layer(const bf16x8* w5, int lane, const bf16x8 in[2]) {
    f32x4 acc{0.0f, 0.0f, 0.0f, 0.0f};
    for (int kc = 0; kc < 2; ++kc) acc = mfma(w5[kc * 64 + lane], in[kc], acc);
    return acc;
}

__device__ __forceinline__ void load_weights_lds(bf16x8* lds, const bf16x8* __restrict__ packed, int n_frag) {
    for (int i = threadIdx.x; i < n_frag; i += blockDim.x) lds[i] = packed[i];
    __syncthreads();
}

// inference over n queries of `stride` floats (first 16 = inputs). kAccumulate: out_rgb[path] +=
// thr * max(pred, 0) with thr = record[16..18] and path = bits of record[19]; else out[q*3 + c] = pred.
template <bool kAccumulate>
__global__ __launch_bounds__(kNrcBlock) void k_nrc_infer(const bf16x8* __restrict__ packed, NrcBox box,
                                                         const float* __restrict__ rec, uint32_t n,
                                                         const uint32_t* __restrict__ n_dev, uint32_t stride,
                                                         float* __restrict__ out,
                                                         const uint32_t* __restrict__ begin_dev) {
    __shared__ bf16x8 w[kFwdSize / 8];
    load_weights_lds(w, packed, kFwdSize / 8);
    if (n_dev) n = min(n, *n_dev);  // count written by an earlier kernel on the stream
    const uint32_t begin = begin_dev ? min(*begin_dev, n) : 0u;
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t pairs = (n - begin + 31) / 32;  // two tiles of 16 queries per wave and step
    for (uint32_t t = wave; t < pairs; t += n_waves) {
        uint32_t q[2];
        bf16x8 a[2][2], b[2][2];
        for (int h = 0; h < 2; ++h) {
            q[h] = begin + t * 32 + 16 * h + (lane & 15);
            float raw[16];
            if (q[h] < n) {
                const float4* r4 = reinterpret_cast<const float4*>(rec + (size_t)q[h] * stride);
                for (int k = 0; k < 4; ++k) {
                    const float4 v = r4[k];
                    raw[4 * k] = v.x;
                    raw[4 * k + 1] = v.y;
                    raw[4 * k + 2] = v.z;
                    raw[4 * k + 3] = v.w;
                }
            } else {
                for (int k = 0; k < 16; ++k) raw[k] = 0.0f;
                raw[5] = 1.0f;
                raw[8] = 1.0f;
            }
            encode(box, raw, g, a[h]);
        }
        hidden_layer2(w + 0 * (kFwdLayer / 8), lane, a[0], a[1], b[0], b[1]);
        hidden_layer2(w + 1 * (kFwdLayer / 8), lane, b[0], b[1], a[0], a[1]);
        hidden_layer2(w + 2 * (kFwdLayer / 8), lane, a[0], a[1], b[0], b[1]);
        hidden_layer2(w + 3 * (kFwdLayer / 8), lane, b[0], b[1], a[0], a[1]);
        for (int h = 0; h < 2; ++h) {
            const f32x4 o = output_layer(w + 4 * (kFwdLayer / 8), lane, a[h]);
            if (g == 0 && q[h] < n) {
                if (kAccumulate) {
                    const float* r = rec + (size_t)q[h] * stride;
                    const uint32_t path = __float_as_uint(r[19]);
                    float* px = out + 4 * (size_t)path;
                    px[0] += r[16] * fmaxf(o[0], 0.0f);
                    px[1] += r[17] * fmaxf(o[1], 0.0f);
                    px[2] += r[18] * fmaxf(o[2], 0.0f);
                } else {
                    out[3 * (size_t)q[h] + 0] = o[0];
                    out[3 * (size_t)q[h] + 1] = o[1];
                    out[3 * (size_t)q[h] + 2] = o[2];
                }
            }
        }
    }
}

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// activations of the batch for the weight gradients: X[layer][feature][query] (inputs of layers
// 1..5) and G[layer][feature][query] (loss gradient at the outputs of layers 1..5, after the ReLU
// mask), both BF16; batch row b uses training row rows[(hash(seed + b)) % n_rows]
__device__ __forceinline__ void store_feats(__bf16* base, uint32_t B, uint32_t q, int g, const bf16x8 v[2]) {
    for (int kc = 0; kc < 2; ++kc)
        for (int j = 0; j < 8; ++j) base[(size_t)((2 * kc + (j >> 2)) * 16 + g * 4 + (j & 3)) * B + q] = v[kc][j];
}

__global__ __launch_bounds__(kNrcBlock) void k_nrc_forward_backward(const bf16x8* __restrict__ pf,
                                                                    const bf16x8* __restrict__ pb, NrcBox box,
                                                                    const float* __restrict__ rows, uint32_t n_rows,
                                                                    uint32_t B, uint32_t seed, __bf16* __restrict__ X,
                                                                    __bf16* __restrict__ G, float* loss) {
    __shared__ bf16x8 wf[kFwdSize / 8];
    __shared__ bf16x8 wb[kBwdSize / 8];
    for (int i = threadIdx.x; i < kBwdSize / 8; i += blockDim.x) wb[i] = pb[i];
    load_weights_lds(wf, pf, kFwdSize / 8);
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
    const size_t layer = (size_t)64 * B;
    float loss_acc = 0.0f;
    for (uint32_t t = wave; t < B / 16; t += n_waves) {
        const uint32_t q = t * 16 + (lane & 15);
        const uint32_t row = hash32(seed + q) % n_rows;
        const float* r = rows + (size_t)row * kNrcRowStride;
        float raw[16];
        for (int k = 0; k < 16; ++k) raw[k] = r[k];
        bf16x8 x[5][2];
        f32x4 pre[4][4];
        encode(box, raw, g, x[0]);
        for (int L = 0; L < 4; ++L) hidden_layer(wf + L * (kFwdLayer / 8), lane, x[L], x[L + 1], pre[L]);
        const f32x4 o = output_layer(wf + 4 * (kFwdLayer / 8), lane, x[4]);
        // relative L2: sum_c (o_c - y_c)^2 / (o_c^2 + 0.01), o_c detached in the denominator
        bf16x8 dy[2];
        for (int kc = 0; kc < 2; ++kc)
            for (int j = 0; j < 8; ++j) dy[kc][j] = to_bf16(0.0f);
        float d5[3] = {0.0f, 0.0f, 0.0f};
        if (g == 0) {
            for (int c = 0; c < 3; ++c) {
                const float den = o[c] * o[c] + 0.01f;
                const float e = o[c] - r[16 + c];
                loss_acc += e * e / den;
                d5[c] = 2.0f * e / den;
                dy[0][c] = to_bf16(d5[c]);
            }
        }
        for (int L = 0; L < 5; ++L) store_feats(X + L * layer, B, q, g, x[L]);
        // G[4] holds rows 0..2 of the output gradient (rows 3..63 stay zero)
        if (g == 0)
            for (int c = 0; c < 3; ++c) G[4 * layer + (size_t)c * B + q] = dy[0][c];
        // backward through W5^T, then W4^T..W2^T, masking with the ReLU of each hidden layer
        bf16x8 cur[2] = {dy[0], dy[1]};
        for (int L = 3; L >= 0; --L) {
            f32x4 acc[4];
            for (int it = 0; it < 4; ++it) {
                acc[it] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                if (L == 3) {
                    acc[it] = mfma(wb[3 * (kFwdLayer / 8) + it * 64 + lane], cur[0], acc[it]);
                } else {
                    for (int kc = 0; kc < 2; ++kc)
                        acc[it] = mfma(wb[L * (kFwdLayer / 8) + (it * 2 + kc) * 64 + lane], cur[kc], acc[it]);
                }
            }
            bf16x8 nxt[2];
            for (int kc = 0; kc < 2; ++kc)
                for (int j = 0; j < 8; ++j) {
                    const int ot = 2 * kc + (j >> 2), v = j & 3;
                    nxt[kc][j] = to_bf16(pre[L][ot][v] > 0.0f ? acc[ot][v] : 0.0f);
                }
            store_feats(G + L * layer, B, q, g, nxt);
            cur[0] = nxt[0];
            cur[1] = nxt[1];
        }
    }
    if (loss) {
        for (int off = 32; off > 0; off >>= 1) loss_acc += __shfl_xor(loss_acc, off);
        if (lane == 0) atomicAdd(loss, loss_acc);
    }
}

// weight gradients of one layer over one slice of the batch: dW[o][i] = sum_q G[o][q] X[i][q],
// written (not accumulated) to partial[slice][param offset]
template <int kOutTiles>  // 4: hidden layers W1..W4; 1: output layer W5
__device__ __forceinline__ void weight_grad_layer(int layer, const __bf16* __restrict__ X,
                                                  const __bf16* __restrict__ G, uint32_t B, uint32_t slice,
                                                  float* __restrict__ partial) {
    const uint32_t s = blockIdx.x;          // slice index
    const int lane = threadIdx.x, g = lane >> 4;
    constexpr int n_ot = kOutTiles;
    const __bf16* Xl = X + (size_t)layer * 64 * B;
    const __bf16* Gl = G + (size_t)layer * 64 * B;
    f32x4 acc[kOutTiles][4];
#pragma unroll
    for (int a = 0; a < kOutTiles; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t q0 = s * slice; q0 < (s + 1) * slice; q0 += 32) {
        const uint32_t qq = q0 + 8 * g;
        bf16x8 bx[4];
        for (int it = 0; it < 4; ++it)
            bx[it] = *reinterpret_cast<const bf16x8*>(Xl + (size_t)(it * 16 + (lane & 15)) * B + qq);
        for (int ot = 0; ot < n_ot; ++ot) {
            const bf16x8 ag = *reinterpret_cast<const bf16x8*>(Gl + (size_t)(ot * 16 + (lane & 15)) * B + qq);
            for (int it = 0; it < 4; ++it) acc[ot][it] = mfma(ag, bx[it], acc[ot][it]);
        }
    }
    float* out = partial + (size_t)s * kNrcParams;
    for (int ot = 0; ot < n_ot; ++ot)
        for (int it = 0; it < 4; ++it)
            for (int v = 0; v < 4; ++v) {
                const int o = ot * 16 + g * 4 + v, i = it * 16 + (lane & 15);
                if (layer == 4) {
                    if (o < 3) out[4 * 4096 + o * 64 + i] = acc[ot][it][v];
                } else {
                    out[layer * 4096 + o * 64 + i] = acc[ot][it][v];
                }
            }
}

// all five layers in one launch: blockIdx.x = slice, blockIdx.y = layer (4: the output layer)
__global__ __launch_bounds__(64) void k_nrc_weight_grad(const __bf16* __restrict__ X, const __bf16* __restrict__ G,
                                                        uint32_t B, uint32_t slice, float* __restrict__ partial) {
    if (blockIdx.y < 4)
        weight_grad_layer<4>((int)blockIdx.y, X, G, B, slice, partial);
    else
        weight_grad_layer<1>(4, X, G, B, slice, partial);
}

__device__ __forceinline__ void pack_weight(uint32_t p, float w, __bf16* __restrict__ pf, __bf16* __restrict__ pb);

// sum of the slice partials, Adam, and the packed BF16 copies for the next forward/backward
__global__ void k_nrc_adam(const float* __restrict__ partial, uint32_t n_slices, float inv_b, float* __restrict__ w,
                           float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2, float eps,
                           float bc1, float bc2, __bf16* __restrict__ pf, __bf16* __restrict__ pb) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNrcParams) return;
    float gsum = 0.0f;
    for (uint32_t s = 0; s < n_slices; ++s) gsum += partial[(size_t)s * kNrcParams + p];
    const float gr = gsum * inv_b;
    const float mm = b1 * m[p] + (1.0f - b1) * gr;
    const float vv = b2 * v[p] + (1.0f - b2) * gr * gr;
    m[p] = mm;
    v[p] = vv;
    const float wn = w[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    w[p] = wn;
    pack_weight(p, wn, pf, pb);
}

// the BF16 copies of parameter p = weight W_L[o][i] at its fragment positions (see the file
// comment) in the forward and backward layouts
__device__ __forceinline__ void pack_weight(uint32_t p, float w, __bf16* __restrict__ pf, __bf16* __restrict__ pb) {
    const int L = p < 4 * 4096 ? (int)(p / 4096) : 4;
    const int local = L < 4 ? (int)(p % 4096) : (int)(p - 4 * 4096);
    const int o = local / 64, i = local % 64;
    const __bf16 val = (__bf16)w;
    // forward: A[o][k] with k -> input i
    {
        const int t = i / 16, g = (i % 16) / 4, vv = i % 4, kc = t / 2, j = (t % 2) * 4 + vv;
        const int lane = g * 16 + o % 16, ot = o / 16;
        pf[L * 4096 + ((ot * 2 + kc) * 64 + lane) * 8 + j] = val;
    }
    // backward: A[i][k] with k -> output o (layers 2..5 only)
    if (L >= 1) {
        const int it = i / 16;
        if (L < 4) {
            const int t = o / 16, g = (o % 16) / 4, vv = o % 4, kc = t / 2, j = (t % 2) * 4 + vv;
            const int lane = g * 16 + i % 16;
            pb[(L - 1) * 4096 + ((it * 2 + kc) * 64 + lane) * 8 + j] = val;
        } else {
            const int lane = i % 16, j = o;  // o < 3: group 0, slot o
            pb[3 * 4096 + (it * 64 + lane) * 8 + j] = val;
        }
    }
}

__global__ void k_nrc_pack(const float* __restrict__ w, __bf16* __restrict__ pf, __bf16* __restrict__ pb) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kNrcParams) pack_weight(p, w[p], pf, pb);
}

// D = A * B through one matrix-core instruction with the fragment layout above (A 16x32, B 32x16,
// row-major FP32 in, rounded to BF16): the self-check of that layout
__global__ __launch_bounds__(64) void k_nrc_mfma_product(const float* __restrict__ a, const float* __restrict__ b,
                                                         float* __restrict__ d) {
    const int lane = threadIdx.x, g = lane >> 4, r = lane & 15;
    bf16x8 fa, fb;
    for (int j = 0; j < 8; ++j) {
        fa[j] = to_bf16(a[r * 32 + 8 * g + j]);
        fb[j] = to_bf16(b[(8 * g + j) * 16 + r]);
    }
    const f32x4 acc = mfma(fa, fb, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
    for (int v = 0; v < 4; ++v) d[(4 * g + v) * 16 + r] = acc[v];
}

}  // namespace cupbr

using namespace cupbr;

// ------------------------------------------------------------------ host side
static uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

namespace cupbr {

int nrc_init_weights(cupbr_nrc* c, uint32_t seed) {
    // He-normal hidden layers (std sqrt(2/64)), output layer std 0.01
    std::vector<float> w(kNrcParams);
    uint64_t s = 0x1234ABCDull ^ ((uint64_t)seed << 32);
    for (uint32_t p = 0; p < kNrcParams; p += 2) {
        const double u1 = ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double u2 = ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double r = std::sqrt(-2.0 * std::log(u1)), a = 2.0 * 3.14159265358979323846 * u2;
        const double n0 = r * std::cos(a), n1 = r * std::sin(a);
        for (int k = 0; k < 2 && p + k < kNrcParams; ++k) {
            const uint32_t q = p + k;
            const double sd = q < 4 * 4096 ? std::sqrt(2.0 / 64.0) : 0.01;
            w[q] = (float)((k ? n1 : n0) * sd);
        }
    }
    return nrc_set_weights(c, w.data());
}

int nrc_set_weights(cupbr_nrc* c, const float* host) {
    CUPBR_HIP_CHECK(hipMemcpy(c->w.ptr, host, kNrcParams * 4, hipMemcpyHostToDevice));
    CUPBR_HIP_CHECK(hipMemset(c->m.ptr, 0, kNrcParams * 4));
    CUPBR_HIP_CHECK(hipMemset(c->v.ptr, 0, kNrcParams * 4));
    c->step = 0;
    return nrc_pack(c, nullptr);
}

int nrc_pack(cupbr_nrc* c, hipStream_t stream) {
    hipLaunchKernelGGL(k_nrc_pack, dim3((kNrcParams + 255) / 256), dim3(256), 0, stream, c->w.as<float>(),
                       c->packed_fwd.as<__bf16>(), c->packed_bwd.as<__bf16>());
    CUPBR_HIP_CHECK(hipGetLastError());
    return CUPBR_OK;
}

int nrc_create(const float lo[3], const float hi[3], uint32_t seed, cupbr_nrc** out) {
    auto* c = new cupbr_nrc();
    for (int a = 0; a < 3; ++a) {
        c->box.lo[a] = lo[a];
        const float e = hi[a] - lo[a];
        c->box.inv_ext[a] = e > 0.0f ? 1.0f / e : 1.0f;
    }
    int rc;
    if ((rc = c->w.ensure(kNrcParams * 4)) || (rc = c->m.ensure(kNrcParams * 4)) || (rc = c->v.ensure(kNrcParams * 4)) ||
        (rc = c->packed_fwd.ensure(kFwdSize * 2)) || (rc = c->packed_bwd.ensure(kBwdSize * 2))) {
        delete c;
        return rc;
    }
    // zero fragments that no weight maps to (padded rows/columns)
    if (hipMemset(c->packed_fwd.ptr, 0, kFwdSize * 2) != hipSuccess ||
        hipMemset(c->packed_bwd.ptr, 0, kBwdSize * 2) != hipSuccess) {
        delete c;
        return set_error(CUPBR_ERR_HIP, "hipMemset failed");
    }
    if ((rc = nrc_init_weights(c, seed))) {
        delete c;
        return rc;
    }
    CUPBR_HIP_CHECK(hipDeviceSynchronize());
    *out = c;
    return CUPBR_OK;
}

int nrc_infer(cupbr_nrc* c, const float* rec, uint32_t n, const uint32_t* n_dev, uint32_t stride, float* out,
              bool accumulate, hipStream_t stream, const uint32_t* begin_dev) {
    if (n == 0) return CUPBR_OK;
    const uint32_t pairs = (n + 31) / 32;  // a wave takes two 16-query tiles at a time
    const int blocks = (int)std::max<uint32_t>(1, std::min<uint32_t>((pairs + 3) / 4, (uint32_t)device_cu_count() * 4));
    if (accumulate)
        hipLaunchKernelGGL(k_nrc_infer<true>, dim3(blocks), dim3(kNrcBlock), 0, stream, c->packed_fwd.as<bf16x8>(),
                           c->box, rec, n, n_dev, stride, out, begin_dev);
    else
        hipLaunchKernelGGL(k_nrc_infer<false>, dim3(blocks), dim3(kNrcBlock), 0, stream, c->packed_fwd.as<bf16x8>(),
                           c->box, rec, n, n_dev, stride, out, begin_dev);
    CUPBR_HIP_CHECK(hipGetLastError());
    return CUPBR_OK;
}

__global__ void k_nrc_zero(uint4* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

int nrc_train(cupbr_nrc* c, const float* rows, uint32_t n_rows, uint32_t batch, uint32_t steps, float lr,
              float* loss_dev, hipStream_t stream) {
    if (n_rows == 0 || steps == 0) return CUPBR_OK;
    const uint32_t slice = 512;
    const uint32_t n_slices = batch / slice;
    int rc;
    if ((rc = c->acts.ensure((size_t)2 * 5 * 64 * batch * 2)) ||
        (rc = c->partial.ensure((size_t)n_slices * kNrcParams * 4)))
        return rc;
    __bf16* X = c->acts.as<__bf16>();
    __bf16* G = X + (size_t)5 * 64 * batch;
    // rows 3..63 of the output-layer gradient are never written: clear them once per size (by a
    // kernel rather than a memset, so that a caller may capture the call in a graph)
    if (c->acts_cleared_for != batch) {
        const size_t n4 = (size_t)2 * 5 * 64 * batch * 2 / 16;
        hipLaunchKernelGGL(k_nrc_zero, dim3((unsigned)std::min<size_t>((n4 + 255) / 256, 4096)), dim3(256), 0, stream,
                           c->acts.as<uint4>(), n4);
        CUPBR_HIP_CHECK(hipGetLastError());
        c->acts_cleared_for = batch;
    }
    const int blocks = (int)std::max<uint32_t>(1, std::min<uint32_t>(batch / 64, (uint32_t)device_cu_count() * 2));
    for (uint32_t s = 0; s < steps; ++s) {
        ++c->step;
        const uint32_t seed = c->step * 0x9E3779B9u + 0x51ED270Bu;
        hipLaunchKernelGGL(k_nrc_forward_backward, dim3(blocks), dim3(kNrcBlock), 0, stream, c->packed_fwd.as<bf16x8>(),
                           c->packed_bwd.as<bf16x8>(), c->box, rows, n_rows, batch, seed, X, G,
                           loss_dev ? loss_dev + s : nullptr);
        CUPBR_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_nrc_weight_grad, dim3(n_slices, 5), dim3(64), 0, stream, X, G, batch, slice,
                           c->partial.as<float>());
        CUPBR_HIP_CHECK(hipGetLastError());
        const float b1 = c->beta1, b2 = c->beta2;
        const float bc1 = 1.0f - std::pow(b1, (float)c->step), bc2 = 1.0f - std::pow(b2, (float)c->step);
        hipLaunchKernelGGL(k_nrc_adam, dim3((kNrcParams + 255) / 256), dim3(256), 0, stream, c->partial.as<float>(),
                           n_slices, 1.0f / (float)batch, c->w.as<float>(), c->m.as<float>(), c->v.as<float>(), lr, b1,
                           b2, c->eps, bc1, bc2, c->packed_fwd.as<__bf16>(), c->packed_bwd.as<__bf16>());
        CUPBR_HIP_CHECK(hipGetLastError());
    }
    return CUPBR_OK;
}

int nrc_mfma_product(const float* a, const float* b, float* d) {
    DevBuf da, db, dd;
    int rc;
    if ((rc = da.ensure(16 * 32 * 4)) || (rc = db.ensure(32 * 16 * 4)) || (rc = dd.ensure(16 * 16 * 4))) return rc;
    CUPBR_HIP_CHECK(hipMemcpy(da.ptr, a, 16 * 32 * 4, hipMemcpyHostToDevice));
    CUPBR_HIP_CHECK(hipMemcpy(db.ptr, b, 32 * 16 * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_nrc_mfma_product, dim3(1), dim3(64), 0, nullptr, da.as<float>(), db.as<float>(), dd.as<float>());
    CUPBR_HIP_CHECK(hipGetLastError());
    CUPBR_HIP_CHECK(hipMemcpy(d, dd.ptr, 16 * 16 * 4, hipMemcpyDeviceToHost));
    return CUPBR_OK;
}

}  // namespace cupbr
