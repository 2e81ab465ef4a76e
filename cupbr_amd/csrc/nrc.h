// Neural radiance cache handle (nrc.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/cupbr.h"
#include "device.h"

namespace cupbr {

constexpr uint32_t kNrcParams = 4 * 64 * 64 + 3 * 64;  // W1..W4 (64x64), W5 (3x64), no biases
constexpr uint32_t kNrcInputs = 16;                     // raw inputs per query
constexpr uint32_t kNrcRowStride = 20;                  // training row: 16 inputs, 3 targets, 1 pad

struct NrcBox {
    float lo[3];
    float inv_ext[3];
};

}  // namespace cupbr

struct cupbr_nrc {
    cupbr::NrcBox box{};
    cupbr::DevBuf w, m, v;                   // FP32 weights and Adam moments
    cupbr::DevBuf packed_fwd, packed_bwd;    // BF16 matrix-core fragments of W and W^T
    cupbr::DevBuf acts, partial;             // training scratch: activations, per-slice weight gradients
    uint32_t acts_cleared_for = 0;
    uint32_t step = 0;
    float beta1 = 0.9f, beta2 = 0.99f, eps = 1e-10f;
};

namespace cupbr {

int nrc_create(const float lo[3], const float hi[3], uint32_t seed, cupbr_nrc** out);
int nrc_set_weights(cupbr_nrc* c, const float* host);
int nrc_init_weights(cupbr_nrc* c, uint32_t seed);
int nrc_pack(cupbr_nrc* c, hipStream_t stream);
// predictions for n records of `stride` floats (inputs first); with n_dev, min(n, *n_dev) records
// (a count on the device), and with begin_dev only the records from *begin_dev on; accumulate:
// out[4 path + c] += record[16 + c] * max(pred_c, 0), path = bits of record[19]; else
// out[3 q + c] = pred_c
int nrc_infer(cupbr_nrc* c, const float* rec, uint32_t n, const uint32_t* n_dev, uint32_t stride, float* out,
              bool accumulate, hipStream_t stream, const uint32_t* begin_dev = nullptr);
// D = A B (16x32 by 32x16) through one matrix-core instruction with the cache's fragment layout
int nrc_mfma_product(const float* a, const float* b, float* d);
// `steps` Adam steps on batches of `batch` rows drawn from rows[0, n_rows) (stride kNrcRowStride)
int nrc_train(cupbr_nrc* c, const float* rows, uint32_t n_rows, uint32_t batch, uint32_t steps, float lr,
              float* loss_dev, hipStream_t stream);

}  // namespace cupbr
