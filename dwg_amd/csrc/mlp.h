// Learned normal initialisation: a 48→128→128→3 MLP over k-NN patches.
#pragma once
#include "common.h"

namespace dwg {

constexpr int kPatchK = 16;            // neighbours per patch (the point itself excluded)
constexpr int kIn = 3 * kPatchK;       // 48
constexpr int kHid = 128;
constexpr int kOut = 3;
// parameter offsets in the flat FP32 vector
constexpr int64_t kOffW1 = 0;
constexpr int64_t kOffB1 = kOffW1 + int64_t(kHid) * kIn;
constexpr int64_t kOffW2 = kOffB1 + kHid;
constexpr int64_t kOffB2 = kOffW2 + int64_t(kHid) * kHid;
constexpr int64_t kOffW3 = kOffB2 + kHid;
constexpr int64_t kOffB3 = kOffW3 + int64_t(kOut) * kHid;
constexpr int64_t kParams = kOffB3 + kOut;  // 23,171
// points per training cloud: larger clouds take patches from a subsample of
// this size
constexpr int64_t kTrainPoints = 50000;

}  // namespace dwg

struct dwg_mlp {
    float* params = nullptr;     // FP32 master weights (kParams)
    uint16_t* bf16 = nullptr;    // BF16 copy used by the kernels (kParams)
    float* grad = nullptr;       // FP32 gradient (kParams)
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    int64_t step = 0;
    ~dwg_mlp();
};
