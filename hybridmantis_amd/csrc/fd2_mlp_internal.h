// fd2_mlp_internal.h -- shapes and parameter layout of the surrogate MLP.
#pragma once
#include <stddef.h>

namespace fd2mlp {

constexpr int kFeat = 5;        // depth/H, col_dx/R, H/500, R/5, top reflectivity
constexpr int kInPad = 32;      // input features padded for the matrix cores
constexpr int kIn = kInPad;
constexpr int kHidden = 256;
constexpr int kOut = 65;        // detection logit + 64 radial logits
constexpr int kOutPad = 80;
constexpr int kTarget = 65;     // detected fraction + radial distribution

// FP32 parameter vector: W0 [32x256], b0, W1 [256x256], b1, W2, b2,
// W3 [256x80], b3 [80]
constexpr size_t kWOff[4] = {0, 8448, 74240, 140032};
constexpr size_t kBOff[4] = {8192, 73984, 139776, 160512};
constexpr size_t kParams = 160592;

}  // namespace fd2mlp
