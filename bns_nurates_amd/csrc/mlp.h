// mlp.h -- shapes and parameter layout of the surrogate MLP (one ensemble
// member): 15 features (as 30 hi/lo BF16 inputs, padded to 32) -> 4 x 256
// (SiLU) -> 15 outputs (padded to 16).  Parameters of a member are one FP32 blob:
// W_0, b_0, W_1, b_1, ..., W_l row-major [K_l][N_l].
#ifndef BNS_NURATES_AMD_MLP_H_
#define BNS_NURATES_AMD_MLP_H_

namespace bns {
namespace mlp {

constexpr int kFeat = 15;     // input features
constexpr int kOut = 15;      // outputs
// Each standardised feature x enters the first layer as two BF16 inputs, x_hi
// = bf16(x) in column i and x_lo = bf16(x - x_hi) in column kFeat + i, with
// the same weights (the rows kFeat .. 2 kFeat - 1 of W_0 mirror rows
// 0 .. kFeat - 1 in the BF16 copies; the FP32 master rows stay zero): the
// product W_0 x then carries ~16 significant bits of x instead of 8.
constexpr int kK0 = 32;       // padded input width (one MFMA K step)
constexpr int kNL = 16;       // padded output width
constexpr int kHidden = 256;  // hidden width
constexpr int kLayers = 5;
constexpr int kMaxMembers = 8;
constexpr int kLayerK[kLayers] = {kK0, kHidden, kHidden, kHidden, kHidden};
constexpr int kLayerN[kLayers] = {kHidden, kHidden, kHidden, kHidden, kNL};

constexpr int w_off(int l) { return l == 0 ? 0 : w_off(l - 1) + kLayerK[l - 1] * kLayerN[l - 1] + kLayerN[l - 1]; }
constexpr int b_off(int l) { return w_off(l) + kLayerK[l] * kLayerN[l]; }
constexpr int packed_off(int l) { return l == 0 ? 0 : packed_off(l - 1) + kLayerK[l - 1] * kLayerN[l - 1]; }
constexpr int kWOff[kLayers] = {w_off(0), w_off(1), w_off(2), w_off(3), w_off(4)};
constexpr int kBOff[kLayers] = {b_off(0), b_off(1), b_off(2), b_off(3), b_off(4)};
constexpr int kPOff[kLayers] = {packed_off(0), packed_off(1), packed_off(2), packed_off(3),
                                packed_off(4)};
constexpr int kParams = b_off(kLayers - 1) + kLayerN[kLayers - 1];  // floats per member
constexpr int kPacked = packed_off(kLayers - 1) + kLayerK[kLayers - 1] * kLayerN[kLayers - 1];

static_assert(kParams == 209936, "parameter layout");

}  // namespace mlp
}  // namespace bns

#endif  // BNS_NURATES_AMD_MLP_H_
