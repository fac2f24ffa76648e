// vc1me_nn_internal.h — launchers of the proposal-network kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "vc1me_internal.h"

namespace pnet {

// the jobs of a step that form one batch of the network (by value)
struct JobIndex {
    uint8_t k[256];
};

constexpr int kChannels = 32;
constexpr int kLayers = 4;            // 3x3 convolutions, dilations 1, 2, 4, 8
constexpr int kOut = 5;               // two (dx, dy) candidates in 64-px units and a confidence logit
// parameters: per conv layer 32 x 9 x cin weights then 32 biases; then the
// head's 5 x 32 weights and 5 biases
constexpr int kW1 = 32 * 9 * 2, kW = 32 * 9 * 32;
constexpr int kParams = (kW1 + 32) + 3 * (kW + 32) + 5 * 32 + 5;
inline int layer_cin(int l) { return l == 0 ? 2 : 32; }
inline int layer_dil(int l) { return 1 << l; }
inline int layer_offset(int l) { return l == 0 ? 0 : (kW1 + 32) + (l - 1) * (kW + 32); }   // weights, then biases
inline int head_offset() { return (kW1 + 32) + 3 * (kW + 32); }

hipError_t launch_downsample(const uint8_t* cur, const uint8_t* ref, int W, int H, int h4, int w4, const int32_t* gxy,
                             uint16_t* x0, hipStream_t s);
hipError_t launch_downsample_jobs(const vc1me::JobDesc* jobs, const JobIndex& idx, int nb, int h4, int w4, uint16_t* x0,
                                  hipStream_t s);
hipError_t launch_to_bf16(const float* in, uint16_t* out, int n, hipStream_t s);
hipError_t launch_conv_fwd(const uint16_t* x, const uint16_t* wt, const float* bias, uint16_t* out, int nb, int h, int w,
                           int cin, int dil, hipStream_t s);
hipError_t launch_conv_bwd_data(const uint16_t* dz, const uint16_t* wtt, const uint16_t* xact, uint16_t* dx, int nb,
                                int h, int w, int dil, hipStream_t s);
hipError_t launch_conv_bwd_weight(const uint16_t* dz, const uint16_t* x, float* dw, float* dbias, int nb, int h, int w,
                                  int cin, int dil, hipStream_t s);
hipError_t launch_pool_head(const uint16_t* a, const float* head, float* pooled, float* out, int nb, int h, int w,
                            int mbh, int mbw, hipStream_t s);
hipError_t launch_pool_head_bwd(const float* dout, const float* pooled, const float* head, const uint16_t* a,
                                float* dhead, uint16_t* da, int nb, int h, int w, int mbh, int mbw, hipStream_t s);
hipError_t launch_loss(const float* out, const float* tgt, const uint8_t* valid, float* dout, float* loss, int n,
                       float inv_count, hipStream_t s);
hipError_t launch_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float c1,
                       float c2, float eps, hipStream_t s);
hipError_t launch_pack_weights(const float* wf, uint16_t* wt, uint16_t* wtt, int cin, hipStream_t s);

}  // namespace pnet
