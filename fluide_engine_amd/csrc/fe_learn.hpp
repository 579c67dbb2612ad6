// fe_learn.hpp -- the pressure warm-start network (no counterpart in the
// reference, whose solvers start from zero or from the previous pressure).
//
// A 4-level convolutional encoder-decoder, 16-64 channels, 3x3x3 kernels with
// zero padding, channel-last BF16 activations, FP32 accumulation on MFMA:
//   X0 (16: b / rms(b), FLUID, AIR, zeros) -> E0 16 -> pool -> E1 32 -> pool
//   -> E2 48 -> pool -> E3 64; D2 = conv[up(E3), E2] 48, D1 = conv[up(D2), E1]
//   32, D0 = conv[up(D1), E0] 16; p = D0 . hw + hb (FP32).
// Trained on the device (Adam on FP32 master weights) on pairs of pressure
// systems and their solutions; at a solve it supplies x0 = alpha * p, alpha
// minimising ||b - alpha A p||, and CG runs to the usual tolerance from there.
#pragma once

#include <cstdint>
#include <mutex>
#include <vector>

#include "fe_internal.hpp"

namespace fe {

constexpr int kNetLayers = 7;  // enc0..enc3, dec2, dec1, dec0
constexpr int kNetIn = 16;     // input channels (3 used)
constexpr int kNetLayerIn[kNetLayers] = {16, 16, 32, 48, 112, 80, 48};
constexpr int kNetLayerOut[kNetLayers] = {16, 32, 48, 64, 48, 32, 16};

struct NetLayer {
    int cin = 0, cout = 0;
    float* w = nullptr;  // 27 * cin * cout, (t * cin + ci) * cout + co
    float* b = nullptr;
    float *gw = nullptr, *gb = nullptr;
    float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;  // Adam moments
    uint16_t* wp = nullptr;   // BF16, packed for the forward MFMA B operand
    uint16_t* wpt = nullptr;  // BF16, flipped and transposed, packed for the backward data pass
};

struct NetSample {
    int n[3] = {0, 0, 0};
    int64_t nfluid = 0;
    float* b = nullptr;       // right-hand side as built by the pressure solve
    uint8_t* flags = nullptr;
    float* target = nullptr;  // solution / rms(solution over FLUID cells)
    uint16_t* code = nullptr; // the system's row codes
    double rms_b = 1.0;
    double ih2[3] = {1, 1, 1};
    double energy = 0.0;      // b . solution over FLUID cells (the solution's squared A-norm)
};

// What one loss evaluation reads.
struct LossIn {
    const float* target = nullptr;  // MSE: the normalised solution
    const uint8_t* flags = nullptr;
    int64_t nfluid = 0;
    const uint16_t* code = nullptr;  // energy: the row codes, 1/h^2, b and b . solution
    double ih2[3] = {1, 1, 1};
    const float* b = nullptr;
    double energy = 0.0;
};

struct NetLevel {
    int n[3] = {0, 0, 0};
    int64_t V = 0;
};

}  // namespace fe

struct fe_net {
    fe_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    fe::NetLayer L[fe::kNetLayers];
    float *head = nullptr, *ghead = nullptr, *mhead = nullptr, *vhead = nullptr;  // 1x1 head: 16 weights + bias
    int64_t adam_t = 0;
    std::vector<fe::NetSample> samples;

    // activations and gradients for one grid size
    fe::NetLevel lv[4];
    uint16_t *X0 = nullptr, *E0 = nullptr, *P1 = nullptr, *E1 = nullptr, *P2 = nullptr, *E2 = nullptr, *P3 = nullptr,
             *E3 = nullptr, *C2 = nullptr, *D2 = nullptr, *C1 = nullptr, *D1 = nullptr, *C0 = nullptr, *D0 = nullptr;
    uint16_t* dpre = nullptr;  // BF16 masked output gradient of the layer being differentiated
    float *out = nullptr, *dout = nullptr, *aq = nullptr;
    float* D0f = nullptr;  // the last decoder level in FP32 (the head's input)
    float *gD0 = nullptr, *gC0 = nullptr, *gE0 = nullptr, *gD1 = nullptr, *gC1 = nullptr, *gE1 = nullptr,
          *gP1 = nullptr, *gD2 = nullptr, *gC2 = nullptr, *gE2 = nullptr, *gP2 = nullptr, *gE3 = nullptr,
          *gP3 = nullptr;
    float* wpart = nullptr;  // weight-gradient partials
    int64_t wpart_cap = 0;
    double* red = nullptr;   // reduction partials (FP64)
    int64_t red_cap = 0;
    float* losses = nullptr; // per training step
    int64_t losses_cap = 0;
    double* scal = nullptr;  // small device scalars
    std::mutex mu;           // scenes advanced on several host threads may share the network
    int scale_mode = 1;      // FE_NET_SCALE_ENERGY (1) or FE_NET_SCALE_RESIDUAL (0)
    int loss_mode = 1;       // FE_NET_LOSS_ENERGY (1) or FE_NET_LOSS_MSE (0)
};

namespace fe {

hipError_t net_alloc(fe_net* net, uint64_t seed);
void net_free(fe_net* net);
int64_t net_param_count();
hipError_t net_get_params(fe_net* net, float* host);
hipError_t net_set_params(fe_net* net, const float* host);
// Sizes of the activation buffers for an n^3 grid (each axis a multiple of 8).
hipError_t net_ensure(fe_net* net, const int n[3]);
// One forward + backward from device inputs; loss (mean over FLUID cells) into
// net->scal[0]; gradients into L[].gw / gb, ghw / ghb.
hipError_t net_forward(fe_net* net, const float* b, const uint8_t* flags, double rms_b);
hipError_t net_backward(fe_net* net, const LossIn& in, float* loss_slot);
// row codes of the pressure system from markers (for losses on host inputs)
hipError_t net_codes(fe_net* net, const int n[3], const uint8_t* flags, uint16_t* code);
hipError_t net_adam(fe_net* net, float lr);
hipError_t net_add_sample(fe_net* net, fe_scene* s);
hipError_t net_train(fe_net* net, int steps, float lr, uint64_t seed);
// x0 = alpha * prediction in the scene's pressure (FP64), from its b and flags
hipError_t net_warm_start(fe_net* net, fe_scene* s);
hipError_t net_fluid_count(fe_net* net, const uint8_t* flags, int64_t V, int64_t* count);

}  // namespace fe
