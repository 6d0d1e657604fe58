// mlp.h -- device state of the Fourier-feature MLP control variate (see mlp_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace zg {

constexpr int kMaxHidden = 8;

struct MlpShape {
    int inDim;     // 2 or 3
    int nFreq;     // F: Fourier features = 2F
    int hidden;    // units per hidden layer
    int nHidden;   // hidden layers (SiLU)
};

// Parameters live in one fp32 vector theta: for l = 0..nHidden (the last is the output layer)
// W_l (out x in, row-major) at wOffset[l], then b_l at bOffset[l].  wBf mirrors theta in bf16
// (only the hidden weight matrices are read from it); bfIndex[i] = i for those entries, -1 else.
struct MlpDevice {
    MlpShape shape;
    int64_t nParams;
    int64_t wOffset[kMaxHidden + 1], bOffset[kMaxHidden + 1];
    float* theta;
    float* grad;
    float* adamM;
    float* adamV;
    uint16_t* wBf;
    int64_t* bfIndex;
    float* freq;          // F x inDim
    int64_t maxRows;
    uint16_t* phi;        // maxRows x 2F
    float* dphi;          // maxRows x 2F
    float* pre[kMaxHidden];       // maxRows x hidden, pre-activations
    uint16_t* act[kMaxHidden];    // maxRows x hidden, activations (bf16)
    uint16_t* dz[2];              // maxRows x hidden, ping-pong gradients (bf16)
    float* loss;                  // 1
    int64_t step;
};

hipError_t mlpZero(float* p, int64_t n, hipStream_t s);
hipError_t mlpCast(MlpDevice& m, hipStream_t s);
hipError_t mlpForward(MlpDevice& m, const float* x, int n, float* out, hipStream_t s);
hipError_t mlpBackward(MlpDevice& m, const float* y, int n, int mode, float* loss, float* out, hipStream_t s);
hipError_t mlpInputGrad(MlpDevice& m, const float* x, int n, float* grad, hipStream_t s);
hipError_t mlpAdam(MlpDevice& m, float lr, hipStream_t s);

}  // namespace zg
