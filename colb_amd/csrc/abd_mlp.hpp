// abd_mlp.hpp — host object of the per-body MLP (abd_mlp.hip) and the device entry points the
// solver uses for inference on its own stream.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/abd_hip.h"
#include "../../include/abd_mlp.h"

struct AbdMlp {
    int L = 0;                      // layers
    std::vector<int> d, dp;         // widths and their multiples of 16
    hipStream_t st = nullptr;
    std::string err;
    std::vector<void*> allocs, act_allocs;
    std::vector<float*> W, b, gW, gb, mW, vW, mb, vb, maskW, maskb;  // FP32 masters, grads, Adam, masks
    std::vector<uint16_t*> Wb;      // BF16 copies read by the matrix products
    std::vector<uint16_t*> A;       // A[0] input, A[l] = bf16(relu(Z[l-1]))
    std::vector<float*> Z;          // pre-activations, Z[L-1] is the output
    std::vector<uint16_t*> G;       // dLoss/dZ[l] in BF16
    std::vector<float*> dA;         // dLoss/dA[l+1]
    float* T = nullptr;             // packed targets
    float* row_loss = nullptr;
    int* idx = nullptr;
    float* colw = nullptr;          // per-output loss weights
    float* loss_dev = nullptr;
    int rows_cap = 0;
    long adam_t = 0;
    int dev = -1;                   // device the model was created on
    void* scratch = nullptr;        // that device's address of the static scratch (abd_mlp_forward_static)
};

namespace abd {
using Mlp = ::AbdMlp;

// float -> bf16, round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

int mlp_reserve(Mlp* m, int rows);
void mlp_forward_dev(Mlp* m, int n, hipStream_t s);
// forward over packed rows in A[0] with caller-given activation buffers A[0..L], Z[0..L-1]
void mlp_forward_bufs(Mlp* m, int n, hipStream_t s, uint16_t* const* A, float* const* Z);
void mlp_backward_dev(Mlp* m, int n, float* loss_dev, hipStream_t s);
}  // namespace abd
