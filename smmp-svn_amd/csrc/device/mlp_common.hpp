// The delayed-acceptance surrogate: an MLP over the whole chain, shared by
// the training kernels (mlp.hip) and the sampler's screening stage
// (surrogate.hpp).
//
// Energy model: E~(x) = scale * y(f(x)).  f(x) has D = nres * kMlpResCols
// columns, kMlpResCols = 36 per residue r (all chains in order): sin and cos
// of the residue's (up to kMlpSlots = 10) variables, slot s at columns
// 36 r + 2 s and 36 r + 2 s + 1 (zero where the residue has no slot s), then
// the kMlpEmb = 16 dimensional embedding of the residue's type.  y = w4 .
// silu(W3 silu(W2 silu(W1^T f + b1) + b2) + b3) + b4 with kMlpHid = 256
// units per hidden layer.  Matrix products take BF16 operands (f rounded as
// bf16((float)sin a)) and accumulate in FP32; biases and the output layer
// are FP32.  The parameter count depends on the system (W1 has D rows): for
// 36 residues D = 1296 and the network has 464,400 parameters.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace smmp {

constexpr int kMlpHid = 256;
constexpr int kMlpSlots = 10;   // variables per residue (at most)
constexpr int kMlpEmb = 16;
constexpr int kMlpTypes = 32;   // embedding rows (1..20 the standard residues, 21 any other)
constexpr int kMlpResCols = 2 * kMlpSlots + kMlpEmb;  // 36
constexpr int kMlpEmbCol = 2 * kMlpSlots;             // embedding columns of a residue start here

// Offsets of the parameters in one FP32 vector (theta) for nres residues:
// W1 [D][H] (input-major), b1, W2 [H][H], b2, W3 [H][H], b3, w4, b4 (padded
// to 16), embeddings [kMlpTypes][kMlpEmb].
struct MlpLayout {
  int nres, D;
  int w1, b1, w2, b2, w3, b3, w4, b4, emb, n;
};
__host__ __device__ inline MlpLayout mlp_layout(int nres) {
  MlpLayout L;
  L.nres = nres;
  L.D = nres * kMlpResCols;
  L.w1 = 0;
  L.b1 = L.w1 + L.D * kMlpHid;
  L.w2 = L.b1 + kMlpHid;
  L.b2 = L.w2 + kMlpHid * kMlpHid;
  L.w3 = L.b2 + kMlpHid;
  L.b3 = L.w3 + kMlpHid * kMlpHid;
  L.w4 = L.b3 + kMlpHid;
  L.b4 = L.w4 + kMlpHid;
  L.emb = L.b4 + 16;
  L.n = L.emb + kMlpTypes * kMlpEmb;
  return L;
}

// Residues of one system (device arrays, 0-based): first variable, number of
// variables, embedding row, chain of each residue; residue of each variable.
struct MlpResidues {
  int32_t nres, nvt;
  const int16_t *var0, *nvar, *type, *chain;
  const int16_t* res_of_var;
};

// What the sampler's screen reads: the residue tables, the layout, the FP32
// parameters (biases, w4, b4, embeddings) and BF16 weights, W1 [D][H]
// (input-major) and W2^T, W3^T [out][in].
struct MlpScreen {
  MlpResidues res;
  MlpLayout lo;
  const float* theta;
  const __bf16* w1;
  const __bf16* w2t;
  const __bf16* w3t;
};

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// The angle feature of variable value a (sin for odd = 0, cos for odd = 1),
// as the training kernels round it: bf16 of the FP32 value.
__device__ __forceinline__ __bf16 mlp_angle_feature(double a, int odd) {
  return (__bf16)(float)(odd ? cos(a) : sin(a));
}

}  // namespace smmp
