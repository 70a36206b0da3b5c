// The delayed-acceptance surrogate: a small MLP over residue windows,
// shared by the training kernels (mlp.hip) and the sampler's screening
// stage (surrogate.hpp).
//
// Energy model: E~(x) = scale * sum_r y(f_r(x)), one row per residue r.
// f_r has kIn = 128 columns: sin and cos of the (up to kSlots) variables of
// residues r-1, r, r+1 of the same chain (60 columns, zero where a residue
// or slot is absent), the kEmb-dimensional embeddings of those three
// residues' types (48 columns), then zeros.  y = w4 . silu(W3 silu(W2
// silu(W1^T f + b1) + b2) + b3) + b4 with kHid hidden units per layer.
// Matrix products take BF16 operands and accumulate in FP32; biases, the
// output layer and the sums over residues are FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace smmp {

constexpr int kMlpIn = 128;
constexpr int kMlpHid = 64;
constexpr int kMlpSlots = 10;   // variables per residue (at most)
constexpr int kMlpEmb = 16;
constexpr int kMlpTypes = 32;   // embedding rows; row 0 = no residue
constexpr int kMlpAngleCols = 3 * 2 * kMlpSlots;  // 60
constexpr int kMlpEmbCol = kMlpAngleCols;          // embeddings start here

// Offsets of the parameters in one FP32 vector (theta): W1 is [kIn][kHid]
// (input-major), W2 and W3 [kHid][kHid], embeddings [kTypes][kEmb].
constexpr int kThW1 = 0;
constexpr int kThB1 = kThW1 + kMlpIn * kMlpHid;
constexpr int kThW2 = kThB1 + kMlpHid;
constexpr int kThB2 = kThW2 + kMlpHid * kMlpHid;
constexpr int kThW3 = kThB2 + kMlpHid;
constexpr int kThB3 = kThW3 + kMlpHid * kMlpHid;
constexpr int kThW4 = kThB3 + kMlpHid;
constexpr int kThB4 = kThW4 + kMlpHid;
constexpr int kThEmb = kThB4 + 16;
constexpr int kThN = kThEmb + kMlpTypes * kMlpEmb;

// Residue windows of one system (device arrays, nres entries each, 0-based):
// first variable, number of variables, embedding row, chain; and per
// variable its residue.
struct MlpResidues {
  int32_t nres, nvt;
  const int16_t *var0, *nvar, *type, *chain;
  const int16_t* res_of_var;
};

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// Angle column c (< kMlpAngleCols) of the window of residue r: the value
// of variable vl[...] behind it, or 'none' (returns false).
__device__ __forceinline__ bool window_angle(const MlpResidues& R, int r, int c, int& var) {
  const int d = c / (2 * kMlpSlots) - 1, s = (c % (2 * kMlpSlots)) / 2;
  const int q = r + d;
  if (q < 0 || q >= R.nres || R.chain[q] != R.chain[r] || s >= R.nvar[q]) return false;
  var = R.var0[q] + s;
  return true;
}
// Embedding row of window position d (0..2) of residue r (0 = no residue).
__device__ __forceinline__ int window_type(const MlpResidues& R, int r, int d) {
  const int q = r + d - 1;
  if (q < 0 || q >= R.nres || R.chain[q] != R.chain[r]) return 0;
  return R.type[q];
}

}  // namespace smmp
