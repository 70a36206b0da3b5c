// First stage of delayed-acceptance Metropolis inside the sweep kernel: the
// whole-chain surrogate network (mlp_common.hpp) evaluated on each torsion
// trial.
//
// A torsion trial changes two inputs of the network (sin and cos of one
// variable), so the first layer's sums change by two rank-1 terms.  They are
// kept as integers in units of 2^-40: every product bf16(W1) * bf16(x) is
// exact in FP64 and is rounded to that grid once, so the sums are an exact
// function of the current variables (adding and removing a term restores
// them bit for bit).  The network's output is therefore a function of the
// state alone, which is what keeps delayed acceptance exact (Christen and
// Fox, 2005: any state function may screen).  Layers 2 and 3 (256 x 256)
// run on v_mfma_f32_16x16x32_bf16 with the one real row in row 0 and the
// BF16 weights streamed from global memory (L2-resident, 128 KiB each).
#pragma once
#include "mlp_common.hpp"

namespace smmp {

typedef __bf16 da_bf16x8 __attribute__((ext_vector_type(8)));
typedef float da_f32x4 __attribute__((ext_vector_type(4)));

constexpr double kDaFixScale = 1099511627776.0;      // 2^40: layer-1 sums in units of 2^-40
constexpr double kDaFixInv = 1.0 / 1099511627776.0;  // exact (a power of two)

struct DaLds {
  int64_t* z1;        // layer-1 sums of the current state (kMlpHid)
  int64_t* z1t;       // of the trial state
  __bf16* h1;         // activations of the state being evaluated (kMlpHid)
  __bf16* h2;
  float* red;         // per-wave partial sums of the output (8)
  float* y;           // [0] output at the current state, [1] at the trial
  double* stats;      // screened, rejected by the network, judged exactly, accepted
  int16_t* var0;      // first variable of each residue
  int16_t* res_of_var;
  __bf16* x;          // features of the start state (D), set-up only
};

__host__ __device__ inline size_t da_lds_bytes(int nres, int nvt) {
  const size_t D = (size_t)nres * kMlpResCols;
  size_t b = 2 * 8 * (size_t)kMlpHid;       // z1, z1t
  b += 2 * 2 * (size_t)kMlpHid;             // h1, h2
  b += 4 * 8 + 4 * 4;                       // red, y
  b += 8 * 4;                               // stats
  b += 2 * ((size_t)nres + nvt);            // var0, res_of_var
  b = (b + 15) & ~size_t(15);
  b += 2 * D;                               // x
  return b + 64;
}

__device__ inline DaLds carve_da_lds(char* p, int nres, int nvt) {
  DaLds D;
  auto take = [&](size_t bytes) {
    char* q = p;
    p += bytes;
    return q;
  };
  D.z1 = reinterpret_cast<int64_t*>(take(8 * (size_t)kMlpHid));
  D.z1t = reinterpret_cast<int64_t*>(take(8 * (size_t)kMlpHid));
  D.stats = reinterpret_cast<double*>(take(8 * 4));
  D.h1 = reinterpret_cast<__bf16*>(take(2 * (size_t)kMlpHid));
  D.h2 = reinterpret_cast<__bf16*>(take(2 * (size_t)kMlpHid));
  D.red = reinterpret_cast<float*>(take(4 * 8));
  D.y = reinterpret_cast<float*>(take(4 * 4));
  D.var0 = reinterpret_cast<int16_t*>(take(2 * (size_t)nres));
  D.res_of_var = reinterpret_cast<int16_t*>(take(2 * (size_t)nvt));
  p = reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(p) + 15) & ~uintptr_t(15));
  D.x = reinterpret_cast<__bf16*>(p);
  return D;
}

// One term of a layer-1 sum on the 2^-40 grid: bf16 weight times bf16
// feature (exact in FP64), rounded to the nearest grid point.
__device__ __forceinline__ int64_t da_term(__bf16 w, __bf16 x) {
  return __double2ll_rn((double)(float)w * (double)(float)x * kDaFixScale);
}

// The B fragments of one 16-column block of W^T [out][in] for this lane:
// 8 k-steps of 32 (input rows 8 (lane / 16) .. +7 of each step) of output
// column n.  Issued as one batch of 16-byte loads.
__device__ __forceinline__ void da_load_block(const __bf16* __restrict__ wt, int n, int kq, da_bf16x8 (&bw)[8]) {
#pragma unroll
  for (int kk = 0; kk < 8; ++kk) bw[kk] = *reinterpret_cast<const da_bf16x8*>(&wt[(size_t)n * kMlpHid + 32 * kk + kq]);
}

// Row 0 of in (LDS, kMlpHid BF16) times the loaded block: the result for
// output column n in lanes 0..15 (element 0).
__device__ __forceinline__ float da_mfma_block(const __bf16* in, int lane, int kq, const da_bf16x8 (&bw)[8]) {
  da_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < 8; ++kk) {
    da_bf16x8 a;
    if ((lane & 15) == 0) {
      a = *reinterpret_cast<const da_bf16x8*>(&in[32 * kk + kq]);
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) a[j] = (__bf16)0.0f;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[kk], acc, 0, 0, 0);
  }
  return acc[0];
}

// The network's output for the layer-1 sums z (LDS): h1 from z (threads <
// kMlpHid), then the two 256 x 256 layers, wave w taking outputs 32 w ..
// 32 w + 31 as two 16-column MFMA blocks over 8 k-steps of 32 with the
// weight loads issued before each barrier.  All threads; on return thread
// 0 holds the output (the other threads 0.0).
__device__ __forceinline__ float da_forward(const DaLds& D, const MlpScreen& S, const int64_t* z) {
  const int t = threadIdx.x, wave = t / 64, lane = t & 63, kq = 8 * (lane >> 4);
  const float* th = S.theta;
  if (t < kMlpHid) D.h1[t] = (__bf16)silu((float)((double)z[t] * kDaFixInv) + th[S.lo.b1 + t]);
  da_bf16x8 bw[8];
  const int n0 = 32 * wave + (lane & 15), n1 = n0 + 16;
  da_load_block(S.w2t, n0, kq, bw);
  __syncthreads();  // h1 complete
  float v = da_mfma_block(D.h1, lane, kq, bw);
  if (lane < 16) D.h2[n0] = (__bf16)silu(v + th[S.lo.b2 + n0]);
  da_load_block(S.w2t, n1, kq, bw);
  v = da_mfma_block(D.h1, lane, kq, bw);
  if (lane < 16) D.h2[n1] = (__bf16)silu(v + th[S.lo.b2 + n1]);
  da_load_block(S.w3t, n0, kq, bw);
  __syncthreads();  // h2 complete
  float part = 0.0f;
  v = da_mfma_block(D.h2, lane, kq, bw);
  if (lane < 16) part += (float)(__bf16)silu(v + th[S.lo.b3 + n0]) * th[S.lo.w4 + n0];
  da_load_block(S.w3t, n1, kq, bw);
  v = da_mfma_block(D.h2, lane, kq, bw);
  if (lane < 16) part += (float)(__bf16)silu(v + th[S.lo.b3 + n1]) * th[S.lo.w4 + n1];
  for (int o = 8; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);  // lanes 0..15 of the wave
  if (lane == 0) D.red[wave] = part;
  __syncthreads();
  float y = 0.0f;
  if (t == 0) {
    y = th[S.lo.b4];
    for (int w = 0; w < 8; ++w) y += D.red[w];
  }
  return y;
}

// Tables, features of the start state vl and its layer-1 sums, and the
// output there (all threads; a barrier precedes the use).
__device__ __attribute__((noinline)) void da_setup(size_t off, int nvt, MlpScreen S, const double* vl) {
  extern __shared__ double smem[];
  const int nres = S.res.nres, Dn = S.lo.D;
  const DaLds D = carve_da_lds(reinterpret_cast<char*>(smem) + off, nres, nvt);
  const float* th = S.theta;
  for (int r = threadIdx.x; r < nres; r += blockDim.x) D.var0[r] = S.res.var0[r];
  for (int v = threadIdx.x; v < nvt; v += blockDim.x) D.res_of_var[v] = S.res.res_of_var[v];
  for (int c = threadIdx.x; c < Dn; c += blockDim.x) {
    const int r = c / kMlpResCols, k = c - r * kMlpResCols;
    __bf16 x = (__bf16)0.0f;
    if (k >= kMlpEmbCol)
      x = (__bf16)th[S.lo.emb + S.res.type[r] * kMlpEmb + (k - kMlpEmbCol)];
    else if ((k >> 1) < S.res.nvar[r])
      x = mlp_angle_feature(vl[S.res.var0[r] + (k >> 1)], k & 1);
    D.x[c] = x;
  }
  if (threadIdx.x < 4) D.stats[threadIdx.x] = 0.0;
  __syncthreads();
  // two threads per output, halves of the columns, added as integers (exact)
  const int n = threadIdx.x & (kMlpHid - 1), half = threadIdx.x / kMlpHid;
  int64_t acc = 0;
  if (half < 2) {
    const int c0 = half * (Dn / 2), c1 = half ? Dn : Dn / 2;
    for (int c = c0; c < c1; ++c) acc += da_term(S.w1[(size_t)c * kMlpHid + n], D.x[c]);
    (half ? D.z1t : D.z1)[n] = acc;
  }
  __syncthreads();
  if (threadIdx.x < kMlpHid) D.z1[threadIdx.x] += D.z1t[threadIdx.x];
  __syncthreads();
  const float y = da_forward(D, S, D.z1);
  if (threadIdx.x == 0) D.y[0] = y;
  __syncthreads();
}

// Layer-1 sums of the trial that sets variable jv from vold to vnew (z1t;
// threads < kMlpHid) and its output, which thread 0 returns (and keeps in
// D.y[1]).  All threads.
__device__ __forceinline__ float da_trial(const DaLds& D, const MlpScreen& S, int jv, double vold, double vnew) {
  const int t = threadIdx.x;
  if (t < kMlpHid) {
    int64_t z = D.z1[t];
    const int r = D.res_of_var[jv];
    if (r >= 0) {
      const int c = r * kMlpResCols + 2 * (jv - D.var0[r]);  // sin column; cos follows
      const __bf16 ws = S.w1[(size_t)c * kMlpHid + t], wc = S.w1[(size_t)(c + 1) * kMlpHid + t];
      z += da_term(ws, mlp_angle_feature(vnew, 0)) - da_term(ws, mlp_angle_feature(vold, 0));
      z += da_term(wc, mlp_angle_feature(vnew, 1)) - da_term(wc, mlp_angle_feature(vold, 1));
    }
    D.z1t[t] = z;
  }
  __syncthreads();
  const float y = da_forward(D, S, D.z1t);
  if (t == 0) D.y[1] = y;
  return y;
}

// The trial became the state (threads < kMlpHid; thread 0 the output).
__device__ __forceinline__ void da_commit(const DaLds& D) {
  if (threadIdx.x < kMlpHid) D.z1[threadIdx.x] = D.z1t[threadIdx.x];
  if (threadIdx.x == 0) D.y[0] = D.y[1];
}

}  // namespace smmp
