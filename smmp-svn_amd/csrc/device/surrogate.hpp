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

// One 256 -> 256 SiLU layer on the single row 'in' (LDS): wave w computes
// outputs 32w .. 32w+31 (two 16-column MFMA blocks over 8 k-steps of 32),
// with W^T [out][in] from global memory.  Writes bf16(silu(z + b)) to out,
// or (last layer) returns this wave's partial output sum w4 . h (lane 0).
template <bool LAST>
__device__ __forceinline__ float da_layer(const __bf16* in, const __bf16* __restrict__ wt, const float* __restrict__ b,
                                          const float* __restrict__ w4, __bf16* out) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x & 63;
  const int kq = 8 * (lane >> 4);
  float part = 0.0f;
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    const int n = 16 * (2 * wave + half) + (lane & 15);
    da_bf16x8 bw[8];
#pragma unroll
    for (int kk = 0; kk < 8; ++kk)  // every load issued before the first product
      bw[kk] = *reinterpret_cast<const da_bf16x8*>(&wt[(size_t)n * kMlpHid + 32 * kk + kq]);
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
    // row 0 of the result sits in lanes 0..15, element 0
    if (lane < 16) {
      const __bf16 h = (__bf16)silu(acc[0] + b[n]);
      if (LAST)
        part += (float)h * w4[n];
      else
        out[n] = h;
    }
  }
  if (LAST) {
    for (int o = 8; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);  // lanes 0..15
  }
  return part;
}

// Output of the network for the layer-1 sums z (LDS, kMlpHid) into D.y[slot]
// (all threads; barriers inside; D.y[slot] is valid on return).
__device__ __attribute__((noinline)) void da_output(size_t off, int nres, int nvt, MlpScreen S, int use_trial,
                                                    int slot) {
  extern __shared__ double smem[];
  const DaLds D = carve_da_lds(reinterpret_cast<char*>(smem) + off, nres, nvt);
  const int64_t* z = use_trial ? D.z1t : D.z1;
  const float* th = S.theta;
  const int t = threadIdx.x;
  if (t < kMlpHid) D.h1[t] = (__bf16)silu((float)((double)z[t] * kDaFixInv) + th[S.lo.b1 + t]);
  __syncthreads();
  da_layer<false>(D.h1, S.w2t, th + S.lo.b2, nullptr, D.h2);
  __syncthreads();
  const float p = da_layer<true>(D.h2, S.w3t, th + S.lo.b3, th + S.lo.w4, nullptr);
  if ((t & 63) == 0) D.red[t / 64] = p;
  __syncthreads();
  if (t == 0) {
    float y = th[S.lo.b4];
    for (int w = 0; w < 8; ++w) y += D.red[w];
    D.y[slot] = y;
  }
  __syncthreads();
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
  da_output(off, nres, nvt, S, 0, 0);
}

// Layer-1 sums of the trial that sets variable jv from vold to vnew (z1t;
// threads < kMlpHid), then its output into D.y[1].  All threads.
__device__ __attribute__((noinline)) void da_trial(size_t off, int nvt, MlpScreen S, int jv, double vold,
                                                   double vnew) {
  extern __shared__ double smem[];
  const int nres = S.res.nres;
  const DaLds D = carve_da_lds(reinterpret_cast<char*>(smem) + off, nres, nvt);
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
  da_output(off, nres, nvt, S, 1, 1);
}

// The trial became the state (threads < kMlpHid; thread 0 the output).
__device__ __forceinline__ void da_commit(const DaLds& D) {
  if (threadIdx.x < kMlpHid) D.z1[threadIdx.x] = D.z1t[threadIdx.x];
  if (threadIdx.x == 0) D.y[0] = D.y[1];
}

}  // namespace smmp
