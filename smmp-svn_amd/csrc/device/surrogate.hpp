// First stage of delayed-acceptance Metropolis inside the sweep kernel: the
// surrogate network (mlp_common.hpp) evaluated on the residue windows a
// torsion trial changes, with its weights staged in LDS as BF16 and the
// products on v_mfma_f32_16x16x32_bf16.
#pragma once
#include "mlp_common.hpp"

namespace smmp {

typedef __bf16 da_bf16x8 __attribute__((ext_vector_type(8)));
typedef float da_f32x4 __attribute__((ext_vector_type(4)));

constexpr int kDaRows = 3;               // windows a torsion changes: residues r-1, r, r+1
constexpr int kDaLdIn = kMlpIn + 8;      // LDS row strides in bf16 (padding keeps 16-byte reads aligned)
constexpr int kDaLdHid = kMlpHid + 8;

struct DaLds {
  __bf16 *w1t, *w2t, *w3t;  // [out][in] (transposed: 8 consecutive inputs are contiguous)
  __bf16* emb;               // [kMlpTypes][kMlpEmb]
  __bf16 *x, *ha, *hb;       // kDaRows rows of input / hidden activations
  float *b1, *b2, *b3, *w4, *b4;
  float* ynew;               // outputs of the rows just evaluated
  float* ycache;             // outputs of every residue's window at the current state
  int16_t *var0, *nvar, *type, *chain, *res_of_var;
  int32_t* rows;             // the rows just evaluated and their number ([kDaRows] = n)
  double* stats;             // screened, rejected by the network, judged exactly, accepted
};

__host__ __device__ inline size_t da_lds_bytes(int nres, int nvt) {
  size_t b = 0;
  b += 2 * (size_t)kMlpHid * kDaLdIn;            // w1t
  b += 2 * 2 * (size_t)kMlpHid * kDaLdHid;       // w2t, w3t
  b += 2 * (size_t)kMlpTypes * kMlpEmb;          // emb
  b += 2 * (size_t)kDaRows * kDaLdIn;            // x
  b += 2 * 2 * (size_t)kDaRows * kDaLdHid;       // ha, hb
  b = (b + 15) & ~size_t(15);
  b += 4 * (4 * (size_t)kMlpHid + 4);            // b1, b2, b3, w4, b4 (+pad)
  b += 4 * 4;                                    // ynew
  b += 8 * 4;                                    // stats
  b += 4 * (kDaRows + 1);                        // rows
  b += 4 * (size_t)nres;                         // ycache
  b += 2 * (4 * (size_t)nres + nvt);             // residue tables
  return b + 64;
}

__device__ inline DaLds carve_da_lds(char* p, int nres, int nvt) {
  DaLds D;
  auto take = [&](size_t bytes) {
    char* q = p;
    p += bytes;
    return q;
  };
  D.w1t = reinterpret_cast<__bf16*>(take(2 * (size_t)kMlpHid * kDaLdIn));
  D.w2t = reinterpret_cast<__bf16*>(take(2 * (size_t)kMlpHid * kDaLdHid));
  D.w3t = reinterpret_cast<__bf16*>(take(2 * (size_t)kMlpHid * kDaLdHid));
  D.emb = reinterpret_cast<__bf16*>(take(2 * (size_t)kMlpTypes * kMlpEmb));
  D.x = reinterpret_cast<__bf16*>(take(2 * (size_t)kDaRows * kDaLdIn));
  D.ha = reinterpret_cast<__bf16*>(take(2 * (size_t)kDaRows * kDaLdHid));
  D.hb = reinterpret_cast<__bf16*>(take(2 * (size_t)kDaRows * kDaLdHid));
  p = reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(p) + 15) & ~uintptr_t(15));
  D.b1 = reinterpret_cast<float*>(take(4 * kMlpHid));
  D.b2 = reinterpret_cast<float*>(take(4 * kMlpHid));
  D.b3 = reinterpret_cast<float*>(take(4 * kMlpHid));
  D.w4 = reinterpret_cast<float*>(take(4 * kMlpHid));
  D.b4 = reinterpret_cast<float*>(take(16));
  D.ynew = reinterpret_cast<float*>(take(16));
  D.stats = reinterpret_cast<double*>(take(32));
  D.rows = reinterpret_cast<int32_t*>(take(4 * (kDaRows + 1)));
  D.ycache = reinterpret_cast<float*>(take(4 * (size_t)nres));
  D.var0 = reinterpret_cast<int16_t*>(take(2 * (size_t)nres));
  D.nvar = reinterpret_cast<int16_t*>(take(2 * (size_t)nres));
  D.type = reinterpret_cast<int16_t*>(take(2 * (size_t)nres));
  D.chain = reinterpret_cast<int16_t*>(take(2 * (size_t)nres));
  D.res_of_var = reinterpret_cast<int16_t*>(take(2 * (size_t)nvt));
  return D;
}

// Weights (BF16, transposed) and residue tables into LDS; all threads, a
// barrier must follow.
__device__ void stage_surrogate(const DaLds& D, const float* __restrict__ th, const MlpResidues& R) {
  for (int t = threadIdx.x; t < kMlpIn * kMlpHid; t += blockDim.x) {
    const int k = t / kMlpHid, n = t - k * kMlpHid;  // theta's W1 is [in][out]
    D.w1t[n * kDaLdIn + k] = (__bf16)th[kThW1 + t];
  }
  for (int t = threadIdx.x; t < kMlpHid * kMlpHid; t += blockDim.x) {
    const int k = t / kMlpHid, n = t - k * kMlpHid;
    D.w2t[n * kDaLdHid + k] = (__bf16)th[kThW2 + t];
    D.w3t[n * kDaLdHid + k] = (__bf16)th[kThW3 + t];
  }
  for (int t = threadIdx.x; t < kMlpTypes * kMlpEmb; t += blockDim.x) D.emb[t] = (__bf16)th[kThEmb + t];
  for (int t = threadIdx.x; t < kMlpHid; t += blockDim.x) {
    D.b1[t] = th[kThB1 + t];
    D.b2[t] = th[kThB2 + t];
    D.b3[t] = th[kThB3 + t];
    D.w4[t] = th[kThW4 + t];
  }
  if (threadIdx.x == 0) D.b4[0] = th[kThB4];
  for (int t = threadIdx.x; t < R.nres; t += blockDim.x) {
    D.var0[t] = R.var0[t];
    D.nvar[t] = R.nvar[t];
    D.type[t] = R.type[t];
    D.chain[t] = R.chain[t];
  }
  for (int t = threadIdx.x; t < R.nvt; t += blockDim.x) D.res_of_var[t] = R.res_of_var[t];
  if (threadIdx.x < 4) D.stats[threadIdx.x] = 0.0;
}

// One hidden layer on kDaRows rows: out = silu(in W + b); waves 0..3 take
// the four 16-column blocks; rows beyond n enter as zeros.
__device__ __forceinline__ void da_layer(const __bf16* in, int ldin, int K, const __bf16* wt, int ldw, const float* b,
                                         __bf16* out, int n) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x & 63;
  if (wave < kMlpHid / 16) {
    const int row = lane & 15, col = 16 * wave + (lane & 15), kq = 8 * (lane >> 4);
    da_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 32) {
      da_bf16x8 a;
      if (row < n)
        a = *reinterpret_cast<const da_bf16x8*>(&in[row * ldin + k0 + kq]);
      else
        for (int j = 0; j < 8; ++j) a[j] = (__bf16)0.0f;
      const da_bf16x8 bw = *reinterpret_cast<const da_bf16x8*>(&wt[col * ldw + k0 + kq]);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw, acc, 0, 0, 0);
    }
    for (int i = 0; i < 4; ++i) {
      const int r = 4 * (lane >> 4) + i;
      if (r < n) out[r * kDaLdHid + col] = (__bf16)silu(acc[i] + b[col]);
    }
  }
  __syncthreads();
}

// Network outputs of the windows of residues D.rows[0..n) for the variables
// vl, with variable jv taking the value vnew (jv < 0: none).  All threads;
// the outputs are in D.ynew on return.  One out-of-line copy; it carves the
// surrogate's LDS region (at byte offset off) itself, so no table of LDS
// pointers goes through memory.
__device__ __attribute__((noinline)) void surrogate_rows(size_t off, int nres, int nvt, const double* vl, int jv,
                                                         double vnew) {
  extern __shared__ double smem[];
  const DaLds D = carve_da_lds(reinterpret_cast<char*>(smem) + off, nres, nvt);
  const int n = D.rows[kDaRows];
  for (int t = threadIdx.x; t < n * kMlpIn; t += blockDim.x) {
    const int k = t / kMlpIn, col = t - k * kMlpIn, r = D.rows[k];
    float v = 0.0f;
    if (col < kMlpAngleCols) {
      const int d = col / (2 * kMlpSlots) - 1, s = (col % (2 * kMlpSlots)) / 2, q = r + d;
      if (q >= 0 && q < nres && D.chain[q] == D.chain[r] && s < D.nvar[q]) {
        const int var = D.var0[q] + s;
        const double a = var == jv ? vnew : vl[var];
        v = (float)((col & 1) ? cos(a) : sin(a));
      }
    } else if (col < kMlpEmbCol + 3 * kMlpEmb) {
      const int e = col - kMlpEmbCol, d = e / kMlpEmb, q = r + d - 1;
      const int ty = (q >= 0 && q < nres && D.chain[q] == D.chain[r]) ? D.type[q] : 0;
      v = (float)D.emb[ty * kMlpEmb + (e - d * kMlpEmb)];
    }
    D.x[k * kDaLdIn + col] = (__bf16)v;
  }
  __syncthreads();
  da_layer(D.x, kDaLdIn, kMlpIn, D.w1t, kDaLdIn, D.b1, D.ha, n);
  da_layer(D.ha, kDaLdHid, kMlpHid, D.w2t, kDaLdHid, D.b2, D.hb, n);
  da_layer(D.hb, kDaLdHid, kMlpHid, D.w3t, kDaLdHid, D.b3, D.ha, n);
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    for (int k = 0; k < n; ++k) {
      float v = (float)D.ha[k * kDaLdHid + lane] * D.w4[lane];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) D.ynew[k] = v + D.b4[0];
    }
  }
  __syncthreads();
}

// The windows variable jv changes (thread 0): residues r-1, r, r+1 of its chain.
__device__ inline void surrogate_window_rows(const DaLds& D, int nres, int jv) {
  const int r = D.res_of_var[jv];
  int n = 0;
  if (r >= 0)
    for (int d = -1; d <= 1; ++d) {
      const int q = r + d;
      if (q >= 0 && q < nres && D.chain[q] == D.chain[r]) D.rows[n++] = q;
    }
  D.rows[kDaRows] = n;
}

// Every residue's output at the current variables into D.ycache (all threads).
__device__ void surrogate_fill_cache(const DaLds& D, size_t off, int nres, int nvt, const double* vl) {
  for (int r0 = 0; r0 < nres; r0 += kDaRows) {
    if (threadIdx.x == 0) {
      int n = 0;
      for (int q = r0; q < nres && n < kDaRows; ++q) D.rows[n++] = q;
      D.rows[kDaRows] = n;
    }
    __syncthreads();
    surrogate_rows(off, nres, nvt, vl, -1, 0.0);
    if (threadIdx.x < D.rows[kDaRows]) D.ycache[D.rows[threadIdx.x]] = D.ynew[threadIdx.x];
    __syncthreads();
  }
}

}  // namespace smmp
