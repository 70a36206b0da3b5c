// Training and evaluation of the delayed-acceptance surrogate on the GPU:
// features of sampled conformations, forward and backward passes as BF16
// MFMA matrix products with FP32 accumulation (v_mfma_f32_16x16x32_bf16),
// and Adam.  Network and feature layout: mlp_common.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../host/devtopo.hpp"
#include "../host/engine.hpp"
#include "mlp_common.hpp"
#include "mlp_host.hpp"

namespace smmp {

#define HIPCHK(x)                                                                                   \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess) throw DeviceError(std::string(#x) + ": " + hipGetErrorString(e_));        \
  } while (0)

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------ features

// Angle columns of nconf conformations (vlvr: nconf x nvt FP64) into x
// (nconf x D, BF16); one thread per (conformation, column).  Embedding
// columns are filled by k_mlp_embed from the current parameters.
__global__ void k_mlp_features(MlpResidues R, const double* __restrict__ vlvr, int nconf, bf16* __restrict__ x) {
  const int D = R.nres * kMlpResCols;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nconf * D) return;
  const int64_t c = t / D;
  const int col = (int)(t - c * D), r = col / kMlpResCols, k = col - r * kMlpResCols;
  if (k >= kMlpEmbCol) return;
  const int s = k >> 1;
  bf16 v = (bf16)0.0f;
  if (s < R.nvar[r]) v = mlp_angle_feature(vlvr[(size_t)c * R.nvt + R.var0[r] + s], k & 1);
  x[t] = v;
}

// Embedding columns from the parameters (BF16 copies of theta's table).
__global__ void k_mlp_embed(const float* __restrict__ theta, MlpLayout Lo, const int16_t* __restrict__ type, int nconf,
                            bf16* __restrict__ x) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int per = Lo.nres * kMlpEmb;
  if (t >= (int64_t)nconf * per) return;
  const int64_t c = t / per;
  const int q = (int)(t - c * per), r = q / kMlpEmb, e = q - r * kMlpEmb;
  x[c * Lo.D + r * kMlpResCols + kMlpEmbCol + e] = (bf16)theta[Lo.emb + type[r] * kMlpEmb + e];
}

// BF16 copies of W1 [D][H], W2 and W3 [H][H] (input-major, as in theta).
__global__ void k_mlp_pack(const float* __restrict__ theta, MlpLayout Lo, bf16* __restrict__ w1,
                           bf16* __restrict__ w2, bf16* __restrict__ w3) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < Lo.D * kMlpHid) w1[t] = (bf16)theta[Lo.w1 + t];
  if (t < kMlpHid * kMlpHid) {
    w2[t] = (bf16)theta[Lo.w2 + t];
    w3[t] = (bf16)theta[Lo.w3 + t];
  }
}

// ------------------------------------------------------------ GEMM

// C = op(A) op(B): op(A)[m][k] = transA ? A[k*lda+m] : A[m*lda+k],
// op(B)[k][n] = transB ? B[n*ldb+k] : B[k*ldb+n].  256 threads compute a
// 64 x 64 tile of C: wave w rows 16w..16w+15, all four 16-column blocks,
// with the A and B chunks of 32 k staged in LDS.  blockIdx.z splits the
// k range into chunks of kchunk (partial products, epilogue kPartial).
enum GemmEpi { kEpiHidden = 0, kEpiBackHidden = 1, kEpiStore = 2, kEpiPartial = 3 };
struct GemmArgs {
  const bf16* A;
  const bf16* B;
  int lda, ldb, transA, transB, M, N, K, kchunk, epi;
  const float* bias;  // kEpiHidden: bias[n]
  float* Z;           // kEpiHidden: pre-activation out; kEpiBackHidden: pre-activation in
  bf16* H;            // kEpiHidden: silu(Z) out; kEpiBackHidden: dZ out
  float* C;           // kEpiStore: C out (ldc = N); kEpiPartial: partials [z][M][N]
};

constexpr int kGemmLd = 40;  // LDS row stride in bf16 (32 + 8 padding; 80 bytes keeps 16-byte reads aligned)

__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) bf16 As[64 * kGemmLd];
  __shared__ __attribute__((aligned(16))) bf16 Bs[64 * kGemmLd];
  const int tid = threadIdx.x, wave = tid / 64, lane = tid & 63;
  const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
  const int k_lo = blockIdx.z * g.kchunk, k_hi = min(g.K, k_lo + g.kchunk);
  f32x4 acc[4];
  for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = k_lo; k0 < k_hi; k0 += 32) {
    // stage: thread tid loads 8 elements of A (row tid/4, k 8*(tid%4)..) and of B (column tid/4)
    {
      const int r = tid >> 2, kk = (tid & 3) * 8;
      for (int j = 0; j < 8; ++j) {
        const int m = m0 + r, k = k0 + kk + j;
        bf16 a = (bf16)0.0f, b = (bf16)0.0f;
        if (m < g.M && k < k_hi) a = g.transA ? g.A[(size_t)k * g.lda + m] : g.A[(size_t)m * g.lda + k];
        const int n = n0 + r;
        if (n < g.N && k < k_hi) b = g.transB ? g.B[(size_t)n * g.ldb + k] : g.B[(size_t)k * g.ldb + n];
        As[r * kGemmLd + kk + j] = a;
        Bs[r * kGemmLd + kk + j] = b;
      }
    }
    __syncthreads();
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[(16 * wave + (lane & 15)) * kGemmLd + 8 * (lane >> 4)]);
    for (int t = 0; t < 4; ++t) {
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[(16 * t + (lane & 15)) * kGemmLd + 8 * (lane >> 4)]);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
    }
    __syncthreads();
  }
  for (int t = 0; t < 4; ++t)
    for (int i = 0; i < 4; ++i) {
      const int m = m0 + 16 * wave + 4 * (lane >> 4) + i, n = n0 + 16 * t + (lane & 15);
      if (m >= g.M || n >= g.N) continue;
      const float v = acc[t][i];
      const size_t o = (size_t)m * g.N + n;
      if (g.epi == kEpiHidden) {
        const float z = v + g.bias[n];
        g.Z[o] = z;
        g.H[o] = (bf16)silu(z);
      } else if (g.epi == kEpiBackHidden) {
        g.H[o] = (bf16)(v * silu_grad(g.Z[o]));
      } else if (g.epi == kEpiStore) {
        g.C[o] = v;
      } else {
        g.C[(size_t)blockIdx.z * g.M * g.N + o] = v;
      }
    }
}

// ------------------------------------------------------------ small kernels

// y[m] = w4 . H3[m] + b4
__global__ void k_mlp_out(const bf16* __restrict__ h3, const float* __restrict__ theta, MlpLayout Lo, int64_t rows,
                          float* __restrict__ y) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= rows) return;
  float s = theta[Lo.b4];
  for (int n = 0; n < kMlpHid; ++n) s += (float)h3[m * kMlpHid + n] * theta[Lo.w4 + n];
  y[m] = s;
}

// Predictions and the gradient of the loss with respect to them: with
// per-conformation targets e = y - target, dy = 2 e / nconf; with pairs
// (conformations 2c, 2c+1 and target[c] for y[2c+1] - y[2c]) e = y[2c+1] -
// y[2c] - target[c], dy = +-2 e / npairs; sq = e^2.
__global__ void k_mlp_conf(const float* __restrict__ y, int nconf, const float* __restrict__ target,
                           float* __restrict__ pred, float* __restrict__ dy, float* __restrict__ sq, int pairs) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (!pairs) {
    if (c >= nconf) return;
    if (pred) pred[c] = y[c];
    if (target) {
      const float e = y[c] - target[c];
      sq[c] = e * e;
      dy[c] = 2.0f * e / (float)nconf;
    }
    return;
  }
  const int np = nconf / 2;
  if (c >= np) return;
  if (pred) {
    pred[2 * c] = y[2 * c];
    pred[2 * c + 1] = y[2 * c + 1];
  }
  if (target) {
    const float e = (y[2 * c + 1] - y[2 * c]) - target[c];
    sq[c] = e * e;
    const float g = 2.0f * e / (float)np;
    dy[2 * c] = -g;
    dy[2 * c + 1] = g;
  }
}

// dZ3 = (dy w4^T) * silu'(Z3)
__global__ void k_mlp_dout(const float* __restrict__ dy, const float* __restrict__ theta, MlpLayout Lo,
                           const float* __restrict__ z3, int64_t rows, bf16* __restrict__ dz) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * kMlpHid) return;
  const int64_t m = t / kMlpHid;
  const int n = (int)(t - m * kMlpHid);
  dz[t] = (bf16)(dy[m] * theta[Lo.w4 + n] * silu_grad(z3[t]));
}

// Column sums over row blocks of 256: partial[b][n] = sum over rows m of
// block b of w[m] * A[m][n] (w = 1 when null); A is BF16 or FP32.
__global__ void k_colsum(const bf16* __restrict__ ab, const float* __restrict__ af, const float* __restrict__ w,
                         int64_t rows, int N, float* __restrict__ partial) {
  __shared__ float red[256];
  const int64_t m0 = (int64_t)blockIdx.x * 256;
  for (int n = 0; n < N; ++n) {
    const int64_t m = m0 + threadIdx.x;
    float v = 0.0f;
    if (m < rows) {
      v = ab ? (float)ab[m * N + n] : af[m * N + n];
      if (w) v *= w[m];
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)blockIdx.x * N + n] = red[0];
    __syncthreads();
  }
}

// out[i] = sum over nblk of partial[b][i] (fixed order), i < n
__global__ void k_sum_partials(const float* __restrict__ partial, int nblk, int n, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int b = 0; b < nblk; ++b) s += partial[(size_t)b * n + i];
  out[i] = s;
}

// Embedding gradients: block t sums dX's embedding columns over the
// (conformation, residue) pairs whose residue type is t (fixed order).
__global__ void k_mlp_emb_grad(const float* __restrict__ dx, const int16_t* __restrict__ type, MlpLayout Lo,
                               int64_t rows, float* __restrict__ grad) {
  __shared__ float red[256 * kMlpEmb];
  const int ty = blockIdx.x;
  float acc[kMlpEmb];
  for (int e = 0; e < kMlpEmb; ++e) acc[e] = 0.0f;
  const int64_t items = rows * Lo.nres;
  for (int64_t q = threadIdx.x; q < items; q += 256) {
    const int64_t m = q / Lo.nres;
    const int r = (int)(q - m * Lo.nres);
    if (type[r] != ty) continue;
    const float* d = dx + m * Lo.D + r * kMlpResCols + kMlpEmbCol;
    for (int e = 0; e < kMlpEmb; ++e) acc[e] += d[e];
  }
  for (int e = 0; e < kMlpEmb; ++e) red[threadIdx.x * kMlpEmb + e] = acc[e];
  __syncthreads();
  if (threadIdx.x < kMlpEmb) {
    float s = 0.0f;
    for (int k = 0; k < 256; ++k) s += red[k * kMlpEmb + threadIdx.x];
    grad[Lo.emb + ty * kMlpEmb + threadIdx.x] = s;
  }
}

// Adam (bias-corrected), step >= 1.
__global__ void k_mlp_adam(float* __restrict__ theta, const float* __restrict__ grad, float* __restrict__ m1,
                           float* __restrict__ m2, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = grad[i];
  const float m = b1 * m1[i] + (1.0f - b1) * g;
  const float v = b2 * m2[i] + (1.0f - b2) * g * g;
  m1[i] = m;
  m2[i] = v;
  theta[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// ------------------------------------------------------------ host side

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  T* get(size_t count) {
    const size_t bytes = count * sizeof(T);
    if (bytes > n) {
      if (p) HIPCHK(hipFree(p));
      p = nullptr;
      HIPCHK(hipMalloc(&p, bytes));
      n = bytes;
    }
    return static_cast<T*>(p);
  }
};

inline unsigned nblocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void gemm(hipStream_t st, const GemmArgs& a, int splits) {
  GemmArgs g = a;
  const int chunk = ((g.K + splits - 1) / splits + 31) / 32 * 32;
  g.kchunk = chunk;
  const int nz = (g.K + chunk - 1) / chunk;
  hipLaunchKernelGGL(k_gemm_bf16, dim3(nblocks(g.M, 64), nblocks(g.N, 64), nz), dim3(256), 0, st, g);
  HIPCHK(hipGetLastError());
}

}  // namespace

// Samples idx[0..nb) of the training set into contiguous batch buffers;
// a sample is one conformation, or with pairs two consecutive ones.
__global__ void k_gather(const double* __restrict__ v, const float* __restrict__ t, const int32_t* __restrict__ idx,
                         int nb, int nvt, int per, double* __restrict__ bv, float* __restrict__ bt) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row_len = (int64_t)per * nvt;
  if (k >= (int64_t)nb * row_len) return;
  const int c = (int)(k / row_len);
  const int64_t j = k - (int64_t)c * row_len;
  bv[k] = v[(size_t)idx[c] * row_len + j];
  if (j == 0) bt[c] = t[idx[c]];
}

// acc[0] += sum of sq[0..n) (one thread, fixed order)
__global__ void k_sum_sq(const float* __restrict__ sq, int n, double* __restrict__ acc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += sq[i];
  acc[0] += s;
}

// Workspace of the training passes and the context's network (FP32
// parameters and Adam moments), kept with the context, plus the BF16 copies
// of the weights the sampler's screen reads.
struct MlpWork {
  DevBuf res, x, z1, z2, z3, h1, h2, h3, dz_a, dz_b, dx, y, dy, sq, pred, part, wb;
  DevBuf theta, grad, m1, m2;
  DevBuf data_v, data_t, perm, bv, bt, lossacc, tmp_v, tmp_t;
  DevBuf sw1, sw2t, sw3t;  // screen weights: W1 [D][H], W2^T and W3^T [out][in], BF16
  std::vector<int16_t> host_tables;
  int step = 0;
  int n = 0;               // parameters of the network held (0: none)
  int nres = 0;            // residues it was made for
  bool screen_fresh = false;
};

void mlp_work_destroy(MlpWork* w) { delete w; }

static int residue_type_id(const std::string& name) {
  static const char* kStd[20] = {"ala", "arg", "asn", "asp", "cys", "gln", "glu", "gly", "his", "ile",
                                 "leu", "lys", "met", "phe", "pro", "ser", "thr", "trp", "tyr", "val"};
  std::string n = name.substr(0, 3);
  for (auto& ch : n) ch = (char)std::tolower((unsigned char)ch);
  for (int k = 0; k < 20; ++k)
    if (n == kStd[k]) return k + 1;
  return 21;  // any other residue of the library
}

int mlp_residue_count(const State& s) { return s.ntlml >= 1 ? s.irsml2(s.ntlml) : 0; }

// Residue tables of the system: first variable, variable count, type and
// chain of each residue (all chains), residue of each variable.
MlpResidues mlp_residues(MlpWork*& w, const State& s, hipStream_t st) {
  if (!w) w = new MlpWork();
  const int nml = s.ntlml;
  if (nml < 1) throw Error("surrogate> no molecule initialised");
  const int nres = s.irsml2(nml), nvt = s.nvars();
  std::vector<int16_t> h(4 * (size_t)nres + nvt);
  int16_t *var0 = h.data(), *nvar = var0 + nres, *type = nvar + nres, *chain = type + nres, *rov = chain + nres;
  for (int v = 0; v < nvt; ++v) rov[v] = -1;
  for (int m = 1; m <= nml; ++m)
    for (int r = s.irsml1(m); r <= s.irsml2(m); ++r) {
      const int k = r - 1;
      if (s.nvrrs(r) > kMlpSlots)
        throw Error("surrogate> residue " + std::to_string(r) + " has more than 10 variables");
      var0[k] = (int16_t)(s.ivrrs1(r) - 1);
      nvar[k] = (int16_t)s.nvrrs(r);
      type[k] = (int16_t)residue_type_id(s.seq.get(r));
      chain[k] = (int16_t)(m - 1);
      for (int j = 0; j < s.nvrrs(r); ++j) rov[s.ivrrs1(r) - 1 + j] = (int16_t)k;
    }
  if (w->host_tables != h) {
    int16_t* d = w->res.get<int16_t>(h.size());
    HIPCHK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(int16_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    w->host_tables = h;
  }
  const int16_t* d = static_cast<const int16_t*>(w->res.p);
  MlpResidues R;
  R.nres = nres;
  R.nvt = nvt;
  R.var0 = d;
  R.nvar = d + nres;
  R.type = d + 2 * (size_t)nres;
  R.chain = d + 3 * (size_t)nres;
  R.res_of_var = d + 4 * (size_t)nres;
  return R;
}

static MlpWork& work(MlpWork*& w) {
  if (!w) w = new MlpWork();
  return *w;
}

void mlp_set_theta(MlpWork*& wp, hipStream_t st, const float* host, int nres) {
  MlpWork& w = work(wp);
  const int n = mlp_layout(nres).n;
  float* th = w.theta.get<float>(n);
  HIPCHK(hipMemcpyAsync(th, host, sizeof(float) * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(w.m1.get<float>(n), 0, sizeof(float) * n, st));
  HIPCHK(hipMemsetAsync(w.m2.get<float>(n), 0, sizeof(float) * n, st));
  HIPCHK(hipStreamSynchronize(st));
  w.step = 0;
  w.n = n;
  w.nres = nres;
  w.screen_fresh = false;
}

static void need_theta(const MlpWork& w, int nres) {
  if (w.n == 0) throw Error("surrogate> no network set (smmp_mlp_set)");
  if (w.nres != nres)
    throw Error("surrogate> the network was made for " + std::to_string(w.nres) + " residues, the system has " +
                std::to_string(nres));
}

void mlp_get_theta(MlpWork*& wp, hipStream_t st, float* host, int nres) {
  MlpWork& w = work(wp);
  need_theta(w, nres);
  HIPCHK(hipMemcpyAsync(host, w.theta.p, sizeof(float) * w.n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
}

// W2^T, W3^T [out][in]: 8 consecutive inputs of one output are contiguous,
// the B operand of a 16x16x32 MFMA step in one 16-byte load.
__global__ void k_mlp_pack_screen(const float* __restrict__ theta, MlpLayout Lo, bf16* __restrict__ w1,
                                  bf16* __restrict__ w2t, bf16* __restrict__ w3t) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < Lo.D * kMlpHid) w1[t] = (bf16)theta[Lo.w1 + t];
  if (t < kMlpHid * kMlpHid) {
    const int k = t / kMlpHid, n = t - k * kMlpHid;  // theta: [in k][out n]
    w2t[n * kMlpHid + k] = (bf16)theta[Lo.w2 + t];
    w3t[n * kMlpHid + k] = (bf16)theta[Lo.w3 + t];
  }
}

MlpScreen mlp_screen(MlpWork*& wp, const State& s, hipStream_t st) {
  MlpWork& w = work(wp);
  const MlpResidues R = mlp_residues(wp, s, st);
  need_theta(w, R.nres);
  const MlpLayout Lo = mlp_layout(R.nres);
  bf16* w1 = w.sw1.get<bf16>((size_t)Lo.D * kMlpHid);
  bf16* w2t = w.sw2t.get<bf16>(kMlpHid * kMlpHid);
  bf16* w3t = w.sw3t.get<bf16>(kMlpHid * kMlpHid);
  if (!w.screen_fresh) {
    hipLaunchKernelGGL(k_mlp_pack_screen, dim3(nblocks(std::max(Lo.D, kMlpHid) * (int64_t)kMlpHid, 256)), dim3(256),
                       0, st, static_cast<const float*>(w.theta.p), Lo, w1, w2t, w3t);
    HIPCHK(hipGetLastError());
    w.screen_fresh = true;
  }
  MlpScreen S;
  S.res = R;
  S.lo = Lo;
  S.theta = static_cast<const float*>(w.theta.p);
  S.w1 = w1;
  S.w2t = w2t;
  S.w3t = w3t;
  return S;
}

// Forward pass over nconf conformations (vlvr on the device, FP64), and
// with targets (device) the squared errors (sq, device) and the gradient of
// the mean squared error (grad, device).  Enqueued on st only.
static void mlp_pass(MlpWork& w, const MlpResidues& R, hipStream_t st, const float* theta, const double* vlvr,
                     int nconf, const float* target, float* pred, float* grad, float* sq, int pairs) {
  const MlpLayout Lo = mlp_layout(R.nres);
  const int D = Lo.D;
  const int64_t rows = nconf;
  bf16* x = w.x.get<bf16>(rows * D);
  float *z1 = w.z1.get<float>(rows * kMlpHid), *z2 = w.z2.get<float>(rows * kMlpHid),
        *z3 = w.z3.get<float>(rows * kMlpHid);
  bf16 *h1 = w.h1.get<bf16>(rows * kMlpHid), *h2 = w.h2.get<bf16>(rows * kMlpHid),
       *h3 = w.h3.get<bf16>(rows * kMlpHid);
  float* y = w.y.get<float>(rows);
  if (target) HIPCHK(hipMemsetAsync(grad, 0, sizeof(float) * Lo.n, st));
  bf16* wb = w.wb.get<bf16>((size_t)D * kMlpHid + 2 * kMlpHid * kMlpHid);
  bf16 *w1 = wb, *w2 = wb + (size_t)D * kMlpHid, *w3 = w2 + kMlpHid * kMlpHid;
  hipLaunchKernelGGL(k_mlp_features, dim3(nblocks(rows * D, 256)), dim3(256), 0, st, R, vlvr, nconf, x);
  hipLaunchKernelGGL(k_mlp_embed, dim3(nblocks(rows * R.nres * kMlpEmb, 256)), dim3(256), 0, st, theta, Lo, R.type,
                     nconf, x);
  hipLaunchKernelGGL(k_mlp_pack, dim3(nblocks(std::max(D, kMlpHid) * (int64_t)kMlpHid, 256)), dim3(256), 0, st,
                     theta, Lo, w1, w2, w3);
  HIPCHK(hipGetLastError());
  const int M = (int)rows;
  auto hidden = [&](const bf16* in, int K, const bf16* W, const float* b, float* Z, bf16* H) {
    GemmArgs g{};
    g.A = in; g.lda = K; g.B = W; g.ldb = kMlpHid; g.M = M; g.N = kMlpHid; g.K = K;
    g.epi = kEpiHidden; g.bias = b; g.Z = Z; g.H = H;
    gemm(st, g, 1);
  };
  hidden(x, D, w1, theta + Lo.b1, z1, h1);
  hidden(h1, kMlpHid, w2, theta + Lo.b2, z2, h2);
  hidden(h2, kMlpHid, w3, theta + Lo.b3, z3, h3);
  hipLaunchKernelGGL(k_mlp_out, dim3(nblocks(rows, 256)), dim3(256), 0, st, h3, theta, Lo, rows, y);
  float* dy = target ? w.dy.get<float>(rows) : nullptr;
  hipLaunchKernelGGL(k_mlp_conf, dim3(nblocks(nconf, 256)), dim3(256), 0, st, y, nconf, target, pred, dy, sq, pairs);
  HIPCHK(hipGetLastError());
  if (!target) return;
  // ---- backward
  bf16 *dza = w.dz_a.get<bf16>(rows * kMlpHid), *dzb = w.dz_b.get<bf16>(rows * kMlpHid);
  float* dx = w.dx.get<float>(rows * D);
  const int nb = (int)nblocks(rows, 256);
  const int splits = std::max(1, std::min(64, M / 256));
  // partials of the weight gradients: splits x K x H for K = D (layer 1) and
  // K = H (layers 2, 3), whichever is larger; column sums nb x (H + 1)
  const size_t kmax = (size_t)std::max(D, kMlpHid);
  float* part = w.part.get<float>(std::max<size_t>((size_t)nb * (kMlpHid + 1), (size_t)splits * kmax * kMlpHid));
  auto colsum = [&](const bf16* ab, const float* af, const float* wt, int N, float* out) {
    hipLaunchKernelGGL(k_colsum, dim3(nb), dim3(256), 0, st, ab, af, wt, rows, N, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(nblocks(N, 64)), dim3(64), 0, st, part, nb, N, out);
  };
  auto wgrad = [&](const bf16* in, int K, const bf16* dz, float* out) {  // out[K][N] = in^T dz
    GemmArgs g{};
    g.A = in; g.lda = K; g.transA = 1; g.B = dz; g.ldb = kMlpHid; g.M = K; g.N = kMlpHid; g.K = M;
    g.epi = kEpiPartial; g.C = part;
    gemm(st, g, splits);
    const int chunk = ((M + splits - 1) / splits + 31) / 32 * 32, nz = (M + chunk - 1) / chunk;
    hipLaunchKernelGGL(k_sum_partials, dim3(nblocks((int64_t)K * kMlpHid, 256)), dim3(256), 0, st, part, nz,
                       K * kMlpHid, out);
  };
  auto back = [&](const bf16* dz, const bf16* W, int K, const float* Zprev, bf16* dzprev) {  // (dz W^T) * silu'
    GemmArgs g{};
    g.A = dz; g.lda = kMlpHid; g.B = W; g.ldb = kMlpHid; g.transB = 1; g.M = M; g.N = K; g.K = kMlpHid;
    g.epi = kEpiBackHidden; g.Z = const_cast<float*>(Zprev); g.H = dzprev;
    gemm(st, g, 1);
  };
  colsum(h3, nullptr, dy, kMlpHid, grad + Lo.w4);  // output layer
  colsum(nullptr, dy, nullptr, 1, grad + Lo.b4);
  hipLaunchKernelGGL(k_mlp_dout, dim3(nblocks(rows * kMlpHid, 256)), dim3(256), 0, st, dy, theta, Lo, z3, rows, dza);
  wgrad(h2, kMlpHid, dza, grad + Lo.w3);  // layer 3
  colsum(dza, nullptr, nullptr, kMlpHid, grad + Lo.b3);
  back(dza, w3, kMlpHid, z2, dzb);
  wgrad(h1, kMlpHid, dzb, grad + Lo.w2);  // layer 2
  colsum(dzb, nullptr, nullptr, kMlpHid, grad + Lo.b2);
  back(dzb, w2, kMlpHid, z1, dza);
  wgrad(x, D, dza, grad + Lo.w1);  // layer 1, and the embeddings through the input
  colsum(dza, nullptr, nullptr, kMlpHid, grad + Lo.b1);
  {
    GemmArgs g{};
    g.A = dza; g.lda = kMlpHid; g.B = w1; g.ldb = kMlpHid; g.transB = 1; g.M = M; g.N = D; g.K = kMlpHid;
    g.epi = kEpiStore; g.C = dx;
    gemm(st, g, 1);
  }
  hipLaunchKernelGGL(k_mlp_emb_grad, dim3(kMlpTypes), dim3(256), 0, st, dx, R.type, Lo, rows, grad);
  HIPCHK(hipGetLastError());
}

static void adam_step(MlpWork& w, hipStream_t st, float lr) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  ++w.step;
  const float c1 = 1.0f - std::pow(b1, (float)w.step), c2 = 1.0f - std::pow(b2, (float)w.step);
  hipLaunchKernelGGL(k_mlp_adam, dim3(nblocks(w.n, 256)), dim3(256), 0, st, static_cast<float*>(w.theta.p),
                     static_cast<const float*>(w.grad.p), static_cast<float*>(w.m1.p), static_cast<float*>(w.m2.p),
                     w.n, lr, b1, b2, eps, c1, c2);
  HIPCHK(hipGetLastError());
  w.screen_fresh = false;
}

// Predictions (and with targets the loss and gradient) of the context's
// network for host conformations.
double mlp_grad(MlpWork*& wp, const State& s, hipStream_t st, const double* vlvr, int nconf, const float* target,
                float* pred, float* grad, int pairs) {
  const MlpResidues R = mlp_residues(wp, s, st);
  MlpWork& w = *wp;
  need_theta(w, R.nres);
  double* dv = w.tmp_v.get<double>((size_t)nconf * R.nvt);
  HIPCHK(hipMemcpyAsync(dv, vlvr, sizeof(double) * nconf * R.nvt, hipMemcpyHostToDevice, st));
  if (pairs && nconf % 2) throw Error("surrogate> pairs need an even number of conformations");
  const int nt = pairs ? nconf / 2 : nconf;  // targets
  float* dt = nullptr;
  if (target) {
    dt = w.tmp_t.get<float>(nt);
    HIPCHK(hipMemcpyAsync(dt, target, sizeof(float) * nt, hipMemcpyHostToDevice, st));
  }
  float* dp = w.pred.get<float>(nconf);
  float* sq = w.sq.get<float>(nconf);
  float* g = w.grad.get<float>(w.n);
  mlp_pass(w, R, st, static_cast<const float*>(w.theta.p), dv, nconf, dt, dp, g, sq, pairs);
  if (pred) HIPCHK(hipMemcpyAsync(pred, dp, sizeof(float) * nconf, hipMemcpyDeviceToHost, st));
  std::vector<float> hsq(target ? nt : 0);
  if (target) {
    if (grad) HIPCHK(hipMemcpyAsync(grad, g, sizeof(float) * w.n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hsq.data(), sq, sizeof(float) * nt, hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  double loss = 0.0;
  for (float v : hsq) loss += v;
  return target ? loss / nt : 0.0;
}

// Minibatch Adam on the context's network: n samples (host) with targets,
// 'epochs' passes in random order (std::mt19937 with seed), batches of
// 'batch' (the last partial batch of an epoch is left out).
void mlp_fit(MlpWork*& wp, const State& s, hipStream_t st, const double* vlvr, const float* target, int n, int epochs,
             int batch, float lr, uint32_t seed, double* epoch_loss, int pairs) {
  const MlpResidues R = mlp_residues(wp, s, st);
  MlpWork& w = *wp;
  need_theta(w, R.nres);
  if (batch < 1 || n < batch) throw Error("surrogate> need at least one batch of training data");
  const int per = pairs ? 2 : 1;  // conformations per sample
  double* dv = w.data_v.get<double>((size_t)n * per * R.nvt);
  float* dtg = w.data_t.get<float>(n);
  HIPCHK(hipMemcpyAsync(dv, vlvr, sizeof(double) * (size_t)n * per * R.nvt, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dtg, target, sizeof(float) * n, hipMemcpyHostToDevice, st));
  int32_t* dperm = w.perm.get<int32_t>(n);
  double* bv = w.bv.get<double>((size_t)batch * per * R.nvt);
  float* bt = w.bt.get<float>(batch);
  float* sq = w.sq.get<float>(batch * per);
  float* pred = w.pred.get<float>(batch * per);
  float* g = w.grad.get<float>(w.n);
  double* acc = w.lossacc.get<double>(1);
  std::vector<int32_t> perm(n);
  std::mt19937 rng(seed);
  const int nbatch = n / batch;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int i = 0; i < n; ++i) perm[i] = i;
    std::shuffle(perm.begin(), perm.end(), rng);
    HIPCHK(hipMemcpyAsync(dperm, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(acc, 0, sizeof(double), st));
    for (int b = 0; b < nbatch; ++b) {
      hipLaunchKernelGGL(k_gather, dim3(nblocks((int64_t)batch * per * R.nvt, 256)), dim3(256), 0, st, dv, dtg,
                         dperm + (size_t)b * batch, batch, R.nvt, per, bv, bt);
      mlp_pass(w, R, st, static_cast<const float*>(w.theta.p), bv, batch * per, bt, pred, g, sq, pairs);
      hipLaunchKernelGGL(k_sum_sq, dim3(1), dim3(64), 0, st, sq, batch, acc);
      adam_step(w, st, lr);
    }
    double tot = 0.0;
    HIPCHK(hipMemcpyAsync(&tot, acc, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));  // perm is a host vector reused next epoch
    if (epoch_loss) epoch_loss[ep] = tot / ((double)nbatch * batch);
  }
}

}  // namespace smmp
