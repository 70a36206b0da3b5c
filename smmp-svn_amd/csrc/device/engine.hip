// HIP engine: device mirrors of contexts and the kernel launches behind the
// C-ABI.  Kernels: k_energy (full energy per conformation), the sweep and
// exchange kernels of sampler.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../host/devtopo.hpp"
#include "../host/engine.hpp"
#include "common.hpp"
#include "energy.hpp"
#include "esolan_host.hpp"
#include "gradient.hpp"
#include "mlp_host.hpp"
#include "sampler.hpp"

namespace smmp {

#define HIPCHK(x)                                                                                   \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess) throw DeviceError(std::string(#x) + ": " + hipGetErrorString(e_));        \
  } while (0)

constexpr int kEnergyThreads = 512;
// k_energy is compiled for up to 1024 threads (it needs fewer than 128 VGPRs,
// so four waves fit per SIMD); systems whose LDS allows one workgroup per CU
// take 1024 threads, the rest 512.
constexpr int kEnergyThreadsMax = 1024;
// k_energy values per conformation: total, eysm, eyel, eyvw, eyhb, eyvr,
// eysl, asa, then between chains eysmi, eyeli, eyvwi, eyhbi
constexpr int kEnergyOut = 12;

// ------------------------------------------------------------- kernels

template <bool MIXED>
__global__ void __launch_bounds__(kEnergyThreadsMax) k_energy(const DevSys* __restrict__ systems,
                                                           const int32_t* __restrict__ conf_sys,
                                                           const int64_t* __restrict__ conf_off,
                                                           const double* __restrict__ vlvr, double* __restrict__ out,
                                                           double* __restrict__ xyz_out, int xyz_stride,
                                                           int32_t* __restrict__ err, double* __restrict__ work,
                                                           int bits_lds, const int32_t* __restrict__ order) {
  extern __shared__ double smem[];
  const int c = order ? order[blockIdx.x] : (int)blockIdx.x;  // block b evaluates conformation order[b]
  const DevSys& S = systems[conf_sys[c]];
  const double* vl = vlvr + conf_off[c];
  EnergyLds L = carve_energy_lds(smem, S, blockDim.x / kWave, bits_lds != 0);
  EnergyOut o;
  Work wk;
  energy_stage_static(S, L);
  conformation_energy<MIXED>(S, vl, S.gbpr, L, nullptr, o, wk, 0, 1, true);
  if (threadIdx.x == 0) {
    double eysm = o.el + o.vw + o.hb + o.tor;  // energy.f / enyshe.f:187
    const double eysmi = o.eli + o.vwi + o.hbi;  // eninteract.f:83
    double* r = out + (size_t)c * kEnergyOut;
    r[0] = eysm + o.sol + eysmi;
    r[1] = eysm;
    r[2] = o.el;
    r[3] = o.vw;
    r[4] = o.hb;
    r[5] = o.tor;
    r[6] = o.sol;
    r[7] = o.asa;
    r[8] = eysmi;
    r[9] = o.eli;
    r[10] = o.vwi;
    r[11] = o.hbi;
    if (o.err) atomicOr(err, 1);
  }
  if (xyz_out) {  // conformation c at xyz_out + 3 * xyz_stride * c
    double* xo = xyz_out + (size_t)3 * xyz_stride * c;
    for (int a = threadIdx.x; a < S.nat; a += blockDim.x) {
      xo[3 * a + 0] = L.x[a];
      xo[3 * a + 1] = L.y[a];
      xo[3 * a + 2] = L.z[a];
    }
  }
  if (work) {  // block totals of the work counters (one row per conformation)
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
    double v[3] = {wk.pairs, wk.contacts, wk.builds};
    for (int k = 0; k < 3; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0)
      for (int k = 0; k < 3; ++k) L.red[8 * wave + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
      for (int k = 0; k < 3; ++k) {
        double t = 0.0;
        for (int w = 0; w < nw; ++w) t += L.red[8 * w + k];
        work[4 * (size_t)c + k] = t;
      }
  }
}

// Energy and derivatives by the variables of one conformation per block
// (gradient.f in vacuum): eysm, eyel, eyvw, eyhb, eyvr at out + kGradOut c,
// dE/dv at grad + conf_off[c] (the layout of vlvr).
constexpr int kGradThreads = 256;
__global__ void __launch_bounds__(kGradThreads) k_gradient(const DevSys* __restrict__ systems,
                                                         const int32_t* __restrict__ conf_sys,
                                                         const int64_t* __restrict__ conf_off,
                                                         const double* __restrict__ vlvr, double* __restrict__ out,
                                                         double* __restrict__ grad, double* __restrict__ work,
                                                         int bits_lds, double* __restrict__ xyz_out, int xyz_stride) {
  extern __shared__ double smem[];
  const int c = blockIdx.x;
  const DevSys& S = systems[conf_sys[c]];
  const double* vl = vlvr + conf_off[c];
  const int nw = blockDim.x / kWave, wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  EnergyLds L = carve_energy_lds(smem, S, nw, bits_lds != 0);
  Work wk;
  energy_stage_static(S, L);
  energy_stage_torsions(S, vl, L);
  __syncthreads();
  build_chain(S, L, wk, S.gbpr);
  if (xyz_out) {  // the coordinates for the solvent term (x, y, z per atom)
    double* o = xyz_out + (size_t)3 * xyz_stride * c;
    for (int a = threadIdx.x; a < S.nat; a += blockDim.x) {
      o[3 * a] = L.x[a];
      o[3 * a + 1] = L.y[a];
      o[3 * a + 2] = L.z[a];
    }
  }
  double* f = L.fr;  // the frames are dead once the chain is built
  double el = 0.0, vw = 0.0, hb = 0.0;
  pair_forces(S, L, f, el, vw, hb, wk);
  const double tor = torsion_terms(S, vl);
  __syncthreads();  // forces complete
  torsion_derivatives(S, vl, L, f, grad + conf_off[c]);
  double v[5] = {el, vw, hb, tor, wk.pairs};
#pragma unroll
  for (int k = 0; k < 5; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0)
    for (int k = 0; k < 5; ++k) L.red[8 * wave + k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int w = 0; w < nw; ++w)
      for (int k = 0; k < 5; ++k) t[k] += L.red[8 * w + k];
    double* r = out + (size_t)kGradOut * c;
    r[0] = t[0] + t[1] + t[2] + t[3];  // eysm = eyel + eyvw + eyhb + eyvr (opeshe.f:384)
    r[1] = t[0];
    r[2] = t[1];
    r[3] = t[2];
    r[4] = t[3];
    if (work) {  // pair terms, -, atoms built, - (the energy kernel's counters)
      work[4 * (size_t)c + 0] = t[4];
      work[4 * (size_t)c + 1] = 0.0;
      work[4 * (size_t)c + 2] = (double)S.nat;
      work[4 * (size_t)c + 3] = 0.0;
    }
  }
}

// The analytic solvent's part of dE/dv (opesol.f:74-247), added to grad:
// forces f = -gradan per atom, and per variable, as in the vacuum kernel,
// s [(e x b) . F + e . T] over its moved atoms (F the sum of f, T of f x r;
// the moved range is the part before the axis where var_sign is -1: the
// term does not change when the whole chain turns).  One workgroup per
// conformation; waves over variables, lanes over atoms; coordinates and
// gradan at xyz / gradan + 3 (stride c + atom).
__global__ void __launch_bounds__(kGradThreads) k_sol_project(const DevSys* __restrict__ systems,
                                                            const int32_t* __restrict__ conf_sys,
                                                            const int64_t* __restrict__ conf_off,
                                                            const double* __restrict__ xyz,
                                                            const double* __restrict__ gradan, int stride,
                                                            double* __restrict__ grad) {
  const int c = blockIdx.x;
  const DevSys& S = systems[conf_sys[c]];
  const int nw = blockDim.x / kWave, wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const double* X = xyz + (size_t)3 * stride * c;
  const double* F = gradan + (size_t)3 * stride * c;
  for (int v = wave; v < S.nvt; v += nw) {
    const int m1 = S.var_m1[v], m2 = S.var_m2[v];
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int a = m1 + lane; a <= m2; a += kWave) {
      const double fx = -F[3 * a], fy = -F[3 * a + 1], fz = -F[3 * a + 2];
      const double x = X[3 * a], y = X[3 * a + 1], z = X[3 * a + 2];
      s[0] += fx;
      s[1] += fy;
      s[2] += fz;
      s[3] += fy * z - fz * y;
      s[4] += fz * x - fx * z;
      s[5] += fx * y - fy * x;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] = wave_sum(s[k]);
    if (lane == 0) {
      const int b = S.var_base[v], p = S.var_axis0[v];
      const double xb = X[3 * b], yb = X[3 * b + 1], zb = X[3 * b + 2];
      double ex = xb - X[3 * p], ey = yb - X[3 * p + 1], ez = zb - X[3 * p + 2];
      const double en = rsqrt_f64(ex * ex + ey * ey + ez * ez);
      ex *= en;
      ey *= en;
      ez *= en;
      grad[conf_off[c] + v] += S.var_sign[v] * ((ey * zb - ez * yb) * s[0] + (ez * xb - ex * zb) * s[1] +
                                               (ex * yb - ey * xb) * s[2] + ex * s[3] + ey * s[4] + ez * s[5]);
    }
  }
}

// ------------------------------------------------------ device buffers

struct DevMem {
  void* p = nullptr;
  size_t n = 0;
  DevMem() = default;
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  ~DevMem() {
    if (p) (void)hipFree(p);
  }
  void* ensure(size_t bytes) {
    if (bytes > n) {
      if (p) HIPCHK(hipFree(p));
      p = nullptr;
      n = 0;
      HIPCHK(hipMalloc(&p, bytes));
      n = bytes;
    }
    return p;
  }
};

// Replica state of the samplers (grown on demand, reused across calls).
struct SamplerBuffers {
  DevMem idvr, axvr, vlvr, xyz, nfree, valid, eol, acc, mt, rng_of, slot_of, beta, err, dbg;
  DevMem wb, walpha, hist, traj, ipoi, iswitch, work, trace, trace_n, phase, part, sync, gsync;
  DevMem sweep_base, tune_vl, tune_mt, tune_gb, gb, da_stats, snap, snap_e, obs, varx, upd, spec_nf;
};

// One system uploaded to the device: a single allocation holding every
// array, plus the DevSys with the device pointers.
struct DevSystem {
  HostSys h;
  DevSys d;
  DevMem mem;
};

// Lays out one system's arrays in a blob (256-byte aligned parts) and fills
// d with pointers into the device copy of that blob at 'base'.  base null:
// only the size is computed (returned).
static size_t system_blob(const HostSys& h, char* base, std::vector<uint8_t>& blob, DevSys& d) {
  size_t off = 0;
  std::vector<std::pair<size_t, std::pair<const void*, size_t>>> parts;
  auto add = [&](const void* src, size_t bytes) {
    off = (off + 255) & ~size_t(255);
    size_t o = off;
    parts.push_back({o, {src, bytes}});
    off += std::max<size_t>(bytes, 1);
    return o;
  };
  size_t o_parent = add(h.parent.data(), h.parent.size() * 2);
  size_t o_lev = add(h.lev_atoms.data(), h.lev_atoms.size() * 2);
  size_t o_levoff = add(h.lev_off.data(), h.lev_off.size() * 4);
  size_t o_type = add(h.type.data(), h.type.size());
  size_t o_q = add(h.q.data(), h.q.size() * 8);
  size_t o_bl = add(h.bl.data(), h.bl.size() * 8);
  size_t o_cba = add(h.cba.data(), h.cba.size() * 8);
  size_t o_sba = add(h.sba.data(), h.sba.size() * 8);
  size_t o_to = add(h.toat0.data(), h.toat0.size() * 8);
  size_t o_tref = add(h.tref.data(), h.tref.size() * 8);
  size_t o_rv = add(h.rvdw.data(), h.rvdw.size() * 8);
  size_t o_sg = add(h.sigma.data(), h.sigma.size() * 8);
  size_t o_tvar = add(h.tvar.data(), h.tvar.size() * 2);
  size_t o_va = add(h.var_atom.data(), h.var_atom.size() * 2);
  size_t o_vb = add(h.var_base.data(), h.var_base.size() * 2);
  size_t o_vx = add(h.var_axis0.data(), h.var_axis0.size() * 2);
  size_t o_m1 = add(h.var_m1.data(), h.var_m1.size() * 2);
  size_t o_m2 = add(h.var_m2.data(), h.var_m2.size() * 2);
  size_t o_h1 = add(h.var_h1.data(), h.var_h1.size() * 2);
  size_t o_h2 = add(h.var_h2.data(), h.var_h2.size() * 2);
  size_t o_vsign = add(h.var_sign.data(), h.var_sign.size() * 8);
  size_t o_e0 = add(h.var_e0.data(), h.var_e0.size() * 8);
  size_t o_vs = add(h.var_sg.data(), h.var_sg.size() * 8);
  size_t o_vr = add(h.var_rn.data(), h.var_rn.size() * 8);
  size_t o_bits = add(h.bits.data(), h.bits.size() * 8);
  size_t o_heavy = add(h.heavy.data(), h.heavy.size() * 2);
  size_t o_heavy_of = add(h.heavy_of.data(), h.heavy_of.size() * 2);
  size_t o_mol_of = add(h.mol_of.data(), h.mol_of.size() * 2);
  size_t o_pA = add(h.pA.data(), h.pA.size() * 8);
  size_t o_pA14 = add(h.pA14.data(), h.pA14.size() * 8);
  size_t o_pC = add(h.pC.data(), h.pC.size() * 8);
  size_t o_pHB = add(h.pHB.data(), h.pHB.size() * 4);
  size_t o_sp = add(h.spoint.data(), h.spoint.size() * 8);
  off = (off + 255) & ~size_t(255);
  if (!base) return off;
  blob.assign(off, 0);
  for (auto& pr : parts)
    if (pr.second.second) std::memcpy(blob.data() + pr.first, pr.second.first, pr.second.second);
  d.nat = h.nat; d.nvt = h.nvt; d.nlev = h.nlev; d.words = h.words; d.nheavy = h.nheavy;
  d.npnt = h.npnt; d.epsd = h.epsd; d.itysol = h.itysol;
  d.nml = h.nml;
  for (int m = 0; m < kMaxMol; ++m) {
    d.mol_a1[m] = h.mol_a1[m]; d.mol_a2[m] = h.mol_a2[m];
    d.mol_i2[m] = h.mol_i2[m]; d.mol_i3[m] = h.mol_i3[m];
    d.mol_h1[m] = h.mol_h1[m]; d.mol_h2[m] = h.mol_h2[m];
  }
  d.boxsize = h.boxsize;
  d.ca_first = h.ca_first;
  d.ca_last = h.ca_last;
  d.np = (h.nat + 63) / 64 * 64;
  d.conv = h.conv; d.plt = h.plt; d.slp = h.slp;
  for (int k = 0; k < 6 * kMaxMol; ++k) d.gbpr[k] = h.gbpr[k];
  d.parent = (const int16_t*)(base + o_parent);
  d.lev_atoms = (const int16_t*)(base + o_lev);
  d.lev_off = (const int32_t*)(base + o_levoff);
  d.type = (const int8_t*)(base + o_type);
  d.q = (const double*)(base + o_q);
  d.bl = (const double*)(base + o_bl);
  d.cba = (const double*)(base + o_cba);
  d.sba = (const double*)(base + o_sba);
  d.toat0 = (const double*)(base + o_to);
  d.tref = (const double*)(base + o_tref);
  d.rvdw = (const double*)(base + o_rv);
  d.sigma = (const double*)(base + o_sg);
  d.tvar = (const int16_t*)(base + o_tvar);
  d.var_atom = (const int16_t*)(base + o_va);
  d.var_base = (const int16_t*)(base + o_vb);
  d.var_axis0 = (const int16_t*)(base + o_vx);
  d.var_m1 = (const int16_t*)(base + o_m1);
  d.var_m2 = (const int16_t*)(base + o_m2);
  d.var_h1 = (const int16_t*)(base + o_h1);
  d.var_h2 = (const int16_t*)(base + o_h2);
  d.var_sign = (const double*)(base + o_vsign);
  d.var_e0 = (const double*)(base + o_e0);
  d.var_sg = (const double*)(base + o_vs);
  d.var_rn = (const double*)(base + o_vr);
  d.bits = (const uint64_t*)(base + o_bits);
  d.heavy = (const int16_t*)(base + o_heavy);
  d.heavy_of = (const int16_t*)(base + o_heavy_of);
  d.mol_of = (const int16_t*)(base + o_mol_of);
  d.pA = (const double*)(base + o_pA);
  d.pA14 = (const double*)(base + o_pA14);
  d.pC = (const double*)(base + o_pC);
  d.pHB = (const int32_t*)(base + o_pHB);
  d.spoint = (const double*)(base + o_sp);
  return off;
}

static void upload_system(DevSystem& ds, hipStream_t st) {
  std::vector<uint8_t> blob;
  DevSys probe;
  const size_t bytes = system_blob(ds.h, nullptr, blob, probe);
  char* base = static_cast<char*>(ds.mem.ensure(bytes));
  system_blob(ds.h, base, blob, ds.d);
  HIPCHK(hipMemcpyAsync(base, blob.data(), bytes, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
}

// Kernel timing with HIP events recorded on the engine's stream around each
// launch, plus the algorithmic work the kernels report (see Work).
enum KernelId { kKSweep = 0, kKExchange = 1, kKEnergy = 2, kKEsolan = 3, kKGradient = 4, kNKernels = 5 };
struct KernelStat {
  double ms = 0.0;
  int64_t launches = 0;
  double work[4] = {0, 0, 0, 0};
};
struct Profiler {
  bool on = false;
  bool phases = false;          // level 2: per-phase wall-clock of the sweep kernel
  double phase_ticks[kNPhase] = {};
  double tick_khz = 0.0;
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  KernelStat stat[kNKernels];
};

struct Engine {
  hipStream_t stream = nullptr;
  int device = -1;
  bool valid = false;           // system upload matches the host state
  DevSystem sys;
  DevMem d_sys, d_confsys, d_confoff, d_vlvr, d_out, d_xyz, d_err, d_work;
  DevMem g_cs, g_off, g_vl, g_out, g_grad;  // gradient batches
  DevMem g_xyz, g_rows, g_gradan;           // their analytic solvent: coordinates, energy rows, gradan
  SamplerBuffers smp;           // replica state of the samplers
  EnergyBatch* batch = nullptr;            // kept by energy_batch for repeated calls on the same set
  std::vector<EnergyBatch*> user_batches;  // smmp_batch_create handles running on this engine
  Profiler prof;
  uint64_t sys_gen = 0;         // counts uploads of the context's system
  bool capturing = false;       // launches are being recorded into a graph
  // captured sampler launches and what they were captured for
  hipGraphExec_t gexec = nullptr;
  std::vector<uint8_t> gkey;
  std::map<std::vector<int64_t>, int> tuned_groups;  // (system, R, precision, mode) -> G
  int32_t info[4] = {0, 0, 0, 0};                     // how the last sampler call ran
  MlpWork* mlp = nullptr;                             // surrogate tables and training workspace
  EsolanWork* esolan = nullptr;                       // analytic solvent workspace
  void release_batches();                             // drops 'batch', detaches user batches
  ~Engine() {
    release_batches();
    mlp_work_destroy(mlp);
    esolan_work_destroy(esolan);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    for (hipEvent_t ev : prof.pool) (void)hipEventDestroy(ev);
    for (auto& p : prof.pending) {
      (void)hipEventDestroy(p.second.first);
      (void)hipEventDestroy(p.second.second);
    }
    if (stream) (void)hipStreamDestroy(stream);
  }
};

static hipEvent_t prof_event(Engine& E) {
  if (!E.prof.pool.empty()) {
    hipEvent_t ev = E.prof.pool.back();
    E.prof.pool.pop_back();
    return ev;
  }
  hipEvent_t ev;
  HIPCHK(hipEventCreate(&ev));
  return ev;
}
// Brackets one launch: call with start=true before and start=false after.
static void prof_mark(Engine& E, int kid, bool start) {
  if (!E.prof.on || E.capturing) return;
  hipEvent_t ev = prof_event(E);
  HIPCHK(hipEventRecord(ev, E.stream));
  if (start) {
    E.prof.pending.push_back({kid, {ev, nullptr}});
  } else {
    E.prof.pending.back().second.second = ev;
  }
}
// Folds finished launches into the statistics (after a stream sync).
static void prof_collect(Engine& E) {
  if (E.prof.pending.empty()) return;
  HIPCHK(hipStreamSynchronize(E.stream));
  for (auto& p : E.prof.pending) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, p.second.first, p.second.second));
    E.prof.stat[p.first].ms += ms;
    E.prof.stat[p.first].launches += 1;
    E.prof.pool.push_back(p.second.first);
    E.prof.pool.push_back(p.second.second);
  }
  E.prof.pending.clear();
}
static void prof_add_work(Engine& E, int kid, const double* dev_work, int rows) {
  if (!E.prof.on || rows <= 0) return;
  std::vector<double> h(4 * (size_t)rows);
  HIPCHK(hipMemcpyAsync(h.data(), dev_work, sizeof(double) * h.size(), hipMemcpyDeviceToHost, E.stream));
  HIPCHK(hipStreamSynchronize(E.stream));
  for (int r = 0; r < rows; ++r)
    for (int k = 0; k < 4; ++k) E.prof.stat[kid].work[k] += h[4 * (size_t)r + k];
}

void engine_profile(Engine*& e, int enable);
int engine_profile_read(Engine*& e, int kernel, double* ms, int64_t* launches, double* work);

static Engine* make_engine() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n < 1) throw DeviceError("no HIP device available (hipGetDeviceCount)");
  auto* eng = new Engine();
  HIPCHK(hipGetDevice(&eng->device));
  HIPCHK(hipStreamCreateWithFlags(&eng->stream, hipStreamNonBlocking));
  return eng;
}

static Engine& engine_of(Engine*& e) {
  if (!e) e = make_engine();
  return *e;
}

void engine_destroy(Engine* e) { delete e; }
void engine_invalidate(Engine* e) {
  if (e) e->valid = false;
}

// Upload the context's system when the topology changed, or refresh the
// global parameters gbpr when only those changed.
static DevSystem& current_system(Engine& E, const State& s) {
  bool gb_same = E.valid && E.sys.h.nml == s.ntlml && E.sys.h.boxsize == s.boxsize;
  for (int m = 1; gb_same && m <= s.ntlml; ++m)
    for (int k = 0; k < 6; ++k)
      if (E.sys.h.gbpr[6 * (m - 1) + k] != s.gbpr(k + 1, m)) gb_same = false;
  if (E.valid && !gb_same && E.sys.h.nml == s.ntlml && s.ntlml >= 1) {  // only the global parameters moved
    for (int m = 1; m <= s.ntlml; ++m)
      for (int k = 0; k < 6; ++k) E.sys.h.gbpr[6 * (m - 1) + k] = E.sys.d.gbpr[6 * (m - 1) + k] = s.gbpr(k + 1, m);
    E.sys.h.boxsize = E.sys.d.boxsize = s.boxsize;
    DevSys* dp = static_cast<DevSys*>(E.d_sys.ensure(sizeof(DevSys)));
    HIPCHK(hipMemcpyAsync(dp, &E.sys.d, sizeof(DevSys), hipMemcpyHostToDevice, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
  } else if (!gb_same) {
    E.sys.h = build_host_sys(s);
    upload_system(E.sys, E.stream);
    ++E.sys_gen;
    DevSys* dp = static_cast<DevSys*>(E.d_sys.ensure(sizeof(DevSys)));
    HIPCHK(hipMemcpyAsync(dp, &E.sys.d, sizeof(DevSys), hipMemcpyHostToDevice, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    E.valid = true;
  }
  return E.sys;
}

constexpr size_t kLdsBytes = 160 * 1024;  // per workgroup on gfx950

// LDS of k_energy for a set of systems: the largest carve-up, with the pair
// bit rows staged in LDS when every system's copy fits.
struct LdsPlan {
  size_t bytes = 0;
  int bits_lds = 0;
};
static LdsPlan plan_energy_lds(const DevSys* const* sys, int n, int threads) {
  LdsPlan with, without;
  for (int i = 0; i < n; ++i) {
    const DevSys& d = *sys[i];
    with.bytes = std::max(with.bytes, energy_lds_bytes(d.np, d.npnt, threads / kWave, d.nheavy,
                                                       pair_bit_words(d.nat, d.words)));
    without.bytes = std::max(without.bytes, energy_lds_bytes(d.np, d.npnt, threads / kWave, d.nheavy, 0));
  }
  with.bits_lds = 1;
  return with.bytes <= kLdsBytes ? with : without;
}

static void check_lds(size_t bytes) {
  if (bytes > kLdsBytes) throw DeviceError("system too large for one workgroup's LDS");
}

// Launches k_energy on stream st for nblocks conformations: block b takes
// conformation d_order[b] (all of them in order when d_order is null).
static void launch_energy_on(hipStream_t st, const DevSys* d_systems, const int32_t* d_cs, const int64_t* d_off,
                             const double* d_vl, double* d_out, double* d_xyz, int xyz_stride, int nblocks,
                             const int32_t* d_order, LdsPlan plan, int precision, int32_t* d_err, double* d_work,
                             int threads = kEnergyThreads) {
  const size_t lds = plan.bytes;
  check_lds(lds);
  auto kern = precision == 0 ? k_energy<false> : k_energy<true>;
  HIPCHK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kern, dim3(nblocks), dim3(threads), lds, st, d_systems, d_cs, d_off, d_vl, d_out, d_xyz,
                     xyz_stride, d_err, d_work, plan.bits_lds, d_order);
  HIPCHK(hipGetLastError());
}

// Launches k_energy for nconf conformations of the systems at d_systems.
static void launch_energy(Engine& E, const DevSys* d_systems, const int32_t* d_cs, const int64_t* d_off,
                          const double* d_vl, double* d_out, double* d_xyz, int xyz_stride, int nconf, LdsPlan plan,
                          int precision) {
  int32_t* d_err = static_cast<int32_t*>(E.d_err.ensure(sizeof(int32_t)));
  HIPCHK(hipMemsetAsync(d_err, 0, sizeof(int32_t), E.stream));
  double* d_work = E.prof.on ? static_cast<double*>(E.d_work.ensure(sizeof(double) * 4 * nconf)) : nullptr;
  if (d_work) HIPCHK(hipMemsetAsync(d_work, 0, sizeof(double) * 4 * nconf, E.stream));
  prof_mark(E, kKEnergy, true);
  launch_energy_on(E.stream, d_systems, d_cs, d_off, d_vl, d_out, d_xyz, xyz_stride, nconf, nullptr, plan, precision,
                   d_err, d_work);
  prof_mark(E, kKEnergy, false);
  if (d_work) {
    prof_collect(E);
    prof_add_work(E, kKEnergy, d_work, nconf);
  }
}

void engine_profile(Engine*& e, int enable) {
  Engine& E = engine_of(e);
  prof_collect(E);
  E.prof.on = enable != 0;
  E.prof.phases = enable >= 2;
  for (auto& st : E.prof.stat) st = KernelStat();
  for (double& t : E.prof.phase_ticks) t = 0.0;
  if (E.prof.phases && E.prof.tick_khz == 0.0) {
    int khz = 0;
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, E.device));
    E.prof.tick_khz = khz;
  }
}

int engine_profile_phases(Engine*& e, double* out) {
  Engine& E = engine_of(e);
  for (int k = 0; k < kNPhase; ++k) {
    const double t = E.prof.phase_ticks[k];
    const bool count = k == kPhNFull || k == kPhCoreCycles || k == kPhSolAtoms || k == kPhSolContacts;
    out[k] = (count || E.prof.tick_khz <= 0.0) ? t : t / E.prof.tick_khz;
  }
  return SMMP_OK;
}

void engine_mlp_set(Engine*& e, const State& s, const float* theta) {
  Engine& E = engine_of(e);
  mlp_set_theta(E.mlp, E.stream, theta, mlp_residue_count(s));
}
void engine_mlp_get(Engine*& e, const State& s, float* theta) {
  Engine& E = engine_of(e);
  mlp_get_theta(E.mlp, E.stream, theta, mlp_residue_count(s));
}
double engine_mlp_grad(Engine*& e, const State& s, const double* vlvr, int nconf, const float* target, float* pred,
                       float* grad, int pairs) {
  Engine& E = engine_of(e);
  return mlp_grad(E.mlp, s, E.stream, vlvr, nconf, target, pred, grad, pairs);
}
void engine_mlp_fit(Engine*& e, const State& s, const double* vlvr, const float* target, int n, int epochs, int batch,
                    float lr, uint32_t seed, double* epoch_loss, int pairs) {
  Engine& E = engine_of(e);
  mlp_fit(E.mlp, s, E.stream, vlvr, target, n, epochs, batch, lr, seed, epoch_loss, pairs);
}

int engine_sampler_info(Engine*& e, int32_t* out) {
  Engine& E = engine_of(e);
  for (int k = 0; k < 4; ++k) out[k] = E.info[k];
  return SMMP_OK;
}

int engine_profile_read(Engine*& e, int kernel, double* ms, int64_t* launches, double* work) {
  Engine& E = engine_of(e);
  if (kernel < 0 || kernel >= kNKernels) return SMMP_EINVAL;
  prof_collect(E);
  const KernelStat& st = E.prof.stat[kernel];
  if (ms) *ms = st.ms;
  if (launches) *launches = st.launches;
  if (work)
    for (int k = 0; k < 4; ++k) work[k] = st.work[k];
  return SMMP_OK;
}

static int read_err(Engine& E) {
  int32_t h = 0;
  HIPCHK(hipMemcpyAsync(&h, E.d_err.p, sizeof(int32_t), hipMemcpyDeviceToHost, E.stream));
  HIPCHK(hipStreamSynchronize(E.stream));
  return h;
}

// Single energy of the context's current variables; updates the host state
// (components, coordinates) like energy.f does for the COMMON blocks.
// Analytic solvent (itysol < 0) of conformations whose coordinates k_energy
// left at d_xyz: adds eysl to their rows of d_out, drawing rotations from
// the twister of each conformation's context (esolan.f:199-200).
static void add_esolan(Engine& E, const DevSys* d_systems, const std::vector<const HostSys*>& hsys,
                       const std::vector<int32_t>& conf_sys, const int32_t* d_cs, int nconf, int stride,
                       const double* d_xyz, double* d_out, const std::function<double(int)>& draw) {
  esolan_energies(E.esolan, E.stream, d_systems, hsys, conf_sys, d_cs, nconf, stride, d_xyz, d_out, kEnergyOut, draw);
  if (E.prof.on) {
    const EsolanStats st = esolan_last_stats(E.esolan);
    KernelStat& k = E.prof.stat[kKEsolan];
    k.ms += st.ms;
    k.launches += st.passes + 1;
    for (int w = 0; w < 4; ++w) k.work[w] += st.work[w];
  }
}

static void energy_into_state(Engine& E, State& s) {
  DevSystem& ds = current_system(E, s);
  const int nvt = ds.h.nvt, nat = ds.h.nat;
  double* d_vl = static_cast<double*>(E.d_vlvr.ensure(sizeof(double) * nvt));
  int32_t* d_cs = static_cast<int32_t*>(E.d_confsys.ensure(sizeof(int32_t)));
  int64_t* d_off = static_cast<int64_t*>(E.d_confoff.ensure(sizeof(int64_t)));
  double* d_out = static_cast<double*>(E.d_out.ensure(sizeof(double) * kEnergyOut));
  double* d_xyz = static_cast<double*>(E.d_xyz.ensure(sizeof(double) * 3 * nat));
  const int32_t zero = 0;
  const int64_t zoff = 0;
  HIPCHK(hipMemcpyAsync(d_vl, s.vlvr.data(), sizeof(double) * nvt, hipMemcpyHostToDevice, E.stream));
  HIPCHK(hipMemcpyAsync(d_cs, &zero, sizeof(zero), hipMemcpyHostToDevice, E.stream));
  HIPCHK(hipMemcpyAsync(d_off, &zoff, sizeof(zoff), hipMemcpyHostToDevice, E.stream));
  launch_energy(E, static_cast<const DevSys*>(E.d_sys.p), d_cs, d_off, d_vl, d_out, d_xyz, nat, 1,
                [&] { const DevSys* one = &ds.d; return plan_energy_lds(&one, 1, kEnergyThreads); }(), 0);
  if (s.itysol < 0)
    add_esolan(E, static_cast<const DevSys*>(E.d_sys.p), {&ds.h}, {0}, d_cs, 1, nat, d_xyz, d_out,
               [&](int) { return host_grnd(s); });
  double o[kEnergyOut];
  std::vector<double> xyz(3 * (size_t)nat);
  HIPCHK(hipMemcpyAsync(o, d_out, sizeof o, hipMemcpyDeviceToHost, E.stream));
  HIPCHK(hipMemcpyAsync(xyz.data(), d_xyz, sizeof(double) * 3 * nat, hipMemcpyDeviceToHost, E.stream));
  HIPCHK(hipStreamSynchronize(E.stream));
  if (read_err(E)) throw Error("enysol> too many solvent neighbours for one atom");
  s.eysm = o[1];
  s.eyel = o[2];
  s.eyvw = o[3];
  s.eyhb = o[4];
  s.eyvr = o[5];
  s.eysl = o[6];
  s.asa = o[7];
  s.eysmi = o[8];
  s.eyeli = o[9];
  s.eyvwi = o[10];
  s.eyhbi = o[11];
  for (int a = 0; a < nat; ++a) {
    s.xat(a + 1) = xyz[3 * a + 0];
    s.yat(a + 1) = xyz[3 * a + 1];
    s.zat(a + 1) = xyz[3 * a + 2];
  }
  // keep the host torsion state in step with the variables (setvar.f)
  for (int m = 1; m <= s.ntlml; ++m) setvar(s, m, s.vlvr.data());
}

double engine_energy(Engine*& e, State& s) {
  Engine& E = engine_of(e);
  energy_into_state(E, s);
  return s.eysm + s.eysl + s.eysmi;  // energy.f:114
}

void engine_gradient(Engine*& e, State& s, int nconf, const double* vlvr, double* out_e, double* out_g) {
  gradient_supported(s);
  if (nconf <= 0) return;
  Engine& E = engine_of(e);
  DevSystem& ds = current_system(E, s);
  const int nvt = ds.h.nvt;
  const size_t nv = (size_t)nconf * nvt;
  int32_t* d_cs = static_cast<int32_t*>(E.g_cs.ensure(sizeof(int32_t) * nconf));
  int64_t* d_off = static_cast<int64_t*>(E.g_off.ensure(sizeof(int64_t) * nconf));
  double* d_vl = static_cast<double*>(E.g_vl.ensure(sizeof(double) * nv));
  double* d_out = static_cast<double*>(E.g_out.ensure(sizeof(double) * kGradOut * nconf));
  double* d_g = static_cast<double*>(E.g_grad.ensure(sizeof(double) * nv));
  std::vector<int32_t> cs(nconf, 0);
  std::vector<int64_t> off(nconf);
  for (int c = 0; c < nconf; ++c) off[c] = (int64_t)c * nvt;
  HIPCHK(hipMemcpyAsync(d_cs, cs.data(), sizeof(int32_t) * nconf, hipMemcpyHostToDevice, E.stream));
  HIPCHK(hipMemcpyAsync(d_off, off.data(), sizeof(int64_t) * nconf, hipMemcpyHostToDevice, E.stream));
  HIPCHK(hipMemcpyAsync(d_vl, vlvr, sizeof(double) * nv, hipMemcpyHostToDevice, E.stream));
  // the pair bits go to LDS only where three workgroups per CU still fit
  // (the kernel is latency-bound: occupancy first); else they are read
  // through the caches
  LdsPlan plan;
  {
    const DevSys& d = ds.d;
    const size_t with = energy_lds_bytes(d.np, d.npnt, kGradThreads / kWave, d.nheavy, pair_bit_words(d.nat, d.words));
    const size_t without = energy_lds_bytes(d.np, d.npnt, kGradThreads / kWave, d.nheavy, 0);
    plan.bits_lds = 3 * with <= kLdsBytes ? 1 : 0;
    plan.bytes = plan.bits_lds ? with : without;
  }
  check_lds(plan.bytes);
  double* d_work = E.prof.on ? static_cast<double*>(E.d_work.ensure(sizeof(double) * 4 * nconf)) : nullptr;
  HIPCHK(hipFuncSetAttribute((const void*)k_gradient, hipFuncAttributeMaxDynamicSharedMemorySize, (int)plan.bytes));
  prof_mark(E, kKGradient, true);
  // the analytic solvent (itysol < 0, one chain): coordinates out of the
  // kernel, esolan's term and gradient, its forces projected on the variables
  const bool sol = s.itysol < 0;
  const int nat = ds.h.nat;
  double* d_xyz = sol ? static_cast<double*>(E.g_xyz.ensure(sizeof(double) * 3 * (size_t)nat * nconf)) : nullptr;
  hipLaunchKernelGGL(k_gradient, dim3(nconf), dim3(kGradThreads), plan.bytes, E.stream,
                     static_cast<const DevSys*>(E.d_sys.p), d_cs, d_off, d_vl, d_out, d_g, d_work, plan.bits_lds,
                     d_xyz, nat);
  HIPCHK(hipGetLastError());
  prof_mark(E, kKGradient, false);
  if (d_work) {
    prof_collect(E);
    prof_add_work(E, kKGradient, d_work, nconf);
  }
  std::vector<double> rows;
  if (sol) {
    double* d_rows = static_cast<double*>(E.g_rows.ensure(sizeof(double) * kEnergyOut * (size_t)nconf));
    double* d_ga = static_cast<double*>(E.g_gradan.ensure(sizeof(double) * 3 * (size_t)nat * nconf));
    HIPCHK(hipMemsetAsync(d_rows, 0, sizeof(double) * kEnergyOut * (size_t)nconf, E.stream));
    esolan_energies(E.esolan, E.stream, static_cast<const DevSys*>(E.d_sys.p), {&ds.h}, cs, d_cs, nconf, nat, d_xyz,
                    d_rows, kEnergyOut, [&](int) { return host_grnd(s); }, d_ga);
    hipLaunchKernelGGL(k_sol_project, dim3(nconf), dim3(kGradThreads), 0, E.stream,
                       static_cast<const DevSys*>(E.d_sys.p), d_cs, d_off, d_xyz, d_ga, nat, d_g);
    HIPCHK(hipGetLastError());
    rows.resize((size_t)kEnergyOut * nconf);
    HIPCHK(hipMemcpyAsync(rows.data(), d_rows, sizeof(double) * rows.size(), hipMemcpyDeviceToHost, E.stream));
  }
  HIPCHK(hipMemcpyAsync(out_e, d_out, sizeof(double) * kGradOut * nconf, hipMemcpyDeviceToHost, E.stream));
  HIPCHK(hipMemcpyAsync(out_g, d_g, sizeof(double) * nv, hipMemcpyDeviceToHost, E.stream));
  HIPCHK(hipStreamSynchronize(E.stream));
  if (sol)  // gradient.f:47-48: eysm carries eysl
    for (int c = 0; c < nconf; ++c) out_e[(size_t)kGradOut * c] += rows[(size_t)kEnergyOut * c + 6];
}

double engine_enysol(Engine*& e, State& s) {
  Engine& E = engine_of(e);
  if (s.itysol < 0) throw Error("enysol> the context uses the analytic solvent (itysol < 0): call energy()");
  energy_into_state(E, s);
  return s.eysl;
}

// ------------------------------------------------ resident energy batch
//
// The systems of a batch are uploaded once and its conformations and results
// stay on the device between calls; a system is rebuilt only when its
// context's topology changed (State::gen) and refreshed when only its global
// parameters moved.  Conformations are grouped by the LDS their system needs:
// class 1 holds the systems that fit two 512-thread workgroups per CU, class
// 0 the rest, which run one 1024-thread workgroup per CU (the register budget
// of k_energy allows four waves per SIMD either way).  Each class is one
// launch with its own LDS size, the two on separate streams so that they fill
// the CUs together; within a class the largest systems go first.
constexpr int kBatchClasses = 2;

constexpr int kXcdRun = 16;  // consecutive conformations of the sorted list per XCD

// One launch of k_energy over the conformations order[begin, begin+count).
struct EnergyClass {
  int begin = 0, count = 0, threads = kEnergyThreads;
  LdsPlan plan;
};

// XCD-aware block order of one launch: the hardware deals workgroups to the
// 8 XCDs in turn (block b to XCD b mod 8), so block x + 8 p takes entry
// r 8k + x k + i of the sorted list (p = r k + i): each XCD gets runs of k
// consecutive entries (the conformations of one system, whose pair bits then
// stay in that XCD's L2), and rounds of 8 runs keep the largest systems
// first on every XCD.  The tail past the last full round keeps its order.
static void xcd_interleave(int32_t* seg, int n, int k) {
  const int round = 8 * k, full = n / round * round;
  const std::vector<int32_t> src(seg, seg + full);
  for (int b = 0; b < full; ++b) {
    const int x = b % 8, p = b / 8;
    seg[b] = src[(p / k) * round + x * k + p % k];
  }
}

// The classes of a set of conformations and the order that groups them
// (largest systems first within a class, runs of one system per XCD).
static std::vector<EnergyClass> plan_classes(const std::vector<DevSys>& dsys, const std::vector<int32_t>& conf_sys,
                                             std::vector<int32_t>& order) {
  const int nsys = (int)dsys.size(), nconf = (int)conf_sys.size();
  // class of each system: 1 when it fits half of the LDS at 512 threads
  // (with or without its pair bits), else 0 (1024 threads)
  const int nw = kEnergyThreads / kWave;
  auto waves_of = [](int k) { return (k == 1 ? kEnergyThreads : kEnergyThreadsMax) / kWave; };
  std::vector<int> cls(nsys);
  for (int i = 0; i < nsys; ++i) {
    const DevSys& d = dsys[i];
    const size_t without = energy_lds_bytes(d.np, d.npnt, nw, d.nheavy, 0);
    check_lds(without);
    cls[i] = without <= kLdsBytes / 2 ? 1 : 0;
  }
  order.resize(nconf);
  for (int c = 0; c < nconf; ++c) order[c] = c;
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    const int sa = conf_sys[a], sb = conf_sys[b];
    if (cls[sa] != cls[sb]) return cls[sa] < cls[sb];
    return dsys[sa].nat > dsys[sb].nat;
  });
  std::vector<EnergyClass> classes;
  for (int k = 0; k < kBatchClasses; ++k) {
    EnergyClass C;
    C.begin = -1;
    size_t with = 0, without = 0;
    for (int j = 0; j < nconf; ++j) {
      const int si = conf_sys[order[j]];
      if (cls[si] != k) continue;
      if (C.begin < 0) C.begin = j;
      ++C.count;
      const DevSys& d = dsys[si];
      with = std::max(with, energy_lds_bytes(d.np, d.npnt, waves_of(k), d.nheavy, pair_bit_words(d.nat, d.words)));
      without = std::max(without, energy_lds_bytes(d.np, d.npnt, waves_of(k), d.nheavy, 0));
    }
    if (C.count == 0) continue;
    check_lds(without);
    C.threads = waves_of(k) * kWave;
    const size_t cap = k == 1 ? kLdsBytes / 2 : kLdsBytes;
    C.plan.bits_lds = with <= cap ? 1 : 0;
    C.plan.bytes = with <= cap ? with : without;
    classes.push_back(C);
  }
  for (const EnergyClass& C : classes) xcd_interleave(order.data() + C.begin, C.count, kXcdRun);
  return classes;
}

struct EnergyBatch {
  Engine* E = nullptr;  // null once the engine is gone
  std::vector<State*> states;
  std::vector<uint64_t> gens;
  std::vector<std::unique_ptr<DevSystem>> sys;
  std::vector<int32_t> conf_sys;
  std::vector<int64_t> off;
  int64_t nvals = 0;
  int nconf = 0, max_nat = 0;
  bool analytic = false;
  bool loaded = false;  // conformations on the device match the current layout
  std::vector<EnergyClass> classes;
  // measured launch shapes per precision (engine_batch_tune; empty: the
  // classes' own), dropped when the layout changes
  std::vector<EnergyClass> tuned[2];
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  DevMem d_sys, d_cs, d_off, d_order, d_vl, d_out, d_xyz;
  ~EnergyBatch() {
    if (join) (void)hipEventDestroy(join);
    if (fork) (void)hipEventDestroy(fork);
    if (side) (void)hipStreamDestroy(side);
  }
};

void Engine::release_batches() {
  delete batch;
  batch = nullptr;
  for (EnergyBatch* b : user_batches) b->E = nullptr;
  user_batches.clear();
}

static Engine& batch_engine(EnergyBatch* b) {
  if (!b || !b->E) throw Error("batch> the batch's first context was destroyed");
  return *b->E;
}

// (Re)builds what depends on the systems' topology: offsets, classes, order.
static void batch_layout(EnergyBatch& B) {
  Engine& E = *B.E;
  const int nsys = (int)B.sys.size();
  std::vector<DevSys> dsys(nsys);
  B.max_nat = 0;
  for (int i = 0; i < nsys; ++i) {
    dsys[i] = B.sys[i]->d;
    B.max_nat = std::max(B.max_nat, B.sys[i]->h.nat);
  }
  B.off.resize(B.nconf);
  B.nvals = 0;
  for (int c = 0; c < B.nconf; ++c) {
    B.off[c] = B.nvals;
    B.nvals += dsys[B.conf_sys[c]].nvt;
  }
  std::vector<int32_t> order;
  B.classes = plan_classes(dsys, B.conf_sys, order);
  B.tuned[0].clear();
  B.tuned[1].clear();
  DevSys* d_systems = static_cast<DevSys*>(B.d_sys.ensure(sizeof(DevSys) * nsys));
  int32_t* d_cs = static_cast<int32_t*>(B.d_cs.ensure(sizeof(int32_t) * std::max(B.nconf, 1)));
  int64_t* d_off = static_cast<int64_t*>(B.d_off.ensure(sizeof(int64_t) * std::max(B.nconf, 1)));
  int32_t* d_order = static_cast<int32_t*>(B.d_order.ensure(sizeof(int32_t) * std::max(B.nconf, 1)));
  B.d_vl.ensure(sizeof(double) * std::max<int64_t>(B.nvals, 1));
  B.d_out.ensure(sizeof(double) * kEnergyOut * std::max(B.nconf, 1));
  if (B.analytic) B.d_xyz.ensure(sizeof(double) * 3 * (size_t)B.max_nat * std::max(B.nconf, 1));
  HIPCHK(hipMemcpyAsync(d_systems, dsys.data(), sizeof(DevSys) * nsys, hipMemcpyHostToDevice, E.stream));
  if (B.nconf > 0) {
    HIPCHK(hipMemcpyAsync(d_cs, B.conf_sys.data(), sizeof(int32_t) * B.nconf, hipMemcpyHostToDevice, E.stream));
    HIPCHK(hipMemcpyAsync(d_off, B.off.data(), sizeof(int64_t) * B.nconf, hipMemcpyHostToDevice, E.stream));
    HIPCHK(hipMemcpyAsync(d_order, order.data(), sizeof(int32_t) * B.nconf, hipMemcpyHostToDevice, E.stream));
  }
  HIPCHK(hipStreamSynchronize(E.stream));
  B.loaded = false;
}

// Brings the device systems in line with the contexts: rebuilds a system
// whose topology changed (and then the layout), refreshes global parameters.
static void batch_refresh(EnergyBatch& B) {
  Engine& E = *B.E;
  bool topo = false, params = false;
  for (size_t i = 0; i < B.sys.size(); ++i) {
    State& s = *B.states[i];
    DevSystem& ds = *B.sys[i];
    if (s.ntlml < 1) throw Error("energy_batch> context without a molecule");
    if (B.gens[i] != s.gen || (s.itysol < 0) != B.analytic) {
      if ((s.itysol < 0) != B.analytic)
        throw Error("energy_batch> systems with the analytic solvent (itysol < 0) and others in one batch");
      ds.h = build_host_sys(s);
      upload_system(ds, E.stream);
      B.gens[i] = s.gen;
      topo = true;
      continue;
    }
    bool same = ds.h.nml == s.ntlml && ds.h.boxsize == s.boxsize;
    for (int m = 1; same && m <= s.ntlml; ++m)
      for (int k = 0; k < 6; ++k)
        if (ds.h.gbpr[6 * (m - 1) + k] != s.gbpr(k + 1, m)) same = false;
    if (!same) {
      for (int m = 1; m <= s.ntlml; ++m)
        for (int k = 0; k < 6; ++k) ds.h.gbpr[6 * (m - 1) + k] = ds.d.gbpr[6 * (m - 1) + k] = s.gbpr(k + 1, m);
      ds.h.boxsize = ds.d.boxsize = s.boxsize;
      params = true;
    }
  }
  if (topo) {
    batch_layout(B);
  } else if (params) {
    std::vector<DevSys> dsys(B.sys.size());
    for (size_t i = 0; i < B.sys.size(); ++i) dsys[i] = B.sys[i]->d;
    HIPCHK(hipMemcpyAsync(B.d_sys.p, dsys.data(), sizeof(DevSys) * dsys.size(), hipMemcpyHostToDevice, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
  }
}

static EnergyBatch* batch_new(Engine& E, std::vector<State*>& sys, const int32_t* conf_sys, int nconf) {
  auto B = std::make_unique<EnergyBatch>();
  B->E = &E;
  const int nsys = (int)sys.size();
  B->analytic = sys[0]->itysol < 0;
  for (int i = 1; i < nsys; ++i)
    if ((sys[i]->itysol < 0) != B->analytic)
      throw Error("energy_batch> systems with the analytic solvent (itysol < 0) and others in one batch");
  B->states = sys;
  B->nconf = nconf;
  B->conf_sys.assign(conf_sys, conf_sys + nconf);
  for (int i = 0; i < nsys; ++i) {
    auto p = std::make_unique<DevSystem>();
    p->h = build_host_sys(*sys[i]);
    upload_system(*p, E.stream);
    B->sys.push_back(std::move(p));
    B->gens.push_back(sys[i]->gen);
  }
  batch_layout(*B);
  return B.release();
}

EnergyBatch* engine_batch_create(Engine*& e, std::vector<State*>& sys, const int32_t* conf_sys, int nconf) {
  Engine& E = engine_of(e);
  EnergyBatch* b = batch_new(E, sys, conf_sys, nconf);
  E.user_batches.push_back(b);
  return b;
}

void engine_batch_destroy(EnergyBatch* b) {
  if (!b) return;
  if (b->E) {
    auto& v = b->E->user_batches;
    v.erase(std::remove(v.begin(), v.end(), b), v.end());
  }
  delete b;
}

int64_t engine_batch_values(const EnergyBatch* b) { return b->nvals; }

void engine_batch_load(EnergyBatch* b, const double* vlvr, int64_t nvals) {
  Engine& E = batch_engine(b);
  batch_refresh(*b);
  if (nvals != b->nvals)
    throw Error("energy_batch> " + std::to_string(nvals) + " values given, the batch's systems take " +
                std::to_string(b->nvals));
  if (b->nvals > 0)
    HIPCHK(hipMemcpyAsync(b->d_vl.p, vlvr, sizeof(double) * b->nvals, hipMemcpyHostToDevice, E.stream));
  b->loaded = true;
}

void engine_batch_run(EnergyBatch* b, int precision) {
  Engine& E = batch_engine(b);
  EnergyBatch& B = *b;
  batch_refresh(B);
  if (!B.loaded) throw Error("energy_batch> load the conformations (again) after a topology change");
  if (B.nconf == 0) return;
  const DevSys* d_systems = static_cast<const DevSys*>(B.d_sys.p);
  const int32_t* d_cs = static_cast<const int32_t*>(B.d_cs.p);
  const int64_t* d_off = static_cast<const int64_t*>(B.d_off.p);
  const int32_t* d_order = static_cast<const int32_t*>(B.d_order.p);
  const double* d_vl = static_cast<const double*>(B.d_vl.p);
  double* d_out = static_cast<double*>(B.d_out.p);
  double* d_xyz = B.analytic ? static_cast<double*>(B.d_xyz.p) : nullptr;
  int32_t* d_err = static_cast<int32_t*>(E.d_err.ensure(sizeof(int32_t)));
  HIPCHK(hipMemsetAsync(d_err, 0, sizeof(int32_t), E.stream));
  double* d_work = E.prof.on ? static_cast<double*>(E.d_work.ensure(sizeof(double) * 4 * B.nconf)) : nullptr;
  if (d_work) HIPCHK(hipMemsetAsync(d_work, 0, sizeof(double) * 4 * B.nconf, E.stream));
  prof_mark(E, kKEnergy, true);
  const std::vector<EnergyClass>& classes = B.tuned[precision].empty() ? B.classes : B.tuned[precision];
  const bool two = classes.size() > 1;
  if (two) {
    if (!B.side) {
      HIPCHK(hipStreamCreateWithFlags(&B.side, hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&B.fork, hipEventDisableTiming));
      HIPCHK(hipEventCreateWithFlags(&B.join, hipEventDisableTiming));
    }
    HIPCHK(hipEventRecord(B.fork, E.stream));
    HIPCHK(hipStreamWaitEvent(B.side, B.fork, 0));
  }
  for (size_t k = 0; k < classes.size(); ++k) {
    const EnergyClass& C = classes[k];
    launch_energy_on(k == 0 ? E.stream : B.side, d_systems, d_cs, d_off, d_vl, d_out, d_xyz, B.max_nat, C.count,
                     d_order + C.begin, C.plan, precision, d_err, d_work, C.threads);
  }
  if (two) {
    HIPCHK(hipEventRecord(B.join, B.side));
    HIPCHK(hipStreamWaitEvent(E.stream, B.join, 0));
  }
  prof_mark(E, kKEnergy, false);
  if (d_work) {
    prof_collect(E);
    prof_add_work(E, kKEnergy, d_work, B.nconf);
  }
  if (B.analytic) {
    // rotations draw from each conformation's context, conformations in order
    std::vector<const HostSys*> hsys(B.sys.size());
    for (size_t i = 0; i < B.sys.size(); ++i) hsys[i] = &B.sys[i]->h;
    add_esolan(E, d_systems, hsys, B.conf_sys, d_cs, B.nconf, B.max_nat, d_xyz, d_out,
               [&](int c) { return host_grnd(*B.states[B.conf_sys[c]]); });
  }
}

// Measures, for each class of the batch alone, launches of 256, 512 and
// 1024 threads with the pair bits in LDS (where they fit) or read through
// the caches, and keeps the fastest for this precision (the best of two
// timed launches each, on the batch's loaded conformations).  The shapes
// change the order of the FP64 sums (rounding only).  out (may be null):
// threads and bits flag per class, two ints each.
int engine_batch_tune(EnergyBatch* b, int precision, int32_t* out) {
  Engine& E = batch_engine(b);
  EnergyBatch& B = *b;
  batch_refresh(B);
  if (!B.loaded) throw Error("energy_batch> load the conformations before tuning");
  if (B.nconf == 0) return 0;
  std::vector<DevSys> dsys(B.sys.size());
  for (size_t i = 0; i < B.sys.size(); ++i) dsys[i] = B.sys[i]->d;
  std::vector<int32_t> order(B.nconf);
  HIPCHK(hipMemcpy(order.data(), B.d_order.p, sizeof(int32_t) * B.nconf, hipMemcpyDeviceToHost));
  const DevSys* d_systems = static_cast<const DevSys*>(B.d_sys.p);
  int32_t* d_err = static_cast<int32_t*>(E.d_err.ensure(sizeof(int32_t)));
  HIPCHK(hipMemsetAsync(d_err, 0, sizeof(int32_t), E.stream));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  std::vector<EnergyClass> best = B.classes;
  try {
    for (size_t k = 0; k < B.classes.size(); ++k) {
      const EnergyClass& C0 = B.classes[k];
      float best_ms = -1.0f;
      for (int threads : {256, 512, 1024}) {
        size_t with = 0, without = 0;
        for (int j = 0; j < C0.count; ++j) {
          const DevSys& d = dsys[B.conf_sys[order[C0.begin + j]]];
          with = std::max(with, energy_lds_bytes(d.np, d.npnt, threads / kWave, d.nheavy, pair_bit_words(d.nat, d.words)));
          without = std::max(without, energy_lds_bytes(d.np, d.npnt, threads / kWave, d.nheavy, 0));
        }
        for (int bits : {1, 0}) {
          EnergyClass C = C0;
          C.threads = threads;
          C.plan.bits_lds = bits;
          C.plan.bytes = bits ? with : without;
          if (C.plan.bytes > kLdsBytes) continue;
          float ms = 0.0f;
          for (int rep = 0; rep < 3; ++rep) {  // the first launch warms up
            HIPCHK(hipEventRecord(e0, E.stream));
            launch_energy_on(E.stream, d_systems, static_cast<const int32_t*>(B.d_cs.p),
                             static_cast<const int64_t*>(B.d_off.p), static_cast<const double*>(B.d_vl.p),
                             static_cast<double*>(B.d_out.p), B.analytic ? static_cast<double*>(B.d_xyz.p) : nullptr,
                             B.max_nat, C.count, static_cast<const int32_t*>(B.d_order.p) + C.begin, C.plan,
                             precision, d_err, nullptr, C.threads);
            HIPCHK(hipEventRecord(e1, E.stream));
            HIPCHK(hipEventSynchronize(e1));
            float t = 0.0f;
            HIPCHK(hipEventElapsedTime(&t, e0, e1));
            if (rep == 1 || (rep == 2 && t < ms)) ms = t;
          }
          if (best_ms < 0.0f || ms < best_ms) {
            best_ms = ms;
            best[k] = C;
          }
        }
      }
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  HIPCHK(hipEventDestroy(e0));
  HIPCHK(hipEventDestroy(e1));
  B.tuned[precision] = best;
  if (out)
    for (size_t k = 0; k < best.size(); ++k) {
      out[2 * k] = best[k].threads;
      out[2 * k + 1] = best[k].plan.bits_lds;
    }
  return (int)best.size();
}

void engine_batch_read(EnergyBatch* b, double* out) {
  Engine& E = batch_engine(b);
  EnergyBatch& B = *b;
  if (B.nconf == 0) return;
  std::vector<double> full((size_t)kEnergyOut * B.nconf);
  HIPCHK(hipMemcpyAsync(full.data(), B.d_out.p, sizeof(double) * full.size(), hipMemcpyDeviceToHost, E.stream));
  HIPCHK(hipStreamSynchronize(E.stream));
  if (read_err(E)) throw Error("enysol> too many solvent neighbours for one atom");
  for (int c = 0; c < B.nconf; ++c)
    for (int k = 0; k < 8; ++k) out[8 * (size_t)c + k] = full[(size_t)kEnergyOut * c + k];
}

// energy_batch: the engine keeps the last batch and reuses it while the
// contexts and the conformations' systems are the same.
int engine_energy_batch(Engine*& e, std::vector<State*>& sys, const int32_t* conf_sys, int nconf, const double* vlvr,
                        int precision, double* out) {
  Engine& E = engine_of(e);
  if (nconf == 0) return SMMP_OK;
  EnergyBatch* B = E.batch;
  const bool same = B && B->states == sys && B->nconf == nconf &&
                    std::equal(B->conf_sys.begin(), B->conf_sys.end(), conf_sys);
  if (!same) {
    delete E.batch;
    E.batch = nullptr;
    E.batch = batch_new(E, sys, conf_sys, nconf);
  }
  // a context whose topology changed since the last call is rebuilt first,
  // so the size check below sees the values its systems take now
  batch_refresh(*E.batch);
  engine_batch_load(E.batch, vlvr, E.batch->nvals);
  engine_batch_run(E.batch, precision);
  engine_batch_read(E.batch, out);
  return SMMP_OK;
}

// ------------------------------------------- energies without allocation
//
// engine_energy_batch_noalloc computes what engine_energy_batch does (the
// same kernels, classes and thread counts, so the results are identical)
// but allocates no device memory: the systems, the conformation tables, the
// values, the results and the error flag of a call are laid out in
// 'scratch', one file-scope array per device, which the device holds until
// the process ends.  Calls on one device share that array: concurrent calls
// on the same GPU race, and each call overwrites what the last one left.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;  // 32 MiB per device
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};  // device address of 'scratch', per device

// The current device's copy of 'scratch', looked up once per device with
// that device current.  Host code passes this address, never the symbol.
static char* noalloc_scratch() {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw DeviceError("energy_batch_noalloc> device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
  }
  return static_cast<char*>(g_scratch_ptr[dev]);
}

int engine_energy_batch_noalloc(std::vector<State*>& sys, const int32_t* conf_sys, int nconf, const double* vlvr,
                                int precision, double* out) {
  if (nconf == 0) return SMMP_OK;
  char* base = noalloc_scratch();
  const int nsys = (int)sys.size();
  std::vector<HostSys> hs(nsys);
  for (int i = 0; i < nsys; ++i) {
    if (sys[i]->itysol < 0)
      throw Error("energy_batch_noalloc> the analytic solvent (itysol < 0) needs its own workspace: "
                  "use smmp_energy_batch");
    hs[i] = build_host_sys(*sys[i]);
  }
  // layout of this call in the scratch array
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = (off + 255) & ~size_t(255);
    off = o + std::max<size_t>(bytes, 1);
    return o;
  };
  std::vector<DevSys> dsys(nsys);
  std::vector<std::vector<uint8_t>> blobs(nsys);
  std::vector<size_t> o_sys(nsys);
  for (int i = 0; i < nsys; ++i) o_sys[i] = take(system_blob(hs[i], nullptr, blobs[i], dsys[i]));
  std::vector<int32_t> cs(conf_sys, conf_sys + nconf);
  std::vector<int64_t> coff(nconf);
  int64_t nvals = 0;
  for (int c = 0; c < nconf; ++c) {
    coff[c] = nvals;
    nvals += hs[cs[c]].nvt;
  }
  const size_t o_dsys = take(sizeof(DevSys) * nsys), o_cs = take(sizeof(int32_t) * nconf),
               o_off = take(sizeof(int64_t) * nconf), o_order = take(sizeof(int32_t) * nconf),
               o_vl = take(sizeof(double) * nvals), o_out = take(sizeof(double) * kEnergyOut * nconf),
               o_err = take(sizeof(int32_t));
  if (off > kNoAllocScratchBytes)
    throw Error("energy_batch_noalloc> this batch needs " + std::to_string(off) + " bytes of device scratch; the " +
                "library holds " + std::to_string(kNoAllocScratchBytes) + " per device (use smmp_energy_batch)");
  for (int i = 0; i < nsys; ++i) system_blob(hs[i], base + o_sys[i], blobs[i], dsys[i]);
  std::vector<int32_t> order;
  const std::vector<EnergyClass> classes = plan_classes(dsys, cs, order);
  hipStream_t st = nullptr;  // the device's null stream: nothing to create
  auto put = [&](size_t o, const void* src, size_t bytes) {
    if (bytes) HIPCHK(hipMemcpyAsync(base + o, src, bytes, hipMemcpyHostToDevice, st));
  };
  for (int i = 0; i < nsys; ++i) put(o_sys[i], blobs[i].data(), blobs[i].size());
  const int32_t zero = 0;
  put(o_dsys, dsys.data(), sizeof(DevSys) * nsys);
  put(o_cs, cs.data(), sizeof(int32_t) * nconf);
  put(o_off, coff.data(), sizeof(int64_t) * nconf);
  put(o_order, order.data(), sizeof(int32_t) * nconf);
  put(o_vl, vlvr, sizeof(double) * nvals);
  put(o_err, &zero, sizeof zero);  // the error flag starts at zero
  const DevSys* d_systems = reinterpret_cast<const DevSys*>(base + o_dsys);
  const int32_t* d_order = reinterpret_cast<const int32_t*>(base + o_order);
  for (const EnergyClass& C : classes)
    launch_energy_on(st, d_systems, reinterpret_cast<const int32_t*>(base + o_cs),
                     reinterpret_cast<const int64_t*>(base + o_off), reinterpret_cast<const double*>(base + o_vl),
                     reinterpret_cast<double*>(base + o_out), nullptr, 0, C.count, d_order + C.begin, C.plan,
                     precision, reinterpret_cast<int32_t*>(base + o_err), nullptr, C.threads);
  std::vector<double> full((size_t)kEnergyOut * nconf);
  int32_t err = 0;
  HIPCHK(hipMemcpyAsync(full.data(), base + o_out, sizeof(double) * full.size(), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&err, base + o_err, sizeof err, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (err) throw Error("enysol> too many solvent neighbours for one atom");
  for (int c = 0; c < nconf; ++c)
    for (int k = 0; k < 8; ++k) out[8 * (size_t)c + k] = full[(size_t)kEnergyOut * c + k];
  return SMMP_OK;
}

// --------------------------------------------------- host MT (twister.f)

double host_grnd(State& s) {
  if (s.mti >= 624) {
    if (s.mti == 625) {
      s.mt(0) = 4357;
      for (int i = 1; i < 624; ++i) s.mt(i) = (int32_t)(69069u * (uint32_t)s.mt(i - 1));
    }
    uint32_t* mt = reinterpret_cast<uint32_t*>(s.mt.data());
    mt_twist(mt);
    s.mti = 0;
  }
  uint32_t y = (uint32_t)s.mt(s.mti++);
  return (double)mt_temper(y) / 4294967295.0;
}

const char* engine_build_info() {
  return "smmp_mi355x: gfx950 HIP engine; energy kernel 512 threads, 64x64 pair tiles, FP64 and mixed FP32/FP64";
}

}  // namespace smmp

#include "sampler_host.inc"
