// HIP engine: device mirrors of contexts and the kernel launches behind the
// C-ABI.  Kernels: k_energy (full energy per conformation), the sweep and
// exchange kernels of sampler.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../host/devtopo.hpp"
#include "../host/engine.hpp"
#include "common.hpp"
#include "energy.hpp"
#include "esolan_host.hpp"
#include "mlp_host.hpp"
#include "sampler.hpp"

namespace smmp {

#define HIPCHK(x)                                                                                   \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess) throw DeviceError(std::string(#x) + ": " + hipGetErrorString(e_));        \
  } while (0)

constexpr int kEnergyThreads = 512;
// k_energy values per conformation: total, eysm, eyel, eyvw, eyhb, eyvr,
// eysl, asa, then between chains eysmi, eyeli, eyvwi, eyhbi
constexpr int kEnergyOut = 12;

// ------------------------------------------------------------- kernels

template <bool MIXED>
__global__ void __launch_bounds__(kEnergyThreads) k_energy(const DevSys* __restrict__ systems,
                                                           const int32_t* __restrict__ conf_sys,
                                                           const int64_t* __restrict__ conf_off,
                                                           const double* __restrict__ vlvr, double* __restrict__ out,
                                                           double* __restrict__ xyz_out, int xyz_stride,
                                                           int32_t* __restrict__ err, double* __restrict__ work,
                                                           int bits_lds) {
  extern __shared__ double smem[];
  const int c = blockIdx.x;
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
  DevMem sweep_base, tune_vl, tune_mt, tune_gb, gb, da_stats, snap, snap_e, obs;
};

// One system uploaded to the device: a single allocation holding every
// array, plus the DevSys with the device pointers.
struct DevSystem {
  HostSys h;
  DevSys d;
  DevMem mem;
};

static void upload_system(DevSystem& ds, hipStream_t st) {
  const HostSys& h = ds.h;
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
  std::vector<uint8_t> blob(off, 0);
  for (auto& pr : parts)
    if (pr.second.second) std::memcpy(blob.data() + pr.first, pr.second.first, pr.second.second);
  char* base = static_cast<char*>(ds.mem.ensure(off));
  HIPCHK(hipMemcpyAsync(base, blob.data(), off, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  DevSys& d = ds.d;
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
}

// Kernel timing with HIP events recorded on the engine's stream around each
// launch, plus the algorithmic work the kernels report (see Work).
enum KernelId { kKSweep = 0, kKExchange = 1, kKEnergy = 2, kKEsolan = 3, kNKernels = 4 };
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
  SamplerBuffers smp;           // replica state of the samplers
  std::vector<std::unique_ptr<DevSystem>> batch_sys;  // energy_batch systems
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
  ~Engine() {
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

// Launches k_energy for nconf conformations of the systems at d_systems.
static void launch_energy(Engine& E, const DevSys* d_systems, const int32_t* d_cs, const int64_t* d_off,
                          const double* d_vl, double* d_out, double* d_xyz, int xyz_stride, int nconf, LdsPlan plan,
                          int precision) {
  const size_t lds = plan.bytes;
  check_lds(lds);
  int32_t* d_err = static_cast<int32_t*>(E.d_err.ensure(sizeof(int32_t)));
  HIPCHK(hipMemsetAsync(d_err, 0, sizeof(int32_t), E.stream));
  double* d_work = E.prof.on ? static_cast<double*>(E.d_work.ensure(sizeof(double) * 4 * nconf)) : nullptr;
  if (d_work) HIPCHK(hipMemsetAsync(d_work, 0, sizeof(double) * 4 * nconf, E.stream));
  prof_mark(E, kKEnergy, true);
  if (precision == 0) {
    HIPCHK(hipFuncSetAttribute((const void*)k_energy<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_energy<false>, dim3(nconf), dim3(kEnergyThreads), lds, E.stream, d_systems, d_cs, d_off, d_vl,
                       d_out, d_xyz, xyz_stride, d_err, d_work, plan.bits_lds);
  } else {
    HIPCHK(hipFuncSetAttribute((const void*)k_energy<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_energy<true>, dim3(nconf), dim3(kEnergyThreads), lds, E.stream, d_systems, d_cs, d_off, d_vl,
                       d_out, d_xyz, xyz_stride, d_err, d_work, plan.bits_lds);
  }
  HIPCHK(hipGetLastError());
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

void engine_mlp_set(Engine*& e, const float* theta) {
  Engine& E = engine_of(e);
  mlp_set_theta(E.mlp, E.stream, theta);
}
void engine_mlp_get(Engine*& e, float* theta) {
  Engine& E = engine_of(e);
  mlp_get_theta(E.mlp, E.stream, theta);
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

double engine_enysol(Engine*& e, State& s) {
  Engine& E = engine_of(e);
  if (s.itysol < 0) throw Error("enysol> the context uses the analytic solvent (itysol < 0): call energy()");
  energy_into_state(E, s);
  return s.eysl;
}

int engine_energy_batch(Engine*& e, std::vector<State*>& sys, const int32_t* conf_sys, int nconf, const double* vlvr,
                        int precision, double* out) {
  Engine& E = engine_of(e);
  if (nconf == 0) return SMMP_OK;
  const int nsys = (int)sys.size();
  const bool analytic = sys[0]->itysol < 0;
  for (int i = 1; i < nsys; ++i)
    if ((sys[i]->itysol < 0) != analytic)
      throw Error("energy_batch> systems with the analytic solvent (itysol < 0) and others in one batch");
  E.batch_sys.clear();
  std::vector<DevSys> dsys(nsys);
  std::vector<const HostSys*> hsys(nsys);
  int max_nat = 0;
  for (int i = 0; i < nsys; ++i) {
    auto p = std::make_unique<DevSystem>();
    p->h = build_host_sys(*sys[i]);
    upload_system(*p, E.stream);
    dsys[i] = p->d;
    hsys[i] = &p->h;
    max_nat = std::max(max_nat, p->h.nat);
    E.batch_sys.push_back(std::move(p));
  }
  std::vector<int64_t> off(nconf);
  int64_t tot = 0;
  for (int c = 0; c < nconf; ++c) {
    off[c] = tot;
    tot += dsys[conf_sys[c]].nvt;
  }
  DevSys* d_systems = static_cast<DevSys*>(E.d_sys.ensure(sizeof(DevSys) * nsys));
  E.valid = false;  // d_sys now holds the batch systems
  int32_t* d_cs = static_cast<int32_t*>(E.d_confsys.ensure(sizeof(int32_t) * nconf));
  int64_t* d_off = static_cast<int64_t*>(E.d_confoff.ensure(sizeof(int64_t) * nconf));
  double* d_vl = static_cast<double*>(E.d_vlvr.ensure(sizeof(double) * tot));
  double* d_out = static_cast<double*>(E.d_out.ensure(sizeof(double) * kEnergyOut * nconf));
  HIPCHK(hipMemcpyAsync(d_systems, dsys.data(), sizeof(DevSys) * nsys, hipMemcpyHostToDevice, E.stream));
  HIPCHK(hipMemcpyAsync(d_cs, conf_sys, sizeof(int32_t) * nconf, hipMemcpyHostToDevice, E.stream));
  HIPCHK(hipMemcpyAsync(d_off, off.data(), sizeof(int64_t) * nconf, hipMemcpyHostToDevice, E.stream));
  HIPCHK(hipMemcpyAsync(d_vl, vlvr, sizeof(double) * tot, hipMemcpyHostToDevice, E.stream));
  std::vector<const DevSys*> host_sys(nsys);
  for (int i = 0; i < nsys; ++i) host_sys[i] = &dsys[i];
  double* d_xyz = analytic ? static_cast<double*>(E.d_xyz.ensure(sizeof(double) * 3 * max_nat * (size_t)nconf))
                           : nullptr;
  launch_energy(E, d_systems, d_cs, d_off, d_vl, d_out, d_xyz, max_nat, nconf,
                plan_energy_lds(host_sys.data(), nsys, kEnergyThreads), precision);
  if (analytic) {
    // rotations draw from each conformation's context, conformations in order
    std::vector<int32_t> cs(conf_sys, conf_sys + nconf);
    add_esolan(E, d_systems, hsys, cs, d_cs, nconf, max_nat, d_xyz, d_out,
               [&](int c) { return host_grnd(*sys[conf_sys[c]]); });
  }
  std::vector<double> full((size_t)kEnergyOut * nconf);
  HIPCHK(hipMemcpyAsync(full.data(), d_out, sizeof(double) * full.size(), hipMemcpyDeviceToHost, E.stream));
  HIPCHK(hipStreamSynchronize(E.stream));
  if (read_err(E)) throw Error("enysol> too many solvent neighbours for one atom");
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
