// Replica samplers on the GPU: one workgroup per replica runs Metropolis
// sweeps (metropolis.f:13-184) with its own twister.f MT19937 stream, an
// exchange kernel swaps temperatures (partem_s.f:209-232).
//
// A trial move of torsion v rotates the atoms [m1, m2] that hang from the
// base atom rigidly about the bond (axis0 -> base) instead of rebuilding the
// chain, and the energy change is summed over what the rotation changes:
// pairs between moved and fixed atoms, the torsion term of v, and the
// solvation of atoms in contact across the moved/fixed boundary.  Above
// |E| >= 1e5 kcal/mol (the reference's own drift-check threshold) the trial
// energy is recomputed in full instead, so cancellation of clash energies
// cannot leak into the tracked energy.  Every sweep ends with a rebuild from
// the variables and a full energy, checked against the tracked energy with
// the reference's 1e-6 rule.
#pragma once
#include "common.hpp"
#include "energy.hpp"
#include "surrogate.hpp"

namespace smmp {

// -------------------------------------------------- MT19937 (twister.f)
__host__ __device__ inline void mt_twist(uint32_t* mt) {
  const uint32_t mag = 0x9908B0DFu;
  int kk;
  uint32_t y;
  for (kk = 0; kk < 624 - 397; ++kk) {
    y = (mt[kk] & 0x80000000u) | (mt[kk + 1] & 0x7fffffffu);
    mt[kk] = mt[kk + 397] ^ (y >> 1) ^ ((y & 1u) ? mag : 0u);
  }
  for (; kk < 623; ++kk) {
    y = (mt[kk] & 0x80000000u) | (mt[kk + 1] & 0x7fffffffu);
    mt[kk] = mt[kk + 397 - 624] ^ (y >> 1) ^ ((y & 1u) ? mag : 0u);
  }
  y = (mt[623] & 0x80000000u) | (mt[0] & 0x7fffffffu);
  mt[623] = mt[396] ^ (y >> 1) ^ ((y & 1u) ? mag : 0u);
}
__host__ __device__ inline uint32_t mt_temper(uint32_t y) {
  y ^= y >> 11;
  y ^= (y << 7) & 0x9D2C5680u;
  y ^= (y << 15) & 0xEFC60000u;
  y ^= y >> 18;
  return y;
}
__host__ __device__ inline void mt_seed(uint32_t* mt, uint32_t* mti, int32_t seed) {
  mt[0] = (uint32_t)seed;
  for (int i = 1; i < 624; ++i) mt[i] = 69069u * mt[i - 1];
  *mti = 624;
}
// grnd(): u = y / (2^32 - 1) in [0, 1] (twister.f:172-176).  One thread.
__device__ inline double mt_grnd(uint32_t* mt, uint32_t* mti) {
  if (*mti >= 624) {
    if (*mti == 625) mt_seed(mt, mti, 4357);
    mt_twist(mt);
    *mti = 0;
  }
  uint32_t y = mt_temper(mt[(*mti)++]);
  return (double)y / 4294967295.0;
}

// ---------------------------------------------------- launch parameters
struct SamplerParams {
  const DevSys* sys;
  int32_t nrep, nvr, nvt, rndord;
  // Update-type choice of metropolis.f:46-57 (/updchois/ upchswitch,
  // bgsprob; /updtparam/ up2bmax, up2bmin): with upch 1 (updtch1) or 2
  // (updtch2) every trial of a variable whose name does not start with 'x'
  // draws one number first; below bgsprob (1) or (beta-up2bmin)/(up2bmax-
  // up2bmin) (2) the trial is a BGS update, which the reference leaves
  // inert (bgsnvar = 0, bgs.f:86-89): no move, no energy, not accepted.
  int32_t upch;
  double bgsprob, up2bmax, up2bmin;
  double upch_beta;      // the beta updtch2 sees when upch_own is 0 (/bet/ of smmp_metropolis)
  int32_t upch_own;      // 1: updtch2 sees the replica's beta (partem_s sets /bet/ per replica)
  double* upd;           // [R][10] ncalls(1..5), nacalls(1..5) of accanalyze (metropolis.f:186-193), or null
  const int32_t* idvr;   // free variables, 0-based
  const int8_t* var_x;   // [nvt] 1: name starts with 'x' (null: all, no draws)
  const double* axvr;    // nvt step widths
  double* vlvr;          // [R][nvt]
  double* xyz;           // [R][3][np] coordinates matching vlvr
  int32_t* nfree;        // [R][nheavy] free points per heavy atom
  int32_t* valid;        // [R] xyz/nfree/eol hold a consistent state
  double* eol;           // [R]
  double* acc;           // [R]
  uint32_t* mt;          // [streams][625] (624 words + position)
  const int32_t* rng_of; // replica -> stream
  const int32_t* slot_of;// replica -> temperature slot
  const double* beta;    // [slot]
  int32_t wkind, kmin, kmax;
  double ebin;
  const double* wb;
  const double* walpha;
  int32_t* hist;         // multicanonical histogram (kmax-kmin+1) or null
  double* traj;          // [sweep][slot] energies or null
  double* obs;           // [sweep][slot][2] radius of gyration and C-alpha end-to-end distance, or null
  int32_t traj_base;     // first sweep index of this launch
  int32_t nsweep;        // sweeps per launch
  int32_t rep0;          // replica of block 0
  int32_t* err;          // bit 0 drift, bit 1 neighbour overflow
  double* dbg;           // optional per-replica diagnostics
  double* work;          // [R][4] algorithmic work counters (accumulated)
  int32_t verify;        // recompute every trial in full and record the first mismatch in dbg
  double* trace;         // optional: per trial of replica trace_rep, (var, E_trial, E_old, accepted)
  int32_t trace_rep, trace_cap;
  int32_t* trace_n;
  double* phase;         // optional: [R][8] wall-clock ticks per trial phase (see PhaseId)
  int32_t bits_lds;      // stage the pair bit rows in LDS
  int32_t var_global;    // per-variable FP64 tables read from global memory instead of LDS
  // Speculative batches (G > 1): the G blocks of a replica evaluate the next
  // G torsion trials at once, block g trial g as if every earlier trial of
  // the batch is rejected, and then take the decisions in order (see
  // k_sweep).  spec_nf: [R][2][G][1 + nheavy] the free-point counts of each
  // block's trial (count, then h | nf << 16).
  int32_t spec;
  int32_t* spec_nf;
  // Workgroups per replica (G).  With G > 1 the G blocks of a replica hold
  // the same state and split every energy change between them; they meet
  // once per trial in group_sum.  The launch is cooperative (all blocks
  // resident).
  int32_t G;
  int32_t xcd_groups;    // place the G blocks of a replica on one XCD
  double* part;          // [R][2][G][kPartVals] partial sums
  uint32_t* sync;        // [R] arrival counters, zero at launch
  uint64_t wait_ticks;   // bound of one wait, device wall-clock ticks
  // Parallel tempering inside one persistent launch: after every sweep all
  // blocks meet (grid barrier), block 0 attempts the exchanges, all meet
  // again and read their new temperature slot.  Cooperative launch.
  int32_t pt_exchange;
  int32_t* ipoi;         // slot -> replica (1-based), partem_s.f:209-232
  int32_t* iswitch;      // which half of the pairs is tried next
  uint32_t* mt_x;        // the exchange stream (625 words)
  uint32_t* gsync;       // grid-barrier arrival counter, zero at launch
  double* gbpr;          // per replica: 6 global parameters per chain (kMaxMol chains)
  // delayed acceptance (da.theta null: plain Metropolis): the surrogate
  // network, kcal/mol per network unit, per-replica counts (screened,
  // rejected by the network, judged exactly, accepted)
  MlpScreen da;
  double da_scale;
  double* da_stats;
  size_t da_lds_off;     // byte offset of the surrogate's LDS region
  // snapshots for training: every snap_every sweeps, each replica's
  // variables (snap[k][r][nvt]) and energy (snap_e[k][r])
  double* snap;
  double* snap_e;
  int32_t snap_every;
  // Sweeps already done when launches are replayed from a graph (a kernel
  // argument cannot change between replays); null: traj_base.
  int32_t* sweep_base;
};

// Fills n words (graph-safe zeroing of counters: no memset nodes).
__global__ void k_fill_u32(uint32_t* p, int n, uint32_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
// No-op: probes whether a cooperative launch is granted to this process.
__global__ void k_coop_probe(int32_t unused) { (void)unused; }
// Advances the sweep counter of graph replays past a finished launch.
__global__ void k_add_i32(int32_t* p, int v) {
  if (threadIdx.x == 0) *p += v;
}
constexpr int kPartVals = 8;
constexpr int kSweepThreads = 512;  // threads of one sweep workgroup (8 waves; 16 spill and run slower)

// Phases of a sweep timed by thread 0 when P.phase is set.
enum PhaseId { kPhPropose = 0, kPhRotate, kPhPairs, kPhSolvent, kPhReduce, kPhAccept, kPhFull, kPhSweepEnd,
               kPhNFull, kPhCoreCycles, kPhRealTicks,
               // solvent change, per trial: slowest wave, and sums over waves of the
               // gather and point-test times, atoms and contacts
               kPhSolMax, kPhSolGather, kPhSolPoints, kPhSolAtoms, kPhSolContacts,
               // pair change, per trial: slowest wave and wave sum of pair_delta,
               // wave sum of the contact marking, thread 0's torsion term
               kPhPairMax, kPhPairSum, kPhMarkSum, kPhTorsion,
               // reduction across the workgroups of a replica, per trial: the
               // wait for the last workgroup's arrival, and the time from that
               // arrival until this workgroup leaves the wait
               kPhSyncWait, kPhSyncLatency, kNPhase };
// broadcast scalars [0,32), phase clocks [32,64), per-wave stats [64,160),
// rigid-move transform [160,176), accanalyze counters [176,186)
constexpr int kScDoubles = 192;
constexpr int kScRigid = 160;
constexpr int kScUpd = 176;
// Speculative batches: per trial of the batch 8 doubles (jv, old value, new
// value, angle change, acceptance number, BGS flag, draws before the
// acceptance number, draws after it), the G gathered results (8 each), this
// block's result (8), batch scalars (8); then a copy of the twister.
constexpr int kSpecMaxG = 8;
constexpr int kSpecDoubles = 16 * kSpecMaxG + 16;

// Extra LDS of the sampler beyond the energy carve-up.  var_global: the
// per-variable FP64 tables stay in global memory (a smaller carve-up, so
// that two workgroups can share a CU).
__host__ __device__ inline size_t sampler_lds_bytes(int np, int npnt, int nheavy, int nvt, int nwaves, size_t nbits,
                                                    bool var_global = false, bool spec = false) {
  size_t b = energy_lds_bytes(np, npnt, nwaves, nheavy, nbits);
  b = (b + 15) & ~size_t(15);
  b += 3 * (size_t)np * sizeof(double);           // backup of moved coordinates
  b += (size_t)nvt * sizeof(double);              // variables
  b += 2 * (size_t)nheavy * sizeof(int32_t);      // nfree current / trial
  b += 625 * sizeof(uint32_t);                    // MT state
  b += kScDoubles * sizeof(double);               // broadcast scalars and phase clocks
  b += 6 * kMaxMol * sizeof(double);              // global parameters of the chains
  if (!var_global) b += 7 * (size_t)nvt * sizeof(double);  // per-variable tables (VarLds)
  b += 8 * (size_t)nvt * sizeof(int16_t);
  b += 2 * (size_t)nheavy;                        // list of the affected heavy atoms
  b += (size_t)nheavy;                            // affected flags
  b += (size_t)nvt;                               // variable names starting with 'x'
  if (spec) b = ((b + 7) & ~size_t(7)) + kSpecDoubles * sizeof(double) + 625 * sizeof(uint32_t);
  return b + 64;
}

// Per-variable tables read by every trial, staged in LDS: step width,
// rotation sign, torsion-energy parameters, the torsion atom's snapshot
// (toat0, tref) and controlling variable, moved range, axis atoms.  With
// the compact carve-up the FP64 tables point at the system's arrays in
// global memory and t0, tref are null (the torsion atom's values are read
// where they are used).
struct VarLds {
  double *ax, *sign, *e0, *sg, *rn, *t0, *tref;
  int16_t *m1, *m2, *base, *axis0, *tv, *idvr, *h1, *h2;
  int8_t* xname;  // 1: the variable's name starts with 'x' (no update-type draw)
};

struct SamplerLds {
  EnergyLds L;
  VarLds V;
  double *bx, *by, *bz, *vl, *sc;
  double* gb;  // global parameters of the replica's chains (6 per chain)
  int32_t *nf, *nf2;
  uint32_t* mt;
  int16_t* alist;
  int8_t* aff;
  double* spec;      // speculative batches (kSpecDoubles; carved only with P.spec)
  uint32_t* mtsave;  // the twister before a batch's draws (625 words)
};

// lds_flags: bit 0 pair bits in LDS, bit 1 per-variable FP64 tables in
// global memory (axvr: the step widths there; may be null where V is unused).
__device__ inline SamplerLds carve_sampler_lds(double* base, const DevSys& S, int nvt, int nwaves, int lds_flags,
                                               const double* axvr) {
  SamplerLds Q;
  const bool bits_lds = (lds_flags & 1) != 0, var_global = (lds_flags & 2) != 0;
  Q.L = carve_energy_lds(base, S, nwaves, bits_lds);
  const size_t nbits = bits_lds ? pair_bit_words(S.nat, S.words) : 0;
  char* p = reinterpret_cast<char*>(base) + ((energy_lds_bytes(S.np, S.npnt, nwaves, S.nheavy, nbits) + 15) & ~size_t(15));
  Q.bx = reinterpret_cast<double*>(p); p += S.np * sizeof(double);
  Q.by = reinterpret_cast<double*>(p); p += S.np * sizeof(double);
  Q.bz = reinterpret_cast<double*>(p); p += S.np * sizeof(double);
  Q.vl = reinterpret_cast<double*>(p); p += nvt * sizeof(double);
  Q.sc = reinterpret_cast<double*>(p); p += kScDoubles * sizeof(double);
  Q.gb = reinterpret_cast<double*>(p); p += 6 * kMaxMol * sizeof(double);
  // doubles first (8-byte aligned), then 32-, 16- and 8-bit arrays
  if (var_global) {
    Q.V.ax = const_cast<double*>(axvr);
    Q.V.sign = const_cast<double*>(S.var_sign);
    Q.V.e0 = const_cast<double*>(S.var_e0);
    Q.V.sg = const_cast<double*>(S.var_sg);
    Q.V.rn = const_cast<double*>(S.var_rn);
    Q.V.t0 = nullptr;
    Q.V.tref = nullptr;
  } else {
    Q.V.ax = reinterpret_cast<double*>(p); p += nvt * sizeof(double);
    Q.V.sign = reinterpret_cast<double*>(p); p += nvt * sizeof(double);
    Q.V.e0 = reinterpret_cast<double*>(p); p += nvt * sizeof(double);
    Q.V.sg = reinterpret_cast<double*>(p); p += nvt * sizeof(double);
    Q.V.rn = reinterpret_cast<double*>(p); p += nvt * sizeof(double);
    Q.V.t0 = reinterpret_cast<double*>(p); p += nvt * sizeof(double);
    Q.V.tref = reinterpret_cast<double*>(p); p += nvt * sizeof(double);
  }
  Q.nf = reinterpret_cast<int32_t*>(p); p += S.nheavy * sizeof(int32_t);
  Q.nf2 = reinterpret_cast<int32_t*>(p); p += S.nheavy * sizeof(int32_t);
  Q.mt = reinterpret_cast<uint32_t*>(p); p += 625 * sizeof(uint32_t);
  Q.V.m1 = reinterpret_cast<int16_t*>(p); p += nvt * sizeof(int16_t);
  Q.V.m2 = reinterpret_cast<int16_t*>(p); p += nvt * sizeof(int16_t);
  Q.V.base = reinterpret_cast<int16_t*>(p); p += nvt * sizeof(int16_t);
  Q.V.axis0 = reinterpret_cast<int16_t*>(p); p += nvt * sizeof(int16_t);
  Q.V.tv = reinterpret_cast<int16_t*>(p); p += nvt * sizeof(int16_t);
  Q.V.idvr = reinterpret_cast<int16_t*>(p); p += nvt * sizeof(int16_t);
  Q.V.h1 = reinterpret_cast<int16_t*>(p); p += nvt * sizeof(int16_t);
  Q.V.h2 = reinterpret_cast<int16_t*>(p); p += nvt * sizeof(int16_t);
  Q.alist = reinterpret_cast<int16_t*>(p); p += S.nheavy * sizeof(int16_t);
  Q.aff = reinterpret_cast<int8_t*>(p); p += S.nheavy;
  Q.V.xname = reinterpret_cast<int8_t*>(p);
  p += nvt;
  p = reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(p) + 7) & ~uintptr_t(7));
  Q.spec = reinterpret_cast<double*>(p);
  Q.mtsave = reinterpret_cast<uint32_t*>(p + kSpecDoubles * sizeof(double));
  return Q;
}

// Stages the per-variable tables (all threads; a barrier must follow).
__device__ void stage_variables(const SamplerParams& P, const DevSys& S, const VarLds& V) {
  const bool in_lds = V.t0 != nullptr;  // else the FP64 tables are the global ones
  for (int k = threadIdx.x; k < P.nvt; k += blockDim.x) {
    const double e0 = S.var_e0[k];
    if (in_lds) {
      V.ax[k] = P.axvr[k];
      V.sign[k] = S.var_sign[k];
      V.e0[k] = e0;
      V.sg[k] = S.var_sg[k];
      V.rn[k] = S.var_rn[k];
    }
    V.m1[k] = S.var_m1[k];
    V.m2[k] = S.var_m2[k];
    V.base[k] = S.var_base[k];
    V.axis0[k] = S.var_axis0[k];
    V.h1[k] = S.var_h1[k];
    V.h2[k] = S.var_h2[k];
    double t0 = 0.0, tref = 0.0;
    int16_t tv = -1;
    if (e0 != 0.0) {
      const int ia = S.var_atom[k];
      t0 = S.toat0[ia];
      tref = S.tref[ia];
      tv = S.tvar[ia];
    }
    if (in_lds) {
      V.t0[k] = t0;
      V.tref[k] = tref;
    }
    V.tv[k] = tv;
  }
  for (int k = threadIdx.x; k < P.nvr; k += blockDim.x) V.idvr[k] = (int16_t)P.idvr[k];
  for (int k = threadIdx.x; k < P.nvt; k += blockDim.x) V.xname[k] = P.var_x ? P.var_x[k] : (int8_t)1;
}


// Weight of an energy (canon.f:135-149, mulcan_par_mod.f90:380-412).
__device__ inline double weight_of(const SamplerParams& P, double beta, double x) {
  if (P.wkind == 0) return beta * x;
  double v = x / P.ebin + copysign(0.5, x);
  v = fmin(fmax(v, (double)P.kmin), (double)P.kmax);
  int k = (int)v - P.kmin;
  if (P.wkind == 1) return (beta + P.wb[k]) * x + P.walpha[k];
  return P.wb[k] * x + P.walpha[k];
}

// ECEPP pair energy split into Coulomb and the 6-12 (or 10-12 H-bond) term;
// A carries a14 for 1-4 pairs.  Branch-free in the pair type.
struct PairE {
  double el, vw, hb;
};
__device__ __forceinline__ void pair_split(const DevSys& S, double dx, double dy, double dz, double qq, double A,
                                           double C, bool hb, double& el, double& lj) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double ri = rsqrt_f64(r2);
  const double r2i = ri * ri;
  const double r6i = r2i * r2i * r2i;
  double ep_inv = 1.0;
  if (S.epsd) {  // sigmoidal dielectric (enyshe.f:129-134)
    const double sr = S.slp * (r2 * ri);
    ep_inv = 1.0 / (S.plt - (sr * sr + 2.0 * sr + 2.0) * (S.plt - 1.0) * exp(-sr) / 2.0);
  }
  el = qq * ri * ep_inv;
  lj = A * (r6i * r6i) - C * (hb ? r6i * r2i * r2i : r6i);
}
// The same pair terms in FP32 from FP64 coordinate differences (the form of
// the full evaluation's mixed tiles).
__device__ __forceinline__ void pair_split_f32(const DevSys& S, double dx64, double dy64, double dz64, float qq,
                                               float A, float C, bool hb, float& el, float& lj) {
  const float dx = (float)dx64, dy = (float)dy64, dz = (float)dz64;
  const float r2 = dx * dx + dy * dy + dz * dz;
  const float ri = __builtin_amdgcn_rsqf(r2);
  const float r2i = ri * ri;
  const float r6i = r2i * r2i * r2i;
  float ep_inv = 1.f;
  if (S.epsd) {  // sigmoidal dielectric (enyshe.f:129-134)
    const float sr = (float)S.slp * (r2 * ri), plt = (float)S.plt;
    ep_inv = 1.f / (plt - (sr * sr + 2.f * sr + 2.f) * (plt - 1.f) * expf(-sr) / 2.f);
  }
  el = qq * ri * ep_inv;
  lj = A * (r6i * r6i) - C * (hb ? r6i * r2i * r2i : r6i);
}

// Energy change of the pairs between the moved range [m1,m2] (new
// coordinates in L, old in the backup) and the fixed atoms.  The smaller
// set is walked uniformly, the larger one across the lanes.  Work items
// item0, item0 + stride, ... of the (row, 64-column block) grid.  MIXED:
// each pair's new and old terms in FP32, their difference and the sums in
// FP64.
template <bool MIXED>
__device__ void pair_delta(const DevSys& S, const SamplerLds& Q, int m1, int m2, int item0, int stride, PairE& d,
                           Work& wk) {
  const EnergyLds& L = Q.L;
  const int lane = threadIdx.x & (kWave - 1);
  const int nat = S.nat, nM = m2 - m1 + 1, nO = nat - nM, words = S.words;
  if (nM <= 0 || nO <= 0) return;
  const bool rowsM = nM <= nO;            // uniform rows from the smaller set
  const int nrow = rowsM ? nM : nO, ncol = rowsM ? nO : nM;
  const int nob = (ncol + 63) / 64;
  for (int item = item0; item < nrow * nob; item += stride) {
    const int r = item / nob, ob = item - r * nob;
    const int c0 = ob * 64 + lane;
    const bool inb = c0 < ncol;
    const int c = inb ? c0 : 0;
    int a, b;  // a in M, b fixed
    if (rowsM) {
      a = m1 + r;
      b = c < m1 ? c : c + nM;
    } else {
      b = r < m1 ? r : r + nM;
      a = m1 + c;
    }
    // every load up front (the bit only masks the sums), so the LDS round
    // trips of an item overlap instead of chaining behind a branch
    const size_t w = ((size_t)a * words + (b >> 6)) * 2;
    const uint64_t w0 = pair_bits(L, w), w1 = pair_bits(L, w + 1);
    const int k = a - m1;
    const double xa = L.x[a], ya = L.y[a], za = L.z[a];
    const double xo = Q.bx[k], yo = Q.by[k], zo = Q.bz[k];
    const double xb = L.x[b], yb = L.y[b], zb = L.z[b];
    const double qa = L.q[a], qb = L.q[b];
    const int tt = L.t[a] * kTy + L.t[b];
    const double2 aa = *reinterpret_cast<const double2*>(L.ptab + 4 * tt);
    const double2 cf = *reinterpret_cast<const double2*>(L.ptab + 4 * tt + 2);
    const bool on = inb && ((w0 >> (b & 63)) & 1ull);
    const bool is14 = (w1 >> (b & 63)) & 1ull;
    const double A = is14 ? aa.y : aa.x, C = cf.x;
    const bool hb = cf.y != 0.0;
    const double qq = S.conv * qa * qb;
    double eln, ljn, elo, ljo;
    if constexpr (MIXED) {
      float fen, fln, feo, flo;
      pair_split_f32(S, xa - xb, ya - yb, za - zb, (float)qq, (float)A, (float)C, hb, fen, fln);
      pair_split_f32(S, xo - xb, yo - yb, zo - zb, (float)qq, (float)A, (float)C, hb, feo, flo);
      eln = fen;
      ljn = fln;
      elo = feo;
      ljo = flo;
    } else {
      pair_split(S, xa - xb, ya - yb, za - zb, qq, A, C, hb, eln, ljn);
      pair_split(S, xo - xb, yo - yb, zo - zb, qq, A, C, hb, elo, ljo);
    }
    const double del = on ? eln - elo : 0.0;
    const double dlj = on ? ljn - ljo : 0.0;
    d.el += del;
    d.hb += hb ? dlj : 0.0;
    d.vw += hb ? 0.0 : dlj;
    wk.pairs += on ? 2.0 : 0.0;
  }
}

// Marks the heavy atoms whose accessible area may change: every moved atom
// (its 122 points stay fixed in the laboratory frame while it rotates, so
// even a rigidly moved environment changes its discretised area) and every
// fixed atom in contact with a moved one before or after the move.  Only
// the fixed atoms of this block's share (h % G == g) are tested.
__device__ void mark_affected(const DevSys& S, const SamplerLds& Q, int h1, int h2, int m1, int g, int G) {
  const int nh = S.nheavy;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave, nwaves = blockDim.x / kWave;
  // h1, h2: heavy atoms of the moved range
  const int nM = h2 - h1, nown = (nh - g + G - 1) / G;
  for (int h = h1 + threadIdx.x; h < h2; h += blockDim.x) Q.aff[h] = 1;
  // a wave per moved heavy atom, lanes over this share's heavy atoms; the
  // packed copy holds the trial positions, the backup the old ones
  const double* hp = Q.L.hp;
  for (int r = wave; r < nM; r += nwaves) {
    const int ha = h1 + r;
    const int k2 = Q.L.hv[ha] - m1;
    const double xn = hp[4 * ha + 0], yn = hp[4 * ha + 1], zn = hp[4 * ha + 2], ra = hp[4 * ha + 3];
    const double xo = Q.bx[k2], yo = Q.by[k2], zo = Q.bz[k2];
    for (int k0 = 0; k0 < nown; k0 += kWave) {
      const int k = k0 + lane;
      const int hb = g + G * k;
      if (k >= nown || (hb >= h1 && hb < h2)) continue;
      const double2 pb = *reinterpret_cast<const double2*>(hp + 4 * hb);
      const double2 qb = *reinterpret_cast<const double2*>(hp + 4 * hb + 2);
      const double lim = (ra + qb.y) * (1.0 + 1e-9);
      const double l2 = lim * lim;
      double dx = xn - pb.x, dy = yn - pb.y, dz = zn - qb.x;
      bool hit = dx * dx + dy * dy + dz * dz <= l2;
      dx = xo - pb.x; dy = yo - pb.y; dz = zo - qb.x;
      hit = hit || (dx * dx + dy * dy + dz * dz <= l2);
      if (hit) Q.aff[hb] = 1;
    }
  }
}

// Wave 0: Q.alist = the affected heavy atoms of share g in increasing
// order; returns their number (valid in wave 0; a barrier must follow).
__device__ int list_affected(const DevSys& S, const SamplerLds& Q, int g, int G) {
  const int lane = threadIdx.x & (kWave - 1);
  const int nown = (S.nheavy - g + G - 1) / G;
  int cnt = 0;
  for (int k0 = 0; k0 < nown; k0 += kWave) {
    const int k = k0 + lane;
    const int h = g + G * k;
    const bool on = k < nown && Q.aff[h] != 0;
    const uint64_t m = __ballot(on);
    if (on) Q.alist[cnt + __popcll(m & ((1ull << lane) - 1ull))] = (int16_t)h;
    cnt += __popcll(m);
  }
  return cnt;
}

// Full energy of the replica in LDS: from the coordinates, or (rebuild) from
// the variables through the chain build.  One out-of-line copy serves every
// call site (inlined, five copies made the kernel twice the size of the
// instruction cache).  It takes no LDS pointers: it carves the layout from
// the dynamic LDS itself, so its accesses stay LDS accesses.
// nf_sel: 0 = no free-point output, 1 = Q.nf, 2 = Q.nf2.
struct ReplicaEnergy {
  EnergyOut o;
  Work w;
};
template <bool MIXED>
__device__ __attribute__((noinline)) ReplicaEnergy replica_energy(const DevSys* sys, int nvt, int lds_flags,
                                                                  int nf_sel, bool rebuild, int part, int nparts) {
  extern __shared__ double smem[];
  const DevSys& S = *sys;
  SamplerLds Q = carve_sampler_lds(smem, S, nvt, blockDim.x / kWave, lds_flags, nullptr);
  int32_t* nfree_out = nf_sel == 1 ? Q.nf : (nf_sel == 2 ? Q.nf2 : nullptr);
  ReplicaEnergy r;
  if (rebuild)
    conformation_energy<MIXED>(S, Q.vl, Q.gb, Q.L, nfree_out, r.o, r.w, part, nparts);
  else
    energy_from_coords<MIXED>(S, Q.vl, Q.L, nfree_out, r.o, r.w, part, nparts);
  return r;
}
__device__ __forceinline__ void add_work(Work& a, const Work& b) {
  a.pairs += b.pairs;
  a.contacts += b.contacts;
  a.builds += b.builds;
  a.moved += b.moved;
}

// Replica and share of block b.  Blocks are dispatched round-robin over
// the 8 XCDs, so with xcd_groups the G blocks of a replica are b = x + 8q
// for one x (one L2 for their reductions).
__device__ __forceinline__ void replica_of_block(const SamplerParams& P, int b, int& r, int& g) {
  const int G = P.G;
  if (G == 1) {
    r = P.rep0 + b;
    g = 0;
  } else if (P.xcd_groups) {
    const int x = b & 7, q = b >> 3;
    r = P.rep0 + (q / G) * 8 + x;
    g = q % G;
  } else {
    r = P.rep0 + b / G;
    g = b % G;
  }
}

// Thread 0 of each of the G blocks of replica r: replaces v[0..nv) (this
// block's partial sums) by the totals over the G blocks.  Every block adds
// the partials in block order, so all get bit-identical totals and take the
// same decisions.  Partials are double-buffered by the parity of the
// reduction index: a block can be at most one reduction ahead of another
// (it waits for all of them at each reduction), so it never overwrites
// values still being read.  The wait is bounded on the device clock; on
// expiry the error flag gets bit 3 and the block goes on.
//
// Every access to the shared partials and counters is an agent-scope atomic
// (stores and loads reach the coherent level), the partials' stores are
// complete (s_waitcnt vmcnt(0)) before the arrival is counted, and a block
// reads the partials only after it has seen every arrival.  Nothing shared
// is read through a non-coherent cache, so the agent-scope release and
// acquire fences are not needed: on gfx950 they write back every dirty line
// of the XCD's L2 (the spill slots of all its workgroups among them) and
// invalidate it, once per trial and workgroup.
__device__ __forceinline__ void part_store(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double part_load(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ void group_wait(const SamplerParams& P, int r, uint32_t epoch);
__device__ void group_sum(const SamplerParams& P, int r, int g, uint32_t& epoch, double* v, int nv,
                          double* ph = nullptr) {
  const int G = P.G;
  if (G == 1) return;
  double* base = P.part + ((size_t)r * 2 + (epoch & 1u)) * G * kPartVals;
  double* mine = base + (size_t)g * kPartVals;
  for (int k = 0; k < nv; ++k) part_store(&mine[k], v[k]);
  const double arrive = ph ? (double)wall_clock64() : 0.0;
  if (ph) part_store(&mine[kPartVals - 1], arrive);  // timing only: the arrival stamp rides in the last slot
  group_wait(P, r, epoch);
  if (ph) {
    const double now = (double)wall_clock64();
    double last = arrive;
    for (int q = 0; q < G; ++q) last = fmax(last, part_load(&base[(size_t)q * kPartVals + kPartVals - 1]));
    ph[kPhSyncWait] += last - arrive;
    ph[kPhSyncLatency] += now - last;
  }
  for (int k = 0; k < nv; ++k) v[k] = 0.0;
  for (int q = 0; q < G; ++q) {
    const double* src = base + (size_t)q * kPartVals;
    for (int k = 0; k < nv; ++k) v[k] += part_load(&src[k]);
  }
  ++epoch;
}

// Like group_sum, but every block receives all G rows of kPartVals values
// (out[kPartVals q + k]) instead of their sum.
__device__ void group_gather(const SamplerParams& P, int r, int g, uint32_t& epoch, const double* v, double* out) {
  const int G = P.G;
  double* base = P.part + ((size_t)r * 2 + (epoch & 1u)) * G * kPartVals;
  double* mine = base + (size_t)g * kPartVals;
  for (int k = 0; k < kPartVals; ++k) part_store(&mine[k], v[k]);
  group_wait(P, r, epoch);
  for (int k = 0; k < G * kPartVals; ++k) out[k] = part_load(&base[k]);
  ++epoch;
}

// Arrival of this block at meeting 'epoch' of replica r and the bounded
// wait for all G blocks (thread 0): on return every block's partials
// stored before its arrival can be read with part_load.
__device__ void group_wait(const SamplerParams& P, int r, uint32_t epoch) {
  const int G = P.G;
  // the signal fences keep the compiler from moving the partials' atomics
  // across the arrival and the polls; they emit no instruction
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this block's partials are stored
  __hip_atomic_fetch_add(&P.sync[r], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t target = (epoch + 1u) * (uint32_t)G;
  const uint64_t t0 = wall_clock64();
  while (__hip_atomic_load(&P.sync[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    __builtin_amdgcn_s_sleep(1);
    // after one expiry anywhere, no block waits again: the launch drains
    if (wall_clock64() - t0 > P.wait_ticks ||
        (__hip_atomic_load(P.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 8)) {
      atomicOr(P.err, 8);
      break;
    }
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Temperature exchanges between neighbouring slots (partem_s.f:209-232):
// pairs (1,2),(3,4).. then (2,3),..,(R,1) alternately; the random number is
// drawn for every pair.  One thread; stream mt_x.
__device__ void exchange_step(int nrep, int32_t* ipoi, int32_t* slot_of, const double* beta, const double* eol,
                              uint32_t* mt_x, int32_t* iswitch) {
  int num_rep1, nu;
  if (*iswitch == 1) {
    num_rep1 = nrep - 1;
    nu = 1;
    *iswitch = 2;
  } else {
    num_rep1 = nrep;
    nu = 2;
    *iswitch = 1;
  }
  for (int i = nu; i <= num_rep1; i += 2) {
    int j = i + 1;
    if (i == nrep) j = 1;
    const int in = ipoi[i - 1], jn = ipoi[j - 1];
    const double bi = beta[i - 1], bj = beta[j - 1];
    const double delta = -bi * eol[jn - 1] - bj * eol[in - 1] + bi * eol[in - 1] + bj * eol[jn - 1];
    const double ra = mt_grnd(mt_x, &mt_x[624]);
    if (ra <= exp(delta)) {
      ipoi[i - 1] = jn;
      ipoi[j - 1] = in;
      slot_of[jn - 1] = i - 1;
      slot_of[in - 1] = j - 1;
    }
  }
}

// All blocks of the launch meet (thread 0 of each; the callers hold block
// barriers on either side).  Bounded on the device clock like group_sum.
__device__ void grid_sync(const SamplerParams& P, uint32_t& gepoch) {
  const uint32_t nb = gridDim.x;
  __hip_atomic_fetch_add(P.gsync, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t target = (gepoch + 1u) * nb;
  const uint64_t t0 = wall_clock64();
  while (__hip_atomic_load(P.gsync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    __builtin_amdgcn_s_sleep(2);
    if (wall_clock64() - t0 > P.wait_ticks ||
        (__hip_atomic_load(P.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 8)) {
      atomicOr(P.err, 8);
      break;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  ++gepoch;
}

// Turns the atoms [m1, m2] by the angle whose sine and cosine are sn, cs
// about the axis through p along the unit vector k (Rodrigues), with their
// packed heavy copies; backup: keep the old coordinates in Q.bx/by/bz.
// All threads.
__device__ __forceinline__ void rotate_range(const SamplerLds& Q, int m1, int m2, double px, double py, double pz,
                                             double kx, double ky, double kz, double sn, double cs, bool backup) {
  const EnergyLds& L = Q.L;
  for (int a = m1 + threadIdx.x; a <= m2; a += blockDim.x) {
    const double vx = L.x[a] - px, vy = L.y[a] - py, vz = L.z[a] - pz;
    if (backup) {
      Q.bx[a - m1] = L.x[a];
      Q.by[a - m1] = L.y[a];
      Q.bz[a - m1] = L.z[a];
    }
    const double kv = kx * vx + ky * vy + kz * vz;
    const double cx = ky * vz - kz * vy, cy = kz * vx - kx * vz, cz = kx * vy - ky * vx;
    const double t = kv * (1.0 - cs);
    const double xn = px + vx * cs + cx * sn + kx * t;
    const double yn = py + vy * cs + cy * sn + ky * t;
    const double zn = pz + vz * cs + cz * sn + kz * t;
    L.x[a] = xn;
    L.y[a] = yn;
    L.z[a] = zn;
    const int h = L.hidx[a];
    if (h >= 0) {
      L.hp[4 * h + 0] = xn;
      L.hp[4 * h + 1] = yn;
      L.hp[4 * h + 2] = zn;
    }
  }
}

// The rotation of a torsion trial of variable jv by dphi (the proposal's
// axis from the current coordinates); every thread computes the same values.
__device__ __forceinline__ void torsion_rotation(const DevSys& S, const SamplerLds& Q, int jv, double dphi, bool backup) {
  const EnergyLds& L = Q.L;
  const VarLds& V = Q.V;
  const int b = V.base[jv], a0 = V.axis0[jv];
  const double ux = L.x[b] - L.x[a0], uy = L.y[b] - L.y[a0], uz = L.z[b] - L.z[a0];
  const double un = rsqrt_f64(ux * ux + uy * uy + uz * uz);
  double sn, cs;
  sincos(V.sign[jv] * dphi, &sn, &cs);
  rotate_range(Q, V.m1[jv], V.m2[jv], L.x[b], L.y[b], L.z[b], ux * un, uy * un, uz * un, sn, cs, backup);
}

// Free-point counts after a full recomputation, where each block computed
// its share (h % G == g): publish (before the meeting that follows: this
// block's share to gnf, fenced) or collect (after it: all of gnf).  All
// threads; ends with a barrier.
__device__ __forceinline__ void share_free_points(const DevSys& S, const SamplerLds& Q, int32_t* gnf, int g, int G,
                                                  bool publish) {
  if (publish) {
    for (int h = threadIdx.x; h < S.nheavy; h += blockDim.x)
      if (h % G == g) gnf[h] = Q.nf[h];
    __threadfence();
  } else {
    __syncthreads();  // thread 0 has passed the meeting
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    for (int h = threadIdx.x; h < S.nheavy; h += blockDim.x)
      Q.nf[h] = __hip_atomic_load(&gnf[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
}

// One trial of a speculative batch, evaluated by this block alone on its
// copy of the replica state, which it leaves as it found it: variable jv
// from vold to vnew (angle change dphi).  Thread 0 receives the energy
// change's parts in res[0..6) (Coulomb, 6-12, H-bond, solvent, torsion
// term, error flag); nf_out (global) receives the trial's free-point counts
// of its affected heavy atoms (count, then h | nf << 16).  All threads.
template <bool MIXED>
__device__ __attribute__((noinline)) Work spec_trial(const DevSys* sys, int nvt, int lds_flags, const double* axvr,
                                                     int jv, double vold, double vnew, double dphi, int32_t* nf_out,
                                                     double* res) {
  extern __shared__ double smem[];
  const DevSys& S = *sys;
  const int tid = threadIdx.x, nwaves = blockDim.x / kWave, wave = tid / kWave, lane = tid & (kWave - 1);
  SamplerLds Q = carve_sampler_lds(smem, S, nvt, nwaves, lds_flags, axvr);
  const EnergyLds& L = Q.L;
  const VarLds& V = Q.V;
  Work wk;
  const int m1 = V.m1[jv], m2 = V.m2[jv], h1 = V.h1[jv], h2 = V.h2[jv], nh = S.nheavy;
  if (tid == 0) wk.moved += m2 - m1 + 1;
  torsion_rotation(S, Q, jv, dphi, true);
  // torsion term of the variable, the new and the old angle, on two lanes
  // of the last wave (enyshe.f:73-79)
  const int tl = tid - ((int)blockDim.x - kWave);
  if (tl == 0 || tl == 1) {
    const double e0 = V.e0[jv];
    double e = 0.0;
    if (e0 != 0.0) {
      const int v = V.tv[jv];
      const int ia = V.t0 ? 0 : S.var_atom[jv];
      double to = V.t0 ? V.t0[jv] : S.toat0[ia];
      if (v >= 0) {
        const double val = v == jv ? (tl == 0 ? vnew : vold) : Q.vl[v];
        to = to + d_difang(V.t0 ? V.tref[jv] : S.tref[ia], val);
        if (fabs(to) > kPi) to = to - copysign(kPi2, to);
      }
      e = e0 * (1.0 + V.sg[jv] * cos(to * V.rn[jv]));
    }
    Q.sc[17 + tl] = e;  // [17] new, [18] old
  }
  for (int h = tid; h < nh; h += blockDim.x) Q.aff[h] = 0;
  __syncthreads();
  PairE d = {0.0, 0.0, 0.0};
  pair_delta<MIXED>(S, Q, m1, m2, wave, nwaves, d, wk);
  if (S.itysol > 0) mark_affected(S, Q, h1, h2, m1, 0, 1);
  __syncthreads();
  int* next_atom = reinterpret_cast<int*>(Q.sc + 30);
  if (S.itysol > 0) {
    const int na = wave == 0 ? list_affected(S, Q, 0, 1) : 0;
    if (tid == 0) {
      Q.sc[29] = (double)na;
      *next_atom = 0;
    }
    __syncthreads();
  }
  int err = 0;
  if (S.itysol > 0) {
    double* nb = L.fr + (size_t)wave * kNbDoubles;
    const int na = (int)Q.sc[29];
    for (;;) {  // waves take the next atom as they finish
      int idx = 0;
      if (lane == 0) idx = atomicAdd(next_atom, 1);
      idx = __shfl(idx, 0, kWave);
      if (idx >= na) break;
      const int h = Q.alist[idx];
      int nfree = solvent_atom(S, L, nb, L.hv[h], wk, lds_flags & 1, nullptr);
      if (nfree < 0) {
        err = 1;
        nfree = Q.nf[h];
      }
      if (lane == 0) Q.nf2[h] = nfree;
    }
  }
  double v[3] = {d.el, d.vw, d.hb};
#pragma unroll
  for (int k = 0; k < 3; ++k) v[k] = wave_sum(v[k]);
  const int errw = __any(err);
  if (lane == 0) {
    for (int k = 0; k < 3; ++k) L.red[8 * wave + k] = v[k];
    L.red[8 * wave + 6] = (double)errw;
  }
  __syncthreads();
  double dsol = 0.0;
  const int na = S.itysol > 0 ? (int)Q.sc[29] : 0;
  if (S.itysol > 0 && wave == 0) {  // in list order: independent of which wave took which atom
    const double c4 = 4.0 * kPi / (double)S.npnt;
    for (int k = lane; k < na; k += kWave) {
      const int h = Q.alist[k];
      const int i = L.hv[h];
      const double trad = L.rv[i];
      dsol += L.sg[i] * (c4 * trad * trad * (double)(Q.nf2[h] - Q.nf[h]));
    }
    dsol = wave_sum(dsol);
  }
  if (tid == 0) {
    double t[6] = {0.0, 0.0, 0.0, dsol, Q.sc[17] - Q.sc[18], 0.0};
    for (int w = 0; w < nwaves; ++w) {
      for (int k = 0; k < 3; ++k) t[k] += L.red[8 * w + k];
      t[5] += L.red[8 * w + 6];
    }
    for (int k = 0; k < 6; ++k) res[k] = t[k];
    nf_out[0] = na;
  }
  for (int k = tid; k < na; k += blockDim.x) {
    const int h = Q.alist[k];
    nf_out[1 + k] = h | (Q.nf2[h] << 16);
  }
  __threadfence();  // the counts are visible before thread 0 reports arrival
  // back to the state before the trial
  for (int a = m1 + tid; a <= m2; a += blockDim.x) {
    const double xo = Q.bx[a - m1], yo = Q.by[a - m1], zo = Q.bz[a - m1];
    L.x[a] = xo;
    L.y[a] = yo;
    L.z[a] = zo;
    const int h = L.hidx[a];
    if (h >= 0) {
      L.hp[4 * h + 0] = xo;
      L.hp[4 * h + 1] = yo;
      L.hp[4 * h + 2] = zo;
    }
  }
  __syncthreads();
  return wk;
}

// The accepted trial of a speculative batch becomes the state of this
// block: the same rotation as in spec_trial (bit-identical coordinates),
// the variable, and the evaluating block's free-point counts nf_in (global,
// visible after the batch's meeting).  All threads.
__device__ __attribute__((noinline)) void spec_apply(const DevSys* sys, int nvt, int lds_flags, const double* axvr,
                                                     int jv, double vnew, double dphi, const int32_t* nf_in) {
  extern __shared__ double smem[];
  const DevSys& S = *sys;
  SamplerLds Q = carve_sampler_lds(smem, S, nvt, blockDim.x / kWave, lds_flags, axvr);
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  torsion_rotation(S, Q, jv, dphi, false);
  if (threadIdx.x == 0) Q.vl[jv] = vnew;
  if (S.itysol > 0) {
    const int na = nf_in[0];
    for (int k = threadIdx.x; k < na; k += blockDim.x) {
      const int32_t e = nf_in[1 + k];
      Q.nf[e & 0xffff] = e >> 16;
    }
  }
  __syncthreads();
}

// Torsion term of variable jv (enyshe.f:73-79) at its trial angle (tl 0,
// Q.sc[2]) and its old angle (tl 1, Q.sc[15]) into Q.sc[17 + tl], on block 0
// of the replica; other lanes and blocks return.  The new value is read
// from Q.sc[2], not Q.vl, which may already hold it.
__device__ __forceinline__ void torsion_pair(const DevSys& S, const SamplerLds& Q, const VarLds& V, int jv, int tl,
                                             int g) {
  if (g != 0 || (tl != 0 && tl != 1)) return;
  const double e0 = V.e0[jv];
  double e = 0.0;
  if (e0 != 0.0) {
    const int v = V.tv[jv];
    const int ia = V.t0 ? 0 : S.var_atom[jv];
    double to = V.t0 ? V.t0[jv] : S.toat0[ia];
    if (v >= 0) {
      const double val = v == jv ? (tl == 0 ? Q.sc[2] : Q.sc[15]) : Q.vl[v];
      to = to + d_difang(V.t0 ? V.tref[jv] : S.tref[ia], val);
      if (fabs(to) > kPi) to = to - copysign(kPi2, to);
    }
    e = e0 * (1.0 + V.sg[jv] * cos(to * V.rn[jv]));
  }
  Q.sc[17 + tl] = e;  // [17] new, [18] old
}

// P.nsweep Metropolis sweeps of one replica per group of G blocks; DA:
// delayed acceptance with the surrogate network (P.da).
// LEAN: one chain, no speculative batches, verify mode or trace (the
// common configuration): their code and the values it keeps live across a
// trial are compiled out.
template <bool MIXED, bool DA, bool LEAN>
__global__ void __launch_bounds__(kSweepThreads) k_sweep(SamplerParams P) {
  const bool verify = !LEAN && P.verify != 0;
  double* const trace = LEAN ? nullptr : P.trace;
  extern __shared__ double smem[];
  const DevSys& S = *P.sys;
  const int G = P.G;
  int r, g;
  replica_of_block(P, blockIdx.x, r, g);
  const int tid = threadIdx.x, nwaves = blockDim.x / kWave, wave = tid / kWave, lane = tid & (kWave - 1);
  const int nvt = P.nvt, np = S.np, nat = S.nat, nh = S.nheavy;
  const int lds_flags = (P.bits_lds ? 1 : 0) | (P.var_global ? 2 : 0);
  SamplerLds Q = carve_sampler_lds(smem, S, nvt, nwaves, lds_flags, P.axvr);
  // speculative batches: every block keeps the free-point counts of all
  // heavy atoms (shared through gnf after each full recomputation)
  const bool spec_on = !LEAN && P.spec != 0 && !DA && !verify && G > 1 && G <= kSpecMaxG;
  Work wk;
  EnergyLds& L = Q.L;
  double* gvl = P.vlvr + (size_t)r * nvt;
  double* gx = P.xyz + (size_t)r * 3 * np;
  int32_t* gnf = P.nfree + (size_t)r * nh;
  uint32_t* gmt = P.mt + (size_t)P.rng_of[r] * 625;
  uint32_t epoch = 0;   // reductions done (thread 0)
  uint32_t gepoch = 0;  // grid barriers done (thread 0)

  energy_stage_static(S, L);
  stage_variables(P, S, Q.V);
  const VarLds& V = Q.V;
  for (int k = tid; k < nvt; k += blockDim.x) Q.vl[k] = gvl[k];
  for (int k = tid; k < 625; k += blockDim.x) Q.mt[k] = gmt[k];
  double* ggb = P.gbpr + (size_t)r * 6 * kMaxMol;
  for (int k = tid; k < 6 * S.nml; k += blockDim.x) Q.gb[k] = ggb[k];
  const bool valid = P.valid[r] != 0;
  if (valid) {
    for (int a = tid; a < nat; a += blockDim.x) {
      L.x[a] = gx[a];
      L.y[a] = gx[np + a];
      L.z[a] = gx[2 * np + a];
    }
    for (int h = tid; h < nh; h += blockDim.x) Q.nf[h] = gnf[h];
  }
  __syncthreads();
  if (valid && S.itysol > 0) {  // the packed heavy copy of the loaded coordinates
    fill_heavy_packed(S, L);
    __syncthreads();
  }
  DaLds D{};
  if (DA) {  // the surrogate's tables, layer-1 sums and output at the start state
    D = carve_da_lds(reinterpret_cast<char*>(smem) + P.da_lds_off, P.da.res.nres, nvt);
    da_setup(P.da_lds_off, nvt, P.da, Q.vl);
  }
  // sc: [0] eol, [1] acz, [2] trial value, [3] dphi, [4] accept, [5] full-trial flag,
  //     [6] jv, [7] drift flag, [8..10] axis point, [11..13] axis unit, [15] old value,
  //     [16] trial energy, [24..28] incremental components (verify), [32..40] phase clocks
  if (!valid) {
    ReplicaEnergy re = replica_energy<MIXED>(P.sys, nvt, lds_flags, 1, true, g, G);
    if (g != 0) re.w.builds = 0.0;
    add_work(wk, re.w);
    if (spec_on) share_free_points(S, Q, gnf, g, G, true);
    if (tid == 0) {
      double v[2] = {re.o.el + re.o.vw + re.o.hb + re.o.tor + re.o.sol, (double)re.o.err};
      group_sum(P, r, g, epoch, v, 2);
      Q.sc[0] = v[0];
      if (v[1] != 0.0) atomicOr(P.err, 2);
    }
    if (spec_on) share_free_points(S, Q, gnf, g, G, false);
  } else if (tid == 0) {
    Q.sc[0] = P.eol[r];
  }
  if (tid == 0) {
    Q.sc[1] = P.acc[r];
    Q.sc[7] = 0.0;
  }
  __syncthreads();
  const double c4 = 4.0 * kPi / (double)(S.npnt > 0 ? S.npnt : 1);
  double* ph = Q.sc + 32;  // phase clocks (thread 0): [0, kNPhase) sums, [kNPhase] last stamp
  const uint64_t core0 = clock64(), real0 = wall_clock64();
  if (tid == 0) {
    for (int k = 0; k < kNPhase; ++k) ph[k] = 0.0;
    ph[kNPhase] = (double)real0;
    for (int k = 64; k < kScDoubles; ++k) Q.sc[k] = 0.0;
  }
  auto lap = [&](int k) {
    if (P.phase && tid == 0) {
      const double now = (double)wall_clock64();
      ph[k] += now - ph[kNPhase];
      ph[kNPhase] = now;
    }
  };

  // first trajectory row of this launch (the counter advances in a launch
  // after this one)
  const int traj_base = P.sweep_base ? *P.sweep_base : P.traj_base;
  for (int sw = 0; sw < P.nsweep; ++sw) {
    if (tid == 0) Q.sc[14] = P.beta[P.slot_of[r]];  // LDS: no global reload inside the loop
    // torsion trials, then with several chains six rigid trials per chain
    const int ntrial = P.nvr + (!LEAN && S.nml > 1 ? 6 * S.nml : 0);
    // speculative batches for this sweep's torsion trials: not from a
    // clashing state (those trials take the full recomputation)
    __syncthreads();
    const bool spec_sweep = spec_on && fabs(Q.sc[0]) < 1e5;
    for (int nsw = 0; nsw < ntrial; ++nsw) {
      const bool rigid = !LEAN && nsw >= P.nvr;  // uniform
      if (spec_sweep && !rigid) {
        // ---- speculative batch of the next nb torsion trials.  The G
        //      blocks hold the same state and twister; each draws all nb
        //      proposals as if every earlier trial of the batch is rejected
        //      (a rejected trial always consumed its acceptance number, so
        //      the positions are fixed), block t evaluates trial t, the
        //      blocks exchange the energy changes and take the decisions in
        //      order with the numbers drawn; the first accepted trial ends
        //      the batch and the twister is set back to the draws the
        //      decisions consumed.  Same decisions, same stream as one
        //      trial at a time.
        double* Z = Q.spec;  // [0,8G) proposals, [8G,16G) results, [16G,16G+8) own, then scalars
        const int nb = min(G, P.nvr - nsw);
        for (int k = tid; k < 625; k += blockDim.x) Q.mtsave[k] = Q.mt[k];
        __syncthreads();
        if (tid == 0) {
          uint32_t* mti = &Q.mt[624];
          int drawn = 0;
          for (int t = 0; t < nb; ++t) {
            double* T = Z + 8 * t;
            int ivar = nsw + t;
            if (P.rndord) {
              ivar = (int)(P.nvr * mt_grnd(Q.mt, mti));
              ++drawn;
              if (ivar >= P.nvr) ivar = P.nvr - 1;
            }
            const int jv = V.idvr[ivar];
            bool bgs = false;
            if (P.upch != 0 && V.xname[jv] == 0) {
              const double bu = P.upch_own ? Q.sc[14] : P.upch_beta;
              const double prob = P.upch == 1 ? P.bgsprob : (bu - P.up2bmin) / (P.up2bmax - P.up2bmin);
              bgs = mt_grnd(Q.mt, mti) < prob;
              ++drawn;
            }
            T[0] = (double)jv;
            T[5] = bgs ? 1.0 : 0.0;
            if (!bgs) {
              const double vold = Q.vl[jv];
              const double dv = V.ax[jv] * (mt_grnd(Q.mt, mti) - 0.5);
              ++drawn;
              const double vnew = d_addang(vold, dv);
              T[1] = vold;
              T[2] = vnew;
              T[3] = d_difang(vold, vnew);
              T[6] = (double)drawn;
              T[4] = mt_grnd(Q.mt, mti);  // its acceptance number, used when the change is uphill
              ++drawn;
              T[7] = (double)drawn;
            } else {
              T[6] = T[7] = (double)drawn;
            }
          }
          Z[16 * G + 8] = (double)drawn;
          Z[16 * G + 9] = (double)(epoch & 1u);
        }
        __syncthreads();
        lap(kPhPropose);
        const bool mine = g < nb && Z[8 * g + 5] == 0.0;
        const int par = (int)Z[16 * G + 9];
        int32_t* const nf_rows = P.spec_nf + (size_t)(r * 2 + par) * G * (nh + 1);
        Work wt;
        if (mine) {
          const double* T = Z + 8 * g;
          wt = spec_trial<MIXED>(P.sys, nvt, lds_flags, P.axvr, (int)T[0], T[1], T[2], T[3],
                                 nf_rows + (size_t)g * (nh + 1), Z + 16 * G);
        }
        lap(kPhSolvent);
        if (tid == 0) {
          double v[kPartVals];
          for (int k = 0; k < kPartVals; ++k) v[k] = (mine && k < 6) ? Z[16 * G + k] : 0.0;
          group_gather(P, r, g, epoch, v, Z + 8 * G);
        }
        __syncthreads();
        lap(kPhReduce);
        if (tid == 0) {  // the decisions in order, identical in the G blocks (metropolis.f:72-87)
          const double eol = Q.sc[0], beta = Q.sc[14];
          int used = nb, kstar = -1, cons = (int)Z[16 * G + 8];
          for (int t = 0; t < nb; ++t) {
            const double* T = Z + 8 * t;
            if (T[5] != 0.0) {  // inert BGS pick: accanalyze(2, 0)
              Q.sc[kScUpd + 1] += 1.0;
              Q.sc[kScUpd + 4] += 1.0;
              continue;
            }
            const double* R = Z + 8 * G + kPartVals * t;
            if (R[5] != 0.0) atomicOr(P.err, 2);
            const double enew = eol + (R[4] + R[0] + R[1] + R[2] + R[3]);
            double delta = weight_of(P, beta, enew) - weight_of(P, beta, eol);
            bool acc, drew = false;
            if (delta <= 0.0) {
              acc = true;
            } else {
              drew = true;
              delta = fmin(delta, 100.0);
              acc = exp(-delta) >= T[4];
            }
            Q.sc[kScUpd + 0] += 1.0;  // accanalyze(1, acc)
            Q.sc[kScUpd + 4] += 1.0;
            if (acc) {
              Q.sc[kScUpd + 5] += 1.0;
              Q.sc[kScUpd + 9] += 1.0;
            }
            if (trace && g == 0 && r == P.trace_rep && *P.trace_n < P.trace_cap) {
              double* tr = trace + 4 * (size_t)(*P.trace_n);
              tr[0] = T[0];
              tr[1] = enew;
              tr[2] = eol;
              tr[3] = acc ? 1.0 : 0.0;
              *P.trace_n += 1;
            }
            if (acc) {
              kstar = t;
              used = t + 1;
              cons = (int)(drew ? T[7] : T[6]);
              Q.sc[0] = enew;
              Q.sc[1] += 1.0;
              break;
            }
          }
          Z[16 * G + 10] = (double)kstar;
          Z[16 * G + 11] = (double)used;
          Z[16 * G + 12] = (double)cons;
        }
        __syncthreads();
        const int kstar = (int)Z[16 * G + 10], used = (int)Z[16 * G + 11], cons = (int)Z[16 * G + 12];
        if (cons < (int)Z[16 * G + 8]) {  // the twister back to the draws the decisions consumed
          for (int k = tid; k < 625; k += blockDim.x) Q.mt[k] = Q.mtsave[k];
          __syncthreads();
          if (tid == 0)
            for (int i = 0; i < cons; ++i) (void)mt_grnd(Q.mt, &Q.mt[624]);
        }
        if (kstar >= 0) {
          const double* T = Z + 8 * kstar;
          spec_apply(P.sys, nvt, lds_flags, P.axvr, (int)T[0], T[2], T[3], nf_rows + (size_t)kstar * (nh + 1));
        }
        if (mine && g < used) add_work(wk, wt);
        __syncthreads();
        lap(kPhAccept);
        nsw += used - 1;
        continue;
      }
      // ---- proposal (leader): metropolis.f:42-66; identical in the G blocks
      if (tid == 0 && rigid) {
        // metropolis.f:94-168: chain m, parameter i of gbpr; the move is
        // x -> c + M (x - c) + t over the chain's atoms
        const int q = nsw - P.nvr, m = q / 6, i = q - 6 * m;
        double* gm = Q.gb + 6 * m;
        const double gbol = gm[i];
        double gnew;
        double* T = Q.sc + kScRigid;  // [0..2] c, [3..11] M (row-major), [12..14] t
        for (int k = 0; k < 15; ++k) T[k] = 0.0;
        if (i < 3) {  // shift of the position, kept in the box
          gnew = gbol + (mt_grnd(Q.mt, &Q.mt[624]) - 0.5);
          if (S.boxsize > 0.0) {
            if (gnew > S.boxsize)
              gnew = S.boxsize;
            else if (gnew < -S.boxsize)
              gnew = -S.boxsize;
          }
          T[3] = T[7] = T[11] = 1.0;
          T[12 + i] = gnew - gbol;
        } else {  // a new Euler angle: turn about the first atom from the old axes to the new
          gnew = (mt_grnd(Q.mt, &Q.mt[624]) - 0.5) * (i == 4 ? kPi : kPi2);
          double eo[3] = {gm[3], gm[4], gm[5]}, en[3] = {gm[3], gm[4], gm[5]};
          en[i - 3] = gnew;
          double ao[3][3], an[3][3];
          euler_frame(eo, ao);
          euler_frame(en, an);
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) T[3 + 3 * a + b] = an[a][0] * ao[b][0] + an[a][1] * ao[b][1] + an[a][2] * ao[b][2];
          const int a1 = S.mol_a1[m];
          T[0] = L.x[a1];
          T[1] = L.y[a1];
          T[2] = L.z[a1];
        }
        Q.sc[2] = gnew;
        Q.sc[15] = gbol;
        Q.sc[6] = (double)(-1 - q);  // trial id: negative for rigid moves
        Q.sc[5] = fabs(Q.sc[0]) >= 1e5 ? 1.0 : 0.0;
        Q.sc[17] = Q.sc[18] = 0.0;  // no torsion term changes
        Q.sc[21] = 0.0;
      }
      if (tid == 0 && !rigid) {
        uint32_t* mti = &Q.mt[624];
        int ivar = nsw;
        if (P.rndord) {  // 1+int(nvr*grnd()); u = 1 exactly would index past the end
          ivar = (int)(P.nvr * mt_grnd(Q.mt, mti));
          if (ivar >= P.nvr) ivar = P.nvr - 1;
        }
        const int jv = V.idvr[ivar];
        // update type (metropolis.f:46-57): a BGS pick is an inert trial
        bool bgs = false;
        if (P.upch != 0 && V.xname[jv] == 0) {
          const double bu = P.upch_own ? Q.sc[14] : P.upch_beta;
          const double prob = P.upch == 1 ? P.bgsprob : (bu - P.up2bmin) / (P.up2bmax - P.up2bmin);
          bgs = mt_grnd(Q.mt, mti) < prob;
        }
        Q.sc[21] = bgs ? 1.0 : 0.0;
        if (bgs) {  // accanalyze(2, 0)
          Q.sc[kScUpd + 1] += 1.0;
          Q.sc[kScUpd + 4] += 1.0;
        }
        const double vold = Q.vl[jv];
        const double dv = bgs ? 0.0 : V.ax[jv] * (mt_grnd(Q.mt, mti) - 0.5);
        const double vnew = d_addang(vold, dv);
        Q.sc[2] = vnew;
        Q.sc[15] = vold;
        Q.sc[3] = d_difang(vold, vnew);
        Q.sc[6] = (double)jv;
        Q.sc[5] = fabs(Q.sc[0]) >= 1e5 ? 1.0 : 0.0;
        const int b = V.base[jv], a0 = V.axis0[jv];
        double ux = L.x[b] - L.x[a0], uy = L.y[b] - L.y[a0], uz = L.z[b] - L.z[a0];
        double un = rsqrt_f64(ux * ux + uy * uy + uz * uz);
        Q.sc[8] = L.x[b]; Q.sc[9] = L.y[b]; Q.sc[10] = L.z[b];
        Q.sc[11] = ux * un; Q.sc[12] = uy * un; Q.sc[13] = uz * un;
      }
      __syncthreads();
      lap(kPhPropose);
      if (Q.sc[21] != 0.0) {  // inert BGS pick (uniform); the barrier keeps thread 0's
        __syncthreads();      // next proposal from overwriting the flag before all read it
        continue;
      }
      const int jv = (int)Q.sc[6];  // variable, or -1 - (6 chain + parameter)
      int m1, m2, h1, h2;
      if (!rigid) {
        m1 = V.m1[jv];
        m2 = V.m2[jv];
        h1 = V.h1[jv];
        h2 = V.h2[jv];
      } else {
        const int m = (-1 - jv) / 6;
        m1 = S.mol_a1[m];
        m2 = S.mol_a2[m];
        h1 = S.mol_h1[m];
        h2 = S.mol_h2[m];
      }
      const bool full = Q.sc[5] != 0.0;
      if (DA && !rigid) {
        // ---- delayed acceptance, first stage: the network's energy change
        const float ytrial = da_trial(D, P.da, jv, Q.sc[15], Q.sc[2]);
        if (tid == 0) {
          const double ds = ((double)ytrial - (double)D.y[0]) * P.da_scale;
          if (trace && g == 0 && r == P.trace_rep && *P.trace_n < P.trace_cap) {
            // delayed acceptance traces its screens: variable, new value,
            // network change, 0 (rejected by the network; the exact stage
            // sets 1 or 2)
            double* t = trace + 4 * (size_t)(*P.trace_n);
            t[0] = jv;
            t[1] = Q.sc[2];
            t[2] = ds;
            t[3] = 0.0;
            *P.trace_n += 1;
          }
          const double d1 = Q.sc[14] * ds;
          bool pass = true;
          if (d1 > 0.0) pass = exp(-fmin(d1, 100.0)) >= mt_grnd(Q.mt, &Q.mt[624]);
          Q.sc[19] = ds;
          Q.sc[20] = pass ? 1.0 : 0.0;
          D.stats[0] += 1.0;
          if (!pass) {  // accanalyze(1, 0)
            D.stats[1] += 1.0;
            Q.sc[kScUpd + 0] += 1.0;
            Q.sc[kScUpd + 4] += 1.0;
          }
        }
        __syncthreads();
        if (Q.sc[20] == 0.0) continue;  // rejected without the exact energy (uniform)
      }
      EnergyOut vref;  // verify mode (G = 1): full energy before the move (thread 0)
      if (verify && !full) {
        const ReplicaEnergy re = replica_energy<MIXED>(P.sys, nvt, lds_flags, 0, false, 0, 1);
        vref = re.o;
        __syncthreads();
      }
      if (rigid) {  // ---- rigid move of a whole chain
        const double* T = Q.sc + kScRigid;
        const double cx = T[0], cy = T[1], cz = T[2];
        if (tid == 0 && g == 0) wk.moved += m2 - m1 + 1;
        for (int a = m1 + tid; a <= m2; a += blockDim.x) {
          const double vx = L.x[a] - cx, vy = L.y[a] - cy, vz = L.z[a] - cz;
          Q.bx[a - m1] = L.x[a];
          Q.by[a - m1] = L.y[a];
          Q.bz[a - m1] = L.z[a];
          const double xn = cx + (T[3] * vx + T[4] * vy + T[5] * vz) + T[12];
          const double yn = cy + (T[6] * vx + T[7] * vy + T[8] * vz) + T[13];
          const double zn = cz + (T[9] * vx + T[10] * vy + T[11] * vz) + T[14];
          L.x[a] = xn;
          L.y[a] = yn;
          L.z[a] = zn;
          const int h = L.hidx[a];
          if (h >= 0) {
            L.hp[4 * h + 0] = xn;
            L.hp[4 * h + 1] = yn;
            L.hp[4 * h + 2] = zn;
          }
        }
        for (int h = tid; h < nh; h += blockDim.x) Q.aff[h] = 0;
        if (tid == 0) Q.gb[-1 - jv] = Q.sc[2];
      } else {
        // ---- rigid rotation of the moved atoms about the bond (Rodrigues)
        double sn, cs;
        sincos(V.sign[jv] * Q.sc[3], &sn, &cs);
        const double px = Q.sc[8], py = Q.sc[9], pz = Q.sc[10];
        const double kx = Q.sc[11], ky = Q.sc[12], kz = Q.sc[13];
        if (tid == 0 && g == 0) wk.moved += m2 - m1 + 1;
        for (int a = m1 + tid; a <= m2; a += blockDim.x) {
          const double vx = L.x[a] - px, vy = L.y[a] - py, vz = L.z[a] - pz;
          Q.bx[a - m1] = L.x[a];
          Q.by[a - m1] = L.y[a];
          Q.bz[a - m1] = L.z[a];
          const double kv = kx * vx + ky * vy + kz * vz;
          const double cx = ky * vz - kz * vy, cy = kz * vx - kx * vz, cz = kx * vy - ky * vx;
          const double t = kv * (1.0 - cs);
          const double xn = px + vx * cs + cx * sn + kx * t;
          const double yn = py + vy * cs + cy * sn + ky * t;
          const double zn = pz + vz * cs + cz * sn + kz * t;
          L.x[a] = xn;
          L.y[a] = yn;
          L.z[a] = zn;
          const int h = L.hidx[a];
          if (h >= 0) {
            L.hp[4 * h + 0] = xn;
            L.hp[4 * h + 1] = yn;
            L.hp[4 * h + 2] = zn;
          }
        }
        // torsion term of the variable, new and old angle: here on two idle
        // lanes of the last wave in vacuo; with the solvent term that wave
        // does it at the start of the solvent phase instead, whose atom queue
        // absorbs it (the rotation phase waits for it otherwise)
        if (S.itysol <= 0) torsion_pair(S, Q, V, jv, tid - ((int)blockDim.x - kWave), g);
        for (int h = tid; h < nh; h += blockDim.x) Q.aff[h] = 0;
        if (tid == 0) Q.vl[jv] = Q.sc[2];
      }
      __syncthreads();
      lap(kPhRotate);
      double enew;
      if (!full) {
        // ---- incremental energy change; block g takes pair items and
        //      heavy atoms k with k % G == g
        PairE d = {0.0, 0.0, 0.0};
        double* pst = P.phase ? Q.sc + 96 + 4 * wave : nullptr;  // this wave's pair timing (lane 0)
        double tpa = pst ? (double)wall_clock64() : 0.0;
        pair_delta<MIXED>(S, Q, m1, m2, g + G * wave, G * nwaves, d, wk);
        if (pst && lane == 0) {
          const double now = (double)wall_clock64();
          pst[0] += now - tpa;
          tpa = now;
        }
        if (S.itysol > 0) mark_affected(S, Q, h1, h2, m1, g, G);
        if (pst && lane == 0) {
          const double now = (double)wall_clock64();
          pst[1] += now - tpa;
          tpa = now;
        }
        __syncthreads();
        if (P.phase && tid == 0) {
          double mx = 0.0;
          for (int w = 0; w < nwaves; ++w) {
            double* st = Q.sc + 96 + 4 * w;
            mx = fmax(mx, st[0]);
            ph[kPhPairSum] += st[0];
            ph[kPhMarkSum] += st[1];
            ph[kPhTorsion] += st[2];
            st[0] = st[1] = st[2] = st[3] = 0.0;
          }
          ph[kPhPairMax] += mx;
        }
        int* next_atom = reinterpret_cast<int*>(Q.sc + 30);  // LDS work counter of the solvent loop
        if (S.itysol > 0) {
          const int na = wave == 0 ? list_affected(S, Q, g, G) : 0;
          if (tid == 0) {
            Q.sc[29] = (double)na;
            *next_atom = 0;
          }
          __syncthreads();
        }
        lap(kPhPairs);
        double dsol = 0.0;
        int err = 0;
        if (S.itysol > 0) {
          double* nb = L.fr + (size_t)wave * kNbDoubles;
          const int na = (int)Q.sc[29];
          double* wst = P.phase ? Q.sc + 64 + 4 * wave : nullptr;  // this wave's timing (lane 0)
          const double tw0 = wst ? (double)wall_clock64() : 0.0;
          if (wave == nwaves - 1 && !rigid) torsion_pair(S, Q, V, jv, lane, g);  // then this wave joins the atom queue
          // waves take the next atom as they finish (atom costs vary with their contacts)
          for (;;) {
            int idx = 0;
            if (lane == 0) idx = atomicAdd(next_atom, 1);
            idx = __shfl(idx, 0, kWave);
            if (idx >= na) break;
            const int h = Q.alist[idx];
            const int i = L.hv[h];
            int nfree = solvent_atom(S, L, nb, i, wk, P.bits_lds, wst);
            if (nfree < 0) {
              err = 1;
              nfree = Q.nf[h];
            }
            if (lane == 0) Q.nf2[h] = nfree;
          }
          if (wst && lane == 0) wst[3] = (double)wall_clock64() - tw0;
        }
        double v[3] = {d.el, d.vw, d.hb};
#pragma unroll
        for (int k = 0; k < 3; ++k) v[k] = wave_sum(v[k]);
        const int errw = __any(err);
        if (lane == 0) {
          for (int k = 0; k < 3; ++k) L.red[8 * wave + k] = v[k];
          L.red[8 * wave + 6] = (double)errw;
        }
        __syncthreads();
        if (S.itysol > 0 && wave == 0) {
          // summed in list order, so the result does not depend on which
          // wave took which atom
          const int na = (int)Q.sc[29];
          for (int k = lane; k < na; k += kWave) {
            const int h = Q.alist[k];
            const int i = L.hv[h];
            const double trad = L.rv[i];
            dsol += L.sg[i] * (c4 * trad * trad * (double)(Q.nf2[h] - Q.nf[h]));
          }
          dsol = wave_sum(dsol);
        }
        lap(kPhSolvent);
        if (P.phase && tid == 0 && S.itysol > 0) {
          double mx = 0.0;
          for (int w = 0; w < nwaves; ++w) {
            double* st = Q.sc + 64 + 4 * w;
            mx = fmax(mx, st[3]);
            ph[kPhSolGather] += st[0];
            ph[kPhSolPoints] += st[1];
            ph[kPhSolAtoms] += st[2];
            st[0] = st[1] = st[2] = st[3] = 0.0;
          }
          ph[kPhSolMax] += mx;
        }
        if (tid == 0) {
          const double dt = g == 0 ? Q.sc[17] - Q.sc[18] : 0.0;  // torsion change (torsion_pair)
          double tot[6] = {0, 0, 0, dsol, dt, 0};  // el, vw, hb, sol, torsion, error
          for (int w = 0; w < nwaves; ++w) {
            for (int k = 0; k < 3; ++k) tot[k] += L.red[8 * w + k];
            tot[5] += L.red[8 * w + 6];
          }
          group_sum(P, r, g, epoch, tot, 6, P.phase ? ph : nullptr);
          if (tot[5] != 0.0) atomicOr(P.err, 2);
          Q.sc[16] = Q.sc[0] + (tot[4] + tot[0] + tot[1] + tot[2] + tot[3]);
          for (int k = 0; k < 4; ++k) Q.sc[24 + k] = tot[k];  // el, vw, hb, sol deltas (verify mode)
          Q.sc[28] = tot[4];
        }
        if (P.phase) __syncthreads();  // uniform: timing only
        lap(kPhReduce);
      } else {
        // ---- full recomputation from the trial coordinates
        const ReplicaEnergy re = replica_energy<MIXED>(P.sys, nvt, lds_flags, 2, false, g, G);
        add_work(wk, re.w);
        if (tid == 0) {
          double v[2] = {re.o.el + re.o.vw + re.o.hb + re.o.tor + re.o.sol, (double)re.o.err};
          group_sum(P, r, g, epoch, v, 2);
          Q.sc[16] = v[0];
          if (v[1] != 0.0) atomicOr(P.err, 2);
          ph[kPhNFull] += 1.0;
        }
        if (P.phase) __syncthreads();
        lap(kPhFull);
      }
      if (verify && !full) {  // compare each incremental component with a full recomputation
        const ReplicaEnergy rn = replica_energy<MIXED>(P.sys, nvt, lds_flags, 0, false, 0, 1);
        const EnergyOut& vnew = rn.o;
        if (tid == 0 && P.dbg[16 * r] == 0.0) {
          const double fd[5] = {vnew.el - vref.el, vnew.vw - vref.vw, vnew.hb - vref.hb, vnew.tor - vref.tor,
                                vnew.sol - vref.sol};
          const double id[5] = {Q.sc[24], Q.sc[25], Q.sc[26], Q.sc[28], Q.sc[27]};
          const double scale = 1.0 + fabs(vref.el) + fabs(vref.vw) + fabs(vref.hb) + fabs(vref.sol);
          // mixed: both sides carry FP32 pair-term rounding (each term within
          // 2^-22 of its magnitude; the full sums' magnitudes bound the moved
          // pairs' too)
          const double mtol = MIXED ? 1e-5 * (vref.mag + vnew.mag) : 0.0;
          bool bad = false;
          for (int k = 0; k < 5; ++k) bad |= fabs(fd[k] - id[k]) > 1e-9 * (scale + fabs(fd[k])) + (k < 3 ? mtol : 0.0);
          if (bad) {
            double* d = P.dbg + 16 * r;
            d[0] = 1.0;
            d[1] = sw;
            d[2] = nsw;
            d[3] = jv;
            for (int k = 0; k < 5; ++k) {
              d[4 + k] = id[k];
              d[9 + k] = fd[k];
            }
            atomicOr(P.err, 4);
          }
        }
        __syncthreads();
      }
      // ---- acceptance (leader): metropolis.f:72-87; identical in the G blocks
      if (tid == 0) {
        enew = Q.sc[16];
        const double eol = Q.sc[0];
        const double beta = Q.sc[14];
        double delta = weight_of(P, beta, enew) - weight_of(P, beta, eol);
        const bool da = DA && !rigid;
        if (da) delta -= beta * Q.sc[19];  // second stage: exact change against the network's
        bool acc;
        if (delta <= 0.0) {
          acc = true;
        } else {
          const double rd = mt_grnd(Q.mt, &Q.mt[624]);
          delta = fmin(delta, 100.0);
          acc = exp(-delta) >= rd;
        }
        Q.sc[4] = acc ? 1.0 : 0.0;
        {  // accanalyze(iupt, iupstate): 1 torsion, 3 shift, 4 new Euler angle
          const int iu = rigid ? (((-1 - jv) % 6) < 3 ? 2 : 3) : 0;
          Q.sc[kScUpd + iu] += 1.0;
          Q.sc[kScUpd + 4] += 1.0;
          if (acc) {
            Q.sc[kScUpd + 5 + iu] += 1.0;
            Q.sc[kScUpd + 9] += 1.0;
          }
        }
        if (trace && g == 0 && r == P.trace_rep) {
          if (da) {  // the screen's record of this trial: 1 rejected exactly, 2 accepted
            if (*P.trace_n > 0) trace[4 * (size_t)(*P.trace_n - 1) + 3] = acc ? 2.0 : 1.0;
          } else if (*P.trace_n < P.trace_cap) {
            double* t = trace + 4 * (size_t)(*P.trace_n);
            t[0] = jv;
            t[1] = enew;
            t[2] = eol;
            t[3] = acc ? 1.0 : 0.0;
            *P.trace_n += 1;
          }
        }
        if (da) {
          D.stats[2] += 1.0;
          if (acc) D.stats[3] += 1.0;
        }
        if (acc) {
          Q.sc[0] = enew;
          Q.sc[1] += 1.0;
        } else if (rigid) {
          Q.gb[-1 - jv] = Q.sc[15];
        } else {
          Q.vl[jv] = Q.sc[15];
        }
      }
      __syncthreads();
      if (Q.sc[4] != 0.0) {
        if (S.itysol > 0) {
          for (int h = tid; h < nh; h += blockDim.x)
            if (full || Q.aff[h]) Q.nf[h] = Q.nf2[h];
        }
        if (DA && !rigid) da_commit(D);  // the trial's layer-1 sums and output become the state's
      } else {
        for (int a = m1 + tid; a <= m2; a += blockDim.x) {
          const double xo = Q.bx[a - m1], yo = Q.by[a - m1], zo = Q.bz[a - m1];
          L.x[a] = xo;
          L.y[a] = yo;
          L.z[a] = zo;
          const int h = L.hidx[a];
          if (h >= 0) {
            L.hp[4 * h + 0] = xo;
            L.hp[4 * h + 1] = yo;
            L.hp[4 * h + 2] = zo;
          }
        }
      }
      __syncthreads();
      lap(kPhAccept);
    }
    // ---- sweep end: rebuild from the variables and check (metropolis.f:172-179)
    ReplicaEnergy re = replica_energy<MIXED>(P.sys, nvt, lds_flags, 1, true, g, G);
    if (g != 0) re.w.builds = 0.0;
    add_work(wk, re.w);
    if (spec_on) share_free_points(S, Q, gnf, g, G, true);
    if (P.obs && g == 0) {  // rgyr(0, ...) of the rebuilt chain (partem_s.f:186-193, rgyr.f:54-140)
      double m[4] = {0.0, 0.0, 0.0, 0.0};
      for (int a = tid; a < S.nat; a += blockDim.x) {
        const double x = L.x[a], y = L.y[a], z = L.z[a];
        m[0] += x;
        m[1] += y;
        m[2] += z;
        m[3] += x * x + y * y + z * z;
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) m[k] = wave_sum(m[k]);
      __syncthreads();
      if (lane == 0)
        for (int k = 0; k < 4; ++k) L.red[8 * wave + k] = m[k];
      __syncthreads();
      if (tid == 0) {
        for (int k = 0; k < 4; ++k) m[k] = 0.0;
        for (int w = 0; w < nwaves; ++w)
          for (int k = 0; k < 4; ++k) m[k] += L.red[8 * w + k];
        const double dn = (double)S.nat, dx = m[0] / dn, dy = m[1] / dn, dz = m[2] / dn, d2 = m[3] / dn;
        const double rg2 = d2 - (dx * dx + dy * dy + dz * dz);
        double ee = 0.0;
        if (S.ca_first >= 0 && S.ca_last >= 0) {
          const double ex = L.x[S.ca_last] - L.x[S.ca_first], ey = L.y[S.ca_last] - L.y[S.ca_first],
                       ez = L.z[S.ca_last] - L.z[S.ca_first];
          ee = sqrt(ex * ex + ey * ey + ez * ez);
        }
        double* o = P.obs + 2 * ((size_t)(traj_base + sw) * P.nrep + P.slot_of[r]);
        o[0] = rg2 > 0.0 ? sqrt(rg2) : 0.0;
        o[1] = ee;
      }
      __syncthreads();
    }
    if (tid == 0) {
      double v[3] = {re.o.el + re.o.vw + re.o.hb + re.o.tor + re.o.sol, (double)re.o.err, re.o.mag};
      group_sum(P, r, g, epoch, v, 3);
      const double enw = v[0];
      const double eol = Q.sc[0];
      // metropolis.f:172-179.  Mixed mode: the tracked energy starts from the
      // previous FP32-term recomputation, so two recomputations' rounding
      // (each within 5e-6 of the summed pair-term magnitudes) are allowed.
      const double tol = MIXED ? 0.000001 + 1e-5 * v[2] : 0.000001;
      if (fabs(eol - enw) > tol && eol < 100000.0) {
        atomicOr(P.err, 1);
        if (g == 0 && P.dbg[16 * r + 14] == 0.0) {  // first drift of this replica
          P.dbg[16 * r + 14] = eol;
          P.dbg[16 * r + 15] = enw;
        }
      }
      if (v[1] != 0.0) atomicOr(P.err, 2);
      Q.sc[0] = enw;
      if (g == 0) {
        const int slot = P.slot_of[r];
        if (P.traj) P.traj[(size_t)(traj_base + sw) * P.nrep + slot] = enw;
        if (P.snap && (traj_base + sw + 1) % P.snap_every == 0) {
          const size_t k = (size_t)((traj_base + sw + 1) / P.snap_every - 1);
          P.snap_e[k * P.nrep + r] = enw;
          for (int v = 0; v < nvt; ++v) P.snap[(k * P.nrep + r) * nvt + v] = Q.vl[v];
        }
        if (P.hist) {
          double hv = enw / P.ebin + copysign(0.5, enw);
          hv = fmin(fmax(hv, (double)P.kmin), (double)P.kmax);
          atomicAdd(&P.hist[(int)hv - P.kmin], 1);
        }
        if (P.pt_exchange) P.eol[r] = enw;
      }
    }
    __syncthreads();
    if (spec_on) share_free_points(S, Q, gnf, g, G, false);
    if (P.pt_exchange) {  // exchange attempt of this sweep (partem_s.f:204-232)
      if (tid == 0) grid_sync(P, gepoch);
      __syncthreads();
      if (blockIdx.x == 0 && tid == 0 && P.nrep > 1)
        exchange_step(P.nrep, P.ipoi, const_cast<int32_t*>(P.slot_of), P.beta, P.eol, P.mt_x, P.iswitch);
      __syncthreads();
      if (tid == 0) grid_sync(P, gepoch);
      __syncthreads();
    }
    lap(kPhSweepEnd);
  }
  if (P.phase && tid == 0 && g == 0) {
    ph[kPhCoreCycles] = (double)(clock64() - core0);  // core clock vs the constant wall clock
    ph[kPhRealTicks] = (double)(wall_clock64() - real0);
    for (int k = 0; k < kNPhase; ++k) P.phase[kNPhase * (size_t)r + k] += ph[k];
  }
  // ---- write the replica state back (block 0 of the group; each block its
  //      own free-point counts)
  if (g == 0) {
    for (int k = tid; k < nvt; k += blockDim.x) gvl[k] = Q.vl[k];
    for (int k = tid; k < 625; k += blockDim.x) gmt[k] = Q.mt[k];
    for (int k = tid; k < 6 * S.nml; k += blockDim.x) ggb[k] = Q.gb[k];
    for (int a = tid; a < nat; a += blockDim.x) {
      gx[a] = L.x[a];
      gx[np + a] = L.y[a];
      gx[2 * np + a] = L.z[a];
    }
    if (tid == 0) {
      P.eol[r] = Q.sc[0];
      P.acc[r] = Q.sc[1];
      P.valid[r] = 1;
      if (DA && P.da_stats)
        for (int k = 0; k < 4; ++k) P.da_stats[4 * (size_t)r + k] += D.stats[k];
      if (P.upd)
        for (int k = 0; k < 10; ++k) P.upd[10 * (size_t)r + k] += Q.sc[kScUpd + k];
    }
  }
  for (int h = tid; h < nh; h += blockDim.x)
    if (h % G == g) gnf[h] = Q.nf[h];
  if (P.work) {
    double v[4] = {wk.pairs, wk.contacts, wk.builds, wk.moved};
    for (int k = 0; k < 4; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0)
      for (int k = 0; k < 4; ++k) L.red[8 * wave + k] = v[k];
    __syncthreads();
    if (tid == 0)
      for (int k = 0; k < 4; ++k) {
        double t = 0.0;
        for (int w = 0; w < nwaves; ++w) t += L.red[8 * w + k];
        atomicAdd(&P.work[4 * r + k], t);
      }
  }
}

// One exchange attempt as its own launch (the per-sweep launch path).
__global__ void k_exchange(int nrep, int32_t* ipoi, int32_t* slot_of, const double* beta, const double* eol,
                           uint32_t* mt_x, int32_t* iswitch) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  exchange_step(nrep, ipoi, slot_of, beta, eol, mt_x, iswitch);
}

}  // namespace smmp
