// Device-side system description and small math helpers for gfx950.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace smmp {

constexpr int kWave = 64;
constexpr int kTy = 18;          // ECEPP atom types
constexpr int kMaxNb = 64;       // solvent neighbours per chunk of one atom's gather
constexpr int kMaxMol = 10;      // chains per system (mxml)
constexpr double kPi = 3.141592653589793;
constexpr double kPi2 = 2.0 * kPi;

// Read-only description of one system of chains in device memory (see HostSys).
struct DevSys {
  int32_t nat, nvt, nlev, words, nheavy, npnt, epsd, itysol, np;  // np: nat rounded up to 64
  int32_t nml;                           // chains
  int16_t mol_a1[kMaxMol], mol_a2[kMaxMol];  // first and last atom of each chain
  int16_t mol_i2[kMaxMol], mol_i3[kMaxMol];  // its second and third backbone atoms
  int16_t mol_h1[kMaxMol], mol_h2[kMaxMol];  // its heavy atoms: heavy indices [h1, h2)
  int16_t ca_first, ca_last;             // C-alpha atoms of the first and last residue, -1 if none
  double conv, plt, slp, boxsize;
  double gbpr[6 * kMaxMol];              // global parameters of the chains (the context's)
  const int16_t* mol_of;                 // per atom: its chain
  const int16_t* parent;
  const int16_t* lev_atoms;
  const int32_t* lev_off;
  const int8_t* type;
  const double *q, *bl, *cba, *sba, *toat0, *tref, *rvdw, *sigma;
  const int16_t* tvar;
  const int16_t *var_atom, *var_base, *var_axis0, *var_m1, *var_m2;
  const int16_t *var_h1, *var_h2;  // heavy-index range [h1, h2) of the moved atoms
  const double* var_sign;  // rotation of [m1,m2] is sign * (change of the torsion)
  const double *var_e0, *var_sg, *var_rn;
  const uint64_t* bits;
  const int16_t* heavy;
  const int16_t* heavy_of;  // per atom: index in heavy, or -1
  const double *pA, *pA14, *pC;
  const int32_t* pHB;
  const double* spoint;  // npnt*3
};

// Work done by one thread, for the algorithmic flop count of a launch:
// pair-term evaluations, solvent contacts (each tested against every
// tessellation point), atoms placed by the chain build, atoms rotated.
struct Work {
  double pairs = 0, contacts = 0, builds = 0, moved = 0;
};

// difang/addang of difang.f with the same fmod/sign steps.
__device__ __forceinline__ double d_difang(double a1, double a2) {
  double d = fmod(a2 - a1, kPi2);
  return fabs(d) <= kPi ? d : d - copysign(kPi2, d);
}
__device__ __forceinline__ double d_addang(double a1, double a2) {
  double d = fmod(a1 + a2, kPi2);
  return fabs(d) <= kPi ? d : d - copysign(kPi2, d);
}

// 1/sqrt(x) in FP64: hardware estimate refined by two Newton steps.
__device__ __forceinline__ double rsqrt_f64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double h = 0.5 * x;
  double e = fma(-h * y, y, 0.5);
  y = fma(y, e, y);
  e = fma(-h * y, y, 0.5);
  y = fma(y, e, y);
  return y;
}

// Wave-wide sum (wave64) by butterfly shuffles.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Orders LDS traffic of one wave (cross-lane hand-off without s_barrier).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace smmp
