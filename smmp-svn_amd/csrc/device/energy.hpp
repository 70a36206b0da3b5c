// Full ECEPP/3 (+OONS) energy of one conformation by one workgroup:
//   1. torsions from the variables (setvar.f:43-57) and their sin/cos,
//   2. Eyring chain build level by level (bldmol.f:13-93, one wave),
//   3. pair terms over 64x64 tiles of the interaction bit rows
//      (enyshe.f:112-184; Coulomb, 6-12 LJ, 10-12 H-bond, 1-4 with a14),
//   4. torsion terms (enyshe.f:73-79),
//   5. OONS solvation by 122-point occlusion (enysol.f:192-324).
// MIXED = FP32 pair terms (per-tile FP32 partial sums, FP64 totals); the
// chain, the solvent term and every total stay FP64.
#pragma once
#include "common.hpp"

namespace smmp {

// A pointer into LDS typed as such, so that loads through it are ds_read
// even where the same data may live in HBM on another path.
typedef const uint64_t __attribute__((address_space(3))) lds_u64;

// LDS carve-up for one conformation (doubles unless noted).
struct EnergyLds {
  double *x, *y, *z, *q;   // np each
  double* fr;              // frames: 6*np (or the solvent scratch)
  double* ptab;            // pair types (18x18) x {A, A14, C, H-bond flag}: 4 * 324
  double* sp;              // npnt*3
  double* red;             // nwaves*8, then one slot holding the solvent loop's atom counter
  int* ctr;                // that counter
  double *rv, *sg;         // solvent radius and sigma per atom (np each)
  double* hp;              // heavy atoms packed by heavy index: x, y, z, radius (4 * nheavy)
  int32_t* nfs;            // free points per heavy atom of a full evaluation (nheavy)
  const uint64_t* bits;    // pair bit rows: an LDS copy, or the system's array in HBM
  const lds_u64* lbits;    // the LDS copy as an LDS pointer (ds_read), or null
  int8_t* t;               // np bytes
  int16_t* hv;             // heavy-atom list (nheavy)
  int16_t* hidx;           // per atom (np): index in hv, or -1
  float *xf, *yf, *zf, *qf;  // mixed-mode copies (alias fr after the build)
};

// Number of u64 words of the pair bit rows (two planes per 64-column word).
__host__ __device__ inline size_t pair_bit_words(int nat, int words) { return (size_t)nat * words * 2; }

// Per-wave solvent scratch: kMaxNb neighbours x kNbFloats floats.
constexpr int kNbFloats = 8;  // per neighbour: 2d (3), 1/M, -t/M, j, 2 unused (see solvent_points)
constexpr int kNbDoubles = kMaxNb * kNbFloats / 2;
// The frames of the chain build (6 per atom), overlaid after the build by
// the solvent scratch of every wave when the system has a numerical solvent
// (npnt > 0).
__host__ __device__ inline int energy_region1(int np, int nwaves, int npnt) {
  int a = 6 * np, b = npnt > 0 ? nwaves * kNbDoubles : 0;
  return a > b ? a : b;
}
// Per-atom solvent radius and sigma: only with a numerical solvent.
__host__ __device__ inline int energy_solvent_atoms(int np, int npnt) { return npnt > 0 ? np : 0; }
// nbits: u64 words of pair bits staged in LDS (0: read from HBM).
__host__ __device__ inline size_t energy_lds_bytes(int np, int npnt, int nwaves, int nheavy, size_t nbits) {
  size_t d = 4 * (size_t)np + 2 * (size_t)energy_solvent_atoms(np, npnt) + energy_region1(np, nwaves, npnt) +
             4 * 324 + 3 * (size_t)npnt + 8 * (size_t)nwaves + 8 + nbits + 4 * (size_t)nheavy +
             ((size_t)nheavy + 1) / 2;
  return d * sizeof(double) + (size_t)np + 2 * (size_t)nheavy + 2 * (size_t)np;
}

__device__ inline EnergyLds carve_energy_lds(double* base, const DevSys& S, int nwaves, bool bits_lds) {
  EnergyLds L;
  const int np = S.np, npnt = S.npnt;
  double* p = base;
  L.x = p; p += np;
  L.y = p; p += np;
  L.z = p; p += np;
  L.q = p; p += np;
  L.fr = p; p += energy_region1(np, nwaves, npnt);
  L.ptab = p; p += 4 * 324;
  L.sp = p; p += 3 * npnt;
  L.red = p; p += 8 * nwaves;
  L.ctr = reinterpret_cast<int*>(p); p += 8;
  L.rv = p; p += energy_solvent_atoms(np, npnt);
  L.sg = p; p += energy_solvent_atoms(np, npnt);
  L.hp = p; p += 4 * S.nheavy;
  L.nfs = reinterpret_cast<int32_t*>(p); p += (S.nheavy + 1) / 2;
  if (bits_lds) {
    L.bits = reinterpret_cast<const uint64_t*>(p);
    L.lbits = (const lds_u64*)(p);
    p += pair_bit_words(S.nat, S.words);
  } else {
    L.bits = S.bits;
    L.lbits = nullptr;
  }
  L.t = reinterpret_cast<int8_t*>(p);
  L.hv = reinterpret_cast<int16_t*>(L.t + np);  // np is a multiple of 64
  L.hidx = L.hv + S.nheavy;
  L.xf = reinterpret_cast<float*>(L.fr);
  L.yf = L.xf + np;
  L.zf = L.yf + np;
  L.qf = L.zf + np;
  return L;
}

// Word k of the pair bit rows.
// The two sources in their own address spaces: a select of the two
// pointers would make every read a flat access.
__device__ __forceinline__ uint64_t pair_bits(const EnergyLds& L, size_t k) {
  if (L.lbits) return ((const __attribute__((address_space(3))) uint64_t*)L.lbits)[k];
  return ((const __attribute__((address_space(1))) uint64_t*)L.bits)[k];
}

// Torsion of atom a for the variable values vl (setvar.f:43-57 from the
// snapshot state toat0/tref).
__device__ __forceinline__ double atom_torsion(const DevSys& S, const double* vl, int a) {
  int v = S.tvar[a];
  double t0 = S.toat0[a];
  if (v < 0) return t0;
  double to = t0 + d_difang(S.tref[a], vl[v]);
  if (fabs(to) > kPi) to = to - copysign(kPi2, to);
  return to;
}

// Stage tables, tessellation points, charges and types (once per launch).
__device__ void energy_stage_static(const DevSys& S, EnergyLds& L) {
  const int tid = threadIdx.x, nt = blockDim.x, np = S.np;
  for (int k = tid; k < 324; k += nt) {
    L.ptab[4 * k + 0] = S.pA[k];
    L.ptab[4 * k + 1] = S.pA14[k];
    L.ptab[4 * k + 2] = S.pC[k];
    L.ptab[4 * k + 3] = (double)S.pHB[k];
  }
  for (int k = tid; k < 3 * S.npnt; k += nt) L.sp[k] = S.spoint[k];
  for (int h = tid; h < S.nheavy; h += nt) L.hv[h] = S.heavy[h];
  if (L.bits != S.bits) {
    uint64_t* b = const_cast<uint64_t*>(L.bits);
    const size_t n = pair_bit_words(S.nat, S.words);
    for (size_t k = tid; k < n; k += nt) b[k] = S.bits[k];
  }
  for (int a = tid; a < np; a += nt) {
    if (a < S.nat) {
      L.q[a] = S.q[a];
      L.t[a] = S.type[a];
      if (S.npnt > 0) {
        L.rv[a] = S.rvdw[a];
        L.sg[a] = S.sigma[a];
      }
      L.hidx[a] = S.heavy_of[a];
    } else {
      L.q[a] = 0.0;
      L.t[a] = 0;
      if (S.npnt > 0) {
        L.rv[a] = 0.0;
        L.sg[a] = 0.0;
      }
      L.hidx[a] = -1;
      L.x[a] = 0.0;
      L.y[a] = 0.0;
      L.z[a] = 0.0;
    }
  }
}

// Torsion sin/cos of every atom into its frame slot (read by eyring_atom).
__device__ void energy_stage_torsions(const DevSys& S, const double* vl, EnergyLds& L) {
  for (int a = threadIdx.x; a < S.nat; a += blockDim.x) {
    double s, c;
    sincos(atom_torsion(S, vl, a), &s, &c);
    L.fr[6 * a + 0] = c;  // cos(toat), replaced by the frame when built
    L.fr[6 * a + 1] = s;
  }
}

// bldmol.f:303-369 (setgbl): the global axes J, K, L (columns of ag) from
// the Euler angles e[0..2] = gbpr(4..6).
__device__ void euler_frame(const double* e, double ag[3][3]) {
  double ca = cos(e[0]), sa = sin(e[0]), cb = cos(e[1]), sb = sin(e[1]), cg = cos(e[2]), sg = sin(e[2]);
  double d;
  double x1 = ca * cg - sa * sb * sg, x2 = sa * cg + ca * sb * sg, x3 = -cb * sg;
  d = sqrt(x1 * x1 + x2 * x2 + x3 * x3);
  ag[0][0] = x1 / d; ag[1][0] = x2 / d; ag[2][0] = x3 / d;
  double y1 = -sa * cb, y2 = ca * cb, y3 = sb;
  d = sqrt(y1 * y1 + y2 * y2 + y3 * y3);
  ag[0][1] = y1 / d; ag[1][1] = y2 / d; ag[2][1] = y3 / d;
  double z1 = ca * sg + sa * sb * cg, z2 = sa * sg - ca * sb * cg, z3 = cb * cg;
  d = sqrt(z1 * z1 + z2 * z2 + z3 * z3);
  ag[0][2] = z1 / d; ag[1][2] = z2 / d; ag[2][2] = z3 / d;
}

// bldmol.f:303-420 (setgbl) and the first atom of chain m from its global
// parameters g[0..5]; one lane.
__device__ void build_root(const DevSys& S, EnergyLds& L, const double* g, int m) {
  double ag[3][3];
  euler_frame(g + 3, ag);
  double x1, x2, x3, y1, y3, z1, z2, z3;
  const int a1 = S.mol_a1[m], i2 = S.mol_i2[m], i3 = S.mol_i3[m];
  double ct2 = L.fr[6 * i2 + 0], st2 = L.fr[6 * i2 + 1], sa2 = S.sba[i2];
  x1 = -S.cba[i2]; x2 = -sa2 * ct2; x3 = -sa2 * st2;
  double dx = sqrt(x1 * x1 + x2 * x2 + x3 * x3);
  x1 /= dx; x2 /= dx; x3 /= dx;
  double ct3 = L.fr[6 * i3 + 0], st3 = L.fr[6 * i3 + 1];
  z1 = -st3 * (st2 * x3 + ct2 * x2);
  z2 = st3 * ct2 * x1 + ct3 * st2;
  z3 = st3 * st2 * x1 - ct3 * ct2;
  double dz = sqrt(z1 * z1 + z2 * z2 + z3 * z3);
  z1 /= dz; z2 /= dz; z3 /= dz;
  y1 = z2 * x3 - z3 * x2;
  y3 = z1 * x2 - z2 * x1;
  double xg[3], zg[3];
  for (int i = 0; i < 3; ++i) {
    xg[i] = ag[i][0] * x1 + ag[i][1] * y1 + ag[i][2] * z1;
    zg[i] = ag[i][0] * x3 + ag[i][1] * y3 + ag[i][2] * z3;
  }
  dx = sqrt(xg[0] * xg[0] + xg[1] * xg[1] + xg[2] * xg[2]);
  dz = sqrt(zg[0] * zg[0] + zg[1] * zg[1] + zg[2] * zg[2]);
  L.x[a1] = g[0];
  L.y[a1] = g[1];
  L.z[a1] = g[2];
  double* f = L.fr + 6 * a1;
  f[0] = xg[0] / dx;
  f[1] = xg[1] / dx;
  f[2] = xg[2] / dx;
  f[3] = zg[0] / dz;
  f[4] = zg[1] / dz;
  f[5] = zg[2] / dz;
}

// bldmol.f:175-242 (eyring) for atom a from its parent's frame.
__device__ __forceinline__ void eyring_atom(const DevSys& S, EnergyLds& L, int a) {
  const int p = S.parent[a];
  double ct = L.fr[6 * a + 0], st = L.fr[6 * a + 1];
  double ca = S.cba[a], sa = S.sba[a], bl = S.bl[a];
  double x1 = L.fr[6 * p + 0], x2 = L.fr[6 * p + 1], x3 = L.fr[6 * p + 2];
  double z1 = L.fr[6 * p + 3], z2 = L.fr[6 * p + 4], z3 = L.fr[6 * p + 5];
  double px = L.x[p], py = L.y[p], pz = L.z[p];
  double y1 = z2 * x3 - z3 * x2, y2 = z3 * x1 - z1 * x3, y3 = z1 * x2 - z2 * x1;
  double h2 = -sa * ct, h3 = -sa * st;
  double u1 = -ca * x1 + h2 * y1 + h3 * z1;
  double u2 = -ca * x2 + h2 * y2 + h3 * z2;
  double u3 = -ca * x3 + h2 * y3 + h3 * z3;
  double dx = 1.0 / sqrt(u1 * u1 + u2 * u2 + u3 * u3);
  u1 *= dx; u2 *= dx; u3 *= dx;
  double w1 = -st * y1 + ct * z1, w2 = -st * y2 + ct * z2, w3 = -st * y3 + ct * z3;
  double dz = 1.0 / sqrt(w1 * w1 + w2 * w2 + w3 * w3);
  L.fr[6 * a + 0] = u1;
  L.fr[6 * a + 1] = u2;
  L.fr[6 * a + 2] = u3;
  L.fr[6 * a + 3] = w1 * dz;
  L.fr[6 * a + 4] = w2 * dz;
  L.fr[6 * a + 5] = w3 * dz;
  L.x[a] = px + u1 * bl;
  L.y[a] = py + u2 * bl;
  L.z[a] = pz + u3 * bl;
}

// Chain build by wave 0, level by level; the other waves wait.  gb holds
// six global parameters per chain; lane m places the first atom of chain m.
__device__ void build_chain(const DevSys& S, EnergyLds& L, Work& wk, const double* gb) {
  if (threadIdx.x < kWave) {
    if (threadIdx.x < S.nml) build_root(S, L, gb + 6 * threadIdx.x, threadIdx.x);
    if (threadIdx.x == 0) wk.builds += S.nat;
    wave_lds_sync();
    const int lane = threadIdx.x;
    for (int lv = 0; lv < S.nlev; ++lv) {
      const int o0 = S.lev_off[lv], o1 = S.lev_off[lv + 1];
      for (int o = o0 + lane; o < o1; o += kWave) eyring_atom(S, L, S.lev_atoms[o]);
      wave_lds_sync();
    }
  }
  __syncthreads();
}

// Upper-triangle tile index -> (bi, bj), bi <= bj, row-major.
__device__ __forceinline__ void tile_coords(int t, int nt, int& bi, int& bj) {
  int b = 0, rem = t;
  while (rem >= nt - b) {
    rem -= nt - b;
    ++b;
  }
  bi = b;
  bj = b + rem;
}

struct PairSums {
  double el, vw, hb;
  double mag;  // mixed mode: sum of |pair term| (scale of the FP32 rounding), else 0
};

// Columns of word bj (atoms 64 bj .. 64 bj + 63) in the chain of atom i.
__device__ __forceinline__ uint64_t chain_cols(const DevSys& S, int i, int bj) {
  const int m = S.mol_of[i];
  const int lo = max(S.mol_a1[m] - 64 * bj, 0), hi = min(S.mol_a2[m] - 64 * bj, 63);
  if (lo > hi) return 0ull;
  const uint64_t upto = hi == 63 ? ~0ull : ((1ull << (hi + 1)) - 1ull);
  return upto & ~((1ull << lo) - 1ull);
}

// Pair terms over the tiles this wave owns (FP64): tiles t0, t0 + stride, ...
// sel: 0 every pair, 1 pairs within a chain, 2 pairs between chains.
__device__ PairSums pair_terms_f64(const DevSys& S, const EnergyLds& L, int t0, int stride, Work& wk,
                                   int sel = 0) {
  const int lane = threadIdx.x & (kWave - 1);
  const int nat = S.nat, nt = (nat + 63) / 64, ntile = nt * (nt + 1) / 2, words = S.words;
  double el = 0.0, vw = 0.0, hb = 0.0;
  for (int t = t0; t < ntile; t += stride) {
    int bi, bj;
    tile_coords(t, nt, bi, bj);
    const int i = bi * 64 + lane;
    const bool vi = i < nat;
    uint64_t w0 = 0, w1 = 0;
    double xi = 0, yi = 0, zi = 0, qi = 0;
    int ti = 0;
    if (vi) {
      w0 = pair_bits(L, ((size_t)i * words + bj) * 2 + 0);
      w1 = pair_bits(L, ((size_t)i * words + bj) * 2 + 1);
      xi = L.x[i]; yi = L.y[i]; zi = L.z[i];
      qi = S.conv * L.q[i];
      ti = L.t[i] * kTy;
    }
    if (bi == bj) {
      uint64_t keep = (lane == 63) ? 0ull : ~((2ull << lane) - 1ull);
      w0 &= keep;
      w1 &= keep;
    }
    if (sel != 0 && vi) {
      const uint64_t in = chain_cols(S, i, bj);
      const uint64_t keep = sel == 1 ? in : ~in;
      w0 &= keep;
      w1 &= keep;
    }
    const int kmax = min(64, nat - bj * 64);
    wk.pairs += __popcll(w0 & (kmax == 64 ? ~0ull : ((1ull << kmax) - 1ull)));
    for (int k = 0; k < kmax; ++k) {
      const bool on = (w0 >> k) & 1ull;
      if (!__any(on)) continue;
      if (on) {
        const int j = bj * 64 + k;
        double dx = L.x[j] - xi, dy = L.y[j] - yi, dz = L.z[j] - zi;
        double r2 = dx * dx + dy * dy + dz * dz;
        double ri = rsqrt_f64(r2);
        double r2i = ri * ri;
        double r6i = r2i * r2i * r2i;
        const int tt = ti + L.t[j];
        const bool is14 = (w1 >> k) & 1ull;
        double A = L.ptab[4 * tt + (is14 ? 1 : 0)];
        double Cc = L.ptab[4 * tt + 2];
        double ep_inv = 1.0;
        if (S.epsd) {
          double rij = r2 * ri;
          double sr = S.slp * rij;
          ep_inv = 1.0 / (S.plt - (sr * sr + 2.0 * sr + 2.0) * (S.plt - 1.0) * exp(-sr) / 2.0);
        }
        el += qi * L.q[j] * ri * ep_inv;
        if (L.ptab[4 * tt + 3] != 0.0)
          hb += A * (r6i * r6i) - Cc * (r6i * r2i * r2i);
        else
          vw += A * (r6i * r6i) - Cc * r6i;
      }
    }
  }
  return {el, vw, hb, 0.0};
}

// Mixed precision: FP32 terms, FP32 sums per tile, FP64 accumulation.
__device__ PairSums pair_terms_mixed(const DevSys& S, const EnergyLds& L, int t0, int stride, Work& wk,
                                     int sel = 0) {
  const int lane = threadIdx.x & (kWave - 1);
  const int nat = S.nat, nt = (nat + 63) / 64, ntile = nt * (nt + 1) / 2, words = S.words;
  double el = 0.0, vw = 0.0, hb = 0.0, mag = 0.0;
  const float conv = (float)S.conv;
  for (int t = t0; t < ntile; t += stride) {
    int bi, bj;
    tile_coords(t, nt, bi, bj);
    const int i = bi * 64 + lane;
    const bool vi = i < nat;
    uint64_t w0 = 0, w1 = 0;
    double xi = 0, yi = 0, zi = 0;
    float qi = 0;
    int ti = 0;
    if (vi) {
      w0 = pair_bits(L, ((size_t)i * words + bj) * 2 + 0);
      w1 = pair_bits(L, ((size_t)i * words + bj) * 2 + 1);
      xi = L.x[i]; yi = L.y[i]; zi = L.z[i];
      qi = conv * L.qf[i];
      ti = L.t[i] * kTy;
    }
    if (bi == bj) {
      uint64_t keep = (lane == 63) ? 0ull : ~((2ull << lane) - 1ull);
      w0 &= keep;
      w1 &= keep;
    }
    if (sel != 0 && vi) {
      const uint64_t in = chain_cols(S, i, bj);
      const uint64_t keep = sel == 1 ? in : ~in;
      w0 &= keep;
      w1 &= keep;
    }
    float fel = 0.f, fvw = 0.f, fhb = 0.f, fmag = 0.f;
    const int kmax = min(64, nat - bj * 64);
    wk.pairs += __popcll(w0 & (kmax == 64 ? ~0ull : ((1ull << kmax) - 1ull)));
    for (int k = 0; k < kmax; ++k) {
      const bool on = (w0 >> k) & 1ull;
      if (!__any(on)) continue;
      if (on) {
        const int j = bj * 64 + k;
        // differences in FP64 (coordinates reach tens of Angstrom), the rest in FP32
        float dx = (float)(L.x[j] - xi), dy = (float)(L.y[j] - yi), dz = (float)(L.z[j] - zi);
        float r2 = dx * dx + dy * dy + dz * dz;
        float ri = __builtin_amdgcn_rsqf(r2);
        float r2i = ri * ri;
        float r6i = r2i * r2i * r2i;
        const int tt = ti + L.t[j];
        const bool is14 = (w1 >> k) & 1ull;
        float A = (float)L.ptab[4 * tt + (is14 ? 1 : 0)];
        float Cc = (float)L.ptab[4 * tt + 2];
        const bool hbt = L.ptab[4 * tt + 3] != 0.0;
        float ep_inv = 1.f;
        if (S.epsd) {  // sigmoidal dielectric (enyshe.f:129-134)
          const float sr = (float)S.slp * (r2 * ri), plt = (float)S.plt;
          ep_inv = 1.f / (plt - (sr * sr + 2.f * sr + 2.f) * (plt - 1.f) * expf(-sr) / 2.f);
        }
        const float tel = qi * L.qf[j] * ri * ep_inv;
        fel += tel;
        const float tlj = A * (r6i * r6i) - Cc * (hbt ? r6i * r2i * r2i : r6i);
        if (hbt)
          fhb += tlj;
        else
          fvw += tlj;
        fmag += fabsf(tel) + fabsf(A * (r6i * r6i)) + fabsf(Cc * r6i);
      }
    }
    el += (double)fel;
    vw += (double)fvw;
    hb += (double)fhb;
    mag += (double)fmag;
  }
  return {el, vw, hb, mag};
}

// enyshe.f:73-79 over all variables; lanes over variables.
__device__ double torsion_terms(const DevSys& S, const double* vl) {
  double e = 0.0;
  for (int v = threadIdx.x; v < S.nvt; v += blockDim.x) {
    double e0 = S.var_e0[v];
    if (e0 != 0.0) {
      double to = atom_torsion(S, vl, S.var_atom[v]);
      e += e0 * (1.0 + S.var_sg[v] * cos(to * S.var_rn[v]));
    }
  }
  return e;
}

constexpr int kMaxPntPerLane = 2;  // tessellations of up to 128 points (tes.dat: 122)

// Solvent point test.  The reference occludes tessellation point p of heavy
// atom i by neighbour j when its FP64
//   sdd = (dx+px)^2 + (dy+py)^2 + (dz+pz)^2   (left to right, no contraction)
// is <= r2 = (rj/ri)^2, with d = (x_i - x_j)/r_i (enysol.f).  In exact
// arithmetic |d+p|^2 = |d|^2 + v, v = 2 d.p + |p|^2, so with t = r2 - |d|^2
// the test is v <= t.  Three levels decide it:
//  1. FP32 (the point loop; 20 bytes of LDS per neighbour, which is what
//     bounds this loop):  s = (v - t)/M32, M32 = 24 u32 (r2 + 21|d|^2 + 20)
//     with u32 = 2^-24, at least twice the FP32 rounding of d, p, |p|^2, v,
//     t and s for |p| < 2: s <= -1 means the reference occludes, and a point
//     whose s exceeds 1 for every neighbour is free.
//  2. FP64 (undecided points only): the same form with M64 = 128 u (r2 +
//     3|d|^2 + 9), u = 2^-53, d and r2 from a reciprocal of r_i.
//  3. The reference's own expression with its divisions, for what is left.
// So every free-point count equals enysol.f's.
constexpr double kU32 = 3.552713678800501e-15;    // 32 * 2^-53
constexpr double kU128 = 1.4210854715202004e-14;  // 128 * 2^-53
constexpr double kF24 = 1.430511474609375e-06;    // 24 * 2^-24

// Minimum of s over the neighbours of one chunk, per owned point (lane l
// owns points l and l+64).  Stops once every valid point is occluded.
__device__ __forceinline__ void solvent_points(const float* nb, int cnt, const float (&px)[kMaxPntPerLane],
                                               const float (&py)[kMaxPntPerLane], const float (&pz)[kMaxPntPerLane],
                                               const float (&pe)[kMaxPntPerLane], const bool (&done)[kMaxPntPerLane],
                                               float (&smin)[kMaxPntPerLane]) {
  for (int k0 = 0; k0 < cnt; k0 += 8) {
#pragma unroll
    for (int u = 0; u < 8; u += 4) {
      if (k0 + u >= cnt) break;
#pragma unroll
      for (int v4 = 0; v4 < 4; ++v4) {
        const float* e = nb + kNbFloats * (k0 + u + v4);
        const float4 a = *reinterpret_cast<const float4*>(e);
        const float c = e[4];
#pragma unroll
        for (int q = 0; q < kMaxPntPerLane; ++q) {
          const float v = fmaf(a.x, px[q], fmaf(a.y, py[q], fmaf(a.z, pz[q], pe[q])));
          smin[q] = fminf(smin[q], fmaf(v, a.w, c));
        }
      }
    }
    bool all = true;
#pragma unroll
    for (int q = 0; q < kMaxPntPerLane; ++q) all = all && (done[q] || smin[q] <= -1.f);
    if (__all(all)) break;
  }
}

// Levels 2 and 3 for one undecided point p of this lane over the
// neighbours of the chunk (rarely run).
__device__ __forceinline__ bool resolve_point(const EnergyLds& L, const float* nb, int cnt, double xi, double yi,
                                              double zi, double trad, double rinv, int p) {
#pragma clang fp contract(off)
  const double px = L.sp[3 * p + 0], py = L.sp[3 * p + 1], pz = L.sp[3 * p + 2];
  const double pe = px * px + py * py + pz * pz;
  for (int k = 0; k < cnt; ++k) {
    const int j = __float_as_int(nb[kNbFloats * k + 5]);
    const double ax = xi - L.x[j], ay = yi - L.y[j], az = zi - L.z[j];
    const double dxa = ax * rinv, dya = ay * rinv, dza = az * rinv;
    const double dda = dxa * dxa + dya * dya + dza * dza;
    const double aka = L.rv[j] * rinv;
    const double r2a = aka * aka;
    const double w = fma(2.0 * dxa, px, fma(2.0 * dya, py, fma(2.0 * dza, pz, pe))) - (r2a - dda);
    const double m = kU128 * (r2a + 3.0 * dda + 9.0);
    if (w <= -m) return true;
    if (w <= m) {  // the reference's expression
      const double dx = ax / trad, dy = ay / trad, dz = az / trad;
      const double akrad = L.rv[j] / trad;
      const double r2b = akrad * akrad;
      const double sx = dx + px, sy = dy + py, sz = dz + pz;
      double sdd = sx * sx;
      sdd = sdd + sy * sy;
      sdd = sdd + sz * sz;
      if (sdd <= r2b) return true;
    }
  }
  return false;
}

// The reference's contact test (enysol.f) for candidate j of atom i:
// dd = |d|^2 <= dr = (1 + rj/ri)^2 with exact divisions, decided from the
// reciprocal form when the two sides are further apart than its rounding.
__device__ __forceinline__ bool contact_exact(double ax, double ay, double az, double rj, double trad) {
#pragma clang fp contract(off)
  const double dx = ax / trad, dy = ay / trad, dz = az / trad;
  const double dd = dx * dx + dy * dy + dz * dz;
  double akrad = rj / trad;
  double dr = 1.0 + akrad;
  dr = dr * dr;
  return dd <= dr;
}

// The packed heavy-atom copy of the current coordinates (all threads; a
// barrier must follow before solvent_atom reads it).
__device__ __forceinline__ void fill_heavy_packed(const DevSys& S, const EnergyLds& L) {
  for (int h = threadIdx.x; h < S.nheavy; h += blockDim.x) {
    const int j = L.hv[h];
    L.hp[4 * h + 0] = L.x[j];
    L.hp[4 * h + 1] = L.y[j];
    L.hp[4 * h + 2] = L.z[j];
    L.hp[4 * h + 3] = L.rv[j];
  }
}

// Point tests of the contacts gathered so far (padded to a multiple of 4),
// then the fallback levels for the undecided points.
__device__ __forceinline__ void solvent_chunk(const EnergyLds& L, float* nb, int cnt, double xi, double yi,
                                              double zi, double trad, double rinv, const float (&px)[kMaxPntPerLane],
                                              const float (&py)[kMaxPntPerLane], const float (&pz)[kMaxPntPerLane],
                                              const float (&pe)[kMaxPntPerLane], bool (&occ)[kMaxPntPerLane]) {
  const int lane = threadIdx.x & (kWave - 1);
  const int padded = (cnt + 3) & ~3;
  if (lane < padded - cnt) {  // s = 4: never occludes, never undecided
    float* e = nb + kNbFloats * (cnt + lane);
    e[0] = e[1] = e[2] = e[3] = 0.f;
    e[4] = 4.f;
    e[5] = __int_as_float(-1);
  }
  wave_lds_sync();
  float smin[kMaxPntPerLane];
#pragma unroll
  for (int q = 0; q < kMaxPntPerLane; ++q) smin[q] = 4.f;
  solvent_points(nb, padded, px, py, pz, pe, occ, smin);
  bool undecided = false;
#pragma unroll
  for (int q = 0; q < kMaxPntPerLane; ++q) {
    if (!occ[q] && smin[q] <= -1.f) occ[q] = true;
    undecided = undecided || (!occ[q] && smin[q] <= 1.f);
  }
  if (__any(undecided)) {
#pragma unroll
    for (int q = 0; q < kMaxPntPerLane; ++q)
      if (!occ[q] && smin[q] <= 1.f) occ[q] = resolve_point(L, nb, cnt, xi, yi, zi, trad, rinv, q * kWave + lane);
  }
  wave_lds_sync();
}

// Per-lane tessellation points (FP32) of the points this lane owns.
struct LanePoints {
  float x[kMaxPntPerLane], y[kMaxPntPerLane], z[kMaxPntPerLane], e[kMaxPntPerLane];
  bool occ[kMaxPntPerLane];
};
__device__ __forceinline__ void lane_points(const EnergyLds& L, int npnt, LanePoints& lp) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int q = 0; q < kMaxPntPerLane; ++q) {
    const int p = q * kWave + lane;
    const bool in = p < npnt;
    lp.x[q] = in ? (float)L.sp[3 * p + 0] : 0.f;
    lp.y[q] = in ? (float)L.sp[3 * p + 1] : 0.f;
    lp.z[q] = in ? (float)L.sp[3 * p + 2] : 0.f;
    lp.e[q] = lp.x[q] * lp.x[q] + lp.y[q] * lp.y[q] + lp.z[q] * lp.z[q];
    lp.occ[q] = !in;
  }
}

// Contact test of heavy atom j against atom i: dd <= (1 + rj/ri)^2 (see
// contact_exact); on a contact, its neighbour entry (see solvent_points).
struct ContactGeom {
  double dxa, dya, dza, dda, r2a;
};
__device__ __forceinline__ bool contact_of(int i, int j, double xi, double yi, double zi, double trad, double rinv,
                                           double xj, double yj, double zj, double rj, ContactGeom& cg) {
#pragma clang fp contract(off)
  if (j < 0 || j == i) return false;
  const double ax = xi - xj, ay = yi - yj, az = zi - zj;
  cg.dxa = ax * rinv;
  cg.dya = ay * rinv;
  cg.dza = az * rinv;
  cg.dda = cg.dxa * cg.dxa + cg.dya * cg.dya + cg.dza * cg.dza;
  const double aka = rj * rinv;
  double dra = 1.0 + aka;
  dra = dra * dra;
  cg.r2a = aka * aka;
  const double mc = kU32 * (cg.dda + dra);
  if (cg.dda <= dra - mc) return true;
  if (cg.dda > dra + mc) return false;
  return contact_exact(ax, ay, az, rj, trad);
}
__device__ __forceinline__ void write_contact(float* e, int j, const ContactGeom& cg) {
  const double t = cg.r2a - cg.dda;
  const float im = (float)__builtin_amdgcn_rcp(kF24 * (cg.r2a + 21.0 * cg.dda + 20.0));  // 1/M32, any rounding
  e[0] = (float)(2.0 * cg.dxa);
  e[1] = (float)(2.0 * cg.dya);
  e[2] = (float)(2.0 * cg.dza);
  e[3] = im;
  e[4] = (float)(-t * (double)im);
  e[5] = __int_as_float(j);
}

// Atoms with more than kMaxNb contacts (not met in proteins of the usual
// packing): the contacts are taken in chunks of kMaxNb, all from LDS.  Out
// of line, with its own LDS carve-up, to keep the common path small.
__device__ __attribute__((noinline)) int solvent_atom_chunked(const DevSys* sys, int nwaves, int bits_lds, int i) {
#pragma clang fp contract(off)
  extern __shared__ double smem[];
  const DevSys& S = *sys;
  const EnergyLds L = carve_energy_lds(smem, S, nwaves, bits_lds != 0);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  float* nb = reinterpret_cast<float*>(L.fr + (size_t)wave * kNbDoubles);
  const int nh = S.nheavy;
  const double xi = L.x[i], yi = L.y[i], zi = L.z[i], trad = L.rv[i], rinv = 1.0 / trad;
  LanePoints lp;
  lane_points(L, S.npnt, lp);
  int cnt = 0;
  for (int h0 = 0; h0 < nh; h0 += kWave) {
    const int h = h0 + lane;
    const int j = h < nh ? L.hv[h] : -1;
    const int jj = j < 0 ? 0 : j;
    ContactGeom cg;
    const bool on = contact_of(i, j, xi, yi, zi, trad, rinv, L.x[jj], L.y[jj], L.z[jj], L.rv[jj], cg);
    const uint64_t m = __ballot(on);
    const int n = __popcll(m);
    if (cnt + n > kMaxNb) {
      solvent_chunk(L, nb, cnt, xi, yi, zi, trad, rinv, lp.x, lp.y, lp.z, lp.e, lp.occ);
      cnt = 0;
    }
    if (on) write_contact(nb + kNbFloats * (cnt + __popcll(m & ((1ull << lane) - 1ull))), j, cg);
    cnt += n;
  }
  solvent_chunk(L, nb, cnt, xi, yi, zi, trad, rinv, lp.x, lp.y, lp.z, lp.e, lp.occ);
  int nfree = 0;
#pragma unroll
  for (int q = 0; q < kMaxPntPerLane; ++q) nfree += __popcll(__ballot(!lp.occ[q]));
  return nfree;
}

// enysol.f:192-324: area of heavy atom i from its free tessellation points,
// by one wave.  The reference's contact test runs on every heavy atom (read
// from the packed copy L.hp, which must be current); the contacts go to the
// wave's LDS scratch (nb, kNbDoubles), and a point is
// free when no contact holds it (the reference's 'last hit' cache only
// reorders this OR).  More than kMaxNb contacts: solvent_atom_chunked.
// Returns the number of free points, or -1 for a tessellation with more
// than 128 points.
__device__ int solvent_atom(const DevSys& S, const EnergyLds& L, double* nbd, int i, Work& wk, int bits_lds,
                            double* tsplit = nullptr) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & (kWave - 1);
  // optional timing (lane 0): tsplit[0] += gather, [1] += point tests, [2] += 1
  double tmark = tsplit ? (double)wall_clock64() : 0.0;
  const int npnt = S.npnt, nh = S.nheavy;
  if (npnt > kWave * kMaxPntPerLane) return -1;  // the host rejects such tessellations first
  float* nb = reinterpret_cast<float*>(nbd);
  const double xi = L.x[i], yi = L.y[i], zi = L.z[i];
  const double trad = L.rv[i];
  const double rinv = 1.0 / trad;
  int cnt = 0;
  bool overflow = false;
  auto visit = [&](int j, double xj, double yj, double zj, double rj) {
    ContactGeom cg;
    const bool on = contact_of(i, j, xi, yi, zi, trad, rinv, xj, yj, zj, rj, cg);
    const uint64_t m = __ballot(on);
    const int n = __popcll(m);
    if (cnt + n > kMaxNb) overflow = true;
    if (on && !overflow) write_contact(nb + kNbFloats * (cnt + __popcll(m & ((1ull << lane) - 1ull))), j, cg);
    cnt += n;
  };
  for (int h0 = 0; h0 < nh; h0 += kWave) {
    const int h = h0 + lane;
    const bool in = h < nh;
    const double2 a = in ? *reinterpret_cast<const double2*>(L.hp + 4 * h) : double2{0.0, 0.0};
    const double2 b = in ? *reinterpret_cast<const double2*>(L.hp + 4 * h + 2) : double2{0.0, 0.0};
    visit(in ? (int)L.hv[h] : -1, a.x, a.y, b.x, b.y);
  }
  if (lane == 0) wk.contacts += cnt;
  if (overflow) {
    wave_lds_sync();
    return solvent_atom_chunked(&S, blockDim.x / kWave, bits_lds, i);
  }
  if (tsplit && lane == 0) {
    const double now = (double)wall_clock64();
    tsplit[0] += now - tmark;
    tmark = now;
  }
  LanePoints lp;
  lane_points(L, npnt, lp);
  solvent_chunk(L, nb, cnt, xi, yi, zi, trad, rinv, lp.x, lp.y, lp.z, lp.e, lp.occ);
  if (tsplit && lane == 0) {
    tsplit[1] += (double)wall_clock64() - tmark;
    tsplit[2] += 1.0;
  }
  int nfree = 0;
#pragma unroll
  for (int q = 0; q < kMaxPntPerLane; ++q) nfree += __popcll(__ballot(!lp.occ[q]));
  return nfree;
}

struct EnergyOut {
  double el, vw, hb, tor, sol, asa;
  double mag;  // mixed mode: sum of |pair term|; 0 in FP64 mode
  int err;
  double eli, vwi, hbi;  // with split: the pair terms between chains (not in el, vw, hb)
};

// Pair, torsion and solvent terms of the coordinates in L (all threads).
// nfree_out (LDS, nheavy) receives the free-point count of every heavy atom
// this block evaluates when not null.  On return thread 0 holds the totals
// in 'out'.  With nparts > 1 the block evaluates share 'part' only (tiles
// and heavy atoms k with k % nparts == part; torsions in part 0), and the
// totals are partial sums.  With split (several chains) the pair terms
// between chains go to eli, vwi, hbi instead of el, vw, hb (eninteract.f).
template <bool MIXED>
__device__ void energy_from_coords(const DevSys& S, const double* vl, EnergyLds& L, int32_t* nfree_out,
                                   EnergyOut& out, Work& wk, int part = 0, int nparts = 1, bool split = false) {
  const int tid = threadIdx.x, nwaves = blockDim.x / kWave, wave = tid / kWave, lane = tid & (kWave - 1);
  if (S.itysol > 0) fill_heavy_packed(S, L);  // read after the barrier below
  if (MIXED) {  // FP32 copies centred on the first atom (frames are dead now)
    const double cx = L.x[0], cy = L.y[0], cz = L.z[0];
    for (int a = tid; a < S.np; a += blockDim.x) {
      L.xf[a] = (float)(L.x[a] - cx);
      L.yf[a] = (float)(L.y[a] - cy);
      L.zf[a] = (float)(L.z[a] - cz);
      L.qf[a] = (float)L.q[a];
    }
    __syncthreads();
  }
  const int t0 = part + nparts * wave, stride = nparts * nwaves;
  split = split && S.nml > 1;
  PairSums ps = MIXED ? pair_terms_mixed(S, L, t0, stride, wk, split ? 1 : 0)
                      : pair_terms_f64(S, L, t0, stride, wk, split ? 1 : 0);
  PairSums pi = {0.0, 0.0, 0.0, 0.0};
  if (split) {
    pi = MIXED ? pair_terms_mixed(S, L, t0, stride, wk, 2) : pair_terms_f64(S, L, t0, stride, wk, 2);
    ps.mag += pi.mag;
  }
  double tor = part == 0 ? torsion_terms(S, vl) : 0.0;
  if (tid == 0) *L.ctr = 0;
  __syncthreads();  // packed heavy atoms written; the solvent scratch overlays the FP32 copies
  double sol = 0.0, asa = 0.0;
  int err = 0;
  if (S.itysol > 0) {
    // This block's heavy atoms are h = part + nparts k.  Waves take the next
    // atom as they finish (an atom's cost varies with its contacts and with
    // how soon its points are all occluded); the areas are then summed with
    // a fixed thread-to-atom map, so the totals do not depend on which wave
    // took which atom.
    double* nb = L.fr + (size_t)wave * kNbDoubles;
    int32_t* nf = nfree_out ? nfree_out : L.nfs;
    const double c4 = 4.0 * kPi / (double)S.npnt;
    const int nmine = S.nheavy > part ? (S.nheavy - part + nparts - 1) / nparts : 0;
    for (;;) {
      int k = 0;
      if (lane == 0) k = atomicAdd(L.ctr, 1);
      k = __shfl(k, 0, kWave);
      if (k >= nmine) break;
      const int h = part + nparts * k;
      const int i = L.hv[h];
      int nfree = solvent_atom(S, L, nb, i, wk, L.lbits != nullptr);
      if (nfree < 0) {
        err = 1;
        nfree = 0;
      }
      if (lane == 0) nf[h] = nfree;
    }
    __syncthreads();
    for (int k = tid; k < nmine; k += blockDim.x) {
      const int h = part + nparts * k;
      const int i = L.hv[h];
      const double trad = L.rv[i];
      const double area = c4 * trad * trad * (double)nf[h];
      sol += area * L.sg[i];
      asa += area;
    }
  }
  double v[7] = {ps.el, ps.vw, ps.hb, tor, sol, asa, ps.mag};
#pragma unroll
  for (int k = 0; k < 7; ++k) v[k] = wave_sum(v[k]);
  const int errw = __any(err);
  __syncthreads();  // L.red may still be read by a previous reduction
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 7; ++k) L.red[8 * wave + k] = v[k];
    L.red[8 * wave + 7] = (double)errw;
  }
  __syncthreads();
  if (tid == 0) {
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int w = 0; w < nwaves; ++w)
      for (int k = 0; k < 8; ++k) s[k] += L.red[8 * w + k];
    out.el = s[0];
    out.vw = s[1];
    out.hb = s[2];
    out.tor = s[3];
    out.sol = s[4];
    out.asa = s[5];
    out.mag = s[6];
    out.err = s[7] != 0.0;
    out.eli = out.vwi = out.hbi = 0.0;
  }
  if (split) {  // second round through L.red: the terms between chains
    double u[3] = {pi.el, pi.vw, pi.hb};
#pragma unroll
    for (int k = 0; k < 3; ++k) u[k] = wave_sum(u[k]);
    __syncthreads();
    if (lane == 0)
      for (int k = 0; k < 3; ++k) L.red[8 * wave + k] = u[k];
    __syncthreads();
    if (tid == 0) {
      double t[3] = {0.0, 0.0, 0.0};
      for (int w = 0; w < nwaves; ++w)
        for (int k = 0; k < 3; ++k) t[k] += L.red[8 * w + k];
      out.eli = t[0];
      out.vwi = t[1];
      out.hbi = t[2];
    }
  }
}

// Whole energy of one conformation from its variables (energy.f:12-119).
template <bool MIXED>
__device__ void conformation_energy(const DevSys& S, const double* vl, const double* gb, EnergyLds& L,
                                    int32_t* nfree_out, EnergyOut& out, Work& wk, int part = 0, int nparts = 1,
                                    bool split = false) {
  energy_stage_torsions(S, vl, L);
  __syncthreads();
  build_chain(S, L, wk, gb);
  energy_from_coords<MIXED>(S, vl, L, nfree_out, out, wk, part, nparts, split);
}

}  // namespace smmp
