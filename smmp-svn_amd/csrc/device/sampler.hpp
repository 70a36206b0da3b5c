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
  const int32_t* idvr;   // free variables, 0-based
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
};

// Phases of a sweep timed by thread 0 when P.phase is set.
enum PhaseId { kPhPropose = 0, kPhRotate, kPhPairs, kPhSolvent, kPhAccept, kPhFull, kPhSweepEnd, kPhNFull };

// Extra LDS of the sampler beyond the energy carve-up.
__host__ __device__ inline size_t sampler_lds_bytes(int np, int npnt, int nheavy, int nvt, int nwaves) {
  size_t b = energy_lds_bytes(np, npnt, nwaves);
  b = (b + 15) & ~size_t(15);
  b += 3 * (size_t)np * sizeof(double);           // backup of moved coordinates
  b += (size_t)nvt * sizeof(double);              // variables
  b += 2 * (size_t)nheavy * sizeof(int32_t);      // nfree current / trial
  b += 625 * sizeof(uint32_t);                    // MT state
  b += 32 * sizeof(double);                       // broadcast scalars
  b += (size_t)nheavy;                            // affected flags
  return b + 64;
}

struct SamplerLds {
  EnergyLds L;
  double *bx, *by, *bz, *vl, *sc;
  int32_t *nf, *nf2;
  uint32_t* mt;
  int8_t* aff;
};

__device__ inline SamplerLds carve_sampler_lds(double* base, const DevSys& S, int nvt, int nwaves) {
  SamplerLds Q;
  Q.L = carve_energy_lds(base, S.np, S.npnt, nwaves);
  char* p = reinterpret_cast<char*>(base) + ((energy_lds_bytes(S.np, S.npnt, nwaves) + 15) & ~size_t(15));
  Q.bx = reinterpret_cast<double*>(p); p += S.np * sizeof(double);
  Q.by = reinterpret_cast<double*>(p); p += S.np * sizeof(double);
  Q.bz = reinterpret_cast<double*>(p); p += S.np * sizeof(double);
  Q.vl = reinterpret_cast<double*>(p); p += nvt * sizeof(double);
  Q.sc = reinterpret_cast<double*>(p); p += 32 * sizeof(double);
  Q.nf = reinterpret_cast<int32_t*>(p); p += S.nheavy * sizeof(int32_t);
  Q.nf2 = reinterpret_cast<int32_t*>(p); p += S.nheavy * sizeof(int32_t);
  Q.mt = reinterpret_cast<uint32_t*>(p); p += 625 * sizeof(uint32_t);
  Q.aff = reinterpret_cast<int8_t*>(p);
  return Q;
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

// ECEPP pair energy (Coulomb, LJ or H-bond; 1-4 repulsion from A14).
struct PairE {
  double el, vw, hb;
};
__device__ __forceinline__ void pair_energy(const DevSys& S, const EnergyLds& L, double dx, double dy, double dz,
                                            double qq, int tt, bool is14, PairE& e, double sgn) {
  double r2 = dx * dx + dy * dy + dz * dz;
  double ri = rsqrt_f64(r2);
  double r2i = ri * ri;
  double r6i = r2i * r2i * r2i;
  double A = is14 ? L.pA14[tt] : L.pA[tt];
  double Cc = L.pC[tt];
  double ep_inv = 1.0;
  if (S.epsd) {
    double rij = r2 * ri;
    double sr = S.slp * rij;
    ep_inv = 1.0 / (S.plt - (sr * sr + 2.0 * sr + 2.0) * (S.plt - 1.0) * exp(-sr) / 2.0);
  }
  e.el += sgn * (qq * ri * ep_inv);
  if (L.pHB[tt] != 0.0)
    e.hb += sgn * (A * (r6i * r6i) - Cc * (r6i * r2i * r2i));
  else
    e.vw += sgn * (A * (r6i * r6i) - Cc * r6i);
}

// Energy change of the pairs between the moved range [m1,m2] (new
// coordinates in L, old in the backup) and the fixed atoms.  The smaller
// set is walked uniformly, the larger one across the lanes.
__device__ void pair_delta(const DevSys& S, const SamplerLds& Q, int m1, int m2, PairE& d, Work& wk) {
  const EnergyLds& L = Q.L;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave, nwaves = blockDim.x / kWave;
  const int nat = S.nat, nM = m2 - m1 + 1, nO = nat - nM, words = S.words;
  if (nM <= 0 || nO <= 0) return;
  const bool rowsM = nM <= nO;            // uniform rows from the smaller set
  const int nrow = rowsM ? nM : nO, ncol = rowsM ? nO : nM;
  const int nob = (ncol + 63) / 64;
  for (int item = wave; item < nrow * nob; item += nwaves) {
    const int r = item / nob, ob = item - r * nob;
    const int c = ob * 64 + lane;
    if (c >= ncol) continue;
    int a, b;  // a in M, b fixed
    if (rowsM) {
      a = m1 + r;
      b = c < m1 ? c : c + nM;
    } else {
      b = r < m1 ? r : r + nM;
      a = m1 + c;
    }
    const uint64_t w0 = S.bits[((size_t)a * words + (b >> 6)) * 2 + 0];
    if (!((w0 >> (b & 63)) & 1ull)) continue;
    wk.pairs += 2.0;
    const bool is14 = (S.bits[((size_t)a * words + (b >> 6)) * 2 + 1] >> (b & 63)) & 1ull;
    const int tt = L.t[a] * kTy + L.t[b];
    const double qq = S.conv * L.q[a] * L.q[b];
    const double xb = L.x[b], yb = L.y[b], zb = L.z[b];
    pair_energy(S, L, L.x[a] - xb, L.y[a] - yb, L.z[a] - zb, qq, tt, is14, d, 1.0);
    const int k = a - m1;
    pair_energy(S, L, Q.bx[k] - xb, Q.by[k] - yb, Q.bz[k] - zb, qq, tt, is14, d, -1.0);
  }
}

// Marks the heavy atoms whose accessible area may change: every moved atom
// (its 122 points stay fixed in the laboratory frame while it rotates, so
// even a rigidly moved environment changes its discretised area) and every
// fixed atom in contact with a moved one before or after the move.
__device__ void mark_affected(const DevSys& S, const SamplerLds& Q, int m1, int m2) {
  const int nh = S.nheavy;
  int h1 = 0, h2 = nh;  // heavy atoms of [m1,m2]: one index range (heavy is sorted)
  {
    int lo = 0, hi = nh;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (S.heavy[mid] < m1) lo = mid + 1; else hi = mid;
    }
    h1 = lo;
    lo = h1, hi = nh;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (S.heavy[mid] <= m2) lo = mid + 1; else hi = mid;
    }
    h2 = lo;
  }
  const int nM = h2 - h1, nO = nh - nM;
  for (int h = h1 + threadIdx.x; h < h2; h += blockDim.x) Q.aff[h] = 1;
  for (int it = threadIdx.x; it < nM * nO; it += blockDim.x) {
    const int r = it / nO, c = it - r * nO;
    const int ha = h1 + r, hb = c < h1 ? c : c + nM;
    const int a = S.heavy[ha], b = S.heavy[hb];
    const double lim = (S.rvdw[a] + S.rvdw[b]) * (1.0 + 1e-9);
    const double l2 = lim * lim;
    const double xb = Q.L.x[b], yb = Q.L.y[b], zb = Q.L.z[b];
    double dx = Q.L.x[a] - xb, dy = Q.L.y[a] - yb, dz = Q.L.z[a] - zb;
    bool hit = dx * dx + dy * dy + dz * dz <= l2;
    const int k = a - m1;
    dx = Q.bx[k] - xb; dy = Q.by[k] - yb; dz = Q.bz[k] - zb;
    hit = hit || (dx * dx + dy * dy + dz * dz <= l2);
    if (hit) {
      Q.aff[ha] = 1;
      Q.aff[hb] = 1;
    }
  }
}

// One block = one replica; P.nsweep Metropolis sweeps.
template <bool MIXED>
__global__ void __launch_bounds__(512) k_sweep(SamplerParams P) {
  extern __shared__ double smem[];
  const DevSys& S = *P.sys;
  const int r = P.rep0 + blockIdx.x;
  const int tid = threadIdx.x, nwaves = blockDim.x / kWave, wave = tid / kWave, lane = tid & (kWave - 1);
  const int nvt = P.nvt, np = S.np, nat = S.nat, nh = S.nheavy;
  SamplerLds Q = carve_sampler_lds(smem, S, nvt, nwaves);
  Work wk;
  EnergyLds& L = Q.L;
  double* gvl = P.vlvr + (size_t)r * nvt;
  double* gx = P.xyz + (size_t)r * 3 * np;
  int32_t* gnf = P.nfree + (size_t)r * nh;
  uint32_t* gmt = P.mt + (size_t)P.rng_of[r] * 625;

  energy_stage_static(S, L);
  for (int k = tid; k < nvt; k += blockDim.x) Q.vl[k] = gvl[k];
  for (int k = tid; k < 625; k += blockDim.x) Q.mt[k] = gmt[k];
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
  // sc: [0] eol, [1] acz, [2] trial value, [3] dphi, [4] accept, [5] full-trial flag,
  //     [6] jv, [7] drift flag, [8..10] axis point, [11..13] axis unit, [14] beta, [15] old value
  if (!valid) {
    EnergyOut o;
    conformation_energy<MIXED>(S, Q.vl, L, Q.nf, o, wk);
    if (tid == 0) {
      Q.sc[0] = o.el + o.vw + o.hb + o.tor + o.sol;
      if (o.err) atomicOr(P.err, 2);
    }
  } else if (tid == 0) {
    Q.sc[0] = P.eol[r];
  }
  if (tid == 0) {
    Q.sc[1] = P.acc[r];
    Q.sc[7] = 0.0;
  }
  __syncthreads();
  const double c4 = 4.0 * kPi / (double)(S.npnt > 0 ? S.npnt : 1);
  double ph[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint64_t tp = wall_clock64();
  auto lap = [&](int k) {
    if (P.phase && tid == 0) {
      const uint64_t now = wall_clock64();
      ph[k] += (double)(now - tp);
      tp = now;
    }
  };

  for (int sw = 0; sw < P.nsweep; ++sw) {
    const double beta = P.beta[P.slot_of[r]];
    for (int nsw = 0; nsw < P.nvr; ++nsw) {
      // ---- proposal (leader): metropolis.f:42-66
      if (tid == 0) {
        uint32_t* mti = &Q.mt[624];
        int ivar = nsw;
        if (P.rndord) {  // 1+int(nvr*grnd()); u = 1 exactly would index past the end
          ivar = (int)(P.nvr * mt_grnd(Q.mt, mti));
          if (ivar >= P.nvr) ivar = P.nvr - 1;
        }
        const int jv = P.idvr[ivar];
        const double vold = Q.vl[jv];
        const double dv = P.axvr[jv] * (mt_grnd(Q.mt, mti) - 0.5);
        const double vnew = d_addang(vold, dv);
        Q.sc[2] = vnew;
        Q.sc[15] = vold;
        Q.sc[3] = d_difang(vold, vnew);
        Q.sc[6] = (double)jv;
        Q.sc[5] = fabs(Q.sc[0]) >= 1e5 ? 1.0 : 0.0;
        const int b = S.var_base[jv], a0 = S.var_axis0[jv];
        double ux = L.x[b] - L.x[a0], uy = L.y[b] - L.y[a0], uz = L.z[b] - L.z[a0];
        double un = rsqrt_f64(ux * ux + uy * uy + uz * uz);
        Q.sc[8] = L.x[b]; Q.sc[9] = L.y[b]; Q.sc[10] = L.z[b];
        Q.sc[11] = ux * un; Q.sc[12] = uy * un; Q.sc[13] = uz * un;
      }
      __syncthreads();
      lap(kPhPropose);
      const int jv = (int)Q.sc[6];
      const int m1 = S.var_m1[jv], m2 = S.var_m2[jv];
      const bool full = Q.sc[5] != 0.0;
      EnergyOut vref;  // verify mode: full energy before the move (thread 0)
      if (P.verify && !full) {
        energy_from_coords<MIXED>(S, Q.vl, L, nullptr, vref, wk);
        __syncthreads();
      }
      // ---- rigid rotation of the moved atoms about the bond (Rodrigues)
      {
        double sn, cs;
        sincos(S.var_sign[jv] * Q.sc[3], &sn, &cs);
        const double px = Q.sc[8], py = Q.sc[9], pz = Q.sc[10];
        const double kx = Q.sc[11], ky = Q.sc[12], kz = Q.sc[13];
        if (tid == 0) wk.moved += m2 - m1 + 1;
        for (int a = m1 + tid; a <= m2; a += blockDim.x) {
          const double vx = L.x[a] - px, vy = L.y[a] - py, vz = L.z[a] - pz;
          Q.bx[a - m1] = L.x[a];
          Q.by[a - m1] = L.y[a];
          Q.bz[a - m1] = L.z[a];
          const double kv = kx * vx + ky * vy + kz * vz;
          const double cx = ky * vz - kz * vy, cy = kz * vx - kx * vz, cz = kx * vy - ky * vx;
          const double t = kv * (1.0 - cs);
          L.x[a] = px + vx * cs + cx * sn + kx * t;
          L.y[a] = py + vy * cs + cy * sn + ky * t;
          L.z[a] = pz + vz * cs + cz * sn + kz * t;
        }
        for (int h = tid; h < nh; h += blockDim.x) Q.aff[h] = 0;
        if (tid == 0) Q.vl[jv] = Q.sc[2];
      }
      __syncthreads();
      lap(kPhRotate);
      double enew;
      if (!full) {
        // ---- incremental energy change
        PairE d = {0.0, 0.0, 0.0};
        pair_delta(S, Q, m1, m2, d, wk);
        if (S.itysol > 0) mark_affected(S, Q, m1, m2);
        double dt = 0.0;
        if (tid == 0) {
          const double e0 = S.var_e0[jv];
          if (e0 != 0.0) {
            const int ia = S.var_atom[jv];
            const double tn = atom_torsion(S, Q.vl, ia);
            Q.vl[jv] = Q.sc[15];
            const double to = atom_torsion(S, Q.vl, ia);
            Q.vl[jv] = Q.sc[2];
            dt = e0 * (1.0 + S.var_sg[jv] * cos(tn * S.var_rn[jv])) - e0 * (1.0 + S.var_sg[jv] * cos(to * S.var_rn[jv]));
          }
        }
        __syncthreads();
        lap(kPhPairs);
        double dsol = 0.0;
        int err = 0;
        if (S.itysol > 0) {
          double* nb = L.fr + (size_t)wave * kMaxNb * 4;
          for (int h = wave; h < nh; h += nwaves) {
            if (!Q.aff[h]) continue;
            const int i = S.heavy[h];
            int nfree = solvent_atom(S, L, nb, i, wk);
            if (nfree < 0) {
              err = 1;
              nfree = Q.nf[h];
            }
            if (lane == 0) {
              Q.nf2[h] = nfree;
              const double trad = S.rvdw[i];
              dsol += S.sigma[i] * (c4 * trad * trad * (double)(nfree - Q.nf[h]));
            }
          }
        }
        double v[4] = {d.el, d.vw, d.hb, dsol};
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = wave_sum(v[k]);
        const int errw = __any(err);
        if (lane == 0) {
          for (int k = 0; k < 4; ++k) L.red[8 * wave + k] = v[k];
          L.red[8 * wave + 6] = (double)errw;
        }
        __syncthreads();
        if (tid == 0) {
          double s = dt, comp[4] = {0, 0, 0, 0};
          int e = 0;
          for (int w = 0; w < nwaves; ++w) {
            for (int k = 0; k < 4; ++k) {
              s += L.red[8 * w + k];
              comp[k] += L.red[8 * w + k];
            }
            e |= L.red[8 * w + 6] != 0.0;
          }
          if (e) atomicOr(P.err, 2);
          Q.sc[16] = Q.sc[0] + s;
          for (int k = 0; k < 4; ++k) Q.sc[24 + k] = comp[k];  // el, vw, hb, sol deltas (verify mode)
          Q.sc[28] = dt;
        }
        if (P.phase) __syncthreads();  // uniform: timing only
        lap(kPhSolvent);
      } else {
        // ---- full recomputation from the trial coordinates
        EnergyOut o;
        energy_from_coords<MIXED>(S, Q.vl, L, Q.nf2, o, wk);
        if (tid == 0) {
          Q.sc[16] = o.el + o.vw + o.hb + o.tor + o.sol;
          if (o.err) atomicOr(P.err, 2);
          ph[kPhNFull] += 1.0;
        }
        if (P.phase) __syncthreads();
        lap(kPhFull);
      }
      if (P.verify && !full) {  // compare each incremental component with a full recomputation
        EnergyOut vnew;
        energy_from_coords<MIXED>(S, Q.vl, L, nullptr, vnew, wk);
        if (tid == 0 && P.dbg[16 * r] == 0.0) {
          const double fd[5] = {vnew.el - vref.el, vnew.vw - vref.vw, vnew.hb - vref.hb, vnew.tor - vref.tor,
                                vnew.sol - vref.sol};
          const double id[5] = {Q.sc[24], Q.sc[25], Q.sc[26], Q.sc[28], Q.sc[27]};
          const double scale = 1.0 + fabs(vref.el) + fabs(vref.vw) + fabs(vref.hb) + fabs(vref.sol);
          bool bad = false;
          for (int k = 0; k < 5; ++k) bad |= fabs(fd[k] - id[k]) > 1e-9 * (scale + fabs(fd[k]));
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
      // ---- acceptance (leader): metropolis.f:72-87
      if (tid == 0) {
        enew = Q.sc[16];
        const double eol = Q.sc[0];
        double delta = weight_of(P, beta, enew) - weight_of(P, beta, eol);
        bool acc;
        if (delta <= 0.0) {
          acc = true;
        } else {
          const double rd = mt_grnd(Q.mt, &Q.mt[624]);
          delta = fmin(delta, 100.0);
          acc = exp(-delta) >= rd;
        }
        Q.sc[4] = acc ? 1.0 : 0.0;
        if (P.trace && r == P.trace_rep && *P.trace_n < P.trace_cap) {
          double* t = P.trace + 4 * (size_t)(*P.trace_n);
          t[0] = jv;
          t[1] = enew;
          t[2] = eol;
          t[3] = acc ? 1.0 : 0.0;
          *P.trace_n += 1;
        }
        if (acc) {
          Q.sc[0] = enew;
          Q.sc[1] += 1.0;
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
      } else {
        for (int a = m1 + tid; a <= m2; a += blockDim.x) {
          L.x[a] = Q.bx[a - m1];
          L.y[a] = Q.by[a - m1];
          L.z[a] = Q.bz[a - m1];
        }
      }
      __syncthreads();
      lap(kPhAccept);
    }
    // ---- sweep end: rebuild from the variables and check (metropolis.f:172-179)
    EnergyOut o;
    conformation_energy<MIXED>(S, Q.vl, L, Q.nf, o, wk);
    if (tid == 0) {
      const double enw = o.el + o.vw + o.hb + o.tor + o.sol;
      const double eol = Q.sc[0];
      if (fabs(eol - enw) > 0.000001 && eol < 100000.0) {
        atomicOr(P.err, 1);
        if (P.dbg[16 * r + 14] == 0.0) {  // first drift of this replica
          P.dbg[16 * r + 14] = eol;
          P.dbg[16 * r + 15] = enw;
        }
      }
      if (o.err) atomicOr(P.err, 2);
      Q.sc[0] = enw;
      const int slot = P.slot_of[r];
      if (P.traj) P.traj[(size_t)(P.traj_base + sw) * P.nrep + slot] = enw;
      if (P.hist) {
        double v = enw / P.ebin + copysign(0.5, enw);
        v = fmin(fmax(v, (double)P.kmin), (double)P.kmax);
        atomicAdd(&P.hist[(int)v - P.kmin], 1);
      }
    }
    __syncthreads();
    lap(kPhSweepEnd);
  }
  if (P.phase && tid == 0)
    for (int k = 0; k < 8; ++k) P.phase[8 * (size_t)r + k] += ph[k];
  // ---- write the replica state back
  for (int k = tid; k < nvt; k += blockDim.x) gvl[k] = Q.vl[k];
  for (int k = tid; k < 625; k += blockDim.x) gmt[k] = Q.mt[k];
  for (int a = tid; a < nat; a += blockDim.x) {
    gx[a] = L.x[a];
    gx[np + a] = L.y[a];
    gx[2 * np + a] = L.z[a];
  }
  for (int h = tid; h < nh; h += blockDim.x) gnf[h] = Q.nf[h];
  if (tid == 0) {
    P.eol[r] = Q.sc[0];
    P.acc[r] = Q.sc[1];
    P.valid[r] = 1;
  }
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
        P.work[4 * r + k] += t;
      }
  }
}

// Temperature exchanges between neighbouring slots (partem_s.f:209-232):
// pairs (1,2),(3,4).. then (2,3),..,(R,1) alternately; the random number is
// drawn for every pair.  One thread; stream 'xs' of P.mt.
__global__ void k_exchange(int nrep, int32_t* ipoi, int32_t* slot_of, const double* beta, const double* eol,
                           uint32_t* mt_x, int32_t* iswitch) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
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

}  // namespace smmp
