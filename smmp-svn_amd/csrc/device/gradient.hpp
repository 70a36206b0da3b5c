// Energy and its derivatives by the torsions of one conformation in vacuum
// (gradient.f:13-70 with opeshe.f:12-388), by one workgroup, FP64:
//   1. per-atom forces f_i = -dE/dr_i over the pairs within each chain:
//      the items (row block of 64 atoms, half of the column words) go to
//      the waves in turn and each lane sums its atom's part of the row into
//      one of two partial force arrays (every pair is evaluated from both
//      ends, so no two lanes write one value and the sums are
//      deterministic); the pair energies are counted from the lower end;
//   2. per variable v, the sums over its moved atoms of f and f x r; then
//      dE/dv = s [(e x b) . F + e . T] - dE_tors/dv, e the unit bond axis,
//      b the base atom and s = +1, or -1 where the moved range is the part
//      of the chain before the axis (the first torsions of a chain: the
//      intramolecular energy does not change when the whole chain turns, so
//      either side of the axis gives the derivative).
// The reference instead visits the variables from the chain's end and adds
// the sums of the variables nested in each moving set; both are the same
// derivative and agree to rounding.  The terms between chains are not part
// of it (gradient.f sums opeshe over the chains).
#pragma once
#include "energy.hpp"

namespace smmp {

// Values per conformation: eysm, eyel, eyvw, eyhb, eyvr.
constexpr int kGradOut = 5;

// The force on atom i (lane) from its partners in columns [64 bj, 64 bj + 64)
// within its chain; adds the energies of the pairs with j > i.
__device__ __forceinline__ void row_forces(const DevSys& S, const EnergyLds& L, int i, int bj, double& fx,
                                           double& fy, double& fz, double& el, double& vw, double& hb, Work& wk) {
  const bool vi = i < S.nat;
  uint64_t w0 = 0, w1 = 0;
  double xi = 0, yi = 0, zi = 0, qi = 0;
  int ti = 0;
  if (vi) {
    w0 = pair_bits(L, ((size_t)i * S.words + bj) * 2 + 0);
    w1 = pair_bits(L, ((size_t)i * S.words + bj) * 2 + 1);
    const uint64_t in = chain_cols(S, i, bj);
    w0 &= in;
    w1 &= in;
    xi = L.x[i];
    yi = L.y[i];
    zi = L.z[i];
    qi = S.conv * L.q[i];
    ti = L.t[i] * kTy;
  }
  const int kmax = min(64, S.nat - bj * 64);
  for (int k = 0; k < kmax; ++k) {
    const bool on = (w0 >> k) & 1ull;
    if (!__any(on)) continue;
    if (on) {
      const int j = bj * 64 + k;
      const double dx = L.x[j] - xi, dy = L.y[j] - yi, dz = L.z[j] - zi;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double ri = rsqrt_f64(r2);
      const double r2i = ri * ri;
      const double r6i = r2i * r2i * r2i;
      const int tt = ti + L.t[j];
      const bool is14 = (w1 >> k) & 1ull;
      const double A = L.ptab[4 * tt + (is14 ? 1 : 0)];
      const double Cc = L.ptab[4 * tt + 2];
      const double qq = qi * L.q[j];
      double eel, deel;
      if (S.epsd) {  // sigmoidal dielectric: d(q/(r eps))/dr = -deel/r
        const double rij = r2 * ri, sr = S.slp * rij;
        const double xsr = (S.plt - 1.0) * exp(-sr) / 2.0;
        const double ep = S.plt - (sr * sr + 2.0 * sr + 2.0) * xsr;
        eel = qq * ri / ep;
        deel = eel + qq * (S.slp * sr * sr * xsr) / (ep * ep);
      } else {
        eel = qq * ri;
        deel = eel;
      }
      const double rp = A * (r6i * r6i);
      double c;
      const bool once = j > i;
      if (once) el += eel;
      if (L.ptab[4 * tt + 3] != 0.0) {  // 10-12 hydrogen bond
        const double ds = Cc * (r6i * r2i * r2i);
        if (once) hb += rp - ds;
        c = (-12.0 * rp + 10.0 * ds - deel) * r2i;
      } else {
        const double ds = Cc * r6i;
        if (once) vw += rp - ds;
        c = (-12.0 * rp + 6.0 * ds - deel) * r2i;
      }
      fx += c * dx;  // c = (dE/dr) / r: c (r_j - r_i) is the force on i
      fy += c * dy;
      fz += c * dz;
      if (once) wk.pairs += 1.0;
    }
  }
}

// Forces into f (LDS, 6 per atom slot: the partial sums over the first and
// the second half of the column words, at f[3 i] and f[3 np + 3 i]) and the
// pair energies of this wave; the waves take the (row block, half) items in
// turn.  All threads; a barrier must follow before f is read.
__device__ __forceinline__ void pair_forces(const DevSys& S, const EnergyLds& L, double* f, double& el, double& vw,
                                            double& hb, Work& wk) {
  const int nwaves = blockDim.x / kWave, wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const int nt = (S.nat + 63) / 64, wh = (S.words + 1) / 2;
  for (int item = wave; item < 2 * nt; item += nwaves) {
    const int bi = item >> 1, h = item & 1;
    const int i = 64 * bi + lane;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const int b0 = h ? wh : 0, b1 = h ? S.words : wh;
    for (int bj = b0; bj < b1; ++bj) row_forces(S, L, i, bj, fx, fy, fz, el, vw, hb, wk);
    if (i < S.nat) {
      double* o = f + (size_t)3 * S.np * h + 3 * i;
      o[0] = fx;
      o[1] = fy;
      o[2] = fz;
    }
  }
}

// dE/dv for every variable into g (global, nvt); wave w takes variables w,
// w + nwaves, ...; the lanes sum over the moved atoms.
__device__ __forceinline__ void torsion_derivatives(const DevSys& S, const double* vl, const EnergyLds& L,
                                                    const double* f, double* __restrict__ g) {
  const int nwaves = blockDim.x / kWave, wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  for (int v = wave; v < S.nvt; v += nwaves) {
    const int m1 = S.var_m1[v], m2 = S.var_m2[v];
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double* f2 = f + (size_t)3 * S.np;
    for (int a = m1 + lane; a <= m2; a += kWave) {
      const double fx = f[3 * a] + f2[3 * a], fy = f[3 * a + 1] + f2[3 * a + 1], fz = f[3 * a + 2] + f2[3 * a + 2];
      const double x = L.x[a], y = L.y[a], z = L.z[a];
      s[0] += fx;
      s[1] += fy;
      s[2] += fz;
      s[3] += fy * z - fz * y;  // f x r
      s[4] += fz * x - fx * z;
      s[5] += fx * y - fy * x;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] = wave_sum(s[k]);
    if (lane == 0) {
      const int b = S.var_base[v], p = S.var_axis0[v];
      const double xb = L.x[b], yb = L.y[b], zb = L.z[b];
      double ex = xb - L.x[p], ey = yb - L.y[p], ez = zb - L.z[p];
      const double en = rsqrt_f64(ex * ex + ey * ey + ez * ez);
      ex *= en;
      ey *= en;
      ez *= en;
      double gv = S.var_sign[v] * ((ey * zb - ez * yb) * s[0] + (ez * xb - ex * zb) * s[1] +
                                   (ex * yb - ey * xb) * s[2] + ex * s[3] + ey * s[4] + ez * s[5]);
      const double e0 = S.var_e0[v];
      if (e0 != 0.0) {  // torsion term e0 (1 + sg cos(rn t)): derivative -e0 sg rn sin(rn t)
        const double rn = S.var_rn[v];
        gv -= e0 * S.var_sg[v] * rn * sin(atom_torsion(S, vl, S.var_atom[v]) * rn);
      }
      g[v] = gv;
    }
  }
}

}  // namespace smmp
