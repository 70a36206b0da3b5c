// Analytic solvent accessible surface and energy (itysol < 0) of esolan.f
// (projection method, J Comp Phys 26, 334-343, 2005) on gfx950, energy part.
//
// One wavefront per (conformation, atom), four independent wavefronts per
// workgroup, pulled from a work counter.  Per atom (esolan.f:122-1006):
//   1. neighbour search over all atoms, in atom order: the first atom that
//      encloses it ends the search (area 0), coinciding centres are an error;
//   2. north-point test; a degenerate north point (esolan.f:188-216) needs
//      random numbers from the twister in atom order, so such atoms are only
//      flagged here and done in a second launch after the host has drawn
//      their rotations in the reference's order;
//   3. coefficients a, b, c, d of every neighbour's circle in the plane of
//      the stereographic projection (esolan.f:238-247);
//   4. the pair loop (esolan.f:251-451): for the circle at list position k
//      (k = n-1 ... 1) every circle at positions n ... k+1 in that order,
//      which may drop the latter, drop circle k and end the inner loop, find
//      the atom covered, or give two intersection points.  The inner loop
//      runs 64 positions at a time; a drop of position L shifts only
//      positions above L, which the loop has passed, so each lane reads the
//      list as it was when the inner loop began, the first break event (in
//      loop order) is found with a ballot, the effects before it are kept,
//      and the list is compacted once per k;
//   5. circles without intersection points (full arcs, esolan.f:529-560) and
//      with them: their points are bucketed by circle, sorted by polar angle
//      about the circle's centre, and every arc whose midpoint no other
//      circle covers adds its term (esolan.f:563-863).
// The branch and list decisions use only +, -, *, /, sqrt with contraction
// off, as in the reference (they match it bit for bit); the arc terms use
// atan2/atan/sin/cos of the device library and are summed by wave trees, so
// areas agree with the reference to rounding.  The Cartesian gradient
// (gradan, esolan.f:533-1016) feeds only the minimiser (opesol.f) and is not
// computed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../host/engine.hpp"
#include "esolan_host.hpp"

#pragma clang fp contract(off)

namespace smmp {

#define HIPCHK(x)                                                                                   \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess) throw DeviceError(std::string(#x) + ": " + hipGetErrorString(e_));        \
  } while (0)

constexpr int kEsThreads = 256;  // four wavefronts per workgroup, each on its own atom
constexpr int kEsWaves = kEsThreads / kWave;
constexpr int kEsSlots = 192;    // circle slots of one atom (at most 191 neighbours)
constexpr int kEsMaxNb = kEsSlots - 1;
constexpr double kEsCoinc = (double)1e-10f;  // esolan.f:127 compares with a default-real literal
constexpr double kPi4 = 4.0 * kPi;

// error bits of a launch
constexpr int kEsErrCoinc = 1, kEsErrNb = 2, kEsErrCap = 4, kEsErrLine = 8;

// pair-loop outcomes (esolan.f:251-451)
constexpr int kNone = 0, kDropL = 1, kVert = 2, kDropK = 4, kZeroAs = 5, kCover = 6;

struct EsLds {
  double A[kEsSlots], B[kEsSlots], C[kEsSlots], D[kEsSlots];  // circle coefficients by slot
  // b/a, c/a, d/a and sqrt((b/a)^2 + (c/a)^2 - 4 d/a) of each circle: the
  // values esolan.f:255-268 forms for every pair, formed once
  double nB[kEsSlots], nC[kEsSlots], nD[kEsSlots], nR[kEsSlots];
  int16_t atom[kEsSlots];                                     // neighbour atom of each slot
  uint8_t lst[kEsSlots], tmp[kEsSlots];                       // circle list (slots) and compaction buffer
  uint8_t rm[kEsSlots];                                       // per list position: removed in this step
  uint8_t hasv[kEsSlots];                                     // per slot: an intersection point lies on it
  uint8_t fin[kEsSlots];                                      // per slot: final circle with points
  int32_t cnt[kEsSlots], off[kEsSlots], cur[kEsSlots];        // point buckets by slot
};

struct EsParams {
  const DevSys* systems;
  const int32_t* conf_sys;
  const double* xyz;     // conformation c at xyz + 3 * stride * c
  int nconf, stride;
  int nitems;            // items of this launch
  const int32_t* items;  // null: item = index; else the listed items
  const double* rot;     // with items: cg, sg, cf, sf per listed item (rotated atoms)
  int32_t* flag;         // per item: 1 = degenerate north point, left for the rotation launch
  int32_t* nflag;
  double* asig;          // per item: sigma * area
  int32_t* err;
  int32_t* next;         // work counter
  char* scratch;         // per wavefront slot: wave_bytes
  size_t wave_bytes;
  int cap;               // intersection points per atom
  double* counters;      // atom pairs scanned, circle pairs classified, arcs tested, coverage tests
};

struct EsScratch {
  double *vt, *vs;   // intersection points (cap)
  uint16_t* vc;      // their two circles: slot1 | slot2 << 8
  double *ang, *srt;  // bucket entries (2 cap): polar angle by entry, sorted within bucket
  uint16_t* ent;     // point of each entry
  uint8_t* entc;     // circle of each entry
};

__host__ __device__ inline size_t es_wave_bytes(int cap) {
  const size_t c = (size_t)cap;
  return c * 8 * 2 + 2 * c * 8 * 2 + c * 2 + 2 * c * 2 + 2 * c + 256;
}

__device__ inline EsScratch es_carve(char* p, int cap) {
  EsScratch s;
  const size_t c = (size_t)cap;
  s.vt = (double*)p;
  s.vs = s.vt + c;
  s.ang = s.vs + c;
  s.srt = s.ang + 2 * c;
  s.vc = (uint16_t*)(s.srt + 2 * c);
  s.ent = s.vc + c;
  s.entc = (uint8_t*)(s.ent + 2 * c);
  return s;
}

__device__ __forceinline__ uint64_t lanes_below(int lane) { return (lane == 0) ? 0ull : (~0ull >> (64 - lane)); }

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, kWave));
  return v;
}

// Two circles of the projection (slot K at list position k, slot Lc at a
// higher position): esolan.f:254-449.
__device__ __forceinline__ int es_classify(const EsLds& L, int K, int Lc, double v[4]) {
  const double ak = L.A[K], al = L.A[Lc];
  int kind = 0;
  if (ak > 0.0 && al > 0.0) kind = 1;
  else if (ak > 0.0 && al < 0.0) kind = 2;
  else if (ak < 0.0 && al > 0.0) kind = 3;
  else if (ak < 0.0 && al < 0.0) kind = 4;
  if (!kind) return kNone;
  const double b1 = L.nB[K], c1 = L.nC[K], d1 = L.nD[K], r1 = L.nR[K];
  const double b2 = L.nB[Lc], c2 = L.nC[Lc], d2 = L.nD[Lc], r2 = L.nR[Lc];
  const double D = 4.0 * ((b1 - b2) * (b2 * d1 - b1 * d2) + (c1 - c2) * (c2 * d1 - c1 * d2) - (d1 - d2) * (d1 - d2)) +
                   (b1 * c2 - b2 * c1) * (b1 * c2 - b2 * c1);
  const double dist = sqrt((b1 - b2) * (b1 - b2) + (c1 - c2) * (c1 - c2));
  int code = kNone;
  if (kind == 1) {  // a01
    if (D <= 0.0 && dist <= r2 - r1) code = kDropK;
    else if (D <= 0.0 && dist <= -r2 + r1) code = kDropL;
    else if (D > 0.0) code = kVert;
  } else if (kind == 2) {  // a03
    if (D <= 0.0 && dist <= -r2 + r1) code = kCover;
    else if (D <= 0.0 && dist >= r2 + r1) code = kDropK;
    else if (D > 0.0) code = kVert;
  } else if (kind == 3) {  // a07
    if (D <= 0.0 && dist <= r2 - r1) code = kZeroAs;
    else if (D <= 0.0 && dist >= r2 + r1) code = kDropL;
    else if (D > 0.0) code = kVert;
  } else {  // a09
    if (D <= 0.0 && dist <= r2 - r1) code = kDropL;
    else if (D <= 0.0 && dist <= -r2 + r1) code = kDropK;
    else if (D <= 0.0) code = kCover;
    else code = kVert;
  }
  if (code == kVert) {
    const double am = 2.0 * ((b1 - b2) * (b1 - b2) + (c1 - c2) * (c1 - c2));
    const double t = -2.0 * (b1 - b2) * (d1 - d2) - (b2 * c1 - b1 * c2) * (c1 - c2);
    const double s = -2.0 * (c1 - c2) * (d1 - d2) + (b2 * c1 - b1 * c2) * (b1 - b2);
    const double pm = sqrt(D), p1 = (c1 - c2) * pm, p2 = (b1 - b2) * pm;
    v[0] = (t + p1) / am;
    v[1] = (s - p2) / am;
    v[2] = (t - p1) / am;
    v[3] = (s + p2) / am;
  }
  return code;
}

// One atom of one conformation.  Returns false if the item was left for the
// rotation launch.
__device__ void es_item(const EsParams& P, EsLds& L, const EsScratch& X, int item, const double* rot, int lane,
                        double* work) {
  const int c = item / P.stride, i = item - c * P.stride;
  const DevSys& S = P.systems[P.conf_sys[c]];
  if (i >= S.nat) {
    if (lane == 0) P.asig[item] = 0.0;
    return;
  }
  const double* xyz = P.xyz + (size_t)3 * P.stride * c;
  const double R = S.rvdw[i];
  const double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
  const uint64_t below = lanes_below(lane);

  // 1. neighbours in atom order (esolan.f:122-150)
  int nb = 0;
  bool enclosed = false;
  for (int base = 0; base < S.nat; base += kWave) {
    const int j = base + lane;
    work[0] += (j < S.nat && j != i) ? 1.0 : 0.0;
    bool coin = false, encl = false, nbr = false;
    if (j < S.nat && j != i) {
      const double dx = xyz[3 * j] - xi, dy = xyz[3 * j + 1] - yi, dz = xyz[3 * j + 2] - zi;
      const double ddp = dx * dx + dy * dy + dz * dz;
      if (ddp < kEsCoinc) {
        coin = true;
      } else {
        const double d = sqrt(ddp), rj = S.rvdw[j];
        if (d + R <= rj) encl = true;
        else if (!(d <= R - rj || d >= R + rj)) nbr = true;
      }
    }
    const uint64_t ev = __ballot(coin || encl);
    const int first = ev ? __ffsll((unsigned long long)ev) - 1 : kWave;
    const bool take = nbr && lane < first;
    const uint64_t nm = __ballot(take);
    if (take) {
      const int pos = nb + __popcll(nm & below);
      if (pos < kEsMaxNb) L.atom[pos] = (int16_t)j;
    }
    nb += __popcll(nm);
    if (ev) {
      if (__shfl(coin ? 1 : 0, first, kWave)) {
        if (lane == 0) atomicOr(P.err, kEsErrCoinc);
        return;
      }
      enclosed = true;
      break;
    }
  }
  if (nb > kEsMaxNb) {
    if (lane == 0) atomicOr(P.err, kEsErrNb);
    return;
  }
  const double pol = R * R;
  if (enclosed || nb == 0) {  // esolan.f:138-140, 154
    if (lane == 0) P.asig[item] = (enclosed ? 0.0 : kPi4 * pol) * S.sigma[i];
    return;
  }
  wave_lds_sync();

  // 2. the north point against the neighbouring surfaces (esolan.f:181-195)
  if (!rot) {
    double mn = INFINITY;
    for (int s = lane; s < nb; s += kWave) {
      const int a = L.atom[s];
      const double x = xyz[3 * a] - xi, y = xyz[3 * a + 1] - yi, z = xyz[3 * a + 2] - zi, rk = S.rvdw[a];
      mn = fmin(mn, fabs(x * x + y * y + (R - z) * (R - z) - rk * rk));
    }
    if (wave_min(mn) < 1.0e-6) {
      if (lane == 0) {
        P.flag[item] = 1;
        atomicAdd(P.nflag, 1);
        P.asig[item] = 0.0;
      }
      return;
    }
  }

  // 3. circle coefficients, after the rotation if any (esolan.f:219-247)
  const double pom = 8.0 * pol;
  for (int s = lane; s < nb; s += kWave) {
    const int a = L.atom[s];
    double x = xyz[3 * a] - xi, y = xyz[3 * a + 1] - yi, z = xyz[3 * a + 2] - zi;
    if (rot) {
      const double cg = rot[0], sg = rot[1], cf = rot[2], sf = rot[3];
      const double sfsg = sf * sg, sfcg = sf * cg, cfsg = cf * sg, cfcg = cf * cg;
      const double xx = x, yy = y, zz = z;
      x = xx * cg - zz * sg;
      y = -xx * sfsg + yy * cf - zz * sfcg;
      z = xx * cfsg + yy * sf + zz * cfcg;
    }
    const double rk = S.rvdw[a], pk = rk * rk;
    const double ca = x * x + y * y + (z - R) * (z - R) - pk, cb = -pom * x, cc = -pom * y;
    const double cd = 4.0 * pol * (x * x + y * y + (R + z) * (R + z) - pk);
    L.A[s] = ca;
    L.B[s] = cb;
    L.C[s] = cc;
    L.D[s] = cd;
    if (ca != 0.0) {
      const double nb = cb / ca, nc = cc / ca, nd = cd / ca;
      L.nB[s] = nb;
      L.nC[s] = nc;
      L.nD[s] = nd;
      L.nR[s] = sqrt(nb * nb + nc * nc - 4.0 * nd);
    }
    L.lst[s] = (uint8_t)s;
    L.rm[s] = 0;
    L.hasv[s] = 0;
    L.fin[s] = 0;
  }
  wave_lds_sync();

  // 4. pairs of circles (esolan.f:251-451)
  int n = nb, k = nb, iv = 0;
  bool covered = false, zero_as = false, overflow = false;
  while (k > 1) {
    k = k - 1;
    const int cnt = n - k;  // positions n ... k+1, fixed when the loop starts
    if (cnt <= 0) continue;
    const int K = L.lst[k - 1];
    bool brk = false, anyrm = false;
    for (int base = 0; base < cnt && !brk; base += kWave) {
      const int lp = n - base - lane;  // 1-based list position of this lane
      const bool valid = base + lane < cnt;
      int code = kNone, Lc = 0;
      double v[4] = {0.0, 0.0, 0.0, 0.0};
      if (valid) {
        Lc = L.lst[lp - 1];
        code = es_classify(L, K, Lc, v);
      }
      work[1] += valid ? 1.0 : 0.0;
      const uint64_t bm = __ballot(code >= kDropK);
      const int first = bm ? __ffsll((unsigned long long)bm) - 1 : kWave;
      const bool live = valid && lane < first;
      if (live && code == kDropL) L.rm[lp - 1] = 1;
      anyrm = anyrm || __ballot(live && code == kDropL) != 0;
      const uint64_t vm = __ballot(live && code == kVert);
      if (vm) {
        if (live && code == kVert) {
          const int o = iv + 2 * __popcll(vm & below);
          if (o + 2 <= P.cap) {
            X.vt[o] = v[0];
            X.vs[o] = v[1];
            X.vt[o + 1] = v[2];
            X.vs[o + 1] = v[3];
            X.vc[o] = X.vc[o + 1] = (uint16_t)(K | (Lc << 8));
          }
          L.hasv[K] = 1;
          L.hasv[Lc] = 1;
        }
        iv += 2 * __popcll(vm);
      }
      if (bm) {
        const int bcode = __shfl(code, first, kWave);
        if (bcode == kCover) {
          covered = true;
        } else if (bcode == kDropK) {
          if (lane == 0) L.rm[k - 1] = 1;
          anyrm = true;
        } else {
          zero_as = true;  // esolan.f:361-364: As = 0, next k
        }
        brk = true;
      }
    }
    if (iv > P.cap) {
      overflow = true;
      break;
    }
    if (covered) break;
    wave_lds_sync();
    if (anyrm) {  // stable removal of the dropped positions
      int nn = k - 1;
      for (int base = k - 1; base < n; base += kWave) {
        const int p = base + lane;
        const bool keep = p < n && !L.rm[p];
        const uint64_t km = __ballot(keep);
        if (keep) L.tmp[nn + __popcll(km & below)] = L.lst[p];
        if (p < n) L.rm[p] = 0;
        nn += __popcll(km);
      }
      wave_lds_sync();
      for (int p = k - 1 + lane; p < nn; p += kWave) L.lst[p] = L.tmp[p];
      n = nn;
      wave_lds_sync();
    }
  }
  if (overflow) {
    if (lane == 0) atomicOr(P.err, kEsErrCap);
    return;
  }
  if (covered) {  // esolan.f:325-326, 428-429
    if (lane == 0) P.asig[item] = 0.0;
    return;
  }

  // 5. areas (esolan.f:453-963)
  const double R2 = R * R, R22 = R2 + R2, R42 = R22 + R22, R22p = R22 * kPi;
  bool line = false, inner = false;
  double full = 0.0;
  for (int p = lane; p < n; p += kWave) {
    const int s = L.lst[p];
    const double a = L.A[s], b = L.B[s], cc = L.C[s], d = L.D[s];
    line = line || a == 0.0;
    inner = inner || a < 0.0;
    if (!L.hasv[s]) {  // a full arc: the cap it cuts off (esolan.f:529-560)
      if (a != 0.0) {
        const double aia = a < 0.0 ? -1.0 : 1.0;
        full += aia * R22p * (1.0 + (-d / a - R42) * fabs(a) / sqrt((R42 * a - d) * (R42 * a - d) + R42 * (b * b + cc * cc)));
      } else {
        full -= R22p * d / sqrt(R42 * (b * b + cc * cc) + d * d);
      }
    } else {
      L.fin[s] = 1;
    }
  }
  if (__ballot(line)) {
    // a == 0 makes the north point lie on that sphere, which the north-point
    // test turns into a rotation first; the lines case of esolan.f:471-500
    // is not reached
    if (lane == 0) atomicOr(P.err, kEsErrLine);
    return;
  }
  double As = zero_as ? 0.0 : kPi4 * pol;
  if ((nb > 0 && n < 1) || __ballot(inner)) As = 0.0;  // esolan.f:501-502
  for (int s = lane; s < nb; s += kWave) L.cnt[s] = 0;
  wave_lds_sync();

  // intersection points bucketed by circle, circles in list order
  for (int v = lane; v < iv; v += kWave) {
    const int w = X.vc[v], c1 = w & 255, c2 = w >> 8;
    if (L.fin[c1]) atomicAdd(&L.cnt[c1], 1);
    if (L.fin[c2]) atomicAdd(&L.cnt[c2], 1);
  }
  wave_lds_sync();
  int E = 0;
  for (int base = 0; base < n; base += kWave) {
    const int p = base + lane;
    const int s = p < n ? L.lst[p] : 0;
    int m = (p < n && L.fin[s]) ? L.cnt[s] : 0;
    int incl = m;  // inclusive prefix over lanes
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const int t = __shfl_up(incl, o, kWave);
      if (lane >= o) incl += t;
    }
    if (p < n && L.fin[s]) {
      L.off[s] = E + incl - m;
      L.cur[s] = E + incl - m;
    }
    E += __shfl(incl, kWave - 1, kWave);
  }
  wave_lds_sync();
  for (int v = lane; v < iv; v += kWave) {
    const int w = X.vc[v], c1 = w & 255, c2 = w >> 8;
    if (L.fin[c1]) {
      const int q = atomicAdd(&L.cur[c1], 1);
      X.ent[q] = (uint16_t)v;
      X.entc[q] = (uint8_t)c1;
    }
    if (L.fin[c2]) {
      const int q = atomicAdd(&L.cur[c2], 1);
      X.ent[q] = (uint16_t)v;
      X.entc[q] = (uint8_t)c2;
    }
  }
  wave_lds_sync();
  // polar angles about each circle's centre (esolan.f:626-640)
  for (int e = lane; e < E; e += kWave) {
    const int s = X.entc[e], v = X.ent[e];
    const double a = L.A[s], aa = 0.5 / a;
    X.ang[e] = atan2(X.vs[v] + L.C[s] * aa, X.vt[v] + L.B[s] * aa);
  }
  wave_lds_sync();
  // ascending within each bucket (esolan.f:642-659)
  for (int e = lane; e < E; e += kWave) {
    const int s = X.entc[e], lo = L.off[s], hi = lo + L.cnt[s];
    const double x = X.ang[e];
    int r = 0;
    for (int q = lo; q < hi; ++q) {
      const double y = X.ang[q];
      r += (y < x || (y == x && q < e)) ? 1 : 0;
    }
    X.srt[lo + r] = x;
  }
  wave_lds_sync();
  // arcs between consecutive points whose midpoint no other circle covers
  double arcs = 0.0;
  for (int p = lane; p < E; p += kWave) {
    const int s = X.entc[p], lo = L.off[s], hi = lo + L.cnt[s];
    const double alp = X.srt[p], bet = (p + 1 < hi) ? X.srt[p + 1] : X.srt[lo] + kPi2;
    if (fabs(bet - alp) < 1e-8) continue;
    const double a = L.A[s], b = L.B[s], cc = L.C[s], d = L.D[s];
    const double ak = a * a, daba = fabs(a), aia = a < 0.0 ? -1.0 : 1.0, b2c2 = b * b + cc * cc;
    const double vv = sqrt((R42 * a - d) * (R42 * a - d) + R42 * b2c2), vv1 = 1.0 / vv;
    const double wv = sqrt(b2c2 - 4.0 * a * d);
    const double aa = 0.5 / a, ba = b * aa, ca = cc * aa, bcd = b2c2 - 2.0 * a * d;
    const double rr = 0.5 * wv / daba;
    const double prob = (alp + bet) / 2.0;
    double sp, cp;
    sincos(prob, &sp, &cp);
    const double ap1 = -ba + rr * cp, ap2 = -ca + rr * sp;
    const double q2 = ap1 * ap1 + ap2 * ap2;
    bool cov = false;
    int q = 0;
    for (; q < n && !cov; ++q) {
      const int t = L.lst[q];
      if (t != s && L.A[t] * q2 + L.B[t] * ap1 + L.C[t] * ap2 + L.D[t] < 0.0) cov = true;
    }
    work[2] += 1.0;
    work[3] += (double)q;
    if (cov) continue;
    const double hm = 5e-1 * (bet - alp), hp = 5e-1 * (bet + alp);
    double dsbetmalp, dcbetmalp, dsbetpalp, dcbetpalp;
    sincos(hm, &dsbetmalp, &dcbetmalp);
    sincos(hp, &dsbetpalp, &dcbetpalp);
    const double uv = daba * (bcd + 2.0 * R42 * ak) * dcbetmalp - a * wv * (b * dcbetpalp + cc * dsbetpalp);
    arcs += R2 * (aia * (alp - bet) + (d + R42 * a) * vv1 * (kPi - 2.0 * atan(uv / (2.0 * ak * vv * dsbetmalp))));
  }
  As = As - wave_sum(full) + wave_sum(arcs);
  if (lane == 0) P.asig[item] = As * S.sigma[i];
}

__global__ void __launch_bounds__(kEsThreads) k_esolan(EsParams P) {
  extern __shared__ double es_smem[];
  const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  EsLds& L = reinterpret_cast<EsLds*>(es_smem)[wave];
  const size_t slot = (size_t)blockIdx.x * kEsWaves + wave;
  const EsScratch X = es_carve(P.scratch + slot * P.wave_bytes, P.cap);
  double n[4] = {0.0, 0.0, 0.0, 0.0};
  for (;;) {
    int idx = 0;
    if (lane == 0) idx = atomicAdd(P.next, 1);
    idx = __shfl(idx, 0, kWave);
    if (idx >= P.nitems) break;
    const int item = P.items ? P.items[idx] : idx;
    es_item(P, L, X, item, P.items ? P.rot + 4 * (size_t)idx : nullptr, lane, n);
    wave_lds_sync();
  }
  for (int k = 0; k < 4; ++k) n[k] = wave_sum(n[k]);
  if (lane == 0 && P.counters)
    for (int k = 0; k < 4; ++k) atomicAdd(&P.counters[k], n[k]);
}

// eysl of each conformation (one wavefront per conformation) and the total
// energy.f:76-82, 114: (eysm + eysl) + eysmi.
__global__ void k_esolan_sum(const DevSys* __restrict__ systems, const int32_t* __restrict__ conf_sys,
                             const double* __restrict__ asig, int nconf, int stride, double* __restrict__ out,
                             int out_stride) {
  const int c = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  if (c >= nconf) return;
  const int nat = systems[conf_sys[c]].nat;
  double s = 0.0;
  for (int a = lane; a < nat; a += kWave) s += asig[(size_t)stride * c + a];
  s = wave_sum(s);
  if (lane == 0) {
    double* r = out + (size_t)out_stride * c;
    r[6] = s;
    r[0] = r[1] + s + r[8];
  }
}

// ------------------------------------------------------------------ host

struct EsBuf {
  void* p = nullptr;
  size_t n = 0;
  ~EsBuf() {
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

struct EsolanWork {
  EsBuf flag, small, asig, items, rot, scratch;
  int cap = 8192;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  EsolanStats last;
  ~EsolanWork() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
  }
};

void esolan_work_destroy(EsolanWork* w) { delete w; }

EsolanStats esolan_last_stats(EsolanWork* W) { return W ? W->last : EsolanStats{}; }

static int es_grid_waves(int nitems, size_t wave_bytes) {
  constexpr size_t kBudget = size_t(3) << 30;  // scratch of one launch
  int waves = std::min(nitems, 256 * 12);
  waves = (int)std::min<size_t>((size_t)waves, std::max<size_t>(kEsWaves, kBudget / wave_bytes));
  return std::max(kEsWaves, (waves + kEsWaves - 1) / kEsWaves * kEsWaves);
}

// One launch over nitems items; returns the error bits.  Grows the point
// capacity and repeats when an atom has more intersection points.
static int es_launch(EsolanWork& W, hipStream_t st, EsParams P, int32_t* d_small, double* d_counters) {
  for (;;) {
    const size_t wb = (es_wave_bytes(W.cap) + 255) / 256 * 256;
    const int waves = es_grid_waves(P.nitems, wb);
    P.scratch = static_cast<char*>(W.scratch.ensure(wb * (size_t)waves));
    P.wave_bytes = wb;
    P.cap = W.cap;
    HIPCHK(hipMemsetAsync(d_small, 0, 3 * sizeof(int32_t), st));  // err, next, nflag
    P.err = d_small;
    P.next = d_small + 1;
    P.nflag = d_small + 2;
    P.counters = d_counters;
    hipLaunchKernelGGL(k_esolan, dim3(waves / kEsWaves), dim3(kEsThreads), sizeof(EsLds) * kEsWaves, st, P);
    HIPCHK(hipGetLastError());
    int32_t h[3];
    HIPCHK(hipMemcpyAsync(h, d_small, sizeof h, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if ((h[0] & kEsErrCap) && W.cap < 65536) {
      W.cap *= 2;
      continue;
    }
    ++W.last.passes;
    return h[0] | (h[2] << 8);
  }
}

static void es_check(int err) {
  if (err & kEsErrCoinc) throw Error("esolan> centres of two atoms coincide");
  if (err & kEsErrNb) throw Error("esolan> more than 191 solvent neighbours of one atom");
  if (err & kEsErrCap) throw Error("esolan> more than 65536 circle intersections at one atom");
  if (err & kEsErrLine) throw Error("esolan> a circle of the projection is a line");
}

void esolan_energies(EsolanWork*& Wp, hipStream_t st, const DevSys* d_systems, const std::vector<const HostSys*>& hsys,
                     const std::vector<int32_t>& conf_sys, const int32_t* d_conf_sys, int nconf, int stride,
                     const double* d_xyz, double* d_out, int out_stride, const std::function<double(int)>& draw) {
  if (!Wp) Wp = new EsolanWork();
  EsolanWork& W = *Wp;
  if (!W.ev0) {
    HIPCHK(hipEventCreate(&W.ev0));
    HIPCHK(hipEventCreate(&W.ev1));
  }
  W.last = EsolanStats{};
  if (nconf == 0) return;
  const int nitems = nconf * stride;
  int32_t* d_flag = static_cast<int32_t*>(W.flag.ensure(sizeof(int32_t) * nitems));
  double* d_asig = static_cast<double*>(W.asig.ensure(sizeof(double) * nitems));
  char* small = static_cast<char*>(W.small.ensure(96));
  int32_t* d_small = reinterpret_cast<int32_t*>(small);
  double* d_counters = reinterpret_cast<double*>(small + 32);
  HIPCHK(hipMemsetAsync(d_flag, 0, sizeof(int32_t) * nitems, st));
  HIPCHK(hipMemsetAsync(d_counters, 0, 4 * sizeof(double), st));
  HIPCHK(hipEventRecord(W.ev0, st));

  EsParams P{};
  P.systems = d_systems;
  P.conf_sys = d_conf_sys;
  P.xyz = d_xyz;
  P.nconf = nconf;
  P.stride = stride;
  P.nitems = nitems;
  P.flag = d_flag;
  P.asig = d_asig;
  int err = es_launch(W, st, P, d_small, d_counters);
  es_check(err & 255);
  const int nflag = err >> 8;

  if (nflag > 0) {
    // degenerate north points: rotations drawn in the reference's order
    // (conformations in order, atoms in order; esolan.f:196-232)
    std::vector<int32_t> flag(nitems);
    HIPCHK(hipMemcpyAsync(flag.data(), d_flag, sizeof(int32_t) * nitems, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<int32_t> items;
    std::vector<double> rot;
    std::vector<double> xyz;
    int xyz_conf = -1;
    for (int it = 0; it < nitems; ++it) {
      if (!flag[it]) continue;
      const int c = it / stride, i = it - c * stride;
      const HostSys& h = *hsys[conf_sys[c]];
      if (c != xyz_conf) {
        xyz.resize((size_t)3 * h.nat);
        HIPCHK(hipMemcpy(xyz.data(), d_xyz + (size_t)3 * stride * c, sizeof(double) * 3 * h.nat,
                         hipMemcpyDeviceToHost));
        xyz_conf = c;
      }
      const double R = h.rvdw[i];
      std::vector<int> nb;
      for (int j = 0; j < h.nat; ++j) {
        if (j == i) continue;
        const double dx = xyz[3 * j] - xyz[3 * i], dy = xyz[3 * j + 1] - xyz[3 * i + 1],
                     dz = xyz[3 * j + 2] - xyz[3 * i + 2];
        const double d = std::sqrt(dx * dx + dy * dy + dz * dz), rj = h.rvdw[j];
        if (!(d <= R - rj || d >= R + rj)) nb.push_back(j);
      }
      double ddp = 0.0, cf = 1.0, sf = 0.0, cg = 1.0, sg = 0.0;
      do {
        double uga = draw(c), ufi = draw(c);
        uga = kPi * uga / 2;
        ufi = kPi * ufi / 2;
        cf = std::cos(ufi);
        sf = std::sin(ufi);
        cg = std::cos(uga);
        sg = std::sin(uga);
        const double xx = R * (cf * sg), yy = R * sf, zz = R * (cf * cg);
        ddp = INFINITY;
        for (int j : nb) {
          const double x = xyz[3 * j] - xyz[3 * i], y = xyz[3 * j + 1] - xyz[3 * i + 1],
                       z = xyz[3 * j + 2] - xyz[3 * i + 2], rk = h.rvdw[j];
          ddp = std::min(ddp, std::fabs((xx - x) * (xx - x) + (yy - y) * (yy - y) + (zz - z) * (zz - z) - rk * rk));
        }
      } while (ddp < 1.0e-6);
      items.push_back(it);
      rot.insert(rot.end(), {cg, sg, cf, sf});
    }
    int32_t* d_items = static_cast<int32_t*>(W.items.ensure(sizeof(int32_t) * items.size()));
    double* d_rot = static_cast<double*>(W.rot.ensure(sizeof(double) * rot.size()));
    HIPCHK(hipMemcpyAsync(d_items, items.data(), sizeof(int32_t) * items.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_rot, rot.data(), sizeof(double) * rot.size(), hipMemcpyHostToDevice, st));
    P.nitems = (int)items.size();
    P.items = d_items;
    P.rot = d_rot;
    err = es_launch(W, st, P, d_small, d_counters);
    es_check(err & 255);
  }
  const int per = 256 / kWave;
  hipLaunchKernelGGL(k_esolan_sum, dim3((nconf + per - 1) / per), dim3(256), 0, st, d_systems, d_conf_sys, d_asig,
                     nconf, stride, d_out, out_stride);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(W.ev1, st));
  double cnt[4];
  HIPCHK(hipMemcpyAsync(cnt, d_counters, sizeof cnt, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, W.ev0, W.ev1));
  W.last.ms = ms;
  for (int k = 0; k < 4; ++k) W.last.work[k] = cnt[k];
}

}  // namespace smmp
