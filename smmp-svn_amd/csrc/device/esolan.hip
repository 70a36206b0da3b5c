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
// areas agree with the reference to rounding.
//
// The gradient form (k_esolan<true>, for the minimiser's opesol.f) also
// forms the Cartesian derivatives of the atom's area by every neighbour
// (esolan.f:547-568 for full arcs, 684-862 for arcs: the coefficients of
// the arc's own circle and of the two circles crossing it at its ends),
// rotates them back with the atom's rotation, takes the atom's own as minus
// their sum over the final circle list (esolan.f:966-1000), and writes
// sigma_i times each as a row of the item; k_esolan_grad_sum then forms
// gradan_j = sum_i sigma_i dAs_i/dr_j in atom order for every atom j.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
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
constexpr int kEsAngLds = 4 * kEsSlots;  // LDS doubles for the bucket angles
constexpr int kEsTri = 12 * kEsSlots;    // LDS bytes for the pair outcomes
constexpr int kEsVp = 512;               // kept pairs held in LDS
constexpr double kEsCoinc = (double)1e-10f;  // esolan.f:127 compares with a default-real literal
constexpr double kPi4 = 4.0 * kPi;

// error bits of a launch
constexpr int kEsErrCoinc = 1, kEsErrNb = 2, kEsErrCap = 4, kEsErrLine = 8;

// pair-loop outcomes (esolan.f:251-451)
constexpr int kNone = 0, kDropL = 1, kVert = 2, kDropK = 4, kZeroAs = 5, kCover = 6;

struct EsLds {
  double A[kEsSlots], B[kEsSlots], C[kEsSlots], D[kEsSlots];  // circle coefficients by slot
  union {
    // b/a, c/a, d/a and sqrt((b/a)^2 + (c/a)^2 - 4 d/a) of each circle: the
    // values esolan.f:255-268 forms for every pair, formed once
    struct {
      double nB[kEsSlots], nC[kEsSlots], nD[kEsSlots], nR[kEsSlots];
    };
    // after the pair loop: the polar angles of the bucket entries and, when
    // both fit, their sorted copy (kEsAngLds entries in all)
    double ang[4 * kEsSlots];
  };
  int16_t atom[kEsSlots];                                     // neighbour atom of each slot
  uint8_t lst[kEsSlots], tmp[kEsSlots];                       // circle list (slots) and compaction buffer
  uint8_t rm[kEsSlots];                                       // per list position: removed in this step
  uint8_t hasv[kEsSlots];                                     // per slot: an intersection point lies on it
  uint8_t fin[kEsSlots];                                      // per slot: final circle with points
  uint16_t vp[kEsVp];        // kept pairs with intersection points (K | Lc << 8), in their order
  uint8_t entc[kEsAngLds];   // circle of each bucket entry (entries in LDS)
  union {
    struct {
      int32_t cnt[kEsSlots], off[kEsSlots], cur[kEsSlots];  // point buckets by slot
    };
    // during the pair loop: the outcome of every pair of slots K < Lc at
    // Lc (Lc - 1) / 2 + K (up to kEsTri pairs)
    uint8_t code[12 * kEsSlots];
  };
};

static_assert(sizeof(EsLds) * kEsWaves <= 80 * 1024, "two workgroups of four wavefronts per CU");

// Gradient form: per slot, the derivative of the atom's area by that
// neighbour's position (frame of the rotated neighbours until the end).
struct EsGradLds {
  double gx[kEsSlots], gy[kEsSlots], gz[kEsSlots];
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
  int forms;             // bit 0: pair loop in list form only; bit 1: bucket entries in the scratch only
  double* counters;      // atom pairs scanned, circle pairs classified, arcs tested, coverage tests
  // gradient form: per item kEsSlots rows (neighbour atom, sigma_i dAs_i/dr
  // of it), slot kEsSlots - 1 the atom's own; grow[item] rows in use (0:
  // none)
  int32_t* gatom;
  double* gval;
  int32_t* grow;
};

struct EsScratch {
  double *vt, *vs;   // intersection points (cap)
  uint16_t* vc;      // their two circles: slot1 | slot2 << 8
  double *ang, *srt;  // bucket entries (2 cap): polar angle by entry, sorted within bucket
  uint16_t* ent;     // point of each entry
  uint8_t* entc;     // circle of each entry
  // gradient form, per sorted entry (2 cap): the other circle through its
  // point; the derivatives of the arc that starts there by its own circle's,
  // its start circle's and its end circle's neighbour (9 values), with the
  // start and end circles (255: no exposed arc)
  uint8_t *oth, *gk1, *gk2;
  double* gq;
};

__host__ __device__ inline size_t es_wave_bytes(int cap, bool grad) {
  const size_t c = (size_t)cap;
  return c * 8 * 2 + 2 * c * 8 * 2 + c * 2 + 2 * c * 2 + 2 * c + 256 + (grad ? 2 * c * (9 * 8 + 3) + 256 : 0);
}

__device__ inline EsScratch es_carve(char* p, int cap, bool grad) {
  EsScratch s;
  const size_t c = (size_t)cap;
  s.vt = (double*)p;
  s.vs = s.vt + c;
  s.ang = s.vs + c;
  s.srt = s.ang + 2 * c;
  s.vc = (uint16_t*)(s.srt + 2 * c);
  s.ent = s.vc + c;
  s.entc = (uint8_t*)(s.ent + 2 * c);
  s.oth = s.gk1 = s.gk2 = nullptr;
  s.gq = nullptr;
  if (grad) {  // 8-byte aligned after the byte arrays
    const size_t o = ((size_t)(s.entc + 2 * c - (uint8_t*)p) + 7) / 8 * 8;
    s.gq = (double*)(p + o);
    s.oth = (uint8_t*)(s.gq + 9 * 2 * c);
    s.gk1 = s.oth + 2 * c;
    s.gk2 = s.gk1 + 2 * c;
  }
  return s;
}

// Rank of entry e (angle x) within its bucket [lo, hi): the entries below
// it, ties broken by entry order (esolan.f:642-659 sorts stably).
template <typename T>
__device__ __forceinline__ int es_rank(const T* ang, int lo, int hi, int e, double x) {
  int r = 0, q = lo;
  for (; q + 4 <= hi; q += 4) {
    const double y0 = ang[q], y1 = ang[q + 1], y2 = ang[q + 2], y3 = ang[q + 3];
    r += (y0 < x || (y0 == x && q < e)) ? 1 : 0;
    r += (y1 < x || (y1 == x && q + 1 < e)) ? 1 : 0;
    r += (y2 < x || (y2 == x && q + 2 < e)) ? 1 : 0;
    r += (y3 < x || (y3 == x && q + 3 < e)) ? 1 : 0;
  }
  for (; q < hi; ++q) {
    const double y = ang[q];
    r += (y < x || (y == x && q < e)) ? 1 : 0;
  }
  return r;
}

__device__ __forceinline__ uint64_t lanes_below(int lane) { return (lane == 0) ? 0ull : (~0ull >> (64 - lane)); }

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, kWave));
  return v;
}

// Two circles of the projection (slot K at list position k, slot Lc at a
// higher position): esolan.f:254-449.  es_code gives the outcome,
// es_vertices the two intersection points of a kVert pair (the same
// operations on the same values, so either order gives the same bits).
__device__ __forceinline__ void es_vertices_of(double b1, double c1, double d1, double b2, double c2, double d2,
                                               double v[4]) {
  const double D = 4.0 * ((b1 - b2) * (b2 * d1 - b1 * d2) + (c1 - c2) * (c2 * d1 - c1 * d2) - (d1 - d2) * (d1 - d2)) +
                   (b1 * c2 - b2 * c1) * (b1 * c2 - b2 * c1);
  const double am = 2.0 * ((b1 - b2) * (b1 - b2) + (c1 - c2) * (c1 - c2));
  const double t = -2.0 * (b1 - b2) * (d1 - d2) - (b2 * c1 - b1 * c2) * (c1 - c2);
  const double s = -2.0 * (c1 - c2) * (d1 - d2) + (b2 * c1 - b1 * c2) * (b1 - b2);
  const double pm = sqrt(D), p1 = (c1 - c2) * pm, p2 = (b1 - b2) * pm;
  v[0] = (t + p1) / am;
  v[1] = (s - p2) / am;
  v[2] = (t - p1) / am;
  v[3] = (s + p2) / am;
}
__device__ __forceinline__ void es_vertices(const EsLds& L, int K, int Lc, double v[4]) {
  es_vertices_of(L.nB[K], L.nC[K], L.nD[K], L.nB[Lc], L.nC[Lc], L.nD[Lc], v);
}
// The same from the circles' own coefficients (the quotients formed again,
// the same bits as those of step 3), for when the LDS of the quotients has
// been reused.
__device__ __forceinline__ void es_vertices_raw(const EsLds& L, int K, int Lc, double v[4]) {
  const double a1 = L.A[K], a2 = L.A[Lc];
  es_vertices_of(L.B[K] / a1, L.C[K] / a1, L.D[K] / a1, L.B[Lc] / a2, L.C[Lc] / a2, L.D[Lc] / a2, v);
}
__device__ __forceinline__ int es_code(const EsLds& L, int K, int Lc) {
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
  return code;
}
__device__ __forceinline__ int es_classify(const EsLds& L, int K, int Lc, double v[4]) {
  const int code = es_code(L, K, Lc);
  if (code == kVert) es_vertices(L, K, Lc, v);
  return code;
}

// ---------------------------------------------------------- gradient form
//
// d(term)/d(a, b, c, d) of one circle -> d/d(x, y, z) of its neighbour at
// (x, y, z) relative to the atom (after the atom's rotation): the circle of
// a neighbour of radius r is a = x^2 + y^2 + (z - R)^2 - r^2, b = -8 R^2 x,
// c = -8 R^2 y, d = 4 R^2 (x^2 + y^2 + (z + R)^2 - r^2).
__device__ __forceinline__ void es_to_xyz(const double dq[4], double x, double y, double z, double R, double& gx,
                                          double& gy, double& gz) {
  const double R42 = 4.0 * R * R, m8 = -(R42 + R42);
  gx += dq[0] * (2.0 * x) + dq[1] * m8 + dq[3] * (2.0 * x * R42);
  gy += dq[0] * (2.0 * y) + dq[2] * m8 + dq[3] * (2.0 * y * R42);
  gz += dq[0] * (2.0 * (z - R)) + dq[3] * (2.0 * (z + R) * R42);
}

// The cap a full arc removes (esolan.f:545, a != 0) by the circle's
// coefficients (esolan.f:559-563).
__device__ __forceinline__ void es_full_grad(double R, double a, double b, double c, double d, double dq[4]) {
  const double R2 = R * R, R22 = R2 + R2, R42 = R22 + R22, R42p = 2.0 * R22 * kPi;
  const double den = (R42 * a - d) * (R42 * a - d) + R42 * (b * b + c * c);
  const double div = R22 * R42p / (den * sqrt(den));
  dq[0] = (R42 * (b * b + c * c - 2.0 * a * d) + 2.0 * d * d) * div;
  dq[1] = -b * (d + R42 * a) * div;
  dq[2] = -c * (d + R42 * a) * div;
  dq[3] = (b * b + c * c - 2.0 * a * d + (R42 + R42) * a * a) * div;
}

// One exposed arc (esolan.f:684-862) on circle q = (a, b, c, d) from polar
// angle alp, where circle q1 crosses it, to bet, where circle q2 does: the
// term's derivatives by q's coefficients (directly and through both end
// angles) and by q1's and q2's (through alp and bet).
__device__ void es_arc_grad(double R, double a, double b, double c, double d, double alp, double bet,
                            const double q1[4], const double q2[4], double dq[4], double dq1[4], double dq2[4]) {
  const double R2 = R * R, R42 = 4.0 * R2;
  const double ak = a * a, daba = fabs(a), aia = a < 0.0 ? -1.0 : 1.0, b2c2 = b * b + c * c;
  const double vv = sqrt((R42 * a - d) * (R42 * a - d) + R42 * b2c2), vv1 = 1.0 / vv, vv2 = vv1 * vv1;
  const double wv = sqrt(b2c2 - 4.0 * a * d), bcd = b2c2 - 2.0 * a * d;
  const double xv = (d + R42 * a) * vv1, yv = (bcd + 2.0 * R42 * ak) / daba * vv1, zv = wv / a * vv1;
  double sa, ca, sb, cb, sm, cm, sp, cp;
  sincos(alp, &sa, &ca);
  sincos(bet, &sb, &cb);
  sincos(5e-1 * (bet - alp), &sm, &cm);
  sincos(5e-1 * (bet + alp), &sp, &cp);
  const double cotm = cm / sm;
  const double bcp = b * cp + c * sp, bsp = -b * sp + c * cp;
  const double tt = yv * cotm - zv * bcp / sm;
  const double a1 = q1[0], b1 = q1[1], c1 = q1[2], d1 = q1[3];
  const double a2 = q2[0], b2 = q2[1], c2 = q2[2], d2 = q2[3];
  const double ab1 = a * b1 - b * a1, ac1 = a * c1 - c * a1, ab2 = a * b2 - b * a2, ac2 = a * c2 - c * a2;
  const double am1 = wv * aia * a1 * (ab1 * sa - ac1 * ca), am2 = wv * aia * a2 * (ab2 * sb - ac2 * cb);
  const double u1 = aia * a1 * (ab1 * ca + ac1 * sa) / wv, u2 = aia * a2 * (ab2 * cb + ac2 * sb) / wv;
  double dal[4], dbe[4];
  dal[0] = (b1 * ab1 + c1 * ac1 - 2.0 * d * a1 * a1 - a * (b1 * b1 + c1 * c1 - 4.0 * a1 * d1) - 2.0 * d * u1 +
            wv * aia * a1 * (b1 * ca + c1 * sa)) / am1;
  dal[1] = (-a1 * ab1 + b * a1 * a1 + b * u1 - wv * aia * a1 * a1 * ca) / am1;
  dal[2] = (-a1 * ac1 + c * a1 * a1 + c * u1 - wv * aia * a1 * a1 * sa) / am1;
  dal[3] = (-2.0 * a * a1 * a1 - 2.0 * daba * a1 * (ab1 * ca + ac1 * sa) / wv) / am1;
  dbe[0] = (b2 * ab2 + c2 * ac2 - 2.0 * d * a2 * a2 - a * (b2 * b2 + c2 * c2 - 4.0 * a2 * d2) - 2.0 * d * u2 +
            wv * aia * a2 * (b2 * cb + c2 * sb)) / am2;
  dbe[1] = (-a2 * ab2 + b * a2 * a2 + b * u2 - wv * aia * a2 * a2 * cb) / am2;
  dbe[2] = (-a2 * ac2 + c * a2 * a2 + c * u2 - wv * aia * a2 * a2 * sb) / am2;
  dbe[3] = (-2.0 * a * a2 * a2 - 2.0 * daba * a2 * (ab2 * cb + ac2 * sb) / wv) / am2;
  const double xn = d + R42 * a, yn = b2c2 - 2.0 * a * d + 2.0 * R42 * ak;
  double dv[4];
  dv[1] = R42 * b * vv1;
  dv[2] = R42 * c * vv1;
  dv[3] = (d - R42 * a) * vv1;
  dv[0] = -dv[3] * R42;
  const double sm2 = sm * sm, w = 4.0 + tt * tt, at = kPi - 2.0 * atan(5e-1 * tt);
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const double dx = (m == 0 ? R42 * vv1 : (m == 3 ? vv1 : 0.0)) - xn * dv[m] * vv2;
    double dy, dz;
    if (m == 0) {
      dy = (-2.0 * d + 4.0 * R42 * a) / daba * vv1 - yn * (aia * vv + daba * dv[0]) / ak * vv2;
      dz = -2.0 * d / wv / a * vv1 - wv * (vv + a * dv[0]) / ak * vv2;
    } else {
      const double p = m == 1 ? 2.0 * b : (m == 2 ? 2.0 * c : -2.0 * a);
      dy = p / daba * vv1 - yn * aia * dv[m] / a * vv2;
      dz = (m == 1 ? b / wv / a : (m == 2 ? c / wv / a : -2.0 / wv)) * vv1 - wv * dv[m] / a * vv2;
    }
    const double extra = m == 1 ? cp * 2.0 * sm : (m == 2 ? sp * 2.0 * sm : 0.0);
    const double dt = cotm * dy - 5e-1 * yv * (dbe[m] - dal[m]) / sm2 - (bcp / sm) * dz -
                      5e-1 * zv * (bsp * sm * (dal[m] + dbe[m]) - bcp * cm * (dbe[m] - dal[m]) + extra) / sm2;
    dq[m] = R2 * (aia * (dal[m] - dbe[m]) + at * dx - 4.0 * xv * dt / w);
  }
  double dal1[4], dbe2[4];
  dal1[0] = (-b * ab1 - c * ac1 + wv * wv * a1 + 2.0 * ak * d1 + wv * aia * ((ab1 - b * a1) * ca + (ac1 - c * a1) * sa)) / am1;
  dal1[1] = (a * ab1 - ak * b1 + wv * daba * a1 * ca) / am1;
  dal1[2] = (a * ac1 - ak * c1 + wv * daba * a1 * sa) / am1;
  dal1[3] = (2.0 * ak * a1) / am1;
  dbe2[0] = (-b * ab2 - c * ac2 + wv * wv * a2 + 2.0 * ak * d2 + wv * aia * ((ab2 - b * a2) * cb + (ac2 - c * a2) * sb)) / am2;
  dbe2[1] = (a * ab2 - ak * b2 + wv * daba * a2 * cb) / am2;
  dbe2[2] = (a * ac2 - ak * c2 + wv * daba * a2 * sb) / am2;
  dbe2[3] = (2.0 * ak * a2) / am2;
  const double ta = (yv - zv * (bsp * sm + bcp * cm)) / sm2, tb = (-yv - zv * (bsp * sm - bcp * cm)) / sm2;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    dq1[m] = R2 * (aia * dal1[m] - 4.0 * xv * (5e-1 * dal1[m] * ta) / w);
    dq2[m] = R2 * (-aia * dbe2[m] - 4.0 * xv * (5e-1 * dbe2[m] * tb) / w);
  }
}

// One atom of one conformation.  Returns false if the item was left for the
// rotation launch.
template <bool GRAD>
__device__ void es_item(const EsParams& P, EsLds& L, EsGradLds& G, const EsScratch& X, int item, const double* rot,
                        int lane, double* work) {
  const int c = item / P.stride, i = item - c * P.stride;
  const DevSys& S = P.systems[P.conf_sys[c]];
  if (GRAD && lane == 0) P.grow[item] = 0;
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

  // 4. pairs of circles (esolan.f:251-451).  List positions keep the slots
  // in increasing order (removal is stable) and only the circle at position
  // k or those above it are ever removed, so position k - 1 always holds
  // slot k - 1, and the circles scanned at step k are the live slots above
  // K = k - 1, from the highest down.  A pair's outcome depends on its two
  // circles only.  With up to kEsTri pairs their outcomes are formed first,
  // all lanes at once, and the loop keeps the live slots as a bit mask (at
  // most 68 slots: two words) instead of a list; the intersection points of
  // the pairs it keeps follow in one parallel pass.  More neighbours: the
  // list form below.
  int n = nb, k = nb, iv = 0;
  bool covered = false, zero_as = false, overflow = false;
  const bool tab = nb * (nb - 1) / 2 <= kEsTri && !(P.forms & 1);
  if (tab) {
    for (int t = lane; t < nb * (nb - 1) / 2; t += kWave) {  // t = Lc (Lc - 1) / 2 + K, K < Lc
      int Lc = (int)(0.5f * (1.0f + sqrtf(1.0f + 8.0f * (float)t)));
      Lc -= Lc * (Lc - 1) / 2 > t ? 1 : 0;  // the float root is within one of the integer one
      Lc += (Lc + 1) * Lc / 2 <= t ? 1 : 0;
      L.code[t] = (uint8_t)es_code(L, t - Lc * (Lc - 1) / 2, Lc);
    }
    wave_lds_sync();
    uint64_t alive[2] = {nb >= kWave ? ~0ull : (1ull << nb) - 1ull, nb > kWave ? (1ull << (nb - kWave)) - 1ull : 0ull};
    uint64_t hv[2] = {0ull, 0ull};  // slots with an intersection point
    int pairs = 0;                  // pairs examined (esolan.f's count: up to the ending one)
    // this lane's slot in each word and the outcomes of step K: read one
    // step ahead (they do not depend on the live set), masked when used
    const int tri0 = lane * (lane - 1) / 2, tri1 = (lane + kWave) * (lane + kWave - 1) / 2;
    auto read_codes = [&](int K, int (&c)[2]) {
      c[0] = L.code[(lane > K && lane < nb) ? tri0 + K : 0];
      c[1] = L.code[(lane + kWave < nb) ? tri1 + K : 0];
    };
    int nxt[2] = {kNone, kNone};
    if (nb >= 2) read_codes(nb - 2, nxt);
    for (int K = nb - 2; K >= 0 && !covered; --K) {
      int code[2] = {nxt[0], nxt[1]};
      if (K > 0) read_codes(K - 1, nxt);
      uint64_t act[2];
#pragma unroll
      for (int w = 0; w < 2; ++w) {  // live slots above K
        const int lo = K + 1 - kWave * w;  // first bit above K in word w
        act[w] = lo <= 0 ? alive[w] : (lo >= kWave ? 0ull : alive[w] & ~((1ull << lo) - 1ull));
      }
      for (int w = 1; w >= 0; --w) {  // the higher word first, each from its top bit
        if (!act[w]) continue;
        const uint64_t ev = __ballot(code[w] >= kDropK) & act[w];
        const int f = ev ? 63 - __clzll(ev) : -1;                          // the first event in scan order
        const uint64_t live = f < 0 ? act[w] : act[w] & ~((2ull << f) - 1ull);  // scanned before it (f = 63: none)
        pairs += __popcll(live) + (f >= 0 ? 1 : 0);
        alive[w] &= ~(__ballot(code[w] == kDropL) & live);
        const uint64_t vm = __ballot(code[w] == kVert) & live;
        if (vm) {
          if ((vm >> lane) & 1ull) {  // descending slots: offset by the kept pairs above this lane
            const uint64_t above = lane == 63 ? 0ull : ~((2ull << lane) - 1ull);
            const int o = iv + 2 * __popcll(vm & above);
            const uint16_t pr = (uint16_t)(K | ((kWave * w + lane) << 8));
            if (o + 2 <= P.cap) X.vc[o] = X.vc[o + 1] = pr;
            if (o / 2 < kEsVp) L.vp[o / 2] = pr;
          }
          hv[w] |= vm;
          hv[K >> 6] |= 1ull << (K & 63);
          iv += 2 * __popcll(vm);
        }
        if (f >= 0) {  // the ending outcome, from the masks (no lane exchange)
          const uint64_t fb = 1ull << f;
          if (__ballot(code[w] == kCover) & fb) covered = true;
          else if (__ballot(code[w] == kDropK) & fb) alive[K >> 6] &= ~(1ull << (K & 63));
          else zero_as = true;  // esolan.f:361-364: As = 0, next k
          break;
        }
      }
      if (iv > P.cap) {
        overflow = true;
        break;
      }
    }
    if (lane == 0) work[1] += (double)pairs;
    if (!covered && !overflow) {  // the final list and the slots with points
      const int n0 = __popcll(alive[0]);
      n = n0 + __popcll(alive[1]);
#pragma unroll
      for (int w = 0; w < 2; ++w) {
        const int sl = kWave * w + lane;
        if (sl < nb) {
          L.hasv[sl] = (uint8_t)((hv[w] >> lane) & 1ull);
          if ((alive[w] >> lane) & 1ull) L.lst[(w ? n0 : 0) + __popcll(alive[w] & below)] = (uint8_t)sl;
        }
      }
      wave_lds_sync();
    }
  } else {
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
  }
  if (overflow) {
    if (lane == 0) atomicOr(P.err, kEsErrCap);
    return;
  }
  if (covered) {  // esolan.f:325-326, 428-429
    if (lane == 0) P.asig[item] = 0.0;
    return;
  }
  // with few kept pairs and entries the bucket step below takes them from
  // LDS and forms each point and its angles at once; else the points go to
  // the scratch here
  const bool fastb = !GRAD && tab && iv / 2 <= kEsVp;
  auto points_to_scratch = [&]() {
    wave_lds_sync();
    for (int o = 2 * lane; o < iv; o += 2 * kWave) {
      const int w = X.vc[o];
      double v[4];
      es_vertices(L, w & 255, w >> 8, v);
      X.vt[o] = v[0];
      X.vs[o] = v[1];
      X.vt[o + 1] = v[2];
      X.vs[o + 1] = v[3];
    }
    wave_lds_sync();
  };
  if (tab && !fastb) points_to_scratch();
  // the position of a neighbour relative to the atom, after its rotation
  auto slot_xyz = [&](int s, double& x, double& y, double& z) {
    const int a = L.atom[s];
    x = xyz[3 * a] - xi;
    y = xyz[3 * a + 1] - yi;
    z = xyz[3 * a + 2] - zi;
    if (rot) {
      const double cg = rot[0], sg = rot[1], cf = rot[2], sf = rot[3];
      const double xx = x, yy = y, zz = z;
      x = xx * cg - zz * sg;
      y = -xx * (sf * sg) + yy * cf - zz * (sf * cg);
      z = xx * (cf * sg) + yy * sf + zz * (cf * cg);
    }
  };
  if constexpr (GRAD) {
    for (int s = lane; s < nb; s += kWave) G.gx[s] = G.gy[s] = G.gz[s] = 0.0;
    wave_lds_sync();
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
        if constexpr (GRAD) {  // each slot once in the list: this lane's own
          double dq[4], x, y, z, gx = 0.0, gy = 0.0, gz = 0.0;
          es_full_grad(R, a, b, cc, d, dq);
          slot_xyz(s, x, y, z);
          es_to_xyz(dq, x, y, z, R, gx, gy, gz);
          G.gx[s] = gx;
          G.gy[s] = gy;
          G.gz[s] = gz;
        }
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
  // intersection points bucketed by circle, circles in list order: bucket
  // offsets from the counts
  auto bucket_offsets = [&]() {
    wave_lds_sync();
    int tot = 0;
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
        L.off[s] = tot + incl - m;
        L.cur[s] = tot + incl - m;
      }
      tot += __shfl(incl, kWave - 1, kWave);
    }
    wave_lds_sync();
    return tot;
  };
  for (int s = lane; s < nb; s += kWave) L.cnt[s] = 0;
  int E = 0;
  bool ent_lds = false;  // entries (circle, angle) in LDS
  if (fastb) {
    // both points of a kept pair lie on both its circles
    wave_lds_sync();
    for (int j = lane; j < iv / 2; j += kWave) {
      const int w = L.vp[j], c1 = w & 255, c2 = w >> 8;
      if (L.fin[c1]) atomicAdd(&L.cnt[c1], 2);
      if (L.fin[c2]) atomicAdd(&L.cnt[c2], 2);
    }
    E = bucket_offsets();
    ent_lds = E <= kEsAngLds && !(P.forms & 2);
    if (ent_lds) {  // each pair's points and their polar angles about both circles (esolan.f:626-640)
      for (int j = lane; j < iv / 2; j += kWave) {
        const int w = L.vp[j], c1 = w & 255, c2 = w >> 8;
        double v[4];
        es_vertices_raw(L, c1, c2, v);  // L.ang overlays the quotients
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            const int c = u ? c2 : c1;
            if (!L.fin[c]) continue;
            const double aa = 0.5 / L.A[c];
            const double t = atan2(v[2 * h + 1] + L.C[c] * aa, v[2 * h] + L.B[c] * aa);
            const int q = atomicAdd(&L.cur[c], 1);
            L.ang[q] = t;
            L.entc[q] = (uint8_t)c;
          }
        }
      }
      wave_lds_sync();
    } else {
      points_to_scratch();
      for (int s = lane; s < nb; s += kWave) L.cnt[s] = 0;
    }
  }
  if (!ent_lds) {
    wave_lds_sync();
    for (int v = lane; v < iv; v += kWave) {
      const int w = X.vc[v], c1 = w & 255, c2 = w >> 8;
      if (L.fin[c1]) atomicAdd(&L.cnt[c1], 1);
      if (L.fin[c2]) atomicAdd(&L.cnt[c2], 1);
    }
    E = bucket_offsets();
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
  }
  // polar angles about each circle's centre (esolan.f:626-640); in LDS
  // over the normalised coefficients (dead after the pair loop) when they
  // fit, the sorted copy too when both do, else in the wave's scratch
  const bool ang_lds = E <= kEsAngLds, srt_lds = 2 * E <= kEsAngLds;
  if (!ent_lds) {
    for (int e = lane; e < E; e += kWave) {
      const int s = X.entc[e], v = X.ent[e];
      const double a = L.A[s], aa = 0.5 / a;
      const double t = atan2(X.vs[v] + L.C[s] * aa, X.vt[v] + L.B[s] * aa);
      if (ang_lds)
        L.ang[e] = t;
      else
        X.ang[e] = t;
    }
    wave_lds_sync();
  }
  auto entry_circle = [&](int e) -> int { return ent_lds ? L.entc[e] : X.entc[e]; };
  // ascending within each bucket (esolan.f:642-659)
  for (int e = lane; e < E; e += kWave) {
    const int s = entry_circle(e), lo = L.off[s], hi = lo + L.cnt[s];
    double x;
    int r;
    if (ang_lds) {
      x = L.ang[e];
      r = es_rank(L.ang, lo, hi, e, x);
    } else {
      x = X.ang[e];
      r = es_rank(X.ang, lo, hi, e, x);
    }
    if (srt_lds)
      L.ang[E + lo + r] = x;
    else
      X.srt[lo + r] = x;
    if constexpr (GRAD) {  // the other circle through this point
      const int w = X.vc[X.ent[e]], c1 = w & 255, c2 = w >> 8;
      X.oth[lo + r] = (uint8_t)(c1 == s ? c2 : c1);
    }
  }
  wave_lds_sync();
  // arcs between consecutive points whose midpoint no other circle covers
  double arcs = 0.0;
  for (int p = lane; p < E; p += kWave) {
    const int s = entry_circle(p), lo = L.off[s], hi = lo + L.cnt[s];
    const int pn = (p + 1 < hi) ? p + 1 : lo;
    double alp, bet;
    if (srt_lds) {
      alp = L.ang[E + p];
      bet = L.ang[E + pn];
    } else {
      alp = X.srt[p];
      bet = X.srt[pn];
    }
    if (pn == lo) bet = bet + kPi2;
    if (GRAD) X.gk1[p] = 255;
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
    // the first circle of the list (other than s) that covers the midpoint;
    // four at a time, so their LDS reads overlap instead of chaining behind
    // the test; q ends one past it (n when none does), as a scan one by one
    bool cov = false;
    int q = 0;
    auto covers = [&](int t) {
      return t != s && L.A[t] * q2 + L.B[t] * ap1 + L.C[t] * ap2 + L.D[t] < 0.0;
    };
    for (; q + 4 <= n; q += 4) {
      const int t0 = L.lst[q], t1 = L.lst[q + 1], t2 = L.lst[q + 2], t3 = L.lst[q + 3];
      const bool c0 = covers(t0), c1 = covers(t1), c2 = covers(t2), c3 = covers(t3);
      if (c0 || c1 || c2 || c3) {
        q += c0 ? 1 : (c1 ? 2 : (c2 ? 3 : 4));
        cov = true;
        break;
      }
    }
    for (; q < n && !cov; ++q) cov = covers(L.lst[q]);
    work[2] += 1.0;
    work[3] += (double)q;
    if (cov) continue;
    // (bet + alp) / 2 is prob to the bit (both halvings exact), so its sine
    // and cosine are sp and cp
    const double hm = 5e-1 * (bet - alp);
    double dsbetmalp, dcbetmalp;
    sincos(hm, &dsbetmalp, &dcbetmalp);
    const double dsbetpalp = sp, dcbetpalp = cp;
    const double uv = daba * (bcd + 2.0 * R42 * ak) * dcbetmalp - a * wv * (b * dcbetpalp + cc * dsbetpalp);
    arcs += R2 * (aia * (alp - bet) + (d + R42 * a) * vv1 * (kPi - 2.0 * atan(uv / (2.0 * ak * vv * dsbetmalp))));
    if constexpr (GRAD) {  // esolan.f:684-862
      const int k1 = X.oth[p], k2 = X.oth[(p + 1 < hi) ? p + 1 : lo];
      const double q1[4] = {L.A[k1], L.B[k1], L.C[k1], L.D[k1]}, q2[4] = {L.A[k2], L.B[k2], L.C[k2], L.D[k2]};
      double dq[4], dq1[4], dq2[4], x, y, z;
      es_arc_grad(R, a, b, cc, d, alp, bet, q1, q2, dq, dq1, dq2);
      double g[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      slot_xyz(s, x, y, z);
      es_to_xyz(dq, x, y, z, R, g[0], g[1], g[2]);
      slot_xyz(k1, x, y, z);
      es_to_xyz(dq1, x, y, z, R, g[3], g[4], g[5]);
      slot_xyz(k2, x, y, z);
      es_to_xyz(dq2, x, y, z, R, g[6], g[7], g[8]);
      for (int m = 0; m < 9; ++m) X.gq[9 * (size_t)p + m] = g[m];
      X.gk1[p] = (uint8_t)k1;
      X.gk2[p] = (uint8_t)k2;
    }
  }
  As = As - wave_sum(full) + wave_sum(arcs);
  if (lane == 0) P.asig[item] = As * S.sigma[i];
  if constexpr (GRAD) {
    wave_lds_sync();
    // every slot's arc contributions in bucket order (the circles in list
    // order, each one's arcs by angle: the order esolan.f adds them in)
    for (int t = lane; t < nb; t += kWave) {
      double gx = G.gx[t], gy = G.gy[t], gz = G.gz[t];
      for (int p = 0; p < E; ++p) {
        const int k1 = X.gk1[p];
        if (k1 == 255) continue;
        const double* q = X.gq + 9 * (size_t)p;
        if (X.entc[p] == t) {
          gx += q[0];
          gy += q[1];
          gz += q[2];
        }
        if (k1 == t) {
          gx += q[3];
          gy += q[4];
          gz += q[5];
        }
        if (X.gk2[p] == t) {
          gx += q[6];
          gy += q[7];
          gz += q[8];
        }
      }
      G.gx[t] = gx;
      G.gy[t] = gy;
      G.gz[t] = gz;
    }
    wave_lds_sync();
    // the atom's own: minus the sum over the final list (esolan.f:966-981)
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int p = lane; p < n; p += kWave) {
      const int t = L.lst[p];
      sx += G.gx[t];
      sy += G.gy[t];
      sz += G.gz[t];
    }
    sx = -wave_sum(sx);
    sy = -wave_sum(sy);
    sz = -wave_sum(sz);
    // back to the frame of the coordinates: the final list's rows and the
    // atom's own (esolan.f:983-1001; a dropped circle's row stays as it is)
    auto unrotate = [&](double& x, double& y, double& z) {
      const double cg = rot[0], sg = rot[1], cf = rot[2], sf = rot[3];
      const double xx = x, yy = y, zz = z;
      x = xx * cg - yy * (sf * sg) + zz * (cf * sg);
      y = yy * cf + zz * sf;
      z = -xx * sg - yy * (sf * cg) + zz * (cf * cg);
    };
    if (rot) {
      for (int p = lane; p < n; p += kWave) {
        const int t = L.lst[p];
        double x = G.gx[t], y = G.gy[t], z = G.gz[t];
        unrotate(x, y, z);
        G.gx[t] = x;
        G.gy[t] = y;
        G.gz[t] = z;
      }
      unrotate(sx, sy, sz);
      wave_lds_sync();
    }
    const double sg = S.sigma[i];
    const size_t row = (size_t)item * kEsSlots;
    for (int t = lane; t < nb; t += kWave) {
      P.gatom[row + t] = L.atom[t];
      double* o = P.gval + 3 * (row + t);
      o[0] = sg * G.gx[t];
      o[1] = sg * G.gy[t];
      o[2] = sg * G.gz[t];
    }
    if (lane == 0) {
      P.gatom[row + kEsSlots - 1] = i;
      double* o = P.gval + 3 * (row + kEsSlots - 1);
      o[0] = sg * sx;
      o[1] = sg * sy;
      o[2] = sg * sz;
      P.grow[item] = nb;
    }
  }
}

template <bool GRAD>
__global__ void __launch_bounds__(kEsThreads) k_esolan(EsParams P) {
  extern __shared__ double es_smem[];
  const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  EsLds& L = reinterpret_cast<EsLds*>(es_smem)[wave];
  EsGradLds& G = reinterpret_cast<EsGradLds*>(reinterpret_cast<EsLds*>(es_smem) + kEsWaves)[GRAD ? wave : 0];
  const size_t slot = (size_t)blockIdx.x * kEsWaves + wave;
  const EsScratch X = es_carve(P.scratch + slot * P.wave_bytes, P.cap, GRAD);
  double n[4] = {0.0, 0.0, 0.0, 0.0};
  for (;;) {
    int idx = 0;
    if (lane == 0) idx = atomicAdd(P.next, 1);
    idx = __shfl(idx, 0, kWave);
    if (idx >= P.nitems) break;
    const int item = P.items ? P.items[idx] : idx;
    es_item<GRAD>(P, L, G, X, item, P.items ? P.rot + 4 * (size_t)idx : nullptr, lane, n);
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

// gradan of every atom j of every conformation (esolan.f:1008-1016): the sum
// over the atoms i of the conformation, in atom order per lane and by a wave
// tree across lanes, of row j of item i (the neighbour rows are in
// increasing atom order: a binary search) or of its own row (i = j).  One
// wavefront per (conformation, atom); out at gradan + 3 (stride c + j).
__global__ void k_esolan_grad_sum(const DevSys* __restrict__ systems, const int32_t* __restrict__ conf_sys,
                                  const int32_t* __restrict__ gatom, const double* __restrict__ gval,
                                  const int32_t* __restrict__ grow, int nconf, int stride, double* __restrict__ gradan) {
  const int w = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const int c = w / stride, j = w - c * stride;
  if (c >= nconf) return;
  const int nat = systems[conf_sys[c]].nat;
  double* o = gradan + 3 * ((size_t)stride * c + j);
  if (j >= nat) {
    if (lane < 3) o[lane] = 0.0;
    return;
  }
  double gx = 0.0, gy = 0.0, gz = 0.0;
  for (int i = lane; i < nat; i += kWave) {
    const size_t item = (size_t)stride * c + i, row = item * kEsSlots;
    const int nr = grow[item];
    if (nr == 0) continue;
    int hit = -1;
    if (i == j) {
      hit = kEsSlots - 1;
    } else {
      int lo = 0, hi = nr - 1;
      while (lo <= hi) {
        const int mid = (lo + hi) >> 1, a = gatom[row + mid];
        if (a == j) {
          hit = mid;
          break;
        }
        if (a < j) lo = mid + 1;
        else hi = mid - 1;
      }
    }
    if (hit >= 0) {
      const double* v = gval + 3 * (row + hit);
      gx += v[0];
      gy += v[1];
      gz += v[2];
    }
  }
  gx = wave_sum(gx);
  gy = wave_sum(gy);
  gz = wave_sum(gz);
  if (lane == 0) {
    o[0] = gx;
    o[1] = gy;
    o[2] = gz;
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
  EsBuf flag, small, asig, items, rot, scratch, gatom, gval, grow;
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
static int es_launch(EsolanWork& W, hipStream_t st, EsParams P, int32_t* d_small, double* d_counters, bool grad) {
  for (;;) {
    const size_t wb = (es_wave_bytes(W.cap, grad) + 255) / 256 * 256;
    const int waves = es_grid_waves(P.nitems, wb);
    P.scratch = static_cast<char*>(W.scratch.ensure(wb * (size_t)waves));
    P.wave_bytes = wb;
    P.cap = W.cap;
    HIPCHK(hipMemsetAsync(d_small, 0, 3 * sizeof(int32_t), st));  // err, next, nflag
    P.err = d_small;
    P.next = d_small + 1;
    P.nflag = d_small + 2;
    P.counters = d_counters;
    if (grad) {
      const size_t lds = (sizeof(EsLds) + sizeof(EsGradLds)) * kEsWaves;
      HIPCHK(hipFuncSetAttribute((const void*)k_esolan<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      hipLaunchKernelGGL(k_esolan<true>, dim3(waves / kEsWaves), dim3(kEsThreads), lds, st, P);
    } else {
      const size_t lds = sizeof(EsLds) * kEsWaves;  // above the default 64 KiB: two workgroups per CU
      HIPCHK(hipFuncSetAttribute((const void*)k_esolan<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      hipLaunchKernelGGL(k_esolan<false>, dim3(waves / kEsWaves), dim3(kEsThreads), lds, st, P);
    }
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
                     const double* d_xyz, double* d_out, int out_stride, const std::function<double(int)>& draw,
                     double* d_gradan) {
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
  // SMMP_ESOLAN_FORMS (tests): force the general forms of the pair loop
  // (bit 0) and of the bucket step (bit 1) that large neighbour sets take
  if (const char* env = std::getenv("SMMP_ESOLAN_FORMS")) P.forms = std::atoi(env) & 3;
  P.systems = d_systems;
  P.conf_sys = d_conf_sys;
  P.xyz = d_xyz;
  P.nconf = nconf;
  P.stride = stride;
  P.nitems = nitems;
  P.flag = d_flag;
  P.asig = d_asig;
  const bool grad = d_gradan != nullptr;
  if (grad) {
    static_assert(kEsSlots - 1 == kEsMaxNb, "the atom's own row follows the neighbour rows");
    P.gatom = static_cast<int32_t*>(W.gatom.ensure(sizeof(int32_t) * kEsSlots * (size_t)nitems));
    P.gval = static_cast<double*>(W.gval.ensure(sizeof(double) * 3 * kEsSlots * (size_t)nitems));
    P.grow = static_cast<int32_t*>(W.grow.ensure(sizeof(int32_t) * (size_t)nitems));
  }
  int err = es_launch(W, st, P, d_small, d_counters, grad);
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
    err = es_launch(W, st, P, d_small, d_counters, grad);
    es_check(err & 255);
  }
  const int per = 256 / kWave;
  hipLaunchKernelGGL(k_esolan_sum, dim3((nconf + per - 1) / per), dim3(256), 0, st, d_systems, d_conf_sys, d_asig,
                     nconf, stride, d_out, out_stride);
  HIPCHK(hipGetLastError());
  if (grad) {
    const int64_t waves = (int64_t)nconf * stride;
    hipLaunchKernelGGL(k_esolan_grad_sum, dim3((unsigned)((waves + per - 1) / per)), dim3(256), 0, st, d_systems,
                       d_conf_sys, P.gatom, P.gval, P.grow, nconf, stride, d_gradan);
    HIPCHK(hipGetLastError());
  }
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
