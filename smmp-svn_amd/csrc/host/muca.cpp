// Multicanonical weight iteration (Berg's recursion) and the final weights,
// mulcan_par_mod.f90:104-155 and :168-176.  Host code: one update every
// 'nup' sweeps touches kmax-kmin+1 bins.
#include <cmath>
#include <cstdint>

#include "../../../include/smmp_mi355x.h"

namespace {

// The reference converts counts with float() and takes REAL logarithms
// (mulcan_par_mod.f90:116-122); the same single-precision steps here.
inline double fcount_product(int32_t a, int32_t b) { return (double)((float)a * (float)b); }
inline double flog(int32_t n) { return (double)std::log((float)n); }

}  // namespace

extern "C" int smmp_muca_update(int kmin, int kmax, double ebin, double beta, int32_t* ihist, double* xhist,
                                double* g1, double* b, double* alpha, double* ent) {
  if (kmax - kmin < 2 || !(ebin > 0.0) || !ihist || !xhist || !g1 || !b || !alpha) return SMMP_EINVAL;
  auto at = [kmin](auto* p, int k) -> auto& { return p[k - kmin]; };
  at(xhist, kmax) += at(ihist, kmax);
  at(xhist, kmax - 1) += at(ihist, kmax - 1);
  for (int i = kmax - 2; i >= kmin; --i) {
    at(xhist, i) += at(ihist, i);
    const int32_t hi = at(ihist, i), hn = at(ihist, i + 1);
    if (hi > 1 && hn > 1) {
      double g0 = fcount_product(hn, hi);
      g0 = g0 / (double)(float)(hi + hn);
      at(g1, i) += g0;
      g0 = g0 / at(g1, i);
      at(b, i) = at(b, i) + g0 * (flog(hn) - flog(hi)) / ebin;
    } else if (at(xhist, i) < 2.0) {
      at(b, i) = std::fmax(0.0, at(b, i + 1));
    }
  }
  at(b, kmax) = at(b, kmax - 2);
  at(b, kmax - 1) = at(b, kmax - 2);
  at(alpha, kmax) = 0.0;
  if (ent) at(ent, kmax) = (beta + at(b, kmax)) * (double)(float)kmax * ebin + at(alpha, kmax);
  at(ihist, kmax) = 0;
  for (int i = kmax - 1; i >= kmin; --i) {
    at(ihist, i) = 0;
    at(alpha, i) = at(alpha, i + 1) + (at(b, i + 1) - at(b, i)) * (i + 0.5) * ebin;
    if (ent) at(ent, i) = (beta + at(b, i)) * (double)(float)i * ebin + at(alpha, i);
  }
  return SMMP_OK;
}

extern "C" int smmp_muca_final(int kmin, int kmax, double ebin, double beta, double* b, double* alpha) {
  if (kmax < kmin || !(ebin > 0.0) || !b || !alpha) return SMMP_EINVAL;
  auto at = [kmin](double* p, int k) -> double& { return p[k - kmin]; };
  at(alpha, kmax) = 0.0;
  at(b, kmax) += beta;
  for (int i = kmax - 1; i >= kmin; --i) {
    at(b, i) += beta;
    at(alpha, i) = at(alpha, i + 1) + (at(b, i + 1) - at(b, i)) * (i + 0.5) * ebin;
  }
  return SMMP_OK;
}
