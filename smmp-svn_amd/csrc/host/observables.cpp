// Observables of the current coordinates that the samplers' outputs carry.
#include <cmath>

#include "state.hpp"

namespace smmp {

// rgyr.f:13-140 (the outputs rgy and ee).  Like the reference, the sums run
// over nat + 1 atoms from the chain's first atom (rgyr.f:72) while the
// means divide by nat; ee joins the C-alpha atoms (ind_alf, c_alfa of
// contacts.f:65-97: the last atom of the residue whose name starts with
// "ca", 0 if none) of the first and the last residue.
void rgyr(const State& s, int nml, double* rgy, double* ee) {
  const int nml1 = nml == 0 ? 1 : nml, nml2 = nml == 0 ? s.ntlml : nml;
  const int a1 = s.iatrs1(s.irsml1(nml1));
  const int nat = s.iatrs2(s.irsml2(nml2)) - a1 + 1;
  *rgy = 0.0;
  *ee = 0.0;
  if (nat <= 0) return;
  const double dn = (double)nat;
  double dx = 0.0, dy = 0.0, dz = 0.0, d2 = 0.0;
  for (int i = a1; i <= a1 + nat && i <= mxat; ++i) {
    const double xi = s.xat(i), yi = s.yat(i), zi = s.zat(i);
    dx = dx + xi;
    dy = dy + yi;
    dz = dz + zi;
    d2 = d2 + xi * xi + yi * yi + zi * zi;
  }
  dx = dx / dn;
  dy = dy / dn;
  dz = dz / dn;
  d2 = d2 / dn;
  const double rg2 = d2 - (dx * dx + dy * dy + dz * dz);
  *rgy = rg2 > 0.0 ? std::sqrt(rg2) : 0.0;
  auto ind_alf = [&](int res) {
    int ia = 0;
    for (int a = s.iatrs1(res); a <= s.iatrs2(res); ++a)
      if (s.nmat.get(a).compare(0, 2, "ca") == 0) ia = a;
    return ia;
  };
  const int i1 = ind_alf(s.irsml1(nml1)), i2 = ind_alf(s.irsml2(nml2));
  const double ex = s.xat(i2) - s.xat(i1), ey = s.yat(i2) - s.yat(i1), ez = s.zat(i2) - s.zat(i1);
  *ee = std::sqrt(ex * ex + ey * ey + ez * ez);
}

// helix.f:13-77: residues whose (phi, psi) lie within 30 degrees of
// (-70, -37) (helical) or (-150, 150) (sheet-like), and the segments they
// form, over the free variables; psi is the free variable after phi.
void helix(const State& s, int32_t out[4]) {
  constexpr double hlim = 30.0, philim = -70.0, psilim = -37.0;
  constexpr double hlim2 = 30.0, philim2 = -150.0, psilim2 = 150.0;
  int nhel = 0, mhel = 0, nbet = 0, mbet = 0;
  bool lhel = false, lbet = false;
  for (int i = 1; i <= s.nvr; ++i) {
    const int iv = s.idvr(i);
    if (s.nmvr.get(iv) != "phi") continue;
    const double xphi = s.vlvr(iv) * crd;
    const double xpsi = s.vlvr(s.idvr(i + 1)) * crd;
    if (std::fabs(xphi - philim) <= hlim) {
      lbet = false;
      if (std::fabs(xpsi - psilim) <= hlim) {
        nhel = nhel + 1;
        if (!lhel) mhel = mhel + 1;
        lhel = true;
      } else {
        lhel = false;
      }
    } else if (std::fabs(xphi - philim2) <= hlim2) {
      lhel = false;
      if (std::fabs(xpsi - psilim2) <= hlim2) {
        nbet = nbet + 1;
        if (!lbet) mbet = mbet + 1;
        lbet = true;
      } else {
        lbet = false;
      }
    } else {
      lhel = false;
      lbet = false;
    }
  }
  out[0] = nhel;
  out[1] = mhel;
  out[2] = nbet;
  out[3] = mbet;
}

}  // namespace smmp
