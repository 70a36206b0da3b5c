// ECEPP/3 parameter tables and OONS solvation data: the reference's BLOCK
// DATA (init_energy.f:315-610), setpar (init_energy.f:106-313) and
// init_energy/tessel (init_energy.f:14-103, enysol.f:341-365).
//
// Literal values are the published ECEPP/3 and OONS parameters.  Where the
// reference initialises a double from a default-real (single precision)
// literal, the value is the float literal widened to double; F() marks those
// so the tables round exactly like the reference's.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

#include "state.hpp"

namespace smmp {

static inline double F(float x) { return (double)x; }
// x**6 for an integer exponent as the reference build evaluates it:
// repeated multiplication from the left (matches its tables bit for bit).
static inline double pow6(double x) { return x * x * x * x * x * x; }

void block_data(State& s) {
  s.conv = 332.0;
  s.eps_s = 2.0;
  s.plt = 78.0;
  s.slp = 0.3;
  s.ehm = 362.55;
  // Atomic polarisabilities (*100), effective electron numbers (*100),
  // minimum pair energies (*1000) and optimal distances (*100), types 1..18.
  const float atpl[18] = {42., 42., 42., 42., 42., 42., 93., 151., 115., 59., 84., 59., 93., 93., 93., 220., 42., 93.};
  const float efel[18] = {85., 85., 85., 85., 85., 85., 520., 520., 520., 700., 700., 700., 610., 610., 610., 1480., 85., 520.};
  const float emin[18] = {37., 36., 61., 61., 44., 36., 38., 140., 99., 94., 200., 94., 107., 107., 107., 223., 99., 38.};
  const float rmin[18] = {292., 293., 268., 268., 283., 293., 412., 374., 370., 324., 312., 324., 351., 351., 351., 415., 248., 412.};
  for (int i = 1; i <= mxtyat; ++i) {
    s.atpl(i) = F(atpl[i - 1]);
    s.efel(i) = F(efel[i - 1]);
    s.emin(i) = F(emin[i - 1]);
    s.rmin(i) = F(rmin[i - 1]);
    s.do_s(i) = (i >= 3 && i <= 5) ? 1 : 0;   // H-bond donors: types 3,4,5
    s.ac_s(i) = (i >= 10 && i <= 15) ? 1 : 0; // acceptors: types 10..15
  }
  // H-bond attraction / repulsion (*1000), given column-major as (do,ac).
  const float chb[24] = {2624., 2624., 4610., .0, 4014., 4014., 5783., .0, 2624., 2624., 4610., .0,
                         8244., 8244., 8244., .0, 8244., 8244., 8244., .0, 8244., 8244., 8244., .0};
  const float ahb[24] = {5890., 5890., 11220., .0, 12040., 12040., 16583., .0, 5890., 5890., 11220., .0,
                         32897., 32897., 32897., .0, 32897., 32897., 32897., .0, 32897., 32897., 32897., .0};
  for (int k = 0; k < 24; ++k) {
    s.chb_s.d[k] = F(chb[k]);
    s.ahb_s.d[k] = F(ahb[k]);
  }
  // Torsional classes 1..19: barrier, sign, periodicity.
  const float e0[19] = {20., 0., 0., 2.7f, .6f, 1.8f, 1.8f, 0., 0., 0., 0., 0., 2., 1.5, 3.5, 8., 18., 20., 15.};
  const float sg[19] = {-1., 0., 0., 1., 1., 1., 1., 0., 0., 0., 0., 0., 1., 1., -1., -1., -1., -1., -1.};
  const float rn[19] = {2., 0., 0., 3., 3., 3., 3., 0., 0., 0., 0., 0., 3., 3., 2., 2., 2., 2., 2.};
  for (int i = 1; i <= mxtyto; ++i) {
    s.e0to_s(i) = F(e0[i - 1]);
    s.sgto_s(i) = F(sg[i - 1]);
    s.rnto_s(i) = F(rn[i - 1]);
  }
  // Solvation coefficients and radii, (set 1..9) x (atom type 1..18).
  // Types 1-6 and 17 (hydrogens) are zero.
  const float cs[10][9] = {
      {0.008f, 0.216f, 0.012f, 0.018f, 0.013f, 0.004f, 0.0325f, 0.016f, 1.000f},      // 7
      {0.427f, -0.732f, 0.012f, 0.018f, 0.013f, 0.004f, 0.0325f, 0.016f, 1.000f},     // 8
      {-.008f, -0.678f, 0.012f, 0.018f, 0.013f, 0.004f, 0.0325f, 0.016f, 1.000f},     // 9
      {-.172f, -0.910f, -0.116f, -0.009f, -0.007f, -0.113f, -0.0175f, -0.006f, 0.f},  // 10
      {-.038f, -0.262f, -0.116f, -0.009f, -0.007f, -0.113f, -0.0175f, -0.006f, 0.f},  // 11
      {-.038f, -0.910f, -0.175f, -0.037f, -0.112f, -0.166f, -0.2800f, -0.024f, 0.f},  // 12
      {-.132f, -0.312f, -0.186f, -0.038f, -0.087f, -0.169f, -0.2175f, -0.05f, 0.f},   // 13
      {-.132f, -0.312f, -0.116f, -0.009f, -0.007f, -0.113f, -0.0175f, -0.006f, 0.f},  // 14
      {-.132f, -0.312f, -0.116f, -0.009f, -0.007f, -0.113f, -0.0175f, -0.006f, 0.f},  // 15
      {-.021f, -0.281f, -0.018f, 0.005f, -.0036f, -0.017f, -0.0090f, 0.021f, 0.f},    // 16
  };
  const float rv[10][9] = {
      {2.0f, 2.0f, 1.9f, 1.9f, 1.9f, 1.9f, 1.9f, 1.9f, 2.0f},        // 7
      {1.55f, 1.55f, 1.9f, 1.9f, 1.9f, 1.9f, 1.9f, 1.9f, 1.5f},      // 8
      {1.75f, 1.75f, 1.9f, 1.9f, 1.9f, 1.9f, 1.9f, 1.9f, 1.85f},     // 9
      {1.4f, 1.4f, 1.4f, 1.4f, 1.4f, 1.4f, 1.4f, 1.4f, 1.4f},        // 10
      {1.4f, 1.4f, 1.4f, 1.4f, 1.4f, 1.4f, 1.4f, 1.4f, 1.4f},        // 11
      {1.4f, 1.4f, 1.4f, 1.4f, 1.4f, 1.4f, 1.4f, 1.4f, 1.4f},        // 12
      {1.55f, 1.55f, 1.7f, 1.7f, 1.7f, 1.7f, 1.7f, 1.7f, 1.5f},      // 13
      {1.55f, 1.55f, 1.7f, 1.7f, 1.7f, 1.7f, 1.7f, 1.7f, 1.5f},      // 14
      {1.55f, 1.55f, 1.7f, 1.7f, 1.7f, 1.7f, 1.7f, 1.7f, 1.5f},      // 15
      {2.0f, 2.0f, 1.8f, 1.8f, 1.8f, 1.8f, 1.8f, 1.8f, 1.85f},       // 16
  };
  for (int t = 1; t <= mxtyat; ++t)
    for (int k = 1; k <= mxtysol; ++k) {
      s.coef_sl(k, t) = 0.0;
      s.rad_vdw(k, t) = 0.0;
    }
  for (int t = 7; t <= 16; ++t)
    for (int k = 1; k <= mxtysol; ++k) {
      s.coef_sl(k, t) = F(cs[t - 7][k - 1]);
      s.rad_vdw(k, t) = F(rv[t - 7][k - 1]);
    }
  for (int k = 1; k <= mxtysol; ++k) {  // type 18 (C-delta of Pro) as type 7
    s.coef_sl(k, 18) = s.coef_sl(k, 7);
    s.rad_vdw(k, 18) = s.rad_vdw(k, 7);
  }
}

// Slater-Kirkwood attraction, combined repulsion, like-pair overrides, 1-4
// halving, H-bond type matrix, torsion tables.  Like the reference it
// rescales its inputs in place, so a second call changes the tables
// (init_energy.f:164-171); init_energy() restores the inputs first.
void setpar(State& s) {
  s.tesgrd = 0;
  if (s.flex) throw Error("setpar> the FLEX force field is not part of this engine (ECEPP/2 and ECEPP/3 only)");
  s.conv = s.conv / s.eps_s;
  for (int i = 1; i <= mxtyat; ++i) {
    s.atpl(i) = s.atpl(i) / 100;
    s.efel(i) = s.efel(i) / 100;
    s.emin(i) = s.emin(i) / 1000;
    s.rmin(i) = s.rmin(i) / 100;
  }
  for (int i = 1; i <= mxtyat; ++i) {
    double ri = s.rmin(i), ai = s.atpl(i);
    double aei = std::sqrt(ai / s.efel(i));
    double aic = ai * s.ehm;
    s.cij(i, i) = aic * ai / (aei + aei);
    double a = .5 * s.cij(i, i) * pow6(ri);
    s.aij(i, i) = a;
    s.a14(i, i) = .5 * a;
    for (int j = i + 1; j <= mxtyat; ++j) {
      double aj = s.atpl(j);
      double c = aic * aj / (aei + std::sqrt(aj / s.efel(j)));
      s.cij(i, j) = c;
      s.cij(j, i) = c;
      double rij = .5 * (ri + s.rmin(j));
      a = .5 * c * pow6(rij);
      s.aij(i, j) = a;
      s.aij(j, i) = a;
      a = .5 * a;
      s.a14(i, j) = a;
      s.a14(j, i) = a;
    }
  }
  // Like-pair parameters fixed by the ECEPP/3 papers.
  const double cd[18] = {45.5, 45.5, 45.5, 45.5, 45.5, 45.5, 370.5, 766.6, 509.5,
                         217.2, 369.0, 217.2, 401.3, 401.3, 401.3, 2274.4, 45.5, 370.5};
  const double ad[18] = {14090.0, 14380.0, 8420.0, 8420.0, 11680.0, 14380.0, 906100.0, 1049000.0, 653600.0,
                         125600.0, 170200.0, 125600.0, 375200.0, 375200.0, 375200.0, 5809000.0, 5340.0, 909000.0};
  for (int i = 1; i <= mxtyat; ++i) {
    s.cij(i, i) = cd[i - 1];
    s.aij(i, i) = ad[i - 1];
    s.a14(i, i) = .5 * s.aij(i, i);
  }
  for (int i = 1; i <= mxtyto; ++i) {
    s.e0to(i) = s.e0to_s(i) / 2.;
    s.sgto(i) = s.sgto_s(i);
    s.rnto(i) = s.rnto_s(i);
  }
  for (int i = 1; i <= mxtyat; ++i)
    for (int j = 1; j <= mxtyat; ++j) {
      s.ihbty(i, j) = 0;
      s.chb(i, j) = 0.;
      s.ahb(i, j) = 0.;
    }
  int iac = 0, ido = 0;
  for (int i = 1; i <= mxtyat; ++i) {
    if (s.do_s(i)) {
      ++ido;
      int jac = 0;
      for (int j = 1; j <= i - 1; ++j)
        if (s.ac_s(j)) {
          ++jac;
          s.ihbty(i, j) = 1;
          s.ihbty(j, i) = -1;
          s.chb(i, j) = s.chb(j, i) = s.chb_s(ido, jac);
          s.ahb(i, j) = s.ahb(j, i) = s.ahb_s(ido, jac);
        }
    } else if (s.ac_s(i)) {
      ++iac;
      int jdo = 0;
      for (int j = 1; j <= i - 1; ++j)
        if (s.do_s(j)) {
          ++jdo;
          s.ihbty(i, j) = -1;
          s.ihbty(j, i) = 1;
          s.chb(i, j) = s.chb(j, i) = s.chb_s(jdo, iac);
          s.ahb(i, j) = s.ahb(j, i) = s.ahb_s(jdo, iac);
        }
    }
  }
  for (int i = 1; i <= mxtyto; ++i) s.esnto(i) = s.e0to(i) * s.sgto(i) * s.rnto(i);
}

// Reads the tessellation points: comment lines start with '!', then the
// count (i5) and one '3f20.10' line per point (enysol.f:341-365).
static void tessel(State& s, const std::string& path) {
  std::ifstream in(path);
  if (!in) throw Error("init_energy> cannot open the file with surface points: " + path);
  std::string line;
  do {
    if (!std::getline(in, line)) throw Error("tessel> unexpected end of " + path);
  } while (!line.empty() && line[0] == '!');
  int n = std::atoi(line.substr(0, 5).c_str());
  if (n <= 0 || n > mxpoint) throw Error("tessel> bad number of surface points in " + path);
  s.npnt = n;
  for (int i = 1; i <= n; ++i) {
    if (!std::getline(in, line)) throw Error("tessel> unexpected end of " + path);
    line.resize(60, ' ');
    for (int k = 0; k < 3; ++k) s.spoint(i, k + 1) = std::strtod(line.substr(20 * k, 20).c_str(), nullptr);
  }
}

void init_energy(State& s, const std::string& libdir) {
  if (s.ientyp == 1) s.flex = 1; else s.flex = 0;
  if (s.ientyp != 0) throw Error("init_energy> only ECEPP (ientyp=0) is implemented by this engine");
  std::string ld = libdir;
  while (!ld.empty() && ld.back() == ' ') ld.pop_back();
  s.libdir = ld;
  s.reslib = ld + (s.sh2 ? "lib.sh2" : "lib.sh3");
  // Idempotent: restore the unscaled inputs before setpar rescales them.
  int32_t flex = s.flex, sh2 = s.sh2, epsd = s.epsd, ientyp = s.ientyp;
  block_data(s);
  s.flex = flex; s.sh2 = sh2; s.epsd = epsd; s.ientyp = ientyp;
  setpar(s);
  int its = std::abs(s.itysol);
  if (its > 0 && its <= mxtysol) {
    tessel(s, ld + "tes.dat");
  } else if (s.itysol != 0) {
    throw Error("init_energy> undefined solvent type");
  }
}

}  // namespace smmp
