// Molecule construction for the standard-geometry model: sequence reader,
// residue library, assembly of atoms/bonds/variables, Eyring chain build,
// terminal groups, moving sets and interaction lists.  Each function restates
// the reference routine named beside it; indices are the reference's 1-based
// atom/variable/residue numbers so the arrays are interchangeable with it.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>

#include "state.hpp"

namespace smmp {

// ---------------------------------------------------------------- helpers

static std::string lower(std::string s) {
  for (auto& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}
static std::string pad(const std::string& s, size_t n) {
  std::string r = s.substr(0, std::min(s.size(), n));
  r.resize(n, ' ');
  return r;
}
static std::string rtrim(std::string s) {
  while (!s.empty() && (s.back() == ' ' || s.back() == '\r' || s.back() == '\n' || s.back() == '\t')) s.pop_back();
  return s;
}

// Fortran 'Iw' input of a fixed field (blanks ignored, empty = 0).
static int read_int_field(const std::string& f) {
  std::string t;
  for (char c : f)
    if (c != ' ') t += c;
  if (t.empty()) return 0;
  return std::atoi(t.c_str());
}
// Fortran 'Dw.d' input: blanks ignored; without a decimal point the last d
// digits are the fraction.
static double read_real_field(const std::string& f, int d) {
  std::string t;
  for (char c : f)
    if (c != ' ') t += c;
  if (t.empty()) return 0.0;
  bool has_point = t.find('.') != std::string::npos;
  bool has_exp = t.find_first_of("eEdD") != std::string::npos;
  for (auto& c : t)
    if (c == 'd' || c == 'D') c = 'e';
  double v = std::strtod(t.c_str(), nullptr);
  if (!has_point && !has_exp) v = v / std::pow(10.0, d);
  return v;
}

double difang(double a1, double a2) {
  double d = std::fmod(a2 - a1, pi2);
  if (std::fabs(d) <= pi) return d;
  return d - std::copysign(pi2, d);
}
double addang(double a1, double a2) {
  double d = std::fmod(a1 + a2, pi2);
  if (std::fabs(d) <= pi) return d;
  return d - std::copysign(pi2, d);
}

// ------------------------------------------------------- residue library

// One residue of the library (the /help_*/ arrays of redres, getmol.f:253-459).
struct LibResidue {
  int nat = 0, nxt = 0, nvr = 0;
  std::vector<std::string> nm;            // atom names (4 chars)
  std::vector<int> ity, iow, iyow, nbd;   // 1-based by atom
  std::vector<int> ibd, iybd;             // [3*(atom-1) + (k-1)]
  std::vector<double> bl, ba, to, cg;     // ba/to in radians
  std::vector<int> vty, vcl, vat;         // variables: type, class, atom
  std::vector<std::string> vnm;
};

// Reads residue 'res' from the library file: the first '#name' header that
// matches, then 'nat' atom lines in the format
// (a4,3(1x,i2,a1,a3,d9.3),d7.4,i4,4i4) (getmol.f:325-328, redres).
static LibResidue redres(const std::string& reslib, const std::string& res_in) {
  std::ifstream in(reslib);
  if (!in) throw Error("getmol> ERROR opening library of residues: " + reslib);
  const std::string resl = pad(lower(rtrim(res_in)), 4);
  std::string line;
  while (std::getline(in, line)) {
    line = rtrim(line);
    if ((int)line.size() < 13 || line[0] != '#' || pad(line.substr(1), 4) != resl) continue;
    LibResidue r;
    r.nat = read_int_field(line.substr(5, 4));
    r.nxt = read_int_field(line.substr(9, 4));
    if (r.nat > mxath) throw Error("redres> too many atoms in residue " + resl);
    r.nm.resize(r.nat + 1); r.ity.resize(r.nat + 1); r.iow.resize(r.nat + 1); r.iyow.resize(r.nat + 1);
    r.nbd.resize(r.nat + 1); r.ibd.assign(3 * (r.nat + 1), 0); r.iybd.assign(3 * (r.nat + 1), 1);
    r.bl.resize(r.nat + 1); r.ba.resize(r.nat + 1); r.to.resize(r.nat + 1); r.cg.resize(r.nat + 1);
    for (int i = 1; i <= r.nat; ++i) {
      if (!std::getline(in, line)) throw Error("redres> ERROR reading " + reslib);
      line = rtrim(line);
      line.resize(80, ' ');
      r.nm[i] = line.substr(0, 4);
      int icl[3];
      char fix[3];
      std::string nm3[3];
      double val[3];
      for (int g = 0; g < 3; ++g) {
        int b = 4 + 16 * g;
        icl[g] = read_int_field(line.substr(b + 1, 2));
        fix[g] = line[b + 3];
        nm3[g] = line.substr(b + 4, 3);
        val[g] = read_real_field(line.substr(b + 7, 9), 3);
      }
      r.bl[i] = val[0];
      double ba = val[1], to = val[2];
      r.cg[i] = read_real_field(line.substr(52, 7), 4);
      int ity = read_int_field(line.substr(59, 4));
      int iow = read_int_field(line.substr(63, 4));
      int ibd[3];
      for (int k = 0; k < 3; ++k) ibd[k] = read_int_field(line.substr(67 + 4 * k, 4));
      if (ity <= 0 || ity > mxtyat) throw Error("redres> unknown type for atom " + r.nm[i] + " in residue " + resl);
      r.ity[i] = ity;
      int jow = std::abs(iow);
      bool iexcp = (resl.substr(0, 3) == "ace" && i == 1);
      if (!iexcp && i <= jow) throw Error("redres> atom " + r.nm[i] + " of " + resl + " placed before its predecessor");
      r.iow[i] = jow;
      r.iyow[i] = iow < 0 ? -1 : 1;
      int ib1 = std::abs(ibd[0]), ib2 = std::abs(ibd[1]), ib3 = std::abs(ibd[2]);
      auto bad = [&]() { return Error("redres> Incorrect order of bonds for atom " + r.nm[i] + " of " + resl); };
      if (ib1 == i || ib2 == i || ib3 == i) throw bad();
      if (ib1 == 0) {
        if (ib2 != 0 || ib3 != 0) throw bad();
        r.nbd[i] = 0;
      } else {
        if (ib1 == jow) throw bad();
        if (ib2 == 0) {
          if (ib3 != 0) throw bad();
          r.nbd[i] = 1;
        } else {
          if (ib2 == jow || ib2 == ib1 || (ib2 > i && ib2 < ib1)) throw bad();
          if (ib3 == 0) {
            r.nbd[i] = 2;
          } else {
            if (ib3 == jow || ib3 == ib1 || ib3 == ib2 || (ib3 > i && (ib3 < ib1 || ib3 < ib2))) throw bad();
            r.nbd[i] = 3;
          }
        }
      }
      for (int k = 0; k < 3; ++k) {
        r.ibd[3 * i + k] = std::abs(ibd[k]);
        r.iybd[3 * i + k] = ibd[k] < 0 ? -1 : 1;
      }
      r.ba[i] = ba * cdr;
      r.to[i] = to * cdr;
      for (int j = 0; j < 3; ++j) {
        if (fix[j] == ' ') continue;
        ++r.nvr;
        if (r.nvr > mxvrh) throw Error("redres> number of variables > mxvrh");
        int ic = icl[j];
        // Only torsions carry potential classes (mxtybl = mxtyba = 0).
        if (ic <= 0 || (j == 2 && ic > mxtyto) || (j == 1) || (j == 0))
          throw Error("redres> unknown class for variable " + nm3[j] + " in residue " + resl);
        r.vty.push_back(j + 1);
        r.vcl.push_back(ic);
        r.vat.push_back(i);
        r.vnm.push_back(nm3[j]);
      }
    }
    return r;
  }
  throw Error("redres> residue >" + resl + "< NOT FOUND in " + reslib);
}

// Library cache keyed by (file, residue): the file is read once per residue.
static const LibResidue& library(const State& s, const std::string& res) {
  static std::map<std::string, LibResidue> cache;
  std::string key = s.reslib + "|" + pad(lower(rtrim(res)), 4);
  auto it = cache.find(key);
  if (it == cache.end()) it = cache.emplace(key, redres(s.reslib, res)).first;
  return it->second;
}

// ------------------------------------------------------------ sequence

// redseq.f:13-213: molecules separated by '#' lines (optional name after it),
// residue names of 1-4 characters separated by blanks.
static void redseq(State& s, const std::string& seqfil) {
  std::ifstream in(seqfil);
  if (!in) throw Error("redseq> ERROR opening sequence file: " + seqfil);
  int nrs = s.ntlml > 0 ? s.irsml2(s.ntlml) : 0;
  std::string line;
  while (std::getline(in, line)) {
    line = rtrim(line);
    if (line.empty()) continue;
    size_t ic = line.find('#');
    if (ic != std::string::npos) {
      if (s.ntlml > 0) {
        s.irsml2(s.ntlml) = nrs;
        if (nrs - s.irsml1(s.ntlml) + 1 == 0) s.ntlml -= 1;
      }
      s.ntlml += 1;
      if (s.ntlml > mxml) throw Error("redseq> NUMBER of molecules > mxml in " + seqfil);
      s.irsml1(s.ntlml) = nrs + 1;
      std::string name = line.substr(ic + 1);
      size_t b = name.find_first_not_of(' ');
      char buf[8];
      if (b != std::string::npos) {
        name = rtrim(name.substr(b));
        s.nmml.set(s.ntlml, name.substr(0, 80));
      } else {
        std::snprintf(buf, sizeof buf, "%4d", s.ntlml);
        s.nmml.set(s.ntlml, buf);
      }
      continue;
    }
    std::istringstream ss(line);
    std::string tok;
    bool any = false;
    while (ss >> tok) {
      if (!any && s.ntlml == 0) {
        s.ntlml = 1;
        s.irsml1(1) = nrs + 1;  // the reference leaves this unset without a '#' line
        s.nmml.set(1, "   1");
      }
      any = true;
      if (tok.size() > 4) throw Error("redseq> residue name >" + tok + "< too long in " + seqfil);
      ++nrs;
      if (nrs > mxrs) throw Error("redseq> NUMBER of residues > mxrs in " + seqfil);
      s.seq.set(nrs, tok);
    }
  }
  if (nrs == 0) throw Error("redseq> no residues found in " + seqfil);
  for (int i = 1; i <= s.ntlml; ++i) {
    int ifirs = s.irsml1(i);
    if (!s.flex && lower(s.seq.get(ifirs)).substr(0, 3) == "pro") s.seq.set(ifirs, "pron");
    if (i == s.ntlml) {
      if (nrs - ifirs + 1 == 0) {
        s.ntlml -= 1;
        if (s.ntlml == 0) throw Error("redseq> no residues found in " + seqfil);
        return;
      }
      s.irsml2(i) = nrs;
    }
  }
}

// ------------------------------------------------------------- assembly

// getmol.f:14-251: atoms, bonds and variables of molecule nml from the
// library, residues joined through the 'next' atom of the previous residue.
static void getmol(State& s, int nml) {
  int ntlat, ntlvr;
  if (nml == 1) {
    ntlat = 0;
    ntlvr = 0;
  } else {
    int i = s.irsml2(nml - 1);
    ntlat = s.iatrs2(i);
    ntlvr = s.ivrrs1(i) + s.nvrrs(i) - 1;
  }
  s.ivrml1(nml) = ntlvr + 1;
  int ilars = s.irsml2(nml), ifirs = s.irsml1(nml);
  for (int nrs = ifirs; nrs <= ilars; ++nrs) {
    std::string res = lower(s.seq.get(nrs));
    if (res.substr(0, 3) == "nme" && nrs != ilars) throw Error("getmol> nme allowed at C-terminus only");
    if (res.substr(0, 3) == "ace" && nrs != ifirs) throw Error("getmol> ace allowed at N-terminus only");
    LibResidue r = library(s, res);  // copy: the first-atom bond edit below is local
    int nat = r.nat;
    s.nvr = r.nvr;  // redres reports through the global nvr (getmol.f:66)
    if (nat + ntlat > mxat) throw Error("getmol> too many atoms (mxat)");
    if (r.nvr + ntlvr > mxvr) throw Error("getmol> too many variables (mxvr)");
    for (int i = 1; i <= nat; ++i) {
      int n = i + ntlat;
      s.nmat.set(n, r.nm[i]);
      s.ityat(n) = r.ity[i];
      s.cgat(n) = r.cg[i];
      s.blat(n) = r.bl[i];
      double ba = r.ba[i];
      s.baat(n) = ba;
      s.csbaat(n) = std::cos(ba);
      s.snbaat(n) = std::sin(ba);
      double to = r.to[i];
      s.toat(n) = to;
      s.cstoat(n) = std::cos(to);
      s.sntoat(n) = std::sin(to);
      int iow = r.iow[i];
      if (iow == 0) {
        if (nrs == ifirs) {
          // first atom of the molecule: iowat is its first bond forward
          s.iowat(n) = r.ibd[3 * i + 0] + ntlat;
          s.iyowat(n) = r.iybd[3 * i + 0];
          for (int j = 0; j < mxbd - 1; ++j) {
            r.ibd[3 * i + j] = r.ibd[3 * i + j + 1];
            r.iybd[3 * i + j] = r.iybd[3 * i + j + 1];
          }
          r.ibd[3 * i + mxbd - 1] = 0;
          r.iybd[3 * i + mxbd - 1] = 1;
          r.nbd[i] -= 1;
        } else {
          int nh = s.ixatrs(nrs - 1);
          s.iowat(n) = nh;
          s.iyowat(n) = 1;
          int nbd = s.nbdat(nh);
          if (nbd == mxbd) throw Error("getmol> need more bonds to connect residues");
          for (int j = 1; j <= nbd; ++j) {
            int nj = s.ibdat(j, nh);
            double t = s.toat(nj);
            if (t == 0.0) throw Error("getmol> DIHEDRAL for atom " + s.nmat.get(nj) + " should be a PHASE angle");
            t = t + to;
            if (std::fabs(t) > pi) t = t - std::copysign(pi2, t);
            s.toat(nj) = t;
            s.cstoat(nj) = std::cos(t);
            s.sntoat(nj) = std::sin(t);
          }
          int nbd1 = nbd + 1;
          s.ibdat(nbd1, nh) = n;
          s.iybdat(nbd1, nh) = 1;
          s.nbdat(nh) = nbd1;
        }
      } else {
        s.iowat(n) = ntlat + iow;
        s.iyowat(n) = r.iyow[i];
      }
      s.nbdat(n) = r.nbd[i];
      for (int j = 1; j <= mxbd; ++j) {
        int ibd = r.ibd[3 * i + j - 1];
        if (ibd != 0) {
          s.ibdat(j, n) = ibd + ntlat;
          s.iybdat(j, n) = r.iybd[3 * i + j - 1];
        } else {
          s.ibdat(j, n) = 0;
          s.iybdat(j, n) = 1;
        }
      }
    }
    s.ivrrs1(nrs) = ntlvr + 1;
    int mvr = 0;
    for (int i = 0; i < r.nvr; ++i) {
      if (nrs == ifirs) {
        int iat = r.vat[i];
        // no variables for the first atom, no torsions of atoms bound to it
        if (iat == 1 || (r.iow[iat] == 1 && r.vty[i] == 3)) continue;
      }
      ++mvr;
      ++ntlvr;
      s.ityvr(ntlvr) = r.vty[i];
      s.iclvr(ntlvr) = r.vcl[i];
      s.iatvr(ntlvr) = r.vat[i] + ntlat;
      s.nmvr.set(ntlvr, r.vnm[i]);
    }
    s.nvrrs(nrs) = mvr;
    s.iatrs1(nrs) = ntlat + 1;
    s.ixatrs(nrs) = ntlat + r.nxt;
    ntlat += nat;
    s.iatrs2(nrs) = ntlat;
  }
  s.nvrml(nml) = (nml == 1) ? ntlvr : ntlvr - s.ivrml1(nml) + 1;
}

// ----------------------------------------------------------- geometry

// bldmol.f:175-242: atom i from the frame (x-axis, z-axis, position) of ia.
static void eyring(State& s, int i, int ia) {
  double ct = s.cstoat(i), st = s.sntoat(i), ca = s.csbaat(i), sa = s.snbaat(i), bl = s.blat(i);
  double x1 = s.xtoat(ia), x2 = s.ytoat(ia), x3 = s.ztoat(ia);
  double z1 = s.xbaat(ia), z2 = s.ybaat(ia), z3 = s.zbaat(ia);
  double y1 = z2 * x3 - z3 * x2, y2 = z3 * x1 - z1 * x3, y3 = z1 * x2 - z2 * x1;
  double h2 = -sa * ct, h3 = -sa * st;
  x1 = -ca * x1 + h2 * y1 + h3 * z1;
  x2 = -ca * x2 + h2 * y2 + h3 * z2;
  x3 = -ca * x3 + h2 * y3 + h3 * z3;
  double dx = 1.0 / std::sqrt(x1 * x1 + x2 * x2 + x3 * x3);
  x1 *= dx; x2 *= dx; x3 *= dx;
  s.xtoat(i) = x1; s.ytoat(i) = x2; s.ztoat(i) = x3;
  z1 = -st * y1 + ct * z1;
  z2 = -st * y2 + ct * z2;
  z3 = -st * y3 + ct * z3;
  double dz = 1.0 / std::sqrt(z1 * z1 + z2 * z2 + z3 * z3);
  s.xbaat(i) = z1 * dz; s.ybaat(i) = z2 * dz; s.zbaat(i) = z3 * dz;
  s.xat(i) = s.xat(ia) + x1 * bl;
  s.yat(i) = s.yat(ia) + x2 * bl;
  s.zat(i) = s.zat(ia) + x3 * bl;
}

// bldmol.f:244-300: right-handed axes from three atom positions.
static void setsys(const State& s, int i1, int i2, int i3, double* x, double* y, double* z) {
  double h1 = s.xat(i2), h2 = s.yat(i2), h3 = s.zat(i2);
  double x1 = h1 - s.xat(i1), x2 = h2 - s.yat(i1), x3 = h3 - s.zat(i1);
  double y1 = h1 - s.xat(i3), y2 = h2 - s.yat(i3), y3 = h3 - s.zat(i3);
  double z1 = x2 * y3 - x3 * y2, z2 = x3 * y1 - x1 * y3, z3 = x1 * y2 - x2 * y1;
  double dx = 1.0 / std::sqrt(x1 * x1 + x2 * x2 + x3 * x3);
  x1 *= dx; x2 *= dx; x3 *= dx;
  double dz = 1.0 / std::sqrt(z1 * z1 + z2 * z2 + z3 * z3);
  z1 *= dz; z2 *= dz; z3 *= dz;
  x[0] = x1; x[1] = x2; x[2] = x3;
  z[0] = z1; z[1] = z2; z[2] = z3;
  y[0] = z2 * x3 - z3 * x2;
  y[1] = z3 * x1 - z1 * x3;
  y[2] = z1 * x2 - z2 * x1;
}

// bldmol.f:303-420: x- and z-axis for the first atom from the global
// orientation angles gbpr(4:6) and the first internal coordinates.
static void setgbl(const State& s, int nml, int i2, int i3, double* xg, double* zg) {
  double ca = std::cos(s.gbpr(4, nml)), sa = std::sin(s.gbpr(4, nml));
  double cb = std::cos(s.gbpr(5, nml)), sb = std::sin(s.gbpr(5, nml));
  double cg = std::cos(s.gbpr(6, nml)), sg = std::sin(s.gbpr(6, nml));
  double ag[3][3];
  double x1 = ca * cg - sa * sb * sg, x2 = sa * cg + ca * sb * sg, x3 = -cb * sg;
  double d = std::sqrt(x1 * x1 + x2 * x2 + x3 * x3);
  ag[0][0] = x1 / d; ag[1][0] = x2 / d; ag[2][0] = x3 / d;
  double y1 = -sa * cb, y2 = ca * cb, y3 = sb;
  d = std::sqrt(y1 * y1 + y2 * y2 + y3 * y3);
  ag[0][1] = y1 / d; ag[1][1] = y2 / d; ag[2][1] = y3 / d;
  double z1 = ca * sg + sa * sb * cg, z2 = sa * sg - ca * sb * cg, z3 = cb * cg;
  d = std::sqrt(z1 * z1 + z2 * z2 + z3 * z3);
  ag[0][2] = z1 / d; ag[1][2] = z2 / d; ag[2][2] = z3 / d;
  double ct2 = s.cstoat(i2), st2 = s.sntoat(i2), sa2 = s.snbaat(i2);
  x1 = -s.csbaat(i2); x2 = -sa2 * ct2; x3 = -sa2 * st2;
  double dx = std::sqrt(x1 * x1 + x2 * x2 + x3 * x3);
  x1 /= dx; x2 /= dx; x3 /= dx;
  double st3 = s.sntoat(i3), ct3 = s.cstoat(i3);
  z1 = -st3 * (st2 * x3 + ct2 * x2);
  z2 = st3 * ct2 * x1 + ct3 * st2;
  z3 = st3 * st2 * x1 - ct3 * ct2;
  double dz = std::sqrt(z1 * z1 + z2 * z2 + z3 * z3);
  z1 /= dz; z2 /= dz; z3 /= dz;
  y1 = z2 * x3 - z3 * x2;
  y3 = z1 * x2 - z2 * x1;
  for (int i = 0; i < 3; ++i) {
    xg[i] = ag[i][0] * x1 + ag[i][1] * y1 + ag[i][2] * z1;
    zg[i] = ag[i][0] * x3 + ag[i][1] * y3 + ag[i][2] * z3;
  }
  dx = std::sqrt(xg[0] * xg[0] + xg[1] * xg[1] + xg[2] * xg[2]);
  dz = std::sqrt(zg[0] * zg[0] + zg[1] * zg[1] + zg[2] * zg[2]);
  for (int i = 0; i < 3; ++i) {
    xg[i] /= dx;
    zg[i] /= dz;
  }
}

// setmvs.f:414-492: atom range [ifirg, ilarg] of the branch starting at
// ifirg; bb if it is a backbone atom; ring-closing bond (irg1, irg2).
static void fndbrn(const State& s, int nml, int nrs, int ifirg, int& ilarg, int& irg1, int& irg2, bool& bb) {
  ilarg = ifirg;
  bb = false;
  irg1 = 0;
  irg2 = 0;
  int ifi = s.iatrs1(nrs), ila = s.iatrs2(nrs), ixt = s.ixatrs(nrs);
  if (ifirg == ifi) {
    bb = true;
    if (nrs != s.irsml1(nml)) {
      ilarg = ila;
    } else {
      int ibd[4] = {s.iowat(ifirg), s.ibdat(1, ifirg), s.ibdat(2, ifirg), s.ibdat(3, ifirg)};
      int il = 0;
      for (int i = 0; i < s.nbdat(ifirg) + 1; ++i) {
        int ib = ibd[i];
        if (ib > il && s.iowat(ib) == ifirg) il = ib;
      }
      if (il > 0) ilarg = il - 1;
    }
  } else {
    if (ifirg == ixt) bb = true;
    for (int i = 1; i <= s.nbdat(ifirg); ++i) {
      int ib = s.ibdat(i, ifirg);
      if (s.iowat(ib) == ifirg) {
        for (int j = ib; j <= ila; ++j) {
          if (j > ib && s.iowat(j) < ib) break;
          if (j == ixt) bb = true;
          for (int k = 1; k <= s.nbdat(j); ++k) {
            int jb = s.ibdat(k, j);
            if (jb < ifirg) {
              irg1 = j;
              irg2 = jb;
            }
          }
          ilarg = j;
        }
      } else if (ib < ifirg) {
        irg1 = ifirg;
        irg2 = ib;
      }
    }
  }
}

// bldmol.f:95-173: the first three backbone atoms of molecule nml.
static void fnd3ba(const State& s, int nml, int& i1, int& i2, int& i3) {
  int irs = s.irsml1(nml);
  i1 = s.iatrs1(irs);
  int i, ix, ig2;
  bool bb;
  fndbrn(s, nml, irs, i1, i, ix, ig2, bb);
  i2 = i + 1;
  int ibd[4] = {s.iowat(i1), s.ibdat(1, i1), s.ibdat(2, i1), s.ibdat(3, i1)};
  ix = 0;
  for (int k = 0; k < s.nbdat(i1) + 1; ++k) {
    if (s.iowat(ibd[k]) != i1) {
      if (ix != 0) throw Error("fnd3ba> Can handle only simple ring at 1st atom of a molecule");
      ix = ibd[k];
    }
  }
  ix = s.ixatrs(irs);
  if (i2 == ix) {
    if (irs < s.irsml2(nml)) {
      i3 = s.iatrs1(irs + 1);
      return;
    }
  } else {
    for (int k = ix; k >= i2 + 1; --k)
      if (s.iowat(k) == i2) {
        i3 = k;
        return;
      }
  }
  throw Error("fnd3ba> Cannot find backbone atom following " + s.nmat.get(i2));
}

void first_backbone_atoms(const State& s, int nml, int& i1, int& i2, int& i3) { fnd3ba(s, nml, i1, i2, i3); }

void bldmol(State& s, int nml) {
  int i1, i2, i3;
  fnd3ba(s, nml, i1, i2, i3);
  s.ixrfpt(1, nml) = i1;
  s.ixrfpt(2, nml) = i2;
  s.ixrfpt(3, nml) = i3;
  s.xat(i1) = s.gbpr(1, nml);
  s.yat(i1) = s.gbpr(2, nml);
  s.zat(i1) = s.gbpr(3, nml);
  s.rfpt(1, nml) = s.xat(i1);
  s.rfpt(2, nml) = s.yat(i1);
  s.rfpt(3, nml) = s.zat(i1);
  double xg[3], zg[3];
  setgbl(s, nml, i2, i3, xg, zg);
  s.xbaat(i1) = zg[0]; s.ybaat(i1) = zg[1]; s.zbaat(i1) = zg[2];
  s.xtoat(i1) = xg[0]; s.ytoat(i1) = xg[1]; s.ztoat(i1) = xg[2];
  int ifirs = s.irsml1(nml);
  for (int i = ifirs; i <= s.irsml2(nml); ++i) {
    int jj = (i == ifirs) ? s.iatrs1(i) + 1 : s.iatrs1(i);
    for (int j = jj; j <= s.iatrs2(i); ++j) eyring(s, j, s.iowat(j));
  }
  double x[3], y[3], z[3];
  setsys(s, i1, i2, i3, x, y, z);
  for (int k = 0; k < 3; ++k) {
    s.xrfax(k + 1, nml) = x[k];
    s.yrfax(k + 1, nml) = -y[k];
    s.zrfax(k + 1, nml) = -z[k];
  }
}

void setvar(State& s, int nml, const double* vlvrx) {
  int i1vr = s.ivrml1(nml);
  for (int i = i1vr; i <= i1vr + s.nvrml(nml) - 1; ++i) {
    int iat = s.iatvr(i), ity = s.ityvr(i);
    if (ity == 3) {  // torsion: keep the phase of every branch of the base
      double tsh = difang(s.toat(iat), vlvrx[i - 1]);
      int iow = s.iowat(iat);
      for (int j = 1; j <= s.nbdat(iow); ++j) {
        int jat = s.ibdat(j, iow);
        if (s.iowat(jat) == iow) {
          double to = s.toat(jat) + tsh;
          if (std::fabs(to) > pi) to = to - std::copysign(pi2, to);
          s.toat(jat) = to;
          s.sntoat(jat) = std::sin(to);
          s.cstoat(jat) = std::cos(to);
        }
      }
    } else if (ity == 2) {
      double ba = vlvrx[i - 1];
      s.baat(iat) = ba;
      s.snbaat(iat) = std::sin(ba);
      s.csbaat(iat) = std::cos(ba);
    } else if (ity == 1) {
      s.blat(iat) = vlvrx[i - 1];
    }
  }
  bldmol(s, nml);
}

// dihedr.f:12-70 and :73-121 (atoms 0 and -1 are the virtual atoms).
static double dihedr(const State& s, int i1, int i2, int i3, int i4) {
  double x1 = s.xat(i2) - s.xat(i1), y1 = s.yat(i2) - s.yat(i1), z1 = s.zat(i2) - s.zat(i1);
  double x2 = s.xat(i3) - s.xat(i2), y2 = s.yat(i3) - s.yat(i2), z2 = s.zat(i3) - s.zat(i2);
  double ux1 = y1 * z2 - z1 * y2, uy1 = z1 * x2 - x1 * z2, uz1 = x1 * y2 - y1 * x2;
  x1 = s.xat(i4) - s.xat(i3); y1 = s.yat(i4) - s.yat(i3); z1 = s.zat(i4) - s.zat(i3);
  double ux2 = z1 * y2 - y1 * z2, uy2 = x1 * z2 - z1 * x2, uz2 = y1 * x2 - x1 * y2;
  double u1 = ux1 * ux1 + uy1 * uy1 + uz1 * uz1, u2 = ux2 * ux2 + uy2 * uy2 + uz2 * uz2;
  double u = u1 * u2;
  if (u == 0.0) throw Error("dihedr> Error in coordinates of atoms");
  double a = (ux1 * ux2 + uy1 * uy2 + uz1 * uz2) / std::sqrt(u);
  a = std::max(a, -1.0);
  a = std::min(a, 1.0);
  double d = std::acos(a);
  if (ux1 * (uy2 * z2 - uz2 * y2) + uy1 * (uz2 * x2 - ux2 * z2) + uz1 * (ux2 * y2 - uy2 * x2) < 0.0) d = -d;
  return d;
}
static double valang(const State& s, int i1, int i2, int i3) {
  double h1 = s.xat(i2), h2 = s.yat(i2), h3 = s.zat(i2);
  double x1 = s.xat(i1) - h1, x2 = s.yat(i1) - h2, x3 = s.zat(i1) - h3;
  double y1 = s.xat(i3) - h1, y2 = s.yat(i3) - h2, y3 = s.zat(i3) - h3;
  double x = x1 * x1 + x2 * x2 + x3 * x3, y = y1 * y1 + y2 * y2 + y3 * y3;
  double u = x * y;
  if (u == 0.0) throw Error("valang> Error in coordinates of atoms");
  double a = (x1 * y1 + x2 * y2 + x3 * y3) / std::sqrt(u);
  a = std::max(a, -1.0);
  a = std::min(a, 1.0);
  return std::acos(a);
}

// ------------------------------------------------------ terminal groups

// addend.f:161-647 (rplgrp): replaces the atoms rooted at 'rpat' in residue
// nrs by those rooted at the atom of the same name in library residue sbrs.
static void rplgrp(State& s, int nml, int nrs, const std::string& rpat_in, const std::string& sbrs) {
  const std::string rpat = pad(rpat_in, 4);
  int ifirs = s.irsml1(nml), ilars = s.irsml2(nml);
  int nxt = s.ixatrs(nrs), nfi = s.iatrs1(nrs), nla = s.iatrs2(nrs);
  int nfirp = 0;
  for (int i = nfi; i <= nla; ++i)
    if (s.nmat.get(i) == rpat) {
      nfirp = i;
      break;
    }
  if (nfirp == 0) throw Error("rplgrp> Cannot find atom >" + rpat + "< to be replaced");
  int nlarp, irng1, irng2;
  bool bb;
  fndbrn(s, nml, nrs, nfirp, nlarp, irng1, irng2, bb);
  auto err10 = [&]() { return Error("rplgrp> Cannot replace atom(s) following " + rpat + " from INSIDE a ring"); };
  auto err11 = [&]() { return Error("rplgrp> Cannot replace atom(s) following " + rpat + ": 2 previous atoms missing"); };
  if (irng1 > 0) throw err10();
  bool ntbb = false;
  int nxtbb1 = 0, nxtbb2 = 0, isgbb1 = 1, ibdrg = 0, iybdrg = 1;
  bool to_label2 = !bb;
  if (bb) {
    if (nfirp == nxt && nrs == ilars) {
      to_label2 = true;
    } else if (nfirp == nfi && nrs == ifirs) {
      ntbb = true;
      int ibd[mxbd + 1] = {s.iowat(nfirp), s.ibdat(1, nfirp), s.ibdat(2, nfirp), s.ibdat(3, nfirp)};
      int iybd[mxbd + 1] = {s.iyowat(nfirp), s.iybdat(1, nfirp), s.iybdat(2, nfirp), s.iybdat(3, nfirp)};
      for (int i = 0; i < s.nbdat(nfirp) + 1; ++i) {
        if (s.iowat(ibd[i]) != nfirp) {
          if (ibdrg != 0) throw Error("rplgrp> Can handle only simple ring at 1st atom of a molecule");
          ibdrg = ibd[i];
          iybdrg = iybd[i];
        }
      }
      nxtbb1 = nlarp + 1;
      isgbb1 = s.iyowat(nxtbb1);
      bool found = false;
      if (nxtbb1 == nxt) {
        if (nrs < ilars) {
          nxtbb2 = s.iatrs1(nrs + 1);
          found = true;
        }
      } else {
        for (int i = nxt; i >= nxtbb1 + 1; --i)
          if (s.iowat(i) == nxtbb1) {
            nxtbb2 = i;
            found = true;
            break;
          }
      }
      if (!found) throw err11();
    } else {
      throw Error("rplgrp> Cannot replace BACKBONE atom " + rpat);
    }
  }
  if (to_label2) {
    if (nfirp == nfi && nrs == ifirs) throw err11();
    nxtbb1 = s.iowat(nfirp);
    if (nxtbb1 == nfi && nrs == ifirs) throw err11();
    nxtbb2 = s.iowat(nxtbb1);
  }
  // substituent from the library
  const LibResidue& sb = library(s, sbrs);
  int natsb = sb.nat, nvrsb = sb.nvr;
  int nfisb = 0;
  for (int i = 1; i <= natsb; ++i)
    if (sb.nm[i] == rpat) {
      nfisb = i;
      break;
    }
  if (nfisb == 0) throw Error("rplgrp> Cannot find atom >" + rpat + "< in substituent residue " + sbrs);
  int nlasb = nfisb;
  for (int i = 1; i <= sb.nbd[nfisb]; ++i) {
    int ib = sb.ibd[3 * nfisb + i - 1];
    if (sb.iow[ib] < nfisb) throw err10();
    for (int j = ib; j <= natsb; ++j) {
      if (j > ib && sb.iow[j] < ib) break;
      for (int k = 1; k <= sb.nbd[j]; ++k)
        if (sb.ibd[3 * j + k - 1] < nfisb) throw err10();
      nlasb = j;
    }
  }
  // local axes at nfirp
  double x[3], y[3], z[3];
  setsys(s, nxtbb1, nfirp, nxtbb2, x, y, z);
  s.xtoat(nfirp) = x[0]; s.ytoat(nfirp) = x[1]; s.ztoat(nfirp) = x[2];
  s.xbaat(nfirp) = z[0]; s.ybaat(nfirp) = z[1]; s.zbaat(nfirp) = z[2];
  if (ntbb) {  // virtual atoms 0 and -1 fix the geometry before the N-terminus
    double ct = s.cstoat(nxtbb2), st = s.sntoat(nxtbb2);
    double ca = s.csbaat(nxtbb1), sa = s.snbaat(nxtbb1);
    double h2 = -sa * ct, h3 = -sa * st;
    double x1 = -ca * x[0] + h2 * y[0] + h3 * z[0];
    double x2 = -ca * x[1] + h2 * y[1] + h3 * z[1];
    double x3 = -ca * x[2] + h2 * y[2] + h3 * z[2];
    double dx = 1.0 / std::sqrt(x1 * x1 + x2 * x2 + x3 * x3);
    x1 *= dx; x2 *= dx; x3 *= dx;
    s.xat(0) = s.xat(nfirp) + x1;
    s.yat(0) = s.yat(nfirp) + x2;
    s.zat(0) = s.zat(nfirp) + x3;
    double z1 = -st * y[0] + ct * z[0], z2 = -st * y[1] + ct * z[1], z3 = -st * y[2] + ct * z[2];
    double dz = 1.0 / std::sqrt(z1 * z1 + z2 * z2 + z3 * z3);
    z1 *= dz; z2 *= dz; z3 *= dz;
    ct = s.cstoat(nxtbb1);
    st = s.sntoat(nxtbb1);
    s.xat(-1) = s.xat(0) - ct * (z2 * x3 - z3 * x2) - st * z1;
    s.yat(-1) = s.yat(0) - ct * (z3 * x1 - z1 * x3) - st * z2;
    s.zat(-1) = s.zat(0) - ct * (z1 * x2 - z2 * x1) - st * z3;
  }
  // shift the atoms after the replaced group
  int nrp = nlarp - nfirp, nsb = nlasb - nfisb, nsh = 0;
  if (nrp != nsb) {
    nsh = nsb - nrp;
    for (int i = nfi; i <= nfirp - 1; ++i)
      for (int j = 1; j <= mxbd; ++j) {
        int ib = s.ibdat(j, i);
        if (ib > nlarp && ib <= nla) s.ibdat(j, i) = ib + nsh;
      }
    int ilaat = s.iatrs2(s.irsml2(s.ntlml));
    int i1, i2, i3;
    if (nrp > nsb) {
      i1 = nlarp + 1; i2 = ilaat; i3 = 1;
    } else {
      if (ilaat + nsh > mxat) throw Error("rplgrp> too many atoms (mxat)");
      i1 = ilaat; i2 = nlarp + 1; i3 = -1;
    }
    for (int i = i1; i3 > 0 ? i <= i2 : i >= i2; i += i3) {
      int ii = i + nsh;
      s.nbdat(ii) = s.nbdat(i);
      int ibd[mxbd + 1] = {s.iowat(i), s.ibdat(1, i), s.ibdat(2, i), s.ibdat(3, i)};
      int iybd[mxbd + 1] = {s.iyowat(i), s.iybdat(1, i), s.iybdat(2, i), s.iybdat(3, i)};
      for (int j = 0; j < mxbd + 1; ++j)
        if (ibd[j] > nfirp) ibd[j] += nsh;
      s.iowat(ii) = ibd[0];
      s.iyowat(ii) = iybd[0];
      for (int j = 1; j <= mxbd; ++j) {
        s.ibdat(j, ii) = ibd[j];
        s.iybdat(j, ii) = iybd[j];
      }
      s.ityat(ii) = s.ityat(i);
      std::memcpy(s.nmat.d[ii - 1], s.nmat.d[i - 1], 4);
      s.cgat(ii) = s.cgat(i); s.blat(ii) = s.blat(i);
      s.baat(ii) = s.baat(i); s.snbaat(ii) = s.snbaat(i); s.csbaat(ii) = s.csbaat(i);
      s.toat(ii) = s.toat(i); s.sntoat(ii) = s.sntoat(i); s.cstoat(ii) = s.cstoat(i);
      s.xat(ii) = s.xat(i); s.yat(ii) = s.yat(i); s.zat(ii) = s.zat(i);
      s.xtoat(ii) = s.xtoat(i); s.ytoat(ii) = s.ytoat(i); s.ztoat(ii) = s.ztoat(i);
      s.xbaat(ii) = s.xbaat(i); s.ybaat(ii) = s.ybaat(i); s.zbaat(ii) = s.zbaat(i);
    }
    for (int i = nrs + 1; i <= s.irsml2(s.ntlml); ++i) {
      s.iatrs1(i) += nsh;
      s.iatrs2(i) += nsh;
      s.ixatrs(i) += nsh;
    }
    s.iatrs2(nrs) = nla + nsh;
    if (nxt > nlarp) s.ixatrs(nrs) = nxt + nsh;
  }
  // the root atom takes type, charge and bonds of the substituent's root
  int ish = nfirp - nfisb;
  s.ityat(nfirp) = sb.ity[nfisb];
  if (!s.sh2) s.cgat(nfirp) = sb.cg[nfisb];
  int nb = sb.nbd[nfisb];
  int ibd[mxbd + 1], iybd[mxbd + 1];
  for (int i = 0; i < mxbd; ++i) {
    int ib = sb.ibd[3 * nfisb + i];
    if (ib != 0) {
      ibd[i] = ib + ish;
      iybd[i] = sb.iybd[3 * nfisb + i];
    } else {
      ibd[i] = 0;
      iybd[i] = 1;
    }
  }
  if (ntbb) {
    ibd[mxbd] = 0;
    iybd[mxbd] = 1;
    ibd[nb] = nxtbb1 + nsh;
    iybd[nb] = isgbb1;
    if (ibdrg != 0) {
      ++nb;
      if (nb > mxbd) throw Error("rplgrp> Cannot add atoms: need more bonds");
      ibd[nb] = ibdrg + nsh;
      iybd[nb] = iybdrg;
    }
    s.iowat(nfirp) = ibd[0];
    s.iyowat(nfirp) = iybd[0];
    for (int i = 1; i <= mxbd; ++i) {
      s.ibdat(i, nfirp) = ibd[i];
      s.iybdat(i, nfirp) = iybd[i];
    }
  } else {
    for (int i = 1; i <= mxbd; ++i) {
      s.ibdat(i, nfirp) = ibd[i - 1];
      s.iybdat(i, nfirp) = iybd[i - 1];
    }
  }
  s.nbdat(nfirp) = nb;
  // substituent atoms, built with Eyring's step from their predecessors
  int ii = nfirp;
  for (int i = nfisb + 1; i <= nlasb; ++i) {
    ++ii;
    s.nbdat(ii) = sb.nbd[i];
    int iow = sb.iow[i] + ish;
    s.iowat(ii) = iow;
    s.iyowat(ii) = sb.iyow[i];
    for (int j = 1; j <= mxbd; ++j) {
      int ib = sb.ibd[3 * i + j - 1];
      s.ibdat(j, ii) = (ib >= nfisb) ? ib + ish : ib;
      s.iybdat(j, ii) = sb.iybd[3 * i + j - 1];
    }
    s.ityat(ii) = sb.ity[i];
    s.nmat.set(ii, sb.nm[i]);
    s.cgat(ii) = sb.cg[i];
    s.blat(ii) = sb.bl[i];
    double ba = sb.ba[i];
    s.baat(ii) = ba; s.csbaat(ii) = std::cos(ba); s.snbaat(ii) = std::sin(ba);
    double to = sb.to[i];
    s.toat(ii) = to; s.cstoat(ii) = std::cos(to); s.sntoat(ii) = std::sin(to);
    eyring(s, ii, iow);
    if (ntbb) {
      if (iow == nfirp) {
        ba = valang(s, 0, nfirp, ii);
        s.baat(ii) = ba; s.csbaat(ii) = std::cos(ba); s.snbaat(ii) = std::sin(ba);
        to = dihedr(s, -1, 0, nfirp, ii);
        s.toat(ii) = to; s.cstoat(ii) = std::cos(to); s.sntoat(ii) = std::sin(to);
      } else if (s.iowat(iow) == nfirp) {
        to = dihedr(s, 0, nfirp, iow, ii);
        s.toat(ii) = to; s.cstoat(ii) = std::cos(to); s.sntoat(ii) = std::sin(to);
      }
    }
  }
  // variables of the replaced group and of the substituent (stored
  // consecutively in both)
  int ilavr = s.ivrml1(s.ntlml) + s.nvrml(s.ntlml) - 1;
  int ifivr = s.ivrrs1(nrs);
  int ivrrp = 0, lvrrp = 0;
  for (int i = ifivr; i <= ifivr + s.nvrrs(nrs) - 1; ++i) {
    int iat = s.iatvr(i);
    if (nfirp < iat && iat <= nlarp) {
      ++ivrrp;
      lvrrp = i;
    }
  }
  if (ivrrp == 0) {
    lvrrp = ilavr;
    for (int i = ifivr; i <= ilavr; ++i)
      if (s.iatvr(i) > nlarp) {
        lvrrp = i - 1;
        break;
      }
  }
  std::vector<int> sty, scl, sat;
  std::vector<std::string> snm;
  for (int i = 0; i < nvrsb; ++i) {
    int iat = sb.vat[i];
    if (nfisb < iat && iat <= nlasb) {
      int ity = sb.vty[i];
      if (ntbb && sb.iow[iat] == nfisb && ity > 2) continue;
      sty.push_back(ity);
      scl.push_back(sb.vcl[i]);
      sat.push_back(iat);
      snm.push_back(sb.vnm[i]);
    }
  }
  int ivrsb = (int)sty.size();
  if (nsh != 0)
    for (int i = lvrrp + 1; i <= ilavr; ++i) s.iatvr(i) += nsh;
  if (ivrrp != ivrsb) {
    int jsh = ivrsb - ivrrp;
    int i1, i2, i3;
    if (ivrrp > ivrsb) {
      i1 = lvrrp + 1; i2 = ilavr; i3 = 1;
    } else {
      if (ilavr + jsh > mxvr) throw Error("rplgrp> number of variables > mxvr");
      i1 = ilavr; i2 = lvrrp + 1; i3 = -1;
    }
    for (int i = i1; i3 > 0 ? i <= i2 : i >= i2; i += i3) {
      int k = i + jsh;
      s.ityvr(k) = s.ityvr(i);
      s.iclvr(k) = s.iclvr(i);
      std::memcpy(s.nmvr.d[k - 1], s.nmvr.d[i - 1], 3);
      s.iatvr(k) = s.iatvr(i);
    }
    for (int i = nrs + 1; i <= s.irsml2(s.ntlml); ++i) s.ivrrs1(i) += jsh;
    s.nvrrs(nrs) += jsh;
    s.nvrml(nml) += jsh;
    for (int i = nml + 1; i <= s.ntlml; ++i) s.ivrml1(i) += jsh;
  }
  ii = lvrrp - ivrrp;
  for (int i = 0; i < ivrsb; ++i) {
    ++ii;
    s.nmvr.set(ii, snm[i]);
    s.ityvr(ii) = sty[i];
    s.iclvr(ii) = scl[i];
    s.iatvr(ii) = sat[i] + ish;
  }
}

// addend.f:13-159: N- and C-terminal groups (nh2 / nh3+, cooh / coo-).
static void addend(State& s, int nml, const std::string& grpn_in, const std::string& grpc_in) {
  std::string grn = pad(lower(rtrim(grpn_in)), 4), grc = pad(lower(rtrim(grpc_in)), 4);
  if (grn.substr(0, 3) == "ace" || grc.substr(0, 3) == "ace" || grn.substr(0, 3) == "nme" || grc.substr(0, 3) == "nme")
    throw Error("addend> N-Acetyl (ace) or N-Methylamide (nme) should be put in SEQUENCE file, not added as end groups");
  int ifirs = s.irsml1(nml);
  std::string res = lower(s.seq.get(ifirs));
  if (res.substr(0, 3) != "ace") {
    std::string sbrs;
    if (grn.substr(0, 3) == "nh2") {
      if (res.substr(0, 3) == "hyp" && !s.flex) {
        if (s.sh2) sbrs = "nh2+";
        else throw Error("addend> No N-terminal Hyp possible with ECEPP/3 dataset");
      } else if (res.substr(0, 3) == "pro") {
        sbrs = "nh1 ";
      } else if (res.substr(0, 3) == "gly" && !s.flex) {
        sbrs = s.sh2 ? "nh2 " : "nh2g";
      } else {
        sbrs = "nh2 ";
      }
    } else if (grn == "nh3+") {
      if (res.substr(0, 3) == "pro" && !s.flex) sbrs = "nh2+";
      else if (res.substr(0, 3) == "gly" && !s.flex) sbrs = "nh3g";
      else sbrs = "nh3+";
    } else {
      throw Error("addend> Can add only nh2 or nh3+ to N-terminus");
    }
    rplgrp(s, nml, ifirs, "n   ", sbrs);
  }
  int ilars = s.irsml2(nml);
  res = lower(s.seq.get(ilars));
  if (res.substr(0, 3) != "nme") {
    std::string sbrs;
    if (grc == "cooh") sbrs = (res.substr(0, 3) == "gly" && !s.sh2) ? "gooh" : "cooh";
    else if (grc == "coo-") sbrs = (res.substr(0, 3) == "gly" && !s.sh2) ? "goo-" : "coo-";
    else throw Error("addend> Can add only cooh or coo- to C-terminus");
    rplgrp(s, nml, ilars, "c   ", sbrs);
  }
}

// ------------------------------------------------------- moving sets

int nursvr(const State& s, int ivr) {
  for (int i = s.ntlml; i >= 1; --i) {
    int ifirs = s.irsml1(i);
    if (ivr >= s.ivrrs1(ifirs) && s.nvrml(i) > 0)
      for (int j = s.irsml2(i); j >= ifirs; --j)
        if (ivr >= s.ivrrs1(j) && s.nvrrs(j) > 0) return j;
  }
  throw Error("nursvr> Cannot find variable");
}

int nursat(const State& s, int iat) {
  for (int i = 1; i <= s.ntlml; ++i) {
    int ifirs = s.irsml1(i), ilars = s.irsml2(i);
    if (iat >= s.iatrs1(ifirs) && iat <= s.iatrs2(ilars))
      for (int j = ifirs; j <= ilars; ++j)
        if (iat >= s.iatrs1(j) && iat <= s.iatrs2(j)) return j;
  }
  throw Error("nursat> Cannot find atom");
}

// setmvs.f:13-412: variable order iorvr (heap sort on primary atom, the same
// sift-down sequence as the reference so equal keys order identically),
// non-overlapping moving sets latms1/2 per variable, ixmsat, added variables.
static void setmvs(State& s, int nml) {
  int ntlvr = s.nvrml(nml);
  int nms, nad;
  if (nml == 1) {
    s.imsml1(1) = 1;
    nms = 0;
    s.iadml1(1) = 1;
    nad = 0;
  } else {
    s.imsml1(nml) = s.imsml1(nml - 1) + s.nmsml(nml - 1);
    nms = s.imsml1(nml) - 1;
    s.iadml1(nml) = s.iadml1(nml - 1) + s.nadml(nml - 1);
    nad = s.iadml1(nml) - 1;
  }
  if (ntlvr == 0) {
    s.nmsml(nml) = 0;
    s.nadml(nml) = 0;
    return;
  }
  int ifirs = s.irsml1(nml), ilars = s.irsml2(nml);
  int ifivr = s.ivrml1(nml), ilavr = ifivr + ntlvr - 1;
  int ifiat = s.iatrs1(ifirs), ilaat = s.iatrs2(ilars);
  auto& key = s.iadvr1;  // sort keys live in iadvr1 until it is recomputed
  for (int n = ifirs; n <= ilars; ++n) {
    int ib = s.ivrrs1(n);
    for (int i = ib; i <= ib + s.nvrrs(n) - 1; ++i) {
      int ia = s.iatvr(i), io = s.iowat(ia), it = s.ityvr(i);
      if (it == 3) {
        for (int j = 1; j <= s.nbdat(io); ++j) {
          int ii = s.ibdat(j, io);
          if (s.iowat(ii) == io) ia = std::min(ia, ii);
        }
        key(i) = ia * 10;
      } else if (it == 2) {
        key(i) = ia * 10 + 1;
      } else if (it == 1) {
        key(i) = ia * 10 + 2;
      }
      s.iorvr(i) = i;
    }
  }
  if (ntlvr > 1) {  // heap sort of iorvr(ifivr..ilavr) by key
    auto a = [&](int r) -> int& { return s.iorvr(ifivr - 1 + r); };  // 1-based within the range
    int n = ntlvr, l = n / 2 + 1, ir = n;
    for (;;) {
      int io, kv;
      if (l > 1) {
        --l;
        io = a(l);
        kv = key(io);
      } else {
        io = a(ir);
        kv = key(io);
        a(ir) = a(1);
        if (--ir == 1) {
          a(1) = io;
          break;
        }
      }
      int i = l, j = l + l;
      while (j <= ir) {
        if (j < ir && key(a(j)) < key(a(j + 1))) ++j;
        if (kv < key(a(j))) {
          a(i) = a(j);
          i = j;
          j += j;
        } else {
          j = ir + 1;
        }
      }
      a(i) = io;
    }
  }
  nms = s.imsml1(nml) - 1;
  for (int io = ifivr; io <= ilavr; ++io) {
    int iv = s.iorvr(io);
    int ir = nursvr(s, iv);
    int ia = s.iatvr(iv), ib = s.iowat(ia), it = s.ityvr(iv);
    int i1 = 0, i2 = 0;
    if (it == 3) {
      for (int i = 1; i <= s.nbdat(ib); ++i) {
        int j = s.ibdat(i, ib);
        if (s.iowat(j) != ib) continue;
        int k, irg1, irg2;
        bool bb;
        fndbrn(s, nml, ir, j, k, irg1, irg2, bb);
        if (bb) k = ilaat;
        if (i1 != 0) {
          if (j > i2 + 1 || k < i1 - 1)
            throw Error("setmvs> Cannot combine disjunct ranges of atom indices for torsion " + s.nmvr.get(iv));
          if (j < i1) i1 = j;
          if (k > i2) i2 = k;
        } else {
          i1 = j;
          i2 = k;
        }
      }
    } else {
      i1 = ia;
      int irg1, irg2;
      bool bb;
      fndbrn(s, nml, ir, i1, i2, irg1, irg2, bb);
      if (bb) i2 = ilaat;
    }
    if (nms + 1 > mxms) throw Error("setmvs> Number of moving sets > mxms");
    s.imsvr1(iv) = nms + 1;
    s.imsvr2(iv) = nms + 1;
    // trim the moving sets of earlier variables so the sets do not overlap
    for (int jo = ifivr; jo <= io - 1; ++jo) {
      int jv = s.iorvr(jo);
      int j1s = s.imsvr1(jv), jns = s.imsvr2(jv) - j1s + 1;
      int j = j1s;
      while (j < j1s + jns) {
        int j1 = s.latms1(j), j2 = s.latms2(j);
        if (i1 <= j2 && i2 >= j1) {
          int ja = 0;
          if (i1 > j1) {
            if (i2 > j2) throw Error("setmvs> atom ranges of two variables overlap only PARTLY");
            ja = 1;
            s.latms2(j) = i1 - 1;
          }
          if (i2 < j2) {
            if (i1 < j1) throw Error("setmvs> atom ranges of two variables overlap only PARTLY");
            if (ja > 0) {
              ++nms;
              if (nms > mxms) throw Error("setmvs> Number of moving sets > mxms");
              ++jns;
              for (int k = nms; k >= j + 2; --k) {
                s.latms1(k) = s.latms1(k - 1);
                s.latms2(k) = s.latms2(k - 1);
              }
              for (int ko = jo + 1; ko <= io; ++ko) {
                int k = s.iorvr(ko);
                s.imsvr1(k) += 1;
                s.imsvr2(k) += 1;
              }
              s.latms2(j + 1) = j2;
            }
            s.latms1(j + ja) = i2 + 1;
            ++ja;
          }
          if (ja == 0) {
            --nms;
            --jns;
            for (int k = j; k <= nms; ++k) {
              s.latms1(k) = s.latms1(k + 1);
              s.latms2(k) = s.latms2(k + 1);
            }
            for (int ko = jo + 1; ko <= io; ++ko) {
              int k = s.iorvr(ko);
              s.imsvr1(k) -= 1;
              s.imsvr2(k) -= 1;
            }
          } else {
            j += ja;
          }
        } else {
          ++j;
        }
      }
      s.imsvr2(jv) = j1s + jns - 1;
    }
    ++nms;
    s.latms1(nms) = i1;
    s.latms2(nms) = i2;
  }
  s.nmsml(nml) = nms - s.imsml1(nml) + 1;
  for (int ia = ifiat; ia <= ilaat; ++ia) s.ixmsat(ia) = 0;
  for (int is = s.imsml1(nml); is <= nms; ++is)
    for (int ia = s.latms1(is); ia <= s.latms2(is); ++ia) s.ixmsat(ia) = is;
  // added variables (setmvs.f:290-369)
  int i = s.iorvr(ifivr);
  int ii = s.imsvr1(i);
  for (int io = ifivr; io <= ilavr - 1; ++io) {
    int ic = i, ia = s.iatvr(i), ib = s.iowat(ia), it = s.ityvr(i), is = ii;
    int n = nad + 1;
    s.iadvr1(i) = n;
    i = s.iorvr(io + 1);
    ii = s.imsvr1(i);
    for (int jo = io + 1; jo <= ilavr; ++jo) {
      int j = s.iorvr(jo), ja = s.iatvr(j), jb = s.iowat(ja);
      auto registered = [&](int from) {
        for (int k = from; k <= nad; ++k)
          if (s.iatvr(s.ladvr(k)) == ja) return true;
        return false;
      };
      if (it == 3 && jb == ib) {
        if (!registered(n)) {
          ++nad;
          if (nad > mxvr) throw Error("setmvs> Number of added variables > mxvr");
          s.ladvr(nad) = j;
        }
      }
      if (is < ii) {
        for (int k = is; k <= ii - 1; ++k) {
          if (s.latms1(k) <= jb && jb <= s.latms2(k)) {
            if (!registered(n)) {
              ++nad;
              if (nad > mxvr) throw Error("setmvs> Number of added variables > mxvr");
              s.ladvr(nad) = j;
            }
          }
        }
      } else if (ja == ia) {
        if (!registered(n)) {
          ++nad;
          if (nad > mxvr) throw Error("setmvs> Number of added variables > mxvr");
          s.ladvr(nad) = j;
        }
      }
    }
    s.iadvr2(ic) = nad;
  }
  s.iadvr1(i) = nad + 1;
  s.iadvr2(i) = nad;
  s.nadml(nml) = nad - s.iadml1(nml) + 1;
}

// ------------------------------------------------- interaction lists

// Sorted list of disjoint atom-index ranges; quench() (mklist.f:305-372)
// removes [i1,i2] and reports whether anything overlapped.
struct RangeList {
  std::vector<int> a, b;
  int n() const { return (int)a.size(); }
  bool quench(int i1, int i2) {
    bool hit = false;
    size_t j = 0;
    while (j < a.size()) {
      int j1 = a[j], j2 = b[j];
      if (i1 <= j2 && i2 >= j1) {
        hit = true;
        bool left = i1 > j1, right = i2 < j2;
        if (left && right) {  // split in two
          a.insert(a.begin() + j + 1, i2 + 1);
          b.insert(b.begin() + j + 1, j2);
          b[j] = i1 - 1;
          j += 2;
        } else if (left) {
          b[j] = i1 - 1;
          ++j;
        } else if (right) {
          a[j] = i2 + 1;
          ++j;
        } else {  // fully covered
          a.erase(a.begin() + j);
          b.erase(b.begin() + j);
        }
      } else {
        ++j;
      }
    }
    return hit;
  }
};

// mklist.f:12-303: per-atom van der Waals domains (ranges of partners) and
// 1-4 partners, built variable by variable from the last to the first.
static void mklist(State& s, int nml) {
  constexpr int mxh = 50, mx2 = 50;
  if (nml == 1) {
    s.ivwml1(1) = 1;
    s.i14ml1(1) = 1;
  } else {
    s.ivwml1(nml) = s.ivwml1(nml - 1) + s.nvwml(nml - 1);
    s.i14ml1(nml) = s.i14ml1(nml - 1) + s.n14ml(nml - 1);
  }
  int ntlms = s.nmsml(nml);
  if (ntlms == 0) {
    s.nvwml(nml) = 0;
    s.n14ml(nml) = 0;
    return;
  }
  int nvw = s.ivwml1(nml) - 1, n14 = s.i14ml1(nml) - 1;
  int ifiat = s.iatrs1(s.irsml1(nml)), ilaat = s.iatrs2(s.irsml2(nml));
  int ifivr = s.ivrml1(nml), ilavr = ifivr + s.nvrml(nml) - 1;
  for (int i = ifiat; i <= ilaat; ++i) {
    s.ivwat1(i) = 0;
    s.i14at1(i) = 0;
  }
  RangeList l1st;
  l1st.a = {ifiat};
  l1st.b = {ilaat};
  int i1s = s.imsml1(nml) + ntlms;
  for (int io = ilavr; io >= ifivr; --io) {
    int iv = s.iorvr(io), it = s.ityvr(iv);
    int i2s = i1s - 1;
    i1s = s.imsvr1(iv);
    if (i2s - i1s + 1 <= 0) continue;
    for (int is = i1s; is <= i2s; ++is) l1st.quench(s.latms1(is), s.latms2(is));
    if (l1st.n() > mxh) throw Error("mklist> too many atom regions");
    RangeList l2nd = l1st;
    int ib = s.iowat(s.iatvr(iv));
    l2nd.quench(ib, ib);
    bool ovlp = false;
    int iob = s.iowat(ib);
    std::vector<int> l2i, l1i;
    auto push2 = [&](int x) {
      if ((int)l2i.size() + 1 > mx2) throw Error("mklist> too many atoms in list L2I");
      l2i.push_back(x);
    };
    if (iob > 0) {
      ovlp = l2nd.quench(iob, iob);
      for (int i = 1; i <= s.nbdat(iob); ++i) {
        int ibd = s.ibdat(i, iob);
        if (ibd != ib && s.iowat(ibd) == iob && l2nd.quench(ibd, ibd)) push2(ibd);
      }
      const size_t n2i = l2i.size();  // a Fortran DO fixes its trip count
      for (size_t i = 0; i < n2i; ++i) {
        int ia = l2i[i], im = s.ixmsat(ia);
        if (im > 0)
          for (int j = s.latms1(im); j <= s.latms2(im); ++j)
            if (ia != j && l2nd.quench(j, j)) push2(j);
      }
      int ioiob = s.iowat(iob);
      if (ioiob > 0) {
        if (l2nd.quench(ioiob, ioiob)) push2(ioiob);
      }
    } else {
      iob = -10;
    }
    for (int i = 1; i <= s.nbdat(ib); ++i) {
      int ibd = s.ibdat(i, ib);
      if (s.iowat(ibd) == ib && l2nd.quench(ibd, ibd)) {
        if ((int)l1i.size() + 1 > mxbd) throw Error("mklist> too many atoms bound to " + s.nmat.get(ib));
        l1i.push_back(ibd);
        for (int j = 1; j <= s.nbdat(ibd); ++j) {
          int jbd = s.ibdat(j, ibd);
          if (s.iowat(jbd) == ibd && l2nd.quench(jbd, jbd)) push2(jbd);
        }
      }
    }
    int im = s.ixmsat(ib);
    if (im > 0)
      for (int i = s.latms1(im); i <= s.latms2(im); ++i)
        if (l2nd.quench(i, i)) push2(i);
    for (int is = i1s; is <= i2s; ++is) {
      for (int i = s.latms1(is); i <= s.latms2(is); ++i) {
        if (nvw + l2nd.n() > mxvw) throw Error("mklist> Number of vdw-domains > mxvw");
        s.ivwat1(i) = nvw + 1;
        s.ivwat2(i) = nvw + l2nd.n();
        for (int j = 0; j < l2nd.n(); ++j) {
          ++nvw;
          s.ixatvw(nvw) = i;
          s.lvwat1(nvw) = l2nd.a[j];
          s.lvwat2(nvw) = l2nd.b[j];
        }
        if (n14 + (int)l2i.size() > mx14) throw Error("mklist> Number of 1-4 interactions > mx14");
        s.i14at1(i) = n14 + 1;
        for (int x : l2i) {
          ++n14;
          s.ixat14(n14) = i;
          s.l14at(n14) = x;
        }
        int iow = s.iowat(i);
        if (iow != ib) {
          if (n14 + (int)l1i.size() > mx14) throw Error("mklist> Number of 1-4 interactions > mx14");
          for (int x : l1i) {
            ++n14;
            s.ixat14(n14) = i;
            s.l14at(n14) = x;
          }
          if (ovlp && (it == 1 || it == 2)) {
            ++n14;
            if (n14 > mx14) throw Error("mklist> Number of 1-4 interactions > mx14");
            s.ixat14(n14) = i;
            s.l14at(n14) = iob;
          }
          if (s.iowat(iow) != ib && it == 1) {
            ++n14;
            if (n14 > mx14) throw Error("mklist> Number of 1-4 interactions > mx14");
            s.ixat14(n14) = i;
            s.l14at(n14) = ib;
          }
        }
        s.i14at2(i) = n14;
      }
    }
  }
  s.nvwml(nml) = nvw - s.ivwml1(nml) + 1;
  s.n14ml(nml) = n14 - s.i14ml1(nml) + 1;
  for (int i = ifiat; i <= ilaat; ++i) {
    if (s.ivwat1(i) <= 0) {
      s.ivwat1(i) = 1;
      s.ivwat2(i) = 0;
    }
    if (s.i14at1(i) <= 0) {
      s.i14at1(i) = 1;
      s.i14at2(i) = 0;
    }
  }
}

PairList expand_pairs(const State& s, int nml) {
  PairList p;
  int ifivr = s.ivrml1(nml), ntlvr = s.nvrml(nml);
  int i1s = s.imsml1(nml) + s.nmsml(nml);
  for (int io = ifivr + ntlvr - 1; io >= ifivr; --io) {
    int iv = s.iorvr(io);
    int i2s = i1s - 1;
    i1s = s.imsvr1(iv);
    for (int ims = i1s; ims <= i2s; ++ims)
      for (int i = s.latms1(ims); i <= s.latms2(ims); ++i) {
        for (int ivw = s.ivwat1(i); ivw <= s.ivwat2(i); ++ivw)
          for (int j = s.lvwat1(ivw); j <= s.lvwat2(ivw); ++j) {
            p.i.push_back(i);
            p.j.push_back(j);
            p.kind.push_back(1);
          }
        for (int k = s.i14at1(i); k <= s.i14at2(i); ++k) {
          p.i.push_back(i);
          p.j.push_back(s.l14at(k));
          p.kind.push_back(2);
        }
      }
  }
  return p;
}

// ------------------------------------------------------- init_molecule

void init_molecule(State& s, int iabin, const std::string& grpn, const std::string& grpc,
                   const std::string& seqfile, const std::string& varfile) {
  if (iabin != 1) throw Error("init_molecule> reading PDB files is not part of this engine (iabin must be 1)");
  if (s.reslib.empty()) throw Error("init_molecule> call init_energy first");
  int ontlml = s.ntlml > 0 ? s.ntlml + 1 : 1;
  std::string sf = rtrim(seqfile);
  if (sf.empty()) throw Error("init_molecule> no sequence file given");
  redseq(s, sf);
  int ntl = s.ntlml;
  for (int i = ontlml; i <= ntl; ++i) {
    getmol(s, i);
    for (int j = 1; j <= 6; ++j) s.gbpr(j, i) = 0.0;
    bldmol(s, i);
    s.ntlml = i;
    addend(s, i, grpn, grpc);
    setmvs(s, i);
    mklist(s, i);
  }
  std::string vf = rtrim(varfile);
  if (!vf.empty()) redvar_file(s, vf);
  s.nvr = 0;
  int ntlvr = s.nvars();
  for (int i = 1; i <= ntlvr; ++i) {
    if (!s.fxvr(i)) {
      ++s.nvr;
      s.idvr(s.nvr) = i;
    }
    int it = s.ityvr(i);
    if (it == 3) s.vlvr(i) = s.toat(s.iatvr(i));
    else if (it == 2) s.vlvr(i) = s.baat(s.iatvr(i));
    else if (it == 1) s.vlvr(i) = s.blat(s.iatvr(i));
    s.olvlvr(i) = s.vlvr(i);
  }
  s.ireg = 0;
  // step widths: every free torsion may move by up to 360 degrees
  for (int i = 1; i <= ntlvr; ++i) s.axvr(i) = (s.ityvr(i) == 3 && !s.fxvr(i)) ? pi2 : 0.0;
  if (s.itysol != 0) {
    int i1 = s.iatrs1(s.irsml1(1)), i2 = s.iatrs2(s.irsml2(s.ntlml));
    int its = std::abs(s.itysol);
    for (int i = i1; i <= i2; ++i) {
      int it = s.ityat(i);
      s.sigma(i) = s.coef_sl(its, it);
      s.rvdw(i) = s.rad_vdw(its, it);
      if (s.nmat.d[i - 1][0] != 'h') s.rvdw(i) = s.rvdw(i) + rwater;
    }
  }
}

void reset_state(State& s) {
  std::memset(static_cast<void*>(&s), 0, offsetof(State, reslib));
  s.reslib.clear();
  s.libdir.clear();
  for (int i = 1; i <= mxml; ++i) s.nmml.set(i, "");
  for (int i = 1; i <= mxrs; ++i) s.seq.set(i, "");
  for (int i = 1; i <= mxat; ++i) s.nmat.set(i, "");
  for (int i = 1; i <= mxvr; ++i) s.nmvr.set(i, "");
  std::memset(s.zimm, ' ', sizeof s.zimm);
  s.mti = 625;
  s.up2bmax = 4.0;  // BLOCK DATA updtchs (metropolis.f:221-229)
  s.up2bmin = 0.5;
  block_data(s);
}

}  // namespace smmp
