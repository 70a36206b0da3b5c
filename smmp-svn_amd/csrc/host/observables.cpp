// Observables of the current coordinates that the samplers' outputs carry.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

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

// valang (dihedr.f:73-118): the angle i1-i2-i3 at i2, in radians.
static double valang(const State& s, int i1, int i2, int i3) {
  const double h1 = s.xat(i2), h2 = s.yat(i2), h3 = s.zat(i2);
  const double x1 = s.xat(i1) - h1, x2 = s.yat(i1) - h2, x3 = s.zat(i1) - h3;
  const double y1 = s.xat(i3) - h1, y2 = s.yat(i3) - h2, y3 = s.zat(i3) - h3;
  const double x = x1 * x1 + x2 * x2 + x3 * x3, y = y1 * y1 + y2 * y2 + y3 * y3, u = x * y;
  if (u == 0.0) throw Error("valang> Error in coordinates");
  double a = (x1 * y1 + x2 * y2 + x3 * y3) / std::sqrt(u);
  a = std::min(std::max(a, -1.0), 1.0);
  return std::acos(a);
}

static double dist(const State& s, int i, int j) {
  const double dx = s.xat(i) - s.xat(j), dy = s.yat(i) - s.yat(j), dz = s.zat(i) - s.zat(j);
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

// ishybd (hbond.f:297-352): McDonald and Thornton's geometric criteria,
// H...A <= 2.5, D...A <= 3.9, angles D-H-A, H-A-B and D-A-B >= 110 degrees.
static bool ishybd(const State& s, int i, int j, int& ih, int& ia) {
  if (i <= 0 || j <= 0) return false;
  const int ihb = s.ihbty(s.ityat(i), s.ityat(j));
  if (ihb == 0) return false;
  if (ihb > 0) {
    ih = i;
    ia = j;
  } else {
    ih = j;
    ia = i;
  }
  const double cahb = 110.0 * cdr;
  if (dist(s, ih, ia) > 2.5) return false;
  const int id = s.iowat(ih);
  if (id <= 0 || dist(s, id, ia) > 3.9 || valang(s, id, ih, ia) < cahb) return false;
  const int ib = s.iowat(ia);
  return ib > 0 && valang(s, ih, ia, ib) >= cahb && valang(s, id, ia, ib) >= cahb;
}

// hbond (hbond.f:14-216): every van der Waals and 1-4 pair of molecule nml
// in the order of the energy's pair loop, H-bonds per (hydrogen, acceptor)
// type; mhb counts those whose donor residue lies more than four residues
// after the acceptor's.  The reference reads its type table with the index
// of the bond, not of the type (hbond.f:172-176), which this keeps.
int hbond(State& s, int nml) {
  if (nml < 1 || nml > s.ntlml)
    throw Error("hbond> nml must name a molecule (the reference leaves its moving-set index unset for 0)");
  constexpr int atbase = mxat;
  for (int i = 1; i <= State::mxtyhb; ++i) {
    s.nutyhb(i) = 0;
    s.ixtyhb(i) = 0;
  }
  s.ntyhb = 0;
  const int ntlvr = s.nvrml(nml);
  if (ntlvr == 0) return 0;
  const int ifivr = s.ivrml1(nml);
  int i1s = s.imsml1(nml) + s.nmsml(nml);
  auto record = [&](int i, int j) {
    int ih = 0, ia = 0;
    if (!ishybd(s, i, j, ih, ia)) return;
    const int ixhb = ih * atbase + ia;
    for (int k = 1; k <= s.ntyhb; ++k)
      if (ixhb == s.ixtyhb(k)) {
        s.nutyhb(k) += 1;
        return;
      }
    if (s.ntyhb >= State::mxtyhb) throw Error("hbond> increase parameter MXTYHB");
    s.ntyhb += 1;
    s.nutyhb(s.ntyhb) = 1;
    s.ixtyhb(s.ntyhb) = ixhb;
  };
  for (int io = ifivr + ntlvr - 1; io >= ifivr; --io) {
    const int iv = s.iorvr(io);
    const int i2s = i1s - 1;
    i1s = s.imsvr1(iv);
    for (int ims = i1s; ims <= i2s; ++ims)
      for (int i = s.latms1(ims); i <= s.latms2(ims); ++i) {
        for (int ivw = s.ivwat1(i); ivw <= s.ivwat2(i); ++ivw)
          for (int j = s.lvwat1(ivw); j <= s.lvwat2(ivw); ++j) record(i, j);
        for (int i14 = s.i14at1(i); i14 <= s.i14at2(i); ++i14) record(i, s.l14at(i14));
      }
  }
  int mhb = 0, ii = 0;
  for (int i = 1; i <= s.ntyhb; ++i)
    for (int j = 1; j <= s.nutyhb(i); ++j) {
      ++ii;
      if (ii > State::mxtyhb) throw Error("hbond> H-bond index beyond MXTYHB");
      const int ix = s.ixtyhb(ii);
      const int id = ix / atbase, ia = ix - id * atbase;
      const int nd = nursat(s, id), na = nursat(s, ia);  // stops (raises) on atom 0, as the reference
      if (nd - na > 4) ++mhb;
    }
  return mhb;
}

// interhbond (hbond.f:228-256): H-bonds between every pair of chains, all
// atom pairs; mmhb(i, j) = mmhb(j, i) per pair of chains.
int interhbond(State& s) {
  int mhb = 0;
  for (int iml = 1; iml <= s.ntlml; ++iml)
    for (int jml = iml + 1; jml <= s.ntlml; ++jml) {
      s.mmhb(iml, jml) = 0;
      for (int ires = s.irsml1(iml); ires <= s.irsml2(iml); ++ires)
        for (int jres = s.irsml1(jml); jres <= s.irsml2(jml); ++jres)
          for (int iat = s.iatrs1(ires); iat <= s.iatrs2(ires); ++iat)
            for (int jat = s.iatrs1(jres); jat <= s.iatrs2(jres); ++jat) {
              int ih = 0, ia = 0;
              if (ishybd(s, iat, jat, ih, ia)) {
                ++mhb;
                s.mmhb(iml, jml) += 1;
              }
            }
      s.mmhb(jml, iml) = s.mmhb(iml, jml);
    }
  return mhb;
}

// zimmer (zimmer.f:15-93): the Zimmerman code of each residue from its
// (phi, psi), '-' where a residue has no free phi; residues 1..nresi.
void zimmer(State& s, int nresi) {
  for (int j = 1; j <= nresi && j <= mxrs; ++j) s.zimm[j - 1] = '-';
  char zim = ' ';  // the reference keeps the last code when no window matches
  for (int i = 1; i <= s.nvr; ++i) {
    const int iv = s.idvr(i);
    if (s.nmvr.get(iv) != "phi") continue;
    const double xphi = s.vlvr(iv) * crd, xpsi = s.vlvr(s.idvr(i + 1)) * crd;
    if (xphi <= -110.0) {
      if (xpsi <= -140.0 || xpsi > 110.0) zim = 'E';
      if (xpsi <= -90.0 && xpsi > -140.0) zim = 'H';
      if (xpsi <= -40.0 && xpsi > -90.0) zim = 'G';
      if (xpsi <= 20.0 && xpsi > -40.0) zim = 'B';
      if (xpsi <= 110.0 && xpsi > 20.0) zim = 'D';
    } else if (xphi <= -40.0) {
      if (xpsi <= -140.0 || xpsi > 130.0) zim = 'F';
      if (xpsi <= -90.0 && xpsi > -140.0) zim = 'H';
      if (xpsi <= -10.0 && xpsi > -90.0) zim = 'A';
      if (xpsi <= 50.0 && xpsi > -10.0) zim = 'B';
      if (xpsi <= 130.0 && xpsi > 50.0) zim = 'C';
    } else if (xphi <= 0.0) {
      zim = 'H';
    } else if (xphi <= 40.0) {
      zim = 'h';
    } else if (xphi <= 110.0) {
      if (xpsi <= -130.0 || xpsi > 140.0) zim = 'f';
      if (xpsi <= 140.0 && xpsi > 90.0) zim = 'h';
      if (xpsi <= 90.0 && xpsi > 10.0) zim = 'a';
      if (xpsi <= 10.0 && xpsi > -50.0) zim = 'b';
      if (xpsi <= -50.0 && xpsi > -130.0) zim = 'c';
    } else {
      if (xpsi <= -110.0 || xpsi > 140.0) zim = 'e';
      if (xpsi <= 140.0 && xpsi > 90.0) zim = 'h';
      if (xpsi <= 90.0 && xpsi > 40.0) zim = 'g';
      if (xpsi <= 40.0 && xpsi > -20.0) zim = 'b';
      if (xpsi <= -20.0 && xpsi > -110.0) zim = 'd';
    }
    const int nres = nursvr(s, iv);
    if (nres >= 1 && nres <= mxrs) s.zimm[nres - 1] = zim;
  }
}

// outpdb (outpdb.f:12-185): the coordinates of molecule nml (0: all) in PDB
// format, hydrogens with their trailing digit moved to the front of the
// name, TER after each chain, CONECT records of the bonds i-j with i < j.
// Eigenvalues d and eigenvectors (columns of v) of a symmetric 4x4 matrix
// by cyclic Jacobi rotations, in the steps of rmsdfun.f's jacobi (the
// sweeps stop once the off-diagonal magnitudes sum to at most 1e-6; the
// upper triangle of a is destroyed).
static void jacobi4(double a[4][4], double d[4], double v[4][4]) {
  constexpr int n = 4;
  double b[n], z[n];
  for (int p = 0; p < n; ++p) {
    b[p] = d[p] = a[p][p];
    z[p] = 0.0;
    for (int q = 0; q < n; ++q) v[p][q] = 0.0;
    v[p][p] = 1.0;
  }
  for (int it = 1; it <= 500; ++it) {
    double sm = 0.0;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) sm += std::fabs(a[p][q]);
    if (sm <= 1e-6) return;
    const double tresh = it < 4 ? 0.2 * sm / (n * n) : 0.0;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double g = 100.0 * std::fabs(a[p][q]);
        if (it > 4 && std::fabs(d[p]) + g == std::fabs(d[p]) && std::fabs(d[q]) + g == std::fabs(d[q])) {
          a[p][q] = 0.0;
        } else if (std::fabs(a[p][q]) > tresh) {
          double h = d[q] - d[p], t;
          if (std::fabs(h) + g == std::fabs(h)) {
            t = a[p][q] / h;
          } else {
            const double theta = 0.5 * h / a[p][q];
            t = 1.0 / (std::fabs(theta) + std::sqrt(1.0 + theta * theta));
            if (theta < 0.0) t = -t;
          }
          const double c = 1.0 / std::sqrt(1.0 + t * t), sn = t * c, tau = sn / (1.0 + c);
          h = t * a[p][q];
          z[p] -= h;
          z[q] += h;
          d[p] -= h;
          d[q] += h;
          a[p][q] = 0.0;
          auto rot = [&](double& x, double& y) {
            const double gx = x, hy = y;
            x = gx - sn * (hy + gx * tau);
            y = hy + sn * (gx - hy * tau);
          };
          for (int j = 0; j < p; ++j) rot(a[j][p], a[j][q]);
          for (int j = p + 1; j < q; ++j) rot(a[p][j], a[j][q]);
          for (int j = q + 1; j < n; ++j) rot(a[p][j], a[q][j]);
          for (int j = 0; j < n; ++j) rot(v[j][p], v[j][q]);
        }
      }
    for (int p = 0; p < n; ++p) {
      b[p] += z[p];
      d[p] = b[p];
      z[p] = 0.0;
    }
  }
  throw Error("jacobi> too many iterations");
}

double rmsdfun(const State& s, int nml, int ir1, int ir2, const int32_t* ixat, const double* xrf, const double* yrf,
               const double* zrf, int isl) {
  if (nml < 1 || nml > s.ntlml) throw Error("rmsdopt> Sorry, there is no molecule #" + std::to_string(nml));
  // atom selection of rmsdopt (rmsdfun.f:64-99): residues ir1..ir2 of chain
  // nml (counted within the chain), atoms with a reference (ixat of the
  // running atom number > 0), no hydrogens; isl 1 backbone n, ca, c, 2 ca
  std::vector<double> x1, x2;
  int na = 0;
  for (int im = 1; im <= s.ntlml; ++im) {
    int nr = 0;
    for (int ir = s.irsml1(im); ir <= s.irsml2(im); ++ir) {
      if (im == nml) ++nr;
      for (int ia = s.iatrs1(ir); ia <= s.iatrs2(ir); ++ia) {
        ++na;
        if (im != nml || nr < ir1 || nr > ir2) continue;
        const int ix = ixat[na - 1];
        const std::string nm = s.nmat.get(ia);
        if (ix <= 0 || nm[0] == 'h') continue;
        const bool take = isl == 0 || (isl == 1 && (nm == "n   " || nm == "ca  " || nm == "c   ")) ||
                          (isl == 2 && nm == "ca  ");
        if (!take) continue;
        x1.insert(x1.end(), {s.xat(ia), s.yat(ia), s.zat(ia)});
        x2.insert(x2.end(), {xrf[ix - 1], yrf[ix - 1], zrf[ix - 1]});
      }
    }
  }
  const int n = (int)x1.size() / 3;
  if (n < 3) throw Error("rmsdopt> <3 atoms selected !");
  // fitmol (rmsdfun.f:104-175): the smallest eigenvalue of Kearsley's
  // quaternion matrix of the centred sets is n rmsd^2
  const double dn = (double)n;
  double a1[3] = {0.0, 0.0, 0.0}, a2[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < n; ++j) {
      a1[i] += x1[3 * j + i];
      a2[i] += x2[3 * j + i];
    }
    a1[i] /= dn;
    a2[i] /= dn;
  }
  double q[4][4] = {};
  for (int j = 0; j < n; ++j) {
    double dm[3], dp[3];
    for (int i = 0; i < 3; ++i) {
      dm[i] = x1[3 * j + i] - a1[i];
      dp[i] = x2[3 * j + i] - a2[i];
    }
    const double xm = dp[0] - dm[0], ym = dp[1] - dm[1], zm = dp[2] - dm[2];
    const double xp = dp[0] + dm[0], yp = dp[1] + dm[1], zp = dp[2] + dm[2];
    q[0][0] += xm * xm + ym * ym + zm * zm;
    q[0][1] += yp * zm - ym * zp;
    q[0][2] += xm * zp - xp * zm;
    q[0][3] += xp * ym - xm * yp;
    q[1][1] += yp * yp + zp * zp + xm * xm;
    q[1][2] += xm * ym - xp * yp;
    q[1][3] += xm * zm - xp * zp;
    q[2][2] += xp * xp + zp * zp + ym * ym;
    q[2][3] += ym * zm - yp * zp;
    q[3][3] += xp * xp + yp * yp + zm * zm;
  }
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 4; ++j) q[j][i] = q[i][j];
  double e[4], v[4][4];
  jacobi4(q, e, v);
  double em = e[0];
  for (int i = 1; i < 4; ++i) em = std::min(em, e[i]);
  return std::sqrt(em / dn);
}

void outpdb(const State& s, int nml, const std::string& path) {
  if (nml < 0 || nml > s.ntlml) throw Error("outpdb> No such molecule #" + std::to_string(nml));
  const int im1 = nml > 0 ? nml : 1, im2 = nml > 0 ? nml : s.ntlml;
  FILE* f = path.empty() ? stdout : std::fopen(path.c_str(), "w");
  if (!f) throw Error("outpdb> cannot open " + path);
  auto line = [&](const std::string& l) { std::fprintf(f, "%-80s\n", l.c_str()); };
  auto upper = [](std::string t) {
    for (auto& ch : t) ch = (char)std::toupper((unsigned char)ch);
    return t;
  };
  char buf[128];
  int iat = 0;
  std::string res = "   ";
  for (int iml = im1; iml <= im2; ++iml) {
    const char chid = s.ntlml == 1 ? ' ' : (char)(64 + iml);
    int irs = 0;
    for (int nrs = s.irsml1(iml); nrs <= s.irsml2(iml); ++nrs) {
      ++irs;
      res = s.seq.get(nrs).substr(0, 3);
      const char* linty = (res != "ace" && res != "nme") ? "ATOM  " : "HETATM";
      for (int i = s.iatrs1(nrs); i <= s.iatrs2(nrs); ++i) {
        ++iat;
        std::string atnm = " " + s.nmat.get(i);  // 5 characters
        if (atnm[1] == 'h') {  // hydrogens by PDB convention: a trailing digit goes first
          int j = 4;
          while (j > 0 && atnm[j] == ' ') --j;
          if (atnm[j] >= '0' && atnm[j] <= '9') {
            atnm[0] = atnm[j];
            atnm[j] = ' ';
          }
        }
        std::snprintf(buf, sizeof buf, "%-6s%5d %-5s%-3s %c%4d%c   %8.3f%8.3f%8.3f %5.2f %5.2f", linty, iat,
                      upper(atnm).c_str(), upper(res).c_str(), chid, irs, ' ', s.xat(i), s.yat(i), s.zat(i), 1.0,
                      0.0);
        line(buf);
      }
    }
    ++iat;
    std::snprintf(buf, sizeof buf, "%-6s%5d      %-3s %c", "TER   ", iat, upper(res).c_str(), chid);
    line(buf);
  }
  for (int iml = im1; iml <= im2; ++iml) {
    const int ifirs = s.irsml1(iml);
    const int base = nml > 0 ? s.iatrs1(ifirs) - 1 : 1 - iml;
    for (int nrs = ifirs; nrs <= s.irsml2(iml); ++nrs) {
      const int nfi = s.iatrs1(nrs);
      for (int i = nfi; i <= s.iatrs2(nrs); ++i) {
        if (s.nbdat(i) <= 0) continue;
        int ibd[4], jj;
        if (nrs == ifirs && i == nfi) {
          ibd[0] = s.iowat(i);
          ibd[1] = s.ibdat(1, i);
          ibd[2] = s.ibdat(2, i);
          ibd[3] = s.ibdat(3, i);
          jj = 4;
        } else {
          ibd[0] = s.ibdat(1, i);
          ibd[1] = s.ibdat(2, i);
          ibd[2] = s.ibdat(3, i);
          jj = 3;
        }
        int nbd = 0;
        for (int j = 0; j < jj; ++j)
          if (ibd[j] > i) ibd[nbd++] = ibd[j];
        if (nbd == 0) continue;
        int n = std::snprintf(buf, sizeof buf, "CONECT%5d", i - base);
        for (int j = 0; j < nbd; ++j) n += std::snprintf(buf + n, sizeof buf - n, "%5d", ibd[j] - base);
        line(buf);
      }
    }
  }
  line("END");
  if (f != stdout) std::fclose(f);
}

}  // namespace smmp
