#include "devtopo.hpp"

#include <algorithm>
#include <string>

namespace smmp {

HostSys build_host_sys(const State& s) {
  if (s.ntlml != 1) throw Error("device engine: one chain per context (ntlml = 1)");
  HostSys h;
  const int nat = s.natoms(), nvt = s.nvars();
  if (nat < 4) throw Error("device engine: chain too short");
  if (nat > 4096) throw Error("device engine: more than 4096 atoms");
  h.nat = nat;
  h.nvt = nvt;
  h.words = (nat + 63) / 64;
  h.parent.assign(nat, -1);
  h.type.resize(nat);
  h.q.resize(nat); h.bl.resize(nat); h.cba.resize(nat); h.sba.resize(nat);
  h.toat0.resize(nat); h.tref.assign(nat, 0.0); h.tvar.assign(nat, -1);
  h.rvdw.resize(nat); h.sigma.resize(nat);
  for (int a = 1; a <= nat; ++a) {
    int i = a - 1;
    h.parent[i] = (a == 1) ? -1 : (int16_t)(s.iowat(a) - 1);
    if (a > 1 && !(s.iowat(a) >= 1 && s.iowat(a) < a))
      throw Error("device engine: atom " + std::to_string(a) + " precedes its predecessor");
    h.type[i] = (int8_t)(s.ityat(a) - 1);
    h.q[i] = s.cgat(a);
    h.bl[i] = s.blat(a);
    h.cba[i] = s.csbaat(a);
    h.sba[i] = s.snbaat(a);
    h.toat0[i] = s.toat(a);
    h.rvdw[i] = s.rvdw(a);
    h.sigma[i] = s.sigma(a);
    if (s.itysol != 0 && s.rvdw(a) > 0.0) h.heavy.push_back((int16_t)i);
  }
  h.nheavy = (int)h.heavy.size();
  h.heavy_of.assign(nat, (int16_t)-1);
  for (int k = 0; k < h.nheavy; ++k) h.heavy_of[h.heavy[k]] = (int16_t)k;
  // build levels: level(a) = level(parent(a)) + 1, first atom at level 0
  std::vector<int> lev(nat, 0);
  int maxlev = 0;
  for (int i = 1; i < nat; ++i) {
    lev[i] = lev[h.parent[i]] + 1;
    maxlev = std::max(maxlev, lev[i]);
  }
  h.nlev = maxlev;
  h.lev_off.assign(maxlev + 1, 0);
  for (int L = 1; L <= maxlev; ++L) {
    h.lev_off[L - 1] = (int)h.lev_atoms.size();
    for (int i = 1; i < nat; ++i)
      if (lev[i] == L) h.lev_atoms.push_back((int16_t)i);
  }
  h.lev_off[maxlev] = (int)h.lev_atoms.size();
  // first three backbone atoms (i1 is atom 1)
  int b1, b2, b3;
  first_backbone_atoms(s, 1, b1, b2, b3);
  if (b1 != 1) throw Error("device engine: the chain must start at atom 1");
  h.i2 = b2 - 1;
  h.i3 = b3 - 1;
  for (int k = 0; k < 6; ++k) h.gbpr[k] = s.gbpr(k + 1, 1);
  // variables: torsion terms, controlled atoms, moved atoms
  h.var_atom.resize(nvt); h.var_base.resize(nvt); h.var_axis0.resize(nvt);
  h.var_m1.resize(nvt); h.var_m2.resize(nvt); h.var_sign.resize(nvt);
  h.var_e0.resize(nvt); h.var_sg.resize(nvt); h.var_rn.resize(nvt); h.var_is_torsion.resize(nvt);
  std::vector<std::vector<int>> children(nat);
  for (int i = 1; i < nat; ++i) children[h.parent[i]].push_back(i);
  for (int v = 1; v <= nvt; ++v) {
    int k = v - 1;
    int iat = s.iatvr(v), ic = s.iclvr(v);
    h.var_atom[k] = (int16_t)(iat - 1);
    h.var_is_torsion[k] = s.ityvr(v) == 3;
    if (s.ityvr(v) != 3) throw Error("device engine: only torsion variables are supported");
    h.var_e0[k] = s.e0to(ic);
    h.var_sg[k] = s.sgto(ic);
    h.var_rn[k] = s.rnto(ic);
    int b = s.iowat(iat);  // base (1-based)
    h.var_base[k] = (int16_t)(b - 1);
    h.var_axis0[k] = (int16_t)(h.parent[b - 1]);
    if (h.var_axis0[k] < 0) throw Error("device engine: torsion about the first atom");
    // torsions set by this variable: the branches of the base (setvar.f:43-57)
    for (int j = 1; j <= s.nbdat(b); ++j) {
      int jat = s.ibdat(j, b);
      if (s.iowat(jat) != b) continue;
      if (h.tvar[jat - 1] >= 0) throw Error("device engine: two variables set the torsion of one atom");
      h.tvar[jat - 1] = (int16_t)k;
      h.tref[jat - 1] = s.toat(iat);
    }
    // moved atoms: every descendant of the base; must be one index range
    std::vector<int> stack(children[b - 1].begin(), children[b - 1].end());
    int m1 = nat, m2 = -1, cnt = 0;
    while (!stack.empty()) {
      int a = stack.back();
      stack.pop_back();
      m1 = std::min(m1, a);
      m2 = std::max(m2, a);
      ++cnt;
      for (int c : children[a]) stack.push_back(c);
    }
    if (cnt == 0 || m2 - m1 + 1 != cnt) throw Error("device engine: moved atoms of a variable are not contiguous");
    h.var_m1[k] = (int16_t)m1;
    h.var_m2[k] = (int16_t)m2;
    h.var_sign[k] = 1.0;
  }
  // The torsions of the second and third backbone atoms also set the frame of
  // the first atom (setgbl keeps bond 1->2 and the plane 1-2-3 fixed in the
  // laboratory), so changing them turns the atoms before the moved range the
  // other way while the moved range stays put.
  for (int k = 0; k < nvt; ++k) {
    if (h.tvar[h.i2] != k && h.tvar[h.i3] != k) continue;
    if (h.var_m2[k] != nat - 1 || h.var_m1[k] < 1)
      throw Error("device engine: a torsion of the first backbone atoms does not move the chain's end");
    h.var_m2[k] = (int16_t)(h.var_m1[k] - 1);
    h.var_m1[k] = 0;
    h.var_sign[k] = -1.0;
  }
  // heavy-index range of each moved range (heavy is sorted by atom)
  h.var_h1.resize(nvt);
  h.var_h2.resize(nvt);
  for (int k = 0; k < nvt; ++k) {
    const auto lo = std::lower_bound(h.heavy.begin(), h.heavy.end(), h.var_m1[k]);
    const auto hi = std::upper_bound(h.heavy.begin(), h.heavy.end(), h.var_m2[k]);
    h.var_h1[k] = (int16_t)(lo - h.heavy.begin());
    h.var_h2[k] = (int16_t)(hi - h.heavy.begin());
  }
  // interaction set from the reference lists (mklist), as symmetric bits
  h.bits.assign((size_t)nat * h.words * 2, 0ull);
  PairList pl = expand_pairs(s, 1);
  for (size_t p = 0; p < pl.i.size(); ++p) {
    int i = pl.i[p] - 1, j = pl.j[p] - 1;
    if (i == j) throw Error("device engine: self pair in the interaction lists");
    uint64_t& bij = h.bits[((size_t)i * h.words + j / 64) * 2 + 0];
    uint64_t m = 1ull << (j % 64);
    if (bij & m) throw Error("device engine: pair listed twice");
    bij |= m;
    h.bits[((size_t)j * h.words + i / 64) * 2 + 0] |= 1ull << (i % 64);
    if (pl.kind[p] == 2) {
      h.bits[((size_t)i * h.words + j / 64) * 2 + 1] |= m;
      h.bits[((size_t)j * h.words + i / 64) * 2 + 1] |= 1ull << (i % 64);
      ++h.n14;
    }
    ++h.npairs;
  }
  // parameter tables
  h.pA.resize(kTypes * kTypes); h.pA14.resize(kTypes * kTypes); h.pC.resize(kTypes * kTypes); h.pHB.resize(kTypes * kTypes);
  for (int ti = 1; ti <= kTypes; ++ti)
    for (int tj = 1; tj <= kTypes; ++tj) {
      int k = (ti - 1) * kTypes + (tj - 1);
      bool hb = s.ihbty(ti, tj) != 0;
      h.pHB[k] = hb;
      h.pA[k] = hb ? s.ahb(ti, tj) : s.aij(ti, tj);
      h.pA14[k] = hb ? s.ahb(ti, tj) : s.a14(ti, tj);
      h.pC[k] = hb ? s.chb(ti, tj) : s.cij(ti, tj);
    }
  h.conv = s.conv;
  h.plt = s.plt;
  h.slp = s.slp;
  h.epsd = s.epsd;
  h.itysol = s.itysol;
  if (s.itysol < 0) throw Error("device engine: the analytic solvent (itysol < 0) is not implemented");
  h.npnt = s.itysol > 0 ? s.npnt : 0;
  if (h.npnt > 128) throw Error("device engine: tessellations of more than 128 points are not supported");
  h.spoint.resize((size_t)h.npnt * 3);
  for (int p = 1; p <= h.npnt; ++p)
    for (int d = 1; d <= 3; ++d) h.spoint[(size_t)(p - 1) * 3 + (d - 1)] = s.spoint(p, d);
  return h;
}

}  // namespace smmp
