#include "devtopo.hpp"

#include <algorithm>
#include <string>

namespace smmp {

HostSys build_host_sys(const State& s) {
  if (s.ntlml < 1 || s.ntlml > mxml) throw Error("device engine: no chain, or more chains than mxml");
  HostSys h;
  const int nat = s.natoms(), nvt = s.nvars();
  if (nat > 4096) throw Error("device engine: more than 4096 atoms");
  h.nat = nat;
  h.nvt = nvt;
  h.words = (nat + 63) / 64;
  h.nml = s.ntlml;
  h.boxsize = s.boxsize;
  {  // ind_alf of c_alfa: the last atom of the residue whose name starts with "ca"
    auto ca_of = [&](int res) {
      int ia = -1;
      for (int a = s.iatrs1(res); a <= s.iatrs2(res); ++a)
        if (s.nmat.get(a).compare(0, 2, "ca") == 0) ia = a - 1;
      return (int16_t)ia;
    };
    h.ca_first = ca_of(s.irsml1(1));
    h.ca_last = ca_of(s.irsml2(s.ntlml));
  }
  h.mol_of.assign(nat, 0);
  for (int m = 1; m <= h.nml; ++m) {
    const int a1 = s.iatrs1(s.irsml1(m)), a2 = s.iatrs2(s.irsml2(m));
    if (a2 - a1 + 1 < 4) throw Error("device engine: chain " + std::to_string(m) + " too short");
    if (a1 != (m == 1 ? 1 : h.mol_a2[m - 2] + 2)) throw Error("device engine: chains are not numbered in order");
    h.mol_a1[m - 1] = (int16_t)(a1 - 1);
    h.mol_a2[m - 1] = (int16_t)(a2 - 1);
    for (int a = a1; a <= a2; ++a) h.mol_of[a - 1] = (int16_t)(m - 1);
    for (int k = 0; k < 6; ++k) h.gbpr[6 * (m - 1) + k] = s.gbpr(k + 1, m);
    int b1, b2, b3;
    first_backbone_atoms(s, m, b1, b2, b3);
    if (b1 != a1) throw Error("device engine: a chain must start with its first backbone atom");
    h.mol_i2[m - 1] = (int16_t)(b2 - 1);
    h.mol_i3[m - 1] = (int16_t)(b3 - 1);
  }
  h.parent.assign(nat, -1);
  h.type.resize(nat);
  h.q.resize(nat); h.bl.resize(nat); h.cba.resize(nat); h.sba.resize(nat);
  h.toat0.resize(nat); h.tref.assign(nat, 0.0); h.tvar.assign(nat, -1);
  h.rvdw.resize(nat); h.sigma.resize(nat);
  for (int a = 1; a <= nat; ++a) {
    int i = a - 1;
    const int a1 = h.mol_a1[h.mol_of[i]] + 1;  // first atom of this chain (1-based)
    h.parent[i] = (a == a1) ? -1 : (int16_t)(s.iowat(a) - 1);
    if (a > a1 && !(s.iowat(a) >= a1 && s.iowat(a) < a))
      throw Error("device engine: atom " + std::to_string(a) + " precedes its predecessor");
    h.type[i] = (int8_t)(s.ityat(a) - 1);
    h.q[i] = s.cgat(a);
    h.bl[i] = s.blat(a);
    h.cba[i] = s.csbaat(a);
    h.sba[i] = s.snbaat(a);
    h.toat0[i] = s.toat(a);
    h.rvdw[i] = s.rvdw(a);
    h.sigma[i] = s.sigma(a);
    if (s.itysol > 0 && s.rvdw(a) > 0.0) h.heavy.push_back((int16_t)i);  // the numerical solvent's atoms
  }
  h.nheavy = (int)h.heavy.size();
  h.heavy_of.assign(nat, (int16_t)-1);
  for (int k = 0; k < h.nheavy; ++k) h.heavy_of[h.heavy[k]] = (int16_t)k;
  for (int m = 0; m < h.nml; ++m) {
    const auto lo = std::lower_bound(h.heavy.begin(), h.heavy.end(), h.mol_a1[m]);
    const auto hi = std::upper_bound(h.heavy.begin(), h.heavy.end(), h.mol_a2[m]);
    h.mol_h1[m] = (int16_t)(lo - h.heavy.begin());
    h.mol_h2[m] = (int16_t)(hi - h.heavy.begin());
  }
  // build levels: level(a) = level(parent(a)) + 1, the first atom of every
  // chain at level 0, so the chains are built side by side
  std::vector<int> lev(nat, 0);
  int maxlev = 0;
  for (int i = 0; i < nat; ++i) {
    if (h.parent[i] < 0) continue;
    lev[i] = lev[h.parent[i]] + 1;
    maxlev = std::max(maxlev, lev[i]);
  }
  h.nlev = maxlev;
  h.lev_off.assign(maxlev + 1, 0);
  for (int L = 1; L <= maxlev; ++L) {
    h.lev_off[L - 1] = (int)h.lev_atoms.size();
    for (int i = 0; i < nat; ++i)
      if (h.parent[i] >= 0 && lev[i] == L) h.lev_atoms.push_back((int16_t)i);
  }
  h.lev_off[maxlev] = (int)h.lev_atoms.size();
  // variables: torsion terms, controlled atoms, moved atoms
  h.var_atom.resize(nvt); h.var_base.resize(nvt); h.var_axis0.resize(nvt);
  h.var_m1.resize(nvt); h.var_m2.resize(nvt); h.var_sign.resize(nvt);
  h.var_e0.resize(nvt); h.var_sg.resize(nvt); h.var_rn.resize(nvt); h.var_is_torsion.resize(nvt);
  std::vector<std::vector<int>> children(nat);
  for (int i = 0; i < nat; ++i)
    if (h.parent[i] >= 0) children[h.parent[i]].push_back(i);
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
  // The torsions of the second and third backbone atoms of a chain also set
  // the frame of its first atom (setgbl keeps bond 1->2 and the plane 1-2-3
  // fixed in the laboratory), so changing them turns the chain's atoms
  // before the moved range the other way while the moved range stays put.
  for (int k = 0; k < nvt; ++k) {
    const int m = h.mol_of[h.var_base[k]];
    if (h.tvar[h.mol_i2[m]] != k && h.tvar[h.mol_i3[m]] != k) continue;
    if (h.var_m2[k] != h.mol_a2[m] || h.var_m1[k] <= h.mol_a1[m])
      throw Error("device engine: a torsion of the first backbone atoms does not move the chain's end");
    h.var_m2[k] = (int16_t)(h.var_m1[k] - 1);
    h.var_m1[k] = h.mol_a1[m];
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
  // interaction set from the reference lists (mklist) of each chain, as
  // symmetric bits, plus every pair between chains (eninteract.f:29-80)
  h.bits.assign((size_t)nat * h.words * 2, 0ull);
  PairList pl;
  for (int m = 1; m <= h.nml; ++m) {
    PairList pm = expand_pairs(s, m);
    pl.i.insert(pl.i.end(), pm.i.begin(), pm.i.end());
    pl.j.insert(pl.j.end(), pm.j.begin(), pm.j.end());
    pl.kind.insert(pl.kind.end(), pm.kind.begin(), pm.kind.end());
  }
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
    if (h.mol_of[i] != h.mol_of[j]) throw Error("device engine: a list pair joins two chains");
    ++h.npairs;
  }
  for (int i = 0; i < nat; ++i)
    for (int j = h.mol_a2[h.mol_of[i]] + 1; j < nat; ++j) {
      h.bits[((size_t)i * h.words + j / 64) * 2 + 0] |= 1ull << (j % 64);
      h.bits[((size_t)j * h.words + i / 64) * 2 + 0] |= 1ull << (i % 64);
      ++h.ninter;
    }
  h.npairs += h.ninter;
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
  h.npnt = s.itysol > 0 ? s.npnt : 0;
  if (h.npnt > 128) throw Error("device engine: tessellations of more than 128 points are not supported");
  if (s.itysol > 0 && h.npnt < 1) throw Error("device engine: the numerical solvent needs tessellation points");
  h.spoint.resize((size_t)h.npnt * 3);
  for (int p = 1; p <= h.npnt; ++p)
    for (int d = 1; d <= 3; ++d) h.spoint[(size_t)(p - 1) * 3 + (d - 1)] = s.spoint(p, d);
  return h;
}

}  // namespace smmp
