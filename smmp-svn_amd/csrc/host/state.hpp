// Host-side state of one SMMP system: the arrays the reference keeps in its
// COMMON blocks (INCL.H), stored with the same capacities and the same
// column-major layout so the Python shim can expose them as numpy views with
// the shapes f2py gives them (smmp.pyf).  Indices follow the reference's
// 1-based convention through the F1/F2 accessors; xat/yat/zat keep their
// lower bound -1 (INCL.H:237, virtual atoms 0 and -1 are used by rplgrp).
#pragma once
#include <atomic>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace smmp {

// Capacities (INCL.H:14-36, 66-70).
constexpr int mxml = 10, mxrs = 100, mxat = 2600, mxbd = 3, mxath = 100;
constexpr int mxvr = mxrs * 8, mxms = mxvr * 3, mxvw = mxat * 4, mx14 = mxat * 4;
constexpr int mxvrh = mxath, mxtyat = 18, mxhbdo = 4, mxhbac = 6, mxtyto = 19;
constexpr int mxpoint = 2000, mxtysol = 9;
constexpr double pi = 3.141592653589793, pi2 = 2.0 * pi, pi4 = 4.0 * pi;
constexpr double crd = 57.2957795130823, cdr = 1.745329251994330e-2;
// The reference declares rwater double precision but initialises it from a
// default-real literal (INCL.H:70), so its value is the float 1.4.
constexpr double rwater = (double)1.4f;

// 1-based (or lower-bound LB) view on a fixed array.
template <typename T, int N, int LB = 1>
struct F1 {
  T d[N - LB + 1];
  T& operator()(int i) { return d[i - LB]; }
  const T& operator()(int i) const { return d[i - LB]; }
  T* data() { return d; }
  static constexpr int size = N - LB + 1;
};
// Column-major 2-D array with 1-based indices (Fortran a(N1,N2)).
template <typename T, int N1, int N2>
struct F2 {
  T d[N1 * N2];
  T& operator()(int i, int j) { return d[(i - 1) + (j - 1) * N1]; }
  const T& operator()(int i, int j) const { return d[(i - 1) + (j - 1) * N1]; }
  T* data() { return d; }
};
// Fixed-width blank-padded Fortran strings, stored as char[L] per element.
template <int L, int N>
struct FStr {
  char d[N][L];
  std::string get(int i) const {  // 1-based
    return std::string(d[i - 1], L);
  }
  void set(int i, const std::string& s) {
    for (int k = 0; k < L; ++k) d[i - 1][k] = k < (int)s.size() ? s[k] : ' ';
  }
  char* data() { return &d[0][0]; }
};

// Everything the energy, the moves and the Python layer read or write.
// A fresh number for every topology a State takes on (process-wide, never
// reused), so device copies of a system can be matched to what they were
// built from.
inline uint64_t next_generation() {
  static std::atomic<uint64_t> g{0};
  return ++g;
}

struct State {
  uint64_t gen = next_generation();  // renewed by init_energy and init_molecule
  // /epar_l/ flex, tesgrd, sh2, epsd (Fortran logicals as int32), ientyp
  int32_t flex = 0, tesgrd = 0, sh2 = 0, epsd = 0, ientyp = 0;
  // /isolty/
  int32_t itysol = 0, isolscl = 0;
  // /updchois/
  int32_t rndord = 0, upchswitch = 0;
  double bgsprob = 0.0;
  // /bet/
  double beta = 0.0;
  // /updstats/ trials and accepted trials per update type (accanalyze,
  // metropolis.f:186-193: 1 torsion, 2 BGS, 3 shift, 4 rotation, 5 all);
  // /updtparam/ the beta range of updtch2 (BLOCK DATA, metropolis.f:221-229)
  F1<int32_t, 5> ncalls, nacalls;
  double up2bmax = 4.0, up2bmin = 0.5;
  // /mol_par/
  int32_t nvr = 0;
  // /mol_i/
  int32_t ntlml = 0;
  F1<int32_t, mxml> irsml1, irsml2, ivrml1, nvrml, imsml1, nmsml, iadml1, nadml;
  F2<int32_t, 3, mxml> ixrfpt;
  int32_t ireg = 0;
  // /mol_c/
  FStr<80, mxml> nmml;
  // /mol_r/
  F2<double, 3, mxml> rfpt;
  F2<double, 6, mxml> gbpr;
  F2<double, 3, mxml> xrfax, yrfax, zrfax;
  // /res_i/, /res_c/
  F1<int32_t, mxrs> iatrs1, iatrs2, ixatrs, ivrrs1, nvrrs;
  FStr<4, mxrs> seq;
  // /atm_i/
  F1<int32_t, mxat> ityat, iowat, iyowat, nbdat, ixmsat;
  F2<int32_t, mxbd, mxat> ibdat, iybdat;
  // /atm_r/
  F1<double, mxat> cgat, xbaat, ybaat, zbaat, baat, snbaat, csbaat;
  F1<double, mxat> xtoat, ytoat, ztoat, toat, sntoat, cstoat, blat;
  F1<double, mxat, -1> xat, yat, zat;
  // /atm_c/
  FStr<4, mxat> nmat;
  // /var_r/, /var_i/, /var_c/, /var_l/
  F1<double, mxvr> olvlvr, vlvr, axvr;
  F1<int32_t, mxvr> iatvr, ityvr, iclvr, iorvr, imsvr1, imsvr2, iadvr1, iadvr2, ladvr, idvr;
  F1<int32_t, mxms> latms1, latms2;
  FStr<3, mxvr> nmvr;
  F1<int32_t, mxvr> fxvr;
  // /epar_i/, /epar_r/
  F2<int32_t, mxtyat, mxtyat> ihbty;
  F2<double, mxtyat, mxtyat> cij, aij, a14, chb, ahb;
  F1<double, mxtyto> e0to, sgto, rnto, esnto;
  double conv = 0.0, plt = 0.0, slp = 0.0, cohb = 0.0;
  // /eny_i/
  F1<int32_t, mxml> nvwml, ivwml1, n14ml, i14ml1;
  F1<int32_t, mxat> ivwat1, ivwat2, i14at1, i14at2;
  F1<int32_t, mxvw> lvwat1, lvwat2, ixatvw;
  F1<int32_t, mx14> l14at, ixat14;
  // /eny_r/ (energies of the last evaluation)
  double eyel = 0, eyvw = 0, eyhb = 0, eyvr = 0, eysm = 0, eysl = 0, eyrg = 0, eyslh = 0, eyslp = 0, eyab = 0;
  double eysmi = 0, eyeli = 0, eyvwi = 0, eyhbi = 0;  // between molecules (eninteract.f)
  F1<double, mxvr> gdeyvr, gdeysl;                    // derivatives by the variables (gradient.f)
  double boxsize = 0;                                 // COMMON /bounds/ (metropolis.f:100-106)
  // /solvent/, /tes/, /pnt/, /sol/, /v_sol/, /ressurf/
  F1<double, mxat> rvdw, sigma;
  F2<double, mxpoint, 3> spoint;
  int32_t npnt = 0;
  F2<double, mxtysol, mxtyat> coef_sl, rad_vdw;
  double vdvol = 0, asa = 0, asap = 0, asah = 0;
  F1<double, mxrs> surfres;
  // /h_bond/ ntyhb, ixtyhb, nutyhb (hbond.f), mmhb (interhbond), /zimme/ zimm,
  // /myblock/ mt state
  static constexpr int mxtyhb = 100;
  int32_t ntyhb = 0;
  F1<int32_t, mxtyhb> ixtyhb, nutyhb;
  F2<int32_t, mxml, mxml> mmhb;
  char zimm[mxrs];
  int32_t mti = 625;
  F1<int32_t, 623, 0> mt;
  // /sche_r/ ECEPP parameter inputs (BLOCK DATA, init_energy.f:315-610)
  double ehm = 0, eps_s = 0;
  F1<double, mxtyat> atpl, efel, emin, rmin;
  F2<double, mxhbdo, mxhbac> chb_s, ahb_s;
  F1<double, mxtyto> e0to_s, sgto_s, rnto_s;
  F1<int32_t, mxtyat> do_s, ac_s;
  // library paths (/fil_c/)
  std::string reslib, libdir;

  int natoms() const { return ntlml > 0 ? iatrs2(irsml2(ntlml)) : 0; }
  int nvars() const { return ntlml > 0 ? ivrml1(ntlml) + nvrml(ntlml) - 1 : 0; }
};

// Error with a message; the C-ABI turns it into a status code (the reference
// stops the process instead).
struct Error {
  std::string msg;
  explicit Error(std::string m) : msg(std::move(m)) {}
};

void reset_state(State& s);
void block_data(State& s);                  // init_energy.f:315-610
void setpar(State& s);                      // init_energy.f:106-313
void init_energy(State& s, const std::string& libdir);  // init_energy.f:14-103
void init_molecule(State& s, int iabin, const std::string& grpn, const std::string& grpc,
                   const std::string& seqfile, const std::string& varfile);  // init_molecule.f:13-219
void setvar(State& s, int nml, const double* vlvrx);  // setvar.f:13-73 (1-based vlvrx(1..))
void rgyr(const State& s, int nml, double* rgy, double* ee);  // rgyr.f:13-140 (rgy, ee)
void helix(const State& s, int32_t out[4]);                   // helix.f:13-77 (nhel, mhel, nbet, mbet)
int hbond(State& s, int nml);                                 // hbond.f:14-216 (mhb; ntyhb, ixtyhb, nutyhb)
int interhbond(State& s);                                     // hbond.f:228-256 (mhb; mmhb)
void zimmer(State& s, int nresi);                             // zimmer.f:15-93 (zimm)
void outpdb(const State& s, int nml, const std::string& path);  // outpdb.f:12-185 ('' = stdout)
// rmsdfun.f:12-261: rmsd of chain nml's residues ir1..ir2 after the optimal
// superposition on the reference (ixat: reference index per atom, 1-based)
double rmsdfun(const State& s, int nml, int ir1, int ir2, const int32_t* ixat, const double* xrf, const double* yrf,
               const double* zrf, int isl);
void bldmol(State& s, int nml);             // bldmol.f:13-93
// fnd3ba (bldmol.f:95-173) on the current topology; ixrfpt keeps the atoms
// of the last bldmol call, which precede addend's renumbering after init.
void first_backbone_atoms(const State& s, int nml, int& i1, int& i2, int& i3);
void redvar_file(State& s, const std::string& varfile);  // redvar.f:13-648
int nursvr(const State& s, int ivr);        // nursvr.f:12-37
int nursat(const State& s, int iat);        // nursvr.f:40-76
double difang(double a1, double a2);        // difang.f:12-36
double addang(double a1, double a2);        // difang.f:38-63
void write_var_file(const State& s, int nml, const std::string& path);  // outvar.f:12-121

// Minimal checker-free consistency: the expanded pair list of one molecule
// as (i, j, kind) with kind 1 = van der Waals domain pair, 2 = 1-4 pair, in
// the summation order of enyshe.f:63-184.
struct PairList {
  std::vector<int32_t> i, j;
  std::vector<int8_t> kind;
};
PairList expand_pairs(const State& s, int nml);

}  // namespace smmp
