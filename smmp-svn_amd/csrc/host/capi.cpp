// C-ABI: context management, COMMON-block views and the host-side entry
// points.  The GPU entry points (energy, metropolis, samplers) live in
// device/engine.hip and are reached through engine.hpp.
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/smmp_mi355x.h"
#include "../device/mlp_common.hpp"
#include "engine.hpp"
#include "minim.hpp"
#include "minimize.hpp"
#include "state.hpp"

using namespace smmp;

struct smmp_ctx {
  State s;
  std::string err;
  Engine* eng = nullptr;  // device mirror, created lazily
};

namespace {

struct ViewDesc {
  const char* name;
  int dtype;
  int ndim;
  int64_t shape[2];
  int64_t lbound[2];
  void* (*ptr)(State&);
};

#define V1(block, f, dt, n, lb) {#block "." #f, dt, 1, {n, 0}, {lb, 0}, [](State& s) -> void* { return (void*)s.f.data(); }}
#define V2(block, f, dt, n1, n2) {#block "." #f, dt, 2, {n1, n2}, {1, 1}, [](State& s) -> void* { return (void*)s.f.data(); }}
#define VS(block, f, dt) {#block "." #f, dt, 0, {1, 0}, {1, 0}, [](State& s) -> void* { return (void*)&s.f; }}
#define VC(block, f, L, n) {#block "." #f, SMMP_CHR, 2, {L, n}, {1, 1}, [](State& s) -> void* { return (void*)s.f.data(); }}

const ViewDesc kViews[] = {
    VS(epar_l, flex, SMMP_I32), VS(epar_l, tesgrd, SMMP_I32), VS(epar_l, sh2, SMMP_I32),
    VS(epar_l, epsd, SMMP_I32), VS(epar_l, ientyp, SMMP_I32),
    VS(isolty, itysol, SMMP_I32), VS(isolty, isolscl, SMMP_I32),
    VS(updchois, rndord, SMMP_I32), VS(updchois, upchswitch, SMMP_I32), VS(updchois, bgsprob, SMMP_F64),
    VS(bet, beta, SMMP_F64),
    V1(updstats, ncalls, SMMP_I32, 5, 1), V1(updstats, nacalls, SMMP_I32, 5, 1),
    VS(updtparam, up2bmax, SMMP_F64), VS(updtparam, up2bmin, SMMP_F64),
    VS(mol_par, nvr, SMMP_I32),
    VS(mol_i, ntlml, SMMP_I32), V1(mol_i, irsml1, SMMP_I32, mxml, 1), V1(mol_i, irsml2, SMMP_I32, mxml, 1),
    V1(mol_i, ivrml1, SMMP_I32, mxml, 1), V1(mol_i, nvrml, SMMP_I32, mxml, 1), V1(mol_i, imsml1, SMMP_I32, mxml, 1),
    V1(mol_i, nmsml, SMMP_I32, mxml, 1), V2(mol_i, ixrfpt, SMMP_I32, 3, mxml), VS(mol_i, ireg, SMMP_I32),
    V1(mol_i, iadml1, SMMP_I32, mxml, 1), V1(mol_i, nadml, SMMP_I32, mxml, 1),
    VC(mol_c, nmml, 80, mxml),
    V2(mol_r, gbpr, SMMP_F64, 6, mxml), V2(mol_r, rfpt, SMMP_F64, 3, mxml),
    V1(res_i, iatrs1, SMMP_I32, mxrs, 1), V1(res_i, iatrs2, SMMP_I32, mxrs, 1), V1(res_i, ixatrs, SMMP_I32, mxrs, 1),
    V1(res_i, ivrrs1, SMMP_I32, mxrs, 1), V1(res_i, nvrrs, SMMP_I32, mxrs, 1),
    VC(res_c, seq, 4, mxrs),
    V1(atm_i, ityat, SMMP_I32, mxat, 1), V1(atm_i, iowat, SMMP_I32, mxat, 1), V1(atm_i, iyowat, SMMP_I32, mxat, 1),
    V1(atm_i, nbdat, SMMP_I32, mxat, 1), V2(atm_i, ibdat, SMMP_I32, mxbd, mxat), V2(atm_i, iybdat, SMMP_I32, mxbd, mxat),
    V1(atm_i, ixmsat, SMMP_I32, mxat, 1),
    V1(atm_r, cgat, SMMP_F64, mxat, 1), V1(atm_r, blat, SMMP_F64, mxat, 1), V1(atm_r, baat, SMMP_F64, mxat, 1),
    V1(atm_r, toat, SMMP_F64, mxat, 1), V1(atm_r, sntoat, SMMP_F64, mxat, 1), V1(atm_r, cstoat, SMMP_F64, mxat, 1),
    V1(atm_r, snbaat, SMMP_F64, mxat, 1), V1(atm_r, csbaat, SMMP_F64, mxat, 1),
    V1(atm_r, xat, SMMP_F64, mxat + 2, -1), V1(atm_r, yat, SMMP_F64, mxat + 2, -1), V1(atm_r, zat, SMMP_F64, mxat + 2, -1),
    VC(atm_c, nmat, 4, mxat),
    V1(var_r, vlvr, SMMP_F64, mxvr, 1), V1(var_r, olvlvr, SMMP_F64, mxvr, 1), V1(var_r, axvr, SMMP_F64, mxvr, 1),
    V1(var_i, iatvr, SMMP_I32, mxvr, 1), V1(var_i, ityvr, SMMP_I32, mxvr, 1), V1(var_i, iclvr, SMMP_I32, mxvr, 1),
    V1(var_i, iorvr, SMMP_I32, mxvr, 1), V1(var_i, imsvr1, SMMP_I32, mxvr, 1), V1(var_i, imsvr2, SMMP_I32, mxvr, 1),
    V1(var_i, latms1, SMMP_I32, mxms, 1), V1(var_i, latms2, SMMP_I32, mxms, 1), V1(var_i, iadvr1, SMMP_I32, mxvr, 1),
    V1(var_i, iadvr2, SMMP_I32, mxvr, 1), V1(var_i, ladvr, SMMP_I32, mxvr, 1), V1(var_i, idvr, SMMP_I32, mxvr, 1),
    VC(var_c, nmvr, 3, mxvr),
    V1(var_l, fxvr, SMMP_I32, mxvr, 1),
    V2(epar_i, ihbty, SMMP_I32, mxtyat, mxtyat),
    V2(epar_r, cij, SMMP_F64, mxtyat, mxtyat), V2(epar_r, aij, SMMP_F64, mxtyat, mxtyat),
    V2(epar_r, a14, SMMP_F64, mxtyat, mxtyat), V2(epar_r, chb, SMMP_F64, mxtyat, mxtyat),
    V2(epar_r, ahb, SMMP_F64, mxtyat, mxtyat), V1(epar_r, e0to, SMMP_F64, mxtyto, 1),
    V1(epar_r, sgto, SMMP_F64, mxtyto, 1), V1(epar_r, rnto, SMMP_F64, mxtyto, 1),
    VS(epar_r, conv, SMMP_F64), VS(epar_r, plt, SMMP_F64), VS(epar_r, slp, SMMP_F64),
    V1(eny_i, ivwat1, SMMP_I32, mxat, 1), V1(eny_i, ivwat2, SMMP_I32, mxat, 1),
    V1(eny_i, i14at1, SMMP_I32, mxat, 1), V1(eny_i, i14at2, SMMP_I32, mxat, 1),
    V1(eny_i, lvwat1, SMMP_I32, mxvw, 1), V1(eny_i, lvwat2, SMMP_I32, mxvw, 1), V1(eny_i, l14at, SMMP_I32, mx14, 1),
    V1(eny_i, nvwml, SMMP_I32, mxml, 1), V1(eny_i, ivwml1, SMMP_I32, mxml, 1),
    V1(eny_i, n14ml, SMMP_I32, mxml, 1), V1(eny_i, i14ml1, SMMP_I32, mxml, 1),
    VS(eny_r, eyel, SMMP_F64), VS(eny_r, eyvw, SMMP_F64), VS(eny_r, eyhb, SMMP_F64), VS(eny_r, eyvr, SMMP_F64),
    VS(eny_r, eysm, SMMP_F64), VS(eny_r, eysl, SMMP_F64),
    V1(eny_r, gdeyvr, SMMP_F64, mxvr, 1), V1(eny_r, gdeysl, SMMP_F64, mxvr, 1),
    VS(eny_r, eysmi, SMMP_F64), VS(eny_r, eyeli, SMMP_F64), VS(eny_r, eyvwi, SMMP_F64), VS(eny_r, eyhbi, SMMP_F64),
    VS(bounds, boxsize, SMMP_F64),
    V1(solvent, rvdw, SMMP_F64, mxat, 1), V1(solvent, sigma, SMMP_F64, mxat, 1),
    V2(tes, spoint, SMMP_F64, mxpoint, 3), VS(pnt, npnt, SMMP_I32),
    VS(v_sol, vdvol, SMMP_F64), VS(v_sol, asa, SMMP_F64),
    V2(h_bond, mmhb, SMMP_I32, mxml, mxml), VS(h_bond, ntyhb, SMMP_I32),
    V1(h_bond, ixtyhb, SMMP_I32, State::mxtyhb, 1), V1(h_bond, nutyhb, SMMP_I32, State::mxtyhb, 1),
    {"zimme.zimm", SMMP_CHR, 2, {mxrs, 1}, {1, 1}, [](State& s) -> void* { return (void*)s.zimm; }},
    VS(myblock, mti, SMMP_I32), V1(myblock, mt, SMMP_I32, 624, 0),
};
constexpr int kNViews = sizeof(kViews) / sizeof(kViews[0]);

int fail(smmp_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

template <typename F>
int guarded(smmp_ctx* c, int code, F&& f) {
  if (!c) return SMMP_EINVAL;
  try {
    c->err.clear();
    return f();
  } catch (const Error& e) {
    return fail(c, code, e.msg);
  } catch (const DeviceError& e) {
    return fail(c, SMMP_EDEVICE, e.msg);
  } catch (const DriftError& e) {
    return fail(c, SMMP_EDRIFT, e.msg);
  } catch (const std::exception& e) {
    return fail(c, SMMP_EINVAL, e.what());
  }
}

}  // namespace

extern "C" {

int smmp_ctx_create(smmp_ctx** out) {
  if (!out) return SMMP_EINVAL;
  auto* c = new (std::nothrow) smmp_ctx();
  if (!c) return SMMP_EINVAL;
  reset_state(c->s);
  *out = c;
  return SMMP_OK;
}

void smmp_ctx_destroy(smmp_ctx* c) {
  if (!c) return;
  engine_destroy(c->eng);
  delete c;
}

const char* smmp_last_error(const smmp_ctx* c) { return c ? c->err.c_str() : "null context"; }

int smmp_state_count(void) { return kNViews; }
const char* smmp_state_name(int i) { return (i >= 0 && i < kNViews) ? kViews[i].name : nullptr; }

int smmp_state_view(smmp_ctx* c, const char* name, void** ptr, int* dtype, int* ndim, int64_t* shape,
                    int64_t* lbound) {
  if (!c || !name) return SMMP_EINVAL;
  for (const auto& v : kViews) {
    if (std::strcmp(v.name, name) != 0) continue;
    if (ptr) *ptr = v.ptr(c->s);
    if (dtype) *dtype = v.dtype;
    if (ndim) *ndim = v.ndim;
    for (int k = 0; k < 2; ++k) {
      if (shape) shape[k] = v.shape[k];
      if (lbound) lbound[k] = v.lbound[k];
    }
    return SMMP_OK;
  }
  return fail(c, SMMP_EINVAL, std::string("unknown state variable ") + name);
}

int smmp_init_energy(smmp_ctx* c, const char* libdir) {
  return guarded(c, SMMP_EINPUT, [&] {
    if (!libdir) throw Error("init_energy> libdir is NULL");
    init_energy(c->s, libdir);
    c->s.gen = next_generation();
    engine_invalidate(c->eng);
    return SMMP_OK;
  });
}

int smmp_init_molecule(smmp_ctx* c, int iabin, const char* grpn, const char* grpc, const char* seqfile,
                       const char* varfile) {
  return guarded(c, SMMP_EINPUT, [&] {
    init_molecule(c->s, iabin, grpn ? grpn : "nh2", grpc ? grpc : "cooh", seqfile ? seqfile : "",
                  varfile ? varfile : "");
    c->s.gen = next_generation();
    engine_invalidate(c->eng);
    return SMMP_OK;
  });
}

int smmp_setvar(smmp_ctx* c, int nml, const double* vlvrx) {
  return guarded(c, SMMP_EINVAL, [&] {
    if (nml < 1 || nml > c->s.ntlml) throw Error("setvar> no such molecule");
    setvar(c->s, nml, vlvrx ? vlvrx : c->s.vlvr.data());
    return SMMP_OK;
  });
}

int smmp_outvar(smmp_ctx* c, int nml, const char* path) {
  return guarded(c, SMMP_EINVAL, [&] {
    write_var_file(c->s, nml, path ? path : "");
    return SMMP_OK;
  });
}

int smmp_sgrnd(smmp_ctx* c, int32_t seed) {
  if (!c) return SMMP_EINVAL;
  c->s.mt(0) = seed;
  for (int i = 1; i < 624; ++i) c->s.mt(i) = (int32_t)(69069u * (uint32_t)c->s.mt(i - 1));
  c->s.mti = 624;
  return SMMP_OK;
}

int smmp_grnd(smmp_ctx* c, double* u) {
  if (!c || !u) return SMMP_EINVAL;
  *u = host_grnd(c->s);
  return SMMP_OK;
}

int smmp_energy(smmp_ctx* c, double* e) {
  return guarded(c, SMMP_ESTATE, [&] {
    if (c->s.ntlml < 1) throw Error("energy> no molecule initialised");
    double v = engine_energy(c->eng, c->s);
    if (e) *e = v;
    return SMMP_OK;
  });
}

int smmp_enysol(smmp_ctx* c, int nmol, double* e) {
  return guarded(c, SMMP_ESTATE, [&] {
    if (c->s.ntlml < 1) throw Error("enysol> no molecule initialised");
    if (nmol != 0 && (nmol < 1 || nmol > c->s.ntlml)) throw Error("enysol> no such molecule");
    if (nmol != 0 && c->s.ntlml > 1)  // enysol.f:53-59: that chain alone, as if the others were absent
      throw Error("enysol> the solvent term of one chain among several is not provided (use nmol = 0)");
    double v = engine_enysol(c->eng, c->s);
    if (e) *e = v;
    return SMMP_OK;
  });
}

int smmp_metropolis(smmp_ctx* c, double* eol, double* acz, const smmp_weight* w) {
  return guarded(c, SMMP_ESTATE, [&] {
    if (!eol || !acz || !w) throw Error("metropolis> NULL argument");
    if (c->s.ntlml < 1) throw Error("metropolis> no molecule initialised");
    return engine_metropolis(c->eng, c->s, eol, acz, *w);
  });
}

int smmp_partem_s(smmp_ctx* c, int num_rep, int nequi, int nswp, int switch_, const double* temps, double* eol,
                  int32_t* ipoi, double* traj, double* obs, const double* start, double* coor) {
  return guarded(c, SMMP_ESTATE, [&] {
    if (num_rep < 1 || num_rep > 2048 || nequi < 0 || nswp < 0 || !temps) throw Error("partem_s> bad arguments");
    if (c->s.ntlml < 1) throw Error("partem_s> no molecule initialised");
    return engine_partem_s(c->eng, c->s, num_rep, nequi, nswp, switch_, temps, eol, ipoi, traj, obs, start, coor);
  });
}

int smmp_redvar(smmp_ctx* c, const char* path, double* vlvr) {
  return guarded(c, SMMP_EINPUT, [&] {
    if (!path || !vlvr) throw Error("redvar> NULL argument");
    if (c->s.ntlml < 1) throw Error("redvar> no molecule initialised");
    auto t = std::make_unique<State>(c->s);
    redvar_file(*t, path);
    for (int i = 1; i <= t->nvars(); ++i) {  // as init_molecule.f:111-131 reads them back
      const int it = t->ityvr(i), a = t->iatvr(i);
      vlvr[i - 1] = it == 3 ? t->toat(a) : it == 2 ? t->baat(a) : t->blat(a);
    }
    return SMMP_OK;
  });
}

int smmp_helix(smmp_ctx* c, int32_t* out) {
  return guarded(c, SMMP_ESTATE, [&] {
    if (!out) throw Error("helix> NULL argument");
    if (c->s.ntlml < 1) throw Error("helix> no molecule initialised");
    helix(c->s, out);
    return SMMP_OK;
  });
}

int smmp_hbond(smmp_ctx* c, int nml, int32_t* mhb) {
  return guarded(c, SMMP_ESTATE, [&] {
    if (!mhb) throw Error("hbond> NULL argument");
    if (c->s.ntlml < 1) throw Error("hbond> no molecule initialised");
    *mhb = hbond(c->s, nml);
    return SMMP_OK;
  });
}

int smmp_interhbond(smmp_ctx* c, int32_t* mhb) {
  return guarded(c, SMMP_ESTATE, [&] {
    if (!mhb) throw Error("interhbond> NULL argument");
    *mhb = interhbond(c->s);
    return SMMP_OK;
  });
}

int smmp_zimmer(smmp_ctx* c, int nresi) {
  return guarded(c, SMMP_ESTATE, [&] {
    if (nresi < 0 || nresi > mxrs) throw Error("zimmer> nresi out of range");
    zimmer(c->s, nresi);
    return SMMP_OK;
  });
}

int smmp_rmsdfun(smmp_ctx* c, int nml, int ir1, int ir2, const int32_t* ixat, const double* xrf, const double* yrf,
                 const double* zrf, int isl, double* rmsd) {
  return guarded(c, SMMP_EINPUT, [&] {
    if (!ixat || !xrf || !yrf || !zrf || !rmsd) throw Error("rmsdfun> NULL argument");
    if (c->s.ntlml < 1) throw Error("rmsdfun> no molecule initialised");
    *rmsd = rmsdfun(c->s, nml, ir1, ir2, ixat, xrf, yrf, zrf, isl);
    return SMMP_OK;
  });
}

int smmp_outpdb(smmp_ctx* c, int nml, const char* path) {
  return guarded(c, SMMP_EINPUT, [&] {
    if (c->s.ntlml < 1) throw Error("outpdb> no molecule initialised");
    outpdb(c->s, nml, path ? std::string(path) : std::string());
    return SMMP_OK;
  });
}

int smmp_rgyr(smmp_ctx* c, int nml, double* rgy, double* ee) {
  return guarded(c, SMMP_ESTATE, [&] {
    if (!rgy || !ee) throw Error("rgyr> NULL argument");
    if (c->s.ntlml < 1) throw Error("rgyr> no molecule initialised");
    if (nml < 0 || nml > c->s.ntlml) throw Error("rgyr> no such molecule");
    rgyr(c->s, nml, rgy, ee);
    return SMMP_OK;
  });
}

int smmp_gradient(smmp_ctx* c) {
  return guarded(c, SMMP_ESTATE, [&] {
    if (c->s.ntlml < 1) throw Error("gradient> no molecule initialised");
    gradient_supported(c->s);
    const int nvt = c->s.nvars();
    double e[kGradEnergies];
    std::vector<double> g(nvt);
    engine_gradient(c->eng, c->s, 1, c->s.vlvr.data(), e, g.data());
    c->s.eysm = e[0];
    c->s.eyel = e[1];
    c->s.eyvw = e[2];
    c->s.eyhb = e[3];
    c->s.eyvr = e[4];
    c->s.eysl = c->s.itysol < 0 ? e[0] - (e[1] + e[2] + e[3] + e[4]) : 0.0;
    for (int k = 0; k < nvt; ++k) c->s.gdeyvr(k + 1) = g[k];
    for (int m = 1; m <= c->s.ntlml; ++m) setvar(c->s, m, c->s.vlvr.data());  // gradient.f:33
    return SMMP_OK;
  });
}

int smmp_gradient_batch(smmp_ctx* c, int nconf, const double* vlvr, double* energies, double* grad) {
  if (nconf < 0 || (nconf > 0 && (!vlvr || !energies || !grad))) return fail(c, SMMP_EINVAL, "gradient_batch> bad arguments");
  return guarded(c, SMMP_ESTATE, [&] {
    if (c->s.ntlml < 1) throw Error("gradient_batch> no molecule initialised");
    engine_gradient(c->eng, c->s, nconf, vlvr, energies, grad);
    return SMMP_OK;
  });
}

int smmp_minim(smmp_ctx* c, int imin, int maxit, double eps, int32_t* nfun) {
  return guarded(c, SMMP_ESTATE, [&] {
    const int k = minim(c->eng, c->s, imin, maxit, eps);
    if (nfun) *nfun = k;
    return SMMP_OK;
  });
}

int smmp_minim_batch(smmp_ctx* c, int imin, int nconf, double* vlvr, int maxit, double eps, double* fmin,
                     int32_t* nfun) {
  if (nconf < 0 || (nconf > 0 && !vlvr)) return fail(c, SMMP_EINVAL, "minim_batch> bad arguments");
  return guarded(c, SMMP_ESTATE, [&] {
    minim_batch(c->eng, c->s, imin, nconf, vlvr, maxit, eps, fmin, nfun);
    return SMMP_OK;
  });
}

int smmp_minqsn(int n, double* x, double* f, double* g, const double* scal, double acur, int maxfun, smmp_fg_fn fg,
                void* user, int32_t* nfun) {
  if (n < 1 || !x || !f || !g || !scal || !fg) return SMMP_EINVAL;
  QuasiNewton m;
  m.start(n, x, *f, g, scal, acur, maxfun);
  std::vector<double> gg(n);
  double ff = 0.0;
  while (m.pending()) {
    fg(user, m.point(), &ff, gg.data());
    m.supply(ff, gg.data());
  }
  for (int i = 0; i < n; ++i) {
    x[i] = m.x()[i];
    g[i] = m.g()[i];
  }
  *f = m.f();
  if (nfun) *nfun = m.nfun();
  return SMMP_OK;
}

int smmp_mincjg(int n, double* x, double* f, double* g, double acur, int maxfun, smmp_fg_fn fg, void* user,
                int32_t* nfun, int32_t* ier) {
  if (n < 1 || !x || !f || !g || !fg) return SMMP_EINVAL;
  ConjugateGradient m;
  m.start(n, x, *f, g, acur, maxfun);
  std::vector<double> gg(n);
  double ff = 0.0;
  while (m.pending()) {
    fg(user, m.point(), &ff, gg.data());
    m.supply(ff, gg.data());
  }
  for (int i = 0; i < n; ++i) {
    x[i] = m.x()[i];
    g[i] = m.g()[i];
  }
  *f = m.f();
  if (nfun) *nfun = m.nfun();
  if (ier) *ier = m.ier();
  return SMMP_OK;
}

int smmp_energy_batch(smmp_ctx* const* ctxs, int nsys, const int32_t* conf_sys, int nconf, const double* vlvr,
                      int precision, double* out) {
  if (!ctxs || nsys < 1 || !conf_sys || nconf < 0 || !vlvr || !out) return SMMP_EINVAL;
  smmp_ctx* c0 = ctxs[0];
  return guarded(c0, SMMP_ESTATE, [&] {
    std::vector<State*> sys(nsys);
    for (int i = 0; i < nsys; ++i) {
      if (!ctxs[i] || ctxs[i]->s.ntlml < 1) throw Error("energy_batch> context without a molecule");
      sys[i] = &ctxs[i]->s;
    }
    for (int k = 0; k < nconf; ++k)
      if (conf_sys[k] < 0 || conf_sys[k] >= nsys) throw Error("energy_batch> conf_sys out of range");
    if (precision != 0 && precision != 1) throw Error("energy_batch> precision must be 0 or 1");
    return engine_energy_batch(c0->eng, sys, conf_sys, nconf, vlvr, precision, out);
  });
}

int smmp_energy_batch_noalloc(smmp_ctx* const* ctxs, int nsys, const int32_t* conf_sys, int nconf,
                              const double* vlvr, int precision, double* out) {
  if (!ctxs || nsys < 1 || !conf_sys || nconf < 0 || !vlvr || !out) return SMMP_EINVAL;
  smmp_ctx* c0 = ctxs[0];
  return guarded(c0, SMMP_EINVAL, [&] {
    std::vector<State*> sys(nsys);
    for (int i = 0; i < nsys; ++i) {
      if (!ctxs[i] || ctxs[i]->s.ntlml < 1) throw Error("energy_batch_noalloc> context without a molecule");
      sys[i] = &ctxs[i]->s;
    }
    for (int k = 0; k < nconf; ++k)
      if (conf_sys[k] < 0 || conf_sys[k] >= nsys) throw Error("energy_batch_noalloc> conf_sys out of range");
    if (precision != 0 && precision != 1) throw Error("energy_batch_noalloc> precision must be 0 or 1");
    return engine_energy_batch_noalloc(sys, conf_sys, nconf, vlvr, precision, out);
  });
}

struct smmp_batch {
  smmp_ctx* c0;
  EnergyBatch* b;
};

int smmp_batch_create(smmp_ctx* const* ctxs, int nsys, const int32_t* conf_sys, int nconf, smmp_batch** out,
                      int64_t* nvals) {
  if (!ctxs || nsys < 1 || !ctxs[0] || nconf < 0 || (nconf > 0 && !conf_sys) || !out) return SMMP_EINVAL;
  *out = nullptr;
  smmp_ctx* c0 = ctxs[0];
  return guarded(c0, SMMP_ESTATE, [&] {
    std::vector<State*> sys(nsys);
    for (int i = 0; i < nsys; ++i) {
      if (!ctxs[i] || ctxs[i]->s.ntlml < 1) throw Error("energy_batch> context without a molecule");
      sys[i] = &ctxs[i]->s;
    }
    for (int k = 0; k < nconf; ++k)
      if (conf_sys[k] < 0 || conf_sys[k] >= nsys) throw Error("energy_batch> conf_sys out of range");
    EnergyBatch* b = engine_batch_create(c0->eng, sys, conf_sys, nconf);
    *out = new smmp_batch{c0, b};
    if (nvals) *nvals = engine_batch_values(b);
    return SMMP_OK;
  });
}

int smmp_batch_load(smmp_batch* b, const double* vlvr, int64_t nvals) {
  if (!b || !vlvr) return SMMP_EINVAL;
  return guarded(b->c0, SMMP_ESTATE, [&] {
    engine_batch_load(b->b, vlvr, nvals);
    return SMMP_OK;
  });
}

int smmp_batch_run(smmp_batch* b, int precision) {
  if (!b) return SMMP_EINVAL;
  return guarded(b->c0, SMMP_ESTATE, [&] {
    if (precision != 0 && precision != 1) throw Error("energy_batch> precision must be 0 or 1");
    engine_batch_run(b->b, precision);
    return SMMP_OK;
  });
}

int smmp_batch_tune(smmp_batch* b, int precision, int32_t* shapes, int32_t* nclasses) {
  if (!b) return SMMP_EINVAL;
  return guarded(b->c0, SMMP_ESTATE, [&] {
    if (precision != 0 && precision != 1) throw Error("energy_batch> precision must be 0 or 1");
    const int n = engine_batch_tune(b->b, precision, shapes);
    if (nclasses) *nclasses = n;
    return SMMP_OK;
  });
}

int smmp_batch_read(smmp_batch* b, double* out) {
  if (!b || !out) return SMMP_EINVAL;
  return guarded(b->c0, SMMP_ESTATE, [&] {
    engine_batch_read(b->b, out);
    return SMMP_OK;
  });
}

void smmp_batch_destroy(smmp_batch* b) {
  if (!b) return;
  engine_batch_destroy(b->b);
  delete b;
}

int smmp_sample(smmp_ctx* c, const smmp_sampler_cfg* cfg, double* vlvr, double* eol, double* acc, int32_t* ipoi,
                int32_t* hist, double* traj) {
  return guarded(c, SMMP_ESTATE, [&] {
    if (!cfg || !vlvr || !eol || !acc || !ipoi) throw Error("sample> NULL argument");
    if (cfg->nrep < 1 || cfg->nsweep < 0 || !cfg->temps || !cfg->seeds) throw Error("sample> bad configuration");
    if (cfg->mode == 2 && !cfg->w) throw Error("sample> multicanonical mode needs a weight");
    if (c->s.ntlml < 1) throw Error("sample> no molecule initialised");
    return engine_sample(c->eng, c->s, *cfg, vlvr, eol, acc, ipoi, hist, traj);
  });
}

const char* smmp_build_info(void) { return engine_build_info(); }

int smmp_profile(smmp_ctx* c, int enable) {
  return guarded(c, SMMP_EDEVICE, [&] {
    engine_profile(c->eng, enable);
    return SMMP_OK;
  });
}

int smmp_profile_read(smmp_ctx* c, int kernel, double* total_ms, int64_t* launches, double* work) {
  return guarded(c, SMMP_EDEVICE, [&] { return engine_profile_read(c->eng, kernel, total_ms, launches, work); });
}

int smmp_mlp_layout(const smmp_ctx* c, int32_t* out) {
  if (!c || !out) return SMMP_EINVAL;
  const State& s = c->s;
  if (s.ntlml < 1) return SMMP_ESTATE;
  const MlpLayout L = mlp_layout(s.irsml2(s.ntlml));
  const int32_t v[16] = {L.n, L.nres, L.D, smmp::kMlpHid, L.w1, L.b1, L.w2, L.b2,
                         L.w3, L.b3, L.w4, L.b4, L.emb, smmp::kMlpEmb, smmp::kMlpTypes, smmp::kMlpSlots};
  for (int k = 0; k < 16; ++k) out[k] = v[k];
  return SMMP_OK;
}

int smmp_mlp_size(const smmp_ctx* c) {
  int32_t out[16];
  return smmp_mlp_layout(c, out) == SMMP_OK ? out[0] : 0;
}

int smmp_mlp_set(smmp_ctx* c, const float* theta) {
  if (!theta) return SMMP_EINVAL;
  return guarded(c, SMMP_EDEVICE, [&] {
    if (c->s.ntlml < 1) throw Error("surrogate> no molecule initialised");
    engine_mlp_set(c->eng, c->s, theta);
    return SMMP_OK;
  });
}

int smmp_mlp_get(smmp_ctx* c, float* theta) {
  if (!theta) return SMMP_EINVAL;
  return guarded(c, SMMP_ESTATE, [&] {
    if (c->s.ntlml < 1) throw Error("surrogate> no molecule initialised");
    engine_mlp_get(c->eng, c->s, theta);
    return SMMP_OK;
  });
}

int smmp_mlp_grad(smmp_ctx* c, const double* vlvr, int nconf, const float* target, int pairs, float* pred,
                  float* grad, double* loss) {
  if (!vlvr || nconf <= 0 || (pairs && nconf % 2)) return SMMP_EINVAL;
  return guarded(c, SMMP_ESTATE, [&] {
    if (c->s.ntlml < 1) throw Error("surrogate> no molecule initialised");
    const double l = engine_mlp_grad(c->eng, c->s, vlvr, nconf, target, pred, grad, pairs);
    if (loss) *loss = l;
    return SMMP_OK;
  });
}

int smmp_mlp_fit(smmp_ctx* c, const double* vlvr, const float* target, int pairs, int n, int epochs, int batch,
                 float lr, uint32_t seed, double* epoch_loss) {
  if (!vlvr || !target || n <= 0 || epochs < 0 || batch <= 0) return SMMP_EINVAL;
  return guarded(c, SMMP_ESTATE, [&] {
    if (c->s.ntlml < 1) throw Error("surrogate> no molecule initialised");
    engine_mlp_fit(c->eng, c->s, vlvr, target, n, epochs, batch, lr, seed, epoch_loss, pairs);
    return SMMP_OK;
  });
}

int smmp_sampler_info(smmp_ctx* c, int32_t* out) {
  if (!out) return SMMP_EINVAL;
  return guarded(c, SMMP_EDEVICE, [&] { return engine_sampler_info(c->eng, out); });
}

int smmp_profile_phases(smmp_ctx* c, double* out) {
  if (!out) return SMMP_EINVAL;
  return guarded(c, SMMP_EDEVICE, [&] { return engine_profile_phases(c->eng, out); });
}

}  // extern "C"
