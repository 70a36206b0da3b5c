/*
 * SMMP oracle: a plain-C, FP64, single-threaded restatement of the SMMP
 * energy-and-move hot path.  TEST INFRASTRUCTURE ONLY: tests/, the smoke()
 * check and bench.py's CPU-baseline leg may load it; the product never does.
 *
 * Arrays use the reference's 1-based indices stored from element 0
 * (a(i) -> p[i-1]); 2-D tables are column-major like Fortran.
 */
#ifndef SMMP_ORACLE_H
#define SMMP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Topology and parameters of one system (inputs, read only).  nml = 0
 * describes one chain through i2, i3 and gbpr; nml >= 1 chains use the
 * mol_* arrays (1-based atoms) and gbm (6 values per chain). */
typedef struct {
  int nat, nvt, nms, npnt;
  const int32_t *ityat, *iowat, *nbdat, *ibdat; /* ibdat: 3*nat */
  const double *cgat, *blat, *baat, *toat;      /* toat: initial torsions */
  const int32_t *iatvr, *ityvr, *iclvr, *iorvr, *imsvr1;
  const int32_t *latms1, *latms2;
  const int32_t *ivwat1, *ivwat2, *lvwat1, *lvwat2, *i14at1, *i14at2, *l14at;
  const double *sigma, *rvdw; /* solvent radii (incl. water probe) */
  const double *cij, *aij, *a14, *chb, *ahb; /* 18x18 */
  const int32_t *ihbty;                      /* 18x18 */
  const double *e0to, *sgto, *rnto;          /* 19 */
  const double *spoint;                      /* npnt x 3, column-major */
  double conv, plt, slp;
  int epsd, itysol;
  int i2, i3;       /* second and third backbone atom (fnd3ba) */
  double gbpr[6];   /* global position/orientation */
  int nml;
  int mol_a1[10], mol_a2[10], mol_i2[10], mol_i3[10];
  double gbm[60];
  double boxsize;   /* COMMON /bounds/: rigid-move positions stay in [-boxsize, boxsize] when > 0 */
  /* variables added to each variable's moving set (setmvs.f): iadvr1 (nvt),
   * ladvr (nad); read by oracle_gradient only */
  const int32_t *iadvr1, *ladvr;
  int nad;
} oracle_topo;

/* Mutable state of one conformation (allocated by the caller). */
typedef struct {
  double *toat, *sntoat, *cstoat, *snbaat, *csbaat; /* nat */
  double *xat, *yat, *zat;                          /* nat */
  double *xtoat, *ytoat, *ztoat, *xbaat, *ybaat, *zbaat;
  /* components of the last energy: total, eysm, eyel, eyvw, eyhb, eyvr, eysl, asa, vdvol */
  double e[9];
  double ei[4];   /* between chains: eysmi, eyeli, eyvwi, eyhbi (eninteract.f) */
  double gb[60];  /* global parameters of the chains (6 per chain), set by oracle_conf_init */
} oracle_conf;

/* Mersenne twister of twister.f (state 624 words + position). */
typedef struct {
  uint32_t mt[624];
  int mti;
} oracle_mt;

void oracle_sgrnd(oracle_mt *g, int32_t seed);
double oracle_grnd(oracle_mt *g);
uint32_t oracle_grnd_u32(oracle_mt *g);
double oracle_difang(double a1, double a2);
double oracle_addang(double a1, double a2);

/* conformation memory helpers */
int oracle_conf_alloc(oracle_conf *c, int nat);
void oracle_conf_free(oracle_conf *c);
void oracle_conf_init(const oracle_topo *t, oracle_conf *c);

/* setvar.f + bldmol.f: apply all variables vlvr[0..nvt-1] and rebuild. */
void oracle_setvar(const oracle_topo *t, oracle_conf *c, const double *vlvr);
/* enyshe.f: fills e[2..5]; returns eysm. */
double oracle_enyshe(const oracle_topo *t, oracle_conf *c);
/* enysol.f: returns eysl; fills e[7], e[8]. */
double oracle_enysol(const oracle_topo *t, oracle_conf *c);
/* esolan.f:12-1050, the energy part: analytic solvent accessible areas
 * (projection method) of all atoms; returns sum sigma_i As_i, As (nat, may
 * be NULL).  Random rotations draw from g; NAN on coinciding atoms. */
double oracle_esolan(const oracle_topo *t, oracle_conf *c, oracle_mt *g, double *as_out);
/* oracle_esolan with the Cartesian gradient of esolan.f:533-1016: gradan
 * (3 nat, atom j at 3 (j - 1), may be NULL) = sum over atoms i of
 * sigma_i dAs_i/dr_j, the reference's gradan. */
double oracle_esolan_grad(const oracle_topo *t, oracle_conf *c, oracle_mt *g, double *as_out, double *gradan);
/* Largest neighbour and vertex counts of the last oracle_esolan call. */
void oracle_esolan_stats(int *max_nb, int *max_iv);
/* The generator oracle_energy gives esolan (itysol < 0); NULL: a draw gives NAN. */
void oracle_set_solvent_mt(oracle_mt *g);
/* Test hook: uga >= 0 forces the rotation of every atom with these numbers. */
void oracle_esolan_force_rotation(double uga, double ufi);
/* eninteract.f:14-86: pairs between chains; fills ei[], returns eysmi. */
double oracle_eninteract(const oracle_topo *t, oracle_conf *c);
/* gradient.f:13-70 with opeshe.f:12-388 and, for itysol < 0 (one chain),
 * opesol.f:12-258 over esolan.f's gradient: applies the variables, fills
 * e[0..6] (eysm + eysl, eysm, eyel, eyvw, eyhb, eyvr, eysl) and the
 * derivatives gd[nvt] of eysm + eysl by every variable; the terms between
 * chains are not part of it (gradient.f sums opeshe over the chains).
 * esolan's random rotations draw from the generator of
 * oracle_set_solvent_mt.  Returns eysm + eysl, or NAN when itysol > 0 (the
 * reference stops), for several chains with itysol < 0, or when esolan
 * fails. */
double oracle_gradient(const oracle_topo *t, oracle_conf *c, const double *vlvr, double *gd);
/* energy.f: setvar + enyshe + solvent + eninteract; returns total. */
double oracle_energy(const oracle_topo *t, oracle_conf *c, const double *vlvr);

/* Weight of an energy (canon.f:135 / mulcan_par_mod.f90:380-412). */
typedef struct {
  int kind;          /* 0 canonical: beta*x; 1 muca_weight; 2 muca_weight2 */
  double beta;
  int kmin, kmax;    /* bins */
  double ebin;
  const double *b, *alpha; /* kmax-kmin+1 */
} oracle_weight;
double oracle_weight_eval(const oracle_weight *w, double x);

/* metropolis.f:13-184.  idvr: free variable indices (1-based), axvr: step
 * widths (nvt).  vlvr (and, with several chains, c->gb through the rigid
 * trials of metropolis.f:94-168) is updated in place.  Returns 0, or 1 if
 * the end-of-sweep drift check fails (the reference stops there). */
int oracle_metropolis(const oracle_topo *t, oracle_conf *c, oracle_mt *g, double *vlvr,
                      int nvr, const int32_t *idvr, const double *axvr,
                      const oracle_weight *w, double *eol, double *acz);

/* Trial order and update-type choice for the calls above (metropolis.f:
 * 42-57; default: in order, upchswitch 0).  xname[nvt]: 1 where the
 * variable's name starts with 'x'; ubeta > 0: the /bet/ beta updtch2 reads,
 * else the weight's.  Clears the accanalyze counts; oracle_updstats gives
 * ncalls(1..5), nacalls(1..5) as 10 doubles. */
void oracle_set_updchois(int rndord, int upchswitch, double bgsprob, double up2bmax, double up2bmin,
                         const int8_t *xname, double ubeta);
void oracle_updstats(double *out);

/* partem_s.f:13-258 (new start), all replicas on one stream.  switch_ > 0:
 * random initial torsions, < 0: all at 180 degrees, 0: all at zero.
 * temps: R temperatures; outputs: eol[R] (by replica), ipoi[R] (slot ->
 * replica, 1-based), coor[R*nvr] final free variables, and per sweep and slot
 * the energy (traj[nswp*R], slot order, may be NULL). */
int oracle_partem(const oracle_topo *t, oracle_mt *g, const double *vlvr0, int nvr,
                  const int32_t *idvr, const double *axvr, int nrep, const double *temps,
                  int nequi, int nswp, int switch_, double *eol, int32_t *ipoi, double *coor, double *traj);

/* mulcan_par_mod.f90:23-190 weight recursion on one histogram update.
 * ihist (kmin..kmax) counts; b, alpha, g1, xhist updated in place. */
/* Records (variable, trial energy, old energy, accepted) of every trial of
 * replica 'rep' of oracle_partem (rep -1: of oracle_metropolis calls). */
void oracle_set_trace(double *buf, int cap, int rep);
int oracle_trace_count(void);

void oracle_muca_update(int kmin, int kmax, double ebin, double beta, const int32_t *ihist,
                        double *xhist, double *g1, double *b, double *alpha);

#ifdef __cplusplus
}
#endif
#endif
