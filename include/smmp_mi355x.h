/*
 * smmp_mi355x: C-ABI of the MI355X-native SMMP energy-and-move engine.
 *
 * Replaces the f2py extension module `smmp` generated from the reference's
 * smmp.pyf (file:line citations are relative to the reference's trunk/).
 * Plain pointers and sizes only.  Every call returns 0 on success or a
 * negative status; smmp_last_error() gives the message (the reference stops
 * the process instead, e.g. mklist.f:112, enysol.f:136, metropolis.f:177).
 *
 * One smmp_ctx holds one system (the reference's COMMON blocks, INCL.H) on
 * the host plus its device mirror on the current HIP device.  Contexts are
 * independent; a context is not thread-safe (as the reference,
 * universe.py:94-95).
 */
#ifndef SMMP_MI355X_H
#define SMMP_MI355X_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct smmp_ctx smmp_ctx;

/* status codes */
#define SMMP_OK 0
#define SMMP_EINVAL -1   /* bad argument */
#define SMMP_EINPUT -2   /* bad input file / topology (a reference 'stop') */
#define SMMP_EDEVICE -3  /* HIP error or no device */
#define SMMP_EDRIFT -4   /* end-of-sweep energy drift > 1e-6 (metropolis.f:172-179) */
#define SMMP_ESTATE -5   /* call out of order (e.g. energy before init_molecule) */

/* dtype codes of state views */
#define SMMP_I32 1
#define SMMP_F64 2
#define SMMP_CHR 3

/* ---- context --------------------------------------------------------- */
int smmp_ctx_create(smmp_ctx **out);
void smmp_ctx_destroy(smmp_ctx *ctx);
const char *smmp_last_error(const smmp_ctx *ctx);
/* Writable view of a COMMON-block variable, named "block.field" as in
 * smmp.pyf (e.g. "atm_r.xat", "bet.beta", "var_r.vlvr").  shape/lbound get
 * up to 2 dims (Fortran order); CHR arrays report (len, n). */
int smmp_state_view(smmp_ctx *ctx, const char *name, void **ptr, int *dtype, int *ndim,
                    int64_t *shape, int64_t *lbound);
/* number of names and the i-th name (for the shim's attribute table) */
int smmp_state_count(void);
const char *smmp_state_name(int i);

/* ---- f2py entry points (smmp.pyf) ------------------------------------ */
/* init_energy(libdir) smmp.pyf:4404, init_energy.f:14 (idempotent here) */
int smmp_init_energy(smmp_ctx *ctx, const char *libdir);
/* init_molecule(iabin,grpn,grpc,seqfile,varfile) smmp.pyf:5011, init_molecule.f:13 */
int smmp_init_molecule(smmp_ctx *ctx, int iabin, const char *grpn, const char *grpc,
                       const char *seqfile, const char *varfile);
/* setvar(nml,vlvrx) smmp.pyf:949, setvar.f:13: host rebuild of molecule nml */
int smmp_setvar(smmp_ctx *ctx, int nml, const double *vlvrx);
/* energy() smmp.pyf:380, energy.f:12: GPU evaluation of the current vlvr;
 * also sets eyel/eyvw/eyhb/eyvr/eysm/eysl and the coordinates */
int smmp_energy(smmp_ctx *ctx, double *e);
/* enysol(nmol) smmp.pyf:664, enysol.f:13 (solvent term of the current coordinates) */
int smmp_enysol(smmp_ctx *ctx, int nmol, double *e);
/* sgrnd(seed) smmp.pyf:5620, twister.f:64; grnd() twister.f:104 */
int smmp_sgrnd(smmp_ctx *ctx, int32_t seed);
int smmp_grnd(smmp_ctx *ctx, double *u);
/* outvar(nml,file) outvar.f:12 ('' = stdout) */
int smmp_outvar(smmp_ctx *ctx, int nml, const char *path);

/* Weight of an energy in a Metropolis step (replaces the `dummy` callback,
 * smmp.pyf:3-10): canonical can_weight (canon.f:135-149), muca_weight /
 * muca_weight2 (mulcan_par_mod.f90:380-412). */
typedef struct {
  int32_t kind;      /* 0 = beta*x, 1 = (beta+b(k))x+alpha(k), 2 = b(k)x+alpha(k) */
  double beta;
  int32_t kmin, kmax;
  double ebin;
  const double *b, *alpha; /* kmax-kmin+1 values, may be NULL for kind 0 */
} smmp_weight;

/* metropolis(eol,acz,dummy) smmp.pyf:22, metropolis.f:13: one sweep on the
 * GPU with the context's MT stream; eol/acz in,out. */
int smmp_metropolis(smmp_ctx *ctx, double *eol, double *acz, const smmp_weight *w);

/* partem_s(num_rep,nequi,nswp,nmes,nsave,newsta,switch) partem_s.f:13.
 * New start (start NULL, newsta): switch > 0 random initial torsions, < 0
 * all 180 degrees, 0 all zero (partem_s.f:114-126), then nequi sweeps of
 * every replica.  Restart (start given, newsta false): start[R*nvr] holds
 * every replica's free variables (idvr order, radians) and the run begins
 * with the parallel-tempering sweeps (partem_s.f:97-111; the reference reads
 * them from conf_NNNN.var but leaves its replicas at zero, which this does
 * not replicate).  nswp sweeps follow.  temps: R temperatures (the
 * 'temperatures' file).  Serial-compatible mode: every replica draws from
 * the context's single MT stream in slot order, as the reference does.
 * Outputs (may be NULL): eol[R] per replica, ipoi[R] slot->replica (1-based),
 * traj[nswp*R] energy per sweep and slot (the energies of time.d),
 * obs[nswp*R*2] per sweep and slot the radius of gyration and the C-alpha
 * end-to-end distance after the slot's sweep (rgyr(0,..) of
 * partem_s.f:186-193, the last two columns of time.d), coor[R*nvr] every
 * replica's final free variables (what partem_s.f:237-255 writes to
 * conf_NNNN.var). */
int smmp_partem_s(smmp_ctx *ctx, int num_rep, int nequi, int nswp, int switch_, const double *temps,
                  double *eol, int32_t *ipoi, double *traj, double *obs, const double *start, double *coor);

/* redvar() redvar.f:13 on a copy of the context: the values of all nvt
 * variables (radians) after reading the .var file at path over the
 * context's current ones; the context itself is not changed. */
int smmp_redvar(smmp_ctx *ctx, const char *path, double *vlvr);

/* rgyr(nml, rgy, ee) rgyr.f:13: radius of gyration over the atoms of chain
 * nml (0: all chains) and the distance between the C-alpha atoms of its first
 * and last residue, for the context's current coordinates. */
int smmp_rgyr(smmp_ctx *ctx, int nml, double *rgy, double *ee);

/* helix(nhel,mhel,nbet,mbet) helix.f:13: out[4] = helical residues, helical
 * segments, sheet-like residues, sheet-like segments of the context's
 * current variables. */
int smmp_helix(smmp_ctx *ctx, int32_t *out);

/* hbond(nml,mhb,ipr) smmp.pyf:1806, hbond.f:14: H-bonds of molecule nml
 * (>= 1) by McDonald and Thornton's criteria over the energy's pair lists,
 * for the context's current coordinates; *mhb counts those between residues
 * more than four apart, h_bond.ntyhb/ixtyhb/nutyhb hold the type table.
 * interhbond(mhb) smmp.pyf:2092, hbond.f:228: H-bonds between chains (all
 * atom pairs), h_bond.mmhb per pair of chains. */
int smmp_hbond(smmp_ctx *ctx, int nml, int32_t *mhb);
int smmp_interhbond(smmp_ctx *ctx, int32_t *mhb);

/* zimmer(nresi) smmp.pyf:3238, zimmer.f:15: Zimmerman code of residues
 * 1..nresi from the current (phi, psi) into zimme.zimm. */
int smmp_zimmer(smmp_ctx *ctx, int nresi);

/* rmsdfun(nml,ir1,ir2,ixat,xrf,yrf,zrf,isl) smmp.pyf:4095, rmsdfun.f:12:
 * rmsd of the heavy atoms of residues ir1..ir2 of molecule nml (counted
 * within it) from the reference coordinates after the optimal
 * superposition (quaternion method, rmsdfun.f's Jacobi), current
 * coordinates; ixat[a-1] (atoms numbered over all molecules): index of atom
 * a's reference (1-based, 0 none); isl 0 all heavy atoms, 1 n/ca/c, 2 ca.
 * Fewer than three atoms selected: SMMP_EINPUT. */
int smmp_rmsdfun(smmp_ctx *ctx, int nml, int ir1, int ir2, const int32_t *ixat, const double *xrf,
                 const double *yrf, const double *zrf, int isl, double *rmsd);

/* outpdb(nml,file) smmp.pyf:3810, outpdb.f:12: PDB file of molecule nml
 * (0: all) from the current coordinates ('' = stdout). */
int smmp_outpdb(smmp_ctx *ctx, int nml, const char *path);

/* gradient.f:13-70 with opeshe.f:12-388 and, for the analytic solvent,
 * opesol.f:12-258 over esolan.f's Cartesian gradient (called by minim.f;
 * smmp.pyf exports the COMMON it fills, eny_r.gdeyvr at smmp.pyf:260): the
 * energy of the context's variables within each chain (eysm, eyel, eyvw,
 * eyhb, eyvr, eysl; with itysol < 0 eysm includes eysl as gradient.f:47
 * leaves it) and its derivatives by every variable into eny_r.gdeyvr, on
 * the device in FP64.  itysol > 0 is refused as the reference stops there
 * (gradient.f:59-62).  itysol < 0 is refused too (SMMP_ESTATE): the device
 * form of the analytic solvent's derivatives is built but switched off until
 * a fault of its first run is explained; when on, it takes one chain
 * (esolan.f numbers a molecule's atoms from 1) and draws the rotations of
 * degenerate north points from the context's twister like esolan.f:199-200. */
int smmp_gradient(smmp_ctx *ctx);

/* minim(imin,maxit,eps) smmp.pyf:1234, minim.f:13: local minimisation of
 * the energy by the free variables, imin 1 quasi-Newton (minqsn.f), 2
 * conjugate gradients (mincjg.f), at most maxit function calls, eps the
 * precision; derivatives from smmp_gradient (vacuum, ireg = 0).  Leaves the
 * best variables (wrapped into [-pi, pi]) in var_r.vlvr and the energies of
 * energy() at the last evaluated point in eny_r, like minim.f.  nfun (may
 * be NULL): function calls used (maxit: the step limit was reached). */
int smmp_minim(smmp_ctx *ctx, int imin, int maxit, double eps, int32_t *nfun);

/* minqsn.f:12 and mincjg.f:12 as library routines for any function:
 * fg(user, x, &f, g) evaluates f and its gradient at x (the reference's
 * MOVE).  x, f, g: the start on entry, the result on return (minqsn: the
 * best point; mincjg: as mincjg.f leaves them); scal (minqsn): per
 * variable scale; acur: precision; maxfun: most calls; nfun: calls used;
 * ier (mincjg, may be NULL): 0 normal end, 1 call limit, 2 step too small,
 * 3 uphill direction, 4 no reduction. */
typedef void (*smmp_fg_fn)(void *user, const double *x, double *f, double *g);
int smmp_minqsn(int n, double *x, double *f, double *g, const double *scal, double acur, int maxfun,
                smmp_fg_fn fg, void *user, int32_t *nfun);
int smmp_mincjg(int n, double *x, double *f, double *g, double acur, int maxfun, smmp_fg_fn fg, void *user,
                int32_t *nfun, int32_t *ier);

/* ---- batched entry points (no reference counterpart: the configs'
 *      many-replica and many-conformation workloads) ------------------- */

/* Energies of many conformations of one or several systems in one launch.
 * ctxs[s] are initialised contexts; conf_sys[c] picks the system of
 * conformation c; vlvr holds for each conformation the values of ALL
 * variables of its system (offsets: sum of nvt of the previous ones).
 * precision: 0 = FP64 throughout, 1 = FP32 pair terms with FP64
 * accumulation and FP64 chain rebuild.
 * out[c*8 + k] = total, eysm, eyel, eyvw, eyhb, eyvr, eysl, asa. */
int smmp_energy_batch(smmp_ctx *const *ctxs, int nsys, const int32_t *conf_sys, int nconf,
                      const double *vlvr, int precision, double *out);
/* gradient.f (vacuum) for nconf full variable vectors of the context's
 * system, vlvr[nconf*nvt], in one launch: energies[5c..] = eysm, eyel,
 * eyvw, eyhb, eyvr; grad[nvt c..] = derivatives by every variable. */
int smmp_gradient_batch(smmp_ctx *ctx, int nconf, const double *vlvr, double *energies, double *grad);
/* minim.f for nconf conformations side by side: each takes exactly the
 * steps smmp_minim would take alone from that start, and every round of
 * function calls is one device launch over the conformations still
 * running.  vlvr[nconf*nvt] in and out (the best points, free variables
 * wrapped into [-pi, pi]); fmin[c]: energy of gradient.f there; nfun[c]:
 * calls used (either may be NULL). */
int smmp_minim_batch(smmp_ctx *ctx, int imin, int nconf, double *vlvr, int maxit, double eps, double *fmin,
                     int32_t *nfun);
/* The same energies as smmp_energy_batch (identical results) from an entry
 * point that allocates no device memory.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Everything the
 * call needs on the device (its systems, conformation tables, values,
 * results, error flag) is laid out in a fixed 32 MiB array per device that
 * the library holds until the process ends; a batch that needs more is
 * refused with SMMP_EINVAL and a message naming the size.  Runs on the
 * current device's null stream.  Calls on one GPU share the array: calls
 * that overlap on the same GPU race.  Not for the analytic solvent
 * (itysol < 0, which needs its own workspace): SMMP_EINVAL. */
int smmp_energy_batch_noalloc(smmp_ctx *const *ctxs, int nsys, const int32_t *conf_sys, int nconf,
                              const double *vlvr, int precision, double *out);
/* The call above keeps its systems on the device and reuses them while it is
 * called again with the same contexts and conf_sys (a system is rebuilt when
 * its context's topology changes: init_energy, init_molecule).
 *
 * Resident batch: the same computation with the conformations and results
 * kept in device memory between calls, for callers that evaluate a fixed set
 * of conformations repeatedly or refill it in place.  The contexts must
 * outlive the batch; it runs on the stream of ctxs[0].
 *   smmp_batch_create: uploads the systems; *nvals = doubles of vlvr it takes
 *   smmp_batch_load:   conformations (layout as smmp_energy_batch) -> device;
 *                      nvals must equal what the batch takes now (it
 *                      changes when a context's topology does)
 *   smmp_batch_run:    energies of the loaded conformations (enqueued; with
 *                      the analytic solvent it waits for the kernels)
 *   smmp_batch_read:   waits, out[c*8 + k] as smmp_energy_batch
 *   smmp_batch_destroy */
typedef struct smmp_batch smmp_batch;
int smmp_batch_create(smmp_ctx *const *ctxs, int nsys, const int32_t *conf_sys, int nconf, smmp_batch **out,
                      int64_t *nvals);
int smmp_batch_load(smmp_batch *b, const double *vlvr, int64_t nvals);
int smmp_batch_run(smmp_batch *b, int precision);
int smmp_batch_read(smmp_batch *b, double *out);
/* The energy kernel's launch shapes for this batch picked by measurement
 * (the project's autotuner; no counterpart in the reference): for each of
 * the batch's LDS classes, 256, 512 and 1024 threads per workgroup with the
 * pair bit rows staged in LDS (where they fit) or read through the caches
 * are timed on the loaded conformations and the fastest is kept for later
 * smmp_batch_run calls at this precision (until the systems' layout
 * changes).  The shapes change the order of the FP64 sums, so energies may
 * differ from untuned runs in the last bits.  shapes (may be NULL): threads
 * and bits flag per class (2 ints each, at most 2 classes); nclasses (may
 * be NULL): the number of classes. */
int smmp_batch_tune(smmp_batch *b, int precision, int32_t *shapes, int32_t *nclasses);
void smmp_batch_destroy(smmp_batch *b);

/* Independent-stream replica sampler: R replicas of the context's system,
 * each with its own twister.f MT19937 stream seeded sgrnd(seeds[r]).
 *   mode 0: canonical at temps[r]; mode 1: parallel tempering over temps
 *   (exchange attempt after every sweep, partem_s.f:209-232 rule, exchange
 *   draws from its own stream seeded sgrnd(seeds[R])); mode 2:
 *   multicanonical with the weight w (muca_weight2) shared by all replicas.
 * With the analytic solvent (itysol < 0; one chain) the replicas take their
 * trials in lockstep with one device energy pass per trial step, and each
 * replica's solvent rotations draw from its own stream; delayed acceptance,
 * verify and graph_sweeps do not apply there.
 * vlvr[r*nvt..] in: start conformations (all variables), out: final.
 * eol[r] out: final energies; acc[r] out: accepted moves; ipoi[R] out.
 * hist (mode 2, may be NULL): kmax-kmin+1 int32 counts of sweep-end bins.
 * traj (may be NULL): nsweep*R energies at sweep ends, slot order.
 * Returns SMMP_EDRIFT if a sweep-end recomputation differs by > 1e-6. */
typedef struct {
  int32_t nrep, nsweep, mode, precision;
  const double *temps;
  const uint32_t *seeds;
  const smmp_weight *w;
  int32_t resync_every;   /* full recomputation cadence in sweeps (1 = every sweep) */
  int32_t verify;         /* 1: recompute every trial in full and fail on the first
                             incremental component that disagrees (slow; tests) */
  int32_t groups;         /* workgroups per replica (upper bound; 0 = automatic,
                             -1 = measured: 1, 2, 4, 8 are timed on the call's own
                             replicas and the fastest is kept for later calls on the
                             same system, replica count, precision and mode).  With
                             more than one, the workgroups of a replica split each
                             energy change and reduce it on the device every trial */
  int32_t graph_sweeps;   /* > 0: the launches of that many sweeps (and their
                             exchanges) are captured once into a HIP graph and
                             replayed nsweep / graph_sweeps times (nsweep must be a
                             multiple); the graph stays with the context and later
                             calls with the same configuration replay it without
                             capturing.  0: launched directly. */
  double *gbpr;           /* several chains: NULL, or R x ntlml x 6 global parameters
                             (gbpr(1..6, chain) of each replica; in: start, out: final).
                             NULL: every replica starts from the context's gbpr.  The
                             sweeps end with the rigid trials of metropolis.f:94-168. */
  int32_t da;             /* 1: delayed acceptance (modes 0 and 1): the context's surrogate
                             network (smmp_mlp_set) screens every torsion trial before its
                             exact energy; 0: plain Metropolis */
  double da_scale;        /* kcal/mol per network unit */
  double *da_stats;       /* NULL or R x 4 out: trials screened by the network, rejected by
                             it, judged with the exact energy, accepted */
  double *snap_vlvr;      /* NULL or (nsweep / snap_every) x R x nvt out: every replica's
                             variables after every snap_every-th sweep (training data) */
  double *snap_e;         /* with snap_vlvr: (nsweep / snap_every) x R energies */
  int32_t snap_every;
  double *obs;            /* NULL or nsweep x R x 2 out: radius of gyration and C-alpha
                             end-to-end distance after every sweep, slot order (rgyr(0,..)) */
} smmp_sampler_cfg;
int smmp_sample(smmp_ctx *ctx, const smmp_sampler_cfg *cfg, double *vlvr, double *eol,
                double *acc, int32_t *ipoi, int32_t *hist, double *traj);

/* ---- the delayed-acceptance surrogate (no reference counterpart) -------
 * An MLP over the whole system: per residue (all chains in order) sin and
 * cos of its up to 10 variables and a 16-dimensional embedding of its type
 * (36 inputs per residue), three SiLU hidden layers of 256, one output;
 * E~ = scale * output.  For 36 residues that is 1296 inputs and 464,400
 * parameters.  The context holds one network (smmp_mlp_size(ctx) FP32
 * parameters, laid out as smmp_mlp_layout gives) and its Adam moments on the
 * device; training and evaluation run BF16 MFMA kernels with FP32
 * accumulation.  All arrays below are host arrays.
 * smmp_mlp_layout: out[16] = parameter count, residues, inputs, hidden
 * width, offsets of W1 (inputs x hidden, input-major), b1, W2 (hidden x
 * hidden), b2, W3, b3, w4, b4, embeddings (types x 16), embedding width,
 * embedding rows, variables per residue. */
int smmp_mlp_layout(const smmp_ctx *ctx, int32_t *out);
int smmp_mlp_size(const smmp_ctx *ctx);
/* Sets the context's network (and clears its Adam moments) / reads it. */
int smmp_mlp_set(smmp_ctx *ctx, const float *theta);
int smmp_mlp_get(smmp_ctx *ctx, float *theta);
/* Outputs for nconf conformations of the context's system (vlvr: nconf x
 * nvt): pred (nconf, may be NULL) in network units.  With targets: *loss =
 * mean squared error and grad (smmp_mlp_size(ctx), may be NULL) its gradient;
 * pairs = 0: one target per conformation (its output), pairs = 1: the
 * conformations come in pairs (x, x') and target[c] is for the difference
 * of the outputs of conformations 2c+1 and 2c (nconf even). */
int smmp_mlp_grad(smmp_ctx *ctx, const double *vlvr, int nconf, const float *target, int pairs, float *pred,
                  float *grad, double *loss);
/* Minibatch training of the context's network on n samples with targets
 * (network units); a sample is one conformation (vlvr: n x nvt), or with
 * pairs two (vlvr: 2n x nvt) as in smmp_mlp_grad.  'epochs' passes in
 * random order (seeded), batches of 'batch' samples, one Adam step each
 * (beta1 0.9, beta2 0.999, eps 1e-8); epoch_loss (epochs, may be NULL):
 * mean squared error of each epoch's batches. */
int smmp_mlp_fit(smmp_ctx *ctx, const double *vlvr, const float *target, int pairs, int n, int epochs, int batch,
                 float lr, uint32_t seed, double *epoch_loss);

/* How the last smmp_sample call ran (no reference counterpart):
 * out[0] workgroups per replica, out[1] 1 if the whole run was one
 * persistent launch per graph replay or call (parallel tempering with the
 * exchanges inside the kernel), out[2] graph replays (0: direct launches),
 * out[3] 1 if the graph was reused from an earlier call.  out holds 4. */
int smmp_sampler_info(smmp_ctx *ctx, int32_t *out);

/* Multicanonical weight iteration (host arithmetic; arrays indexed from
 * kmin, kmax-kmin+1 values each).
 * smmp_muca_update: one update of mulcan_par (mulcan_par_mod.f90:104-155):
 * adds the sweep histogram ihist to xhist, advances Berg's recursion in
 * g1/b, recomputes alpha and (if not NULL) ent = (beta+b)k ebin + alpha,
 * and clears ihist.  beta is mulcan_par's 1/(k_B temp).
 * smmp_muca_final: the weights written to muca.d (mulcan_par_mod.f90:168-176):
 * b += beta and alpha recomputed, in place. */
int smmp_muca_update(int kmin, int kmax, double ebin, double beta, int32_t *ihist, double *xhist, double *g1,
                     double *b, double *alpha, double *ent);
int smmp_muca_final(int kmin, int kmax, double ebin, double beta, double *b, double *alpha);

/* Device / build information string (arch, kernels, tile shapes). */
const char *smmp_build_info(void);

/* Kernel timing (no reference counterpart): when enabled, every launch on
 * the context's stream is bracketed by HIP events.  kernel: 0 = sweep,
 * 1 = exchange, 2 = energy, 3 = analytic solvent (itysol < 0).  work[4]:
 * pair-term evaluations, solvent contacts (each tested against all
 * tessellation points), atoms placed by chain builds, atoms rotated by trial
 * moves, summed over the launches; for kernel 3: atom pairs scanned for
 * neighbours, circle pairs classified, arcs tested, circles tested against
 * arc midpoints. */
int smmp_profile(smmp_ctx *ctx, int enable);
int smmp_profile_read(smmp_ctx *ctx, int kernel, double *total_ms, int64_t *launches, double *work);
/* With smmp_profile(ctx, 2) the sweep kernel also times its phases on the
 * device clock (one replica = one workgroup, thread 0 after each barrier).
 * out[...]: milliseconds summed over replicas for proposal, rotation, pair
 * change (+ contact marking, torsion term), solvent change, the reduction
 * across a replica's workgroups, acceptance and commit, full-recomputation
 * trials, sweep-end rebuild; out[8] = number of trials that took the
 * full-recomputation path (|E| >= 1e5); out[9] = core clock cycles and
 * out[10] = milliseconds of the whole launches (their ratio is the clock);
 * solvent change of the first workgroup of each replica: out[11] ms of its
 * slowest wave, out[12]/out[13] ms of neighbour gathering / point tests
 * summed over waves, out[14] atoms evaluated, out[15] unused; pair change
 * of the same workgroups: out[16] ms of the slowest wave, out[17] ms summed
 * over waves, out[18] ms of contact marking summed over waves, out[19] ms of
 * thread 0's torsion term; reduction across a replica's workgroups (the
 * same workgroups): out[20] ms waiting for the last workgroup to arrive,
 * out[21] ms from that arrival until the wait ends.  out must hold 22
 * values. */
int smmp_profile_phases(smmp_ctx *ctx, double *out);

#ifdef __cplusplus
}
#endif
#endif
