/*
 * SMMP oracle (test infrastructure only; see smmp_oracle.h).
 *
 * A line-by-line-faithful FP64 restatement of the reference's arithmetic, in
 * its evaluation order, so results agree with the reference objects to the
 * last bits where the compilers round alike.  Built with -ffp-contract=off.
 * Each function names the reference routine it restates.
 */
#include "smmp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI 3.141592653589793
#define PI2 (2.0 * PI)
#define GASC 0.00198773

/* ------------------------------------------------ twister.f:64-179 */

void oracle_sgrnd(oracle_mt *g, int32_t seed) {
  g->mt[0] = (uint32_t)seed;
  for (int i = 1; i < 624; ++i) g->mt[i] = 69069u * g->mt[i - 1];
  g->mti = 624;
}

uint32_t oracle_grnd_u32(oracle_mt *g) {
  if (g->mti >= 624) {
    if (g->mti == 625) oracle_sgrnd(g, 4357);
    static const uint32_t mag01[2] = {0u, 0x9908B0DFu};
    int kk;
    uint32_t y;
    for (kk = 0; kk < 624 - 397; ++kk) {
      y = (g->mt[kk] & 0x80000000u) | (g->mt[kk + 1] & 0x7fffffffu);
      g->mt[kk] = g->mt[kk + 397] ^ (y >> 1) ^ mag01[y & 1u];
    }
    for (; kk < 623; ++kk) {
      y = (g->mt[kk] & 0x80000000u) | (g->mt[kk + 1] & 0x7fffffffu);
      g->mt[kk] = g->mt[kk + (397 - 624)] ^ (y >> 1) ^ mag01[y & 1u];
    }
    y = (g->mt[623] & 0x80000000u) | (g->mt[0] & 0x7fffffffu);
    g->mt[623] = g->mt[396] ^ (y >> 1) ^ mag01[y & 1u];
    g->mti = 0;
  }
  uint32_t y = g->mt[g->mti++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9D2C5680u;
  y ^= (y << 15) & 0xEFC60000u;
  y ^= y >> 18;
  return y;
}

double oracle_grnd(oracle_mt *g) { return (double)oracle_grnd_u32(g) / 4294967295.0; }

/* ------------------------------------------------ difang.f:12-63 */

double oracle_difang(double a1, double a2) {
  double d = fmod(a2 - a1, PI2);
  return fabs(d) <= PI ? d : d - copysign(PI2, d);
}
double oracle_addang(double a1, double a2) {
  double d = fmod(a1 + a2, PI2);
  return fabs(d) <= PI ? d : d - copysign(PI2, d);
}

/* ------------------------------------------------ conformation memory */

int oracle_conf_alloc(oracle_conf *c, int nat) {
  memset(c, 0, sizeof *c);
  double **f[] = {&c->toat, &c->sntoat, &c->cstoat, &c->snbaat, &c->csbaat, &c->xat, &c->yat,
                  &c->zat, &c->xtoat, &c->ytoat, &c->ztoat, &c->xbaat, &c->ybaat, &c->zbaat};
  for (size_t k = 0; k < sizeof f / sizeof f[0]; ++k) {
    *f[k] = (double *)calloc((size_t)nat + 1, sizeof(double));
    if (!*f[k]) return -1;
  }
  return 0;
}

void oracle_conf_free(oracle_conf *c) {
  double *f[] = {c->toat, c->sntoat, c->cstoat, c->snbaat, c->csbaat, c->xat, c->yat,
                 c->zat, c->xtoat, c->ytoat, c->ztoat, c->xbaat, c->ybaat, c->zbaat};
  for (size_t k = 0; k < sizeof f / sizeof f[0]; ++k) free(f[k]);
  memset(c, 0, sizeof *c);
}

void oracle_conf_init(const oracle_topo *t, oracle_conf *c) {
  memset(c->ei, 0, sizeof c->ei);
  memset(c->gb, 0, sizeof c->gb);
  if (t->nml > 0)
    memcpy(c->gb, t->gbm, sizeof(double) * 6 * (size_t)t->nml);
  else
    memcpy(c->gb, t->gbpr, sizeof(double) * 6);
  for (int i = 0; i < t->nat; ++i) {
    c->toat[i] = t->toat[i];
    c->sntoat[i] = sin(t->toat[i]);
    c->cstoat[i] = cos(t->toat[i]);
    c->snbaat[i] = sin(t->baat[i]);
    c->csbaat[i] = cos(t->baat[i]);
  }
}

/* ------------------------------------------------ bldmol.f */

#define A(p, i) ((p)[(i) - 1])

/* bldmol.f:175-242 */
static void eyring(const oracle_topo *t, oracle_conf *c, int i, int ia) {
  double ct = A(c->cstoat, i), st = A(c->sntoat, i), ca = A(c->csbaat, i), sa = A(c->snbaat, i);
  double bl = A(t->blat, i);
  double x1 = A(c->xtoat, ia), x2 = A(c->ytoat, ia), x3 = A(c->ztoat, ia);
  double z1 = A(c->xbaat, ia), z2 = A(c->ybaat, ia), z3 = A(c->zbaat, ia);
  double y1 = z2 * x3 - z3 * x2, y2 = z3 * x1 - z1 * x3, y3 = z1 * x2 - z2 * x1;
  double h2 = -sa * ct, h3 = -sa * st;
  x1 = -ca * x1 + h2 * y1 + h3 * z1;
  x2 = -ca * x2 + h2 * y2 + h3 * z2;
  x3 = -ca * x3 + h2 * y3 + h3 * z3;
  double dx = 1.0 / sqrt(x1 * x1 + x2 * x2 + x3 * x3);
  x1 = x1 * dx;
  x2 = x2 * dx;
  x3 = x3 * dx;
  A(c->xtoat, i) = x1;
  A(c->ytoat, i) = x2;
  A(c->ztoat, i) = x3;
  z1 = -st * y1 + ct * z1;
  z2 = -st * y2 + ct * z2;
  z3 = -st * y3 + ct * z3;
  double dz = 1.0 / sqrt(z1 * z1 + z2 * z2 + z3 * z3);
  A(c->xbaat, i) = z1 * dz;
  A(c->ybaat, i) = z2 * dz;
  A(c->zbaat, i) = z3 * dz;
  A(c->xat, i) = A(c->xat, ia) + x1 * bl;
  A(c->yat, i) = A(c->yat, ia) + x2 * bl;
  A(c->zat, i) = A(c->zat, ia) + x3 * bl;
}

/* chain m (0-based): first and last atom, second and third backbone atom */
static int nchains(const oracle_topo *t) { return t->nml > 0 ? t->nml : 1; }
static void chain_atoms(const oracle_topo *t, int m, int *a1, int *a2, int *i2, int *i3) {
  if (t->nml > 0) {
    *a1 = t->mol_a1[m]; *a2 = t->mol_a2[m]; *i2 = t->mol_i2[m]; *i3 = t->mol_i3[m];
  } else {
    *a1 = 1; *a2 = t->nat; *i2 = t->i2; *i3 = t->i3;
  }
}

/* bldmol.f:303-420 for a chain with global parameters g and backbone atoms i2, i3 */
static void setgbl(const oracle_conf *c, const double *g, int i2, int i3, double *xg, double *zg) {
  double ca = cos(g[3]), sa = sin(g[3]), cb = cos(g[4]), sb = sin(g[4]), cg = cos(g[5]), sg = sin(g[5]);
  double ag[3][3], d;
  double x1 = ca * cg - sa * sb * sg, x2 = sa * cg + ca * sb * sg, x3 = -cb * sg;
  d = sqrt(x1 * x1 + x2 * x2 + x3 * x3);
  ag[0][0] = x1 / d; ag[1][0] = x2 / d; ag[2][0] = x3 / d;
  double y1 = -sa * cb, y2 = ca * cb, y3 = sb;
  d = sqrt(y1 * y1 + y2 * y2 + y3 * y3);
  ag[0][1] = y1 / d; ag[1][1] = y2 / d; ag[2][1] = y3 / d;
  double z1 = ca * sg + sa * sb * cg, z2 = sa * sg - ca * sb * cg, z3 = cb * cg;
  d = sqrt(z1 * z1 + z2 * z2 + z3 * z3);
  ag[0][2] = z1 / d; ag[1][2] = z2 / d; ag[2][2] = z3 / d;
  double ct2 = A(c->cstoat, i2), st2 = A(c->sntoat, i2), sa2 = A(c->snbaat, i2);
  x1 = -A(c->csbaat, i2); x2 = -sa2 * ct2; x3 = -sa2 * st2;
  double dx = sqrt(x1 * x1 + x2 * x2 + x3 * x3);
  x1 = x1 / dx; x2 = x2 / dx; x3 = x3 / dx;
  double st3 = A(c->sntoat, i3), ct3 = A(c->cstoat, i3);
  z1 = -st3 * (st2 * x3 + ct2 * x2);
  z2 = st3 * ct2 * x1 + ct3 * st2;
  z3 = st3 * st2 * x1 - ct3 * ct2;
  double dz = sqrt(z1 * z1 + z2 * z2 + z3 * z3);
  z1 = z1 / dz; z2 = z2 / dz; z3 = z3 / dz;
  y1 = z2 * x3 - z3 * x2;
  y3 = z1 * x2 - z2 * x1;
  for (int i = 0; i < 3; ++i) {
    xg[i] = ag[i][0] * x1 + ag[i][1] * y1 + ag[i][2] * z1;
    zg[i] = ag[i][0] * x3 + ag[i][1] * y3 + ag[i][2] * z3;
  }
  dx = sqrt(xg[0] * xg[0] + xg[1] * xg[1] + xg[2] * xg[2]);
  dz = sqrt(zg[0] * zg[0] + zg[1] * zg[1] + zg[2] * zg[2]);
  for (int i = 0; i < 3; ++i) {
    xg[i] = xg[i] / dx;
    zg[i] = zg[i] / dz;
  }
}

/* setvar.f:13-73 followed by bldmol.f:13-93 for every chain (energy.f:40-46) */
void oracle_setvar(const oracle_topo *t, oracle_conf *c, const double *vlvr) {
  for (int i = 1; i <= t->nvt; ++i) {
    int iat = A(t->iatvr, i), ity = A(t->ityvr, i);
    if (ity != 3) continue; /* the ECEPP/3 library defines torsion variables only */
    double tsh = oracle_difang(A(c->toat, iat), vlvr[i - 1]);
    int iow = A(t->iowat, iat);
    for (int j = 1; j <= A(t->nbdat, iow); ++j) {
      int jat = t->ibdat[(j - 1) + 3 * (iow - 1)];
      if (A(t->iowat, jat) == iow) {
        double to = A(c->toat, jat) + tsh;
        if (fabs(to) > PI) to = to - copysign(PI2, to);
        A(c->toat, jat) = to;
        A(c->sntoat, jat) = sin(to);
        A(c->cstoat, jat) = cos(to);
      }
    }
  }
  for (int m = 0; m < nchains(t); ++m) {
    int a1, a2, i2, i3;
    chain_atoms(t, m, &a1, &a2, &i2, &i3);
    const double *g = c->gb + 6 * m;
    double xg[3], zg[3];
    A(c->xat, a1) = g[0];
    A(c->yat, a1) = g[1];
    A(c->zat, a1) = g[2];
    setgbl(c, g, i2, i3, xg, zg);
    A(c->xbaat, a1) = zg[0]; A(c->ybaat, a1) = zg[1]; A(c->zbaat, a1) = zg[2];
    A(c->xtoat, a1) = xg[0]; A(c->ytoat, a1) = xg[1]; A(c->ztoat, a1) = xg[2];
    for (int j = a1 + 1; j <= a2; ++j) eyring(t, c, j, A(t->iowat, j));
  }
}

/* ------------------------------------------------ enyshe.f:13-195 */

static inline double dielectric(const oracle_topo *t, double rij) {
  if (!t->epsd) return 1.0;
  double sr = t->slp * rij;
  return t->plt - (sr * sr + 2.0 * sr + 2.0) * (t->plt - 1.0) * exp(-sr) / 2.0;
}

double oracle_enyshe(const oracle_topo *t, oracle_conf *c) {
  double eyel = 0.0, eyvw = 0.0, eyhb = 0.0, eyvr = 0.0;
  int i1s = t->nms + 1;
  for (int io = t->nvt; io >= 1; --io) {
    int iv = A(t->iorvr, io), ia = A(t->iatvr, iv), it = A(t->ityvr, iv), ic = A(t->iclvr, iv);
    if (it == 3) {
      double e0 = A(t->e0to, ic);
      if (e0 != 0.0) eyvr = eyvr + e0 * (1.0 + A(t->sgto, ic) * cos(A(c->toat, ia) * A(t->rnto, ic)));
    }
    int i2s = i1s - 1;
    i1s = A(t->imsvr1, iv);
    for (int ims = i1s; ims <= i2s; ++ims) {
      for (int i = A(t->latms1, ims); i <= A(t->latms2, ims); ++i) {
        int ity = A(t->ityat, i);
        double cqi = t->conv * A(t->cgat, i);
        double xi = A(c->xat, i), yi = A(c->yat, i), zi = A(c->zat, i);
        for (int ivw = A(t->ivwat1, i); ivw <= A(t->ivwat2, i); ++ivw) {
          for (int j = A(t->lvwat1, ivw); j <= A(t->lvwat2, ivw); ++j) {
            int jty = A(t->ityat, j);
            int k = (ity - 1) + 18 * (jty - 1);
            double xij = A(c->xat, j) - xi, yij = A(c->yat, j) - yi, zij = A(c->zat, j) - zi;
            double rij2 = xij * xij + yij * yij + zij * zij;
            double rij4 = rij2 * rij2, rij6 = rij4 * rij2, rij = sqrt(rij2);
            double ep = dielectric(t, rij);
            eyel = eyel + cqi * A(t->cgat, j) / (rij * ep);
            if (t->ihbty[k] == 0)
              eyvw = eyvw + t->aij[k] / (rij6 * rij6) - t->cij[k] / rij6;
            else
              eyhb = eyhb + t->ahb[k] / (rij6 * rij6) - t->chb[k] / (rij6 * rij4);
          }
        }
        for (int i14 = A(t->i14at1, i); i14 <= A(t->i14at2, i); ++i14) {
          int j = A(t->l14at, i14);
          int jty = A(t->ityat, j);
          int k = (ity - 1) + 18 * (jty - 1);
          double xij = A(c->xat, j) - xi, yij = A(c->yat, j) - yi, zij = A(c->zat, j) - zi;
          double rij2 = xij * xij + yij * yij + zij * zij;
          double rij4 = rij2 * rij2, rij6 = rij4 * rij2, rij = sqrt(rij2);
          double ep = dielectric(t, rij);
          eyel = eyel + cqi * A(t->cgat, j) / (rij * ep);
          if (t->ihbty[k] == 0)
            eyvw = eyvw + t->a14[k] / (rij6 * rij6) - t->cij[k] / rij6;
          else
            eyhb = eyhb + t->ahb[k] / (rij6 * rij6) - t->chb[k] / (rij6 * rij4);
        }
      }
    }
  }
  double eysm = eyel + eyvw + eyhb + eyvr;
  c->e[1] = eysm;
  c->e[2] = eyel;
  c->e[3] = eyvw;
  c->e[4] = eyhb;
  c->e[5] = eyvr;
  return eysm;
}

/* ------------------------------------------------ enysol.f:13-339 */

double oracle_enysol(const oracle_topo *t, oracle_conf *c) {
  const int nlow = 1, nup = t->nat, numat = nup - nlow + 1;
  double asa = 0.0, vdvol = 0.0, eysl = 0.0;
  double avr_x = A(c->xat, nlow), avr_y = A(c->yat, nlow), avr_z = A(c->zat, nlow);
  double xmin = avr_x, ymin = avr_y, zmin = avr_z, xmax = xmin, ymax = ymin, zmax = zmin;
  double rmax = A(t->rvdw, nlow);
  for (int j = nlow + 1; j <= nup; ++j) {
    double x = A(c->xat, j), y = A(c->yat, j), z = A(c->zat, j);
    if (x <= xmin) xmin = x; else if (x >= xmax) xmax = x;
    avr_x = avr_x + x;
    if (y <= ymin) ymin = y; else if (y >= ymax) ymax = y;
    avr_y = avr_y + y;
    if (z <= zmin) zmin = z; else if (z >= zmax) zmax = z;
    avr_z = avr_z + z;
    if (A(t->rvdw, j) >= rmax) rmax = A(t->rvdw, j);
  }
  avr_x = avr_x / (double)numat;
  avr_y = avr_y / (double)numat;
  avr_z = avr_z / (double)numat;
  double diamax = 2.0 * rmax;
  double sizex = xmax - xmin, sizey = ymax - ymin, sizez = zmax - zmin;
  int ndx = (int)(sizex / diamax + 1), ndy = (int)(sizey / diamax + 1), ndz = (int)(sizez / diamax + 1);
  int nqxy = ndx * ndy, ncbox = nqxy * ndz;
  if (ncbox >= 80000) return NAN; /* the reference stops here */
  double shiftx = ((double)ndx * diamax - sizex) / 2.0;
  double shifty = ((double)ndy * diamax - sizey) / 2.0;
  double shiftz = ((double)ndz * diamax - sizez) / 2.0;
  xmin = xmin - shiftx;
  ymin = ymin - shifty;
  zmin = zmin - shiftz;
  int *inbox = (int *)calloc((size_t)ncbox + 2, sizeof(int));
  int *numbox = (int *)calloc((size_t)numat + 1, sizeof(int));
  int *indsort = (int *)calloc((size_t)numat + 1, sizeof(int));
  int *look = (int *)calloc((size_t)numat + 1, sizeof(int));
  double *xyz = (double *)calloc(3 * ((size_t)numat + 1), sizeof(double));
  double *radb2 = (double *)calloc((size_t)numat + 1, sizeof(double));
  int npnt = t->npnt;
  char *surfc = (char *)calloc((size_t)npnt + 1, 1);
  for (int j = nlow; j <= nup; ++j) {
    double fx = (A(c->xat, j) - xmin) / diamax, fy = (A(c->yat, j) - ymin) / diamax, fz = (A(c->zat, j) - zmin) / diamax;
    int mx = (int)(fx > 0.0 ? fx : 0.0), my = (int)(fy > 0.0 ? fy : 0.0), mz = (int)(fz > 0.0 ? fz : 0.0);
    if (mx > ndx - 1) mx = ndx - 1;
    if (my > ndy - 1) my = ndy - 1;
    if (mz > ndz - 1) mz = ndz - 1;
    int nboxj = mx + my * ndx + mz * nqxy + 1;
    numbox[j] = nboxj;
    inbox[nboxj] += 1;
  }
  for (int i = 1; i <= ncbox; ++i) inbox[i + 1] = inbox[i + 1] + inbox[i];
  for (int i = nlow; i <= nup; ++i) {
    int j = numbox[i], jj = inbox[j];
    indsort[jj] = i;
    inbox[j] = jj - 1;
  }
  for (int iz = 0; iz <= ndz - 1; ++iz)
    for (int iy = 0; iy <= ndy - 1; ++iy)
      for (int ix = 0; ix <= ndx - 1; ++ix) {
        int ibox = ix + iy * ndx + iz * nqxy + 1;
        int lbn = inbox[ibox + 1] - inbox[ibox];
        if (lbn <= 0) continue;
        int nsx = ix - 1 > 0 ? ix - 1 : 0, nsy = iy - 1 > 0 ? iy - 1 : 0, nsz = iz - 1 > 0 ? iz - 1 : 0;
        int nex = ix + 1 < ndx - 1 ? ix + 1 : ndx - 1, ney = iy + 1 < ndy - 1 ? iy + 1 : ndy - 1;
        int nez = iz + 1 < ndz - 1 ? iz + 1 : ndz - 1;
        int jcnt = 1;
        for (int jz = nsz; jz <= nez; ++jz)
          for (int jy = nsy; jy <= ney; ++jy)
            for (int jx = nsx; jx <= nex; ++jx) {
              int jbox = jx + jy * ndx + jz * nqxy + 1;
              for (int ii = inbox[jbox] + 1; ii <= inbox[jbox + 1]; ++ii)
                if (A(t->rvdw, indsort[ii]) > 0.0) look[jcnt++] = indsort[ii];
            }
        for (int ia = inbox[ibox] + 1; ia <= inbox[ibox + 1]; ++ia) {
          int jbi = indsort[ia];
          double trad = A(t->rvdw, jbi);
          if (!(trad > 0.0)) continue;
          int nnei = 0;
          for (int ib = 1; ib <= jcnt - 1; ++ib) {
            int jtk = look[ib];
            if (jtk == jbi) continue;
            double dx = (A(c->xat, jbi) - A(c->xat, jtk)) / trad;
            double dy = (A(c->yat, jbi) - A(c->yat, jtk)) / trad;
            double dz = (A(c->zat, jbi) - A(c->zat, jtk)) / trad;
            double dd = dx * dx + dy * dy + dz * dz;
            double akrad = A(t->rvdw, jtk) / trad;
            double dr = 1.0 + akrad;
            dr = dr * dr;
            if (dd <= dr) {
              ++nnei;
              xyz[3 * nnei + 0] = dx;
              xyz[3 * nnei + 1] = dy;
              xyz[3 * nnei + 2] = dz;
              radb2[nnei] = akrad * akrad;
            }
          }
          /* a surface point is free when no neighbour sphere holds it */
          int icount = 0;
          double sx = 0.0, sy = 0.0, sz = 0.0;
          for (int il = 1; il <= npnt; ++il) {
            double px = t->spoint[il - 1], py = t->spoint[npnt + il - 1], pz = t->spoint[2 * npnt + il - 1];
            int free_pt = 1;
            for (int ik = 1; ik <= nnei; ++ik) {
              double ax = xyz[3 * ik] + px, ay = xyz[3 * ik + 1] + py, az = xyz[3 * ik + 2] + pz;
              double sdd = 0.0;
              sdd = sdd + ax * ax;
              sdd = sdd + ay * ay;
              sdd = sdd + az * az;
              if (sdd <= radb2[ik]) {
                free_pt = 0;
                break;
              }
            }
            surfc[il] = (char)free_pt;
          }
          for (int il = 1; il <= npnt; ++il)
            if (surfc[il]) {
              ++icount;
              sx = sx + t->spoint[il - 1];
              sy = sy + t->spoint[npnt + il - 1];
              sz = sz + t->spoint[2 * npnt + il - 1];
            }
          double sdr = 4.0 * PI * trad * trad / (double)npnt;
          double area = sdr * (double)icount;
          double volume = sdr / 3.0 * (trad * (double)icount + (A(c->xat, jbi) - avr_x) * sx +
                                       (A(c->yat, jbi) - avr_y) * sy + (A(c->zat, jbi) - avr_z) * sz);
          asa = asa + area;
          vdvol = vdvol + volume;
          eysl = eysl + area * A(t->sigma, jbi);
        }
      }
  free(inbox); free(numbox); free(indsort); free(look); free(xyz); free(radb2); free(surfc);
  c->e[6] = eysl;
  c->e[7] = asa;
  c->e[8] = vdvol;
  return eysl;
}

/* eninteract.f:14-86: every atom pair between two chains with the
 * reference's terms; the sums run over atoms, not over residue blocks, so
 * they may differ from the reference's in the last bits */
double oracle_eninteract(const oracle_topo *t, oracle_conf *c) {
  double eyeli = 0.0, eyvwi = 0.0, eyhbi = 0.0;
  const int n = nchains(t);
  for (int im = 0; im < n; ++im)
    for (int jm = im + 1; jm < n; ++jm) {
      int a1, a2, b1, b2, u, v;
      chain_atoms(t, im, &a1, &a2, &u, &v);
      chain_atoms(t, jm, &b1, &b2, &u, &v);
      for (int i = a1; i <= a2; ++i) {
        const int ity = A(t->ityat, i);
        const double cqi = t->conv * A(t->cgat, i);
        const double xi = A(c->xat, i), yi = A(c->yat, i), zi = A(c->zat, i);
        for (int j = b1; j <= b2; ++j) {
          const int jty = A(t->ityat, j);
          const int k = (ity - 1) + 18 * (jty - 1);
          const double xij = A(c->xat, j) - xi, yij = A(c->yat, j) - yi, zij = A(c->zat, j) - zi;
          const double rij2 = xij * xij + yij * yij + zij * zij;
          const double rij4 = rij2 * rij2, rij6 = rij4 * rij2, rij = sqrt(rij2);
          const double ep = dielectric(t, rij);
          eyeli = eyeli + cqi * A(t->cgat, j) / (rij * ep);
          if (t->ihbty[k] == 0)
            eyvwi = eyvwi + t->aij[k] / (rij6 * rij6) - t->cij[k] / rij6;
          else
            eyhbi = eyhbi + t->ahb[k] / (rij6 * rij6) - t->chb[k] / (rij6 * rij4);
        }
      }
    }
  c->ei[0] = eyeli + eyvwi + eyhbi;
  c->ei[1] = eyeli;
  c->ei[2] = eyvwi;
  c->ei[3] = eyhbi;
  return c->ei[0];
}

/* energy.f:12-119: intra-chain terms of all chains, solvent over all atoms,
 * then the terms between chains (added to the total only, energy.f:114) */
/* The generator of esolan.f's random rotations in oracle_energy (the
 * reference draws them from the one global twister). */
static oracle_mt *g_sol_mt = NULL;
void oracle_set_solvent_mt(oracle_mt *g) { g_sol_mt = g; }

double oracle_esolan(const oracle_topo *t, oracle_conf *c, oracle_mt *g, double *as_out);

/* ------------------------------------------------ gradient.f, opeshe.f */

/* One pair of opeshe.f:150-262: adds its terms to el, vw, hb and returns
 * c = (dE/dr) / r, so that c (r_j - r_i) is the force on atom i. */
static double pair_force(const oracle_topo *t, double cqi, double qj, double rij2, double A, double Cc, int hb,
                         double *el, double *vw, double *ehb) {
  double rij4 = rij2 * rij2, rij6 = rij4 * rij2, rij = sqrt(rij2);
  double eel, deel;
  if (t->epsd) { /* opeshe.f:168-176: d(q/(r eps(r)))/dr = -deel/r */
    double sr = t->slp * rij, sr2 = sr * sr;
    double xsr = (t->plt - 1.0) * exp(-sr) / 2.0;
    double ep = t->plt - (sr2 + 2.0 * sr + 2.0) * xsr;
    eel = cqi * qj / (rij * ep);
    deel = eel + cqi * qj * (t->slp * sr2 * xsr) / (ep * ep);
  } else {
    eel = cqi * qj / rij;
    deel = eel;
  }
  *el = *el + eel;
  double rp = A / (rij6 * rij6), ds;
  if (hb) { /* 10-12 hydrogen bond */
    ds = Cc / (rij6 * rij4);
    *ehb = *ehb + rp - ds;
    return (-12.0 * rp + 10.0 * ds - deel) / rij2;
  }
  ds = Cc / rij6;
  *vw = *vw + rp - ds;
  return (-12.0 * rp + 6.0 * ds - deel) / rij2;
}

/* opeshe.f:12-388 over every chain at once (the moving sets of all chains
 * follow each other, as enyshe.f's loop above uses them): variables in
 * descending processing order; the forces on the atoms of a variable's own
 * moving set are completed by their partners' earlier subtractions, and the
 * sums of force and torque of the variables added to its set come from
 * their own rows (the recursion of Abe, Braun, Noguti and Go 1984). */
/* opesol.f:12-258 (one chain): the analytic solvent term (esolan.f, its
 * random rotations drawn from the solvent generator) and its derivatives
 * by every variable, added to gd.  Forces f = -gradan and torques f x r per
 * atom; variables in descending processing order summing their moving sets
 * and added variables, except where the variable's base atom (or, for a
 * bond length, its own atom) is the second reference atom of the chain
 * frame: there the sum runs, with the opposite sign, over the atom's vdW
 * and 1-4 partners (opesol.f:174-199).  Returns eysl (NAN on an esolan
 * failure). */
static double opesol(const oracle_topo *t, oracle_conf *c, double *gd) {
  const int nat = t->nat, nvt = t->nvt;
  double *ga = calloc(3 * (size_t)nat, sizeof(double));
  double *fr = calloc(3 * (size_t)nat, sizeof(double));
  double *fvr = calloc(6 * (size_t)nvt, sizeof(double));
  if (!ga || !fr || !fvr) {
    free(ga);
    free(fr);
    free(fvr);
    return NAN;
  }
  const double eysl = oracle_esolan_grad(t, c, g_sol_mt, NULL, ga);
  if (isnan(eysl)) {
    free(ga);
    free(fr);
    free(fvr);
    return NAN;
  }
  for (int i = 1; i <= nat; ++i) { /* f = -gradan, g = f x r */
    double *f = ga + 3 * (i - 1), *g = fr + 3 * (i - 1);
    f[0] = -f[0];
    f[1] = -f[1];
    f[2] = -f[2];
    const double xi = A(c->xat, i), yi = A(c->yat, i), zi = A(c->zat, i);
    g[0] = f[1] * zi - f[2] * yi;
    g[1] = f[2] * xi - f[0] * zi;
    g[2] = f[0] * yi - f[1] * xi;
  }
  const int ix2 = t->i2;
  int i1s = t->nms + 1, i1a = t->nad + 1;
  for (int io = nvt; io >= 1; --io) {
    const int iv = A(t->iorvr, io), it = A(t->ityvr, iv), ia = A(t->iatvr, iv), ib = A(t->iowat, ia);
    const double xb = A(c->xat, ib), yb = A(c->yat, ib), zb = A(c->zat, ib);
    double ex = 0.0, ey = 0.0, ez = 0.0;
    int lnb = 0;
    if (it == 3) {
      ex = A(c->xtoat, ib); ey = A(c->ytoat, ib); ez = A(c->ztoat, ib);
      lnb = ib == ix2;
    } else if (it == 2) {
      ex = A(c->xbaat, ia); ey = A(c->ybaat, ia); ez = A(c->zbaat, ia);
      lnb = ib == ix2;
    } else if (it == 1) {
      ex = A(c->xtoat, ia); ey = A(c->ytoat, ia); ez = A(c->ztoat, ia);
      lnb = ia == ix2;
    }
    double F[3] = {0.0, 0.0, 0.0}, T[3] = {0.0, 0.0, 0.0};
    if (!lnb) {
      const int i2s = i1s - 1;
      i1s = A(t->imsvr1, iv);
      for (int ims = i1s; ims <= i2s; ++ims)
        for (int i = A(t->latms1, ims); i <= A(t->latms2, ims); ++i)
          for (int d = 0; d < 3; ++d) {
            F[d] = F[d] + ga[3 * (i - 1) + d];
            T[d] = T[d] + fr[3 * (i - 1) + d];
          }
      const int i2a = i1a - 1;
      i1a = A(t->iadvr1, iv);
      for (int iad = i1a; iad <= i2a; ++iad) {
        const double *s = fvr + 6 * (size_t)(A(t->ladvr, iad) - 1);
        for (int d = 0; d < 3; ++d) {
          F[d] = F[d] + s[d];
          T[d] = T[d] + s[3 + d];
        }
      }
    } else {
      for (int ivw = A(t->ivwat1, ia); ivw <= A(t->ivwat2, ia); ++ivw)
        for (int j = A(t->lvwat1, ivw); j <= A(t->lvwat2, ivw); ++j)
          for (int d = 0; d < 3; ++d) {
            F[d] = F[d] - ga[3 * (j - 1) + d];
            T[d] = T[d] - fr[3 * (j - 1) + d];
          }
      for (int i14 = A(t->i14at1, ia); i14 <= A(t->i14at2, ia); ++i14) {
        const int j = A(t->l14at, i14);
        for (int d = 0; d < 3; ++d) {
          F[d] = F[d] - ga[3 * (j - 1) + d];
          T[d] = T[d] - fr[3 * (j - 1) + d];
        }
      }
    }
    double *o = fvr + 6 * (size_t)(iv - 1);
    for (int d = 0; d < 3; ++d) {
      o[d] = F[d];
      o[3 + d] = T[d];
    }
    double gs = 0.0;
    if (it == 3 || it == 2)
      gs = (ey * zb - ez * yb) * F[0] + (ez * xb - ex * zb) * F[1] + (ex * yb - ey * xb) * F[2] + ex * T[0] +
           ey * T[1] + ez * T[2];
    else if (it == 1)
      gs = -(ex * F[0] + ey * F[1] + ez * F[2]);
    gd[iv - 1] = gd[iv - 1] + gs;
  }
  free(ga);
  free(fr);
  free(fvr);
  return eysl;
}

double oracle_gradient(const oracle_topo *t, oracle_conf *c, const double *vlvr, double *gd) {
  if (t->itysol > 0 || (t->itysol < 0 && t->nml > 1)) return NAN;
  oracle_setvar(t, c, vlvr);
  const int nat = t->nat, nvt = t->nvt;
  double *fat = calloc(3 * (size_t)nat, sizeof(double));
  double *fvr = calloc(6 * (size_t)nvt, sizeof(double)); /* force and torque sums per variable */
  if (!fat || !fvr) {
    free(fat);
    free(fvr);
    return NAN;
  }
  double eyel = 0.0, eyvw = 0.0, eyhb = 0.0, eyvr = 0.0;
  int i1s = t->nms + 1, i1a = t->nad + 1;
  for (int io = nvt; io >= 1; --io) {
    const int iv = A(t->iorvr, io), ia = A(t->iatvr, iv), ib = A(t->iowat, ia);
    const int it = A(t->ityvr, iv), ic = A(t->iclvr, iv);
    const double xb = A(c->xat, ib), yb = A(c->yat, ib), zb = A(c->zat, ib);
    double ex = 0.0, ey = 0.0, ez = 0.0, fv = 0.0;
    if (it == 3) { /* torsion: axis along the bond to the base */
      ex = A(c->xtoat, ib); ey = A(c->ytoat, ib); ez = A(c->ztoat, ib);
      const double e0 = A(t->e0to, ic);
      if (e0 != 0.0) {
        const double vrn = A(c->toat, ia) * A(t->rnto, ic);
        eyvr = eyvr + e0 * (1.0 + A(t->sgto, ic) * cos(vrn));
        fv = e0 * A(t->sgto, ic) * A(t->rnto, ic) * sin(vrn); /* esnto (init_energy.f:309) */
      }
    } else if (it == 2) { /* bond angle */
      ex = A(c->xbaat, ia); ey = A(c->ybaat, ia); ez = A(c->zbaat, ia);
    } else if (it == 1) { /* bond length */
      ex = A(c->xtoat, ia); ey = A(c->ytoat, ia); ez = A(c->ztoat, ia);
    }
    double F[3] = {0.0, 0.0, 0.0}, T[3] = {0.0, 0.0, 0.0};
    const int i2s = i1s - 1;
    i1s = A(t->imsvr1, iv);
    for (int ims = i1s; ims <= i2s; ++ims) {
      for (int i = A(t->latms1, ims); i <= A(t->latms2, ims); ++i) {
        const int ity = A(t->ityat, i);
        const double cqi = t->conv * A(t->cgat, i);
        const double xi = A(c->xat, i), yi = A(c->yat, i), zi = A(c->zat, i);
        double fi[3] = {fat[3 * (i - 1)], fat[3 * (i - 1) + 1], fat[3 * (i - 1) + 2]};
        for (int pass = 0; pass < 2; ++pass) { /* vdW domains, then 1-4 partners */
          const int k1 = pass == 0 ? A(t->ivwat1, i) : A(t->i14at1, i);
          const int k2 = pass == 0 ? A(t->ivwat2, i) : A(t->i14at2, i);
          for (int k = k1; k <= k2; ++k) {
            const int j1 = pass == 0 ? A(t->lvwat1, k) : A(t->l14at, k);
            const int j2 = pass == 0 ? A(t->lvwat2, k) : A(t->l14at, k);
            for (int j = j1; j <= j2; ++j) {
              const int jty = A(t->ityat, j), q = (ity - 1) + 18 * (jty - 1);
              const double xij = A(c->xat, j) - xi, yij = A(c->yat, j) - yi, zij = A(c->zat, j) - zi;
              const double rij2 = xij * xij + yij * yij + zij * zij;
              const int hb = t->ihbty[q] != 0;
              const double Aq = hb ? t->ahb[q] : (pass == 0 ? t->aij[q] : t->a14[q]);
              const double Cq = hb ? t->chb[q] : t->cij[q];
              const double cf = pair_force(t, cqi, A(t->cgat, j), rij2, Aq, Cq, hb, &eyel, &eyvw, &eyhb);
              const double fjx = cf * xij, fjy = cf * yij, fjz = cf * zij;
              fi[0] = fi[0] + fjx;
              fi[1] = fi[1] + fjy;
              fi[2] = fi[2] + fjz;
              fat[3 * (j - 1)] = fat[3 * (j - 1)] - fjx;
              fat[3 * (j - 1) + 1] = fat[3 * (j - 1) + 1] - fjy;
              fat[3 * (j - 1) + 2] = fat[3 * (j - 1) + 2] - fjz;
            }
          }
        }
        fat[3 * (i - 1)] = fi[0];
        fat[3 * (i - 1) + 1] = fi[1];
        fat[3 * (i - 1) + 2] = fi[2];
        F[0] = F[0] + fi[0];
        F[1] = F[1] + fi[1];
        F[2] = F[2] + fi[2];
        T[0] = T[0] + (fi[1] * zi - fi[2] * yi); /* f x r */
        T[1] = T[1] + (fi[2] * xi - fi[0] * zi);
        T[2] = T[2] + (fi[0] * yi - fi[1] * xi);
      }
    }
    const int i2a = i1a - 1;
    i1a = A(t->iadvr1, iv);
    for (int iad = i1a; iad <= i2a; ++iad) {
      const double *s = fvr + 6 * (size_t)(A(t->ladvr, iad) - 1);
      for (int d = 0; d < 3; ++d) {
        F[d] = F[d] + s[d];
        T[d] = T[d] + s[3 + d];
      }
    }
    double *o = fvr + 6 * (size_t)(iv - 1);
    for (int d = 0; d < 3; ++d) {
      o[d] = F[d];
      o[3 + d] = T[d];
    }
    if (it == 3 || it == 2)
      gd[iv - 1] = (ey * zb - ez * yb) * F[0] + (ez * xb - ex * zb) * F[1] + (ex * yb - ey * xb) * F[2] +
                   ex * T[0] + ey * T[1] + ez * T[2] - fv;
    else
      gd[iv - 1] = -(ex * F[0] + ey * F[1] + ez * F[2]) - fv;
  }
  free(fat);
  free(fvr);
  const double eysm = eyel + eyvw + eyhb + eyvr;
  double eysl = 0.0;
  if (t->itysol < 0) { /* gradient.f:42-55 */
    eysl = opesol(t, c, gd);
    if (isnan(eysl)) return NAN;
  }
  c->e[1] = eysm;
  c->e[2] = eyel;
  c->e[3] = eyvw;
  c->e[4] = eyhb;
  c->e[5] = eyvr;
  c->e[6] = eysl;
  c->e[0] = eysm + eysl;
  return eysm + eysl;
}

double oracle_energy(const oracle_topo *t, oracle_conf *c, const double *vlvr) {
  oracle_setvar(t, c, vlvr);
  double esm = 0.0;
  esm = esm + oracle_enyshe(t, c);
  if (t->itysol > 0) {
    esm = esm + oracle_enysol(t, c);
  } else if (t->itysol < 0) { /* energy.f:79-82: esolan(0) over all chains */
    c->e[6] = oracle_esolan(t, c, g_sol_mt, NULL);
    esm = esm + c->e[6];
  } else {
    c->e[6] = 0.0;
  }
  esm = esm + oracle_eninteract(t, c);
  c->e[0] = esm;
  return esm;
}

/* ------------------------------------------------ weights */

double oracle_weight_eval(const oracle_weight *w, double x) {
  if (w->kind == 0) return w->beta * x; /* canon.f:135-149 */
  double xmin = (double)w->kmin, xmax = (double)w->kmax;
  double v = x / w->ebin + copysign(0.5, x);
  if (v < xmin) v = xmin;
  if (v > xmax) v = xmax;
  int k = (int)v - w->kmin;
  if (w->kind == 1) return (w->beta + w->b[k]) * x + w->alpha[k]; /* muca_weight */
  return w->b[k] * x + w->alpha[k];                                 /* muca_weight2 */
}

/* ------------------------------------------------ per-trial trace (tests) */
static double *g_trace = NULL;
static int g_trace_cap = 0, g_trace_n = 0, g_trace_rep = -1, g_cur_rep = -1;

void oracle_set_trace(double *buf, int cap, int rep) {
  g_trace = buf;
  g_trace_cap = cap;
  g_trace_n = 0;
  g_trace_rep = rep;
}
int oracle_trace_count(void) { return g_trace_n; }

static void trace_trial(int jv, double enw, double eol, int acc) {
  if (!g_trace || g_cur_rep != g_trace_rep || g_trace_n >= g_trace_cap) return;
  double *t = g_trace + 4 * g_trace_n++;
  t[0] = jv - 1;
  t[1] = enw;
  t[2] = eol;
  t[3] = acc;
}

/* ------------------------------------------------ metropolis.f:13-184 */

/* Trial order and update-type choice of metropolis.f:42-57 (/updchois/):
 * rndord draws the variable; with upchswitch 1 (updtch1, metropolis.f:
 * 196-207) or 2 (updtch2, :209-220) a variable whose name does not start
 * with 'x' draws once more and, below bgsprob or (beta - up2bmin) /
 * (up2bmax - up2bmin), takes a BGS update, which the reference leaves inert
 * (bgsnvar = 0, bgs.f:86-89: no move, no energy).  g_upd holds accanalyze's
 * ncalls(1..5), nacalls(1..5) (metropolis.f:186-193). */
static int g_rndord = 0, g_upch = 0;
static double g_bgsprob = 0.0, g_up2bmax = 4.0, g_up2bmin = 0.5, g_ubeta = -1.0;
static const int8_t *g_xname = NULL;
static double g_upd[10];

void oracle_set_updchois(int rndord, int upchswitch, double bgsprob, double up2bmax, double up2bmin,
                         const int8_t *xname, double ubeta) {
  g_rndord = rndord;
  g_upch = upchswitch;
  g_bgsprob = bgsprob;
  g_up2bmax = up2bmax;
  g_up2bmin = up2bmin;
  g_xname = xname;
  g_ubeta = ubeta;
  memset(g_upd, 0, sizeof g_upd);
}
void oracle_updstats(double *out) { memcpy(out, g_upd, sizeof g_upd); }

static void accanalyze(int iupt, int st) {
  g_upd[4] += 1.0;
  g_upd[iupt - 1] += 1.0;
  g_upd[9] += st;
  g_upd[4 + iupt] += st;
}

int oracle_metropolis(const oracle_topo *t, oracle_conf *c, oracle_mt *g, double *vlvr,
                      int nvr, const int32_t *idvr, const double *axvr,
                      const oracle_weight *w, double *eol1, double *acz) {
  double *vrol = (double *)malloc(sizeof(double) * (size_t)t->nvt);
  oracle_mt *sol_prev = g_sol_mt;
  g_sol_mt = g; /* esolan's rotations draw from the trials' twister */
  double eol = oracle_energy(t, c, vlvr);
  for (int nsw = 1; nsw <= nvr; ++nsw) {
    int ivar = nsw;
    if (g_rndord) { /* 1 + int(nvr*grnd()) (metropolis.f:42-44); u = 1 would pass the end */
      ivar = 1 + (int)(nvr * oracle_grnd(g));
      if (ivar > nvr) ivar = nvr;
    }
    int jv = idvr[ivar - 1];
    if (g_upch != 0 && !(g_xname && g_xname[jv - 1])) {
      const double bu = g_ubeta > 0.0 ? g_ubeta : w->beta;
      const double prob = g_upch == 1 ? g_bgsprob : (bu - g_up2bmin) / (g_up2bmax - g_up2bmin);
      if (oracle_grnd(g) < prob) { /* iupt = 2: bgs() returns 0 at once */
        accanalyze(2, 0);
        continue;
      }
    }
    memcpy(vrol, vlvr, sizeof(double) * (size_t)t->nvt);
    double dv = axvr[jv - 1] * (oracle_grnd(g) - 0.5);
    vlvr[jv - 1] = oracle_addang(vrol[jv - 1], dv);
    double enw = oracle_energy(t, c, vlvr);
    double delta = oracle_weight_eval(w, enw) - oracle_weight_eval(w, eol);
    double up = 0.0;
    const double eprev = eol;
    if (delta <= 0.0) {
      eol = enw;
      up = 1.0;
    } else {
      double rd = oracle_grnd(g);
      if (delta > 100.0) delta = 100.0;
      double ex = exp(-delta);
      if (ex >= rd) {
        eol = enw;
        up = 1.0;
      } else {
        memcpy(vlvr, vrol, sizeof(double) * (size_t)t->nvt);
      }
    }
    trace_trial(jv, enw, eprev, up != 0.0);
    *acz = *acz + up;
    accanalyze(1, up != 0.0);
  }
  /* rigid moves of each chain when there are several (metropolis.f:94-168):
   * three shifts of its position, then three new Euler angles */
  if (t->nml > 1) {
    for (int m = 0; m < t->nml; ++m) {
      for (int i = 0; i < 6; ++i) {
        double *gp = &c->gb[6 * m + i];
        const double gbol = *gp;
        if (i < 3) {
          *gp = *gp + (oracle_grnd(g) - 0.5);
          if (t->boxsize > 0.0) {
            if (*gp > t->boxsize)
              *gp = t->boxsize;
            else if (*gp < -t->boxsize)
              *gp = -t->boxsize;
          }
        } else {
          *gp = (oracle_grnd(g) - 0.5) * (i == 4 ? PI : PI2);
        }
        double enw = oracle_energy(t, c, vlvr);
        double delta = oracle_weight_eval(w, enw) - oracle_weight_eval(w, eol);
        double up = 0.0;
        const double eprev = eol;
        if (delta <= 0.0) {
          eol = enw;
          up = 1.0;
        } else {
          double rd = oracle_grnd(g);
          if (delta > 100.0) delta = 100.0;
          double ex = exp(-delta);
          if (ex >= rd) {
            eol = enw;
            up = 1.0;
          } else {
            *gp = gbol;
          }
        }
        trace_trial(-1 - (6 * m + i), enw, eprev, up != 0.0);
        *acz = *acz + up;
        accanalyze(i < 3 ? 3 : 4, up != 0.0);
      }
    }
  }
  double enw = oracle_energy(t, c, vlvr);
  int bad = fabs(eol - enw) > 0.000001 && eol < 100000;
  *eol1 = eol;
  free(vrol);
  g_sol_mt = sol_prev;
  return bad;
}

/* ------------------------------------------------ partem_s.f:13-258 */

int oracle_partem(const oracle_topo *t, oracle_mt *g, const double *vlvr0, int nvr,
                  const int32_t *idvr, const double *axvr, int nrep, const double *temps,
                  int nequi, int nswp, int switch_, double *eol, int32_t *ipoi, double *coor, double *traj) {
  oracle_conf c;
  if (oracle_conf_alloc(&c, t->nat)) return -1;
  oracle_conf_init(t, &c);
  double *vlvr = (double *)malloc(sizeof(double) * (size_t)t->nvt);
  double *pbe = (double *)malloc(sizeof(double) * (size_t)nrep);
  memcpy(vlvr, vlvr0, sizeof(double) * (size_t)t->nvt);
  for (int j = 0; j < nrep; ++j) {
    pbe[j] = 1.0 / (temps[j] * GASC);
    ipoi[j] = j + 1;
  }
  for (int i = 0; i < nrep; ++i)
    for (int j = 0; j < nvr; ++j) {
      int iv = idvr[j];
      double vr = 0.0; /* switch = 0 leaves the zero-initialised start */
      if (switch_ > 0) {
        double dv = axvr[iv - 1] * (oracle_grnd(g) - 0.5);
        vr = oracle_addang(PI, dv);
      } else if (switch_ < 0) {
        vr = PI;
      }
      coor[(size_t)i * nvr + j] = vr;
    }
  oracle_weight w = {0};
  int bad = 0;
  double acz = 0.0;
  oracle_mt *sol_prev = g_sol_mt;
  g_sol_mt = g;
  for (int k = 0; k < nrep; ++k) {
    g_cur_rep = k;
    w.beta = pbe[k];
    for (int i = 0; i < nvr; ++i) vlvr[idvr[i] - 1] = coor[(size_t)k * nvr + i];
    double energ = oracle_energy(t, &c, vlvr);
    for (int nsw = 1; nsw <= nequi; ++nsw) bad |= oracle_metropolis(t, &c, g, vlvr, nvr, idvr, axvr, &w, &energ, &acz);
    for (int i = 0; i < nvr; ++i) coor[(size_t)k * nvr + i] = vlvr[idvr[i] - 1];
    eol[k] = energ;
  }
  int iswitch = 1;
  for (int nsw = 1; nsw <= nswp; ++nsw) {
    for (int k1 = 1; k1 <= nrep; ++k1) {
      int k = ipoi[k1 - 1] - 1;
      g_cur_rep = k;
      w.beta = pbe[k1 - 1];
      double energ = eol[k];
      for (int i = 0; i < nvr; ++i) vlvr[idvr[i] - 1] = coor[(size_t)k * nvr + i];
      oracle_setvar(t, &c, vlvr);
      bad |= oracle_metropolis(t, &c, g, vlvr, nvr, idvr, axvr, &w, &energ, &acz);
      if (traj) traj[(size_t)(nsw - 1) * nrep + (k1 - 1)] = energ;
      for (int i = 0; i < nvr; ++i) coor[(size_t)k * nvr + i] = vlvr[idvr[i] - 1];
      eol[k] = energ;
    }
    if (!(nrep > 1)) continue;
    int num_rep1, nu;
    if (iswitch == 1) {
      num_rep1 = nrep - 1;
      nu = 1;
      iswitch = 2;
    } else {
      num_rep1 = nrep;
      nu = 2;
      iswitch = 1;
    }
    for (int i = nu; i <= num_rep1; i += 2) {
      int j = i + 1;
      if (i == nrep) j = 1;
      int in = ipoi[i - 1], jn = ipoi[j - 1];
      double delta = -pbe[i - 1] * eol[jn - 1] - pbe[j - 1] * eol[in - 1] + pbe[i - 1] * eol[in - 1] +
                     pbe[j - 1] * eol[jn - 1];
      double ra = oracle_grnd(g);
      if (ra <= exp(delta)) {
        ipoi[i - 1] = jn;
        ipoi[j - 1] = in;
      }
    }
  }
  free(vlvr);
  free(pbe);
  oracle_conf_free(&c);
  g_sol_mt = sol_prev;
  return bad;
}

/* ------------------------------------------------ mulcan_par_mod.f90:122-155
 * The reference converts the counts with FLOAT() and takes LOG() of those
 * single-precision values; the restatement does the same. */
void oracle_muca_update(int kmin, int kmax, double ebin, double beta, const int32_t *ihist,
                        double *xhist, double *g1, double *b, double *alpha) {
#define H(k) ihist[(k) - kmin]
#define X(k) xhist[(k) - kmin]
#define G(k) g1[(k) - kmin]
#define B(k) b[(k) - kmin]
#define AL(k) alpha[(k) - kmin]
  (void)beta;
  X(kmax) = X(kmax) + H(kmax);
  X(kmax - 1) = X(kmax - 1) + H(kmax - 1);
  for (int i = kmax - 2; i >= kmin; --i) {
    X(i) = X(i) + H(i);
    if (H(i) > 1 && H(i + 1) > 1) {
      double g0 = (double)((float)H(i + 1) * (float)H(i));
      g0 = g0 / (double)(float)(H(i) + H(i + 1));
      G(i) = G(i) + g0;
      g0 = g0 / G(i);
      double bi_1 = (double)logf((float)H(i + 1));
      double bi = (double)logf((float)H(i));
      B(i) = B(i) + g0 * (bi_1 - bi) / ebin;
    } else if (X(i) < 2.0) {
      B(i) = B(i + 1) > 0.0 ? B(i + 1) : 0.0;
    }
  }
  B(kmax) = B(kmax - 2);
  B(kmax - 1) = B(kmax - 2);
  AL(kmax) = 0;
  for (int i = kmax - 1; i >= kmin; --i) AL(i) = AL(i + 1) + (B(i + 1) - B(i)) * (i + 0.5) * ebin;
#undef H
#undef X
#undef G
#undef B
#undef AL
}

/* ------------------------------------------------ esolan.f:12-1050 (energy) */

/* Largest neighbour count and vertex count of the last oracle_esolan call
 * (sizes the device's per-atom buffers in the tests). */
static int g_es_max_nb = 0, g_es_max_iv = 0;
void oracle_esolan_stats(int *max_nb, int *max_iv) {
  *max_nb = g_es_max_nb;
  *max_iv = g_es_max_iv;
}

/* Test hook: >= 0 forces the random rotation of esolan.f:200-232 for every
 * atom with these (u_gamma, u_phi) in [0, 1) instead of drawing them. */
static double g_es_force_uga = -1.0, g_es_force_ufi = -1.0;
void oracle_esolan_force_rotation(double uga, double ufi) {
  g_es_force_uga = uga;
  g_es_force_ufi = ufi;
}

#define ES_MAXNB 256
#define ES_MAXV (4 * ES_MAXNB * ES_MAXNB)

/* Analytic solvent accessible surface of every atom (projection method,
 * esolan.f), energy part: returns sum sigma_i As_i and writes As (nat, may be
 * NULL).  The random rotations of esolan.f:200-232 draw from g.  Returns NAN
 * if two atom centres coincide (esolan.f:127-133 stops there). */
double oracle_esolan(const oracle_topo *t, oracle_conf *c, oracle_mt *g, double *as_out) {
  return oracle_esolan_grad(t, c, g, as_out, NULL);
}

/* d(area term)/d(a, b, c, d) of one circle -> d/d(x, y, z) of its atom
 * (ddat of that atom after the rotation): the coefficients are
 * a = x^2 + y^2 + (z - R)^2 - r^2, b = -8 R^2 x, c = -8 R^2 y,
 * d = 4 R^2 (x^2 + y^2 + (z + R)^2 - r^2) (esolan.f:545-570, 813-862). */
static void es_chain(const double dq[4], double x, double y, double z, double R, double out[3]) {
  const double R2 = R * R, R42 = 4.0 * R2, m8 = -(R42 + R42);
  out[0] += dq[0] * (2.0 * x) + dq[1] * m8 + dq[3] * (2.0 * x * R42);
  out[1] += dq[0] * (2.0 * y) + dq[2] * m8 + dq[3] * (2.0 * y * R42);
  out[2] += dq[0] * (2.0 * (z - R)) + dq[3] * (2.0 * (z + R) * R42);
}

/* Derivatives of one exposed arc's term (esolan.f:684-862).  The arc runs
 * on circle q = (a, b, c, d) from polar angle alp, where circle q1 crosses it,
 * to bet, where circle q2 does; its term depends on q's coefficients
 * directly and through both end angles, and on q1's and q2's through alp and
 * bet.  Returns d(term)/d(coefficients) of q, q1 and q2 in dq, dq1, dq2. */
static void es_arc_grad(double R, double a, double b, double c, double d, double alp, double bet, const double q1[4],
                        const double q2[4], double dq[4], double dq1[4], double dq2[4]) {
  const double R2 = R * R, R42 = 4.0 * R2;
  const double ak = a * a, daba = fabs(a), aia = a < 0.0 ? -1.0 : 1.0, b2c2 = b * b + c * c;
  const double vv = sqrt((R42 * a - d) * (R42 * a - d) + R42 * b2c2), vv1 = 1.0 / vv, vv2 = vv1 * vv1;
  const double wv = sqrt(b2c2 - 4.0 * a * d), bcd = b2c2 - 2.0 * a * d;
  const double xv = (d + R42 * a) * vv1, yv = (bcd + 2.0 * R42 * ak) / daba * vv1, zv = wv / a * vv1;
  const double sa = sin(alp), ca = cos(alp), sb = sin(bet), cb = cos(bet);
  const double hm = 5e-1 * (bet - alp), hp = 5e-1 * (bet + alp);
  const double cm = cos(hm), cp = cos(hp), sm = sin(hm), sp = sin(hp), cotm = 1.0 / tan(hm);
  const double bcp = b * cp + c * sp, bsp = -b * sp + c * cp; /* and their derivatives by hp */
  const double tt = yv * cotm - zv * bcp / sm;
  const double a1 = q1[0], b1 = q1[1], c1 = q1[2], d1 = q1[3];
  const double a2 = q2[0], b2 = q2[1], c2 = q2[2], d2 = q2[3];
  const double ab1 = a * b1 - b * a1, ac1 = a * c1 - c * a1, ab2 = a * b2 - b * a2, ac2 = a * c2 - c * a2;
  const double am1 = wv * aia * a1 * (ab1 * sa - ac1 * ca), am2 = wv * aia * a2 * (ab2 * sb - ac2 * cb);
  /* the end angles by the arc circle's coefficients */
  const double u1 = aia * a1 * (ab1 * ca + ac1 * sa) / wv, u2 = aia * a2 * (ab2 * cb + ac2 * sb) / wv;
  double dal[4], dbe[4], dal1[4], dbe2[4];
  dal[0] = (b1 * ab1 + c1 * ac1 - 2.0 * d * a1 * a1 - a * (b1 * b1 + c1 * c1 - 4.0 * a1 * d1) - 2.0 * d * u1 +
            wv * aia * a1 * (b1 * ca + c1 * sa)) / am1;
  dal[1] = (-a1 * ab1 + b * a1 * a1 + b * u1 - wv * aia * a1 * a1 * ca) / am1;
  dal[2] = (-a1 * ac1 + c * a1 * a1 + c * u1 - wv * aia * a1 * a1 * sa) / am1;
  dal[3] = (-2.0 * a * a1 * a1 - 2.0 * daba * a1 * (ab1 * ca + ac1 * sa) / wv) / am1;
  dbe[0] = (b2 * ab2 + c2 * ac2 - 2.0 * d * a2 * a2 - a * (b2 * b2 + c2 * c2 - 4.0 * a2 * d2) - 2.0 * d * u2 +
            wv * aia * a2 * (b2 * cb + c2 * sb)) / am2;
  dbe[1] = (-a2 * ab2 + b * a2 * a2 + b * u2 - wv * aia * a2 * a2 * cb) / am2;
  dbe[2] = (-a2 * ac2 + c * a2 * a2 + c * u2 - wv * aia * a2 * a2 * sb) / am2;
  dbe[3] = (-2.0 * a * a2 * a2 - 2.0 * daba * a2 * (ab2 * cb + ac2 * sb) / wv) / am2;
  /* the end angles by the crossing circles' coefficients */
  dal1[0] = (-b * ab1 - c * ac1 + wv * wv * a1 + 2.0 * ak * d1 + wv * aia * ((ab1 - b * a1) * ca + (ac1 - c * a1) * sa)) / am1;
  dal1[1] = (a * ab1 - ak * b1 + wv * daba * a1 * ca) / am1;
  dal1[2] = (a * ac1 - ak * c1 + wv * daba * a1 * sa) / am1;
  dal1[3] = (2.0 * ak * a1) / am1;
  dbe2[0] = (-b * ab2 - c * ac2 + wv * wv * a2 + 2.0 * ak * d2 + wv * aia * ((ab2 - b * a2) * cb + (ac2 - c * a2) * sb)) / am2;
  dbe2[1] = (a * ab2 - ak * b2 + wv * daba * a2 * cb) / am2;
  dbe2[2] = (a * ac2 - ak * c2 + wv * daba * a2 * sb) / am2;
  dbe2[3] = (2.0 * ak * a2) / am2;
  /* vv, xv, yv, zv by (a, b, c, d) */
  double dv[4], dx[4], dy[4], dz[4], dt[4];
  dv[1] = R42 * b * vv1;
  dv[2] = R42 * c * vv1;
  dv[3] = (d - R42 * a) * vv1;
  dv[0] = -dv[3] * R42;
  const double xn = d + R42 * a, yn = b2c2 - 2.0 * a * d + 2.0 * R42 * ak;
  dx[0] = R42 * vv1 - xn * dv[0] * vv2;
  dx[1] = -xn * dv[1] * vv2;
  dx[2] = -xn * dv[2] * vv2;
  dx[3] = vv1 - xn * dv[3] * vv2;
  dy[0] = (-2.0 * d + 4.0 * R42 * a) / daba * vv1 - yn * (aia * vv + daba * dv[0]) / ak * vv2;
  dy[1] = 2.0 * b / daba * vv1 - yn * aia * dv[1] / a * vv2;
  dy[2] = 2.0 * c / daba * vv1 - yn * aia * dv[2] / a * vv2;
  dy[3] = -2.0 * a / daba * vv1 - yn * aia * dv[3] / a * vv2;
  dz[0] = -2.0 * d / wv / a * vv1 - wv * (vv + a * dv[0]) / ak * vv2;
  dz[1] = b / wv / a * vv1 - wv * dv[1] / a * vv2;
  dz[2] = c / wv / a * vv1 - wv * dv[2] / a * vv2;
  dz[3] = -2.0 / wv * vv1 - wv * dv[3] / a * vv2;
  const double sm2 = sm * sm;
  for (int m = 0; m < 4; ++m) {
    const double extra = m == 1 ? cp * 2.0 * sm : (m == 2 ? sp * 2.0 * sm : 0.0); /* d(bcp)/db, d(bcp)/dc */
    dt[m] = cotm * dy[m] - 5e-1 * yv * (dbe[m] - dal[m]) / sm2 - (bcp / sm) * dz[m] -
            5e-1 * zv * (bsp * sm * (dal[m] + dbe[m]) - bcp * cm * (dbe[m] - dal[m]) + extra) / sm2;
  }
  const double w = 4.0 + tt * tt, at = PI - 2.0 * atan(5e-1 * tt);
  for (int m = 0; m < 4; ++m) dq[m] = R2 * (aia * (dal[m] - dbe[m]) + at * dx[m] - 4.0 * xv * dt[m] / w);
  for (int m = 0; m < 4; ++m) {
    const double ta = 5e-1 * dal1[m] * ((yv - zv * (bsp * sm + bcp * cm)) / sm2);
    const double tb = 5e-1 * dbe2[m] * ((-yv - zv * (bsp * sm - bcp * cm)) / sm2);
    dq1[m] = R2 * (aia * dal1[m] - 4.0 * xv * ta / w);
    dq2[m] = R2 * (-aia * dbe2[m] - 4.0 * xv * tb / w);
  }
}

/* The same with the Cartesian gradient of esolan.f:533-1016: gradan (3 nat,
 * atom j at 3 (j - 1); may be NULL) = sum over atoms i of sigma_i dAs_i/dr_j,
 * accumulated over i in atom order as the reference sums it. */
double oracle_esolan_grad(const oracle_topo *t, oracle_conf *c, oracle_mt *g, double *as_out, double *gradan) {
  const int numat = t->nat;
  static int nb0[ES_MAXNB + 2];
  double *gr = gradan ? (double *)calloc(3 * (size_t)(numat + 1), sizeof(double)) : NULL;
  if (gradan) memset(gradan, 0, sizeof(double) * 3 * (size_t)numat);
  static int neib[ES_MAXNB + 2], neibp[ES_MAXNB + 2], fullarc[ES_MAXNB + 2], al[2 * ES_MAXNB + 2];
  static int ta2[ES_MAXNB + 2], ivrx[ES_MAXV + 2];
  static double vertex[ES_MAXV + 1][4];
  static double ayx[2 * ES_MAXNB + 2][2], ayx1[2 * ES_MAXNB + 2], ax[2 * ES_MAXNB + 2][2], probe[2 * ES_MAXNB + 2];
  double *ddat = (double *)malloc(sizeof(double) * 4 * (size_t)(numat + 1));
  double *neibor = (double *)malloc(sizeof(double) * 4 * (size_t)(numat + 1));
  double sss = 0.0;
  g_es_max_nb = 0;
  g_es_max_iv = 0;
#define DD(k, m) ddat[4 * (k) + (m) - 1]
#define NB(k, m) neibor[4 * (k) + (m) - 1]
  for (int i = 1; i <= numat; ++i) {
    const double R = A(t->rvdw, i), poli = R * R;
    double As = 4.0 * PI * poli;
    const double xi = A(c->xat, i), yi = A(c->yat, i), zi = A(c->zat, i);
    int jj = 0, bad = 0, es_cov = 0, ifu_f = 0, ine_f = 0;
    for (int j = 1; j <= numat; ++j) {
      if (j == i) continue;
      double ddp = (A(c->xat, j) - xi) * (A(c->xat, j) - xi) + (A(c->yat, j) - yi) * (A(c->yat, j) - yi) +
                   (A(c->zat, j) - zi) * (A(c->zat, j) - zi);
      if (ddp < (double)1e-10f) { /* a default-real literal in esolan.f:127 */
        free(ddat);
        free(neibor);
        free(gr);
        return NAN;
      }
      ddp = sqrt(ddp);
      if (ddp + R <= A(t->rvdw, j)) {  /* atom i inside atom j */
        As = 0.0;
        bad = 1;
        break;
      }
      if (!(ddp <= R - A(t->rvdw, j) || ddp >= R + A(t->rvdw, j))) {
        if (jj >= ES_MAXNB) {
          free(ddat);
          free(neibor);
          free(gr);
          return NAN;
        }
        neib[++jj] = j;
      }
    }
    if (bad || jj == 0) goto atom_done;
    neib[0] = jj;
    nb0[0] = jj;
    for (int q = 1; q <= jj; ++q) nb0[q] = neib[q];
    if (gr) {
      for (int q = 0; q <= jj; ++q) {
        const int j = q == 0 ? i : nb0[q];
        gr[3 * j] = gr[3 * j + 1] = gr[3 * j + 2] = 0.0;
      }
    }
    if (jj > g_es_max_nb) g_es_max_nb = jj;
    {
      const double R2 = R * R, R22 = R2 + R2, R42 = R22 + R22, R22p = R22 * PI;
      for (int q = 1; q <= neib[0]; ++q) {
        const int k = neib[q];
        DD(k, 2) = A(c->xat, k) - xi;
        DD(k, 3) = A(c->yat, k) - yi;
        DD(k, 4) = A(c->zat, k) - zi;
        DD(k, 1) = A(t->rvdw, k);
      }
      /* the north point of sphere i against the neighbouring surfaces (esolan.f:188-200) */
      int jjj = 0;
      double cf = 1.0, sf = 0.0, cg = 1.0, sg = 0.0;
      {
        const int k = neib[1];
        const double polk = A(t->rvdw, k) * A(t->rvdw, k);
        double ddp = fabs(DD(k, 2) * DD(k, 2) + DD(k, 3) * DD(k, 3) + (R - DD(k, 4)) * (R - DD(k, 4)) - polk);
        for (int q = 2; q <= neib[0]; ++q) {
          const int kk = neib[q];
          const double pk = A(t->rvdw, kk) * A(t->rvdw, kk);
          const double d2 = fabs(DD(kk, 2) * DD(kk, 2) + DD(kk, 3) * DD(kk, 3) + (R - DD(kk, 4)) * (R - DD(kk, 4)) - pk);
          if (d2 < ddp) ddp = d2;
        }
        while (ddp < 1.0e-6 || (g_es_force_uga >= 0.0 && jjj == 0)) {
          ++jjj;
          if (g_es_force_uga < 0.0 && !g) {
            free(ddat);
            free(neibor);
            free(gr);
            return NAN;
          }
          double uga = g_es_force_uga >= 0.0 ? g_es_force_uga : oracle_grnd(g);
          double ufi = g_es_force_uga >= 0.0 ? g_es_force_ufi : oracle_grnd(g);
          uga = PI * uga / 2;
          ufi = PI * ufi / 2;
          cf = cos(ufi);
          sf = sin(ufi);
          cg = cos(uga);
          sg = sin(uga);
          const double cfsg = cf * sg, cfcg = cf * cg;
          const double xx = R * cfsg, yy = R * sf, zz = R * cfcg;
          const int k1 = neib[1];
          const double p1 = A(t->rvdw, k1) * A(t->rvdw, k1);
          ddp = fabs((xx - DD(k1, 2)) * (xx - DD(k1, 2)) + (yy - DD(k1, 3)) * (yy - DD(k1, 3)) +
                     (zz - DD(k1, 4)) * (zz - DD(k1, 4)) - p1);
          for (int q = 2; q <= neib[0]; ++q) {
            const int kk = neib[q];
            const double pk = A(t->rvdw, kk) * A(t->rvdw, kk);
            const double d2 = fabs((xx - DD(kk, 2)) * (xx - DD(kk, 2)) + (yy - DD(kk, 3)) * (yy - DD(kk, 3)) +
                                   (zz - DD(kk, 4)) * (zz - DD(kk, 4)) - pk);
            if (d2 < ddp) ddp = d2;
          }
          if (g_es_force_uga >= 0.0) break;
        }
      }
      if (jjj != 0) { /* rotate the neighbours (esolan.f:219-232) */
        const double sfsg = sf * sg, sfcg = sf * cg, cfsg = cf * sg, cfcg = cf * cg;
        for (int q = 1; q <= neib[0]; ++q) {
          const int k = neib[q];
          const double xx = DD(k, 2), yy = DD(k, 3), zz = DD(k, 4);
          DD(k, 2) = xx * cg - zz * sg;
          DD(k, 3) = -xx * sfsg + yy * cf - zz * sfcg;
          DD(k, 4) = xx * cfsg + yy * sf + zz * cfcg;
        }
      }
      /* circles a (t^2 + s^2) + b t + c s + d = 0 of the projection (esolan.f:238-247) */
      const double pom8 = 8.0 * poli;
      for (int q = 1; q <= neib[0]; ++q) {
        const int j = neib[q];
        const double pj = A(t->rvdw, j) * A(t->rvdw, j);
        NB(j, 1) = DD(j, 2) * DD(j, 2) + DD(j, 3) * DD(j, 3) + (DD(j, 4) - R) * (DD(j, 4) - R) - pj;
        NB(j, 2) = -pom8 * DD(j, 2);
        NB(j, 3) = -pom8 * DD(j, 3);
        NB(j, 4) = 4.0 * poli * (DD(j, 2) * DD(j, 2) + DD(j, 3) * DD(j, 3) + (R + DD(j, 4)) * (R + DD(j, 4)) - pj);
      }
      /* pairs of circles (esolan.f:251-451): enclosed ones dropped, crossing ones give vertices */
      int iv = 0;
      const int nb = neib[0];
      int k = neib[0];
      while (k > 1) {
        k = k - 1;
        const int lstart = neib[0];
        const int count = lstart - k > 0 ? lstart - k : 0; /* trip count fixed at entry */
        for (int it = 0, L = lstart; it < count; ++it, --L) {
          const double ak = NB(neib[k], 1), aL = NB(neib[L], 1);
          int kind = 0;
          if (ak > 0.0 && aL > 0.0) kind = 1;
          else if (ak > 0.0 && aL < 0.0) kind = 2;
          else if (ak < 0.0 && aL > 0.0) kind = 3;
          else if (ak < 0.0 && aL < 0.0) kind = 4;
          if (!kind) continue;
          const double b1 = NB(neib[k], 2) / ak, c1 = NB(neib[k], 3) / ak, d1 = NB(neib[k], 4) / ak;
          const double b2 = NB(neib[L], 2) / aL, c2 = NB(neib[L], 3) / aL, d2 = NB(neib[L], 4) / aL;
          const double D = 4.0 * ((b1 - b2) * (b2 * d1 - b1 * d2) + (c1 - c2) * (c2 * d1 - c1 * d2) - (d1 - d2) * (d1 - d2)) +
                           (b1 * c2 - b2 * c1) * (b1 * c2 - b2 * c1);
          const double dist = sqrt((b1 - b2) * (b1 - b2) + (c1 - c2) * (c1 - c2));
          const double r1 = sqrt(b1 * b1 + c1 * c1 - 4.0 * d1), r2 = sqrt(b2 * b2 + c2 * c2 - 4.0 * d2);
          int drop = 0; /* 1: drop k and leave the inner loop, 2: drop L */
          int vert = 0, covered = 0, leave = 0;
          if (kind == 1) {
            if (D <= 0.0 && dist <= r2 - r1) drop = 1;
            else if (D <= 0.0 && dist <= -r2 + r1) drop = 2;
            else if (D > 0.0) vert = 1;
          } else if (kind == 2) {
            if (D <= 0.0 && dist <= -r2 + r1) covered = 1;
            else if (D <= 0.0 && dist >= r2 + r1) drop = 1;
            else if (D > 0.0) vert = 1;
          } else if (kind == 3) {
            if (D <= 0.0 && dist <= r2 - r1) {
              As = 0.0; /* esolan.f:361-364: and the loop goes on with the next k */
              leave = 1;
            } else if (D <= 0.0 && dist >= r2 + r1) drop = 2;
            else if (D > 0.0) vert = 1;
          } else {
            if (D <= 0.0 && dist <= r2 - r1) drop = 2;
            else if (D <= 0.0 && dist <= -r2 + r1) drop = 1;
            else if (D <= 0.0) covered = 1;
            else vert = 1;
          }
          if (covered) {
            As = 0.0;
            es_cov = 1;
            goto arcs_done;
          }
          if (drop == 1) {
            neib[0] = neib[0] - 1;
            for (int q = k; q <= neib[0]; ++q) neib[q] = neib[q + 1];
            break;
          }
          if (drop == 2) {
            neib[0] = neib[0] - 1;
            for (int q = L; q <= neib[0]; ++q) neib[q] = neib[q + 1];
          }
          if (leave) break;
          if (vert) {
            if (iv + 2 > ES_MAXV) {
              free(ddat);
              free(neibor);
              free(gr);
              return NAN;
            }
            const double am = 2.0 * ((b1 - b2) * (b1 - b2) + (c1 - c2) * (c1 - c2));
            const double tt = -2.0 * (b1 - b2) * (d1 - d2) - (b2 * c1 - b1 * c2) * (c1 - c2);
            const double ss = -2.0 * (c1 - c2) * (d1 - d2) + (b2 * c1 - b1 * c2) * (b1 - b2);
            const double pm = sqrt(D), p1 = (c1 - c2) * pm, p2 = (b1 - b2) * pm;
            ++iv;
            vertex[iv][0] = (tt + p1) / am;
            vertex[iv][1] = (ss - p2) / am;
            vertex[iv][2] = neib[k];
            vertex[iv][3] = neib[L];
            ++iv;
            vertex[iv][0] = (tt - p1) / am;
            vertex[iv][1] = (ss + p2) / am;
            vertex[iv][2] = neib[k];
            vertex[iv][3] = neib[L];
          }
        }
      }
      if (iv > g_es_max_iv) g_es_max_iv = iv;
      /* lines through the north point, inner circles (esolan.f:453-503) */
      int ita2 = 0, ita3 = 0;
      for (int q = 1; q <= neib[0]; ++q) {
        if (NB(neib[q], 1) == 0.0) ta2[++ita2] = q;
        if (NB(neib[q], 1) < 0.0) ++ita3;
      }
      if (ita2 > 0 && ita3 < 1) {
        As = R22p;
        if (ita2 > 1) {
          double amial = PI, amaal = -PI, amibe = PI, amabe = -PI;
          for (int q = 1; q <= ita2; ++q) {
            const int jq = ta2[q];
            const double alp = atan2(NB(neib[jq], 3), NB(neib[jq], 2));
            double bet = alp;
            if (alp <= 0.0) bet = bet + PI2;
            if (amial > alp) amial = alp;
            if (amaal < alp) amaal = alp;
            if (amibe > bet) amibe = bet;
            if (amabe < bet) amabe = bet;
          }
          const double dal = amaal - amial, dbe = amabe - amibe;
          if (dal < PI) As = R22 * (PI - dal);
          else if (dbe < PI) As = R22 * (PI - dbe);
          else As = 0.0;
        }
      } else if ((nb > 0 && neib[0] < 1) || ita3 > 0) {
        As = 0.0;
      }
      /* circles without vertices (full arcs) and with them (esolan.f:505-526) */
      neibp[0] = neib[0];
      for (int q = 1; q <= neib[0]; ++q) neibp[q] = neib[q];
      int ifu = 0, ine = 0;
      for (int q = 1; q <= neibp[0]; ++q) {
        int has = 0;
        for (int v = 1; v <= iv; ++v)
          if (neibp[q] == (int)vertex[v][2] || neibp[q] == (int)vertex[v][3]) {
            has = 1;
            break;
          }
        if (has) neib[++ine] = neibp[q];
        else {
          ++ifu;
          fullarc[ifu] = neibp[q];
          al[ifu] = neibp[q];
        }
      }
      neib[0] = ine;
      fullarc[0] = ifu;
      al[0] = ifu;
      ifu_f = ifu;
      ine_f = ine;
      for (int q = 1; q <= ifu; ++q) { /* esolan.f:529-560 */
        const double a = NB(fullarc[q], 1), b = NB(fullarc[q], 2), cc = NB(fullarc[q], 3), d = NB(fullarc[q], 4);
        const double aia = a < 0.0 ? -1.0 : 1.0;
        if (a != 0.0) {
          As = As - aia * R22p * (1.0 + (-d / a - R42) * fabs(a) / sqrt((R42 * a - d) * (R42 * a - d) + R42 * (b * b + cc * cc)));
          if (gr) { /* esolan.f:547-568 */
            const int s = fullarc[q];
            const double div = R22 * (R22p + R22p) / pow((R42 * a - d) * (R42 * a - d) + R42 * (b * b + cc * cc), 1.5);
            const double gp[4] = {(R42 * (b * b + cc * cc - 2.0 * a * d) + 2.0 * d * d) * div, -b * (d + R42 * a) * div,
                                  -cc * (d + R42 * a) * div, (b * b + cc * cc - 2.0 * a * d + (R42 + R42) * a * a) * div};
            es_chain(gp, DD(s, 2), DD(s, 3), DD(s, 4), R, gr + 3 * s);
          }
        } else
          As = As + R22p * d / sqrt(R42 * (b * b + cc * cc) + d * d);
      }
      for (int q = 1; q <= ine; ++q) { /* esolan.f:563-957 */
        const int jk = neib[q];
        const double a = NB(jk, 1), b = NB(jk, 2), cc = NB(jk, 3), d = NB(jk, 4);
        const double ak2 = a * a, daba = fabs(a), aia = a < 0.0 ? -1.0 : 1.0, b2c2 = b * b + cc * cc;
        const double vv = sqrt((R42 * a - d) * (R42 * a - d) + R42 * b2c2), vv1 = 1.0 / vv;
        const double wv = sqrt(b2c2 - 4.0 * a * d);
        int nyx = 0, ivr = 0;
        for (int v = 1; v <= iv; ++v) {
          if ((int)vertex[v][2] == jk) {
            ++nyx;
            ayx[nyx][0] = vertex[v][0];
            ayx[nyx][1] = vertex[v][1];
            ivrx[++ivr] = v;
          }
          if ((int)vertex[v][3] == jk) {
            ++nyx;
            ayx[nyx][0] = vertex[v][0];
            ayx[nyx][1] = vertex[v][1];
            const double sw = vertex[v][2];
            vertex[v][2] = vertex[v][3];
            vertex[v][3] = sw;
            ivrx[++ivr] = v;
          }
        }
        int ial = ifu;
        for (int u = 1; u <= ine; ++u)
          if (neib[u] != jk) al[++ial] = neib[u];
        al[0] = ial;
        if (a != 0.0) {
          const double aa = 0.5 / a, ba = b * aa, ca = cc * aa, bcd = b2c2 - 2.0 * a * d;
          for (int u = 1; u <= nyx; ++u) {
            ayx[u][0] += ba;
            ayx[u][1] += ca;
          }
          const double ct = -ba, cs = -ca, rr = 0.5 * wv / daba;
          for (int u = 1; u <= nyx; ++u) ayx1[u] = atan2(ayx[u][1], ayx[u][0]);
          for (int u = 1; u <= nyx - 1; ++u) { /* selection sort of the polar angles */
            double a1 = ayx1[u];
            int jm = u;
            for (int w = u + 1; w <= nyx; ++w)
              if (a1 > ayx1[w]) {
                a1 = ayx1[w];
                jm = w;
              }
            if (jm != u) {
              a1 = ayx1[jm];
              ayx1[jm] = ayx1[u];
              ayx1[u] = a1;
              const int tmp = ivrx[jm];
              ivrx[jm] = ivrx[u];
              ivrx[u] = tmp;
            }
          }
          ayx1[nyx + 1] = ayx1[1] + PI2;
          for (int u = 1; u <= nyx; ++u) {
            if (fabs(ayx1[u + 1] - ayx1[u]) < 1e-8) continue;
            const double prob = (ayx1[u] + ayx1[u + 1]) / 2.0;
            const double ap1 = ct + rr * cos(prob), ap2 = cs + rr * sin(prob);
            int cov = 0;
            for (int w = 1; w <= ial; ++w)
              if (NB(al[w], 1) * (ap1 * ap1 + ap2 * ap2) + NB(al[w], 2) * ap1 + NB(al[w], 3) * ap2 + NB(al[w], 4) < 0.0) {
                cov = 1;
                break;
              }
            if (cov) continue;
            const double alp = ayx1[u], bet = ayx1[u + 1];
            const double dcbetmalp = cos(5e-1 * (bet - alp)), dcbetpalp = cos(5e-1 * (bet + alp));
            const double dsbetmalp = sin(5e-1 * (bet - alp)), dsbetpalp = sin(5e-1 * (bet + alp));
            const double uv = daba * (bcd + 2.0 * R42 * ak2) * dcbetmalp -
                              a * sqrt(b2c2 - 4.0 * a * d) * (b * dcbetpalp + cc * dsbetpalp);
            As = As + R2 * (aia * (alp - bet) + (d + R42 * a) * vv1 * (PI - 2.0 * atan(uv / (2.0 * ak2 * vv * dsbetmalp))));
            if (gr) { /* esolan.f:684-862: the circles crossing at the arc's ends */
              const int k1 = (int)vertex[ivrx[u]][3], k2 = (int)vertex[ivrx[u == nyx ? 1 : u + 1]][3];
              const double q1[4] = {NB(k1, 1), NB(k1, 2), NB(k1, 3), NB(k1, 4)};
              const double q2[4] = {NB(k2, 1), NB(k2, 2), NB(k2, 3), NB(k2, 4)};
              double dq[4], dq1[4], dq2[4];
              es_arc_grad(R, a, b, cc, d, alp, bet, q1, q2, dq, dq1, dq2);
              es_chain(dq, DD(jk, 2), DD(jk, 3), DD(jk, 4), R, gr + 3 * jk);
              es_chain(dq1, DD(k1, 2), DD(k1, 3), DD(k1, 4), R, gr + 3 * k1);
              es_chain(dq2, DD(k2, 2), DD(k2, 3), DD(k2, 4), R, gr + 3 * k2);
            }
          }
        } else { /* the circle is a line (esolan.f:858-957) */
          const double ang = atan2(b, cc), s1 = sin(ang), s2 = cos(ang);
          for (int u = 1; u <= nyx; ++u) {
            ax[u][0] = ayx[u][0] * s2 - ayx[u][1] * s1;
            ax[u][1] = ayx[u][0] * s1 + ayx[u][1] * s2;
          }
          for (int u = 1; u <= nyx - 1; ++u) {
            double a1 = ax[u][0];
            int jm = u;
            for (int w = u + 1; w <= nyx; ++w)
              if (a1 > ax[w][0]) {
                a1 = ax[w][0];
                jm = w;
              }
            if (jm != u) {
              double t0 = ax[jm][0], t1 = ax[jm][1];
              ax[jm][0] = ax[u][0];
              ax[jm][1] = ax[u][1];
              ax[u][0] = t0;
              ax[u][1] = t1;
              t0 = ayx[jm][0];
              t1 = ayx[jm][1];
              ayx[jm][0] = ayx[u][0];
              ayx[jm][1] = ayx[u][1];
              ayx[u][0] = t0;
              ayx[u][1] = t1;
            }
          }
          if (cc != 0.0)
            for (int u = 1; u <= nyx; ++u) ayx1[u] = (ayx[u][0] - ayx[1][0]) / cc;
          else
            for (int u = 1; u <= nyx; ++u) ayx1[u] = -(ayx[u][1] - ayx[1][1]) / b;
          probe[1] = -1.0;
          for (int u = 1; u <= nyx - 1; ++u) probe[u + 1] = (ayx1[u] + ayx1[u + 1]) / 2.0;
          probe[nyx + 1] = ayx1[nyx] + 1.0;
          const double bc = b * b + cc * cc, ccs = sqrt(R42 * bc + d * d), caba = cc * ayx[1][0] - b * ayx[1][1];
          for (int u = 1; u <= nyx + 1; ++u) {
            const double p1 = ayx[1][0] + cc * probe[u], p2 = ayx[1][1] - b * probe[u];
            int cov = 0;
            for (int w = 1; w <= ial; ++w)
              if (NB(al[w], 1) * (p1 * p1 + p2 * p2) + NB(al[w], 2) * p1 + NB(al[w], 3) * p2 + NB(al[w], 4) < 0.0) {
                cov = 1;
                break;
              }
            if (cov) continue;
            if (u == 1) {
              const double arg2 = bc * ayx1[1] + caba;
              As = As + R22 * d * (atan(arg2 / ccs) + PI / 2.0) / ccs;
            } else if (u <= nyx) {
              const double arg1 = bc * ayx1[u - 1] + caba, arg2 = bc * ayx1[u] + caba;
              As = As + R22 * d * (atan(arg2 / ccs) - atan(arg1 / ccs)) / ccs;
            } else {
              const double arg1 = bc * ayx1[nyx] + caba;
              As = As + R22 * d * (PI / 2.0 - atan(arg1 / ccs)) / ccs;
            }
          }
        }
      }
    arcs_done:
      /* esolan.f:966-1000: the atom's own derivative is minus the sum over
       * the final list (full arcs, then circles with points); with a
       * rotation, the final list's rows and its own turn back.  A circle
       * that crossed others and was dropped afterwards keeps its rotated
       * row (outside the list, as in the reference).  A covered atom has
       * no rows. */
      if (gr && !es_cov) {
        double *self = gr + 3 * i;
        for (int k3 = 0; k3 < 3; ++k3) {
          double ss = 0.0;
          for (int q = 1; q <= ifu_f; ++q) ss = ss + gr[3 * fullarc[q] + k3];
          for (int q = 1; q <= ine_f; ++q) ss = ss + gr[3 * neib[q] + k3];
          self[k3] = -ss;
        }
        if (jjj != 0) {
          const double sfsg = sf * sg, sfcg = sf * cg, cfsg = cf * sg, cfcg = cf * cg;
          for (int q = 0; q <= ifu_f + ine_f; ++q) {
            double *v = q == 0 ? self : (q <= ifu_f ? gr + 3 * fullarc[q] : gr + 3 * neib[q - ifu_f]);
            const double xx = v[0], yy = v[1], zz = v[2];
            v[0] = xx * cg - yy * sfsg + zz * cfsg;
            v[1] = yy * cf + zz * sf;
            v[2] = -xx * sg - yy * sfcg + zz * cfcg;
          }
        }
        const double sgi = A(t->sigma, i);
        for (int q = 0; q <= nb0[0]; ++q) {
          const int j = q == 0 ? i : nb0[q];
          for (int k3 = 0; k3 < 3; ++k3) gradan[3 * (j - 1) + k3] = gradan[3 * (j - 1) + k3] + sgi * gr[3 * j + k3];
        }
      }
    }
  atom_done:
    if (as_out) A(as_out, i) = As;
    sss = sss + As * A(t->sigma, i);
  }
#undef DD
#undef NB
  free(ddat);
  free(neibor);
  return sss;
}
