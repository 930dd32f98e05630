/* Oracle: EOS, viscosity, local time step, Roe-MUSCL residual, explicit update.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Every loop keeps the reference's
 * traversal order and operation order so the result can be compared with the
 * flang build of the reference routines bit for bit where the compilers agree.
 */
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define NID(b) ((b)->id2 - (b)->id1 + 1)
#define NJD(b) ((b)->jd2 - (b)->jd1 + 1)
static inline long IDX(const orc_block *b, int i, int j, int k) {
  long nid = NID(b), nijd = nid * NJD(b);
  return b->npc + (i - b->id1) + (long)(j - b->jd1) * nid + (long)(k - b->kd1) * nijd;
}
static inline long INC(const orc_block *b, int d) {
  long nid = NID(b);
  return d == 0 ? 1 : (d == 1 ? nid : nid * NJD(b));
}
/* sn(m, kdir, comp) of the block's own metric slice (snorm.f90:65-71) */
#define SN(sn, b, m, d, c) (sn)[((long)(c) * 3 + (d)) * (b)->nnn + (m)]

/* Index ranges [lo[x], hi[x]] of a loop along direction d: the along-d index
 * runs over [alo, ahi], the two others over the real cells. */
static void ranges(const orc_block *b, int d, int alo, int ahi, int lo[3], int hi[3]) {
  lo[0] = b->ii1; hi[0] = b->ii2 - 1;
  lo[1] = b->jj1; hi[1] = b->jj2 - 1;
  lo[2] = b->kk1; hi[2] = b->kk2 - 1;
  lo[d] = alo; hi[d] = ahi;
}
static inline int d1of(const orc_block *b, int d) { return d == 0 ? b->ii1 : (d == 1 ? b->jj1 : b->kk1); }
static inline int d2of(const orc_block *b, int d) { return d == 0 ? b->ii2 : (d == 1 ? b->jj2 : b->kk2); }

#define FOR3(lo, hi) \
  for (int k = (lo)[2]; k <= (hi)[2]; k++) \
    for (int j = (lo)[1]; j <= (hi)[1]; j++) \
      for (int i = (lo)[0]; i <= (hi)[0]; i++)

/* The same loop with the k planes shared among OpenMP threads: for bodies
 * that write only their own cell (or face). */
#define FOR3P(lo, hi) \
  _Pragma("omp parallel for schedule(static)") \
  FOR3(lo, hi)


/* zpres.f90:66-79 */
void orc_zpres(const orc_block *b, const orc_phys *ph, const double *v, long ldv,
               double *ps, double *temp, double *cson) {
  const double *v1 = v, *v2 = v + ldv, *v3 = v + 2 * ldv, *v4 = v + 3 * ldv, *v5 = v + 4 * ldv;
  int lo[3], hi[3];
  ranges(b, 0, b->ii1, b->ii2 - 1, lo, hi);
  FOR3P(lo, hi) {
    long n = IDX(b, i, j, k);
    double rhoe = v5[n] - 0.5 * (v2[n] * v2[n] + v3[n] * v3[n] + v4[n] * v4[n]) / v1[n];
    ps[n] = ph->gam1 * (rhoe - ph->pinfl);
    temp[n] = ph->gam * ps[n] / v1[n];
    cson[n] = sqrt(temp[n]);
  }
}

/* zvismo.f90:84-104 (iflu=1: Sutherland, iflu>=2: constant).  Returns the
 * number of cells with T <= 0, where the reference aborts (:122-127). */
int orc_zvismo(const orc_block *b, const orc_phys *ph, const double *temp, double *mu) {
  int bad = 0;
  double a = 110.4 / ph->tnz, usrey = 1. / ph->reynz;
  int lo[3], hi[3];
  ranges(b, 0, b->ii1, b->ii2 - 1, lo, hi);
  _Pragma("omp parallel for schedule(static) reduction(+ : bad)")
  FOR3(lo, hi) {
    long n = IDX(b, i, j, k);
    double t = temp[n];
    if (t <= 0.) { bad++; mu[n] = 0.; }
    else if (ph->iflu == 1) mu[n] = usrey * t * sqrt(t) * (1. + a) / (t + a);
    else mu[n] = usrey;
  }
  return bad;
}

/* chronos.f90:110-194: local time step of one block. */
void orc_chronos(const orc_block *b, const orc_phys *ph, double etal, int ns,
                 const double *v, long ldv, const double *mu, const double *mut,
                 const double *sn, const double *vol, const double *cson, double *dt) {
  int i1 = b->ii1, i2 = b->ii2, j1 = b->jj1, j2 = b->jj2, k1 = b->kk1, k2 = b->kk2;
  long nnn = b->nnn, n0 = b->npc, nci = INC(b, 0), ncj = INC(b, 1), nck = INC(b, 2);
  double *sfsi = orc_zeros(nnn), *sfsj = orc_zeros(nnn);
  double *sfsk = orc_zeros(nnn), *dism = orc_zeros(nnn);
  for (long n = IDX(b, i1, j1, k1); n <= IDX(b, i2, j2, k2); n++) dt[n] = 0.;
#pragma omp parallel for schedule(static)
  for (int k = k1; k <= k2 - 1; k++) {
    for (int j = j1; j <= j2 - 1; j++)
      for (int i = i1; i <= i2; i++) {
        long m = IDX(b, i, j, k) - n0;
        sfsi[m] = SN(sn, b, m, 0, 0) * SN(sn, b, m, 0, 0) + SN(sn, b, m, 0, 1) * SN(sn, b, m, 0, 1) +
                  SN(sn, b, m, 0, 2) * SN(sn, b, m, 0, 2);
      }
    for (int j = j1; j <= j2; j++)
      for (int i = i1; i <= i2 - 1; i++) {
        long m = IDX(b, i, j, k) - n0;
        sfsj[m] = SN(sn, b, m, 1, 0) * SN(sn, b, m, 1, 0) + SN(sn, b, m, 1, 1) * SN(sn, b, m, 1, 1) +
                  SN(sn, b, m, 1, 2) * SN(sn, b, m, 1, 2);
      }
  }
#pragma omp parallel for schedule(static)
  for (int k = k1; k <= k2; k++)
    for (int j = j1; j <= j2 - 1; j++)
      for (int i = i1; i <= i2 - 1; i++) {
        long m = IDX(b, i, j, k) - n0;
        sfsk[m] = SN(sn, b, m, 2, 0) * SN(sn, b, m, 2, 0) + SN(sn, b, m, 2, 1) * SN(sn, b, m, 2, 1) +
                  SN(sn, b, m, 2, 2) * SN(sn, b, m, 2, 2);
      }
#pragma omp parallel for schedule(static)
  for (int k = k1; k <= k2 - 1; k++)
    for (int j = j1; j <= j2 - 1; j++)
      for (int i = i1; i <= i2 - 1; i++) {
        long mc = IDX(b, i, j, k) - n0;
        double simax = fmax(sfsi[mc + nci], sfsi[mc]);
        double sjmax = fmax(sfsj[mc + ncj], sfsj[mc]);
        double skmax = fmax(sfsk[mc + nck], sfsk[mc]);
        double smoy = sqrt(simax + sjmax + skmax);
        dism[mc] = vol[mc + n0] / smoy;
      }
  const double *t1 = v, *t2 = v + ldv, *t3 = v + 2 * ldv, *t4 = v + 3 * ldv;
#pragma omp parallel for schedule(static)
  for (int k = k1; k <= k2 - 1; k++)
    for (int j = j1; j <= j2 - 1; j++)
      for (int i = i1; i <= i2 - 1; i++) {
        long nc = IDX(b, i, j, k), mc = nc - n0;
        double ro = t1[nc];
        double qq = (t2[nc] * t2[nc] + t3[nc] * t3[nc] + t4[nc] * t4[nc]) / (ro * ro);
        double q = sqrt(qq);
        double dte = dism[mc] / (q + cson[nc]);
        dt[nc] = etal * dte;
      }
  if (ns) {
#pragma omp parallel for schedule(static)
    for (int k = k1; k <= k2 - 1; k++)
      for (int j = j1; j <= j2 - 1; j++)
        for (int i = i1; i <= i2 - 1; i++) {
          long nc = IDX(b, i, j, k), mc = nc - n0;
          double rv1 = ph->gam * (mu[nc] / ph->pr + mut[nc] / ph->prt) / t1[nc];
          double rv = fmax(rv1, 4. / 3. * (mu[nc] + mut[nc]) / t1[nc]);
          double dtv = etal * 0.25 * (dism[mc] * dism[mc]) / rv;
          dt[nc] = fmin(dt[nc], dtv);
        }
  }
  free(sfsi); free(sfsj); free(sfsk); free(dism);
}

/* chrono.f90:105-153: with kdtl=0 the step is global.  dtmin over the real
 * cells of all blocks; the reference stops when it is below dt1min, else every
 * stored entry of every block gets min(dt1min, dtmin) = dt1min.  Returns 1 on
 * the stop condition (dt untouched), 0 otherwise. */
int orc_chrono_global(const orc_block *blocks, int nb, double dt1min, double *dt, double *dtmin_out) {
  double dtmin = DBL_MAX;
  for (int l = 0; l < nb; l++) {
    const orc_block *b = &blocks[l];
    for (int k = b->kk1; k <= b->kk2 - 1; k++)
      for (int j = b->jj1; j <= b->jj2 - 1; j++)
        for (int i = b->ii1; i <= b->ii2 - 1; i++) dtmin = fmin(dtmin, dt[IDX(b, i, j, k)]);
  }
  if (dtmin_out) *dtmin_out = dtmin;
  dtmin = fmin(dt1min, dtmin);
  if (dtmin < dt1min) return 1;
  for (int l = 0; l < nb; l++) {
    const orc_block *b = &blocks[l];
    for (long n = IDX(b, b->id1, b->jd1, b->kd1); n <= IDX(b, b->id2, b->jd2, b->kd2); n++) dt[n] = dtmin;
  }
  return 0;
}

/* sch_expli.f90:53-66: v += dt/vol * u on the real cells (the reference's
 * sign convention: u is the net outgoing flux). */
void orc_sch_expli(const orc_block *b, const double *u, double *v, long ldv, const double *dt,
                   const double *vol) {
  int lo[3], hi[3];
  ranges(b, 0, b->ii1, b->ii2 - 1, lo, hi);
  FOR3P(lo, hi) {
    long n = IDX(b, i, j, k);
    double dtv = dt[n] / vol[n];
    for (int e = 0; e < 5; e++) v[e * ldv + n] = v[e * ldv + n] + dtv * u[e * ldv + n];
  }
}

/* One stage of the 4-stage explicit scheme composed on the residual (the
 * reference has the single-stage update above only): v = v0 - alpha dt/vol u
 * on the real cells, nvar planes, with the physical sign. */
void orc_rk_stage(const orc_block *b, double alpha, const double *v0, const double *u, double *v, long ldv,
                  int nvar, const double *dt, const double *vol) {
  int lo[3], hi[3];
  ranges(b, 0, b->ii1, b->ii2 - 1, lo, hi);
  FOR3P(lo, hi) {
    long n = IDX(b, i, j, k);
    double c = alpha * (dt[n] / vol[n]);
    for (int e = 0; e < nvar; e++) v[e * ldv + n] = v0[e * ldv + n] - c * u[e * ldv + n];
  }
}

/* ------------------------------------------------------------------------ */
/* Roe flux of one face from the left/right primitive states, in the explicit
 * P |Lambda| Q form of sch_roe.f90:327-473 / sch_roe_euler.f90:181-320.
 * tl/tr: stresses (xx,xy,xz,yy,yz,zz) and heat flux (x,y,z) of the two cells,
 * or NULL for the Euler form.  Returns f[5] = 2x the numerical flux . S. */
static void roe_face(const orc_phys *ph, const double s[3], double rhol, double ul, double vl,
                     double wl, double pl, double rhor, double ur, double vr, double wr, double prr,
                     const double *tl, const double *tr, double f[5]) {
  const double gam = ph->gam, gam1 = ph->gam1, pinfl = ph->pinfl;
  double cnds = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  double nx = s[0] / cnds, ny = s[1] / cnds, nz = s[2] / cnds;
  double al = sqrt(gam * pl / rhol), ar = sqrt(gam * prr / rhor);
  double q2l = ul * ul + vl * vl + wl * wl, q2r = ur * ur + vr * vr + wr * wr;
  double hl = al * al / gam1 + 0.5 * q2l, hr = ar * ar / gam1 + 0.5 * q2r;
  double el = pl / (gam1 * rhol) + 0.5 * q2l + pinfl / rhol;
  double er = prr / (gam1 * rhor) + 0.5 * q2r + pinfl / rhor;
  double gd = sqrt(rhor / rhol), gd1 = 1. / (1. + gd), gd2 = gd * gd1;
  double rhom = sqrt(rhol * rhor), rhomi = 1. / rhom;
  double um = gd1 * ul + gd2 * ur, vm = gd1 * vl + gd2 * vr, wm = gd1 * wl + gd2 * wr;
  double hm = gd1 * hl + gd2 * hr;
  double vitm2 = 0.5 * (um * um + vm * vm + wm * wm);
  double am = sqrt(gam1 * (hm - vitm2));
  double am2i = 1. / (am * am);
  double vn = um * nx + vm * ny + wm * nz;
  double rhoiam = rhom / am, rhoami = am2i / rhoiam;
  double lam[5];
  lam[0] = lam[1] = lam[2] = fabs(vn * cnds);
  lam[3] = fabs(vn * cnds + am * cnds);
  lam[4] = fabs(vn * cnds - am * cnds);
  double Q[5][5], P[5][5];
  double g = gam1 * vitm2 * am2i;
  Q[0][0] = (1. - g) * nx - (vm * nz - wm * ny) * rhomi;
  Q[0][1] = gam1 * um * nx * am2i;
  Q[0][2] = gam1 * vm * nx * am2i + nz * rhomi;
  Q[0][3] = gam1 * wm * nx * am2i - ny * rhomi;
  Q[0][4] = -gam1 * nx * am2i;
  Q[1][0] = (1. - g) * ny - (wm * nx - um * nz) * rhomi;
  Q[1][1] = gam1 * um * ny * am2i - nz * rhomi;
  Q[1][2] = gam1 * vm * ny * am2i;
  Q[1][3] = gam1 * wm * ny * am2i + nx * rhomi;
  Q[1][4] = -gam1 * ny * am2i;
  Q[2][0] = (1. - g) * nz - (um * ny - vm * nx) * rhomi;
  Q[2][1] = gam1 * um * nz * am2i + ny * rhomi;
  Q[2][2] = gam1 * vm * nz * am2i - nx * rhomi;
  Q[2][3] = gam1 * wm * nz * am2i;
  Q[2][4] = -gam1 * nz * am2i;
  Q[3][0] = gam1 * vitm2 * rhoami - vn * rhomi;
  Q[3][1] = nx * rhomi - gam1 * um * rhoami;
  Q[3][2] = ny * rhomi - gam1 * vm * rhoami;
  Q[3][3] = nz * rhomi - gam1 * wm * rhoami;
  Q[3][4] = gam1 * rhoami;
  Q[4][0] = gam1 * vitm2 * rhoami + vn * rhomi;
  Q[4][1] = -nx * rhomi - gam1 * um * rhoami;
  Q[4][2] = -ny * rhomi - gam1 * vm * rhoami;
  Q[4][3] = -nz * rhomi - gam1 * wm * rhoami;
  Q[4][4] = gam1 * rhoami;
  P[0][0] = nx; P[0][1] = ny; P[0][2] = nz; P[0][3] = 0.5 * rhoiam; P[0][4] = 0.5 * rhoiam;
  P[1][0] = um * nx; P[1][1] = um * ny - rhom * nz; P[1][2] = um * nz + rhom * ny;
  P[1][3] = 0.5 * rhoiam * (um + nx * am); P[1][4] = 0.5 * rhoiam * (um - nx * am);
  P[2][0] = vm * nx + rhom * nz; P[2][1] = vm * ny; P[2][2] = vm * nz - rhom * nx;
  P[2][3] = 0.5 * rhoiam * (vm + ny * am); P[2][4] = 0.5 * rhoiam * (vm - ny * am);
  P[3][0] = wm * nx - rhom * ny; P[3][1] = wm * ny + rhom * nx; P[3][2] = wm * nz;
  P[3][3] = 0.5 * rhoiam * (wm + nz * am); P[3][4] = 0.5 * rhoiam * (wm - nz * am);
  P[4][0] = vitm2 * nx + rhom * (vm * nz - wm * ny);
  P[4][1] = vitm2 * ny + rhom * (wm * nx - um * nz);
  P[4][2] = vitm2 * nz + rhom * (um * ny - vm * nx);
  P[4][3] = 0.5 * rhoiam * (hm + am * vn); P[4][4] = 0.5 * rhoiam * (hm - am * vn);
  double dw[5];
  dw[0] = rhor - rhol;
  dw[1] = rhor * ur - rhol * ul;
  dw[2] = rhor * vr - rhol * vl;
  dw[3] = rhor * wr - rhol * wl;
  dw[4] = rhor * er - rhol * el;
  double di[5];
  for (int r = 0; r < 5; r++) {
    double acc = 0.;
    for (int c = 0; c < 5; c++) {
      double a = P[r][0] * lam[0] * Q[0][c];
      for (int s_ = 1; s_ < 5; s_++) a = a + P[r][s_] * lam[s_] * Q[s_][c];
      acc = c == 0 ? a * dw[c] : acc + a * dw[c];
    }
    di[r] = acc;
  }
  static const double z9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const double *L = tl ? tl : z9, *R = tr ? tr : z9;
  /* L/R[0..5] = txx txy txz tyy tyz tzz, [6..8] = qx qy qz */
  double dfxx = rhor * (ur * ur) + prr - pinfl - R[0] + rhol * (ul * ul) + pl - pinfl - L[0];
  double dfxy = rhor * ur * vr - R[1] + rhol * ul * vl - L[1];
  double dfxz = rhor * ur * wr - R[2] + rhol * ul * wl - L[2];
  double dfyy = rhor * (vr * vr) + prr - pinfl - R[3] + rhol * (vl * vl) + pl - pinfl - L[3];
  double dfyz = rhor * vr * wr - R[4] + rhol * vl * wl - L[4];
  double dfzz = rhor * (wr * wr) + prr - pinfl - R[5] + rhol * (wl * wl) + pl - pinfl - L[5];
  double dfex = (rhor * er + prr - pinfl) * ur + (rhol * el + pl - pinfl) * ul -
                (R[0] * ur + R[1] * vr + R[2] * wr) - (L[0] * ul + L[1] * vl + L[2] * wl) - R[6] - L[6];
  double dfey = (rhor * er + prr - pinfl) * vr + (rhol * el + pl - pinfl) * vl -
                (R[1] * ur + R[3] * vr + R[4] * wr) - (L[1] * ul + L[3] * vl + L[4] * wl) - R[7] - L[7];
  double dfez = (rhor * er + prr - pinfl) * wr + (rhol * el + pl - pinfl) * wl -
                (R[2] * ur + R[4] * vr + R[5] * wr) - (L[2] * ul + L[4] * vl + L[5] * wl) - R[8] - L[8];
  f[0] = (rhor * ur + rhol * ul) * s[0] + (rhor * vr + rhol * vl) * s[1] + (rhor * wr + rhol * wl) * s[2] - di[0];
  f[1] = dfxx * s[0] + dfxy * s[1] + dfxz * s[2] - di[1];
  f[2] = dfxy * s[0] + dfyy * s[1] + dfyz * s[2] - di[2];
  f[3] = dfxz * s[0] + dfyz * s[1] + dfzz * s[2] - di[3];
  f[4] = dfex * s[0] + dfey * s[1] + dfez * s[2] - di[4];
}

/* Physical flux . S of a boundary (ghost) state, sch_roe.f90:497-578.
 * rho_div is the density the energy flux is divided by: the ghost's own in
 * the Navier-Stokes form; sch_roe_euler.f90:630-632,912-914 divide by the
 * adjacent real cell on the low j and k faces. */
static void phys_flux(const orc_phys *ph, const double w[5], double p, const double *t,
                      double rho_div, int euler_form, const double s[3], double f[5]) {
  const double pinfl = ph->pinfl;
  static const double z9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const double *T = t ? t : z9;
  double fxx = w[1] * (w[1] / w[0]) + p - pinfl - T[0];
  double fxy = w[2] * (w[1] / w[0]) - T[1];
  double fxz = w[3] * (w[1] / w[0]) - T[2];
  double fyy = w[2] * (w[2] / w[0]) + p - pinfl - T[3];
  double fyz = w[3] * (w[2] / w[0]) - T[4];
  double fzz = w[3] * (w[3] / w[0]) + p - pinfl - T[5];
  double fex, fey, fez;
  if (euler_form) {
    fex = (w[4] + p - pinfl) * w[1] / rho_div;
    fey = (w[4] + p - pinfl) * w[2] / rho_div;
    fez = (w[4] + p - pinfl) * w[3] / rho_div;
  } else {
    fex = ((w[4] + p - pinfl - T[0]) * w[1] - T[1] * w[2] - T[2] * w[3]) / rho_div - T[6];
    fey = ((w[4] + p - pinfl - T[3]) * w[2] - T[1] * w[1] - T[4] * w[3]) / rho_div - T[7];
    fez = ((w[4] + p - pinfl - T[5]) * w[3] - T[2] * w[1] - T[4] * w[2]) / rho_div - T[8];
  }
  f[0] = w[1] * s[0] + w[2] * s[1] + w[3] * s[2];
  f[1] = fxx * s[0] + fxy * s[1] + fxz * s[2];
  f[2] = fxy * s[0] + fyy * s[1] + fyz * s[2];
  f[3] = fxz * s[0] + fyz * s[1] + fzz * s[2];
  f[4] = fex * s[0] + fey * s[1] + fez * s[2];
}

static inline double phi_minmod(double a) { /* sch_roe.f90:1500-1504 */
  double m = 1. < a ? 1. : a;
  return 0. > m ? 0. : m;
}

/* MUSCL primitive states (rho, u, v, w, p) left/right of the face between
 * n-inc and n, unlimited kappa scheme (sch_roe_euler.f90:143-171). */
static void muscl_unlimited(const double *v, long ldv, const double *ps, long n, long inc,
                            double muscl, double xk, double L[5], double R[5]) {
  const double *v1 = v;
  double q[4][5]; /* q[s][e] at n-2inc, n-inc, n, n+inc */
  for (int s_ = 0; s_ < 4; s_++) {
    long nn = n + (s_ - 2) * inc;
    q[s_][0] = v1[nn];
    for (int e = 1; e < 4; e++) q[s_][e] = v[e * ldv + nn] / v1[nn];
    q[s_][4] = ps[nn];
  }
  for (int e = 0; e < 5; e++) {
    L[e] = q[1][e] + 0.25 * muscl * ((1. - xk) * (q[1][e] - q[0][e]) + (1. + xk) * (q[2][e] - q[1][e]));
    R[e] = q[2][e] - 0.25 * muscl * ((1. + xk) * (q[2][e] - q[1][e]) + (1. - xk) * (q[3][e] - q[2][e]));
  }
}

/* Limited states with the chained slope ratios r[e][m] (sch_roe.f90:196-241). */
static void muscl_minmod(const double *v, long ldv, const double *ps, long n, long inc, long m,
                         double *const r[5], double muscl, double xk, double L[5], double R[5]) {
  const double *v1 = v;
  double q[4][5];
  for (int s_ = 0; s_ < 4; s_++) {
    long nn = n + (s_ - 2) * inc;
    q[s_][0] = v1[nn];
    for (int e = 1; e < 4; e++) q[s_][e] = v[e * ldv + nn] / v1[nn];
    q[s_][4] = ps[nn];
  }
  double c = 0.25 * muscl;
  for (int e = 0; e < 5; e++) {
    double rl = r[e][m], rr = r[e][m + inc];
    if (e == 0 || e == 4) {
      L[e] = q[1][e] + c * phi_minmod(rl) * (1. - xk) * (q[1][e] - q[0][e]) +
             c * phi_minmod(1. / rl) * (1. + xk) * (q[2][e] - q[1][e]);
      R[e] = q[2][e] - c * phi_minmod(rr) * (1. + xk) * (q[2][e] - q[1][e]) -
             c * phi_minmod(1. / rr) * (1. - xk) * (q[3][e] - q[2][e]);
    } else {
      L[e] = q[1][e] + c * phi_minmod(rl) * (1. - xk) * (q[1][e] - q[0][e]) +
             (1. + xk) * (q[2][e] - q[1][e]) * 0.25 * muscl * phi_minmod(1. / rl);
      R[e] = q[2][e] - c * phi_minmod(rr) * (1. + xk) * (q[2][e] - q[1][e]) -
             (1. - xk) * (q[3][e] - q[2][e]) * 0.25 * muscl * phi_minmod(1. / rr);
    }
  }
}

/* Shared body of sch_roe (ns=1) and sch_roe_euler (ns=0).
 * tq: 9 arrays (txx txy txz tyy tyz tzz qx qy qz) of length ldv, or NULL. */
static void roe_residual(const orc_block *b, const orc_phys *ph, const orc_sch *sc, int ns,
                         int ityprk, double *u, const double *v, long ldv, double *ff,
                         const double *tq, const double *sn, const double *ps) {
  long n0 = b->npc, nnn = b->nnn;
  double muscl = (double)sc->muscl, xk = sc->xk;
  int ilim = ns ? sc->ilim : 0;
  /* zero u over the block storage (sch_roe.f90:113-125) */
  for (long n = IDX(b, b->ii1 - 1, b->jj1 - 1, b->kk1 - 1); n <= IDX(b, b->ii2, b->jj2, b->kk2); n++)
    for (int e = 0; e < 5; e++) u[e * ldv + n] = 0.;
  double *r[5];
  for (int e = 0; e < 5; e++) r[e] = orc_zeros(nnn);
  for (int d = 0; d < 3; d++) {
    long inc = INC(b, d);
    int a1 = d1of(b, d), a2 = d2of(b, d);
    int lo[3], hi[3];
    if (ilim == 1) {
      /* slope differences on faces a1..a2 (sch_roe.f90:146-158) */
      ranges(b, d, a1, a2, lo, hi);
      FOR3P(lo, hi) {
        long n = IDX(b, i, j, k), m = n - n0;
        r[0][m] = v[n] - v[n - inc];
        for (int e = 1; e < 4; e++) r[e][m] = v[e * ldv + n] / v[n] - v[e * ldv + n - inc] / v[n - inc];
        r[4][m] = ps[n] - ps[n - inc];
      }
      /* in-place ratios on faces a1+1..a2-1 in storage order: the value read
       * at m-inc has already been replaced by its own ratio (:161-194) */
      ranges(b, d, a1 + 1, a2 - 1, lo, hi);
      FOR3D(lo, hi, d) {
        long m = IDX(b, i, j, k) - n0;
        for (int e = 0; e < 5; e++) {
          if (fabs(r[e][m - inc] * r[e][m] * r[e][m + inc]) <= DBL_MIN) r[e][m] = 1.;
          else r[e][m] = r[e][m] / r[e][m - inc];
        }
      }
    }
    /* interior faces a1+1..a2-1 (:322-495) */
    ranges(b, d, a1 + 1, a2 - 1, lo, hi);
    FOR3D(lo, hi, d) {
      long n = IDX(b, i, j, k), m = n - n0;
      int along = d == 0 ? i : (d == 1 ? j : k);
      double L[5], R[5];
      if (ilim == 1 && along <= a2 - 2) muscl_minmod(v, ldv, ps, n, inc, m, r, muscl, xk, L, R);
      else muscl_unlimited(v, ldv, ps, n, inc, muscl, xk, L, R);
      double s[3] = {SN(sn, b, m, d, 0), SN(sn, b, m, d, 1), SN(sn, b, m, d, 2)};
      double tl[9], tr[9], f[5];
      if (tq) for (int c = 0; c < 9; c++) { tl[c] = tq[c * ldv + n - inc]; tr[c] = tq[c * ldv + n]; }
      roe_face(ph, s, L[0], L[1], L[2], L[3], L[4], R[0], R[1], R[2], R[3], R[4],
               tq ? tl : NULL, tq ? tr : NULL, f);
      for (int e = 0; e < 5; e++) u[e * ldv + n] = u[e * ldv + n] - 0.5 * f[e];
      for (int e = 0; e < 5; e++) u[e * ldv + n - inc] = u[e * ldv + n - inc] + 0.5 * f[e];
    }
    /* boundary faces: physical flux of the ghost state (:497-578) */
    for (int side = 0; side < 2; side++) {
      int a = side == 0 ? a1 : a2;
      ranges(b, d, a, a, lo, hi);
      FOR3(lo, hi) {
        long n = IDX(b, i, j, k), m = n - n0;
        long ng = side == 0 ? n - inc : n;       /* ghost cell */
        long nr = side == 0 ? n : n - inc;       /* real cell */
        double w[5], s[3] = {SN(sn, b, m, d, 0), SN(sn, b, m, d, 1), SN(sn, b, m, d, 2)};
        for (int e = 0; e < 5; e++) w[e] = v[e * ldv + ng];
        double t[9], f[5];
        if (tq) for (int c = 0; c < 9; c++) t[c] = tq[c * ldv + ng];
        double rdiv = (!ns && side == 0 && d > 0) ? v[nr] : w[0];
        phys_flux(ph, w, ps[ng], tq ? t : NULL, rdiv, !ns, s, f);
        if (side == 0) for (int e = 0; e < 5; e++) u[e * ldv + nr] = u[e * ldv + nr] - f[e];
        else for (int e = 0; e < 5; e++) u[e * ldv + nr] = u[e * ldv + nr] + f[e];
      }
    }
  }
  if (ityprk != 0) {
    int lo[3], hi[3];
    ranges(b, 0, b->ii1, b->ii2 - 1, lo, hi);
    FOR3(lo, hi) {
      long n = IDX(b, i, j, k);
      for (int e = 0; e < 5; e++) ff[e * ldv + n] = ff[e * ldv + n] - u[e * ldv + n];
    }
  }
  for (int e = 0; e < 5; e++) free(r[e]);
}

/* sch_roe_euler.f90:4 */
void orc_sch_roe_euler(const orc_block *b, const orc_phys *ph, const orc_sch *sc, int ityprk,
                       double *u, const double *v, long ldv, double *ff, const double *sn,
                       const double *ps) {
  roe_residual(b, ph, sc, 0, ityprk, u, v, ldv, ff, NULL, sn, ps);
}

/* sch_roe.f90:4 */
void orc_sch_roe(const orc_block *b, const orc_phys *ph, const orc_sch *sc, int ityprk, double *u,
                 const double *v, long ldv, double *ff, const double *tq, const double *sn,
                 const double *ps) {
  roe_residual(b, ph, sc, 1, ityprk, u, v, ldv, ff, tq, sn, ps);
}

/* Test hook: the Roe face flux of two primitive states (rho,u,v,w,p) with
 * the Euler form (no stresses); f = 2x numerical flux . S. */
void orc_roe_face(const orc_phys *ph, const double *s, const double *L, const double *R, double *f) {
  roe_face(ph, s, L[0], L[1], L[2], L[3], L[4], R[0], R[1], R[2], R[3], R[4], NULL, NULL, f);
}
