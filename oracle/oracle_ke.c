/* CPU restatement of the k-epsilon part of the turbulence stage (met_num.f90
 * ksecmb <= 1, :352-367, and the eddy viscosity met_kemut): met_smke with
 * met_exgr, met_bark, met_bare and its variants met_smkes (' ', 'S'),
 * met_smker ('R'), met_smkec ('C') and met_smkesas ('L').
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  v holds 7 planes of ldv; g9
 * (velocity gradients), tprod, bark and bare are block-local (m = n - npc);
 * cfke, ts6, ts7, mut are storage arrays.  kc = cke1, cke2, cmu
 * (atlecdon.f90:124-128).  Integer powers x**2 are x*x, as flang evaluates
 * them; x**0.5 and x**1.5 are pow(). */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

#define NID(b) ((long)((b)->id2 - (b)->id1 + 1))
#define NIJD(b) (NID(b) * (long)((b)->jd2 - (b)->jd1 + 1))
#define IDX(b, i, j, k) \
  ((long)(b)->npc + ((i) - (b)->id1) + (long)((j) - (b)->jd1) * NID(b) + (long)((k) - (b)->kd1) * NIJD(b))

/* restatements in the other oracle files */
void orc_teq_grads(const orc_block *b, const double *sn, const double *vol, const double *s, double *g);
void orc_teq_gradv(const orc_block *b, const double *sn, const double *vol, double *v, long ldv, double *g9);
void orc_met_laplaciens(const orc_block *b, int ischema, const double *sn, const double *vol, const double *dsx,
                        const double *dsy, const double *dsz, double *dx, double *dy, double *dz);
void orc_met_bark(const orc_block *b, const double *sn, const double *vol, const double *v, long ldv,
                  const double *mu, double *bark);

static double fmax2(double a, double b) { return a > b ? a : b; }

/* met_exgr.f90:47-87 -- ghost layers of three block-local arrays copied from
 * the adjacent layer (ex1 = 1, ex2 = 0): the k faces over the real i and j,
 * the i faces over k1-1..k2 and the real j, the j faces over k1-1..k2 and
 * i1-1..i2, in that order. */
void orc_met_exgr(const orc_block *b, double *d1, double *d2, double *d3) {
  long n0 = b->npc, nid = NID(b), nijd = NIJD(b);
  double *ds[3] = {d1, d2, d3};
  int i1 = b->ii1, i2 = b->ii2, j1 = b->jj1, j2 = b->jj2, k1 = b->kk1, k2 = b->kk2;
  for (int side = 0; side < 2; side++) {
    int k = side == 0 ? k1 - 1 : k2, ks = side == 0 ? 1 : -1;
    for (int j = j1; j <= j2 - 1; j++)
      for (int i = i1; i <= i2 - 1; i++) {
        long m = IDX(b, i, j, k) - n0;
        for (int c = 0; c < 3; c++) ds[c][m] = 1. * ds[c][m + ks * nijd] + 0. * ds[c][m + 2 * ks * nijd];
      }
  }
  for (int k = k1 - 1; k <= k2; k++)
    for (int j = j1; j <= j2 - 1; j++)
      for (int side = 0; side < 2; side++) {
        int i = side == 0 ? i1 - 1 : i2, is = side == 0 ? 1 : -1;
        long m = IDX(b, i, j, k) - n0;
        for (int c = 0; c < 3; c++) ds[c][m] = 1. * ds[c][m + is] + 0. * ds[c][m + 2 * is];
      }
  for (int k = k1 - 1; k <= k2; k++)
    for (int side = 0; side < 2; side++) {
      int j = side == 0 ? j1 - 1 : j2, js = side == 0 ? 1 : -1;
      for (int i = i1 - 1; i <= i2; i++) {
        long m = IDX(b, i, j, k) - n0;
        for (int c = 0; c < 3; c++) ds[c][m] = 1. * ds[c][m + js * nid] + 0. * ds[c][m + 2 * js * nid];
      }
    }
}

/* met_bare.f90:4 -- bare = 2 mut mu / rho x the sum over the nine velocity
 * gradients of |grad(dv_ab)|^2 (teq_grads of each, accumulated in the order
 * dvxx, dvxy, ..., dvzz) on the real cells. */
void orc_met_bare(const orc_block *b, const double *sn, const double *vol, const double *v, long ldv,
                  const double *mu, const double *mut, const double *g9, double *bare) {
  long nnn = b->nnn, n0 = b->npc;
  double *g = calloc(3 * nnn, sizeof(double));
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) bare[IDX(b, i, j, k) - n0] = 0.;
  for (int c = 0; c < 9; c++) {
    orc_teq_grads(b, sn, vol, g9 + c * nnn, g);
    for (int k = b->kk1; k <= b->kk2 - 1; k++)
      for (int j = b->jj1; j <= b->jj2 - 1; j++)
        for (int i = b->ii1; i <= b->ii2 - 1; i++) {
          long m = IDX(b, i, j, k) - n0;
          bare[m] = bare[m] + g[m] * g[m] + g[nnn + m] * g[nnn + m] + g[2 * nnn + m] * g[2 * nnn + m];
        }
  }
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - n0;
        bare[m] = 2. * mut[n] * mu[n] * bare[m] / v[n];
      }
  free(g);
}

/* the 2x2 source Jacobian's larger eigenvalue modulus shared by the
 * variants (met_smkes.f90:92-101 and siblings) */
static double ke_cfke(double c1f1, double c2f2, double tp, double s6, double s7, double bark) {
  double d = -bark;
  double x = c1f1 * tp - c2f2 * s7;
  double g = fmax2(0., d * (d - 2. * (x - c2f2 * s7)));
  double y = x - c2f2 * s7;
  double rdelta = sqrt(y * y + 4 * x * s7 + g);
  double bb = x - c2f2 * s7 - d;
  double xl1 = 0.5 * (bb + rdelta) / s6, xl2 = 0.5 * (bb - rdelta) / s6;
  return fmax2(fabs(xl1), fabs(xl2));
}

/* met_smke.f90:4 -- met_exgr on the three rows of g9, met_bark, met_bare,
 * then the variant: 0 met_smkes (Jones-Launder), 1 met_smker, 2 met_smkec
 * (compressible dissipation), 3 met_smkesas (scale-adaptive, Laplacians of
 * u and v by met_laplaciens with scheme ischema).  lparoi = 1 zeroes bark and
 * bare (and f2 = 1) where the variant does. */
void orc_met_smke(const orc_block *b, const double *kc, int variant, int lparoi, int ischema, const double *sn,
                  const double *vol, const double *v, long ldv, const double *mu, const double *mut,
                  const double *tprod, double *g9, const double *cson, double *bark, double *bare, double *cfke,
                  double *ts6, double *ts7) {
  long nnn = b->nnn, n0 = b->npc;
  const double cke1 = kc[0], cke2 = kc[1];
  orc_met_exgr(b, g9, g9 + nnn, g9 + 2 * nnn);
  orc_met_exgr(b, g9 + 3 * nnn, g9 + 4 * nnn, g9 + 5 * nnn);
  orc_met_exgr(b, g9 + 6 * nnn, g9 + 7 * nnn, g9 + 8 * nnn);
  orc_met_bark(b, sn, vol, v, ldv, mu, bark);
  orc_met_bare(b, sn, vol, v, ldv, mu, mut, g9, bare);
  double *lap = NULL;
  if (variant == 3) {
    lap = calloc(6 * nnn, sizeof(double));
    orc_met_laplaciens(b, ischema, sn, vol, g9, g9 + nnn, g9 + 2 * nnn, lap, lap + 2 * nnn, lap + 3 * nnn);
    orc_met_laplaciens(b, ischema, sn, vol, g9 + 3 * nnn, g9 + 4 * nnn, g9 + 5 * nnn, lap + nnn, lap + 4 * nnn,
                       lap + 5 * nnn);
  }
  const double *s1 = v, *s6 = v + 5 * ldv, *s7 = v + 6 * ldv;
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - n0;
        double esk = s7[n] / s6[n];
        double roe2sk = (s7[n] * s7[n]) / s6[n];
        double f1 = 1.;
        if (variant == 1) { /* met_smker.f90:66-84 */
          ts6[n] = tprod[m] - s7[n];
          ts7[n] = cke1 * tprod[m] * esk - cke2 * roe2sk;
          cfke[n] = ke_cfke(cke1, cke2, tprod[m], s6[n], s7[n], 0.);
          continue;
        }
        double retur = (s6[n] * s6[n]) / (s7[n] * mu[n]);
        double arg = -(retur * retur);
        double f2 = (1. - 0.3 * exp(arg));
        double xmt = 0., epsc = 0., pid = 0.;
        if (variant == 2) { /* met_smkec.f90:74-79 */
          xmt = sqrt(2. * s6[n] / s1[n]) / cson[n];
          epsc = 0.1 * s7[n] * (xmt * xmt);
        }
        if (lparoi == 1) {
          bark[m] = 0.;
          bare[m] = 0.;
          f2 = 1.;
        }
        if (variant == 2)
          ts6[n] = tprod[m] + pid - s7[n] - epsc + bark[m];
        else
          ts6[n] = tprod[m] - s7[n] + bark[m];
        if (variant == 3) { /* met_smkesas.f90:140-165 */
          const double *g = g9;
          double xx = g[m], xy = g[nnn + m], xz = g[2 * nnn + m], yx = g[3 * nnn + m], yy = g[4 * nnn + m];
          double yz = g[5 * nnn + m], zx = g[6 * nnn + m], zy = g[7 * nnn + m], zz = g[8 * nnn + m];
          const double qsi2 = 1.47, xkappa = 0.41, cc43 = 4. / 3.;
          double a = zy + yz, c = xz + zx, e = yx + xy;
          double ss = sqrt(cc43 * (xx * xx + yy * yy + zz * zz) + a * a + c * c + e * e);
          double l1 = lap[m], l2 = lap[nnn + m];
          double nlaplu = sqrt(l1 * l1 + l2 * l2);
          double xlvk = xkappa * ss / nlaplu;
          double xlvk2 = xlvk * xlvk;
          double xl = pow(s6[n] / s1[n], 0.5) / (esk);
          double xl2 = xl * xl;
          double psas = xl2 / xlvk2;
          ts7[n] = cke1 * f1 * tprod[m] * esk + qsi2 * psas * tprod[m] * esk - cke2 * f2 * roe2sk + bare[m];
        } else {
          ts7[n] = cke1 * f1 * tprod[m] * esk - cke2 * f2 * roe2sk + bare[m];
        }
        cfke[n] = ke_cfke(cke1 * f1, cke2 * f2, tprod[m], s6[n], s7[n], bark[m]);
      }
  free(lap);
}

/* met_kemut.f90:4 -- form 0: Jones-Launder ('2JL', kesst = 0); 1: with the
 * SST limiter (kesst = 1; orc_teq_gradv into g9 first, as the reference
 * calls teq_gradv there); 2: Launder-Sharma ('2LS').  fmu = 1 when lparoi is
 * 1 or 2. */
void orc_met_kemut(const orc_block *b, double cmu, int form, int lparoi, const double *sn, const double *vol,
                   double *v, long ldv, const double *mu, double *mut, const double *dist, double *g9) {
  long nnn = b->nnn, n0 = b->npc;
  if (form == 1) orc_teq_gradv(b, sn, vol, v, ldv, g9);
  const double a1 = sqrt(cmu);
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - n0;
        double v1 = v[n], v6 = v[5 * ldv + n], v7 = v[6 * ldv + n];
        double retur = (v6 * v6) / (v7 * mu[n]);
        double mut0 = cmu * retur * mu[n];
        double fmu;
        if (lparoi == 1 || lparoi == 2)
          fmu = 1.;
        else if (form == 2) {
          double q = 1. + retur / 50.;
          fmu = exp(-3.4 / (q * q));
        } else
          fmu = exp(-2.5 / (1. + retur / 50.));
        if (form != 1) {
          mut[n] = fmu * mut0;
          continue;
        }
        double coef1 = 2. * pow(v6, 1.5) / (v7 * sqrt(v1) * dist[n]);
        double coef2 = 500. * mu[n] * cmu * v6 / (v7 * (dist[n] * dist[n]) * v1);
        const double *g = g9;
        double x = g[7 * nnn + m] - g[5 * nnn + m], y = g[2 * nnn + m] - g[6 * nnn + m];
        double z = g[3 * nnn + m] - g[nnn + m];
        double rota = sqrt(x * x + y * y + z * z);
        double zeta = fmax2(coef1, coef2);
        double z2 = 2. * (zeta * zeta);
        double exp2x = exp(z2 < 25. ? z2 : 25.);
        double f2 = (exp2x - 1.) / (exp2x + 1.);
        mut[n] = fmu * mut0 / fmax2(1., rota * f2 * v6 * a1 * fmu / v7);
      }
}
