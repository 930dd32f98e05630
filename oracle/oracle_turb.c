/* Oracle: turbulence-transport kernels of met_num (k/omega, nu~ equations).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  t is the state v(ip11,7): the
 * transported variables are t[5*ldv+n] and t[6*ldv+n]; d likewise. */
#include <math.h>
#include <stdlib.h>
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
#define SN(sn, b, m, d, c) (sn)[((long)(c) * 3 + (d)) * (b)->nnn + (m)]

/* met_roe2o.f90:281-289: minmod of two and of three arguments */
static double amimd(double a, double b) {
  double s = copysign(1.0, a);
  double m = fabs(a) < b * s ? fabs(a) : b * s;
  return s * (0. > m ? 0. : m);
}
static double fi1(double x, double y, double z) { return amimd(x, amimd(y, z)); }

/* met_roe2o.f90:88-267 (klroe=2): second-order upwind dissipation of the two
 * transported variables, dd = 0.5 |V.S| (delta - minmod(delta-, delta,
 * delta+)), face differences mirrored one face beyond the block; the k
 * faces, then j, then i.  del6, del7: scratch of nnn entries. */
void orc_met_roe2o(const orc_block *b, const double *t, long ldv, double *d, const double *sn) {
  long n0 = b->npc, nnn = b->nnn;
  int i1 = b->ii1, i2 = b->ii2, j1 = b->jj1, j2 = b->jj2, k1 = b->kk1, k2 = b->kk2;
  const double *t6 = t + 5 * ldv, *t7 = t + 6 * ldv;
  double *d6 = d + 5 * ldv, *d7 = d + 6 * ldv;
  double *del6 = calloc(nnn, 8), *del7 = calloc(nnn, 8);
  for (int k = k1 - 1; k <= k2; k++)
    for (int j = j1 - 1; j <= j2; j++)
      for (int i = i1 - 1; i <= i2; i++) {
        long n = IDX(b, i, j, k);
        d6[n] = 0.; d7[n] = 0.;
      }
  for (int dd = 2; dd >= 0; dd--) {
    long inc = INC(b, dd);
    int a1 = dd == 0 ? i1 : (dd == 1 ? j1 : k1), a2 = dd == 0 ? i2 : (dd == 1 ? j2 : k2);
    int lo[3] = {i1, j1, k1}, hi[3] = {i2 - 1, j2 - 1, k2 - 1};
    lo[dd] = a1; hi[dd] = a2;
    for (int k = lo[2]; k <= hi[2]; k++)
      for (int j = lo[1]; j <= hi[1]; j++)
        for (int i = lo[0]; i <= hi[0]; i++) {
          long n = IDX(b, i, j, k), m = n - n0;
          del6[m] = t6[n] - t6[n - inc];
          del7[m] = t7[n] - t7[n - inc];
        }
    /* mirror: face a1-1 copies face a1, face a2+1 copies face a2 */
    for (int side = 0; side < 2; side++) {
      int l2[3] = {lo[0], lo[1], lo[2]}, h2[3] = {hi[0], hi[1], hi[2]};
      int a = side == 0 ? a1 - 1 : a2 + 1;
      l2[dd] = h2[dd] = a;
      long is = side == 0 ? inc : -inc;
      for (int k = l2[2]; k <= h2[2]; k++)
        for (int j = l2[1]; j <= h2[1]; j++)
          for (int i = l2[0]; i <= h2[0]; i++) {
            long m = IDX(b, i, j, k) - n0;
            del6[m] = del6[m + is];
            del7[m] = del7[m + is];
          }
    }
    for (int k = lo[2]; k <= hi[2]; k++)
      for (int j = lo[1]; j <= hi[1]; j++)
        for (int i = lo[0]; i <= hi[0]; i++) {
          long n = IDX(b, i, j, k), m = n - n0, n1 = n - inc;
          double rro = 2 * t[n], rro1 = 2 * t[n1];
          double u = t[ldv + n] / rro + t[ldv + n1] / rro1;
          double v = t[2 * ldv + n] / rro + t[2 * ldv + n1] / rro1;
          double w = t[3 * ldv + n] / rro + t[3 * ldv + n1] / rro1;
          double rlam = u * SN(sn, b, m, dd, 0) + v * SN(sn, b, m, dd, 1) + w * SN(sn, b, m, dd, 2);
          double x = 0.5 * fabs(rlam) * (del6[m] - fi1(del6[m - inc], del6[m], del6[m + inc]));
          d6[n] = d6[n] - x;
          d6[n - inc] = d6[n - inc] + x;
          x = 0.5 * fabs(rlam) * (del7[m] - fi1(del7[m - inc], del7[m], del7[m + inc]));
          d7[n] = d7[n] - x;
          d7[n - inc] = d7[n - inc] + x;
        }
  }
  free(del6); free(del7);
}

/* met_roe.f90:85-180 (klroe=1): first-order upwind dissipation |V.n| dt6,
 * unit normal, no face area.  Its index function omits the +1 of every other
 * kernel (met_roe.f90:184-188), so every access lands one storage entry
 * before the intended cell (and face).  Reproduced here, except the zeroing
 * of the entry just before the block (outside this block's storage). */
void orc_met_roe(const orc_block *b, const double *t, long ldv, double *d, const double *sn) {
  long n0 = b->npc;
  int i1 = b->ii1, i2 = b->ii2, j1 = b->jj1, j2 = b->jj2, k1 = b->kk1, k2 = b->kk2;
  const double *t6 = t + 5 * ldv, *t7 = t + 6 * ldv;
  double *d6 = d + 5 * ldv, *d7 = d + 6 * ldv;
  for (long n = IDX(b, i1 - 1, j1 - 1, k1 - 1) - 1; n <= IDX(b, i2, j2, k2) - 1; n++) {
    if (n < n0) continue;
    d6[n] = 0.; d7[n] = 0.;
  }
  for (int dd = 0; dd < 3; dd++) {
    long inc = INC(b, dd);
    int lo[3] = {i1, j1, k1}, hi[3] = {i2 - 1, j2 - 1, k2 - 1};
    hi[dd] += 1;
    for (int k = lo[2]; k <= hi[2]; k++)
      for (int j = lo[1]; j <= hi[1]; j++)
        for (int i = lo[0]; i <= hi[0]; i++) {
          long n = IDX(b, i, j, k) - 1, m = n - n0, n1 = n - inc;
          double s1 = SN(sn, b, m, dd, 0), s2 = SN(sn, b, m, dd, 1), s3 = SN(sn, b, m, dd, 2);
          double cnds = sqrt(s1 * s1 + s2 * s2 + s3 * s3);
          double um = 0.5 * (t[ldv + n] / t[n] + t[ldv + n1] / t[n1]);
          double vm = 0.5 * (t[2 * ldv + n] / t[n] + t[2 * ldv + n1] / t[n1]);
          double wm = 0.5 * (t[3 * ldv + n] / t[n] + t[3 * ldv + n1] / t[n1]);
          double vn = fabs(um * s1 + vm * s2 + wm * s3) / cnds;
          double x6 = vn * (t6[n] - t6[n1]), x7 = vn * (t7[n] - t7[n1]);
          d6[n] = d6[n] - x6;
          d7[n] = d7[n] - x7;
          d6[n - inc] = d6[n - inc] + x6;
          d7[n - inc] = d7[n - inc] + x7;
        }
  }
}

/* ------------------------------------------------------------------------ */
/* teq_grads.f90:74-420 (weighted-scheme branch with unit weights, the one
 * every non-central scheme takes): Green-Gauss gradient of a block-local
 * scalar s(m); g = dsdx, dsdy, dsdz (3 x nnn, block-local). */
void orc_teq_grads(const orc_block *b, const double *sn, const double *vol, const double *s, double *g) {
  long n0 = b->npc, nnn = b->nnn;
  int i1 = b->ii1, i2 = b->ii2, j1 = b->jj1, j2 = b->jj2, k1 = b->kk1, k2 = b->kk2;
  for (long m = 0; m < 3 * nnn; m++) g[m] = 0.;
  double *G[3] = {g, g + nnn, g + 2 * nnn};
  const double eps = 0.000001;
  for (int d = 2; d >= 0; d--) {
    long inc = INC(b, d);
    int lo[3] = {i1, j1, k1}, hi[3] = {i2 - 1, j2 - 1, k2 - 1};
    int a1 = d == 0 ? i1 : (d == 1 ? j1 : k1), a2 = d == 0 ? i2 : (d == 1 ? j2 : k2);
    lo[d] = a1 + 1; hi[d] = a2 - 1;
    int jk = (d == 1);  /* the j sweep runs j outer, k inner */
    int olo = jk ? lo[1] : lo[2], ohi = jk ? hi[1] : hi[2], qlo = jk ? lo[2] : lo[1], qhi = jk ? hi[2] : hi[1];
    for (int o = olo; o <= ohi; o++)
      for (int q = qlo; q <= qhi; q++) {
        int j = jk ? o : q, k = jk ? q : o;
        for (int i = lo[0]; i <= hi[0]; i++) {
          long m = IDX(b, i, j, k) - n0;
          for (int c = 0; c < 3; c++) G[c][m] = G[c][m] - (s[m] + s[m - inc]) * SN(sn, b, m, d, c);
        }
        for (int i = lo[0]; i <= hi[0]; i++) {
          long m = IDX(b, i, j, k) - n0;
          for (int c = 0; c < 3; c++) G[c][m - inc] = G[c][m - inc] + (s[m] + s[m - inc]) * SN(sn, b, m, d, c);
        }
      }
    for (int side = 0; side < 2; side++) {
      int l2[3] = {i1, j1, k1}, h2[3] = {i2 - 1, j2 - 1, k2 - 1};
      l2[d] = h2[d] = side == 0 ? a1 : a2;
      for (int k = l2[2]; k <= h2[2]; k++)
        for (int j = l2[1]; j <= h2[1]; j++)
          for (int i = l2[0]; i <= h2[0]; i++) {
            long m = IDX(b, i, j, k) - n0;
            for (int c = 0; c < 3; c++) {
              if (side == 0) G[c][m] = G[c][m] - SN(sn, b, m, d, c) * 2 * s[m - inc];
              else G[c][m - inc] = G[c][m - inc] + SN(sn, b, m, d, c) * 2 * s[m];
            }
          }
    }
  }
  for (long n = IDX(b, i1, j1, k1); n <= IDX(b, i2 - 1, j2 - 1, k2 - 1); n++) {
    long m = n - n0;
    double ts = copysign(0.5, -vol[n]);
    double vols = (0.5 + ts) * eps + (0.5 - ts) * vol[n];
    double c0 = 0.5 / (vols + 2.2250738585072014e-308);
    for (int c = 0; c < 3; c++) G[c][m] = G[c][m] * c0;
  }
}

/* scalar of block-local storage over the box i1-1..i2, j1-1..j2, k1-1..k2 */
static void fill_box(const orc_block *b, double *s, const double *num, const double *den) {
  for (int k = b->kk1 - 1; k <= b->kk2; k++)
    for (int j = b->jj1 - 1; j <= b->jj2; j++)
      for (int i = b->ii1 - 1; i <= b->ii2; i++) {
        long n = IDX(b, i, j, k);
        s[n - b->npc] = den ? num[n] / den[n] : num[n];
      }
}

/* teq_gradv.f90:79-162: clamps rho >= 1e-20 on the box (a side effect on v
 * that the reference has), then grad u, grad v, grad w into g9 (9 x nnn,
 * block-local, order dvxx dvxy dvxz dvyx ... dvzz). */
void orc_teq_gradv(const orc_block *b, const double *sn, const double *vol, double *v, long ldv, double *g9) {
  long nnn = b->nnn;
  for (int k = b->kk1 - 1; k <= b->kk2; k++)
    for (int j = b->jj1 - 1; j <= b->jj2; j++)
      for (int i = b->ii1 - 1; i <= b->ii2; i++) {
        long n = IDX(b, i, j, k);
        v[n] = fmax(v[n], 1.e-20);
      }
  double *s = calloc(nnn, 8);
  for (int a = 0; a < 3; a++) {
    fill_box(b, s, v + (a + 1) * ldv, v);
    orc_teq_grads(b, sn, vol, s, g9 + 3 * a * nnn);
  }
  free(s);
}

/* met_prod.f90:54-72: production tprod(m) = tau^R : grad u with the
 * Boussinesq stress of mut and the -2/3 rho k term (v6 = rho k). */
void orc_met_prod(const orc_block *b, const double *mut, const double *v, long ldv, const double *g9, double *tprod) {
  long n0 = b->npc, nnn = b->nnn;
  const double ds3 = 2. / 3.;
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - n0;
        double xx = g9[m], xy = g9[nnn + m], xz = g9[2 * nnn + m], yx = g9[3 * nnn + m], yy = g9[4 * nnn + m];
        double yz = g9[5 * nnn + m], zx = g9[6 * nnn + m], zy = g9[7 * nnn + m], zz = g9[8 * nnn + m];
        double s6 = v[5 * ldv + n];
        double torxx = ds3 * (-s6 - mut[n] * (-2 * xx + yy + zz));
        double toryy = ds3 * (-s6 - mut[n] * (-2 * yy + xx + zz));
        double torzz = ds3 * (-s6 - mut[n] * (-2 * zz + yy + xx));
        double torxy = mut[n] * (xy + yx);
        double torxz = mut[n] * (xz + zx);
        double toryz = mut[n] * (zy + yz);
        tprod[m] = xx * torxx + yy * toryy + zz * torzz + torxy * (xy + yx) + torxz * (xz + zx) + toryz * (zy + yz);
      }
}

/* met_gradtr.f90:107-171: fd (6 arrays of ldv) = grad(v6/rho), grad(v7/rho)
 * over the contiguous storage range first..last real cell. */
void orc_met_gradtr(const orc_block *b, const double *sn, const double *vol, const double *v, long ldv, double *fd) {
  long n0 = b->npc, nnn = b->nnn;
  double *s = calloc(nnn, 8), *g = calloc(3 * nnn, 8);
  for (int a = 0; a < 2; a++) {
    fill_box(b, s, v + (5 + a) * ldv, v);
    orc_teq_grads(b, sn, vol, s, g);
    for (long n = IDX(b, b->ii1, b->jj1, b->kk1); n <= IDX(b, b->ii2 - 1, b->jj2 - 1, b->kk2 - 1); n++)
      for (int c = 0; c < 3; c++) fd[(3 * a + c) * ldv + n] = g[c * nnn + n - n0];
  }
  free(s); free(g);
}

/* x**6 as the reference build evaluates an integer power: ((((x x) x) x) x) x */
static inline double pow6(double x) { return ((((x * x) * x) * x) * x) * x; }

/* met_difsa.f90 + met_smsa.f90:91-222: Spalart-Allmaras source ts6 = qcxts5
 * (production, destruction and the cb2/sigma grad(rho nu~).grad(nu~) term),
 * ts7 = 0 and the implicit coefficient cfke.  fd5: grad(nu~) (3 arrays of
 * ldv from met_gradtr); g9: velocity gradients (block-local). */
/* sas: NULL, or the two Laplacian components of met_smsasas (block-local) */
static void sa_sources(const orc_block *b, const orc_turb *tc, const double *sn, const double *vol, const double *v,
                       long ldv, const double *mu, const double *dist, const double *g9, const double *fd5,
                       const double *delta, const double *const *sas, double *cfke, double *ts6, double *ts7) {
  long n0 = b->npc, nnn = b->nnn;
  double *s = calloc(nnn, 8), *gd = calloc(3 * nnn, 8), *gr = calloc(3 * nnn, 8), *sdif = calloc(nnn, 8);
  /* met_difsa: gradient of rho nu~, dotted with grad(nu~), times cb2/sigma */
  fill_box(b, s, v + 5 * ldv, NULL);
  orc_teq_grads(b, sn, vol, s, gd);
  for (long n = IDX(b, b->ii1, b->jj1, b->kk1); n <= IDX(b, b->ii2 - 1, b->jj2 - 1, b->kk2 - 1); n++) {
    long m = n - n0;
    sdif[m] = gd[m] * fd5[n] + gd[nnn + m] * fd5[ldv + n] + gd[2 * nnn + m] * fd5[2 * ldv + n];
    sdif[m] = sdif[m] * tc->cb2 / tc->sigma;
  }
  /* gradient of rho (met_smsa.f90:140-159) */
  fill_box(b, s, v, NULL);
  orc_teq_grads(b, sn, vol, s, gr);
  double kappa2 = tc->kappa * tc->kappa, uns6 = 1. / 6.;
  double cv13 = tc->cv1 * tc->cv1 * tc->cv1, cw36 = pow6(tc->cw3), cv133 = 3. * cv13, ct42 = 2. * tc->ct4;
  double cb2sig = tc->cb2 / tc->sigma;
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - n0;
        double xy = g9[nnn + m], xz = g9[2 * nnn + m], yx = g9[3 * nnn + m];
        double yz = g9[5 * nnn + m], zx = g9[6 * nnn + m], zy = g9[7 * nnn + m];
        double a = zy - yz, bb = xz - zx, c = yx - xy;
        double vort = sqrt(a * a + bb * bb + c * c);
        double v6 = v[5 * ldv + n], v1 = v[n];
        double xkhi = v6 / mu[n];
        double xkhi2 = xkhi * xkhi, xkhi3 = xkhi2 * xkhi;
        double dd = dist[n];
        if (delta) {  /* met_smdes.f90:125: min(d, ctdes delta), ctdes = 0.65 */
          double cd = 0.65 * delta[m];
          dd = dd < cd ? dd : cd;
        }
        double dist2 = dd * dd;
        double nutilde = v6 / v1;
        double kapd2 = kappa2 * dist2;
        double fv1 = xkhi3 / (xkhi3 + cv13);
        double fv2 = 1. - (xkhi / (1 + xkhi * fv1));
        if (delta || sas) fv2 = fabs(fv2); /* met_smdes.f90:128, met_smsasas.f90:214 */
        double stilde = vort + nutilde * fv2 / kapd2;
        double rtilde = nutilde / (stilde * kapd2);
        if (!sas) rtilde = fmin(rtilde, 1.); /* met_smsasas has no clamp */
        double rtil6 = pow6(rtilde);
        double ft2 = tc->ct3 * exp(-tc->ct4 * xkhi2);
        double sm1 = tc->cb1 * (1. - ft2) * stilde * v6;
        double xg = rtilde + tc->cw2 * (rtil6 - rtilde);
        double xg6 = pow6(xg);
        double fw = xg * pow((1. + cw36) / (xg6 + cw36), uns6);
        double sm2 = -(tc->cw1 * fw - tc->cb1 * ft2 / kappa2) * v6 * nutilde / dist2;
        if (sas) { /* met_smsasas.f90:218-237: the SAS destruction where fd = 1 - tanh((8 rtilde)^3) > 0.8 */
          double r8 = 8. * rtilde;
          double fdd = 1. - tanh((r8 * r8) * r8);
          if (fdd > 0.8) {
            double xx = g9[m], yy = g9[4 * nnn + m], zz = g9[8 * nnn + m];
            double s1 = zy + yz, s2 = xz + zx, s3 = yx + xy;
            double ss = sqrt(4. / 3. * (xx * xx + yy * yy + zz * zz) + s1 * s1 + s2 * s2 + s3 * s3);
            double t1 = sas[0][m], t2 = sas[1][m];
            double nlaplu = sqrt(t1 * t1 + t2 * t2);
            double xlvk = tc->kappa * ss / nlaplu;
            double xlvk2 = xlvk * xlvk;
            sm2 = -tc->cw1 * fw * v6 * 1. * nutilde / xlvk2 + tc->cb1 * ft2 * v6 * nutilde / (kappa2 * dist2);
          }
        }
        ts6[n] = sm1 + sm2 + sdif[m];
        ts7[n] = 0.;
        double xkhi4 = xkhi3 * xkhi;
        double dfv1 = cv133 * (fv1 * fv1) / (mu[n] * xkhi4);
        double dfv2 = (xkhi2 * dfv1 - 1. / mu[n]) / ((1. + xkhi * fv1) * (1. + xkhi * fv1));
        double dst = (fv2 / v1 + nutilde * dfv2) / kapd2;
        double dpr = (1. / v1 - nutilde * dst / stilde) / (stilde * kapd2);
        double dxg = dpr * (1. + tc->cw2 * (6. * rtil6 / rtilde - 1.));
        double dfw = (1. - xg6 / (xg6 + cw36)) * fw * dxg / xg;
        double dft2 = -ct42 * xkhi * ft2;
        double dsm1 = tc->cb1 * ((1. - ft2) * (stilde + dst * v6) - stilde * v6 * dft2);
        double dsm2 = 2. * sm2 / v6 - (tc->cw1 * dfw - tc->cb1 * dft2 / kappa2) * nutilde * v6 / dist2;
        /* the reference indexes the grad(nu~) array by the block-local m here
         * (met_smsa.f90:216-217), which is the entry n only for the first block */
        long mf = m;
        double dsdif = -cb2sig * (fd5[mf] * gr[m] + fd5[ldv + mf] * gr[nnn + m] + fd5[2 * ldv + mf] * gr[2 * nnn + m]) /
                       (v1 * v1);
        cfke[n] = fabs(dsm1 + dsm2 + dsdif);
      }
  free(s); free(gd); free(gr); free(sdif);
}

void orc_met_smsa(const orc_block *b, const orc_turb *tc, const double *sn, const double *vol, const double *v,
                  long ldv, const double *mu, const double *dist, const double *g9, const double *fd5,
                  double *cfke, double *ts6, double *ts7) {
  sa_sources(b, tc, sn, vol, v, ldv, mu, dist, g9, fd5, NULL, NULL, cfke, ts6, ts7);
}

/* met_dist.f90:4 (3-D branch, :66-103): delta(m) = the longest of the twelve
 * edges of real cell m, from the node coordinates x, y, z (stored like the
 * cells). */
void orc_met_dist(const orc_block *b, const double *x, const double *y, const double *z, double *delta) {
  long ci = 1, cj = NID(b), ck = (long)NID(b) * NJD(b);
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - b->npc;
#define EDGE(p, q) sqrt((x[p] - x[q]) * (x[p] - x[q]) + (y[p] - y[q]) * (y[p] - y[q]) + (z[p] - z[q]) * (z[p] - z[q]))
        double e[12] = {EDGE(n, n + ci), EDGE(n + cj, n + ci + cj), EDGE(n + ck, n + ci + ck),
                        EDGE(n + cj + ck, n + ci + cj + ck), EDGE(n, n + cj), EDGE(n + ci, n + ci + cj),
                        EDGE(n + ck, n + cj + ck), EDGE(n + ci + ck, n + ci + cj + ck), EDGE(n, n + ck),
                        EDGE(n + ci, n + ci + ck), EDGE(n + cj, n + cj + ck), EDGE(n + ci + cj, n + ci + cj + ck)};
#undef EDGE
        double mx = e[0];
        for (int q = 1; q < 12; q++) mx = e[q] > mx ? e[q] : mx;
        delta[m] = mx;
      }
}

/* met_smdes.f90:4 -- the Spalart-Allmaras sources of met_smsa with the DES
 * length min(d, 0.65 delta) for the wall distance (delta from met_dist,
 * written to the caller's delta as the reference does) and |fv2|. */
void orc_met_smdes(const orc_block *b, const orc_turb *tc, const double *x, const double *y, const double *z,
                   double *delta, const double *sn, const double *vol, const double *v, long ldv, const double *mu,
                   const double *dist, const double *g9, const double *fd5, double *cfke, double *ts6, double *ts7) {
  orc_met_dist(b, x, y, z, delta);
  sa_sources(b, tc, sn, vol, v, ldv, mu, dist, g9, fd5, delta, NULL, cfke, ts6, ts7);
}

/* met_laplaciens.f90:4 -- face sums of (dsx, dsy, dsz) into (dsd2x, dsd2y,
 * dsd2z): each zeroed on indc(i1-1,j1-1,k1-1)..indc(i2+1,j2+1,k2+1), then
 * for the k, j, i faces in that order the interior faces (both neighbours'
 * values) and the boundary faces (twice the ghost's value), scattered as the
 * reference loops do.  ischema = 1 is the plain form, whose j faces read
 * dsy(m-ncj) for dsd2x(m) (:112); other schemes the weighted form with
 * weights 1 (block-local arrays, sn the block's slice). */
void orc_met_laplaciens(const orc_block *b, int ischema, const double *sn, const double *vol, const double *dsx,
                        const double *dsy, const double *dsz, double *dx, double *dy, double *dz) {
  long n0 = b->npc, nid = NID(b), nijd = (long)NID(b) * NJD(b);
  const double *ds[3] = {dsx, dsy, dsz};
  double *dd[3] = {dx, dy, dz};
  long lo = IDX(b, b->ii1 - 1, b->jj1 - 1, b->kk1 - 1) - n0, hi = IDX(b, b->ii2 + 1, b->jj2 + 1, b->kk2 + 1) - n0;
  for (long m = lo; m <= hi; m++) dx[m] = dy[m] = dz[m] = 0.;
  int i1 = b->ii1, i2 = b->ii2, j1 = b->jj1, j2 = b->jj2, k1 = b->kk1, k2 = b->kk2;
  for (int dir = 2; dir >= 0; dir--) {
    long inc = dir == 0 ? 1 : (dir == 1 ? nid : nijd);
    int a1 = dir == 0 ? i1 : (dir == 1 ? j1 : k1), a2 = dir == 0 ? i2 : (dir == 1 ? j2 : k2);
    /* interior faces a = a1+1 .. a2-1 */
    for (int k = k1; k <= k2 - 1; k++)
      for (int j = j1; j <= j2 - 1; j++)
        for (int i = i1; i <= i2 - 1; i++) {
          int a = dir == 0 ? i : (dir == 1 ? j : k);
          if (a < a1 + 1) continue;
          long m = IDX(b, i, j, k) - n0;
          for (int c = 0; c < 3; c++) {
            double s = SN(sn, b, m, dir, c);
            double own = (ischema == 1 && dir == 1 && c == 0) ? ds[0][m] + ds[1][m - inc] : ds[c][m] + ds[c][m - inc];
            double t = ischema == 1 ? s * own : (ds[c][m] + ds[c][m - inc]) * s;
            double t2 = ischema == 1 ? s * (ds[c][m] + ds[c][m - inc]) : t;
            dd[c][m] = dd[c][m] - t;
            dd[c][m - inc] = dd[c][m - inc] + t2;
          }
        }
    /* boundary faces a1 (ghost below) then a2 (ghost above) */
    for (int side = 0; side < 2; side++)
      for (int k = k1; k <= k2 - 1; k++)
        for (int j = j1; j <= j2 - 1; j++)
          for (int i = i1; i <= i2 - 1; i++) {
            int a = dir == 0 ? i : (dir == 1 ? j : k);
            if (a != a1) continue;
            long m0 = IDX(b, i, j, k) - n0, m = side == 0 ? m0 : m0 + (long)(a2 - a1) * inc;
            for (int c = 0; c < 3; c++) {
              double s = SN(sn, b, m, dir, c);
              if (side == 0) dd[c][m] = dd[c][m] - s * 2 * ds[c][m - inc];
              else dd[c][m - inc] = dd[c][m - inc] + s * 2 * ds[c][m];
            }
          }
  }
  /* 0.5/vol (vol <= 0 guarded by 1e-6) and dsd2x = the sum of the three
   * (:314-325), over indc(i1,j1,k1)..indc(i2-1,j2-1,k2-1) */
  for (long n = IDX(b, i1, j1, k1); n <= IDX(b, i2 - 1, j2 - 1, k2 - 1); n++) {
    long m = n - n0;
    double ts = -vol[n] >= 0. && !signbit(-vol[n]) ? 0.5 : -0.5;
    double vols = (0.5 + ts) * 0.000001 + (0.5 - ts) * vol[n];
    double c0 = 0.5 / vols;
    dx[m] = dx[m] * c0;
    dy[m] = dy[m] * c0;
    dz[m] = dz[m] * c0;
    dx[m] = dx[m] + dy[m] + dz[m];
  }
}

/* met_smsasas.f90:4 -- Spalart-Allmaras with the SAS destruction: the
 * sources of met_smsa without the rtilde clamp and with |fv2|, and where
 * fd = 1 - tanh((8 rtilde)^3) > 0.8 the destruction from the von Karman
 * length kappa S / sqrt(tn1^2 + tn2^2), tn1 and tn2 the Laplacians of u
 * and v from met_laplaciens of (du/dx, du/dy, du/dz) and (dv/dx, dv/dy,
 * dv/dz) (:173-193; the one of w is formed but not used).  g9 holds the velocity gradients (block-local). */
void orc_met_smsasas(const orc_block *b, const orc_turb *tc, int ischema, const double *sn, const double *vol,
                     const double *v, long ldv, const double *mu, const double *dist, const double *g9,
                     const double *fd5, double *cfke, double *ts6, double *ts7) {
  long nnn = b->nnn;
  double *w = calloc(6 * nnn, sizeof(double));
  orc_met_laplaciens(b, ischema, sn, vol, g9, g9 + nnn, g9 + 2 * nnn, w, w + 2 * nnn, w + 3 * nnn);
  orc_met_laplaciens(b, ischema, sn, vol, g9 + 3 * nnn, g9 + 4 * nnn, g9 + 5 * nnn, w + nnn, w + 4 * nnn,
                     w + 5 * nnn);
  const double *sas[2] = {w, w + nnn};
  sa_sources(b, tc, sn, vol, v, ldv, mu, dist, g9, fd5, NULL, sas, cfke, ts6, ts7);
  free(w);
}

/* k-l model constants (module modeleturb, atlecdon.f90:226-231) */
typedef struct {
  double cklb1, ckle2, sigmal, xkappa;
} orc_kl;

/* met_bark.f90:4 -- bark = -2 mu |grad sqrt(k/rho)|^2 on the storage range
 * indc(i1,j1,k1)..indc(i2-1,j2-1,k2-1) (block-local), the gradient by
 * teq_grads of sqrt(v6/v1) filled on the box i1-1..i2 etc. */
static void bark_of(const orc_block *b, const double *sn, const double *vol, const double *v, long ldv,
                    const double *mu, double *bark) {
  long nnn = b->nnn, n0 = b->npc;
  double *t = calloc(nnn, 8), *g = calloc(3 * nnn, 8);
  for (int k = b->kk1 - 1; k <= b->kk2; k++)
    for (int j = b->jj1 - 1; j <= b->jj2; j++)
      for (int i = b->ii1 - 1; i <= b->ii2; i++) {
        long n = IDX(b, i, j, k);
        t[n - n0] = sqrt(v[5 * ldv + n] / v[n]);
      }
  orc_teq_grads(b, sn, vol, t, g);
  for (long n = IDX(b, b->ii1, b->jj1, b->kk1); n <= IDX(b, b->ii2 - 1, b->jj2 - 1, b->kk2 - 1); n++) {
    long m = n - n0;
    double x = g[m] * g[m] + g[nnn + m] * g[nnn + m] + g[2 * nnn + m] * g[2 * nnn + m];
    bark[m] = -2. * mu[n] * x;
  }
  free(t); free(g);
}

/* met_bark.f90:4 as a routine of its own (the k-epsilon sources call it) */
void orc_met_bark(const orc_block *b, const double *sn, const double *vol, const double *v, long ldv,
                  const double *mu, double *bark) {
  bark_of(b, sn, vol, v, ldv, mu, bark);
}

/* met_smkl.f90:4 -- Smith k-l sources: ts6 = P - csk2 sqrt(k) rho k / l +
 * bark, ts7 = csl1 rho sqrt(k) (1 - l^2/(kappa d)^2) + rho l div u - mut
 * |grad l|^2 l/(kappa d)^2/(sigmal l) + (2/sigmal) mut grad l.grad k / k,
 * and cfke the larger modulus of the eigenvalues of the 2x2 source Jacobian
 * (:163-226; bark from met_bark, grad rho by teq_grads).  fd: gradients of
 * rho k and rho l (six planes of ldv); g9, tprod block-local. */
/* Shared body of met_smkl and met_smklsas: lap = the Laplacians of u and v
 * (block-local, two planes) selects the scale-adaptive term sl5 of
 * met_smklsas.f90:238-251, subtracted from ts7 where the shielding function
 * fd = 1 - tanh((8 rtilde)^3) exceeds 0.8. */
static void kl_sources(const orc_block *b, const orc_kl *kc, const double *sn, const double *vol, const double *v,
                       long ldv, const double *mu, const double *mut, const double *dist, const double *fd,
                       const double *g9, const double *tprod, const double *lap, double *cfke, double *ts6,
                       double *ts7) {
  long nnn = b->nnn, n0 = b->npc;
  double *bark = calloc(nnn, 8), *t = calloc(nnn, 8), *gr = calloc(3 * nnn, 8);
  bark_of(b, sn, vol, v, ldv, mu, bark);
  fill_box(b, t, v, NULL);
  orc_teq_grads(b, sn, vol, t, gr);
  double csk2 = 2. * sqrt(2.) / kc->cklb1, csl1 = sqrt(2.) * (2. - kc->ckle2) / kc->cklb1;
  double dssigl = 2. / kc->sigmal, c132 = 1.5 * csk2, csl12 = 2. * csl1;
  const double *f5x = fd, *f5y = fd + ldv, *f5z = fd + 2 * ldv, *f6x = fd + 3 * ldv, *f6y = fd + 4 * ldv,
               *f6z = fd + 5 * ldv;
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - n0;
        double v1 = v[n], v6 = v[5 * ldv + n], v7 = v[6 * ldv + n];
        double xk = v6 / v1, xl = v7 / v1;
        double kd = kc->xkappa * dist[n];
        double xkapad2 = kd * kd;
        double xlskap2 = xl * xl / xkapad2;
        double rack = sqrt(xk), racrk = sqrt(v6), racro = sqrt(v1);
        double divv = g9[m] + g9[4 * nnn + m] + g9[8 * nnn + m];
        double glgl = f6x[n] * f6x[n] + f6y[n] * f6y[n] + f6z[n] * f6z[n];
        double glgk = (f6x[n] * f5x[n] + f6y[n] * f5y[n] + f6z[n] * f5z[n]) / xk;
        double sk2 = csk2 * rack * v6 / xl;
        double sl1 = csl1 * v1 * rack * (1. - xlskap2);
        double sl2 = v7 * divv;
        double sl3 = mut[n] * glgl * xlskap2 / (kc->sigmal * xl);
        double sl4 = dssigl * mut[n] * glgk;
        double sl5 = 0.;
        if (lap) {
          double nutilde = mut[n] / v1;
          double xx = g9[m], xy = g9[nnn + m], xz = g9[2 * nnn + m], yx = g9[3 * nnn + m], yy = g9[4 * nnn + m];
          double yz = g9[5 * nnn + m], zx = g9[6 * nnn + m], zy = g9[7 * nnn + m], zz = g9[8 * nnn + m];
          double dvdv = xx * xx + xy * xy + yx * yx + yy * yy;
          double rtilde = nutilde / (sqrt(dvdv) * xkapad2);
          double r8 = 8. * rtilde;
          double rtil3 = (r8 * r8) * r8;
          double fdd = 1. - tanh(rtil3);
          if (fdd > 0.8) {
            const double qsi2 = 1.47, cc43 = 4. / 3.;
            double a = zy + yz, c = xz + zx, e = yx + xy;
            double ss = sqrt(cc43 * (xx * xx + yy * yy + zz * zz) + a * a + c * c + e * e);
            double l1 = lap[m], l2 = lap[nnn + m];
            double nlaplu = sqrt(l1 * l1 + l2 * l2);
            double xlvk = kc->xkappa * ss / nlaplu;
            sl5 = qsi2 * tprod[m] * ((xl * xl) * xl) / (xk * (xlvk * xlvk));
          }
        }
        ts6[n] = tprod[m] - sk2 + bark[m];
        ts7[n] = lap ? sl1 + sl2 - sl3 + sl4 - sl5 : sl1 + sl2 - sl3 + sl4;
        double gkgro = (f5x[n] * gr[m] + f5y[n] * gr[nnn + m] + f5z[n] * gr[2 * nnn + m]) / v6;
        double glgro = (f6x[n] * gr[m] + f6y[n] * gr[nnn + m] + f6z[n] * gr[2 * nnn + m]) / v1;
        double dqkdk = -c132 * racro * racrk / v7 - bark[m] / v6 + mu[n] * gkgro / v1;
        double dqkdl = sk2 / v7;
        double dqldl = -csl12 * rack * xl / xkapad2 + divv - mut[n] * (glgl - 2 * xl * glgro) / (kc->sigmal * v1 * xkapad2);
        double dqldk = dssigl * mut[n] * (-glgk + glgro) / v6;
        double xdelta = dqkdk * dqkdk + dqldl * dqldl + fabs(-2. * dqkdk * dqldl + 4. * dqkdl * dqldk);
        double rdelta = sqrt(xdelta);
        double xl1 = (dqkdk + dqldl + rdelta) * 0.5, xl2 = (dqkdk + dqldl - rdelta) * 0.5;
        double a1 = fabs(xl1), a2 = fabs(xl2);
        cfke[n] = a1 > a2 ? a1 : a2;
      }
  free(bark); free(t); free(gr);
}

void orc_met_smkl(const orc_block *b, const orc_kl *kc, const double *sn, const double *vol, const double *v, long ldv,
                  const double *mu, const double *mut, const double *dist, const double *fd, const double *g9,
                  const double *tprod, double *cfke, double *ts6, double *ts7) {
  kl_sources(b, kc, sn, vol, v, ldv, mu, mut, dist, fd, g9, tprod, NULL, cfke, ts6, ts7);
}

/* met_smklsas.f90:4 -- met_smkl with the scale-adaptive term; the Laplacians
 * of u and v by met_laplaciens with scheme ischema, of the gradients as given
 * (their ghost entries read as they are). */
void orc_met_smklsas(const orc_block *b, const orc_kl *kc, int ischema, const double *sn, const double *vol,
                     const double *v, long ldv, const double *mu, const double *mut, const double *dist,
                     const double *fd, const double *g9, const double *tprod, double *cfke, double *ts6,
                     double *ts7) {
  long nnn = b->nnn;
  double *w = calloc(6 * nnn, sizeof(double));
  orc_met_laplaciens(b, ischema, sn, vol, g9, g9 + nnn, g9 + 2 * nnn, w, w + 2 * nnn, w + 3 * nnn);
  orc_met_laplaciens(b, ischema, sn, vol, g9 + 3 * nnn, g9 + 4 * nnn, g9 + 5 * nnn, w + nnn, w + 4 * nnn,
                     w + 5 * nnn);
  kl_sources(b, kc, sn, vol, v, ldv, mu, mut, dist, fd, g9, tprod, w, cfke, ts6, ts7);
  free(w);
}

/* met_klmut.f90:4 -- Smith k-l eddy viscosity mut = fmu mu chi, chi =
 * sqrt(2)/cklb1^(1/3) sqrt(k/rho) rho l / mu, fmu = ((c1^4 f1 + s)/(c1^4 +
 * s))^(1/4), s = (c2^2 + chi^2) chi^2, f1 = exp(-50 (l/(kappa d))^2),
 * c1 = 25.5, c2 = 2 (:38-62). */
void orc_met_klmut(const orc_block *b, const orc_kl *kc, const double *v, long ldv, const double *mu, double *mut,
                   const double *dist) {
  double c1 = 25.5, c2 = 2.;
  double r2sb1 = sqrt(2.) / pow(kc->cklb1, 1. / 3.);
  double c14 = ((c1 * c1) * c1) * c1, c22 = c2 * c2;
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k);
        double v1 = v[n], v6 = v[5 * ldv + n], v7 = v[6 * ldv + n];
        double rack = sqrt(v6 / v1);
        double xxi = r2sb1 * rack * v7 / mu[n];
        double xxi2 = xxi * xxi;
        double q = v7 / (kc->xkappa * v1 * dist[n]);
        double xl2 = q * q;
        double f1 = exp(-50. * xl2);
        double ss2 = (c22 + xxi2) * xxi2;
        double fmu = pow((c14 * f1 + ss2) / (c14 + ss2), 0.25);
        mut[n] = fmu * mu[n] * xxi;
      }
}

/* met_smko.f90:80-110: compressible Wilcox k-omega sources. */
void orc_met_smko(const orc_block *b, const orc_turb *tc, const double *v, long ldv, const double *tprod,
                  double *cfke, double *ts6, double *ts7, const double *cson) {
  long n0 = b->npc;
  const double xw = 5. / 9., xmt0 = 0.1;
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - n0;
        double v1 = v[n], v6 = v[5 * ldv + n], v7 = v[6 * ldv + n];
        double omeg = v7 / v1;
        double cd = 0.;
        double xmt = sqrt(2. * v6 / v1) / cson[n];
        double f = 0.5 * (1. + copysign(1.0, xmt - xmt0)) * (xmt * xmt - xmt0 * xmt0);
        double betasc = tc->betas * (1. + 1.5 * f);
        double betac = tc->beta - tc->betas * 1.5 * f;
        ts6[n] = tprod[m] - betasc * v6 * omeg;
        ts7[n] = xw * tprod[m] * v7 / v6 - betac * v7 * omeg + cd;
        cfke[n] = 0.18 * omeg;
      }
}

/* met_fludcsa.f90:96-107 (kfludis=2), met_fludko.f90 (kfludis=5) and
 * met_fludc.f90:100-114 (kfludis=1): the gradients fd (6 arrays of ldv) become
 * diffusive flux densities, over the contiguous range first..last real cell. */
void orc_met_flud(const orc_block *b, int kfludis, const orc_turb *tc, const double *v, long ldv, const double *mu,
                  const double *mut, double *fd) {
  long n0 = IDX(b, b->ii1, b->jj1, b->kk1), n1 = IDX(b, b->ii2 - 1, b->jj2 - 1, b->kk2 - 1);
  double sigma1 = 1. / tc->sigma;
  for (long n = n0; n <= n1; n++) {
    if (kfludis == 2) {
      double smutot = sigma1 * (mu[n] + v[5 * ldv + n]);
      for (int c = 0; c < 3; c++) fd[c * ldv + n] = smutot * fd[c * ldv + n];
      for (int c = 3; c < 6; c++) fd[c * ldv + n] = 0.;
    } else if (kfludis == 5) {
      for (int c = 0; c < 3; c++) fd[c * ldv + n] = (mu[n] + mut[n] * tc->sigmk) * fd[c * ldv + n];
      for (int c = 3; c < 6; c++) fd[c * ldv + n] = (mu[n] + mut[n] * tc->sigmw) * fd[c * ldv + n];
    } else {
      for (int c = 0; c < 3; c++) fd[c * ldv + n] = (mu[n] + mut[n] / tc->alfak) * fd[c * ldv + n];
    }
  }
  if (kfludis == 1)
    for (long n = n0; n <= n1; n++)
      for (int c = 3; c < 6; c++) fd[c * ldv + n] = (mu[n] + mut[n] / tc->alfae) * fd[c * ldv + n];
}

/* sch_turb.f90:92-301 (without met_pardis, applied by the caller): centred
 * convective minus diffusive face fluxes of v6, v7, doubled ghost values on
 * the boundary faces; u6 = 0.5 sum - vol ts6 - d6 (same for 7). */
void orc_sch_turb(const orc_block *b, double *u, const double *v, long ldv, const double *d, const double *fd,
                  const double *ts6, const double *ts7, const double *sn, const double *vol) {
  long n0 = b->npc, nnn = b->nnn;
  int i1 = b->ii1, i2 = b->ii2, j1 = b->jj1, j2 = b->jj2, k1 = b->kk1, k2 = b->kk2;
  double *f = calloc(6 * nnn, 8);
  double *u6 = u + 5 * ldv, *u7 = u + 6 * ldv;
  for (long n = IDX(b, i1 - 1, j1 - 1, k1 - 1); n <= IDX(b, i2, j2, k2); n++) { u6[n] = 0.; u7[n] = 0.; }
  for (long n = IDX(b, i1 - 1, j1 - 1, k1 - 1); n <= IDX(b, i2, j2, k2); n++) {
    long m = n - n0;
    for (int c = 0; c < 3; c++) {
      f[c * nnn + m] = v[5 * ldv + n] * (v[(c + 1) * ldv + n] / v[n]) - fd[c * ldv + n];
      f[(3 + c) * nnn + m] = v[6 * ldv + n] * (v[(c + 1) * ldv + n] / v[n]) - fd[(3 + c) * ldv + n];
    }
  }
  for (int d_ = 0; d_ < 3; d_++) {
    long inc = INC(b, d_);
    int a1 = d_ == 0 ? i1 : (d_ == 1 ? j1 : k1), a2 = d_ == 0 ? i2 : (d_ == 1 ? j2 : k2);
    int lo[3] = {i1, j1, k1}, hi[3] = {i2 - 1, j2 - 1, k2 - 1};
    lo[d_] = a1 + 1; hi[d_] = a2 - 1;
    for (int k = lo[2]; k <= hi[2]; k++)
      for (int j = lo[1]; j <= hi[1]; j++)
        for (int i = lo[0]; i <= hi[0]; i++) {
          long n = IDX(b, i, j, k), m = n - n0;
          double s1 = SN(sn, b, m, d_, 0), s2 = SN(sn, b, m, d_, 1), s3 = SN(sn, b, m, d_, 2);
          double x6 = (f[m] + f[m - inc]) * s1 + (f[nnn + m] + f[nnn + m - inc]) * s2 + (f[2 * nnn + m] + f[2 * nnn + m - inc]) * s3;
          double x7 = (f[3 * nnn + m] + f[3 * nnn + m - inc]) * s1 + (f[4 * nnn + m] + f[4 * nnn + m - inc]) * s2 +
                      (f[5 * nnn + m] + f[5 * nnn + m - inc]) * s3;
          u6[n] = u6[n] - x6; u7[n] = u7[n] - x7;
          u6[n - inc] = u6[n - inc] + x6; u7[n - inc] = u7[n - inc] + x7;
        }
    for (int side = 0; side < 2; side++) {
      int l2[3] = {i1, j1, k1}, h2[3] = {i2 - 1, j2 - 1, k2 - 1};
      l2[d_] = h2[d_] = side == 0 ? a1 : a2;
      for (int k = l2[2]; k <= h2[2]; k++)
        for (int j = l2[1]; j <= h2[1]; j++)
          for (int i = l2[0]; i <= h2[0]; i++) {
            long n = IDX(b, i, j, k), m = n - n0;
            long mg = side == 0 ? m - inc : m;
            double s1 = SN(sn, b, m, d_, 0), s2 = SN(sn, b, m, d_, 1), s3 = SN(sn, b, m, d_, 2);
            double x6 = 2 * f[mg] * s1 + 2 * f[nnn + mg] * s2 + 2 * f[2 * nnn + mg] * s3;
            double x7 = 2 * f[3 * nnn + mg] * s1 + 2 * f[4 * nnn + mg] * s2 + 2 * f[5 * nnn + mg] * s3;
            if (side == 0) { u6[n] = u6[n] - x6; u7[n] = u7[n] - x7; }
            else { u6[n - inc] = u6[n - inc] + x6; u7[n - inc] = u7[n - inc] + x7; }
          }
    }
  }
  for (int k = k1; k <= k2 - 1; k++)
    for (int j = j1; j <= j2 - 1; j++)
      for (int i = i1; i <= i2 - 1; i++) {
        long n = IDX(b, i, j, k);
        u6[n] = 0.5 * u6[n] - vol[n] * ts6[n] - d[5 * ldv + n];
        u7[n] = 0.5 * u7[n] - vol[n] * ts7[n] - d[6 * ldv + n];
      }
  free(f);
}
