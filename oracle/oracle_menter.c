/* CPU restatement of the Menter k-omega model's part of the turbulence stage
 * (met_num.f90:449-456, 542-546, 815-850): the sources met_smmt with its
 * blending function, the wall correction met_mtcorf1 of that function, the
 * flux densities met_fludmt and the eddy viscosity met_komut.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Storage as in oracle.h; v holds
 * 7 planes of ldv; fd holds the six arrays txxf5x..tzzf6z (gradients of k
 * and omega on entry of the sources and of the flux densities) in planes of
 * ldv; g9 / gkgo / tprod are block-local (index m = n - npc). */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

#define NID(b) ((long)((b)->id2 - (b)->id1 + 1))
#define NIJD(b) (NID(b) * (long)((b)->jd2 - (b)->jd1 + 1))
#define IDX(b, i, j, k) \
  ((long)(b)->npc + ((i) - (b)->id1) + (long)((j) - (b)->jd1) * NID(b) + (long)((k) - (b)->kd1) * NIJD(b))

/* met_mtcorf1.f90:4 -- along the wall-normal line of every wall point of
 * every listed wall, in list order: frac = 1 on the first imxclko+2 cells
 * (:89-93); the last cell whose nearest wall point belongs to this wall
 * bounds the line (m3mxx), the last such cell with frac > 0.99 is m3lim
 * (:95-104); then frac = 1 up to m3lim and the running minimum along the
 * line beyond it (:106-131).  A line is given by its first cell (0-based
 * storage index), its m3 range and direction, the cell stride dm3 and the
 * wall-point window (mpn0, mpn1] (atindnor.f90:4). */
void orc_met_mtcorf1(const orc_wall_line *lines, long nlines, int imxclko, const int32_t *mnpar, double *frac) {
  for (long L = 0; L < nlines; L++) {
    const orc_wall_line *w = &lines[L];
    int m3min = w->m3min, m3max = w->m3max, is = w->isens3, dm3 = w->dm3;
    long nmin = w->nmin, isdm3 = (long)is * dm3;
    int m3mxx = m3min + is * (imxclko + 1);
    for (int m3 = m3min; is > 0 ? m3 <= m3mxx : m3 >= m3mxx; m3 += is)
      frac[nmin + (long)(m3 - m3min) * dm3 * is] = 1.;
    int m3lim = 0;
    for (int m3 = m3min; is > 0 ? m3 <= m3max : m3 >= m3max; m3 += is) {
      long n = nmin + (long)(m3 - m3min) * isdm3;
      int npar = mnpar[n];
      if (npar > w->mpn0 && npar <= w->mpn1) {
        m3mxx = m3;
        if (frac[n] > 0.99) m3lim = m3;
      }
    }
    if (is > 0) {
      if (m3mxx > m3min)
        for (int m3 = m3min; m3 <= m3mxx; m3++) {
          long n = nmin + (long)(m3 - m3min) * isdm3;
          frac[n] = m3 <= m3lim ? 1. : fmin(frac[n], frac[n - isdm3]);
        }
    } else {
      if (m3mxx < m3min)
        for (int m3 = m3min; m3 >= m3mxx; m3--) {
          long n = nmin + (long)(m3 - m3min) * isdm3;
          frac[n] = m3 >= m3lim ? 1. : fmin(frac[n], frac[n - isdm3]);
        }
    }
  }
}

/* met_smmt.f90:4 -- Menter k-omega sources.  kfracom = 0: frac = 1 and
 * gkgo = 0 (:63-74); else gkgo = grad k . grad omega and the blending
 * function frac = tanh(zeta^4) for zeta <= 2.5, 1 above (:76-100), then the
 * wall correction (:101-104, lines as in orc_met_mtcorf1); then with the
 * blended constants ts6 = P - betae rho k omega, ts7 = P wgam rho/mut -
 * beta rho omega^2 + 2 wsig rho gkgo/omega, cfke = 0.18 omega (:106-124). */
void orc_met_smmt(const orc_block *b, const orc_menter *mc, int kfracom, int imxclko, const double *v, long ldv,
                  const double *mu, const double *mut, const double *dist, const double *fd, const double *tprod,
                  double *cfke, double *gkgo, double *frac, double *ts6, double *ts7, const orc_wall_line *lines,
                  long nlines, const int32_t *mnpar) {
  long n0 = b->npc;
  double xgam1 = mc->beta1 / mc->betae - mc->sigma1 * (mc->okappa * mc->okappa) / sqrt(mc->betae);
  double xgam2 = mc->beta2 / mc->betae - mc->sigma2 * (mc->okappa * mc->okappa) / sqrt(mc->betae);
  double w4sig2 = 4. * mc->wsig2;
  const double *f5x = fd, *f5y = fd + ldv, *f5z = fd + 2 * ldv, *f6x = fd + 3 * ldv, *f6y = fd + 4 * ldv,
               *f6z = fd + 5 * ldv;
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - n0;
        if (kfracom == 0) {
          frac[n] = 1.;
          gkgo[m] = 0.;
          continue;
        }
        gkgo[m] = f6x[n] * f5x[n] + f6y[n] * f5y[n] + f6z[n] * f5z[n];
        double omeg = v[6 * ldv + n] / v[n];
        double coef1 = sqrt(v[5 * ldv + n] / v[n]) / (mc->betae * omeg * dist[n]);
        double dist2 = dist[n] * dist[n];
        double coef2 = 500. * mu[n] / (v[6 * ldv + n] * dist2);
        double coef4 = mc->wsig2 * v[n] * gkgo[m] / omeg;
        double dkomg = coef4 > 1e-20 ? coef4 : 1e-20;
        double coef3 = w4sig2 * v[5 * ldv + n] / (dist2 * dkomg);
        double zeta = coef1 > coef2 ? coef1 : coef2;
        zeta = zeta < coef3 ? zeta : coef3;
        if (zeta <= 2.5) {
          double z4 = ((zeta * zeta) * zeta) * zeta;  // zeta**4 as flang evaluates it
          double exp2x = exp(2. * z4);
          frac[n] = (exp2x - 1.) / (exp2x + 1.);
        } else {
          frac[n] = 1.;
        }
      }
  if (kfracom != 0) orc_met_mtcorf1(lines, nlines, imxclko, mnpar, frac);
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - n0;
        double fm1 = frac[n], fm2 = 1. - fm1;
        double wgam = fm1 * xgam1 + fm2 * xgam2;
        double wsig = fm1 * mc->wsig1 + fm2 * mc->wsig2;
        double beta = fm1 * mc->beta1 + fm2 * mc->beta2;
        double omeg = v[6 * ldv + n] / v[n];
        double sl3 = 2. * wsig * v[n] * gkgo[m] / omeg;
        ts6[n] = tprod[m] - mc->betae * v[5 * ldv + n] * omeg;
        ts7[n] = tprod[m] * wgam * v[n] / mut[n] - beta * v[6 * ldv + n] * omeg + sl3;
        cfke[n] = 0.18 * omeg;
      }
}

/* met_fludmt.f90:4 -- flux densities (mu + mut sigma) grad of k (sigme) and
 * omega (sigma), over the storage range indc(i1,j1,k1)..indc(i2-1,j2-1,k2-1);
 * the sigmas blended with frac when kfracom != 0 (:57-78). */
void orc_met_fludmt(const orc_block *b, const orc_menter *mc, int kfracom, const double *mu, const double *mut,
                    const double *frac, double *fd, long ldv) {
  long ind1 = IDX(b, b->ii1, b->jj1, b->kk1), ind2 = IDX(b, b->ii2 - 1, b->jj2 - 1, b->kk2 - 1);
  for (long n = ind1; n <= ind2; n++) {
    double se = mc->sigme1, sa = mc->sigma1;
    if (kfracom != 0) {
      double fm1 = frac[n], fm2 = 1. - fm1;
      se = fm1 * mc->sigme1 + fm2 * mc->sigme2;
      sa = fm1 * mc->sigma1 + fm2 * mc->sigma2;
    }
    double ce = mu[n] + mut[n] * se, ca = mu[n] + mut[n] * sa;
    for (int c = 0; c < 3; c++) fd[c * ldv + n] = ce * fd[c * ldv + n];
    for (int c = 3; c < 6; c++) fd[c * ldv + n] = ca * fd[c * ldv + n];
  }
}

/* met_komut.f90:4 -- eddy viscosity of the Menter models: rho k / omega
 * (komsst = 0, :46-55); with the SST limiter k / max(omega, |rot u| F2/a1),
 * a1 = sqrt(betae), F2 = tanh(zeta^2) from zeta = max(sqrt(k/rho)/(betae/2
 * omega d), 500 mu/(rho omega d^2)), exponent capped at 25 (komsst = 1,
 * :56-85); g9 holds the velocity gradients teq_gradv computes first. */
void orc_met_komut(const orc_block *b, double betae, int komsst, const double *v, long ldv, const double *mu,
                   double *mut, const double *dist, const double *g9, long ldg) {
  double betae2 = betae / 2., a1 = sqrt(betae);
  const double *dxy = g9 + ldg, *dxz = g9 + 2 * ldg, *dyx = g9 + 3 * ldg;
  const double *dyz = g9 + 5 * ldg, *dzx = g9 + 6 * ldg, *dzy = g9 + 7 * ldg;
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - b->npc;
        if (komsst == 0) {
          mut[n] = v[n] * v[5 * ldv + n] / v[6 * ldv + n];
          continue;
        }
        double omeg = v[6 * ldv + n] / v[n];
        double coef1 = sqrt(v[5 * ldv + n] / v[n]) / (betae2 * omeg * dist[n]);
        double coef2 = 500. * mu[n] / (v[6 * ldv + n] * (dist[n] * dist[n]));
        double a = dzy[m] - dyz[m], bb = dxz[m] - dzx[m], c = dyx[m] - dxy[m];
        double rota = sqrt(a * a + bb * bb + c * c);
        double zeta = coef1 > coef2 ? coef1 : coef2;
        double z2 = 2. * (zeta * zeta);
        double exp2x = exp(z2 < 25. ? z2 : 25.);
        double f2 = (exp2x - 1.) / (exp2x + 1);
        double lim = rota * f2 / a1;
        mut[n] = v[5 * ldv + n] / (omeg > lim ? omeg : lim);
      }
}

/* met_smmtr.f90:4 -- Menter sources with Durbin's realizability limit: the
 * blending function always (as met_smmt with kfracom = 1, gkgo recomputed
 * in both loops, :60-89) with the wall correction (:90-92), then ts7's
 * production divided by alpha = min(1, (1/sqrt(0.09)) / (S/omega)), S from
 * the velocity gradients g9 (:94-119). */
void orc_met_smmtr(const orc_block *b, const orc_menter *mc, int imxclko, const double *v, long ldv,
                   const double *mu, const double *mut, const double *dist, const double *fd, const double *g9,
                   long ldg, const double *tprod, double *cfke, double *frac, double *ts6, double *ts7,
                   const orc_wall_line *lines, long nlines, const int32_t *mnpar) {
  long n0 = b->npc;
  double *gk = calloc(b->nnn, sizeof(double));
  double rcmu = 1. / sqrt(0.09);
  orc_met_smmt(b, mc, 1, imxclko, v, ldv, mu, mut, dist, fd, tprod, cfke, gk, frac, ts6, ts7, lines, nlines, mnpar);
  double xgam1 = mc->beta1 / mc->betae - mc->sigma1 * (mc->okappa * mc->okappa) / sqrt(mc->betae);
  double xgam2 = mc->beta2 / mc->betae - mc->sigma2 * (mc->okappa * mc->okappa) / sqrt(mc->betae);
  const double *dxx = g9, *dxy = g9 + ldg, *dxz = g9 + 2 * ldg, *dyx = g9 + 3 * ldg, *dyy = g9 + 4 * ldg;
  const double *dyz = g9 + 5 * ldg, *dzx = g9 + 6 * ldg, *dzy = g9 + 7 * ldg, *dzz = g9 + 8 * ldg;
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - n0;
        double fm1 = frac[n], fm2 = 1. - fm1;
        double wgam = fm1 * xgam1 + fm2 * xgam2;
        double beta = fm1 * mc->beta1 + fm2 * mc->beta2;
        double omeg = v[6 * ldv + n] / v[n];
        double wsig = fm1 * mc->wsig1 + fm2 * mc->wsig2;
        double sl3 = 2. * wsig * v[n] * gk[m] / omeg;
        double a = dxx[m] * dxx[m] + dyy[m] * dyy[m] + dzz[m] * dzz[m];
        double s1 = dzy[m] + dyz[m], s2 = dxz[m] + dzx[m], s3 = dyx[m] + dxy[m];
        double ss = sqrt(4. * a / 3. + s1 * s1 + s2 * s2 + s3 * s3) / omeg;
        double alpha = rcmu / ss < 1. ? rcmu / ss : 1.;
        ts7[n] = tprod[m] * wgam * v[n] / mut[n] / alpha - beta * v[6 * ldv + n] * omeg + sl3;
      }
  free(gk);
}
