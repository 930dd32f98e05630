/* CPU restatement of the closing part of the turbulence stage of met_num
 * (met_num.f90:697-930): the line-implicit phase of the transport equations,
 * the limiters of at_cutke and the eddy viscosity of the SA and k-omega
 * models.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Storage as in oracle.h; v holds 7 planes of ldv. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NID(b) ((long)((b)->id2 - (b)->id1 + 1))
#define NIJD(b) (NID(b) * (long)((b)->jd2 - (b)->jd1 + 1))
#define IDX(b, i, j, k) \
  ((long)(b)->npc + ((i) - (b)->id1) + (long)((j) - (b)->jd1) * NID(b) + (long)((k) - (b)->kd1) * NIJD(b))
#define SN(sn, b, m, d, c) ((sn)[((long)(c) * 3 + (d)) * (b)->nnn + (m)])

/* impli2_eqt_3d.f90:4 -- implicit phase of the two transport equations:
 * line-implicit (Thomas) sweeps along i, then j, then k, each repeated
 * ibalai = 2 times, with the transverse neighbours' increments of the last
 * sweep on the right-hand side; then v(6:7) += dw.  cmt is the model's
 * diffusion constant (:82-95).  Wall-law zeroing (:384-404) is out of
 * scope (lparoi = 0). */
void orc_impli2_eqt_3d(const orc_block *b, const double *u, long ldv, const double *dt, double *v,
                       const double *mu, const double *mut, const double *cfke, const double *sn,
                       const double *vol, double cmt) {
  long n0 = b->npc, nnn = b->nnn;
  int i1 = b->ii1, i2 = b->ii2, j1 = b->jj1, j2 = b->jj2, k1 = b->kk1, k2 = b->kk2;
  long nci = 1, ncj = NID(b), nck = NIJD(b);
  double *dw6 = calloc(nnn, 8), *dw7 = calloc(nnn, 8), *rv = calloc(nnn, 8), *cd = calloc(nnn, 8);
  double *uf[3], *ce[3], *al = calloc(nnn, 8), *b6 = calloc(nnn, 8), *b7 = calloc(nnn, 8);
  for (int d = 0; d < 3; d++) {
    uf[d] = calloc(nnn, 8);
    ce[d] = calloc(nnn, 8);
  }
  for (int k = k1; k <= k2 - 1; k++)
    for (int j = j1; j <= j2 - 1; j++)
      for (int i = i1; i <= i2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - n0;
        rv[m] = (mu[n] + mut[n] / cmt) / v[n];
        cd[m] = vol[n] / dt[n] + cfke[n] * vol[n];
      }
  /* face coefficients (:126-208): faces i1..i2 (resp. j, k), all real in the
   * transverse directions */
  for (int d = 0; d < 3; d++) {
    long inc = d == 0 ? nci : (d == 1 ? ncj : nck);
    int hi[3] = {i2 - 1, j2 - 1, k2 - 1};
    hi[d] += 1;
    for (int k = k1; k <= hi[2]; k++)
      for (int j = j1; j <= hi[1]; j++)
        for (int i = i1; i <= hi[0]; i++) {
          long n = IDX(b, i, j, k), m = n - n0;
          double s1 = SN(sn, b, m, d, 0), s2 = SN(sn, b, m, d, 1), s3 = SN(sn, b, m, d, 2);
          double cnds = s1 * s1 + s2 * s2 + s3 * s3;
          double uu = 0.5 * (v[ldv + n] / v[n] + v[ldv + n - inc] / v[n - inc]);
          double vv = 0.5 * (v[2 * ldv + n] / v[n] + v[2 * ldv + n - inc] / v[n - inc]);
          double ww = 0.5 * (v[3 * ldv + n] / v[n] + v[3 * ldv + n - inc] / v[n - inc]);
          uf[d][m] = 0.5 * (uu * s1 + vv * s2 + ww * s3);
          ce[d][m] = fabs(uf[d][m]) + (rv[m] + rv[m - inc]) * cnds / (vol[n] + vol[n - inc]);
        }
  }
  for (int k = k1; k <= k2 - 1; k++)
    for (int j = j1; j <= j2 - 1; j++)
      for (int i = i1; i <= i2 - 1; i++) {
        long m = IDX(b, i, j, k) - n0;
        cd[m] = cd[m] + ce[0][m] + ce[0][m + nci] + ce[1][m] + ce[1][m + ncj] + ce[2][m] + ce[2][m + nck];
      }
  const double *u6 = u + 5 * ldv, *u7 = u + 6 * ldv;
  for (int d = 0; d < 3; d++) {
    /* the two transverse directions p, q of the sweep direction d */
    int p = d == 0 ? 1 : 0, q = d == 2 ? 1 : 2;
    long inc = d == 0 ? nci : (d == 1 ? ncj : nck);
    long ip = p == 0 ? nci : ncj, iq = q == 1 ? ncj : nck;
    if (d > 0) {
      memset(al, 0, nnn * 8);
      memset(b6, 0, nnn * 8);
      memset(b7, 0, nnn * 8);
    }
    for (int sweep = 0; sweep < 2; sweep++) {
      /* elimination from the high end of every line (:238-263, :292-317,
       * :346-371); the transverse increments are those of the last sweep */
      int lo[3] = {i1, j1, k1}, hi[3] = {i2 - 1, j2 - 1, k2 - 1};
      for (int a = hi[d]; a >= lo[d]; a--)
        for (int t2 = lo[q]; t2 <= hi[q]; t2++)
          for (int t1 = lo[p]; t1 <= hi[p]; t1++) {
            int c[3];
            c[d] = a; c[p] = t1; c[q] = t2;
            long n = IDX(b, c[0], c[1], c[2]), m = n - n0;
            double td = -uf[p][m + ip] + uf[p][m] - uf[q][m + iq] + uf[q][m];
            double tpp = ce[p][m + ip] - uf[p][m + ip], tmp_ = ce[p][m] + uf[p][m];
            double tpq = ce[q][m + iq] - uf[q][m + iq], tmq = ce[q][m] + uf[q][m];
            double r6 = -u6[n] + td * dw6[m] + tmp_ * dw6[m - ip] + tpp * dw6[m + ip] + tmq * dw6[m - iq] +
                        tpq * dw6[m + iq];
            double r7 = -u7[n] + td * dw7[m] + tmp_ * dw7[m - ip] + tpp * dw7[m + ip] + tmq * dw7[m - iq] +
                        tpq * dw7[m + iq];
            double ai = ce[d][m] + uf[d][m];
            double bi = cd[m] + uf[d][m + inc] - uf[d][m];
            double ci = -ce[d][m + inc] + uf[d][m + inc];
            al[m] = ai / (bi + ci * al[m + inc]);
            b6[m] = (r6 - ci * b6[m + inc]) / (bi + ci * al[m + inc]);
            b7[m] = (r7 - ci * b7[m + inc]) / (bi + ci * al[m + inc]);
          }
      /* substitution from the low end */
      for (int a = lo[d]; a <= hi[d]; a++)
        for (int t2 = lo[q]; t2 <= hi[q]; t2++)
          for (int t1 = lo[p]; t1 <= hi[p]; t1++) {
            int c[3];
            c[d] = a; c[p] = t1; c[q] = t2;
            long m = IDX(b, c[0], c[1], c[2]) - n0;
            dw6[m] = al[m] * dw6[m - inc] + b6[m];
            dw7[m] = al[m] * dw7[m - inc] + b7[m];
          }
    }
  }
  for (int k = k1; k <= k2 - 1; k++)
    for (int j = j1; j <= j2 - 1; j++)
      for (int i = i1; i <= i2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - n0;
        v[5 * ldv + n] = v[5 * ldv + n] + dw6[m];
        v[6 * ldv + n] = v[6 * ldv + n] + dw7[m];
      }
  free(dw6); free(dw7); free(rv); free(cd); free(al); free(b6); free(b7);
  for (int d = 0; d < 3; d++) {
    free(uf[d]);
    free(ce[d]);
  }
}

/* met_samut.f90:4 -- mut = rho nu~ fv1, fv1 = X^3/(X^3 + cv1^3), X = rho nu~/mu */
void orc_met_samut(const orc_block *b, const orc_turb *tc, const double *v, long ldv, const double *mu,
                   double *mut) {
  double cv13 = tc->cv1 * tc->cv1 * tc->cv1;
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k);
        double x = v[5 * ldv + n] / mu[n];
        double x3 = x * x * x;
        double fv1 = x3 / (x3 + cv13);
        mut[n] = v[5 * ldv + n] * fv1;
      }
}

/* met_kocmut.f90:4 -- k-omega eddy viscosity: mut = rho k / omega
 * (kwsst = 0), or the SST limiter with the vorticity of the velocity
 * gradients g9 (teq_gradv order xx xy xz yx yy yz zx zy zz, block-local m)
 * and the wall distance (kwsst = 1). */
void orc_met_kocmut(const orc_block *b, const orc_turb *tc, int kwsst, const double *v, long ldv,
                    const double *mu, double *mut, const double *dist, const double *g9, long ldg) {
  double betas = tc->betas, a1 = sqrt(betas);
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - b->npc;
        if (kwsst == 0) {
          mut[n] = v[n] * v[5 * ldv + n] / v[6 * ldv + n];
        } else {
          const double *dxx = g9, *dxy = g9 + ldg, *dxz = g9 + 2 * ldg, *dyx = g9 + 3 * ldg;
          const double *dyz = g9 + 5 * ldg, *dzx = g9 + 6 * ldg, *dzy = g9 + 7 * ldg;
          (void)dxx;
          double omeg = v[6 * ldv + n] / v[n];
          double coef1 = sqrt(v[5 * ldv + n] / v[n]) / (betas * omeg * dist[n]);
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
}

/* met_cutsa.f90:4 -- SA limiter: nu~ < 0 -> epsk, and v(7) = 0 */
void orc_met_cutsa(const orc_block *b, double epsk, double *v, long ldv) {
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k);
        if (v[5 * ldv + n] < 0.) v[5 * ldv + n] = epsk;
        v[6 * ldv + n] = 0.;
      }
}

/* met_cut.f90:4 -- both variables reset to (epsk, epse) unless both exceed
 * them, in the reference's sign arithmetic */
void orc_met_cut(const orc_block *b, double epsk, double epse, double *v, long ldv) {
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k);
        double t6 = copysign(1.0, v[5 * ldv + n] - epsk);
        double t7 = copysign(1.0, v[6 * ldv + n] - epse);
        double ts = copysign(1.0, t6 + t7 - 1.5);
        v[5 * ldv + n] = 0.5 * ((1. + ts) * v[5 * ldv + n] + (1. - ts) * epsk);
        v[6 * ldv + n] = 0.5 * ((1. + ts) * v[6 * ldv + n] + (1. - ts) * epse);
      }
}

/* met_rbse.f90:26-38 -- the 6 flux-density planes of boundary ghosts */
void orc_met_rbse(long nf, const long *ghost, const long *inner, double *fd, long ldv) {
  for (long f = 0; f < nf; f++)
    for (int q = 0; q < 6; q++) fd[q * ldv + ghost[f]] = fd[q * ldv + inner[f]];
}

/* met_rbsc.f90:84-95 (without periodic rotation): every ghost from the
 * inner values of both sides, gathered before any ghost is written */
void orc_met_rbsc(long nf, const long *ghost, const long *inner, const long *remote, double *fd, long ldv) {
  double *buf = malloc(sizeof(double) * 6 * (nf > 0 ? nf : 1));
  for (long f = 0; f < nf; f++)
    for (int q = 0; q < 6; q++) buf[6 * f + q] = fd[q * ldv + remote[f]];
  for (long f = 0; f < nf; f++)
    for (int q = 0; q < 6; q++) fd[q * ldv + ghost[f]] = 0.5 * (fd[q * ldv + inner[f]] + buf[6 * f + q]);
  free(buf);
}

/* met_bord.f90:4 -- ghost values of the two transported variables (v
 * planes 6, 7; 0-based 5, 6) on one frontier list, by its kind:
 *   0 coincident ('rc'): 0.5 (inner + remote), met_rbvc.f90:30-65 (all
 *     remote values read before any ghost is written);
 *   1 wall ('para') with kparoi = 0: 0, 0 (met_bord.f90:118-125);
 *   2 wall ('para') with kparoi = 1: 0, the inner cell's v7 (:127-134);
 *   3 far field ('nrd'): rokinf = a, roeinf = b (met_rbve's copy is
 *     overwritten, :204-214);
 *   4 symmetry ('sym'): the inner cell's values (:228-236). */
void orc_met_bord(int kind, long nf, const long *ghost, const long *inner, const long *remote, double a, double b,
                  double *v, long ldv) {
  double *k = v + 5 * ldv, *w = v + 6 * ldv;
  if (kind == 0) {
    double *buf = malloc(sizeof(double) * 2 * (nf > 0 ? nf : 1));
    for (long f = 0; f < nf; f++) { buf[2 * f] = k[remote[f]]; buf[2 * f + 1] = w[remote[f]]; }
    for (long f = 0; f < nf; f++) {
      k[ghost[f]] = 0.5 * (k[inner[f]] + buf[2 * f]);
      w[ghost[f]] = 0.5 * (w[inner[f]] + buf[2 * f + 1]);
    }
    free(buf);
    return;
  }
  for (long f = 0; f < nf; f++) {
    long g = ghost[f], i = inner[f];
    switch (kind) {
      case 1: k[g] = 0.; w[g] = 0.; break;
      case 2: k[g] = 0.; w[g] = w[i]; break;
      case 3: k[g] = a; w[g] = b; break;
      default: k[g] = k[i]; w[g] = w[i]; break;
    }
  }
}

/* met_parko.f90:4 -- omega condition of the k-omega models at a wall (kparoi
 * = 1): per wall point (line as met_mtcorf1's: first cell nmin, cells
 * nmin + s isens3 dm3 for s = 0.. along the wall normal, wall point numbers
 * (mpn0, mpn1] in mnpar), rho omega = 6 mu / (beta1 d^2) in the wall cell;
 * in the next min(10, m3tot) cells whose wall point lies on this wall, the
 * value blended with the current one over cells 2..4 (pond = (d - d2) / (d4
 * - d2), the reference's imxclko = 3) and floored by 6 mu / (beta1 d^2); the
 * ghost takes the wall cell's value.  Returns 1 (and writes nothing) when a
 * line is shorter than 10 cells, where the reference stops. */
int orc_met_parko(long nl, const orc_wall_line *ln, const long *ghost, const int *mnpar, const double *mu,
                  const double *dist, double beta1, double *v, long ldv) {
  double *w = v + 6 * ldv;
  const int m3clpar = 10, imxclko = 3;
  for (long p = 0; p < nl; p++)
    if (abs(ln[p].m3max - ln[p].m3min) + 1 < m3clpar) return 1;
  for (long p = 0; p < nl; p++) {
    const orc_wall_line *L = &ln[p];
    long st = (long)L->isens3 * L->dm3;
    long nmin = L->nmin, nminp = nmin + st, nmaxp = nmin + st * (imxclko + 1);
    int m3tot = abs(L->m3max - L->m3min) + 1;
    int nmore = m3tot < m3clpar ? m3tot : m3clpar;
    w[nmin] = 6. * mu[nmin] / (beta1 * (dist[nmin] * dist[nmin]));
    for (int s = 1; s <= nmore; s++) {
      long n = nmin + s * st;
      int npar = mnpar[n];
      if (npar > L->mpn0 && npar <= L->mpn1) {
        double rom = 6. * mu[n] / (beta1 * (dist[n] * dist[n]));
        if ((n - nmaxp) * (n - nminp) <= 0) {
          double pond = (dist[n] - dist[nminp]) / (dist[nmaxp] - dist[nminp]);
          w[n] = rom * (1. - pond) + pond * w[n];
        }
        w[n] = fmax(w[n], rom);
      }
    }
    w[ghost[p]] = w[nmin];
  }
  return 0;
}

/* met_smkor.f90:4 -- Wilcox k-omega sources with cross diffusion and the
 * strain-rate limiter (equatt '2KOR', met_num.f90:509): fd holds the
 * gradients of k and omega (met_gradtr), g9 the velocity gradients
 * (block-local, teq_gradv order), tprod the production (block-local). */
void orc_met_smkor(const orc_block *b, const orc_turb *tc, double sigmd, const double *v, long ldv,
                   const double *fd, const double *g9, long ldg, const double *tprod, double *cfke, double *ts6,
                   double *ts7) {
  double xw = tc->beta / tc->betas - tc->sigmw * (0.41 * 0.41) / sqrt(tc->betas);
  double rcmu = 1. / sqrt(0.09);
  const double *dxx = g9, *dxy = g9 + ldg, *dxz = g9 + 2 * ldg, *dyx = g9 + 3 * ldg, *dyy = g9 + 4 * ldg;
  const double *dyz = g9 + 5 * ldg, *dzx = g9 + 6 * ldg, *dzy = g9 + 7 * ldg, *dzz = g9 + 8 * ldg;
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - b->npc;
        double gkgo = fd[3 * ldv + n] * fd[n] + fd[4 * ldv + n] * fd[ldv + n] + fd[5 * ldv + n] * fd[2 * ldv + n];
        double omeg = v[6 * ldv + n] / v[n];
        double cd = sigmd * v[n] * (gkgo > 0. ? gkgo : 0.) / omeg;
        double a = dzy[m] + dyz[m], bb = dxz[m] + dzx[m], c = dyx[m] + dxy[m];
        double ss = sqrt(4. * (dxx[m] * dxx[m] + dyy[m] * dyy[m] + dzz[m] * dzz[m]) / 3. + a * a + bb * bb + c * c) /
                    omeg;
        double r = rcmu / ss;
        double alpha = 1. < r ? 1. : r;
        ts6[n] = tprod[m] - tc->betas * v[5 * ldv + n] * omeg;
        ts7[n] = xw * tprod[m] * v[6 * ldv + n] / v[5 * ldv + n] / alpha - tc->beta * v[6 * ldv + n] * omeg + cd;
        cfke[n] = 0.18 * omeg;
      }
}

/* met_smch.f90:4 -- Chien low-Reynolds k-epsilon sources; y+ from the
 * friction velocity utau of the cell's wall point mnpar (1-based, 0 =
 * none: y+ = 1000), floored at utaumin. */
void orc_met_smch(const orc_block *b, double cke1, double cke2, double utaumin, const double *v, long ldv,
                  const double *mu, const double *dist, const int *mnpar, const double *utau, const double *tprod,
                  double *ts6, double *ts7) {
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - b->npc;
        int mp = mnpar[n];
        double yplus = 1000.;
        if (mp >= 1) {
          double ut = fabs(utau[mp - 1]);
          yplus = (ut > utaumin ? ut : utaumin) * dist[n] * v[n] / mu[n];
        }
        double dist2 = dist[n] * dist[n];
        double rodst = v[n] * dist2;
        ts6[n] = tprod[m] - v[6 * ldv + n] - 2. * mu[n] * v[5 * ldv + n] / rodst;
        double epssk = v[6 * ldv + n] / v[5 * ldv + n];
        double rtur = v[5 * ldv + n] / (mu[n] * epssk);
        double f2 = 1. - .22222222 * exp(-rtur * rtur / 36.);
        double sch1 = cke1 * epssk * tprod[m];
        double sch2 = cke2 * f2 * epssk * v[6 * ldv + n];
        double sch3 = 2. * mu[n] * (v[6 * ldv + n] / rodst) * exp(-yplus / 2.);
        ts7[n] = sch1 - sch2 - sch3;
      }
}
