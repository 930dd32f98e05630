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
