/* Oracle: JST artificial dissipation (dissip_jameson.f90:4).
 * TEST INFRASTRUCTURE ONLY (see oracle.h). */
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

/* spectral radius |V.S| + c|S| of a face with the average state of its two
 * cells (dissip_jameson.f90:118-127) */
static double radius(const orc_block *b, const double *v, long ldv, const double *sn, const double *cson,
                     long n, long inc, long m, int d) {
  double s1 = SN(sn, b, m, d, 0), s2 = SN(sn, b, m, d, 1), s3 = SN(sn, b, m, d, 2);
  double cnds = s1 * s1 + s2 * s2 + s3 * s3;
  double uu = 0.5 * (v[ldv + n] / v[n] + v[ldv + n - inc] / v[n - inc]);
  double vv = 0.5 * (v[2 * ldv + n] / v[n] + v[2 * ldv + n - inc] / v[n - inc]);
  double ww = 0.5 * (v[3 * ldv + n] / v[n] + v[3 * ldv + n - inc] / v[n - inc]);
  double cc = 0.5 * (cson[n] + cson[n - inc]);
  return fabs(uu * s1 + vv * s2 + ww * s3) + sqrt(cnds) * cc;
}

/* dissip_jameson.f90:73-431: d(1:5) = sum over faces of the 2nd/4th
 * difference fluxes, pressure sensor per direction, first-order boundary
 * faces.  snu: scratch of nnn entries. */
void orc_dissip_jameson(const orc_block *b, double ck2, double ck4, const double *v, long ldv, double *d,
                        const double *sn, const double *ps, const double *cson) {
  long n0 = b->npc;
  double *snu = calloc(b->nnn, sizeof(double));
  for (long n = IDX(b, b->ii1 - 1, b->jj1 - 1, b->kk1 - 1); n <= IDX(b, b->ii2, b->jj2, b->kk2); n++)
    for (int e = 0; e < 5; e++) d[e * ldv + n] = 0.;
  for (int dd = 0; dd < 3; dd++) {
    long inc = INC(b, dd);
    int a1 = dd == 0 ? b->ii1 : (dd == 1 ? b->jj1 : b->kk1);
    int a2 = dd == 0 ? b->ii2 : (dd == 1 ? b->jj2 : b->kk2);
    int lo[3] = {b->ii1, b->jj1, b->kk1}, hi[3] = {b->ii2 - 1, b->jj2 - 1, b->kk2 - 1};
    for (int k = lo[2]; k <= hi[2]; k++)
      for (int j = lo[1]; j <= hi[1]; j++)
        for (int i = lo[0]; i <= hi[0]; i++) {
          long n = IDX(b, i, j, k);
          snu[n - n0] = fabs(ps[n + inc] - 2. * ps[n] + ps[n - inc]) / (ps[n + inc] + 2. * ps[n] + ps[n - inc]);
        }
    int l2[3] = {lo[0], lo[1], lo[2]}, h2[3] = {hi[0], hi[1], hi[2]};
    l2[dd] = a1 + 1; h2[dd] = a2 - 1;
    for (int k = l2[2]; k <= h2[2]; k++)
      for (int j = l2[1]; j <= h2[1]; j++)
        for (int i = l2[0]; i <= h2[0]; i++) {
          long n = IDX(b, i, j, k), m = n - n0;
          double eps2 = ck2 * fmax(snu[m - inc], snu[m]);
          double eps4 = fmax(0., ck4 - eps2);
          double rl = radius(b, v, ldv, sn, cson, n, inc, m, dd);
          for (int e = 0; e < 5; e++) {
            const double *q = v + e * ldv;
            double ds = eps2 * rl * (q[n] - q[n - inc]) - eps4 * rl * (q[n + inc] + 3. * (q[n - inc] - q[n]) - q[n - 2 * inc]);
            d[e * ldv + n] = d[e * ldv + n] - ds;
            d[e * ldv + n - inc] = d[e * ldv + n - inc] + ds;
          }
        }
    for (int side = 0; side < 2; side++) {
      int l3[3] = {lo[0], lo[1], lo[2]}, h3[3] = {hi[0], hi[1], hi[2]};
      l3[dd] = h3[dd] = side == 0 ? a1 : a2;
      for (int k = l3[2]; k <= h3[2]; k++)
        for (int j = l3[1]; j <= h3[1]; j++)
          for (int i = l3[0]; i <= h3[0]; i++) {
            long n = IDX(b, i, j, k), m = n - n0;
            double eps2 = side == 0 ? ck2 * snu[m] : ck2 * snu[m - inc];
            double rl = radius(b, v, ldv, sn, cson, n, inc, m, dd);
            for (int e = 0; e < 5; e++) {
              const double *q = v + e * ldv;
              double ds = eps2 * rl * (q[n] - q[n - inc]);
              if (side == 0) d[e * ldv + n] = d[e * ldv + n] - ds;
              else d[e * ldv + n - inc] = d[e * ldv + n - inc] + ds;
            }
          }
    }
  }
  free(snu);
}
