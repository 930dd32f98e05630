/* Oracle: viscous stresses (zgrad2 + zfluto), matrix-free point-Jacobi
 * implicit phase (implimf_3d), boundary ghost filling.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <float.h>
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

/* Green-Gauss gradients of (u, v, w, T): zgrad2.f90:148-485, with the
 * face weights cmu* = 1 (inivec.f90:107-115).  g: 12 block-local arrays of
 * nnn entries, order dvxx dvxy dvxz dvyx dvyy dvyz dvzx dvzy dvzz dtx dty dtz
 * (dv<a><b> = d v_a / d x_b). */
void orc_zgrad2(const orc_block *b, const double *sn, const double *vol, const double *s, long ldv,
                const double *temp, double *g) {
  long n0 = b->npc, nnn = b->nnn;
  int i1 = b->ii1, i2 = b->ii2, j1 = b->jj1, j2 = b->jj2, k1 = b->kk1, k2 = b->kk2;
  double *vx = orc_zeros(nnn), *vy = orc_zeros(nnn), *vz = orc_zeros(nnn);
#pragma omp parallel for schedule(static)
  for (int k = k1 - 1; k <= k2; k++)
    for (int j = j1 - 1; j <= j2; j++)
      for (int i = i1 - 1; i <= i2; i++) {
        long n = IDX(b, i, j, k), m = n - n0;
        vx[m] = s[ldv + n] / s[n];
        vy[m] = s[2 * ldv + n] / s[n];
        vz[m] = s[3 * ldv + n] / s[n];
      }
  for (long m = 0; m < 12 * nnn; m++) g[m] = 0.;
  double *G[12];
  for (int c = 0; c < 12; c++) G[c] = g + c * nnn;
  const double eps = 0.00001;
  for (int d = 2; d >= 0; d--) { /* k faces, then j, then i (zgrad2.f90:181-457) */
    long inc = INC(b, d);
    int lo[3] = {i1, j1, k1}, hi[3] = {i2 - 1, j2 - 1, k2 - 1};
    int a1 = d == 0 ? i1 : (d == 1 ? j1 : k1), a2 = d == 0 ? i2 : (d == 1 ? j2 : k2);
    lo[d] = a1 + 1; hi[d] = a2 - 1;
    /* interior faces (the reference's j sweep runs j outer, k inner,
     * :277-278; each cell receives its two j faces in increasing j either
     * way, so the threads share k there): FOR3D splits a direction other
     * than d, the result is the serial one bit for bit */
    FOR3D(lo, hi, d) {
          long n = IDX(b, i, j, k), m = n - n0;
          double sx = SN(sn, b, m, d, 0), sy = SN(sn, b, m, d, 1), sz = SN(sn, b, m, d, 2);
          double fv[4] = {vx[m] + vx[m - inc], vy[m] + vy[m - inc], vz[m] + vz[m - inc],
                          temp[n] + temp[n - inc]};
          for (int a = 0; a < 4; a++) {
            double t0 = fv[a] * sx, t1 = fv[a] * sy, t2 = fv[a] * sz;
            G[3 * a][m] = G[3 * a][m] - t0;
            G[3 * a + 1][m] = G[3 * a + 1][m] - t1;
            G[3 * a + 2][m] = G[3 * a + 2][m] - t2;
            G[3 * a][m - inc] = G[3 * a][m - inc] + t0;
            G[3 * a + 1][m - inc] = G[3 * a + 1][m - inc] + t1;
            G[3 * a + 2][m - inc] = G[3 * a + 2][m - inc] + t2;
          }
        }
    /* boundary faces carry twice the ghost value (:231-269 and siblings) */
    for (int side = 0; side < 2; side++) {
      int l2[3] = {i1, j1, k1}, h2[3] = {i2 - 1, j2 - 1, k2 - 1};
      l2[d] = h2[d] = side == 0 ? a1 : a2;
#pragma omp parallel for schedule(static)
      for (int k = l2[2]; k <= h2[2]; k++)
        for (int j = l2[1]; j <= h2[1]; j++)
          for (int i = l2[0]; i <= h2[0]; i++) {
            long n = IDX(b, i, j, k), m = n - n0;
            long mg = side == 0 ? m - inc : m, mr = side == 0 ? m : m - inc;
            double gv[4] = {vx[mg], vy[mg], vz[mg], temp[mg + n0]};
            for (int a = 0; a < 4; a++)
              for (int c = 0; c < 3; c++) {
                double t = SN(sn, b, m, d, c) * 2 * gv[a];
                if (side == 0) G[3 * a + c][mr] = G[3 * a + c][mr] - t;
                else G[3 * a + c][mr] = G[3 * a + c][mr] + t;
              }
          }
    }
  }
  /* scaling by 1/(2 vol), guarded for vol <= 0 (:463-485), over the
   * contiguous storage range first..last real cell */
#pragma omp parallel for schedule(static)
  for (long n = IDX(b, i1, j1, k1); n <= IDX(b, i2 - 1, j2 - 1, k2 - 1); n++) {
    long m = n - n0;
    double ts = copysign(0.5, -vol[n]);
    double vols = (0.5 + ts) * eps + (0.5 - ts) * vol[n];
    double c0 = .5 / (vols + DBL_MIN);
    for (int c = 0; c < 12; c++) G[c][m] = G[c][m] * c0;
  }
  free(vx); free(vy); free(vz);
}

/* zfluto.f90:104-121.  tq: 9 arrays of ldv (txx txy txz tyy tyz tzz qx qy qz). */
void orc_zfluto(const orc_block *b, const orc_phys *ph, const double *mu, const double *mut,
                double *tq, long ldv, const double *g) {
  long n0 = b->npc, nnn = b->nnn;
  const double ds3 = 2. / 3.;
#pragma omp parallel for schedule(static)
  for (int k = b->kk1; k <= b->kk2 - 1; k++)
    for (int j = b->jj1; j <= b->jj2 - 1; j++)
      for (int i = b->ii1; i <= b->ii2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - n0;
        double xx = g[m], xy = g[nnn + m], xz = g[2 * nnn + m], yx = g[3 * nnn + m], yy = g[4 * nnn + m];
        double yz = g[5 * nnn + m], zx = g[6 * nnn + m], zy = g[7 * nnn + m], zz = g[8 * nnn + m];
        double mt = mu[n] + mut[n];
        tq[n] = ds3 * mt * (2 * xx - yy - zz);
        tq[3 * ldv + n] = ds3 * mt * (2 * yy - xx - zz);
        tq[5 * ldv + n] = ds3 * mt * (2 * zz - xx - yy);
        tq[ldv + n] = mt * (xy + yx);
        tq[2 * ldv + n] = mt * (xz + zx);
        tq[4 * ldv + n] = mt * (yz + zy);
        double kq = ph->cp * (mu[n] / ph->pr + mut[n] / ph->prt);
        tq[6 * ldv + n] = kq * g[9 * nnn + m];
        tq[7 * ldv + n] = kq * g[10 * nnn + m];
        tq[8 * ldv + n] = kq * g[11 * nnn + m];
      }
}

/* ------------------------------------------------------------------------ */
/* Boundary ghosts.  Face lists hold flat indices of ghost / adjacent real
 * cell / coincident remote real cell, in the reference's listing order. */

/* rbvc.f90:84-117 (no periodic rotation): ghost = face average of the two
 * real cells, also for p, T, c. */
void orc_bc_coincident(long nf, const long *ghost, const long *inner, const long *remote, double *v,
                       long ldv, double *ps, double *temp, double *cson) {
  for (long f = 0; f < nf; f++) {
    long nd = ghost[f], ndm = inner[f], nr = remote[f];
    for (int e = 0; e < 5; e++) v[e * ldv + nd] = 0.5 * (v[e * ldv + ndm] + v[e * ldv + nr]);
    ps[nd] = 0.5 * (ps[ndm] + ps[nr]);
    temp[nd] = 0.5 * (temp[ndm] + temp[nr]);
    cson[nd] = 0.5 * (cson[ndm] + cson[nr]);
  }
}

/* clpara.f90:44-58: adiabatic no-slip wall. */
void orc_bc_wall(long nf, const long *ghost, const long *inner, double *v, long ldv, double *ps,
                 double *temp, double *cson) {
  for (long f = 0; f < nf; f++) {
    long nc = ghost[f], ni = inner[f];
    double rhoe1 = v[4 * ldv + ni] - 0.5 * (v[ldv + ni] * v[ldv + ni] + v[2 * ldv + ni] * v[2 * ldv + ni] +
                                            v[3 * ldv + ni] * v[3 * ldv + ni]) / v[ni];
    v[nc] = v[ni];
    v[ldv + nc] = 0.;
    v[2 * ldv + nc] = 0.;
    v[3 * ldv + nc] = 0.;
    v[4 * ldv + nc] = rhoe1;
    ps[nc] = ps[ni];
    temp[nc] = temp[ni];
    cson[nc] = cson[ni];
  }
}

/* rbve.f90 copy then clnrd.f90:60-134: non-reflecting far field against the
 * state w0 = (rho, rho u, rho v, rho w, rho E) (utnrd.f90 without rotation).
 * nx, ny, nz: unit normals pointing into the domain. */
void orc_bc_farfield(long nf, const long *ghost, const long *inner, const double *nxn,
                     const double *nyn, const double *nzn, const double *w0, const orc_phys *ph,
                     double *v, long ldv, double *ps, double *temp, double *cson) {
  const double gam = ph->gam, gam1 = ph->gam1, pinfl = ph->pinfl;
  for (long f = 0; f < nf; f++) {
    long nl = ghost[f], ni = inner[f];
    for (int e = 0; e < 5; e++) v[e * ldv + nl] = v[e * ldv + ni];
    double nx = nxn[f], ny = nyn[f], nz = nzn[f];
    double ro0 = v[nl];
    double ps0 = gam1 * (v[4 * ldv + nl] - 0.5 * (v[ldv + nl] * v[ldv + nl] + v[2 * ldv + nl] * v[2 * ldv + nl] +
                                                  v[3 * ldv + nl] * v[3 * ldv + nl]) / v[nl] - pinfl);
    double roc0 = sqrt(gam * ps0 * ro0);
    double roqn0 = v[ldv + nl] * nx + v[2 * ldv + nl] * ny + v[3 * ldv + nl] * nz;
    double epsm = .5 + copysign(.5, roc0 - roqn0);
    double eps0 = .5 + copysign(.5, -roqn0);
    double epsp = .5 + copysign(.5, -roc0 - roqn0);
    double rod = w0[0];
    double qxd = w0[1] / rod, qyd = w0[2] / rod, qzd = w0[3] / rod;
    double pd = gam1 * (w0[4] - .5 * rod * (qxd * qxd + qyd * qyd + qzd * qzd) - pinfl);
    double qnd = qxd * nx + qyd * ny + qzd * nz;
    double qtxd = qxd - qnd * nx, qtyd = qyd - qnd * ny, qtzd = qzd - qnd * nz;
    double ros = v[nl];
    double qxs = v[ldv + nl] / ros, qys = v[2 * ldv + nl] / ros, qzs = v[3 * ldv + nl] / ros;
    double p_s = gam1 * (v[4 * ldv + nl] - .5 * ros * (qxs * qxs + qys * qys + qzs * qzs) - pinfl);
    double qns = qxs * nx + qys * ny + qzs * nz;
    double qtxs = qxs - qns * nx, qtys = qys - qns * ny, qtzs = qzs - qns * nz;
    double qtx = eps0 * qtxs + (1. - eps0) * qtxd;
    double qty = eps0 * qtys + (1. - eps0) * qtyd;
    double qtz = eps0 * qtzs + (1. - eps0) * qtzd;
    double am = epsm * (p_s - roc0 * qns) + (1. - epsm) * (pd - roc0 * qnd);
    double ap = epsp * (p_s + roc0 * qns) + (1. - epsp) * (pd + roc0 * qnd);
    double qn = 0.5 * (ap - am) / roc0;
    double p = 0.5 * (ap + am);
    double bs = (p - p_s) * (ro0 * ro0) / (roc0 * roc0) + ros;
    double b0 = (p - pd) * (ro0 * ro0) / (roc0 * roc0) + rod;
    double ro = eps0 * bs + (1. - eps0) * b0;
    v[nl] = ro;
    v[ldv + nl] = ro * (qtx + qn * nx);
    v[2 * ldv + nl] = ro * (qty + qn * ny);
    v[3 * ldv + nl] = ro * (qtz + qn * nz);
    v[4 * ldv + nl] = p / gam1 + pinfl + 0.5 * (v[ldv + nl] * v[ldv + nl] + v[2 * ldv + nl] * v[2 * ldv + nl] +
                                                  v[3 * ldv + nl] * v[3 * ldv + nl]) / ro;
    ps[nl] = p;
    temp[nl] = gam * ps[nl] / ro;
    cson[nl] = sqrt(temp[nl]);
  }
}

/* utnrd.f90:59-105 in a non-rotating frame (omg = 0, so the node
 * coordinates drop out): the far-field conservative state from the Mach
 * number, the incidence angles alpha, beta (degrees) and the stagnation
 * state roa1, aa1, pa1 (c_nzst.f90:56-59). */
void orc_utnrd(double rmach, double alpha0, double beta0, double roa1, double aa1, double pa1,
               const orc_phys *ph, double *w0) {
  const double gam = ph->gam, gam1 = ph->gam1, gam2 = 0.5 * ph->gam1, gam4 = 1. / ph->gam1;
  double pis2 = atan2(1.0, 0.0), degrad = pis2 / 90.0;
  double ro = roa1 / pow(1. + gam2 * (rmach * rmach), gam4);
  double a = aa1 / pow(1. + gam2 * (rmach * rmach), .5);
  double p = pa1 / pow(1. + gam2 * (rmach * rmach), gam / gam1);
  double q = rmach * a;
  double alphar = alpha0 * degrad, betar = beta0 * degrad;
  w0[0] = ro;
  w0[1] = ro * q * cos(alphar) * cos(betar);
  w0[2] = -ro * q * sin(betar);
  w0[3] = ro * q * sin(alphar) * cos(betar);
  w0[4] = p / gam1 + ph->pinfl + 0.5 * ro * (q * q);
}

/* rfve.f90:24-41: ghost = adjacent real cell for v (5 planes), p, T, c. */
void orc_rfve(long nf, const long *ghost, const long *inner, double *v, long ldv, double *ps, double *temp,
              double *cson) {
  for (long f = 0; f < nf; f++) {
    long nd = ghost[f], ni = inner[f];
    for (int e = 0; e < 5; e++) v[e * ldv + nd] = v[e * ldv + ni];
    ps[nd] = ps[ni];
    temp[nd] = temp[ni];
    cson[nd] = cson[ni];
  }
}

/* clsym.f90:56-79: symmetry plane.  The ghost takes the adjacent cell's
 * density and energy and its velocity without the normal component; the
 * ghost's p, T and c are not written (clsym has no such arguments), so they
 * keep whatever they held. */
void orc_bc_sym(long nf, const long *ghost, const long *inner, const double *nxn, const double *nyn,
                const double *nzn, double *v, long ldv) {
  for (long f = 0; f < nf; f++) {
    long nl = ghost[f], ni = inner[f];
    double ro1 = v[ni];
    double qx1 = v[ldv + ni] / ro1, qy1 = v[2 * ldv + ni] / ro1, qz1 = v[3 * ldv + ni] / ro1;
    double roe1 = v[4 * ldv + ni];
    double qn1 = qx1 * nxn[f] + qy1 * nyn[f] + qz1 * nzn[f];
    double qtx = qx1 - qn1 * nxn[f], qty = qy1 - qn1 * nyn[f], qtz = qz1 - qn1 * nzn[f];
    v[nl] = ro1;
    v[ldv + nl] = ro1 * qtx;
    v[2 * ldv + nl] = ro1 * qty;
    v[3 * ldv + nl] = ro1 * qtz;
    v[4 * ldv + nl] = roe1;
  }
}

/* rbte.f90:46-59 (tq ghost = adjacent real cell) for non-coincident faces and
 * rbtc.f90:104-117 (average with the remote cell) for coincident faces. */
void orc_bc_tq_copy(long nf, const long *ghost, const long *inner, double *tq, long ldv) {
  for (long f = 0; f < nf; f++)
    for (int c = 0; c < 9; c++) tq[c * ldv + ghost[f]] = tq[c * ldv + inner[f]];
}
void orc_bc_tq_coincident(long nf, const long *ghost, const long *inner, const long *remote,
                          double *tq, long ldv) {
  for (long f = 0; f < nf; f++)
    for (int c = 0; c < 9; c++)
      tq[c * ldv + ghost[f]] = 0.5 * (tq[c * ldv + inner[f]] + tq[c * ldv + remote[f]]);
}

/* ------------------------------------------------------------------------ */
/* implimf_3d.f90:87-486: coefficients, lmx point-Jacobi sub-iterations with
 * exact flux-difference off-diagonals, then v += d.  ns=0 drops the viscous
 * spectral radius (mu, mut unused).  d: 5 arrays of ldv (the increment). */
void orc_implimf_3d(const orc_block *b, const orc_phys *ph, int ns, const double *u, const double *dt,
                    double *v, long ldv, double *d, const double *ff, const double *mu,
                    const double *mut, int lmx, const double *sn, const double *vol, double dtpas,
                    int kfmg, int ityprk, const double *ps, const double *cson) {
  const double gam = ph->gam, gam1 = ph->gam1, pinfl = ph->pinfl;
  long n0 = b->npc, nnn = b->nnn;
  int i1 = b->ii1, i2 = b->ii2, j1 = b->jj1, j2 = b->jj2, k1 = b->kk1, k2 = b->kk2;
  long nci = INC(b, 0), ncj = INC(b, 1), nck = INC(b, 2);
  double *df[9], *coefe[3], *rv = orc_zeros(nnn), *diag = orc_zeros(nnn), *d2w[5];
  for (int c = 0; c < 9; c++) df[c] = orc_zeros(nnn);
  for (int c = 0; c < 3; c++) coefe[c] = orc_zeros(nnn);
  for (int c = 0; c < 5; c++) d2w[c] = orc_zeros(nnn);
  const double fact = 1.5;
  for (long n = IDX(b, i1 - 1, j1 - 1, k1 - 1); n <= IDX(b, i2 + 1, j2 + 1, k2 + 1); n++)
    for (int e = 0; e < 5; e++) d[e * ldv + n] = 0.;
#pragma omp parallel for schedule(static)
  for (int k = k1; k <= k2 - 1; k++)
    for (int j = j1; j <= j2 - 1; j++)
      for (int i = i1; i <= i2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - n0;
        rv[m] = ns ? gam * (mu[n] / ph->pr + mut[n] / ph->prt) / v[n] : 0.;
        diag[m] = vol[n] / dt[n];
      }
  for (int dd = 0; dd < 3; dd++) {
    long inc = INC(b, dd);
    int lo[3] = {i1, j1, k1}, hi[3] = {i2 - 1, j2 - 1, k2 - 1};
    hi[dd] += 1; /* faces lo..hi of the direction, incl. the two boundaries */
#pragma omp parallel for schedule(static)
    for (int k = lo[2]; k <= hi[2]; k++)
      for (int j = lo[1]; j <= hi[1]; j++)
        for (int i = lo[0]; i <= hi[0]; i++) {
          long n = IDX(b, i, j, k), m = n - n0;
          double sx = SN(sn, b, m, dd, 0), sy = SN(sn, b, m, dd, 1), sz = SN(sn, b, m, dd, 2);
          double cnds = sx * sx + sy * sy + sz * sz;
          double uu = 0.5 * (v[ldv + n] / v[n] + v[ldv + n - inc] / v[n - inc]);
          double vv = 0.5 * (v[2 * ldv + n] / v[n] + v[2 * ldv + n - inc] / v[n - inc]);
          double ww = 0.5 * (v[3 * ldv + n] / v[n] + v[3 * ldv + n - inc] / v[n - inc]);
          double vn = uu * sx + vv * sy + ww * sz;
          double cc = 0.5 * (cson[n] + cson[n - inc]);
          coefe[dd][m] = 0.5 * (fabs(vn) + sqrt(cnds) * cc) + (rv[m] + rv[m - inc]) * cnds / (vol[n] + vol[n - inc]);
        }
  }
#pragma omp parallel for schedule(static)
  for (int k = k1; k <= k2 - 1; k++)
    for (int j = j1; j <= j2 - 1; j++)
      for (int i = i1; i <= i2 - 1; i++) {
        long m = IDX(b, i, j, k) - n0;
        diag[m] = diag[m] + coefe[0][m] + coefe[0][m + nci] + coefe[1][m] + coefe[1][m + ncj] +
                  coefe[2][m] + coefe[2][m + nck];
        if (kfmg == 3) diag[m] = diag[m] + fact * vol[m + n0] / dtpas;
      }
  for (int ls = 1; ls <= lmx; ls++) {
#pragma omp parallel for schedule(static)
    for (int k = k1; k <= k2 - 1; k++)
      for (int j = j1; j <= j2 - 1; j++)
        for (int i = i1; i <= i2 - 1; i++) {
          long n = IDX(b, i, j, k), m = n - n0;
          for (int e = 0; e < 5; e++)
            d2w[e][m] = ityprk == 0 ? -u[e * ldv + n] : -u[e * ldv + n] - ff[e * ldv + n];
        }
    for (int dd = 0; dd < 3; dd++) {
      long inc = INC(b, dd);
      int lo[3] = {i1, j1, k1}, hi[3] = {i2 - 1, j2 - 1, k2 - 1};
      hi[dd] += 1;
      FOR3D(lo, hi, dd) {
            long n = IDX(b, i, j, k), m = n - n0;
            double s1 = SN(sn, b, m, dd, 0), s2 = SN(sn, b, m, dd, 1), s3 = SN(sn, b, m, dd, 2);
            /* df order: xx xy xz ex yy yz ey zz ez (implimf_3d.f90:97-105) */
            double *dxx = df[0], *dxy = df[1], *dxz = df[2], *dex = df[3], *dyy = df[4], *dyz = df[5];
            double *dey = df[6], *dzz = df[7], *dez = df[8];
            double tn[5];
            tn[0] = 0.5 * (d[ldv + n] + d[ldv + n - inc]) * s1 + 0.5 * (d[2 * ldv + n] + d[2 * ldv + n - inc]) * s2 +
                    0.5 * (d[3 * ldv + n] + d[3 * ldv + n - inc]) * s3;
            tn[1] = 0.5 * (dxx[m] + dxx[m - inc]) * s1 + 0.5 * (dxy[m] + dxy[m - inc]) * s2 +
                    0.5 * (dxz[m] + dxz[m - inc]) * s3;
            tn[2] = 0.5 * (dxy[m] + dxy[m - inc]) * s1 + 0.5 * (dyy[m] + dyy[m - inc]) * s2 +
                    0.5 * (dyz[m] + dyz[m - inc]) * s3;
            tn[3] = 0.5 * (dxz[m] + dxz[m - inc]) * s1 + 0.5 * (dyz[m] + dyz[m - inc]) * s2 +
                    0.5 * (dzz[m] + dzz[m - inc]) * s3;
            tn[4] = 0.5 * (dex[m] + dex[m - inc]) * s1 + 0.5 * (dey[m] + dey[m - inc]) * s2 +
                    0.5 * (dez[m] + dez[m - inc]) * s3;
            for (int e = 0; e < 5; e++) {
              d2w[e][m] = d2w[e][m] + tn[e] + coefe[dd][m] * d[e * ldv + n - inc] +
                          coefe[dd][m + inc] * d[e * ldv + n + inc];
            }
            for (int e = 0; e < 5; e++) d2w[e][m - inc] = d2w[e][m - inc] - tn[e];
          }
    }
#pragma omp parallel for schedule(static)
    for (int k = k1; k <= k2 - 1; k++)
      for (int j = j1; j <= j2 - 1; j++)
        for (int i = i1; i <= i2 - 1; i++) {
          long n = IDX(b, i, j, k), m = n - n0;
          for (int e = 0; e < 5; e++) d[e * ldv + n] = d2w[e][m] / diag[m];
          double wi1 = v[n] + d[n], wi2 = v[ldv + n] + d[ldv + n], wi3 = v[2 * ldv + n] + d[2 * ldv + n];
          double wi4 = v[3 * ldv + n] + d[3 * ldv + n], wi5 = v[4 * ldv + n] + d[4 * ldv + n];
          double ui = wi2 / wi1, vi = wi3 / wi1, wi = wi4 / wi1;
          double pres = gam1 * (wi5 - .5 * wi1 * (ui * ui + vi * vi + wi * wi) - pinfl);
          double p = ps[n];
          double V1 = v[n], V2 = v[ldv + n], V3 = v[2 * ldv + n], V4 = v[3 * ldv + n], V5 = v[4 * ldv + n];
          double fxx = V2 * (V2 / V1) + p, fxy = V3 * (V2 / V1), fxz = V4 * (V2 / V1);
          double fyy = V3 * (V3 / V1) + p, fyz = V4 * (V3 / V1), fzz = V4 * (V4 / V1) + p;
          double fex = (V5 + p - pinfl) * V2 / V1, fey = (V5 + p - pinfl) * V3 / V1;
          double fez = (V5 + p - pinfl) * V4 / V1;
          df[0][m] = wi1 * (ui * ui) + pres - fxx;
          df[1][m] = wi1 * ui * vi - fxy;
          df[2][m] = wi1 * ui * wi - fxz;
          df[3][m] = ui * (wi5 + pres - pinfl) - fex;
          df[4][m] = wi1 * (vi * vi) + pres - fyy;
          df[5][m] = wi1 * vi * wi - fyz;
          df[6][m] = vi * (wi5 + pres - pinfl) - fey;
          df[7][m] = wi1 * (wi * wi) + pres - fzz;
          df[8][m] = wi * (wi5 + pres - pinfl) - fez;
        }
  }
#pragma omp parallel for schedule(static)
  for (int k = k1; k <= k2 - 1; k++)
    for (int j = j1; j <= j2 - 1; j++)
      for (int i = i1; i <= i2 - 1; i++) {
        long n = IDX(b, i, j, k);
        for (int e = 0; e < 5; e++) v[e * ldv + n] = v[e * ldv + n] + d[e * ldv + n];
      }
  for (int c = 0; c < 9; c++) free(df[c]);
  for (int c = 0; c < 3; c++) free(coefe[c]);
  for (int c = 0; c < 5; c++) free(d2w[c]);
  free(rv); free(diag);
}

/* implimf_eu.f90:6 -- point-Jacobi implicit phase of the Euler equations.
 * Differs from implimf_3d in its coefficients and flux increments:
 *   diag  = vol (fact + 1/dt), fact = 1.5/dt1min when kfmg = 3 (:85,131)
 *           + sum over directions of coefe(m) + coefe(m+inc) (:168),
 *   coefe = 0.5 |V.S| + |S| c (no viscous part, 0.5 on |V.S| only) with
 *           V from reciprocal densities (:147-158),
 *   dF    from momenta with wi1 = 1/(rho+d1) and the caller's ps; the
 *           z energy increment lacks the w factor exactly as in :312;
 *   the final update adds all nvar planes of d to v (v(n,:) = v(n,:) +
 *           d(n,:), :329), planes 6.. included although this routine leaves
 *           them as the caller passed them. */
void orc_implimf_eu(const orc_block *b, const orc_phys *ph, const double *u, const double *dt, double *v,
                    long ldv, int nvar, double *d, const double *ff, int lmx, const double *sn, const double *vol,
                    double dt1min, int kfmg, int ityprk, const double *ps, const double *cson) {
  const double gam1 = ph->gam1, pinfl = ph->pinfl;
  long n0 = b->npc, nnn = b->nnn;
  int i1 = b->ii1, i2 = b->ii2, j1 = b->jj1, j2 = b->jj2, k1 = b->kk1, k2 = b->kk2;
  double *df[9], *coefe[3], *diag = orc_zeros(nnn), *d2w[5];
  for (int c = 0; c < 9; c++) df[c] = orc_zeros(nnn);
  for (int c = 0; c < 3; c++) coefe[c] = orc_zeros(nnn);
  for (int c = 0; c < 5; c++) d2w[c] = orc_zeros(nnn);
  double fact = kfmg == 3 ? 1.5 / dt1min : 0.;
  for (int e = 0; e < 5; e++)
    for (long n = IDX(b, i1 - 1, j1 - 1, k1 - 1); n <= IDX(b, i2 + 1, j2 + 1, k2 + 1); n++) d[e * ldv + n] = 0.;
  for (int k = k1; k <= k2 - 1; k++)
    for (int j = j1; j <= j2 - 1; j++)
      for (int i = i1; i <= i2 - 1; i++) {
        long n = IDX(b, i, j, k), m = n - n0;
        diag[m] = vol[n] * (fact + 1. / dt[n]);
      }
  for (int dd = 0; dd < 3; dd++) {
    long inc = INC(b, dd);
    int lo[3] = {i1, j1, k1}, hi[3] = {i2 - 1, j2 - 1, k2 - 1};
    hi[dd] += 1; /* faces of the direction, both boundaries included */
    for (int k = lo[2]; k <= hi[2]; k++)
      for (int j = lo[1]; j <= hi[1]; j++)
        for (int i = lo[0]; i <= hi[0]; i++) {
          long n = IDX(b, i, j, k), m = n - n0;
          double s1 = SN(sn, b, m, dd, 0), s2 = SN(sn, b, m, dd, 1), s3 = SN(sn, b, m, dd, 2);
          double cnds = s1 * s1 + s2 * s2 + s3 * s3;
          double norm = 1. / v[n], norm1 = 1. / v[n - inc];
          double uu = 0.5 * (v[ldv + n] * norm + v[ldv + n - inc] * norm1);
          double vv = 0.5 * (v[2 * ldv + n] * norm + v[2 * ldv + n - inc] * norm1);
          double ww = 0.5 * (v[3 * ldv + n] * norm + v[3 * ldv + n - inc] * norm1);
          double cc = 0.5 * (cson[n] + cson[n - inc]);
          coefe[dd][m] = 0.5 * fabs(uu * s1 + vv * s2 + ww * s3) + sqrt(cnds) * cc;
        }
    for (int k = lo[2]; k <= hi[2]; k++)
      for (int j = lo[1]; j <= hi[1]; j++)
        for (int i = lo[0]; i <= hi[0]; i++) {
          long m = IDX(b, i, j, k) - n0;
          diag[m] = diag[m] + coefe[dd][m] + coefe[dd][m + inc];
        }
  }
  for (int ls = 1; ls <= lmx; ls++) {
    for (int k = k1; k <= k2 - 1; k++)
      for (int j = j1; j <= j2 - 1; j++)
        for (int i = i1; i <= i2 - 1; i++) {
          long n = IDX(b, i, j, k), m = n - n0;
          for (int e = 0; e < 5; e++) d2w[e][m] = -u[e * ldv + n];
        }
    if (ityprk != 0)
      for (int k = k1; k <= k2 - 1; k++)
        for (int j = j1; j <= j2 - 1; j++)
          for (int i = i1; i <= i2 - 1; i++) {
            long n = IDX(b, i, j, k), m = n - n0;
            for (int e = 0; e < 5; e++) d2w[e][m] = d2w[e][m] - ff[e * ldv + n];
          }
    for (int dd = 0; dd < 3; dd++) {
      long inc = INC(b, dd);
      int lo[3] = {i1, j1, k1}, hi[3] = {i2 - 1, j2 - 1, k2 - 1};
      hi[dd] += 1;
      for (int k = lo[2]; k <= hi[2]; k++)
        for (int j = lo[1]; j <= hi[1]; j++)
          for (int i = lo[0]; i <= hi[0]; i++) {
            long n = IDX(b, i, j, k), m = n - n0;
            double s1 = SN(sn, b, m, dd, 0), s2 = SN(sn, b, m, dd, 1), s3 = SN(sn, b, m, dd, 2);
            double *dxx = df[0], *dxy = df[1], *dxz = df[2], *dex = df[3], *dyy = df[4], *dyz = df[5];
            double *dey = df[6], *dzz = df[7], *dez = df[8];
            double tn[5];
            tn[0] = 0.5 * ((d[ldv + n] + d[ldv + n - inc]) * s1 + (d[2 * ldv + n] + d[2 * ldv + n - inc]) * s2 +
                           (d[3 * ldv + n] + d[3 * ldv + n - inc]) * s3);
            tn[1] = 0.5 * ((dxx[m] + dxx[m - inc]) * s1 + (dxy[m] + dxy[m - inc]) * s2 + (dxz[m] + dxz[m - inc]) * s3);
            tn[2] = 0.5 * ((dxy[m] + dxy[m - inc]) * s1 + (dyy[m] + dyy[m - inc]) * s2 + (dyz[m] + dyz[m - inc]) * s3);
            tn[3] = 0.5 * ((dxz[m] + dxz[m - inc]) * s1 + (dyz[m] + dyz[m - inc]) * s2 + (dzz[m] + dzz[m - inc]) * s3);
            tn[4] = 0.5 * ((dex[m] + dex[m - inc]) * s1 + (dey[m] + dey[m - inc]) * s2 + (dez[m] + dez[m - inc]) * s3);
            for (int e = 0; e < 5; e++)
              d2w[e][m] = d2w[e][m] + tn[e] + coefe[dd][m] * d[e * ldv + n - inc] +
                          coefe[dd][m + inc] * d[e * ldv + n + inc];
            for (int e = 0; e < 5; e++) d2w[e][m - inc] = d2w[e][m - inc] - tn[e];
          }
    }
    for (int k = k1; k <= k2 - 1; k++)
      for (int j = j1; j <= j2 - 1; j++)
        for (int i = i1; i <= i2 - 1; i++) {
          long n = IDX(b, i, j, k), m = n - n0;
          for (int e = 0; e < 5; e++) d[e * ldv + n] = d2w[e][m] / diag[m];
          double wi1 = 1. / (v[n] + d[n]);
          double ui = v[ldv + n] + d[ldv + n], vi = v[2 * ldv + n] + d[2 * ldv + n];
          double wi = v[3 * ldv + n] + d[3 * ldv + n], wi5 = v[4 * ldv + n] + d[4 * ldv + n];
          double pres = gam1 * (wi5 - .5 * wi1 * (ui * ui + vi * vi + wi * wi) - pinfl);
          double norm = 1. / v[n], p = ps[n];
          double V2 = v[ldv + n], V3 = v[2 * ldv + n], V4 = v[3 * ldv + n], V5 = v[4 * ldv + n];
          double fxx = V2 * (V2 * norm) + p, fxy = V3 * (V2 * norm), fxz = V4 * (V2 * norm);
          double fyy = V3 * (V3 * norm) + p, fyz = V4 * (V3 * norm), fzz = V4 * (V4 * norm) + p;
          double fex = (V5 + p - pinfl) * V2 * norm, fey = (V5 + p - pinfl) * V3 * norm;
          double fez = (V5 + p - pinfl) * V4 * norm;
          df[0][m] = wi1 * (ui * ui) + pres - fxx;
          df[1][m] = wi1 * ui * vi - fxy;
          df[2][m] = wi1 * ui * wi - fxz;
          df[3][m] = wi1 * ui * (wi5 + pres - pinfl) - fex;
          df[4][m] = wi1 * (vi * vi) + pres - fyy;
          df[5][m] = wi1 * vi * wi - fyz;
          df[6][m] = wi1 * vi * (wi5 + pres - pinfl) - fey;
          df[7][m] = wi1 * (wi * wi) + pres - fzz;
          df[8][m] = wi1 * (wi5 + pres - pinfl) - fez;
        }
  }
  for (int k = k1; k <= k2 - 1; k++)
    for (int j = j1; j <= j2 - 1; j++)
      for (int i = i1; i <= i2 - 1; i++) {
        long n = IDX(b, i, j, k);
        for (int e = 0; e < nvar; e++) v[e * ldv + n] = v[e * ldv + n] + d[e * ldv + n];
      }
  for (int c = 0; c < 9; c++) free(df[c]);
  for (int c = 0; c < 3; c++) free(coefe[c]);
  for (int c = 0; c < 5; c++) free(d2w[c]);
  free(diag);
}

/* ------------------------------------------------------------------------ */
/* residu.f90:93-187 for each block, then the aggregate of rscpsv.f90:
 * 168-192.  r = (v - u) / dt on the real cells for the first neq variables.
 * out: (nb + 1) records of neq x 7 doubles (L1, L2, Linf, block, i, j, k),
 * as the device entry point writes them; the per-block sums run in the
 * reference's loop order, the location is the last cell attaining Linf
 * (residu's s = 0.5 (sign(1, |r| - dumax) + 1)). */
void orc_residu(const orc_block *blocks, int nb, int neq, const double *v, const double *u, long ldv,
                const double *dt, double *out) {
  double d1g[7] = {0}, d2g[7] = {0}, dmg[7] = {0};
  long npts = 0;
  int gl[7] = {0};
  for (int l = 0; l < nb; l++) {
    const orc_block *b = &blocks[l];
    long nmax = (long)(b->ii2 - b->ii1) * (b->jj2 - b->jj1) * (b->kk2 - b->kk1);
    double coef = 1. / (double)nmax;
    double d1[7] = {0}, d2[7] = {0}, dm[7] = {0};
    int im[7] = {0}, jm[7] = {0}, km[7] = {0};
    for (int k = b->kk1; k <= b->kk2 - 1; k++)
      for (int j = b->jj1; j <= b->jj2 - 1; j++)
        for (int i = b->ii1; i <= b->ii2 - 1; i++) {
          long n = IDX(b, i, j, k);
          for (int e = 0; e < neq; e++) {
            double r = (v[e * ldv + n] - u[e * ldv + n]) / dt[n];
            d1[e] = d1[e] + fabs(r);
            d2[e] = d2[e] + r * r;
            dm[e] = fmax(dm[e], fabs(r));
            double s = 0.5 * (copysign(1.0, fabs(r) - dm[e]) + 1.);
            im[e] = (int)lround((1. - s) * im[e] + s * i);
            jm[e] = (int)lround((1. - s) * jm[e] + s * j);
            km[e] = (int)lround((1. - s) * km[e] + s * k);
          }
        }
    for (int e = 0; e < neq; e++) {
      double *o = out + ((long)l * neq + e) * 7;
      o[0] = coef * d1[e];
      o[1] = sqrt(coef * d2[e]);
      o[2] = dm[e];
      o[3] = l;
      o[4] = im[e];
      o[5] = jm[e];
      o[6] = km[e];
      d1g[e] = d1g[e] + nmax * o[0];
      d2g[e] = d2g[e] + nmax * (o[1] * o[1]);
      if (l == 0 || fabs(o[2]) >= dmg[e]) gl[e] = l;
      dmg[e] = fmax(dmg[e], fabs(o[2]));
    }
    npts += nmax;
  }
  for (int e = 0; e < neq; e++) {
    double *o = out + ((long)nb * neq + e) * 7;
    double *w = out + ((long)gl[e] * neq + e) * 7;
    o[0] = d1g[e] / npts;
    o[1] = sqrt(d2g[e] / npts);
    o[2] = dmg[e];
    o[3] = w[3];
    o[4] = w[4];
    o[5] = w[5];
    o[6] = w[6];
  }
}
