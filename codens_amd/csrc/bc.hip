// Boundary ghost filling on gfx950, one thread per boundary face cell.
// Face lists come from the block topology (initis.f90, initbs.f90): ghost
// cell, adjacent real cell, and for coincident faces the neighbour's real cell.
#include "common.hpp"

// No FMA contraction: ghost states round exactly as the reference's (they
// feed the chained slope ratios of sch_roe).
#pragma clang fp contract(off)

using namespace codens;

constexpr int BT = 256;

// rbvc.f90:84-117 without periodic rotation
__global__ void __launch_bounds__(BT) k_bc_coincident(long nf, const int64_t *__restrict__ g,
                                                      const int64_t *__restrict__ in,
                                                      const int64_t *__restrict__ rm, double *v, long ldv,
                                                      double *ps, double *temp, double *cson) {
  long f = blockIdx.x * (long)BT + threadIdx.x;
  if (f >= nf) return;
  long nd = g[f], ni = in[f], nr = rm[f];
#pragma unroll
  for (int e = 0; e < 5; e++) v[e * ldv + nd] = 0.5 * (v[e * ldv + ni] + v[e * ldv + nr]);
  ps[nd] = 0.5 * (ps[ni] + ps[nr]);
  temp[nd] = 0.5 * (temp[ni] + temp[nr]);
  cson[nd] = 0.5 * (cson[ni] + cson[nr]);
}

// clpara.f90:44-58
__global__ void __launch_bounds__(BT) k_bc_wall(long nf, const int64_t *__restrict__ g,
                                                const int64_t *__restrict__ in, double *v, long ldv, double *ps,
                                                double *temp, double *cson) {
  long f = blockIdx.x * (long)BT + threadIdx.x;
  if (f >= nf) return;
  long nc = g[f], ni = in[f];
  double r = v[ni], a = v[ldv + ni], bb = v[2 * ldv + ni], c = v[3 * ldv + ni];
  double rhoe1 = v[4 * ldv + ni] - 0.5 * (a * a + bb * bb + c * c) / r;
  v[nc] = r;
  v[ldv + nc] = 0.;
  v[2 * ldv + nc] = 0.;
  v[3 * ldv + nc] = 0.;
  v[4 * ldv + nc] = rhoe1;
  ps[nc] = ps[ni];
  temp[nc] = temp[ni];
  cson[nc] = cson[ni];
}

// rbve.f90 then clnrd.f90:60-134 (characteristic far field)
__global__ void __launch_bounds__(BT) k_bc_farfield(long nf, const int64_t *__restrict__ g,
                                                    const int64_t *__restrict__ in,
                                                    const double *__restrict__ nrm, double w00, double w01,
                                                    double w02, double w03, double w04, Phys ph, double *v,
                                                    long ldv, double *ps, double *temp, double *cson) {
  long f = blockIdx.x * (long)BT + threadIdx.x;
  if (f >= nf) return;
  const double gam = ph.gam, gam1 = ph.gam1, pinfl = ph.pinfl;
  long nl = g[f], ni = in[f];
  double nx = nrm[f], ny = nrm[nf + f], nz = nrm[2 * nf + f];
  double s0 = v[ni], s1 = v[ldv + ni], s2 = v[2 * ldv + ni], s3 = v[3 * ldv + ni], s4 = v[4 * ldv + ni];
  double ro0 = s0;
  double ps0 = gam1 * (s4 - 0.5 * (s1 * s1 + s2 * s2 + s3 * s3) / s0 - pinfl);
  double roc0 = sqrt(gam * ps0 * ro0);
  double roqn0 = s1 * nx + s2 * ny + s3 * nz;
  double epsm = .5 + copysign(.5, roc0 - roqn0);
  double eps0 = .5 + copysign(.5, -roqn0);
  double epsp = .5 + copysign(.5, -roc0 - roqn0);
  double rod = w00;
  double qxd = w01 / rod, qyd = w02 / rod, qzd = w03 / rod;
  double pd = gam1 * (w04 - .5 * rod * (qxd * qxd + qyd * qyd + qzd * qzd) - pinfl);
  double qnd = qxd * nx + qyd * ny + qzd * nz;
  double qtxd = qxd - qnd * nx, qtyd = qyd - qnd * ny, qtzd = qzd - qnd * nz;
  double ros = s0;
  double qxs = s1 / ros, qys = s2 / ros, qzs = s3 / ros;
  double p_s = gam1 * (s4 - .5 * ros * (qxs * qxs + qys * qys + qzs * qzs) - pinfl);
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
  double m1 = ro * (qtx + qn * nx), m2 = ro * (qty + qn * ny), m3 = ro * (qtz + qn * nz);
  v[nl] = ro;
  v[ldv + nl] = m1;
  v[2 * ldv + nl] = m2;
  v[3 * ldv + nl] = m3;
  v[4 * ldv + nl] = p / gam1 + pinfl + 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / ro;
  double t = gam * p / ro;
  ps[nl] = p;
  temp[nl] = t;
  cson[nl] = sqrt(t);
}

// rfve.f90:24-41: ghost = adjacent real cell for v (5 planes), p, T and c
// (atsch_num.f90:450-458 runs it on the non-coincident frontiers in the
// first cycle of a computation, before rbord)
__global__ void __launch_bounds__(BT) k_rfve(long nf, const int64_t *__restrict__ g,
                                             const int64_t *__restrict__ in, double *v, long ldv, double *ps,
                                             double *temp, double *cson) {
  long f = blockIdx.x * (long)BT + threadIdx.x;
  if (f >= nf) return;
  long nd = g[f], ni = in[f];
#pragma unroll
  for (int e = 0; e < 5; e++) v[e * ldv + nd] = v[e * ldv + ni];
  ps[nd] = ps[ni];
  temp[nd] = temp[ni];
  cson[nd] = cson[ni];
}

// clsym.f90:56-79: symmetry plane.  The ghost takes the adjacent cell's
// density and energy and its velocity without the normal component; the
// ghost's p, T and c are left as they are (clsym has no such arguments).
__global__ void __launch_bounds__(BT) k_bc_sym(long nf, const int64_t *__restrict__ g,
                                               const int64_t *__restrict__ in, const double *__restrict__ nrm,
                                               double *v, long ldv) {
  long f = blockIdx.x * (long)BT + threadIdx.x;
  if (f >= nf) return;
  long nl = g[f], ni = in[f];
  double nx = nrm[f], ny = nrm[nf + f], nz = nrm[2 * nf + f];
  double ro1 = v[ni];
  double qx1 = v[ldv + ni] / ro1, qy1 = v[2 * ldv + ni] / ro1, qz1 = v[3 * ldv + ni] / ro1;
  double roe1 = v[4 * ldv + ni];
  double qn1 = qx1 * nx + qy1 * ny + qz1 * nz;
  double qtx = qx1 - qn1 * nx, qty = qy1 - qn1 * ny, qtz = qz1 - qn1 * nz;
  v[nl] = ro1;
  v[ldv + nl] = ro1 * qtx;
  v[2 * ldv + nl] = ro1 * qty;
  v[3 * ldv + nl] = ro1 * qtz;
  v[4 * ldv + nl] = roe1;
}

// rbte.f90:46-59 (9 planes of tau, q) and met_rbse.f90:26-38 (6 planes of
// the turbulent diffusive flux densities): ghost = adjacent real cell
template <int NP>
__global__ void __launch_bounds__(BT) k_bc_planes_copy(long nf, const int64_t *__restrict__ g,
                                                       const int64_t *__restrict__ in, double *tq, long ldv) {
  long f = blockIdx.x * (long)BT + threadIdx.x;
  if (f >= nf) return;
  long nd = g[f], ni = in[f];
  double x[NP];  // every load before the first store (a ghost is never a source)
#pragma unroll
  for (int q = 0; q < NP; q++) x[q] = tq[q * ldv + ni];
#pragma unroll
  for (int q = 0; q < NP; q++) tq[q * ldv + nd] = x[q];
}

// rbtc.f90:104-117 and met_rbsc.f90:84-95 without periodic rotation:
// ghost = average of both real cells
template <int NP>
__global__ void __launch_bounds__(BT) k_bc_planes_coincident(long nf, const int64_t *__restrict__ g,
                                                             const int64_t *__restrict__ in,
                                                             const int64_t *__restrict__ rm, double *tq, long ldv) {
  long f = blockIdx.x * (long)BT + threadIdx.x;
  if (f >= nf) return;
  long nd = g[f], ni = in[f], nr = rm[f];
  double x[NP];  // every load before the first store (a ghost is never a source)
#pragma unroll
  for (int q = 0; q < NP; q++) x[q] = 0.5 * (tq[q * ldv + ni] + tq[q * ldv + nr]);
#pragma unroll
  for (int q = 0; q < NP; q++) tq[q * ldv + nd] = x[q];
}

static inline dim3 fgrid(int64_t nf) { return dim3((unsigned)((nf + BT - 1) / BT)); }

extern "C" int codens_bc_coincident(int64_t nf, const int64_t *g, const int64_t *in, const int64_t *rm, double *v,
                                    int64_t ldv, double *ps, double *temp, double *cson, void *stream) {
  if (nf == 0) return CODENS_OK;
  if (nf < 0 || !g || !in || !rm || !v || !ps || !temp || !cson) {
    codens_set_error("codens_bc_coincident: invalid argument");
    return CODENS_EINVAL;
  }
  hipLaunchKernelGGL(k_bc_coincident, fgrid(nf), dim3(BT), 0, (hipStream_t)stream, (long)nf, g, in, rm, v,
                     (long)ldv, ps, temp, cson);
  return codens_check_launch("codens_bc_coincident");
}

extern "C" int codens_bc_wall(int64_t nf, const int64_t *g, const int64_t *in, double *v, int64_t ldv, double *ps,
                              double *temp, double *cson, void *stream) {
  if (nf == 0) return CODENS_OK;
  if (nf < 0 || !g || !in || !v || !ps || !temp || !cson) {
    codens_set_error("codens_bc_wall: invalid argument");
    return CODENS_EINVAL;
  }
  hipLaunchKernelGGL(k_bc_wall, fgrid(nf), dim3(BT), 0, (hipStream_t)stream, (long)nf, g, in, v, (long)ldv, ps,
                     temp, cson);
  return codens_check_launch("codens_bc_wall");
}

extern "C" int codens_bc_farfield(int64_t nf, const int64_t *g, const int64_t *in, const double *nrm,
                                  const double *w0, const codens_phys *cp, double *v, int64_t ldv, double *ps,
                                  double *temp, double *cson, void *stream) {
  if (nf == 0) return CODENS_OK;
  if (nf < 0 || !g || !in || !nrm || !w0 || !cp || !v || !ps || !temp || !cson) {
    codens_set_error("codens_bc_farfield: invalid argument");
    return CODENS_EINVAL;
  }
  hipLaunchKernelGGL(k_bc_farfield, fgrid(nf), dim3(BT), 0, (hipStream_t)stream, (long)nf, g, in, nrm, w0[0],
                     w0[1], w0[2], w0[3], w0[4], make_phys(cp), v, (long)ldv, ps, temp, cson);
  return codens_check_launch("codens_bc_farfield");
}

extern "C" int codens_rfve(int64_t nf, const int64_t *g, const int64_t *in, double *v, int64_t ldv, double *ps,
                           double *temp, double *cson, void *stream) {
  if (nf == 0) return CODENS_OK;
  if (nf < 0 || !g || !in || !v || !ps || !temp || !cson) {
    codens_set_error("codens_rfve: invalid argument");
    return CODENS_EINVAL;
  }
  hipLaunchKernelGGL(k_rfve, fgrid(nf), dim3(BT), 0, (hipStream_t)stream, (long)nf, g, in, v, (long)ldv, ps, temp,
                     cson);
  return codens_check_launch("codens_rfve");
}

extern "C" int codens_bc_sym(int64_t nf, const int64_t *g, const int64_t *in, const double *nrm, double *v,
                             int64_t ldv, void *stream) {
  if (nf == 0) return CODENS_OK;
  if (nf < 0 || !g || !in || !nrm || !v) {
    codens_set_error("codens_bc_sym: invalid argument");
    return CODENS_EINVAL;
  }
  hipLaunchKernelGGL(k_bc_sym, fgrid(nf), dim3(BT), 0, (hipStream_t)stream, (long)nf, g, in, nrm, v, (long)ldv);
  return codens_check_launch("codens_bc_sym");
}

// utnrd.f90:59-105 in a non-rotating frame (omg = 0): the far-field state
// from the Mach number, the incidence angles (degrees) and the stagnation
// state, with the reference's libm calls and operation order.  Host only.
extern "C" int codens_utnrd(double rmach, double alpha0, double beta0, double roa1, double aa1, double pa1,
                            const codens_phys *cp, double *w0) {
  if (!cp || !w0 || !(cp->gam > 1.0)) {
    codens_set_error("codens_utnrd: invalid argument");
    return CODENS_EINVAL;
  }
  const double gam = cp->gam, gam1 = cp->gam - 1.0, gam2 = 0.5 * gam1, gam4 = 1. / gam1;
  double degrad = atan2(1.0, 0.0) / 90.0;
  double f = 1. + gam2 * (rmach * rmach);
  double ro = roa1 / pow(f, gam4);
  double a = aa1 / pow(f, .5);
  double p = pa1 / pow(f, gam / gam1);
  double q = rmach * a;
  double alphar = alpha0 * degrad, betar = beta0 * degrad;
  w0[0] = ro;
  w0[1] = ro * q * cos(alphar) * cos(betar);
  w0[2] = -ro * q * sin(betar);
  w0[3] = ro * q * sin(alphar) * cos(betar);
  w0[4] = p / gam1 + cp->pinfl + 0.5 * ro * (q * q);
  return CODENS_OK;
}

extern "C" int codens_bc_tq_copy(int64_t nf, const int64_t *g, const int64_t *in, double *tq, int64_t ldv,
                                 void *stream) {
  if (nf == 0) return CODENS_OK;
  if (nf < 0 || !g || !in || !tq) {
    codens_set_error("codens_bc_tq_copy: invalid argument");
    return CODENS_EINVAL;
  }
  hipLaunchKernelGGL(k_bc_planes_copy<9>, fgrid(nf), dim3(BT), 0, (hipStream_t)stream, (long)nf, g, in, tq, (long)ldv);
  return codens_check_launch("codens_bc_tq_copy");
}

extern "C" int codens_bc_tq_coincident(int64_t nf, const int64_t *g, const int64_t *in, const int64_t *rm,
                                       double *tq, int64_t ldv, void *stream) {
  if (nf == 0) return CODENS_OK;
  if (nf < 0 || !g || !in || !rm || !tq) {
    codens_set_error("codens_bc_tq_coincident: invalid argument");
    return CODENS_EINVAL;
  }
  hipLaunchKernelGGL(k_bc_planes_coincident<9>, fgrid(nf), dim3(BT), 0, (hipStream_t)stream, (long)nf, g, in, rm, tq,
                     (long)ldv);
  return codens_check_launch("codens_bc_tq_coincident");
}

extern "C" int codens_met_rbse(int64_t nf, const int64_t *g, const int64_t *in, double *fd, int64_t ldv,
                               void *stream) {
  if (nf == 0) return CODENS_OK;
  if (nf < 0 || !g || !in || !fd) {
    codens_set_error("codens_met_rbse: invalid argument");
    return CODENS_EINVAL;
  }
  hipLaunchKernelGGL(k_bc_planes_copy<6>, fgrid(nf), dim3(BT), 0, (hipStream_t)stream, (long)nf, g, in, fd,
                     (long)ldv);
  return codens_check_launch("codens_met_rbse");
}

extern "C" int codens_met_rbsc(int64_t nf, const int64_t *g, const int64_t *in, const int64_t *rm, double *fd,
                               int64_t ldv, void *stream) {
  if (nf == 0) return CODENS_OK;
  if (nf < 0 || !g || !in || !rm || !fd) {
    codens_set_error("codens_met_rbsc: invalid argument");
    return CODENS_EINVAL;
  }
  hipLaunchKernelGGL(k_bc_planes_coincident<6>, fgrid(nf), dim3(BT), 0, (hipStream_t)stream, (long)nf, g, in, rm,
                     fd, (long)ldv);
  return codens_check_launch("codens_met_rbsc");
}

// met_bord.f90:4 with met_rbvc.f90:30-65 and met_rbve.f90:26-36: the two
// transported turbulence variables (planes 5, 6) of the ghosts of one
// frontier list, by kind (codens_met_bord's codes).  A coincident ghost
// reads two real cells, which no ghost list writes, so the reference's
// buffer of remote values is not needed.
__global__ void __launch_bounds__(BT) k_met_bord(int kind, long nf, const int64_t *__restrict__ g,
                                                 const int64_t *__restrict__ in, const int64_t *__restrict__ rm,
                                                 double a, double b, double *v, long ldv) {
  long f = blockIdx.x * (long)BT + threadIdx.x;
  if (f >= nf) return;
  double *k = v + 5 * ldv, *w = v + 6 * ldv;
  long nd = g[f], ni = in[f];
  switch (kind) {
    case 0: {
      long nr = rm[f];
      k[nd] = 0.5 * (k[ni] + k[nr]);
      w[nd] = 0.5 * (w[ni] + w[nr]);
      break;
    }
    case 1: k[nd] = 0.; w[nd] = 0.; break;
    case 2: k[nd] = 0.; w[nd] = w[ni]; break;
    case 3: k[nd] = a; w[nd] = b; break;
    default: k[nd] = k[ni]; w[nd] = w[ni]; break;
  }
}

extern "C" int codens_met_bord(int32_t kind, int64_t nf, const int64_t *g, const int64_t *in, const int64_t *rm,
                               double rokinf, double roeinf, double *v, int64_t ldv, void *stream) {
  if (nf == 0) return CODENS_OK;
  if (nf < 0 || kind < 0 || kind > 4 || !g || (kind != 1 && kind != 3 && !in) || (kind == 0 && !rm) || !v ||
      ldv <= 0) {
    codens_set_error("codens_met_bord: invalid argument (kind %d)", kind);
    return CODENS_EINVAL;
  }
  hipLaunchKernelGGL(k_met_bord, fgrid(nf), dim3(BT), 0, (hipStream_t)stream, (int)kind, (long)nf, g,
                     in ? in : g, rm ? rm : g, rokinf, roeinf, v, (long)ldv);
  return codens_check_launch("codens_met_bord");
}

// met_parko.f90:4, one thread per wall point: the wall cell's rho omega =
// 6 mu / (beta1 d^2); along the wall normal, the next min(10, m3tot) cells
// whose wall point lies on this wall are blended with it over cells 2..4
// (imxclko = 3) and floored by it; the ghost takes the wall cell's value.
__global__ void __launch_bounds__(BT) k_met_parko(long nl, const codens_wall_line *__restrict__ ln,
                                                  const int64_t *__restrict__ ghost,
                                                  const int32_t *__restrict__ mnpar,
                                                  const double *__restrict__ mu, const double *__restrict__ dist,
                                                  double beta1, double *v, long ldv) {
  long p = blockIdx.x * (long)BT + threadIdx.x;
  if (p >= nl) return;
  double *w = v + 6 * ldv;
  const codens_wall_line L = ln[p];
  const int m3clpar = 10, imxclko = 3;
  const long st = (long)L.isens3 * L.dm3;
  const long nmin = L.nmin, nminp = nmin + st, nmaxp = nmin + st * (imxclko + 1);
  const int m3tot = abs(L.m3max - L.m3min) + 1;
  const int nmore = min(m3tot, m3clpar);
  double dn = dist[nmin];
  w[nmin] = 6. * mu[nmin] / (beta1 * (dn * dn));
  for (int s = 1; s <= nmore; s++) {
    long n = nmin + s * st;
    int npar = mnpar[n];
    if (npar > L.mpn0 && npar <= L.mpn1) {
      double d = dist[n];
      double rom = 6. * mu[n] / (beta1 * (d * d));
      if ((n - nmaxp) * (n - nminp) <= 0) {
        double pond = (d - dist[nminp]) / (dist[nmaxp] - dist[nminp]);
        w[n] = rom * (1. - pond) + pond * w[n];
      }
      w[n] = fmax(w[n], rom);
    }
  }
  w[ghost[p]] = w[nmin];
}

extern "C" int codens_met_parko(int32_t nfaces, const int64_t *face_start, const codens_wall_line *lines,
                                const int64_t *ghost, const int32_t *mnpar, const double *mu, const double *dist,
                                double beta1, double *v, int64_t ldv, void *stream) {
  if (nfaces == 0) return CODENS_OK;
  if (nfaces < 0 || !face_start || !lines || !ghost || !mnpar || !mu || !dist || !v || ldv <= 0) {
    codens_set_error("codens_met_parko: invalid argument");
    return CODENS_EINVAL;
  }
  // walls in order (a cell reached from two walls gets both, as the
  // reference's loop over walls gives it); the points of one wall are
  // distinct cell columns
  for (int f = 0; f < nfaces; f++) {
    long a = face_start[f], e = face_start[f + 1];
    if (e < a) {
      codens_set_error("codens_met_parko: face_start is not increasing");
      return CODENS_EINVAL;
    }
    if (e > a)
      hipLaunchKernelGGL(k_met_parko, fgrid(e - a), dim3(BT), 0, (hipStream_t)stream, e - a, lines + a, ghost + a,
                         mnpar, mu, dist, beta1, v, (long)ldv);
  }
  return codens_check_launch("codens_met_parko");
}
