// Viscous stresses and heat flux (zvisqc -> zgrad2 + zfluto) on gfx950.
// The velocities (rho u)/rho of every storage cell are formed once (as
// zgrad2.f90:161-180 forms vx, vy, vz); then one thread per real cell
// gathers the Green-Gauss sums of its six faces (instead of the reference's
// +/- face scatter), scales them by 1/(2 vol) and forms tau and q in
// registers: the twelve gradient arrays never reach HBM.
#include "common.hpp"

using namespace codens;

__global__ void __launch_bounds__(256) k_vel(long npc, long nnn, const double *__restrict__ v, long ldv,
                                             double *__restrict__ uvw) {
  long m = blockIdx.x * 256L + threadIdx.x;
  if (m >= nnn) return;
  long n = npc + m;
  double r = v[n];
  uvw[m] = v[ldv + n] / r;
  uvw[nnn + m] = v[2 * ldv + n] / r;
  uvw[2 * nnn + m] = v[3 * ldv + n] / r;
}

__global__ void __launch_bounds__(TX *TY) k_zvisqc(Blk b, Phys ph, const double *__restrict__ sn,
                                                   const double *__restrict__ vol,
                                                   const double *__restrict__ uvw,
                                                   const double *__restrict__ temp,
                                                   const double *__restrict__ mu,
                                                   const double *__restrict__ mut, double *__restrict__ tq,
                                                   long ldv) {
  const dim3 xb = xcd_block();
  int c[3];
  c[0] = b.i1 + xb.x * TX + threadIdx.x;
  c[1] = b.j1 + xb.y * TY + threadIdx.y;
  c[2] = b.k1 + xb.z;
  if (c[0] >= b.i2 || c[1] >= b.j2) return;
  long n = b.idx(c[0], c[1], c[2]), m = n - b.npc;
  const long N = b.nnn;
  double qc[4];
  qc[0] = uvw[m];
  qc[1] = uvw[N + m];
  qc[2] = uvw[2 * N + m];
  qc[3] = temp[n];
  double G[12];
#pragma unroll
  for (int q = 0; q < 12; q++) G[q] = 0.;
#pragma unroll
  for (int dd = 0; dd < 3; dd++) {
    int d = 2 - dd;  // k faces, then j, then i (zgrad2.f90:181-457)
    long inc = b.inc(d);
    int a = c[d], a1 = b.lo(d), a2 = b.hi(d);
    // low face (S at m): subtract
    {
      long nn = n - inc, mm = m - inc;
      double qn[4] = {uvw[mm], uvw[N + mm], uvw[2 * N + mm], temp[nn]};
      bool bnd = (a == a1);
#pragma unroll
      for (int cc = 0; cc < 3; cc++) {
        double s = snv(sn, b.nnn, m, d, cc);
#pragma unroll
        for (int q = 0; q < 4; q++) {
          double t = bnd ? s * 2 * qn[q] : (qc[q] + qn[q]) * s;
          G[3 * q + cc] -= t;
        }
      }
    }
    // high face (S at m+inc): add
    {
      long nn = n + inc, mm = m + inc;
      double qn[4] = {uvw[mm], uvw[N + mm], uvw[2 * N + mm], temp[nn]};
      bool bnd = (a + 1 == a2);
#pragma unroll
      for (int cc = 0; cc < 3; cc++) {
        double s = snv(sn, b.nnn, m + inc, d, cc);
#pragma unroll
        for (int q = 0; q < 4; q++) {
          double t = bnd ? s * 2 * qn[q] : (qn[q] + qc[q]) * s;
          G[3 * q + cc] += t;
        }
      }
    }
  }
  // zgrad2.f90:463-485 (guard for non-positive volumes)
  double vn = vol[n];
  double ts = copysign(0.5, -vn);
  double vols = (0.5 + ts) * 0.00001 + (0.5 - ts) * vn;
  double c0 = .5 / (vols + DBL_MIN);
#pragma unroll
  for (int q = 0; q < 12; q++) G[q] *= c0;
  // zfluto.f90:104-121
  double mt = mu[n] + mut[n];
  const double ds3 = 2. / 3.;
  double xx = G[0], xy = G[1], xz = G[2], yx = G[3], yy = G[4], yz = G[5], zx = G[6], zy = G[7], zz = G[8];
  tq[n] = ds3 * mt * (2 * xx - yy - zz);
  tq[ldv + n] = mt * (xy + yx);
  tq[2 * ldv + n] = mt * (xz + zx);
  tq[3 * ldv + n] = ds3 * mt * (2 * yy - xx - zz);
  tq[4 * ldv + n] = mt * (yz + zy);
  tq[5 * ldv + n] = ds3 * mt * (2 * zz - xx - yy);
  double kq = ph.cp * (mu[n] / ph.pr + mut[n] / ph.prt);
  tq[6 * ldv + n] = kq * G[9];
  tq[7 * ldv + n] = kq * G[10];
  tq[8 * ldv + n] = kq * G[11];
}

extern "C" int codens_zvisqc(const codens_block *cb, const codens_phys *cp, const double *sn, const double *vol,
                             const double *v, int64_t ldv, const double *temp, const double *mu,
                             const double *mut, double *tq, void *stream) {
  int rc = codens_check_block(cb, "codens_zvisqc");
  if (rc) return rc;
  if (!cp || !sn || !vol || !v || !temp || !mu || !mut || !tq) {
    codens_set_error("codens_zvisqc: null argument");
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  double *uvw = (double *)codens_workspace(sizeof(double) * 3 * b.nnn, 4);
  if (!uvw) {
    codens_set_error("codens_zvisqc: workspace allocation failed");
    return CODENS_EHIP;
  }
  {
    CodensTimed _tm("zvisqc", (hipStream_t)stream);
    hipLaunchKernelGGL(k_vel, dim3((unsigned)((b.nnn + 255) / 256)), dim3(256), 0, (hipStream_t)stream, (long)b.npc,
                       (long)b.nnn, v, (long)ldv, uvw);
    hipLaunchKernelGGL(k_zvisqc, cell_grid(b), dim3(TX, TY), 0, (hipStream_t)stream, b, make_phys(cp), sn, vol,
                       (const double *)uvw, temp, mu, mut, tq, (long)ldv);
  }
  return codens_check_launch("codens_zvisqc");
}
