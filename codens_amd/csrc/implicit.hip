// Matrix-free point-Jacobi implicit phase (implimf_3d) on gfx950.
//
// implimf_3d.f90 keeps nine flux-increment arrays dF = F(v+d) - F(v) between
// sub-iterations.  Here every sub-iteration is one kernel that re-derives dF
// of a cell and of its six neighbours from v and the previous increment d:
// per sub-iteration a cell moves u, v, d_old, sn, diag, coefe and d_new through
// HBM (33 doubles) instead of also streaming nine dF arrays in and out.
// Because every read of d and dF in a sub-iteration sees the previous
// sub-iteration's values (implimf_3d.f90:230-431), the cells are independent
// and any traversal order gives the reference's result.
#include "common.hpp"

using namespace codens;

// Increment buffer view: element (e, n) at p[e*ld + n - off].
struct DView {
  double *p;
  long ld, off;
  __device__ inline double &at(int e, long n) const { return p[e * ld + n - off]; }
};

struct DFlux {  // mass-flux increment (d of rho u, rho v, rho w) and dF
  double mx, my, mz, xx, xy, xz, yy, yz, zz, ex, ey, ez;
};

// implimf_3d.f90:434-465 for one cell: dF = F(v + d) - F(v), p(v) from the EOS
// exactly as zpres.f90:72-73 computes it.
__device__ inline DFlux dflux(const Phys &ph, const double V[5], const double D[5]) {
  const double gam1 = ph.gam1, pinfl = ph.pinfl;
  double wi1 = V[0] + D[0], wi2 = V[1] + D[1], wi3 = V[2] + D[2], wi4 = V[3] + D[3], wi5 = V[4] + D[4];
  double ui = wi2 / wi1, vi = wi3 / wi1, wi = wi4 / wi1;
  double pres = gam1 * (wi5 - .5 * wi1 * (ui * ui + vi * vi + wi * wi) - pinfl);
  double rhoe = V[4] - 0.5 * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3]) / V[0];
  double p = gam1 * (rhoe - pinfl);
  double a = V[1] / V[0], bb = V[2] / V[0], c = V[3] / V[0];
  double fxx = V[1] * a + p, fxy = V[2] * a, fxz = V[3] * a;
  double fyy = V[2] * bb + p, fyz = V[3] * bb, fzz = V[3] * c + p;
  double e5 = V[4] + p - pinfl;
  double fex = e5 * V[1] / V[0], fey = e5 * V[2] / V[0], fez = e5 * V[3] / V[0];
  DFlux r;
  r.mx = D[1]; r.my = D[2]; r.mz = D[3];
  r.xx = wi1 * (ui * ui) + pres - fxx;
  r.xy = wi1 * ui * vi - fxy;
  r.xz = wi1 * ui * wi - fxz;
  r.ex = ui * (wi5 + pres - pinfl) - fex;
  r.yy = wi1 * (vi * vi) + pres - fyy;
  r.yz = wi1 * vi * wi - fyz;
  r.ey = vi * (wi5 + pres - pinfl) - fey;
  r.zz = wi1 * (wi * wi) + pres - fzz;
  r.ez = wi * (wi5 + pres - pinfl) - fez;
  return r;
}

// tn of a face (implimf_3d.f90:276-292): half-sum of both cells' increments . S
__device__ inline void face_tn(const DFlux &A, const DFlux &B, double s1, double s2, double s3, double tn[5]) {
  tn[0] = 0.5 * (A.mx + B.mx) * s1 + 0.5 * (A.my + B.my) * s2 + 0.5 * (A.mz + B.mz) * s3;
  tn[1] = 0.5 * (A.xx + B.xx) * s1 + 0.5 * (A.xy + B.xy) * s2 + 0.5 * (A.xz + B.xz) * s3;
  tn[2] = 0.5 * (A.xy + B.xy) * s1 + 0.5 * (A.yy + B.yy) * s2 + 0.5 * (A.yz + B.yz) * s3;
  tn[3] = 0.5 * (A.xz + B.xz) * s1 + 0.5 * (A.yz + B.yz) * s2 + 0.5 * (A.zz + B.zz) * s3;
  tn[4] = 0.5 * (A.ex + B.ex) * s1 + 0.5 * (A.ey + B.ey) * s2 + 0.5 * (A.ez + B.ez) * s3;
}

#define CELL3(b)                                     \
  int c[3];                                          \
  c[0] = (b).i1 + blockIdx.x * TX + threadIdx.x;     \
  c[1] = (b).j1 + blockIdx.y * TY + threadIdx.y;     \
  c[2] = (b).k1 + blockIdx.z;                        \
  if (c[0] >= (b).i2 || c[1] >= (b).j2) return;

__device__ inline double rv_of(const Phys &ph, const double *mu, const double *mut, const double *v, long n) {
  return ph.gam * (mu[n] / ph.pr + mut[n] / ph.prt) / v[n];
}

// implimf_3d.f90:113-224: spectral-radius coefficients of the faces and the
// diagonal.  coefe[d] holds the low face of every real cell.
__global__ void __launch_bounds__(TX *TY) k_impl_coef(Blk b, Phys ph, const double *__restrict__ v, long ldv,
                                                      const double *__restrict__ mu,
                                                      const double *__restrict__ mut,
                                                      const double *__restrict__ dt,
                                                      const double *__restrict__ vol,
                                                      const double *__restrict__ cson,
                                                      const double *__restrict__ sn, double dtpas, int kfmg,
                                                      double *__restrict__ coefe, double *__restrict__ diag) {
  CELL3(b)
  long n = b.idx(c[0], c[1], c[2]), m = n - b.npc;
  double rvc = rv_of(ph, mu, mut, v, n);
  double dg = vol[n] / dt[n];
  double u0 = v[ldv + n] / v[n], v0 = v[2 * ldv + n] / v[n], w0 = v[3 * ldv + n] / v[n];
  double cf[6];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    long inc = b.inc(d);
    int a = c[d], a1 = b.lo(d), a2 = b.hi(d);
#pragma unroll
    for (int side = 0; side < 2; side++) {
      long nn = side == 0 ? n - inc : n + inc;
      long mf = side == 0 ? m : m + inc;
      bool nreal = side == 0 ? (a > a1) : (a + 1 < a2);
      double s1 = snv(sn, b.nnn, mf, d, 0), s2 = snv(sn, b.nnn, mf, d, 1), s3 = snv(sn, b.nnn, mf, d, 2);
      double cnds = s1 * s1 + s2 * s2 + s3 * s3;
      double un = v[ldv + nn] / v[nn], vn_ = v[2 * ldv + nn] / v[nn], wn = v[3 * ldv + nn] / v[nn];
      double rvn = nreal ? rv_of(ph, mu, mut, v, nn) : 0.;
      double uu, vv, ww, cc, rsum, vsum;
      if (side == 0) {  // face m: cells n (own) and n-inc
        uu = 0.5 * (u0 + un); vv = 0.5 * (v0 + vn_); ww = 0.5 * (w0 + wn);
        cc = 0.5 * (cson[n] + cson[nn]); rsum = rvc + rvn; vsum = vol[n] + vol[nn];
      } else {          // face m+inc: cells n+inc and n (own)
        uu = 0.5 * (un + u0); vv = 0.5 * (vn_ + v0); ww = 0.5 * (wn + w0);
        cc = 0.5 * (cson[nn] + cson[n]); rsum = rvn + rvc; vsum = vol[nn] + vol[n];
      }
      double vnf = uu * s1 + vv * s2 + ww * s3;
      cf[2 * d + side] = 0.5 * (fabs(vnf) + sqrt(cnds) * cc) + rsum * cnds / vsum;
    }
    coefe[d * b.nnn + m] = cf[2 * d];
  }
  dg = dg + cf[0] + cf[1] + cf[2] + cf[3] + cf[4] + cf[5];
  if (kfmg == 3) dg = dg + 1.5 * vol[n] / dtpas;
  diag[m] = dg;
}

// One point-Jacobi sub-iteration (implimf_3d.f90:230-465).  first != 0 for
// ls = 1, where d and dF are still zero.
__global__ void __launch_bounds__(TX *TY) k_impl_sweep(Blk b, Phys ph, int first, int ityprk,
                                                       const double *__restrict__ u,
                                                       const double *__restrict__ ff,
                                                       const double *__restrict__ v, long ldv,
                                                       const double *__restrict__ sn,
                                                       const double *__restrict__ coefe,
                                                       const double *__restrict__ diag, DView dold, DView dnew) {
  CELL3(b)
  long n = b.idx(c[0], c[1], c[2]), m = n - b.npc;
  double d2w[5];
#pragma unroll
  for (int e = 0; e < 5; e++) d2w[e] = ityprk == 0 ? -u[e * ldv + n] : -u[e * ldv + n] - ff[e * ldv + n];
  if (!first) {
    double Vc[5], Dc[5];
#pragma unroll
    for (int e = 0; e < 5; e++) {
      Vc[e] = v[e * ldv + n];
      Dc[e] = dold.at(e, n);
    }
    DFlux Fc = dflux(ph, Vc, Dc);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      long inc = b.inc(d);
      int a = c[d], a1 = b.lo(d), a2 = b.hi(d);
      double DL[5] = {0, 0, 0, 0, 0}, DR[5] = {0, 0, 0, 0, 0};
      DFlux FL = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, FR = FL;
      bool lreal = a > a1, rreal = a + 1 < a2;
      if (lreal) {
        double VL[5];
#pragma unroll
        for (int e = 0; e < 5; e++) {
          VL[e] = v[e * ldv + n - inc];
          DL[e] = dold.at(e, n - inc);
        }
        FL = dflux(ph, VL, DL);
      }
      if (rreal) {
        double VR[5];
#pragma unroll
        for (int e = 0; e < 5; e++) {
          VR[e] = v[e * ldv + n + inc];
          DR[e] = dold.at(e, n + inc);
        }
        FR = dflux(ph, VR, DR);
      }
      double tlo[5], thi[5];
      face_tn(Fc, FL, snv(sn, b.nnn, m, d, 0), snv(sn, b.nnn, m, d, 1), snv(sn, b.nnn, m, d, 2), tlo);
      face_tn(FR, Fc, snv(sn, b.nnn, m + inc, d, 0), snv(sn, b.nnn, m + inc, d, 1), snv(sn, b.nnn, m + inc, d, 2),
              thi);
      double cl = coefe[d * b.nnn + m];
      double ch = rreal ? coefe[d * b.nnn + m + inc] : 0.;
#pragma unroll
      for (int e = 0; e < 5; e++) {
        double t = d2w[e] + tlo[e] + cl * DL[e];
        if (rreal) t = t + ch * DR[e];
        d2w[e] = t - thi[e];
      }
    }
  }
  double dg = diag[m];
#pragma unroll
  for (int e = 0; e < 5; e++) dnew.at(e, n) = d2w[e] / dg;
}

// implimf_3d.f90:473-486
__global__ void __launch_bounds__(TX *TY) k_impl_update(Blk b, double *v, long ldv, const double *__restrict__ d) {
  CELL3(b)
  long n = b.idx(c[0], c[1], c[2]);
#pragma unroll
  for (int e = 0; e < 5; e++) v[e * ldv + n] = v[e * ldv + n] + d[e * ldv + n];
}

__global__ void k_zero5(double *d, long ldv, long n0, long n1) {
  long n = n0 + blockIdx.x * 256L + threadIdx.x;
  if (n > n1) return;
#pragma unroll
  for (int e = 0; e < 5; e++) d[e * ldv + n] = 0.;
}

extern "C" int codens_implimf_3d(const codens_block *cb, const codens_phys *cp, const double *u, const double *dt,
                                 double *v, int64_t ldv, double *d, const double *ff, const double *mu,
                                 const double *mut, int32_t lmx, const double *sn, const double *vol, double dtpas,
                                 int32_t kfmg, int32_t ityprk, const double *ps, const double *cson,
                                 void *stream) {
  (void)ps;  // p of the real cells is re-derived from v with the zpres formula
  int rc = codens_check_block(cb, "codens_implimf_3d");
  if (rc) return rc;
  if (!cp || !u || !dt || !v || !d || !mu || !mut || !sn || !vol || !cson || (ityprk && !ff) || lmx < 0) {
    codens_set_error("codens_implimf_3d: invalid argument");
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  Phys ph = make_phys(cp);
  hipStream_t st = (hipStream_t)stream;
  double *ws = (double *)codens_workspace(sizeof(double) * 9 * b.nnn, 2);
  if (!ws) {
    codens_set_error("codens_implimf_3d: workspace allocation failed");
    return CODENS_EHIP;
  }
  double *coefe = ws, *diag = ws + 3 * b.nnn, *W = ws + 4 * b.nnn;
  long n0 = b.npc, n1 = b.npc + b.nnn - 1;
  hipLaunchKernelGGL(k_zero5, dim3((b.nnn + 255) / 256), dim3(256), 0, st, d, (long)ldv, n0, n1);
  dim3 g = cell_grid(b), t(TX, TY);
  {
    CodensTimed _tm("implimf_coef", st);
    hipLaunchKernelGGL(k_impl_coef, g, t, 0, st, b, ph, v, (long)ldv, mu, mut, dt, vol, cson, sn, dtpas, (int)kfmg,
                     coefe, diag);
  }
  DView A = {d, (long)ldv, 0}, B = {W, b.nnn, b.npc};
  for (int s = 1; s <= lmx; s++) {
    DView out = ((lmx - s) % 2 == 0) ? A : B;
    DView in = ((lmx - s) % 2 == 0) ? B : A;
    {
    CodensTimed _tm("implimf_sweep", st);
    hipLaunchKernelGGL(k_impl_sweep, g, t, 0, st, b, ph, s == 1 ? 1 : 0, (int)ityprk, u, ityprk ? ff : u, v,
                       (long)ldv, sn, coefe, diag, in, out);
  }
  }
  {
    CodensTimed _tm("implimf_update", st);
    hipLaunchKernelGGL(k_impl_update, g, t, 0, st, b, v, (long)ldv, d);
  }
  return codens_check_launch("codens_implimf_3d");
}
