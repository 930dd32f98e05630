// Turbulence-transport kernels of met_num on gfx950.  The state t is
// v(ip11,7); the two transported variables are t[5*ldv+n], t[6*ldv+n] and
// their dissipation d[5*ldv+n], d[6*ldv+n].
#include "common.hpp"

using namespace codens;

__device__ inline double amimd(double a, double b) {
  double s = copysign(1.0, a);
  return s * fmax(0., fmin(fabs(a), b * s));
}
__device__ inline double fi1(double x, double y, double z) { return amimd(x, amimd(y, z)); }

// ---------------------------------------------------------------------------
// met_roe2o (klroe=2): one thread per entry of the zeroed box
// (i1-1..i2, j1-1..j2, k1-1..k2); each face's dissipation is evaluated by
// the two cells that share it.
struct Roe2oFace {
  double d6, d7;
};

__device__ inline double del_at(const double *__restrict__ tv, long n, long inc, int a, int a1, int a2) {
  // face difference of face `a` (between n-inc and n), mirrored one face
  // beyond the block (met_roe2o.f90:107-120)
  if (a < a1) { n += inc; }
  else if (a > a2) { n -= inc; }
  return tv[n] - tv[n - inc];
}

__device__ Roe2oFace roe2o_face(const Blk &b, const double *__restrict__ t, long ldv, const double *__restrict__ sn,
                                long n, int d, int a) {
  long inc = b.inc(d), m = n - b.npc, n1 = n - inc;
  int a1 = b.lo(d), a2 = b.hi(d);
  double rro = 2 * t[n], rro1 = 2 * t[n1];
  double u = t[ldv + n] / rro + t[ldv + n1] / rro1;
  double v = t[2 * ldv + n] / rro + t[2 * ldv + n1] / rro1;
  double w = t[3 * ldv + n] / rro + t[3 * ldv + n1] / rro1;
  double rlam = u * snv(sn, b.nnn, m, d, 0) + v * snv(sn, b.nnn, m, d, 1) + w * snv(sn, b.nnn, m, d, 2);
  Roe2oFace r;
  const double *t6 = t + 5 * ldv, *t7 = t + 6 * ldv;
  double dl = del_at(t6, n - inc, inc, a - 1, a1, a2), dc = t6[n] - t6[n1], dr = del_at(t6, n + inc, inc, a + 1, a1, a2);
  r.d6 = 0.5 * fabs(rlam) * (dc - fi1(dl, dc, dr));
  dl = del_at(t7, n - inc, inc, a - 1, a1, a2); dc = t7[n] - t7[n1]; dr = del_at(t7, n + inc, inc, a + 1, a1, a2);
  r.d7 = 0.5 * fabs(rlam) * (dc - fi1(dl, dc, dr));
  return r;
}

__global__ void __launch_bounds__(TX *TY) k_met_roe2o(Blk b, const double *__restrict__ t, long ldv,
                                                      double *__restrict__ d, const double *__restrict__ sn) {
  const dim3 xb = xcd_block();
  int c[3];
  c[0] = b.i1 - 1 + xb.x * TX + threadIdx.x;
  c[1] = b.j1 - 1 + xb.y * TY + threadIdx.y;
  c[2] = b.k1 - 1 + xb.z;
  if (c[0] > b.i2 || c[1] > b.j2) return;
  long n = b.idx(c[0], c[1], c[2]);
  double s6 = 0., s7 = 0.;
#pragma unroll
  for (int dd = 2; dd >= 0; dd--) {
    int ta = dd == 0 ? 1 : 0, tb = dd == 2 ? 1 : 2;
    bool trans = c[ta] >= b.lo(ta) && c[ta] < b.hi(ta) && c[tb] >= b.lo(tb) && c[tb] < b.hi(tb);
    if (!trans) continue;
    int a = c[dd], a1 = b.lo(dd), a2 = b.hi(dd);
    long inc = b.inc(dd);
    if (a >= a1 && a <= a2) {  // this entry is the high side of face a
      Roe2oFace f = roe2o_face(b, t, ldv, sn, n, dd, a);
      s6 -= f.d6; s7 -= f.d7;
    }
    if (a + 1 >= a1 && a + 1 <= a2) {  // low side of face a+1
      Roe2oFace f = roe2o_face(b, t, ldv, sn, n + inc, dd, a + 1);
      s6 += f.d6; s7 += f.d7;
    }
  }
  d[5 * ldv + n] = s6;
  d[6 * ldv + n] = s7;
}

// ---------------------------------------------------------------------------
// met_roe (klroe=1) with the reference's index function, which omits the +1
// of the other kernels (met_roe.f90:184-188): every access is one storage
// entry before the intended one.  One thread per entry of the zeroed range.
__device__ inline bool roe1_face_valid(const Blk &b, long p, int d) {
  long q = p + 1 - b.npc;  // entry p is face (i,j,k) with idx(i,j,k) = p + 1
  if (q < 0 || q >= b.nnn) return false;
  int c[3];
  c[0] = b.id1 + (int)(q % b.nid);
  c[1] = b.jd1 + (int)((q / b.nid) % (b.nijd / b.nid));
  c[2] = b.kd1 + (int)(q / b.nijd);
  for (int x = 0; x < 3; x++) {
    int hi = b.hi(x) - 1 + (x == d ? 1 : 0);
    if (c[x] < b.lo(x) || c[x] > hi) return false;
  }
  return true;
}

__device__ inline void roe1_face(const Blk &b, const double *__restrict__ t, long ldv, const double *__restrict__ sn,
                                 long p, int d, double &x6, double &x7) {
  long inc = b.inc(d), m = p - b.npc, n1 = p - inc;
  double s1 = snv(sn, b.nnn, m, d, 0), s2 = snv(sn, b.nnn, m, d, 1), s3 = snv(sn, b.nnn, m, d, 2);
  double cnds = sqrt(s1 * s1 + s2 * s2 + s3 * s3);
  double um = 0.5 * (t[ldv + p] / t[p] + t[ldv + n1] / t[n1]);
  double vm = 0.5 * (t[2 * ldv + p] / t[p] + t[2 * ldv + n1] / t[n1]);
  double wm = 0.5 * (t[3 * ldv + p] / t[p] + t[3 * ldv + n1] / t[n1]);
  double vn = fabs(um * s1 + vm * s2 + wm * s3) / cnds;
  x6 = vn * (t[5 * ldv + p] - t[5 * ldv + n1]);
  x7 = vn * (t[6 * ldv + p] - t[6 * ldv + n1]);
}

__global__ void __launch_bounds__(256) k_met_roe(Blk b, long n0, long n1, const double *__restrict__ t, long ldv,
                                                 double *__restrict__ d, const double *__restrict__ sn) {
  long p = n0 + blockIdx.x * 256L + threadIdx.x;
  if (p > n1) return;
  double s6 = 0., s7 = 0.;
  for (int dd = 0; dd < 3; dd++) {
    long inc = b.inc(dd);
    double x6, x7;
    if (roe1_face_valid(b, p, dd)) {
      roe1_face(b, t, ldv, sn, p, dd, x6, x7);
      s6 -= x6; s7 -= x7;
    }
    if (roe1_face_valid(b, p + inc, dd)) {
      roe1_face(b, t, ldv, sn, p + inc, dd, x6, x7);
      s6 += x6; s7 += x7;
    }
  }
  d[5 * ldv + p] = s6;
  d[6 * ldv + p] = s7;
}

extern "C" int codens_met_roe2o(const codens_block *cb, const double *t, int64_t ldv, double *d, const double *sn,
                                void *stream) {
  int rc = codens_check_block(cb, "codens_met_roe2o");
  if (rc) return rc;
  if (!t || !d || !sn) {
    codens_set_error("codens_met_roe2o: null argument");
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  hipStream_t st = (hipStream_t)stream;
  int nx = b.i2 - b.i1 + 2, ny = b.j2 - b.j1 + 2, nz = b.k2 - b.k1 + 2;
  {
    CodensTimed _tm("met_roe2o", st);
    hipLaunchKernelGGL(k_met_roe2o, dim3((nx + TX - 1) / TX, (ny + TY - 1) / TY, nz), dim3(TX, TY), 0, st, b, t,
                       (long)ldv, d, sn);
  }
  return codens_check_launch("codens_met_roe2o");
}

extern "C" int codens_met_roe(const codens_block *cb, const double *t, int64_t ldv, double *d, const double *sn,
                              void *stream) {
  int rc = codens_check_block(cb, "codens_met_roe");
  if (rc) return rc;
  if (!t || !d || !sn) {
    codens_set_error("codens_met_roe: null argument");
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  hipStream_t st = (hipStream_t)stream;
  long n0 = b.idx(b.i1 - 1, b.j1 - 1, b.k1 - 1) - 1, n1 = b.idx(b.i2, b.j2, b.k2) - 1;
  if (n0 < b.npc) n0 = b.npc;  // the entry before the block is not this block's
  {
    CodensTimed _tm("met_roe", st);
    hipLaunchKernelGGL(k_met_roe, dim3((n1 - n0 + 256) / 256), dim3(256), 0, st, b, n0, n1, t, (long)ldv, d, sn);
  }
  return codens_check_launch("codens_met_roe");
}

// ---------------------------------------------------------------------------
// Green-Gauss gradient of a cell scalar s = num/den (den may be null) at a real
// cell, gathered from its six faces (teq_grads.f90:74-420, unit weights):
// interior faces (s_c + s_nbr).S, boundary faces 2 s_ghost.S, then 1/(2 vol)
// with the guard for non-positive volumes.
__device__ inline double sval(const double *__restrict__ num, const double *__restrict__ den, long n) {
  return den ? num[n] / den[n] : num[n];
}

__device__ void grad_scalar(const Blk &b, const double *__restrict__ sn, const double *__restrict__ vol,
                            const double *__restrict__ num, const double *__restrict__ den, long n, const int c[3],
                            double g[3]) {
  long m = n - b.npc;
  double sc = sval(num, den, n);
  g[0] = g[1] = g[2] = 0.;
#pragma unroll
  for (int dd = 0; dd < 3; dd++) {
    int d = 2 - dd;
    long inc = b.inc(d);
    int a = c[d], a1 = b.lo(d), a2 = b.hi(d);
    double sl = sval(num, den, n - inc), sh = sval(num, den, n + inc);
#pragma unroll
    for (int q = 0; q < 3; q++) {
      double s = snv(sn, b.nnn, m, d, q);
      g[q] -= (a == a1) ? s * 2 * sl : (sc + sl) * s;
      double s2 = snv(sn, b.nnn, m + inc, d, q);
      g[q] += (a + 1 == a2) ? s2 * 2 * sh : (sh + sc) * s2;
    }
  }
  double ts = copysign(0.5, -vol[n]);
  double vols = (0.5 + ts) * 0.000001 + (0.5 - ts) * vol[n];
  double c0 = 0.5 / (vols + DBL_MIN);
  g[0] *= c0; g[1] *= c0; g[2] *= c0;
}

#define REAL_CELL(b)                                 \
  int c[3];                                          \
  const dim3 xb = xcd_block();                       \
  c[0] = (b).i1 + xb.x * TX + threadIdx.x;     \
  c[1] = (b).j1 + xb.y * TY + threadIdx.y;     \
  c[2] = (b).k1 + xb.z;                        \
  if (c[0] >= (b).i2 || c[1] >= (b).j2) return;      \
  long n = (b).idx(c[0], c[1], c[2]), m = n - (b).npc;

// teq_gradv.f90:79-86: rho >= 1e-20 on the box i1-1..i2, j1-1..j2, k1-1..k2
__global__ void __launch_bounds__(TX *TY) k_rho_clamp(Blk b, double *v) {
  const dim3 xb = xcd_block();
  int i = b.i1 - 1 + xb.x * TX + threadIdx.x, j = b.j1 - 1 + xb.y * TY + threadIdx.y;
  int k = b.k1 - 1 + xb.z;
  if (i > b.i2 || j > b.j2) return;
  long n = b.idx(i, j, k);
  v[n] = fmax(v[n], 1.e-20);
}

// teq_gradv.f90:87-162: grad u, grad v, grad w (block-local, 9 x nnn)
__global__ void __launch_bounds__(TX *TY) k_teq_gradv(Blk b, const double *__restrict__ sn,
                                                      const double *__restrict__ vol, const double *__restrict__ v,
                                                      long ldv, double *__restrict__ g9) {
  REAL_CELL(b)
#pragma unroll
  for (int a = 0; a < 3; a++) {
    double g[3];
    grad_scalar(b, sn, vol, v + (a + 1) * ldv, v, n, c, g);
#pragma unroll
    for (int q = 0; q < 3; q++) g9[(3 * a + q) * b.nnn + m] = g[q];
  }
}

// met_prod.f90:54-72
__global__ void __launch_bounds__(TX *TY) k_met_prod(Blk b, const double *__restrict__ mut,
                                                     const double *__restrict__ v, long ldv,
                                                     const double *__restrict__ g9, double *__restrict__ tprod) {
  REAL_CELL(b)
  long N = b.nnn;
  double xx = g9[m], xy = g9[N + m], xz = g9[2 * N + m], yx = g9[3 * N + m], yy = g9[4 * N + m];
  double yz = g9[5 * N + m], zx = g9[6 * N + m], zy = g9[7 * N + m], zz = g9[8 * N + m];
  double s6 = v[5 * ldv + n], mt = mut[n];
  const double ds3 = 2. / 3.;
  double torxx = ds3 * (-s6 - mt * (-2 * xx + yy + zz));
  double toryy = ds3 * (-s6 - mt * (-2 * yy + xx + zz));
  double torzz = ds3 * (-s6 - mt * (-2 * zz + yy + xx));
  double torxy = mt * (xy + yx), torxz = mt * (xz + zx), toryz = mt * (zy + yz);
  tprod[m] = xx * torxx + yy * toryy + zz * torzz + torxy * (xy + yx) + torxz * (xz + zx) + toryz * (zy + yz);
}

// met_gradtr.f90:107-171: fd = grad(v6/rho), grad(v7/rho) over the contiguous
// storage range first..last real cell (ghost entries inside it get 0).
__global__ void __launch_bounds__(256) k_met_gradtr(Blk b, long n0, long n1, const double *__restrict__ sn,
                                                    const double *__restrict__ vol, const double *__restrict__ v,
                                                    long ldv, double *__restrict__ fd) {
  long n = n0 + blockIdx.x * 256L + threadIdx.x;
  if (n > n1) return;
  long m = n - b.npc;
  int c[3];
  c[0] = b.id1 + (int)(m % b.nid);
  c[1] = b.jd1 + (int)((m / b.nid) % (b.nijd / b.nid));
  c[2] = b.kd1 + (int)(m / b.nijd);
  bool real = b.real(c[0], c[1], c[2]);
#pragma unroll
  for (int a = 0; a < 2; a++) {
    double g[3] = {0., 0., 0.};
    if (real) grad_scalar(b, sn, vol, v + (5 + a) * ldv, v, n, c, g);
#pragma unroll
    for (int q = 0; q < 3; q++) fd[(3 * a + q) * ldv + n] = g[q];
  }
}

struct Turb {
  double cb1, cb2, sigma, kappa, cw1, cw2, cw3, cv1, ct3, ct4, betas, beta, sigmk, sigmw, alfak, alfae;
};

// x**6 as the reference build evaluates an integer power: ((((x x) x) x) x) x
__device__ inline double pow6(double x) { return ((((x * x) * x) * x) * x) * x; }

// met_difsa.f90 + met_smsa.f90:91-222 (Spalart-Allmaras)
__global__ void __launch_bounds__(TX *TY) k_met_smsa(Blk b, Turb tc, const double *__restrict__ sn,
                                                     const double *__restrict__ vol, const double *__restrict__ v,
                                                     long ldv, const double *__restrict__ mu,
                                                     const double *__restrict__ dist, const double *__restrict__ g9,
                                                     const double *__restrict__ fd,
                                                     const double *__restrict__ delta,
                                                     const double *__restrict__ tn1, const double *__restrict__ tn2,
                                                     double *__restrict__ cfke, double *__restrict__ ts6,
                                                     double *__restrict__ ts7) {
  REAL_CELL(b)
  long N = b.nnn;
  double gd[3], gr[3];
  grad_scalar(b, sn, vol, v + 5 * ldv, nullptr, n, c, gd);  // grad(rho nu~)
  grad_scalar(b, sn, vol, v, nullptr, n, c, gr);             // grad(rho)
  double sdif = gd[0] * fd[n] + gd[1] * fd[ldv + n] + gd[2] * fd[2 * ldv + n];
  sdif = sdif * tc.cb2 / tc.sigma;
  double kappa2 = tc.kappa * tc.kappa, uns6 = 1. / 6.;
  double cv13 = tc.cv1 * tc.cv1 * tc.cv1, cw36 = pow6(tc.cw3), cv133 = 3. * cv13, ct42 = 2. * tc.ct4;
  double cb2sig = tc.cb2 / tc.sigma;
  double xy = g9[N + m], xz = g9[2 * N + m], yx = g9[3 * N + m], yz = g9[5 * N + m], zx = g9[6 * N + m];
  double zy = g9[7 * N + m];
  double a = zy - yz, bb = xz - zx, cc = yx - xy;
  double vort = sqrt(a * a + bb * bb + cc * cc);
  double v6 = v[5 * ldv + n], v1 = v[n];
  double xkhi = v6 / mu[n];
  double xkhi2 = xkhi * xkhi, xkhi3 = xkhi2 * xkhi;
  double dd = dist[n];
  if (delta) {  // met_smdes.f90:125: the DES length min(d, 0.65 delta)
    double cd = 0.65 * delta[m];
    dd = dd < cd ? dd : cd;
  }
  double dist2 = dd * dd;
  double nutilde = v6 / v1;
  double kapd2 = kappa2 * dist2;
  double fv1 = xkhi3 / (xkhi3 + cv13);
  double fv2 = 1. - (xkhi / (1 + xkhi * fv1));
  const bool sas = tn1 != nullptr;
  if (delta || sas) fv2 = fabs(fv2);  // met_smdes.f90:128, met_smsasas.f90:214
  double stilde = vort + nutilde * fv2 / kapd2;
  double rtilde = nutilde / (stilde * kapd2);
  if (!sas) rtilde = fmin(rtilde, 1.);  // met_smsasas has no clamp
  double rtil6 = pow6(rtilde);
  double ft2 = tc.ct3 * exp(-tc.ct4 * xkhi2);
  double sm1 = tc.cb1 * (1. - ft2) * stilde * v6;
  double xg = rtilde + tc.cw2 * (rtil6 - rtilde);
  double xg6 = pow6(xg);
  double fw = xg * pow((1. + cw36) / (xg6 + cw36), uns6);
  double sm2 = -(tc.cw1 * fw - tc.cb1 * ft2 / kappa2) * v6 * nutilde / dist2;
  if (sas) {  // met_smsasas.f90:218-237: the SAS destruction where fd > 0.8
    double r8 = 8. * rtilde;
    double fdd = 1. - tanh((r8 * r8) * r8);
    if (fdd > 0.8) {
      double xx = g9[m], yy = g9[4 * N + m], zz = g9[8 * N + m];
      double s1 = zy + yz, s2 = xz + zx, s3 = yx + xy;
      double ss = sqrt(4. / 3. * (xx * xx + yy * yy + zz * zz) + s1 * s1 + s2 * s2 + s3 * s3);
      double t1 = tn1[m], t2 = tn2[m];
      double xlvk = tc.kappa * ss / sqrt(t1 * t1 + t2 * t2);
      double xlvk2 = xlvk * xlvk;
      sm2 = -tc.cw1 * fw * v6 * 1. * nutilde / xlvk2 + tc.cb1 * ft2 * v6 * nutilde / (kappa2 * dist2);
    }
  }
  ts6[n] = sm1 + sm2 + sdif;
  ts7[n] = 0.;
  double xkhi4 = xkhi3 * xkhi;
  double dfv1 = cv133 * (fv1 * fv1) / (mu[n] * xkhi4);
  double dfv2 = (xkhi2 * dfv1 - 1. / mu[n]) / ((1. + xkhi * fv1) * (1. + xkhi * fv1));
  double dst = (fv2 / v1 + nutilde * dfv2) / kapd2;
  double dpr = (1. / v1 - nutilde * dst / stilde) / (stilde * kapd2);
  double dxg = dpr * (1. + tc.cw2 * (6. * rtil6 / rtilde - 1.));
  double dfw = (1. - xg6 / (xg6 + cw36)) * fw * dxg / xg;
  double dft2 = -ct42 * xkhi * ft2;
  double dsm1 = tc.cb1 * ((1. - ft2) * (stilde + dst * v6) - stilde * v6 * dft2);
  double dsm2 = 2. * sm2 / v6 - (tc.cw1 * dfw - tc.cb1 * dft2 / kappa2) * nutilde * v6 / dist2;
  // the reference reads grad(nu~) at the block-local index here
  // (met_smsa.f90:216-217), the entry n only for the first block
  double dsdif = -cb2sig * (fd[m] * gr[0] + fd[ldv + m] * gr[1] + fd[2 * ldv + m] * gr[2]) / (v1 * v1);
  cfke[n] = fabs(dsm1 + dsm2 + dsdif);
}

// ---- Smith k-l model: constants kc = cklb1, ckle2, sigmal, xkappa
// (atlecdon.f90:226-231) ----

// sqrt(v6/v1) over the block's storage (met_bark.f90:79-87 fills the box
// i1-1..i2, j1-1..j2, k1-1..k2 of it), block-local
__global__ void __launch_bounds__(256) k_sqrt_k(long npc, long nnn, const double *__restrict__ v, long ldv,
                                                double *__restrict__ tk) {
  long m = blockIdx.x * 256L + threadIdx.x;
  if (m >= nnn) return;
  long n = npc + m;
  tk[m] = sqrt(v[5 * ldv + n] / v[n]);
}

// met_smkl.f90:163-226 with met_bark.f90:88-116: ts6 = P - csk2 sqrt(k) rho
// k / l + bark, ts7 = csl1 rho sqrt(k)(1 - l^2/(kappa d)^2) + rho l div u -
// mut |grad l|^2 l^2/(kappa d)^2 / (sigmal l) + (2/sigmal) mut grad l .
// grad k / k, cfke the larger modulus of the eigenvalues of the source
// Jacobian; bark = -2 mu |grad sqrt(k/rho)|^2, grad rho, both Green-Gauss
__global__ void __launch_bounds__(TX *TY) k_met_smkl(Blk b, double cklb1, double ckle2, double sigmal, double xkappa,
                                                     const double *__restrict__ sn, const double *__restrict__ vol,
                                                     const double *__restrict__ v, long ldv,
                                                     const double *__restrict__ mu, const double *__restrict__ mut,
                                                     const double *__restrict__ dist, const double *__restrict__ fd,
                                                     const double *__restrict__ g9, const double *__restrict__ tprod,
                                                     const double *__restrict__ tk,
                                                     const double *__restrict__ lap, double *__restrict__ cfke,
                                                     double *__restrict__ ts6, double *__restrict__ ts7) {
  REAL_CELL(b)
  const long N = b.nnn;
  double gt[3], gr[3];
  grad_scalar(b, sn, vol, tk - b.npc, nullptr, n, c, gt);
  grad_scalar(b, sn, vol, v, nullptr, n, c, gr);
  double bark = -2. * mu[n] * (gt[0] * gt[0] + gt[1] * gt[1] + gt[2] * gt[2]);
  double csk2 = 2. * sqrt(2.) / cklb1, csl1 = sqrt(2.) * (2. - ckle2) / cklb1;
  double dssigl = 2. / sigmal, c132 = 1.5 * csk2, csl12 = 2. * csl1;
  double v1 = v[n], v6 = v[5 * ldv + n], v7 = v[6 * ldv + n];
  double f5x = fd[n], f5y = fd[ldv + n], f5z = fd[2 * ldv + n];
  double f6x = fd[3 * ldv + n], f6y = fd[4 * ldv + n], f6z = fd[5 * ldv + n];
  double xk = v6 / v1, xl = v7 / v1;
  double kd = xkappa * dist[n];
  double xkapad2 = kd * kd;
  double xlskap2 = xl * xl / xkapad2;
  double rack = sqrt(xk), racrk = sqrt(v6), racro = sqrt(v1);
  double divv = g9[m] + g9[4 * N + m] + g9[8 * N + m];
  double glgl = f6x * f6x + f6y * f6y + f6z * f6z;
  double glgk = (f6x * f5x + f6y * f5y + f6z * f5z) / xk;
  double sk2 = csk2 * rack * v6 / xl;
  double sl1 = csl1 * v1 * rack * (1. - xlskap2);
  double sl2 = v7 * divv;
  double sl3 = mut[n] * glgl * xlskap2 / (sigmal * xl);
  double sl4 = dssigl * mut[n] * glgk;
  ts6[n] = tprod[m] - sk2 + bark;
  if (lap) {
    // met_smklsas.f90:238-253: the scale-adaptive term where the shielding
    // function 1 - tanh((8 rtilde)^3) exceeds 0.8
    double sl5 = 0.;
    double xx = g9[m], xy = g9[N + m], xz = g9[2 * N + m], yx = g9[3 * N + m], yy = g9[4 * N + m];
    double yz = g9[5 * N + m], zx = g9[6 * N + m], zy = g9[7 * N + m], zz = g9[8 * N + m];
    double nutilde = mut[n] / v1;
    double dvdv = xx * xx + xy * xy + yx * yx + yy * yy;
    double rtilde = nutilde / (sqrt(dvdv) * xkapad2);
    double r8 = 8. * rtilde;
    double rtil3 = (r8 * r8) * r8;
    if (1. - tanh(rtil3) > 0.8) {
      const double qsi2 = 1.47, cc43 = 4. / 3.;
      double a = zy + yz, cc = xz + zx, e = yx + xy;
      double ss = sqrt(cc43 * (xx * xx + yy * yy + zz * zz) + a * a + cc * cc + e * e);
      double l1 = lap[m], l2 = lap[N + m];
      double nlaplu = sqrt(l1 * l1 + l2 * l2);
      double xlvk = xkappa * ss / nlaplu;
      sl5 = qsi2 * tprod[m] * ((xl * xl) * xl) / (xk * (xlvk * xlvk));
    }
    ts7[n] = sl1 + sl2 - sl3 + sl4 - sl5;
  } else {
    ts7[n] = sl1 + sl2 - sl3 + sl4;
  }
  double gkgro = (f5x * gr[0] + f5y * gr[1] + f5z * gr[2]) / v6;
  double glgro = (f6x * gr[0] + f6y * gr[1] + f6z * gr[2]) / v1;
  double dqkdk = -c132 * racro * racrk / v7 - bark / v6 + mu[n] * gkgro / v1;
  double dqkdl = sk2 / v7;
  double dqldl = -csl12 * rack * xl / xkapad2 + divv - mut[n] * (glgl - 2 * xl * glgro) / (sigmal * v1 * xkapad2);
  double dqldk = dssigl * mut[n] * (-glgk + glgro) / v6;
  double xdelta = dqkdk * dqkdk + dqldl * dqldl + fabs(-2. * dqkdk * dqldl + 4. * dqkdl * dqldk);
  double rdelta = sqrt(xdelta);
  double a1 = fabs((dqkdk + dqldl + rdelta) * 0.5), a2 = fabs((dqkdk + dqldl - rdelta) * 0.5);
  cfke[n] = a1 > a2 ? a1 : a2;
}

// met_klmut.f90:38-62: mut = fmu mu chi
__global__ void __launch_bounds__(TX *TY) k_met_klmut(Blk b, double cklb1, double xkappa, const double *__restrict__ v,
                                                      long ldv, const double *__restrict__ mu,
                                                      double *__restrict__ mut, const double *__restrict__ dist) {
  REAL_CELL(b)
  const double c1 = 25.5, c2 = 2.;
  double r2sb1 = sqrt(2.) / pow(cklb1, 1. / 3.);
  double c14 = ((c1 * c1) * c1) * c1, c22 = c2 * c2;
  double v1 = v[n], v6 = v[5 * ldv + n], v7 = v[6 * ldv + n];
  double xxi = r2sb1 * sqrt(v6 / v1) * v7 / mu[n];
  double xxi2 = xxi * xxi;
  double q = v7 / (xkappa * v1 * dist[n]);
  double f1 = exp(-50. * (q * q));
  double ss2 = (c22 + xxi2) * xxi2;
  double fmu = pow((c14 * f1 + ss2) / (c14 + ss2), 0.25);
  mut[n] = fmu * mu[n] * xxi;
}

// met_smko.f90:80-110 (compressible Wilcox k-omega)
__global__ void __launch_bounds__(TX *TY) k_met_smko(Blk b, Turb tc, const double *__restrict__ v, long ldv,
                                                     const double *__restrict__ tprod, double *__restrict__ cfke,
                                                     double *__restrict__ ts6, double *__restrict__ ts7,
                                                     const double *__restrict__ cson) {
  REAL_CELL(b)
  const double xw = 5. / 9., xmt0 = 0.1;
  double v1 = v[n], v6 = v[5 * ldv + n], v7 = v[6 * ldv + n];
  double omeg = v7 / v1;
  double xmt = sqrt(2. * v6 / v1) / cson[n];
  double f = 0.5 * (1. + copysign(1.0, xmt - xmt0)) * (xmt * xmt - xmt0 * xmt0);
  double betasc = tc.betas * (1. + 1.5 * f);
  double betac = tc.beta - tc.betas * 1.5 * f;
  ts6[n] = tprod[m] - betasc * v6 * omeg;
  ts7[n] = xw * tprod[m] * v7 / v6 - betac * v7 * omeg + 0.;
  cfke[n] = 0.18 * omeg;
}

// met_fludcsa (kfludis=2), met_fludko (5), met_fludc (1): flux densities
// over the contiguous storage range first..last real cell
__global__ void __launch_bounds__(256) k_met_flud(long n0, long n1, int kfludis, Turb tc, const double *__restrict__ v,
                                                  long ldv, const double *__restrict__ mu,
                                                  const double *__restrict__ mut, double *__restrict__ fd) {
  long n = n0 + blockIdx.x * 256L + threadIdx.x;
  if (n > n1) return;
  double f5, f6;
  if (kfludis == 2) {
    f5 = (1. / tc.sigma) * (mu[n] + v[5 * ldv + n]);
    f6 = 0.;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      fd[q * ldv + n] = f5 * fd[q * ldv + n];
      fd[(3 + q) * ldv + n] = 0.;
    }
    return;
  }
  if (kfludis == 5) {
    f5 = mu[n] + mut[n] * tc.sigmk;
    f6 = mu[n] + mut[n] * tc.sigmw;
  } else {
    f5 = mu[n] + mut[n] / tc.alfak;
    f6 = mu[n] + mut[n] / tc.alfae;
  }
#pragma unroll
  for (int q = 0; q < 3; q++) {
    fd[q * ldv + n] = f5 * fd[q * ldv + n];
    fd[(3 + q) * ldv + n] = f6 * fd[(3 + q) * ldv + n];
  }
}

// sch_turb.f90:92-301 (met_pardis left to the caller): one thread per entry of
// the zeroed range; real cells gather their six face fluxes.
__device__ inline void turb_flux(const double *__restrict__ v, long ldv, const double *__restrict__ fd, long n,
                                 double f[6]) {
  double r = v[n];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    double vq = v[(q + 1) * ldv + n] / r;
    f[q] = v[5 * ldv + n] * vq - fd[q * ldv + n];
    f[3 + q] = v[6 * ldv + n] * vq - fd[(3 + q) * ldv + n];
  }
}

__global__ void __launch_bounds__(256) k_sch_turb(Blk b, long n0, long n1, double *__restrict__ u,
                                                  const double *__restrict__ v, long ldv,
                                                  const double *__restrict__ d, const double *__restrict__ fd,
                                                  const double *__restrict__ ts6, const double *__restrict__ ts7,
                                                  const double *__restrict__ sn, const double *__restrict__ vol) {
  long n = n0 + blockIdx.x * 256L + threadIdx.x;
  if (n > n1) return;
  long m = n - b.npc;
  int c[3];
  c[0] = b.id1 + (int)(m % b.nid);
  c[1] = b.jd1 + (int)((m / b.nid) % (b.nijd / b.nid));
  c[2] = b.kd1 + (int)(m / b.nijd);
  if (!b.real(c[0], c[1], c[2])) {
    u[5 * ldv + n] = 0.;
    u[6 * ldv + n] = 0.;
    return;
  }
  double fc[6], s6 = 0., s7 = 0.;
  turb_flux(v, ldv, fd, n, fc);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    long inc = b.inc(d);
    int a = c[d], a1 = b.lo(d), a2 = b.hi(d);
    double fl[6], fh[6];
    turb_flux(v, ldv, fd, n - inc, fl);
    turb_flux(v, ldv, fd, n + inc, fh);
    double s1 = snv(sn, b.nnn, m, d, 0), s2 = snv(sn, b.nnn, m, d, 1), s3 = snv(sn, b.nnn, m, d, 2);
    double x6, x7;
    if (a == a1) {
      x6 = 2 * fl[0] * s1 + 2 * fl[1] * s2 + 2 * fl[2] * s3;
      x7 = 2 * fl[3] * s1 + 2 * fl[4] * s2 + 2 * fl[5] * s3;
    } else {
      x6 = (fc[0] + fl[0]) * s1 + (fc[1] + fl[1]) * s2 + (fc[2] + fl[2]) * s3;
      x7 = (fc[3] + fl[3]) * s1 + (fc[4] + fl[4]) * s2 + (fc[5] + fl[5]) * s3;
    }
    s6 -= x6; s7 -= x7;
    s1 = snv(sn, b.nnn, m + inc, d, 0); s2 = snv(sn, b.nnn, m + inc, d, 1); s3 = snv(sn, b.nnn, m + inc, d, 2);
    if (a + 1 == a2) {
      x6 = 2 * fh[0] * s1 + 2 * fh[1] * s2 + 2 * fh[2] * s3;
      x7 = 2 * fh[3] * s1 + 2 * fh[4] * s2 + 2 * fh[5] * s3;
    } else {
      x6 = (fh[0] + fc[0]) * s1 + (fh[1] + fc[1]) * s2 + (fh[2] + fc[2]) * s3;
      x7 = (fh[3] + fc[3]) * s1 + (fh[4] + fc[4]) * s2 + (fh[5] + fc[5]) * s3;
    }
    s6 += x6; s7 += x7;
  }
  u[5 * ldv + n] = 0.5 * s6 - vol[n] * ts6[n] - d[5 * ldv + n];
  u[6 * ldv + n] = 0.5 * s7 - vol[n] * ts7[n] - d[6 * ldv + n];
}

static Turb make_turb(const double *tc) {
  Turb t;
  double *p = &t.cb1;
  for (int i = 0; i < 16; i++) p[i] = tc[i];
  return t;
}

#define CHECK_BLOCK(name)                              \
  int rc = codens_check_block(cb, name);               \
  if (rc) return rc;                                   \
  Blk b = make_blk(cb);                                \
  hipStream_t st = (hipStream_t)stream;

extern "C" int codens_teq_gradv(const codens_block *cb, const double *sn, const double *vol, double *v, int64_t ldv,
                                double *g9, void *stream) {
  CHECK_BLOCK("codens_teq_gradv")
  if (!sn || !vol || !v || !g9) { codens_set_error("codens_teq_gradv: null argument"); return CODENS_EINVAL; }
  int nx = b.i2 - b.i1 + 2, ny = b.j2 - b.j1 + 2, nz = b.k2 - b.k1 + 2;
  {
    CodensTimed _tm("teq_gradv", st);
    hipLaunchKernelGGL(k_rho_clamp, dim3((nx + TX - 1) / TX, (ny + TY - 1) / TY, nz), dim3(TX, TY), 0, st, b, v);
    hipLaunchKernelGGL(k_teq_gradv, cell_grid(b), dim3(TX, TY), 0, st, b, sn, vol, v, (long)ldv, g9);
  }
  return codens_check_launch("codens_teq_gradv");
}

extern "C" int codens_met_prod(const codens_block *cb, const double *mut, const double *v, int64_t ldv,
                               const double *g9, double *tprod, void *stream) {
  CHECK_BLOCK("codens_met_prod")
  if (!mut || !v || !g9 || !tprod) { codens_set_error("codens_met_prod: null argument"); return CODENS_EINVAL; }
  hipLaunchKernelGGL(k_met_prod, cell_grid(b), dim3(TX, TY), 0, st, b, mut, v, (long)ldv, g9, tprod);
  return codens_check_launch("codens_met_prod");
}

extern "C" int codens_met_gradtr(const codens_block *cb, const double *sn, const double *vol, const double *v,
                                 int64_t ldv, double *fd, void *stream) {
  CHECK_BLOCK("codens_met_gradtr")
  if (!sn || !vol || !v || !fd) { codens_set_error("codens_met_gradtr: null argument"); return CODENS_EINVAL; }
  long n0 = b.idx(b.i1, b.j1, b.k1), n1 = b.idx(b.i2 - 1, b.j2 - 1, b.k2 - 1);
  hipLaunchKernelGGL(k_met_gradtr, dim3((n1 - n0 + 256) / 256), dim3(256), 0, st, b, n0, n1, sn, vol, v, (long)ldv,
                     fd);
  return codens_check_launch("codens_met_gradtr");
}

extern "C" int codens_met_smsa(const codens_block *cb, const double *tc, const double *sn, const double *vol,
                               const double *v, int64_t ldv, const double *mu, const double *dist, const double *g9,
                               const double *fd, double *cfke, double *ts6, double *ts7, void *stream) {
  CHECK_BLOCK("codens_met_smsa")
  if (!tc || !sn || !vol || !v || !mu || !dist || !g9 || !fd || !cfke || !ts6 || !ts7) {
    codens_set_error("codens_met_smsa: null argument");
    return CODENS_EINVAL;
  }
  hipLaunchKernelGGL(k_met_smsa, cell_grid(b), dim3(TX, TY), 0, st, b, make_turb(tc), sn, vol, v, (long)ldv, mu, dist,
                     g9, fd, (const double *)nullptr, (const double *)nullptr, (const double *)nullptr, cfke, ts6,
                     ts7);
  return codens_check_launch("codens_met_smsa");
}

// met_dist.f90:66-103: the longest edge of each real cell from the node
// coordinates (stored like the cells), block-local
__global__ void __launch_bounds__(TX *TY) k_met_dist(Blk b, const double *__restrict__ x, const double *__restrict__ y,
                                                     const double *__restrict__ z, double *__restrict__ delta) {
  REAL_CELL(b)
  const long ci = 1, cj = b.nid, ck = b.nijd;
  auto edge = [&](long p, long q) {
    double a = x[p] - x[q], c = y[p] - y[q], e = z[p] - z[q];
    return sqrt(a * a + c * c + e * e);
  };
  const long P[12] = {n, n + cj, n + ck, n + cj + ck, n, n + ci, n + ck, n + ci + ck, n, n + ci, n + cj, n + ci + cj};
  const long Q[12] = {n + ci, n + ci + cj, n + ci + ck, n + ci + cj + ck, n + cj, n + ci + cj, n + cj + ck,
                      n + ci + cj + ck, n + ck, n + ci + ck, n + cj + ck, n + ci + cj + ck};
  double mx = edge(P[0], Q[0]);
#pragma unroll
  for (int q = 1; q < 12; q++) {
    double e = edge(P[q], Q[q]);
    mx = e > mx ? e : mx;
  }
  delta[m] = mx;
}

extern "C" int codens_met_smdes(const codens_block *cb, const double *tc, const double *x, const double *y,
                                const double *z, double *delta, const double *sn, const double *vol, const double *v,
                                int64_t ldv, const double *mu, const double *dist, const double *g9, const double *fd,
                                double *cfke, double *ts6, double *ts7, void *stream) {
  CHECK_BLOCK("codens_met_smdes")
  if (!tc || !x || !y || !z || !delta || !sn || !vol || !v || !mu || !dist || !g9 || !fd || !cfke || !ts6 || !ts7) {
    codens_set_error("codens_met_smdes: null argument");
    return CODENS_EINVAL;
  }
  hipLaunchKernelGGL(k_met_dist, cell_grid(b), dim3(TX, TY), 0, st, b, x, y, z, delta);
  hipLaunchKernelGGL(k_met_smsa, cell_grid(b), dim3(TX, TY), 0, st, b, make_turb(tc), sn, vol, v, (long)ldv, mu, dist,
                     g9, fd, (const double *)delta, (const double *)nullptr, (const double *)nullptr, cfke, ts6,
                     ts7);
  return codens_check_launch("codens_met_smdes");
}

// met_laplaciens.f90:4 on the real cells: the face sums of (dsx, dsy, dsz)
// gathered per cell in the reference's accumulation order (directions k,
// j, i; in each the low face then the high face, except on the low boundary
// layer of a line longer than one cell, where the interior loop's high face
// comes before the boundary loop's low face; interior faces carry both
// neighbours' values, boundary faces twice the ghost's), scaled by 0.5/vol
// and summed (:314-325).  ischema = 1 is the plain form, whose j faces read
// dsy(m-ncj) for the x sum of the cell above (:112); other schemes the
// weighted form with unit weights.
__global__ void __launch_bounds__(TX *TY) k_laplacian(Blk b, int ischema, const double *__restrict__ sn,
                                                      const double *__restrict__ vol, const double *__restrict__ dsx,
                                                      const double *__restrict__ dsy, const double *__restrict__ dsz,
                                                      double *__restrict__ out) {
  REAL_CELL(b)
  const double *ds[3] = {dsx, dsy, dsz};
  double acc[3] = {0., 0., 0.};
#pragma unroll
  for (int dir = 2; dir >= 0; dir--) {
    long inc = b.inc(dir);
    int a = c[dir], a1 = b.lo(dir), a2 = b.hi(dir);
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const double *d = ds[q];
      double sl = snv(sn, b.nnn, m, dir, q), sh = snv(sn, b.nnn, m + inc, dir, q);
      double lo, hi;
      if (a > a1) {
        double below = (ischema == 1 && dir == 1 && q == 0) ? dsy[m - inc] : d[m - inc];
        lo = ischema == 1 ? sl * (d[m] + below) : (d[m] + d[m - inc]) * sl;
      } else {
        lo = sl * 2 * d[m - inc];
      }
      if (a + 1 < a2) hi = ischema == 1 ? sh * (d[m + inc] + d[m]) : (d[m + inc] + d[m]) * sh;
      else hi = sh * 2 * d[m + inc];
      if (a == a1 && a + 1 < a2) {
        acc[q] = acc[q] + hi;
        acc[q] = acc[q] - lo;
      } else {
        acc[q] = acc[q] - lo;
        acc[q] = acc[q] + hi;
      }
    }
  }
  double vl = vol[n];
  double ts = (-vl >= 0. && !signbit(-vl)) ? 0.5 : -0.5;  // sign(0.5, -vol)
  double vols = (0.5 + ts) * 0.000001 + (0.5 - ts) * vl;
  double c0 = 0.5 / vols;
  out[m] = acc[0] * c0 + acc[1] * c0 + acc[2] * c0;
}

extern "C" int codens_met_smsasas(const codens_block *cb, const double *tc, int32_t ischema, const double *sn,
                                  const double *vol, const double *v, int64_t ldv, const double *mu,
                                  const double *dist, const double *g9, const double *fd, double *cfke, double *ts6,
                                  double *ts7, void *stream) {
  CHECK_BLOCK("codens_met_smsasas")
  if (!tc || !sn || !vol || !v || !mu || !dist || !g9 || !fd || !cfke || !ts6 || !ts7) {
    codens_set_error("codens_met_smsasas: null argument");
    return CODENS_EINVAL;
  }
  // the Laplacians of u and v (met_smsasas.f90:173-193)
  double *tn = (double *)codens_workspace(sizeof(double) * 2 * b.nnn, 7);
  if (!tn) {
    codens_set_error("codens_met_smsasas: workspace allocation failed");
    return CODENS_EHIP;
  }
  const long N = b.nnn;
  hipLaunchKernelGGL(k_laplacian, cell_grid(b), dim3(TX, TY), 0, st, b, (int)ischema, sn, vol, g9, g9 + N, g9 + 2 * N,
                     tn);
  hipLaunchKernelGGL(k_laplacian, cell_grid(b), dim3(TX, TY), 0, st, b, (int)ischema, sn, vol, g9 + 3 * N, g9 + 4 * N,
                     g9 + 5 * N, tn + N);
  hipLaunchKernelGGL(k_met_smsa, cell_grid(b), dim3(TX, TY), 0, st, b, make_turb(tc), sn, vol, v, (long)ldv, mu, dist,
                     g9, fd, (const double *)nullptr, (const double *)tn, (const double *)(tn + N), cfke, ts6, ts7);
  return codens_check_launch("codens_met_smsasas");
}

extern "C" int codens_met_smko(const codens_block *cb, const double *tc, const double *v, int64_t ldv,
                               const double *tprod, double *cfke, double *ts6, double *ts7, const double *cson,
                               void *stream) {
  CHECK_BLOCK("codens_met_smko")
  if (!tc || !v || !tprod || !cfke || !ts6 || !ts7 || !cson) {
    codens_set_error("codens_met_smko: null argument");
    return CODENS_EINVAL;
  }
  hipLaunchKernelGGL(k_met_smko, cell_grid(b), dim3(TX, TY), 0, st, b, make_turb(tc), v, (long)ldv, tprod, cfke, ts6,
                     ts7, cson);
  return codens_check_launch("codens_met_smko");
}

extern "C" int codens_met_flud(const codens_block *cb, int32_t kfludis, const double *tc, const double *v,
                               int64_t ldv, const double *mu, const double *mut, double *fd, void *stream) {
  CHECK_BLOCK("codens_met_flud")
  if (!tc || (kfludis == 2 && !v) || !mu || (kfludis != 2 && !mut) || !fd ||
      (kfludis != 1 && kfludis != 2 && kfludis != 5)) {
    codens_set_error("codens_met_flud: invalid argument (kfludis %d)", kfludis);
    return CODENS_EINVAL;
  }
  long n0 = b.idx(b.i1, b.j1, b.k1), n1 = b.idx(b.i2 - 1, b.j2 - 1, b.k2 - 1);
  hipLaunchKernelGGL(k_met_flud, dim3((n1 - n0 + 256) / 256), dim3(256), 0, st, n0, n1, (int)kfludis, make_turb(tc),
                     v, (long)ldv, mu, mut, fd);
  return codens_check_launch("codens_met_flud");
}

extern "C" int codens_sch_turb(const codens_block *cb, double *u, const double *v, int64_t ldv, const double *d,
                               const double *fd, const double *ts6, const double *ts7, const double *sn,
                               const double *vol, void *stream) {
  CHECK_BLOCK("codens_sch_turb")
  if (!u || !v || !d || !fd || !ts6 || !ts7 || !sn || !vol) {
    codens_set_error("codens_sch_turb: null argument");
    return CODENS_EINVAL;
  }
  long n0 = b.idx(b.i1 - 1, b.j1 - 1, b.k1 - 1), n1 = b.idx(b.i2, b.j2, b.k2);
  {
    CodensTimed _tm("sch_turb", st);
    hipLaunchKernelGGL(k_sch_turb, dim3((n1 - n0 + 256) / 256), dim3(256), 0, st, b, n0, n1, u, v, (long)ldv, d, fd,
                       ts6, ts7, sn, vol);
  }
  return codens_check_launch("codens_sch_turb");
}

extern "C" int codens_met_smkl(const codens_block *cb, const double *kc, const double *sn, const double *vol,
                               const double *v, int64_t ldv, const double *mu, const double *mut, const double *dist,
                               const double *fd, const double *g9, const double *tprod, double *cfke, double *ts6,
                               double *ts7, void *stream) {
  CHECK_BLOCK("codens_met_smkl")
  if (!kc || !sn || !vol || !v || !mu || !mut || !dist || !fd || !g9 || !tprod || !cfke || !ts6 || !ts7) {
    codens_set_error("codens_met_smkl: null argument");
    return CODENS_EINVAL;
  }
  double *tk = (double *)codens_workspace(sizeof(double) * b.nnn, 7);
  if (!tk) {
    codens_set_error("codens_met_smkl: workspace allocation failed");
    return CODENS_EHIP;
  }
  hipLaunchKernelGGL(k_sqrt_k, dim3((unsigned)((b.nnn + 255) / 256)), dim3(256), 0, st, (long)b.npc, (long)b.nnn, v,
                     (long)ldv, tk);
  hipLaunchKernelGGL(k_met_smkl, cell_grid(b), dim3(TX, TY), 0, st, b, kc[0], kc[1], kc[2], kc[3], sn, vol, v,
                     (long)ldv, mu, mut, dist, fd, g9, tprod, (const double *)tk, (const double *)nullptr, cfke, ts6,
                     ts7);
  return codens_check_launch("codens_met_smkl");
}

extern "C" int codens_met_smklsas(const codens_block *cb, const double *kc, int32_t ischema, const double *sn,
                                  const double *vol, const double *v, int64_t ldv, const double *mu, const double *mut,
                                  const double *dist, const double *fd, const double *g9, const double *tprod,
                                  double *cfke, double *ts6, double *ts7, void *stream) {
  CHECK_BLOCK("codens_met_smklsas")
  if (!kc || !sn || !vol || !v || !mu || !mut || !dist || !fd || !g9 || !tprod || !cfke || !ts6 || !ts7) {
    codens_set_error("codens_met_smklsas: null argument");
    return CODENS_EINVAL;
  }
  double *ws = (double *)codens_workspace(sizeof(double) * 3 * b.nnn, 7);
  if (!ws) {
    codens_set_error("codens_met_smklsas: workspace allocation failed");
    return CODENS_EHIP;
  }
  const long N = b.nnn;
  double *tk = ws, *lap = ws + N;
  hipLaunchKernelGGL(k_laplacian, cell_grid(b), dim3(TX, TY), 0, st, b, (int)ischema, sn, vol, g9, g9 + N, g9 + 2 * N,
                     lap);
  hipLaunchKernelGGL(k_laplacian, cell_grid(b), dim3(TX, TY), 0, st, b, (int)ischema, sn, vol, g9 + 3 * N, g9 + 4 * N,
                     g9 + 5 * N, lap + N);
  hipLaunchKernelGGL(k_sqrt_k, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, (long)b.npc, N, v, (long)ldv, tk);
  hipLaunchKernelGGL(k_met_smkl, cell_grid(b), dim3(TX, TY), 0, st, b, kc[0], kc[1], kc[2], kc[3], sn, vol, v,
                     (long)ldv, mu, mut, dist, fd, g9, tprod, (const double *)tk, (const double *)lap, cfke, ts6, ts7);
  return codens_check_launch("codens_met_smklsas");
}

extern "C" int codens_met_klmut(const codens_block *cb, const double *kc, const double *v, int64_t ldv,
                                const double *mu, double *mut, const double *dist, void *stream) {
  CHECK_BLOCK("codens_met_klmut")
  if (!kc || !v || !mu || !mut || !dist) {
    codens_set_error("codens_met_klmut: null argument");
    return CODENS_EINVAL;
  }
  hipLaunchKernelGGL(k_met_klmut, cell_grid(b), dim3(TX, TY), 0, st, b, kc[0], kc[3], v, (long)ldv, mu, mut, dist);
  return codens_check_launch("codens_met_klmut");
}

// ---- k-epsilon family (met_num.f90 ksecmb <= 1): met_smke with met_exgr,
// met_bark, met_bare and the variants; met_kemut ----

// met_exgr.f90:47-87: ghost layers of the nine block-local velocity
// gradients copied from the adjacent layer (ex1 = 1, ex2 = 0), in three
// passes: pass 0 the k faces over the real i, j; pass 1 the i faces over
// k1-1..k2 and the real j; pass 2 the j faces over k1-1..k2 and i1-1..i2.
// One thread per (face cell, side).
__global__ void __launch_bounds__(256) k_exgr(Blk b, int pass, double *__restrict__ g9) {
  long t = blockIdx.x * 256L + threadIdx.x;
  int nx, ny;
  if (pass == 0) { nx = b.i2 - b.i1; ny = b.j2 - b.j1; }
  else if (pass == 1) { nx = b.j2 - b.j1; ny = b.k2 - b.k1 + 2; }
  else { nx = b.i2 - b.i1 + 2; ny = b.k2 - b.k1 + 2; }
  if (t >= 2L * nx * ny) return;
  int side = (int)(t / ((long)nx * ny));
  long r = t % ((long)nx * ny);
  int p = (int)(r % nx), q = (int)(r / nx);
  int i, j, k;
  long step;
  if (pass == 0) { i = b.i1 + p; j = b.j1 + q; k = side == 0 ? b.k1 - 1 : b.k2; step = b.nijd; }
  else if (pass == 1) { j = b.j1 + p; k = b.k1 - 1 + q; i = side == 0 ? b.i1 - 1 : b.i2; step = 1; }
  else { i = b.i1 - 1 + p; k = b.k1 - 1 + q; j = side == 0 ? b.j1 - 1 : b.j2; step = b.nid; }
  long s1 = side == 0 ? step : -step;
  long m = b.idx(i, j, k) - b.npc;
#pragma unroll
  for (int c = 0; c < 9; c++) {
    double *d = g9 + (long)c * b.nnn;
    d[m] = 1. * d[m + s1] + 0. * d[m + 2 * s1];
  }
}

// met_bark.f90:88-116: bark = -2 mu |grad sqrt(k/rho)|^2 (block-local)
__global__ void __launch_bounds__(TX *TY) k_met_bark(Blk b, const double *__restrict__ sn,
                                                     const double *__restrict__ vol, const double *__restrict__ mu,
                                                     const double *__restrict__ tk, double *__restrict__ bark) {
  REAL_CELL(b)
  double gt[3];
  grad_scalar(b, sn, vol, tk - b.npc, nullptr, n, c, gt);
  bark[m] = -2. * mu[n] * (gt[0] * gt[0] + gt[1] * gt[1] + gt[2] * gt[2]);
}

// met_bare.f90:57-312: bare = 2 mut mu / rho x sum over the nine velocity
// gradients (dvxx, dvxy, ..., dvzz) of |grad(dv_ab)|^2 (block-local)
__global__ void __launch_bounds__(TX *TY) k_met_bare(Blk b, const double *__restrict__ sn,
                                                     const double *__restrict__ vol, const double *__restrict__ v,
                                                     const double *__restrict__ mu, const double *__restrict__ mut,
                                                     const double *__restrict__ g9, double *__restrict__ bare) {
  REAL_CELL(b)
  double acc = 0.;
  for (int q = 0; q < 9; q++) {
    double g[3];
    grad_scalar(b, sn, vol, g9 + (long)q * b.nnn - b.npc, nullptr, n, c, g);
    acc = acc + g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
  }
  bare[m] = 2. * mut[n] * mu[n] * acc / v[n];
}

__device__ inline double ke_cfke(double c1f1, double c2f2, double tp, double s6, double s7, double bark) {
  double d = -bark;
  double x = c1f1 * tp - c2f2 * s7;
  double g = fmax(0., d * (d - 2. * (x - c2f2 * s7)));
  double y = x - c2f2 * s7;
  double rdelta = sqrt(y * y + 4 * x * s7 + g);
  double bb = x - c2f2 * s7 - d;
  double xl1 = 0.5 * (bb + rdelta) / s6, xl2 = 0.5 * (bb - rdelta) / s6;
  return fmax(fabs(xl1), fabs(xl2));
}

// the variants of met_smke: 0 met_smkes.f90:62-101 (Jones-Launder), 1
// met_smker.f90:66-84, 2 met_smkec.f90:69-110 (compressible dissipation), 3
// met_smkesas.f90:126-165 (scale-adaptive; lap: the Laplacians of u and v)
__global__ void __launch_bounds__(TX *TY) k_met_smke(Blk b, double cke1, double cke2, int variant, int lparoi,
                                                     const double *__restrict__ v, long ldv,
                                                     const double *__restrict__ mu, const double *__restrict__ tprod,
                                                     const double *__restrict__ g9, const double *__restrict__ cson,
                                                     const double *__restrict__ lap, double *__restrict__ bark,
                                                     double *__restrict__ bare, double *__restrict__ cfke,
                                                     double *__restrict__ ts6, double *__restrict__ ts7) {
  REAL_CELL(b)
  const long N = b.nnn;
  double s1 = v[n], s6 = v[5 * ldv + n], s7 = v[6 * ldv + n], tp = tprod[m];
  double esk = s7 / s6;
  double roe2sk = (s7 * s7) / s6;
  double f1 = 1.;
  if (variant == 1) {
    ts6[n] = tp - s7;
    ts7[n] = cke1 * tp * esk - cke2 * roe2sk;
    cfke[n] = ke_cfke(cke1, cke2, tp, s6, s7, 0.);
    return;
  }
  double retur = (s6 * s6) / (s7 * mu[n]);
  double arg = -(retur * retur);
  double f2 = (1. - 0.3 * exp(arg));
  double epsc = 0., pid = 0.;
  if (variant == 2) {
    double xmt = sqrt(2. * s6 / s1) / cson[n];
    epsc = 0.1 * s7 * (xmt * xmt);
  }
  double bk = bark[m], be = bare[m];
  if (lparoi == 1) {
    bk = 0.; be = 0.; f2 = 1.;
    bark[m] = 0.;
    bare[m] = 0.;
  }
  ts6[n] = variant == 2 ? tp + pid - s7 - epsc + bk : tp - s7 + bk;
  if (variant == 3) {
    double xx = g9[m], xy = g9[N + m], xz = g9[2 * N + m], yx = g9[3 * N + m], yy = g9[4 * N + m];
    double yz = g9[5 * N + m], zx = g9[6 * N + m], zy = g9[7 * N + m], zz = g9[8 * N + m];
    const double qsi2 = 1.47, xkappa = 0.41, cc43 = 4. / 3.;
    double a = zy + yz, cc = xz + zx, e = yx + xy;
    double ss = sqrt(cc43 * (xx * xx + yy * yy + zz * zz) + a * a + cc * cc + e * e);
    double l1 = lap[m], l2 = lap[N + m];
    double nlaplu = sqrt(l1 * l1 + l2 * l2);
    double xlvk = xkappa * ss / nlaplu;
    double xlvk2 = xlvk * xlvk;
    double xl = pow(s6 / s1, 0.5) / (esk);
    double xl2 = xl * xl;
    double psas = xl2 / xlvk2;
    ts7[n] = cke1 * f1 * tp * esk + qsi2 * psas * tp * esk - cke2 * f2 * roe2sk + be;
  } else {
    ts7[n] = cke1 * f1 * tp * esk - cke2 * f2 * roe2sk + be;
  }
  cfke[n] = ke_cfke(cke1 * f1, cke2 * f2, tp, s6, s7, bk);
}

extern "C" int codens_met_smke(const codens_block *cb, const double *kc, int32_t variant, int32_t lparoi,
                               int32_t ischema, const double *sn, const double *vol, const double *v, int64_t ldv,
                               const double *mu, const double *mut, const double *tprod, double *g9,
                               const double *cson, double *bark, double *bare, double *cfke, double *ts6,
                               double *ts7, void *stream) {
  CHECK_BLOCK("codens_met_smke")
  if (!kc || !sn || !vol || !v || !mu || !mut || !tprod || !g9 || !bark || !bare || !cfke || !ts6 || !ts7 ||
      (variant == 2 && !cson) || variant < 0 || variant > 3) {
    codens_set_error("codens_met_smke: null or invalid argument");
    return CODENS_EINVAL;
  }
  double *ws = (double *)codens_workspace(sizeof(double) * 3 * b.nnn, 7);
  if (!ws) {
    codens_set_error("codens_met_smke: workspace allocation failed");
    return CODENS_EHIP;
  }
  const long N = b.nnn;
  double *tk = ws, *lap = ws + N;
  long nf[3] = {2L * (b.i2 - b.i1) * (b.j2 - b.j1), 2L * (b.j2 - b.j1) * (b.k2 - b.k1 + 2),
                2L * (b.i2 - b.i1 + 2) * (b.k2 - b.k1 + 2)};
  for (int pass = 0; pass < 3; pass++)
    hipLaunchKernelGGL(k_exgr, dim3((unsigned)((nf[pass] + 255) / 256)), dim3(256), 0, st, b, pass, g9);
  hipLaunchKernelGGL(k_sqrt_k, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, (long)b.npc, N, v, (long)ldv, tk);
  hipLaunchKernelGGL(k_met_bark, cell_grid(b), dim3(TX, TY), 0, st, b, sn, vol, mu, (const double *)tk, bark);
  hipLaunchKernelGGL(k_met_bare, cell_grid(b), dim3(TX, TY), 0, st, b, sn, vol, v, mu, mut, (const double *)g9,
                     bare);
  if (variant == 3) {
    hipLaunchKernelGGL(k_laplacian, cell_grid(b), dim3(TX, TY), 0, st, b, (int)ischema, sn, vol, g9, g9 + N,
                       g9 + 2 * N, lap);
    hipLaunchKernelGGL(k_laplacian, cell_grid(b), dim3(TX, TY), 0, st, b, (int)ischema, sn, vol, g9 + 3 * N,
                       g9 + 4 * N, g9 + 5 * N, lap + N);
  }
  hipLaunchKernelGGL(k_met_smke, cell_grid(b), dim3(TX, TY), 0, st, b, kc[0], kc[1], (int)variant, (int)lparoi, v,
                     (long)ldv, mu, tprod, (const double *)g9, cson, (const double *)lap, bark, bare, cfke, ts6, ts7);
  return codens_check_launch("codens_met_smke");
}

// met_kemut.f90:44-130: form 0 Jones-Launder, 1 with the SST limiter (g9
// from teq_gradv), 2 Launder-Sharma; fmu = 1 for lparoi 1 or 2
__global__ void __launch_bounds__(TX *TY) k_met_kemut(Blk b, double cmu, int form, int lparoi,
                                                      const double *__restrict__ v, long ldv,
                                                      const double *__restrict__ mu, double *__restrict__ mut,
                                                      const double *__restrict__ dist,
                                                      const double *__restrict__ g9) {
  REAL_CELL(b)
  const long N = b.nnn;
  double v1 = v[n], v6 = v[5 * ldv + n], v7 = v[6 * ldv + n];
  double retur = (v6 * v6) / (v7 * mu[n]);
  double mut0 = cmu * retur * mu[n];
  double fmu;
  if (lparoi == 1 || lparoi == 2) {
    fmu = 1.;
  } else if (form == 2) {
    double q = 1. + retur / 50.;
    fmu = exp(-3.4 / (q * q));
  } else {
    fmu = exp(-2.5 / (1. + retur / 50.));
  }
  if (form != 1) {
    mut[n] = fmu * mut0;
    return;
  }
  double a1 = sqrt(cmu);
  double coef1 = 2. * pow(v6, 1.5) / (v7 * sqrt(v1) * dist[n]);
  double coef2 = 500. * mu[n] * cmu * v6 / (v7 * (dist[n] * dist[n]) * v1);
  double x = g9[7 * N + m] - g9[5 * N + m], y = g9[2 * N + m] - g9[6 * N + m], z = g9[3 * N + m] - g9[N + m];
  double rota = sqrt(x * x + y * y + z * z);
  double zeta = coef1 > coef2 ? coef1 : coef2;
  double z2 = 2. * (zeta * zeta);
  double exp2x = exp(z2 < 25. ? z2 : 25.);
  double f2 = (exp2x - 1.) / (exp2x + 1.);
  double lim = rota * f2 * v6 * a1 * fmu / v7;
  mut[n] = fmu * mut0 / (1. > lim ? 1. : lim);
}

extern "C" int codens_met_kemut(const codens_block *cb, double cmu, int32_t form, int32_t lparoi, const double *sn,
                                const double *vol, double *v, int64_t ldv, const double *mu, double *mut,
                                const double *dist, double *g9, void *stream) {
  CHECK_BLOCK("codens_met_kemut")
  if (!v || !mu || !mut || form < 0 || form > 2 || (form == 1 && (!sn || !vol || !dist || !g9))) {
    codens_set_error("codens_met_kemut: null or invalid argument");
    return CODENS_EINVAL;
  }
  if (form == 1) {
    int r = codens_teq_gradv(cb, sn, vol, v, ldv, g9, stream);
    if (r) return r;
  }
  hipLaunchKernelGGL(k_met_kemut, cell_grid(b), dim3(TX, TY), 0, st, b, cmu, (int)form, (int)lparoi, v, (long)ldv, mu,
                     mut, dist, (const double *)g9);
  return codens_check_launch("codens_met_kemut");
}
