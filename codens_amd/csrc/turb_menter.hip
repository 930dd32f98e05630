// Menter k-omega model's part of the turbulence stage of met_num on gfx950:
// the sources with their blending function (met_smmt), the wall correction
// of the blending function along wall-normal lines (met_mtcorf1), the flux
// densities (met_fludmt) and the eddy viscosity (met_komut).
#include "common.hpp"

using namespace codens;

#define REAL3(b)                                     \
  const dim3 xb = xcd_block();                       \
  int c[3];                                          \
  c[0] = (b).i1 + xb.x * TX + threadIdx.x;           \
  c[1] = (b).j1 + xb.y * TY + threadIdx.y;           \
  c[2] = (b).k1 + xb.z;                              \
  if (c[0] >= (b).i2 || c[1] >= (b).j2) return;

namespace {
// constants as codens_menter orders them
struct Menter {
  double betae, okappa, sigme1, sigma1, beta1, wsig1, sigme2, sigma2, beta2, wsig2;
};

Menter make_menter(const double *mc) {
  return Menter{mc[0], mc[1], mc[2], mc[3], mc[4], mc[5], mc[6], mc[7], mc[8], mc[9]};
}
}  // namespace

// met_smmt.f90:63-100: gkgo = grad k . grad omega and the blending function
// frac = tanh(zeta^4) (1 above zeta = 2.5); kfracom = 0: frac = 1, gkgo = 0
__global__ void __launch_bounds__(TX *TY) k_smmt_blend(Blk b, Menter mc, int kfracom, const double *__restrict__ v,
                                                       long ldv, const double *__restrict__ mu,
                                                       const double *__restrict__ dist,
                                                       const double *__restrict__ fd, double *__restrict__ gkgo,
                                                       double *__restrict__ frac) {
  REAL3(b)
  long n = b.idx(c[0], c[1], c[2]), m = n - b.npc;
  if (kfracom == 0) {
    frac[n] = 1.;
    gkgo[m] = 0.;
    return;
  }
  double g = fd[3 * ldv + n] * fd[n] + fd[4 * ldv + n] * fd[ldv + n] + fd[5 * ldv + n] * fd[2 * ldv + n];
  gkgo[m] = g;
  double rho = v[n], rk = v[5 * ldv + n], rw = v[6 * ldv + n];
  double omeg = rw / rho;
  double coef1 = sqrt(rk / rho) / (mc.betae * omeg * dist[n]);
  double dist2 = dist[n] * dist[n];
  double coef2 = 500. * mu[n] / (rw * dist2);
  double coef4 = mc.wsig2 * rho * g / omeg;
  double dkomg = coef4 > 1e-20 ? coef4 : 1e-20;
  double coef3 = (4. * mc.wsig2) * rk / (dist2 * dkomg);
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

// met_mtcorf1.f90:81-131 for the lines of one wall: one thread per line (the
// lines of one wall are distinct cell columns).  Indexes outside [0, nfrac)
// are skipped.
__global__ void __launch_bounds__(64) k_mtcorf1(const codens_wall_line *__restrict__ lines, long nlines, int imxclko,
                                                const int32_t *__restrict__ mnpar, double *__restrict__ frac,
                                                long nfrac) {
  long t = blockIdx.x * 64L + threadIdx.x;
  if (t >= nlines) return;
  const codens_wall_line w = lines[t];
  int m3min = w.m3min, m3max = w.m3max, is = w.isens3;
  long nmin = w.nmin, isdm3 = (long)is * w.dm3;
  auto at = [&](int m3) { return nmin + (long)(m3 - m3min) * isdm3; };
  auto ok = [&](long n) { return n >= 0 && n < nfrac; };
  int m3mxx = m3min + is * (imxclko + 1);
  for (int m3 = m3min; is > 0 ? m3 <= m3mxx : m3 >= m3mxx; m3 += is)
    if (ok(at(m3))) frac[at(m3)] = 1.;
  int m3lim = 0;
  for (int m3 = m3min; is > 0 ? m3 <= m3max : m3 >= m3max; m3 += is) {
    long n = at(m3);
    if (!ok(n)) continue;
    int npar = mnpar[n];
    if (npar > w.mpn0 && npar <= w.mpn1) {
      m3mxx = m3;
      if (frac[n] > 0.99) m3lim = m3;
    }
  }
  if (is > 0 ? m3mxx > m3min : m3mxx < m3min)
    for (int m3 = m3min; is > 0 ? m3 <= m3mxx : m3 >= m3mxx; m3 += is) {
      long n = at(m3);
      if (!ok(n) || !ok(n - isdm3)) continue;
      bool one = is > 0 ? m3 <= m3lim : m3 >= m3lim;
      frac[n] = one ? 1. : fmin(frac[n], frac[n - isdm3]);
    }
}

// met_smmt.f90:106-124: the sources with the blended constants; with the
// velocity gradients g9 (met_smmtr.f90:94-119) the production of ts7 is
// divided by Durbin's alpha = min(1, (1/sqrt(0.09)) / (S/omega)) (without
// them alpha = 1, an exact division)
__global__ void __launch_bounds__(TX *TY) k_smmt_sources(Blk b, Menter mc, const double *__restrict__ v, long ldv,
                                                         const double *__restrict__ mut,
                                                         const double *__restrict__ tprod,
                                                         const double *__restrict__ gkgo,
                                                         const double *__restrict__ frac,
                                                         const double *__restrict__ g9, double *__restrict__ cfke,
                                                         double *__restrict__ ts6, double *__restrict__ ts7) {
  REAL3(b)
  long n = b.idx(c[0], c[1], c[2]), m = n - b.npc;
  double alpha = 1.;
  if (g9) {
    const long q = b.nnn;
    double dxx = g9[m], dxy = g9[q + m], dxz = g9[2 * q + m], dyx = g9[3 * q + m], dyy = g9[4 * q + m];
    double dyz = g9[5 * q + m], dzx = g9[6 * q + m], dzy = g9[7 * q + m], dzz = g9[8 * q + m];
    double a = dxx * dxx + dyy * dyy + dzz * dzz;
    double s1 = dzy + dyz, s2 = dxz + dzx, s3 = dyx + dxy;
    double ss = sqrt(4. * a / 3. + s1 * s1 + s2 * s2 + s3 * s3) / (v[6 * ldv + n] / v[n]);
    double r = (1. / sqrt(0.09)) / ss;
    alpha = r < 1. ? r : 1.;
  }
  double sb = sqrt(mc.betae), ok2 = mc.okappa * mc.okappa;
  double xgam1 = mc.beta1 / mc.betae - mc.sigma1 * ok2 / sb;
  double xgam2 = mc.beta2 / mc.betae - mc.sigma2 * ok2 / sb;
  double fm1 = frac[n], fm2 = 1. - fm1;
  double wgam = fm1 * xgam1 + fm2 * xgam2;
  double wsig = fm1 * mc.wsig1 + fm2 * mc.wsig2;
  double beta = fm1 * mc.beta1 + fm2 * mc.beta2;
  double rho = v[n], rk = v[5 * ldv + n], rw = v[6 * ldv + n];
  double omeg = rw / rho;
  double sl3 = 2. * wsig * rho * gkgo[m] / omeg;
  double p = tprod[m];
  ts6[n] = p - mc.betae * rk * omeg;
  ts7[n] = p * wgam * rho / mut[n] / alpha - beta * rw * omeg + sl3;
  cfke[n] = 0.18 * omeg;
}

// met_fludmt.f90:57-78 over indc(i1,j1,k1)..indc(i2-1,j2-1,k2-1)
__global__ void __launch_bounds__(256) k_fludmt(Menter mc, int kfracom, long n0, long n1,
                                                const double *__restrict__ mu, const double *__restrict__ mut,
                                                const double *__restrict__ frac, double *__restrict__ fd, long ldv) {
  long n = n0 + blockIdx.x * 256L + threadIdx.x;
  if (n > n1) return;
  double se = mc.sigme1, sa = mc.sigma1;
  if (kfracom != 0) {
    double fm1 = frac[n], fm2 = 1. - fm1;
    se = fm1 * mc.sigme1 + fm2 * mc.sigme2;
    sa = fm1 * mc.sigma1 + fm2 * mc.sigma2;
  }
  double ce = mu[n] + mut[n] * se, ca = mu[n] + mut[n] * sa;
#pragma unroll
  for (int q = 0; q < 3; q++) fd[q * ldv + n] = ce * fd[q * ldv + n];
#pragma unroll
  for (int q = 3; q < 6; q++) fd[q * ldv + n] = ca * fd[q * ldv + n];
}

// met_komut.f90:46-85 (komsst = 0: rho k / omega; 1: SST limiter from the
// velocity gradients g9)
__global__ void __launch_bounds__(TX *TY) k_komut(Blk b, int komsst, double betae, const double *__restrict__ v,
                                                  long ldv, const double *__restrict__ mu, double *__restrict__ mut,
                                                  const double *__restrict__ dist, const double *__restrict__ g9) {
  REAL3(b)
  long n = b.idx(c[0], c[1], c[2]), m = n - b.npc, nnn = b.nnn;
  if (komsst == 0) {
    mut[n] = v[n] * v[5 * ldv + n] / v[6 * ldv + n];
    return;
  }
  double betae2 = betae / 2., a1 = sqrt(betae);
  double omeg = v[6 * ldv + n] / v[n];
  double coef1 = sqrt(v[5 * ldv + n] / v[n]) / (betae2 * omeg * dist[n]);
  double coef2 = 500. * mu[n] / (v[6 * ldv + n] * (dist[n] * dist[n]));
  double x = g9[7 * nnn + m] - g9[5 * nnn + m], y = g9[2 * nnn + m] - g9[6 * nnn + m];
  double z = g9[3 * nnn + m] - g9[nnn + m];
  double rota = sqrt(x * x + y * y + z * z);
  double zeta = coef1 > coef2 ? coef1 : coef2;
  double z2 = 2. * (zeta * zeta);
  double exp2x = exp(z2 < 25. ? z2 : 25.);
  double f2 = (exp2x - 1.) / (exp2x + 1);
  double lim = rota * f2 / a1;
  mut[n] = v[5 * ldv + n] / (omeg > lim ? omeg : lim);
}

static int mtcorf1_launch(int32_t nfaces, const int64_t *face_start, const codens_wall_line *lines, int32_t imxclko,
                          const int32_t *mnpar, double *frac, int64_t nfrac, hipStream_t st, const char *what) {
  if (nfaces < 0 || (nfaces > 0 && (!face_start || !lines || !mnpar)) || !frac || nfrac <= 0 || imxclko < 0) {
    codens_set_error("%s: invalid wall-line arguments", what);
    return CODENS_EINVAL;
  }
  for (int f = 0; f < nfaces; f++) {
    long a = face_start[f], e = face_start[f + 1];
    if (a < 0 || e < a) {
      codens_set_error("%s: wall %d: bad line range", what, f);
      return CODENS_EINVAL;
    }
    if (e == a) continue;
    // walls in list order: a later wall's lines may cross an earlier one's
    hipLaunchKernelGGL(k_mtcorf1, dim3((unsigned)((e - a + 63) / 64)), dim3(64), 0, st, lines + a, e - a,
                       (int)imxclko, mnpar, frac, (long)nfrac);
  }
  return CODENS_OK;
}

extern "C" int codens_met_mtcorf1(int32_t nfaces, const int64_t *face_start, const codens_wall_line *lines,
                                  int32_t imxclko, const int32_t *mnpar, double *frac, int64_t nfrac, void *stream) {
  int rc = mtcorf1_launch(nfaces, face_start, lines, imxclko, mnpar, frac, nfrac, (hipStream_t)stream,
                          "codens_met_mtcorf1");
  if (rc) return rc;
  return codens_check_launch("codens_met_mtcorf1");
}

extern "C" int codens_met_smmt(const codens_block *cb, const double *mc, int32_t kfracom, int32_t imxclko,
                               const double *v, int64_t ldv, const double *mu, const double *mut, const double *dist,
                               const double *fd, const double *tprod, double *cfke, double *gkgo, double *frac,
                               double *ts6, double *ts7, int32_t nfaces, const int64_t *face_start,
                               const codens_wall_line *lines, const int32_t *mnpar, void *stream) {
  int rc = codens_check_block(cb, "codens_met_smmt");
  if (rc) return rc;
  if (!mc || !v || !mu || !mut || !dist || (kfracom && !fd) || !tprod || !cfke || !gkgo || !frac || !ts6 || !ts7) {
    codens_set_error("codens_met_smmt: null argument");
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  Menter m = make_menter(mc);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_smmt_blend, cell_grid(b), dim3(TX, TY), 0, st, b, m, (int)kfracom, v, (long)ldv, mu, dist,
                     fd, gkgo, frac);
  if (kfracom != 0 && nfaces > 0) {
    rc = mtcorf1_launch(nfaces, face_start, lines, imxclko, mnpar, frac, ldv, st, "codens_met_smmt");
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_smmt_sources, cell_grid(b), dim3(TX, TY), 0, st, b, m, v, (long)ldv, mut, tprod, gkgo, frac,
                     (const double *)nullptr, cfke, ts6, ts7);
  return codens_check_launch("codens_met_smmt");
}

extern "C" int codens_met_smmtr(const codens_block *cb, const double *mc, int32_t imxclko, const double *v,
                                int64_t ldv, const double *mu, const double *mut, const double *dist,
                                const double *fd, const double *g9, const double *tprod, double *cfke, double *frac,
                                double *ts6, double *ts7, int32_t nfaces, const int64_t *face_start,
                                const codens_wall_line *lines, const int32_t *mnpar, void *stream) {
  int rc = codens_check_block(cb, "codens_met_smmtr");
  if (rc) return rc;
  if (!mc || !v || !mu || !mut || !dist || !fd || !g9 || !tprod || !cfke || !frac || !ts6 || !ts7) {
    codens_set_error("codens_met_smmtr: null argument");
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  Menter m = make_menter(mc);
  hipStream_t st = (hipStream_t)stream;
  double *gkgo = (double *)codens_workspace(sizeof(double) * b.nnn, 5);  // grad k . grad omega per cell
  if (!gkgo) {
    codens_set_error("codens_met_smmtr: workspace allocation failed");
    return CODENS_EHIP;
  }
  hipLaunchKernelGGL(k_smmt_blend, cell_grid(b), dim3(TX, TY), 0, st, b, m, 1, v, (long)ldv, mu, dist, fd, gkgo,
                     frac);
  if (nfaces > 0) {
    rc = mtcorf1_launch(nfaces, face_start, lines, imxclko, mnpar, frac, ldv, st, "codens_met_smmtr");
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_smmt_sources, cell_grid(b), dim3(TX, TY), 0, st, b, m, v, (long)ldv, mut, tprod, gkgo, frac,
                     g9, cfke, ts6, ts7);
  return codens_check_launch("codens_met_smmtr");
}

extern "C" int codens_met_fludmt(const codens_block *cb, const double *mc, int32_t kfracom, const double *mu,
                                 const double *mut, const double *frac, double *fd, int64_t ldv, void *stream) {
  int rc = codens_check_block(cb, "codens_met_fludmt");
  if (rc) return rc;
  if (!mc || !mu || !mut || (kfracom && !frac) || !fd) {
    codens_set_error("codens_met_fludmt: null argument");
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  long n0 = b.idx(b.i1, b.j1, b.k1), n1 = b.idx(b.i2 - 1, b.j2 - 1, b.k2 - 1);
  hipLaunchKernelGGL(k_fludmt, dim3((unsigned)((n1 - n0 + 256) / 256)), dim3(256), 0, (hipStream_t)stream,
                     make_menter(mc), (int)kfracom, n0, n1, mu, mut, frac, fd, (long)ldv);
  return codens_check_launch("codens_met_fludmt");
}

extern "C" int codens_teq_gradv(const codens_block *cb, const double *sn, const double *vol, double *v, int64_t ldv,
                                double *g9, void *stream);

extern "C" int codens_met_komut(const codens_block *cb, double betae, int32_t komsst, const double *sn_blk,
                                const double *vol, double *v, int64_t ldv, const double *mu, double *mut,
                                const double *dist, double *g9, void *stream) {
  int rc = codens_check_block(cb, "codens_met_komut");
  if (rc) return rc;
  if (!v || !mu || !mut || (komsst == 1 && (!sn_blk || !vol || !dist || !g9)) || (komsst != 0 && komsst != 1)) {
    codens_set_error("codens_met_komut: invalid argument");
    return CODENS_EINVAL;
  }
  if (komsst == 1) {
    rc = codens_teq_gradv(cb, sn_blk, vol, v, ldv, g9, stream);
    if (rc) return rc;
  }
  Blk b = make_blk(cb);
  hipLaunchKernelGGL(k_komut, cell_grid(b), dim3(TX, TY), 0, (hipStream_t)stream, b, (int)komsst, betae, v,
                     (long)ldv, mu, mut, dist, g9);
  return codens_check_launch("codens_met_komut");
}
