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
  int c[3];
  c[0] = b.i1 - 1 + blockIdx.x * TX + threadIdx.x;
  c[1] = b.j1 - 1 + blockIdx.y * TY + threadIdx.y;
  c[2] = b.k1 - 1 + blockIdx.z;
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
