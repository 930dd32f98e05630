// JST artificial dissipation (dissip_jameson) on gfx950.  One thread per
// stored entry of the reference's zeroed range: real cells gather the
// second/fourth-difference fluxes of their six faces (each face evaluated by
// the two cells that share it), ghost entries are zeroed.
#include "common.hpp"

using namespace codens;

// pressure sensor of cell n along a direction (dissip_jameson.f90:100-110)
__device__ inline double sensor(const double *__restrict__ ps, long n, long inc) {
  return fabs(ps[n + inc] - 2. * ps[n] + ps[n - inc]) / (ps[n + inc] + 2. * ps[n] + ps[n - inc]);
}

// spectral radius of the face between n-inc and n (dissip_jameson.f90:118-127)
__device__ inline double face_radius(const double *__restrict__ v, long ldv, const double *__restrict__ sn, long nnn,
                                     const double *__restrict__ cson, long n, long inc, long m, int d) {
  double s1 = snv(sn, nnn, m, d, 0), s2 = snv(sn, nnn, m, d, 1), s3 = snv(sn, nnn, m, d, 2);
  double cnds = s1 * s1 + s2 * s2 + s3 * s3;
  double uu = 0.5 * (v[ldv + n] / v[n] + v[ldv + n - inc] / v[n - inc]);
  double vv = 0.5 * (v[2 * ldv + n] / v[n] + v[2 * ldv + n - inc] / v[n - inc]);
  double ww = 0.5 * (v[3 * ldv + n] / v[n] + v[3 * ldv + n - inc] / v[n - inc]);
  double cc = 0.5 * (cson[n] + cson[n - inc]);
  return fabs(uu * s1 + vv * s2 + ww * s3) + sqrt(cnds) * cc;
}

__global__ void __launch_bounds__(256) k_jst(Blk b, long n0, long n1, double ck2, double ck4,
                                             const double *__restrict__ v, long ldv, double *__restrict__ d,
                                             const double *__restrict__ sn, const double *__restrict__ ps,
                                             const double *__restrict__ cson) {
  long n = n0 + blockIdx.x * 256L + threadIdx.x;
  if (n > n1) return;
  long m = n - b.npc;
  int c[3];
  c[0] = b.id1 + (int)(m % b.nid);
  c[1] = b.jd1 + (int)((m / b.nid) % (b.nijd / b.nid));
  c[2] = b.kd1 + (int)(m / b.nijd);
  if (!b.real(c[0], c[1], c[2])) {
#pragma unroll
    for (int e = 0; e < 5; e++) d[e * ldv + n] = 0.;
    return;
  }
  double acc[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int dd = 0; dd < 3; dd++) {
    long inc = b.inc(dd);
    int a = c[dd], a1 = b.lo(dd), a2 = b.hi(dd);
    double sc = sensor(ps, n, inc);
    // low face (between n-inc and n): subtracted from this cell
    {
      double rl = face_radius(v, ldv, sn, b.nnn, cson, n, inc, m, dd);
      if (a == a1) {
        double eps2 = ck2 * sc;
#pragma unroll
        for (int e = 0; e < 5; e++) acc[e] -= eps2 * rl * (v[e * ldv + n] - v[e * ldv + n - inc]);
      } else {
        double eps2 = ck2 * fmax(sensor(ps, n - inc, inc), sc);
        double eps4 = fmax(0., ck4 - eps2);
#pragma unroll
        for (int e = 0; e < 5; e++) {
          const double *q = v + e * ldv;
          acc[e] -= eps2 * rl * (q[n] - q[n - inc]) - eps4 * rl * (q[n + inc] + 3. * (q[n - inc] - q[n]) - q[n - 2 * inc]);
        }
      }
    }
    // high face (between n and n+inc): added to this cell
    {
      long nh = n + inc;
      double rl = face_radius(v, ldv, sn, b.nnn, cson, nh, inc, m + inc, dd);
      if (a + 1 == a2) {
        double eps2 = ck2 * sc;
#pragma unroll
        for (int e = 0; e < 5; e++) acc[e] += eps2 * rl * (v[e * ldv + nh] - v[e * ldv + n]);
      } else {
        double eps2 = ck2 * fmax(sc, sensor(ps, nh, inc));
        double eps4 = fmax(0., ck4 - eps2);
#pragma unroll
        for (int e = 0; e < 5; e++) {
          const double *q = v + e * ldv;
          acc[e] += eps2 * rl * (q[nh] - q[n]) - eps4 * rl * (q[nh + inc] + 3. * (q[n] - q[nh]) - q[n - inc]);
        }
      }
    }
  }
#pragma unroll
  for (int e = 0; e < 5; e++) d[e * ldv + n] = acc[e];
}

extern "C" int codens_dissip_jameson(const codens_block *cb, double ki2, double ki4, const double *v, int64_t ldv,
                                     double *d, const double *sn, const double *ps, const double *cson,
                                     void *stream) {
  int rc = codens_check_block(cb, "codens_dissip_jameson");
  if (rc) return rc;
  if (!v || !d || !sn || !ps || !cson) {
    codens_set_error("codens_dissip_jameson: null argument");
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  hipStream_t st = (hipStream_t)stream;
  long n0 = b.idx(b.i1 - 1, b.j1 - 1, b.k1 - 1), n1 = b.idx(b.i2, b.j2, b.k2);
  {
    CodensTimed _tm("dissip_jameson", st);
    hipLaunchKernelGGL(k_jst, dim3((n1 - n0 + 256) / 256), dim3(256), 0, st, b, n0, n1, ki2, ki4, v, (long)ldv, d,
                       sn, ps, cson);
  }
  return codens_check_launch("codens_dissip_jameson");
}
