// EOS, molecular viscosity, local time step and explicit update on gfx950.
// One thread per real cell, 64 (i) x 4 (j) threads per workgroup, so each
// wavefront reads one contiguous i-row of every SoA array.
#include "common.hpp"

// No FMA contraction here: p, T, c and the ghost states feed the chained
// slope ratios of sch_roe, which amplify any rounding difference, so these
// kernels round exactly as the reference does.
#pragma clang fp contract(off)

using namespace codens;

#define CELL_IJK(b)                                   \
  const dim3 xb = xcd_block();                        \
  int i = (b).i1 + xb.x * TX + threadIdx.x;     \
  int j = (b).j1 + xb.y * TY + threadIdx.y;     \
  int k = (b).k1 + xb.z;                        \
  if (i >= (b).i2 || j >= (b).j2) return;

// zpres.f90:66-79
__global__ void __launch_bounds__(TX *TY) k_zpres(Blk b, Phys ph, const double *__restrict__ v, long ldv,
                                                  double *__restrict__ ps, double *__restrict__ temp,
                                                  double *__restrict__ cson) {
  CELL_IJK(b)
  long n = b.idx(i, j, k);
  double v1 = v[n], v2 = v[ldv + n], v3 = v[2 * ldv + n], v4 = v[3 * ldv + n], v5 = v[4 * ldv + n];
  double rhoe = v5 - 0.5 * (v2 * v2 + v3 * v3 + v4 * v4) / v1;
  double p = ph.gam1 * (rhoe - ph.pinfl);
  double t = ph.gam * p / v1;
  ps[n] = p;
  temp[n] = t;
  cson[n] = sqrt(t);
}

// zvismo.f90:84-127
__global__ void __launch_bounds__(TX *TY) k_zvismo(Blk b, Phys ph, const double *__restrict__ temp,
                                                   double *__restrict__ mu, int *nbad) {
  CELL_IJK(b)
  long n = b.idx(i, j, k);
  double t = temp[n];
  double a = 110.4 / ph.tnz, usrey = 1. / ph.reynz;
  double m;
  if (t <= 0.) {
    m = 0.;
    if (nbad) atomicAdd(nbad, 1);
  } else if (ph.iflu == 1) {
    m = usrey * t * sqrt(t) * (1. + a) / (t + a);
  } else {
    m = usrey;
  }
  mu[n] = m;
}

__device__ inline double face_s2(const double *__restrict__ sn, long nnn, long m, int d) {
  double a = snv(sn, nnn, m, d, 0), b_ = snv(sn, nnn, m, d, 1), c = snv(sn, nnn, m, d, 2);
  return a * a + b_ * b_ + c * c;
}

// chronos.f90:116-194: dism = vol / sqrt(sum over directions of the larger
// squared face area), dt = eta*dism/(|u|+c), min with the viscous limit.
__global__ void __launch_bounds__(TX *TY) k_chronos(Blk b, Phys ph, double etal, int ns,
                                                    const double *__restrict__ v, long ldv,
                                                    const double *__restrict__ mu,
                                                    const double *__restrict__ mut,
                                                    const double *__restrict__ sn,
                                                    const double *__restrict__ vol,
                                                    const double *__restrict__ cson,
                                                    double *__restrict__ dt) {
  CELL_IJK(b)
  long n = b.idx(i, j, k), m = n - b.npc;
  double simax = fmax(face_s2(sn, b.nnn, m + 1, 0), face_s2(sn, b.nnn, m, 0));
  double sjmax = fmax(face_s2(sn, b.nnn, m + b.nid, 1), face_s2(sn, b.nnn, m, 1));
  double skmax = fmax(face_s2(sn, b.nnn, m + b.nijd, 2), face_s2(sn, b.nnn, m, 2));
  double dism = vol[n] / sqrt(simax + sjmax + skmax);
  double ro = v[n], t2 = v[ldv + n], t3 = v[2 * ldv + n], t4 = v[3 * ldv + n];
  double qq = (t2 * t2 + t3 * t3 + t4 * t4) / (ro * ro);
  double d = etal * (dism / (sqrt(qq) + cson[n]));
  if (ns) {
    double rv1 = ph.gam * (mu[n] / ph.pr + mut[n] / ph.prt) / ro;
    double rv = fmax(rv1, 4. / 3. * (mu[n] + mut[n]) / ro);
    double dtv = etal * 0.25 * (dism * dism) / rv;
    d = fmin(d, dtv);
  }
  dt[n] = d;
}

// sch_expli.f90:53-66
__global__ void __launch_bounds__(TX *TY) k_sch_expli(Blk b, const double *__restrict__ u, double *v, long ldv,
                                                      const double *__restrict__ dt,
                                                      const double *__restrict__ vol) {
  CELL_IJK(b)
  long n = b.idx(i, j, k);
  double dtv = dt[n] / vol[n];
#pragma unroll
  for (int e = 0; e < 5; e++) v[e * ldv + n] = v[e * ldv + n] + dtv * u[e * ldv + n];
}

// One stage of the 4-stage explicit scheme composed from the single-stage
// residual: v = v0 - alpha (dt / vol) u on the real cells, nvar planes.
// The reference's only explicit update is forward Euler, v += dt/vol u with
// u the net outgoing flux (sch_expli.f90:53-66); the stage takes the
// physical sign.  No contraction: the product and the difference round
// separately, as the C restatement does.
__global__ void __launch_bounds__(TX *TY) k_rk_stage(Blk b, double alpha, const double *__restrict__ v0,
                                                     const double *__restrict__ u, double *__restrict__ v, long ldv,
                                                     const double *__restrict__ dt, const double *__restrict__ vol,
                                                     int nvar) {
#pragma clang fp contract(off)
  CELL_IJK(b)
  long n = b.idx(i, j, k);
  double c = alpha * (dt[n] / vol[n]);
  for (int e = 0; e < nvar; e++) v[e * ldv + n] = v0[e * ldv + n] - c * u[e * ldv + n];
}

extern "C" int codens_zpres(const codens_block *cb, const codens_phys *cp, const double *v, int64_t ldv,
                            double *ps, double *temp, double *cson, void *stream) {
  int rc = codens_check_block(cb, "codens_zpres");
  if (rc) return rc;
  if (!cp || !v || !ps || !temp || !cson) {
    codens_set_error("codens_zpres: null argument");
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  {
    CodensTimed _tm("zpres", (hipStream_t)stream);
    hipLaunchKernelGGL(k_zpres, cell_grid(b), dim3(TX, TY), 0, (hipStream_t)stream, b, make_phys(cp), v,
                     (long)ldv, ps, temp, cson);
  }
  return codens_check_launch("codens_zpres");
}

extern "C" int codens_zvismo(const codens_block *cb, const codens_phys *cp, const double *temp, double *mu,
                             int32_t *nbad, void *stream) {
  int rc = codens_check_block(cb, "codens_zvismo");
  if (rc) return rc;
  if (!cp || !temp || !mu) {
    codens_set_error("codens_zvismo: null argument");
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  {
    CodensTimed _tm("zvismo", (hipStream_t)stream);
    hipLaunchKernelGGL(k_zvismo, cell_grid(b), dim3(TX, TY), 0, (hipStream_t)stream, b, make_phys(cp), temp,
                     mu, (int *)nbad);
  }
  return codens_check_launch("codens_zvismo");
}

extern "C" int codens_chronos(const codens_block *cb, const codens_phys *cp, double eta, int32_t ns,
                              const double *v, int64_t ldv, const double *mu, const double *mut,
                              const double *sn, const double *vol, const double *cson, double *dt,
                              void *stream) {
  int rc = codens_check_block(cb, "codens_chronos");
  if (rc) return rc;
  if (!cp || !v || !sn || !vol || !cson || !dt || (ns && (!mu || !mut))) {
    codens_set_error("codens_chronos: null argument");
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  hipStream_t st = (hipStream_t)stream;
  long n0 = b.idx(b.i1, b.j1, b.k1), n1 = b.idx(b.i2, b.j2, b.k2);
  // chronos.f90:110-114 zeroes dt over indc(i1,j1,k1)..indc(i2,j2,k2): the
  // ghost entries of that range here, the real cells are written below
  codens_zero_shell(b, n0, n1, dt, 0, 1, st);
  {
    CodensTimed _tm("chronos", st);
    hipLaunchKernelGGL(k_chronos, cell_grid(b), dim3(TX, TY), 0, st, b, make_phys(cp), eta, (int)ns, v,
                     (long)ldv, mu, mut, sn, vol, cson, dt);
  }
  return codens_check_launch("codens_chronos");
}

extern "C" int codens_sch_expli(const codens_block *cb, const double *u, double *v, int64_t ldv,
                                const double *dt, const double *vol, void *stream) {
  int rc = codens_check_block(cb, "codens_sch_expli");
  if (rc) return rc;
  if (!u || !v || !dt || !vol) {
    codens_set_error("codens_sch_expli: null argument");
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  {
    CodensTimed _tm("sch_expli", (hipStream_t)stream);
    hipLaunchKernelGGL(k_sch_expli, cell_grid(b), dim3(TX, TY), 0, (hipStream_t)stream, b, u, v, (long)ldv,
                     dt, vol);
  }
  return codens_check_launch("codens_sch_expli");
}

extern "C" int codens_rk_stage(const codens_block *cb, double alpha, const double *v0, const double *u, double *v,
                               int64_t ldv, int32_t nvar, const double *dt, const double *vol, void *stream) {
  int rc = codens_check_block(cb, "codens_rk_stage");
  if (rc) return rc;
  if (!v0 || !u || !v || !dt || !vol || nvar < 1 || nvar > 7) {
    codens_set_error("codens_rk_stage: invalid argument (null pointer or nvar %d outside 1..7)", nvar);
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  {
    CodensTimed _tm("rk_stage", (hipStream_t)stream);
    hipLaunchKernelGGL(k_rk_stage, cell_grid(b), dim3(TX, TY), 0, (hipStream_t)stream, b, alpha, v0, u, v,
                       (long)ldv, dt, vol, (int)nvar);
  }
  return codens_check_launch("codens_rk_stage");
}
