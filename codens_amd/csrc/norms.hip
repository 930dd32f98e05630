// Steadiness norms of an update, residu.f90:131-187 per block and the
// all-block aggregate of rscpsv.f90:168-192, on gfx950, in a fixed reduction
// order so that the result does not depend on timing.
//
// For each equation e < neq and real cell n: r = (v(n,e) - u(n,e)) / dt(n).
// Per block: L1 = (1/N) sum |r|, L2 = sqrt((1/N) sum r^2), Linf = max |r| and
// the (i, j, k) of the last cell in storage order that attains it (residu's
// s = 0.5 (sign(1, |r| - max) + 1) moves the location on ties too).  All
// blocks: sum N_l L1_l / N, sqrt(sum N_l L2_l^2 / N), max Linf_l.
//
// Three passes, each a plain launch (no atomics, no host round trip):
//  1. per block, NRM_G workgroups stride over the real cells and write one
//     partial (sum |r|, sum r^2, max |r|, its flat index) per equation;
//  2. one workgroup per block folds its partials in workgroup order;
//  3. one workgroup folds the blocks in block order.
// Output: (nb + 1) records of neq x 7 doubles (L1, L2, Linf, block, i, j, k),
// the blocks then the aggregate (whose block/i/j/k locate its Linf).
#include <algorithm>

#include "common.hpp"

using namespace codens;

namespace {
constexpr int NRM_T = 256;   // threads per workgroup of pass 1
constexpr int NRM_G = 256;   // workgroups per block in pass 1
constexpr int NRM_EQ = 7;    // at most 7 equations (residu with equat(6:7)='ke')
constexpr int NRM_HDR = 8;   // per-block header: N, npc, id1, jd1, kd1, nid, nijd, spare
constexpr long NRM_BLOCK = NRM_HDR + (long)NRM_G * 4 * NRM_EQ;  // doubles of one block's scratch
constexpr int NRM_REC = 7;   // doubles of one output record per equation

struct Acc {
  double s1, s2, mx;
  long n;
};

__device__ __forceinline__ void fold(Acc &a, const Acc &b) {
  a.s1 += b.s1;
  a.s2 += b.s2;
  if (b.mx > a.mx || (b.mx == a.mx && b.n > a.n)) {
    a.mx = b.mx;
    a.n = b.n;
  }
}
}  // namespace

// Pass 1: partials of one block (grid NRM_G x 1, NRM_T threads).
__global__ void __launch_bounds__(NRM_T) k_norms_partial(Blk b, int neq, const double *__restrict__ v,
                                                         const double *__restrict__ u, long ldv,
                                                         const double *__restrict__ dt, double *scr) {
  __shared__ double sh[4][NRM_T / 64][NRM_EQ];
  Acc a[NRM_EQ];
#pragma unroll
  for (int e = 0; e < NRM_EQ; e++) a[e] = {0.0, 0.0, 0.0, -1};
  long nx = b.i2 - b.i1, ny = b.j2 - b.j1, nz = b.k2 - b.k1;
  long total = nx * ny * nz;
  for (long t = blockIdx.x * (long)NRM_T + threadIdx.x; t < total; t += (long)NRM_G * NRM_T) {
    int i = b.i1 + (int)(t % nx);
    long q = t / nx;
    int j = b.j1 + (int)(q % ny);
    int k = b.k1 + (int)(q / ny);
    long n = b.idx(i, j, k);
    double rdt = dt[n];
#pragma unroll
    for (int e = 0; e < NRM_EQ; e++) {
      if (e < neq) {
        double r = (v[e * ldv + n] - u[e * ldv + n]) / rdt;
        double ar = fabs(r);
        a[e].s1 += ar;
        a[e].s2 += r * r;
        // cells visited in increasing n by a thread: a tie moves the location on
        if (ar >= a[e].mx) {
          a[e].mx = ar;
          a[e].n = n;
        }
      }
    }
  }
  // wavefront then workgroup, in a fixed order
#pragma unroll
  for (int e = 0; e < NRM_EQ; e++) {
    if (e >= neq) continue;
    for (int off = 32; off > 0; off >>= 1) {
      Acc o;
      o.s1 = __shfl_down(a[e].s1, off, 64);
      o.s2 = __shfl_down(a[e].s2, off, 64);
      o.mx = __shfl_down(a[e].mx, off, 64);
      o.n = __shfl_down(a[e].n, off, 64);
      fold(a[e], o);
    }
  }
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int e = 0; e < NRM_EQ; e++) {
      sh[0][w][e] = a[e].s1;
      sh[1][w][e] = a[e].s2;
      sh[2][w][e] = a[e].mx;
      sh[3][w][e] = (double)a[e].n;
    }
  }
  __syncthreads();
  double *blk = scr;
  if (threadIdx.x < neq) {
    int e = threadIdx.x;
    Acc s = {sh[0][0][e], sh[1][0][e], sh[2][0][e], (long)sh[3][0][e]};
    for (int q = 1; q < NRM_T / 64; q++) {
      Acc o = {sh[0][q][e], sh[1][q][e], sh[2][q][e], (long)sh[3][q][e]};
      fold(s, o);
    }
    double *p = blk + NRM_HDR + ((long)blockIdx.x * NRM_EQ + e) * 4;
    p[0] = s.s1;
    p[1] = s.s2;
    p[2] = s.mx;
    p[3] = (double)s.n;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    blk[0] = (double)total;
    blk[1] = (double)b.npc;
    blk[2] = (double)b.id1;
    blk[3] = (double)b.jd1;
    blk[4] = (double)b.kd1;
    blk[5] = (double)b.nid;
    blk[6] = (double)b.nijd;
    blk[7] = 0.0;
  }
}

// Pass 2: one workgroup per block; thread e folds equation e's partials.
__global__ void __launch_bounds__(64) k_norms_block(int neq, const double *__restrict__ scr, double *out) {
  int l = blockIdx.x, e = threadIdx.x;
  if (e >= neq) return;
  const double *blk = scr + (long)l * NRM_BLOCK;
  Acc s = {0.0, 0.0, 0.0, -1};
  for (int g = 0; g < NRM_G; g++) {
    const double *p = blk + NRM_HDR + ((long)g * NRM_EQ + e) * 4;
    Acc o = {p[0], p[1], p[2], (long)p[3]};
    fold(s, o);
  }
  double nmax = blk[0];
  long npc = (long)blk[1], nid = (long)blk[5], nijd = (long)blk[6];
  int id1 = (int)blk[2], jd1 = (int)blk[3], kd1 = (int)blk[4];
  long m = s.n - npc;
  double coef = 1. / nmax;  // residu.f90:111 (coef = 1./real(nmax))
  double *o = out + ((long)l * neq + e) * NRM_REC;
  o[0] = coef * s.s1;
  o[1] = sqrt(coef * s.s2);
  o[2] = s.mx;
  o[3] = (double)l;
  o[4] = s.n < 0 ? 0.0 : (double)(id1 + m % nid);
  o[5] = s.n < 0 ? 0.0 : (double)(jd1 + (m % nijd) / nid);
  o[6] = s.n < 0 ? 0.0 : (double)(kd1 + m / nijd);
}

// Pass 3: the aggregate of rscpsv.f90:168-192, blocks in order.
__global__ void __launch_bounds__(64) k_norms_all(int nb, int neq, const double *__restrict__ scr, double *out) {
  int e = threadIdx.x;
  if (e >= neq) return;
  double d1 = 0.0, d2 = 0.0, dm = 0.0, npts = 0.0;
  const double *loc = nullptr;
  for (int l = 0; l < nb; l++) {
    double nmax = scr[(long)l * NRM_BLOCK];
    const double *o = out + ((long)l * neq + e) * NRM_REC;
    d1 = d1 + nmax * o[0];
    d2 = d2 + nmax * (o[1] * o[1]);
    if (!loc || fabs(o[2]) >= dm) loc = o;
    dm = fmax(dm, fabs(o[2]));
    npts += nmax;
  }
  double *g = out + ((long)nb * neq + e) * NRM_REC;
  g[0] = d1 / npts;
  g[1] = sqrt(d2 / npts);
  g[2] = dm;
  g[3] = loc[3];
  g[4] = loc[4];
  g[5] = loc[5];
  g[6] = loc[6];
}

size_t codens_norms_scratch_doubles(int nb) { return (size_t)nb * NRM_BLOCK; }

// The three passes with the caller's scratch (codens_norms_scratch_doubles(nb)
// doubles).
int codens_norms_run(const codens_block *blocks, int nb, int neq, const double *v, const double *u, long ldv,
                     const double *dt, double *out, double *scr, hipStream_t st, const char *what) {
  for (int l = 0; l < nb; l++) {
    int rc = codens_check_block(&blocks[l], what);
    if (rc) return rc;
  }
  for (int l = 0; l < nb; l++) {
    Blk b = make_blk(&blocks[l]);
    hipLaunchKernelGGL(k_norms_partial, dim3(NRM_G), dim3(NRM_T), 0, st, b, neq, v, u, ldv, dt,
                       scr + (long)l * NRM_BLOCK);
  }
  hipLaunchKernelGGL(k_norms_block, dim3(nb), dim3(64), 0, st, neq, (const double *)scr, out);
  hipLaunchKernelGGL(k_norms_all, dim3(1), dim3(64), 0, st, nb, neq, (const double *)scr, out);
  return codens_check_launch(what);
}

extern "C" int codens_residu(const codens_block *blocks, int32_t nb, int32_t neq, const double *v, const double *u,
                             int64_t ldv, const double *dt, double *out, void *stream) {
  const char *what = "codens_residu";
  if (!blocks || nb <= 0 || neq < 1 || neq > NRM_EQ || !v || !u || !dt || !out || ldv <= 0) {
    codens_set_error("%s: invalid argument (nb=%d neq=%d)", what, nb, neq);
    return CODENS_EINVAL;
  }
  double *scr = (double *)codens_workspace(sizeof(double) * codens_norms_scratch_doubles(nb), 14);
  if (!scr) {
    codens_set_error("%s: workspace allocation failed", what);
    return CODENS_EHIP;
  }
  return codens_norms_run(blocks, nb, neq, v, u, (long)ldv, dt, out, scr, (hipStream_t)stream, what);
}

// ---------------------------------------------------------------------------
// Non-physical states: cells with rho <= 0, p <= 0 or a non-finite value
// among the first neq variables (the conditions under which the reference's
// zvismo.f90:122-127 and chrono.f90:141 abort, checked before they are
// reached).  Adds the count to *nbad (device int).
__global__ void __launch_bounds__(256) k_check_state(Blk b, int neq, Phys ph, const double *__restrict__ v,
                                                     long ldv, int *nbad) {
  long nx = b.i2 - b.i1, ny = b.j2 - b.j1, nz = b.k2 - b.k1;
  long total = nx * ny * nz;
  int bad = 0;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < total; t += (long)gridDim.x * 256L) {
    int i = b.i1 + (int)(t % nx);
    long q = t / nx;
    int j = b.j1 + (int)(q % ny);
    int k = b.k1 + (int)(q / ny);
    long n = b.idx(i, j, k);
    double r = v[n], a = v[ldv + n], c = v[2 * ldv + n], d = v[3 * ldv + n], en = v[4 * ldv + n];
    double p = ph.gam1 * (en - 0.5 * (a * a + c * c + d * d) / r - ph.pinfl);
    bool fin = true;
    for (int e = 0; e < neq; e++) fin = fin && isfinite(v[e * ldv + n]);
    if (!(r > 0.0) || !(p > 0.0) || !fin) bad++;
  }
  for (int off = 32; off > 0; off >>= 1) bad += __shfl_down(bad, off, 64);
  if ((threadIdx.x & 63) == 0 && bad) atomicAdd(nbad, bad);
}

extern "C" int codens_check_state(const codens_block *blocks, int32_t nb, int32_t neq, const codens_phys *ph,
                                  const double *v, int64_t ldv, int32_t *nbad, void *stream) {
  const char *what = "codens_check_state";
  if (!blocks || nb <= 0 || neq < 5 || !ph || !v || !nbad || ldv <= 0) {
    codens_set_error("%s: invalid argument", what);
    return CODENS_EINVAL;
  }
  Phys p = make_phys(ph);
  for (int l = 0; l < nb; l++) {
    int rc = codens_check_block(&blocks[l], what);
    if (rc) return rc;
    Blk b = make_blk(&blocks[l]);
    long cells = (long)(b.i2 - b.i1) * (b.j2 - b.j1) * (b.k2 - b.k1);
    unsigned g = (unsigned)std::min<long>(1024, (cells + 255) / 256);
    hipLaunchKernelGGL(k_check_state, dim3(g), dim3(256), 0, (hipStream_t)stream, b, neq, p, v, (long)ldv,
                       (int *)nbad);
  }
  return codens_check_launch(what);
}
