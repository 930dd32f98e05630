// Viscous stresses and heat flux (zvisqc -> zgrad2 + zfluto) on gfx950.
// A workgroup owns a TX x TY tile of real cells and marches a slab of up to
// ZV_KS k-planes (fewer on small blocks, slab_depth).
// Per plane, the velocities (rho u)/rho (one division each, as
// zgrad2.f90:161-180 forms vx, vy, vz) and T of the tile and of its 1-cell
// ring go to LDS for the i and j neighbours; the own column's planes k-1
// and k+1 stay in registers.  Each cell gathers the Green-Gauss sums of its
// six faces (instead of the reference's +/- face scatter) in zgrad2's order
// (k, j, i; low face, then high), scales them by 1/(2 vol) and forms tau
// and q in registers: neither the velocity planes nor the twelve gradient
// arrays reach HBM.
#include "common.hpp"

using namespace codens;

constexpr int ZV_KS = 8;
constexpr int ZV_PL = (TX + 2) * (TY + 2);

__device__ __forceinline__ void zv_q(const double *__restrict__ v, long ldv, const double *__restrict__ temp, long n,
                                     double q[4]) {
  double r = v[n];
  q[0] = v[ldv + n] / r;
  q[1] = v[2 * ldv + n] / r;
  q[2] = v[3 * ldv + n] / r;
  q[3] = temp[n];
}

__global__ void __launch_bounds__(TX *TY) k_zvisqc_tiled(Blk b, Phys ph, const double *__restrict__ sn,
                                                         const double *__restrict__ vol,
                                                         const double *__restrict__ v, long ldv,
                                                         const double *__restrict__ temp,
                                                         const double *__restrict__ mu,
                                                         const double *__restrict__ mut, double *__restrict__ tq,
                                                         int ks) {
  __shared__ double Lq[4][ZV_PL];
  const dim3 xb = xcd_block();
  const int tx = threadIdx.x, ty = threadIdx.y, t = tx + TX * ty;
  const int i0 = b.i1 + xb.x * TX, j0 = b.j1 + xb.y * TY;
  const int kb = b.k1 + xb.z * ks, ke = min(kb + ks, b.k2);
  const int i = i0 + tx, j = j0 + ty;
  const bool own = i < b.i2 && j < b.j2;
  // ring cell of this thread (t < 2 TX + 2 TY), loaded where it lies in the
  // block's storage (the faces of real cells reach only the ghost layer)
  int ri = 0, rj = 0, rslot = 0;
  const bool ring = t < 2 * TX + 2 * TY;
  if (ring) {
    int rx, ry;
    if (t < TX) { ri = i0 + t; rj = j0 - 1; rx = t + 1; ry = 0; }
    else if (t < 2 * TX) { ri = i0 + t - TX; rj = j0 + TY; rx = t - TX + 1; ry = TY + 1; }
    else if (t < 2 * TX + TY) { ri = i0 - 1; rj = j0 + t - 2 * TX; rx = 0; ry = t - 2 * TX + 1; }
    else { ri = i0 + TX; rj = j0 + t - 2 * TX - TY; rx = TX + 1; ry = t - 2 * TX - TY + 1; }
    rslot = ry * (TX + 2) + rx;
  }
  const int id2 = b.id1 + (int)b.nid - 1, jd2 = b.jd1 + (int)(b.nijd / b.nid) - 1;
  const bool rneed = ring && ri >= b.id1 && ri <= id2 && rj >= b.jd1 && rj <= jd2;
  const int slot = (ty + 1) * (TX + 2) + (tx + 1);
  // own cell's slot also holds a cell past a partial tile's edge (the ghost
  // layer), which the neighbouring real cell reads
  const bool load = i <= b.i2 && j <= b.j2;
  double qm[4] = {0, 0, 0, 0}, qc[4] = {0, 0, 0, 0}, qp[4] = {0, 0, 0, 0};
  if (own) {
    zv_q(v, ldv, temp, b.idx(i, j, kb - 1), qm);
    zv_q(v, ldv, temp, b.idx(i, j, kb), qc);
  } else if (load) {
    zv_q(v, ldv, temp, b.idx(i, j, kb), qc);
  }
  const long N = b.nnn;
  for (int k = kb; k < ke; k++) {
    if (load) {
#pragma unroll
      for (int q = 0; q < 4; q++) Lq[q][slot] = qc[q];
    }
    if (rneed) {
      double qr[4];
      zv_q(v, ldv, temp, b.idx(ri, rj, k), qr);
#pragma unroll
      for (int q = 0; q < 4; q++) Lq[q][rslot] = qr[q];
    }
    const long n = b.idx(min(i, b.i2), min(j, b.j2), k), m = n - b.npc;
    // the cell's face metrics, volume and viscosities, loaded before the
    // barrier so that their latency overlaps the plane exchange
    double sl[3][3], sh[3][3], vn = 1., mun = 0., mutn = 0.;
    if (own) {
      zv_q(v, ldv, temp, n + b.nijd, qp);
#pragma unroll
      for (int d = 0; d < 3; d++) {
#pragma unroll
        for (int cc = 0; cc < 3; cc++) {
          sl[d][cc] = snv(sn, N, m, d, cc);
          sh[d][cc] = snv(sn, N, m + b.inc(d), d, cc);
        }
      }
      vn = vol[n];
      mun = mu[n];
      mutn = mut[n];
    }
    __syncthreads();
    if (own) {
      double G[12];
#pragma unroll
      for (int q = 0; q < 12; q++) G[q] = 0.;
#pragma unroll
      for (int dd = 0; dd < 3; dd++) {
        const int d = 2 - dd;  // k faces, then j, then i (zgrad2.f90:181-457)
        const int a = d == 0 ? i : (d == 1 ? j : k), a1 = b.lo(d), a2 = b.hi(d);
        const int so = d == 0 ? 1 : (TX + 2);
        double qlo[4], qhi[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          qlo[q] = d == 2 ? qm[q] : Lq[q][slot - so];
          qhi[q] = d == 2 ? qp[q] : Lq[q][slot + so];
        }
        {  // low face (S at m): subtract
          bool bnd = (a == a1);
#pragma unroll
          for (int cc = 0; cc < 3; cc++) {
            double s = sl[d][cc];
#pragma unroll
            for (int q = 0; q < 4; q++) {
              double tt = bnd ? s * 2 * qlo[q] : (qc[q] + qlo[q]) * s;
              G[3 * q + cc] -= tt;
            }
          }
        }
        {  // high face (S at m+inc): add
          bool bnd = (a + 1 == a2);
#pragma unroll
          for (int cc = 0; cc < 3; cc++) {
            double s = sh[d][cc];
#pragma unroll
            for (int q = 0; q < 4; q++) {
              double tt = bnd ? s * 2 * qhi[q] : (qhi[q] + qc[q]) * s;
              G[3 * q + cc] += tt;
            }
          }
        }
      }
      // zgrad2.f90:463-485 (guard for non-positive volumes)
      double ts = copysign(0.5, -vn);
      double vols = (0.5 + ts) * 0.00001 + (0.5 - ts) * vn;
      double c0 = .5 / (vols + DBL_MIN);
#pragma unroll
      for (int q = 0; q < 12; q++) G[q] *= c0;
      // zfluto.f90:104-121
      double mt = mun + mutn;
      const double ds3 = 2. / 3.;
      double xx = G[0], xy = G[1], xz = G[2], yx = G[3], yy = G[4], yz = G[5], zx = G[6], zy = G[7], zz = G[8];
      tq[n] = ds3 * mt * (2 * xx - yy - zz);
      tq[ldv + n] = mt * (xy + yx);
      tq[2 * ldv + n] = mt * (xz + zx);
      tq[3 * ldv + n] = ds3 * mt * (2 * yy - xx - zz);
      tq[4 * ldv + n] = mt * (yz + zy);
      tq[5 * ldv + n] = ds3 * mt * (2 * zz - xx - yy);
      double kq = ph.cp * (mun / ph.pr + mutn / ph.prt);
      tq[6 * ldv + n] = kq * G[9];
      tq[7 * ldv + n] = kq * G[10];
      tq[8 * ldv + n] = kq * G[11];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; q++) {
      qm[q] = qc[q];
      qc[q] = qp[q];
    }
    if (!own && load && k + 1 < ke) zv_q(v, ldv, temp, b.idx(i, j, k + 1), qc);
  }
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
  CodensTimed _tm("zvisqc", (hipStream_t)stream);
  dim3 g = cell_grid(b);
  const int ks = slab_depth((long)g.x * g.y, b.k2 - b.k1, ZV_KS);
  g.z = (unsigned)((b.k2 - b.k1 + ks - 1) / ks);
  hipLaunchKernelGGL(k_zvisqc_tiled, g, dim3(TX, TY), 0, (hipStream_t)stream, b, make_phys(cp), sn, vol, v, (long)ldv,
                     temp, mu, mut, tq, ks);
  return codens_check_launch("codens_zvisqc");
}
